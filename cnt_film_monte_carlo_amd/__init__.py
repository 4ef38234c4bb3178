"""MI355X-native kinetic Monte Carlo exciton hopping over carbon-nanotube networks.

The product is libcntkmc.so (HIP kernels for gfx950 behind the C ABI of include/cntkmc.h);
this package holds its sources (csrc/), its build script (build.py), the C++ host layer
(host/) and Python bindings (kmc.py, networks.py) used by tests and benchmarks.
"""
__version__ = "0.1.0"
