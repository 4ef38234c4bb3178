import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _ensure_built():
    lib = ROOT / "cnt_film_monte_carlo_amd" / "libcntkmc.so"
    orc = ROOT / "oracle" / "liboracle_kmc.so"
    if not lib.exists():
        from cnt_film_monte_carlo_amd import build as b
        b.build(verbose=False)
    if not orc.exists():
        from oracle import oracle as o
        o.build()


_ensure_built()


@pytest.fixture(scope="session")
def kmc():
    from cnt_film_monte_carlo_amd import kmc as k
    return k


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def nw():
    from cnt_film_monte_carlo_amd import networks as n
    return n


def site_arrays(pos, ori, chir=None, flags=None, left=None, right=None):
    """Dict of site arrays in the generator's format from (N,3) position/orientation arrays."""
    pos = np.asarray(pos, float)
    ori = np.asarray(ori, float)
    n = pos.shape[0]
    return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), ox=ori[:, 0].copy(),
                oy=ori[:, 1].copy(), oz=ori[:, 2].copy(),
                chirality=np.zeros(n, np.uint8) if chir is None else np.asarray(chir, np.uint8),
                flags=np.zeros(n, np.uint8) if flags is None else np.asarray(flags, np.uint8),
                left=np.full(n, -1, np.int32) if left is None else np.asarray(left, np.int32),
                right=np.full(n, -1, np.int32) if right is None else np.asarray(right, np.int32),
                n=n)


def write_mesh(d, pos_nm, orient, sep=" ", fmt="%.17g"):
    """The six single_cnt.{pos,orient}.{x,y,z}.dat matrices of a mesh directory (one row per
    tube, positions in nm), the input of create_scatterers (src/monte_carlo/monte_carlo.h:
    199-271)."""
    names = ("pos.x", "pos.y", "pos.z", "orient.x", "orient.y", "orient.z")
    for name, m in zip(names, list(pos_nm) + list(orient)):
        np.savetxt(Path(d) / f"single_cnt.{name}.dat", np.asarray(m), fmt=fmt, delimiter=sep)
