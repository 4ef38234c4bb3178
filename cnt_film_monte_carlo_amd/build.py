"""Build libcntkmc.so (HIP, gfx950) in-tree with hipcc.

Every HIP source under csrc/ is compiled to an object in _build/ and linked into
cnt_film_monte_carlo_amd/libcntkmc.so, the C-ABI library declared in include/cntkmc.h.  Then the
C++ consumer of that ABI, host/kubo_main.cpp (the reference driver's Kubo loop over the
monte_carlo host layer of host/monte_carlo.hpp), is compiled and linked against it into
host/kubo_main.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcntkmc.so"
HOST = PKG / "host"
KUBO_MAIN = HOST / "kubo_main"

ARCH = os.environ.get("CNTKMC_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def kernel_source_hash() -> str:
    """SHA-256 over the library's sources (csrc/*.hip, csrc/*.h, include/*.h) and the device
    flags they are compiled with (include paths left out, so the hash does not depend on where
    the tree lies).  Hardware-counter records carry it, so that a record taken on other kernels
    is recognised as such."""
    import hashlib

    h = hashlib.sha256()
    flags = [f for f in COMMON if not f.startswith("-I")]
    h.update(" ".join(flags).encode())
    files = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.h"), *(ROOT / "include").glob("*.h")])
    for f in files:
        h.update(f.name.encode() + b"\0")
        h.update(f.read_bytes())
    return h.hexdigest()


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every csrc/*.hip and link libcntkmc.so (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_host()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {KUBO_MAIN.name}", file=sys.stderr)
    return LIB


def build_host() -> Path:
    """host/kubo_main: a C++ program over include/cntkmc.h, linked against libcntkmc.so (found
    next to it through its rpath)."""
    tmp = KUBO_MAIN.with_name(KUBO_MAIN.name + ".tmp")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{HOST}",
           str(HOST / "kubo_main.cpp"), f"-L{PKG}", "-lcntkmc", "-Wl,-rpath,$ORIGIN/..",
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"kubo_main failed to build:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, KUBO_MAIN)
    return KUBO_MAIN


if __name__ == "__main__":
    build()
