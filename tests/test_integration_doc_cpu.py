"""Every C++ block of INTEGRATION.md compiles as written against include/cntkmc.h: no elided
arguments, no undeclared names (g++ -std=c++17 -fsyntax-only, the host compiler only)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


def test_integration_has_cpp_blocks():
    assert len(_blocks()) >= 4


@pytest.mark.parametrize("k", range(4))
def test_integration_block_compiles(k, tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++ on this host")
    blocks = _blocks()
    assert k < len(blocks)
    src = tmp_path / f"block{k}.cpp"
    src.write_text(blocks[k])
    res = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                          f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
