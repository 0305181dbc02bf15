"""Compiles tests/index_check.cpp with g++ against the library's index
header and runs it: fast division, coordinates, bitmask neighbour ranks
(periodic and halo modes) and the tile offset bijection are checked on the
CPU against brute force."""
import os
import subprocess

import _pkg


def test_index_arithmetic(tmp_path):
    src = os.path.join(_pkg.ROOT, "tests", "index_check.cpp")
    inc = os.path.join(_pkg.PKG_DIR, "csrc")
    exe = str(tmp_path / "index_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", inc, "-o", exe, src], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
