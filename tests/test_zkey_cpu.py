"""The z sub-slab key of resident grids and the work list's step-window rule
built on it (CPU only, host build of tests/zkey_check.cpp against
desi_amd/csrc/pcf_internal.h):

- the key is non-decreasing in z, also across slab and cell edges, and its
  cell part is the z cell the grid build assigns;
- for random geometries, blocks and z reaches, every point within the reach
  of a block point falls in the window of sub-slabs the fill takes
  ([key(first) - Kf, key(last) + Kf]).

The checks fail on planted mistakes: a window without its slab of slack, or
one bounded by the block's first key alone.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_zkey_and_windows(tmp_path):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    exe = str(tmp_path / "zkey_check")
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "desi_amd", "csrc"),
                    "-o", exe, os.path.join(ROOT, "tests", "zkey_check.cpp")],
                   check=True, cwd=ROOT, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "zkey_check ok" in r.stdout
