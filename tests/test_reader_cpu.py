"""The command line's catalog reader (desi_amd/cli/pcf_io.h): the parallel
text parser gives the same FP32 values as a one-thread parse and as numpy's
conversion of the same text, for any thread count and any whitespace layout
(the reference's `ifstream >> float` ignores line structure,
cuda/create_grid.cuh:51-53); raw .bin catalogs round-trip; short or malformed
files are errors."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("r") / "reader_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-o", exe,
                    os.path.join(ROOT, "tests", "reader_check.cpp")], check=True)
    return exe


def _read(reader, path, n, threads, tmp_path):
    out = str(tmp_path / f"out_{threads}.f32")
    p = subprocess.run([reader, str(path), str(n), str(threads), out], capture_output=True,
                       text=True, timeout=120)
    if p.returncode != 0:
        return None, p.stdout
    return np.fromfile(out, dtype=np.float32).reshape(-1, 4), float(p.stdout)


@pytest.mark.parametrize("name,n", [("data.dat", 32768), ("rand0_1K.dat", 1000),
                                    ("data_500.dat", 400)])
def test_golden_catalogs_any_thread_count(reader, tmp_path, name, n):
    path = os.path.join(GOLDEN, "catalogs", name)
    ref = np.loadtxt(path, dtype=np.float64)[:n].astype(np.float32)
    for t in (1, 2, 7, 16):
        got, mx = _read(reader, path, n, t, tmp_path)
        assert np.array_equal(got, ref), t
        assert mx == float(max(0.0, ref[:, :3].max()))


def test_odd_whitespace(reader, tmp_path):
    rng = np.random.default_rng(3)
    vals = (rng.random((3000, 4)) * 1000).astype(np.float32)
    seps = [" ", "\t", "   ", "\r\n", "\n\n", " \t "]
    parts = []
    for i, v in enumerate(vals.ravel()):
        parts.append(repr(float(v)) if i % 5 else f"{float(v):.9e}")
        parts.append(seps[i % len(seps)])
    path = tmp_path / "odd.dat"
    path.write_text("".join(parts[:-1]))  # no trailing newline
    for t in (1, 3, 16):
        got, _ = _read(reader, path, 3000, t, tmp_path)
        assert np.array_equal(got, vals), t


def test_binary_catalog(reader, tmp_path):
    rng = np.random.default_rng(4)
    vals = (rng.random((5000, 4)) * 250).astype(np.float32)
    path = tmp_path / "cat.bin"
    vals.tofile(path)
    got, mx = _read(reader, path, 4000, 4, tmp_path)
    assert np.array_equal(got, vals[:4000])
    assert mx == float(vals[:4000, :3].max())
    got, msg = _read(reader, path, 6000, 4, tmp_path)
    assert got is None and "fewer than 6000 points" in msg


def test_short_and_malformed(reader, tmp_path):
    path = tmp_path / "short.dat"
    path.write_text("1 2 3 4\n5 6 7 8\n9 10 11\n")
    got, msg = _read(reader, path, 3, 2, tmp_path)
    assert got is None and "fewer than 3 rows" in msg
    path.write_text("1 2 3 4\n5 6 x 8\n9 10 11 12\n")
    got, msg = _read(reader, path, 3, 2, tmp_path)
    assert got is None and "row 2" in msg
    got, _ = _read(reader, path, 1, 2, tmp_path)
    assert np.array_equal(got, np.array([[1, 2, 3, 4]], np.float32))
