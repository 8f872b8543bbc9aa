"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer (CPU only).

- The CPU oracle and the command line's catalog reader and histogram writers,
  driven by tests/sanitize_check.cpp over every mode at small sizes, built
  with -fsanitize=address,undefined -fno-sanitize-recover=all: any report
  aborts the program.
- The multi-threaded catalog reader under ThreadSanitizer (tests/reader_check.cpp
  at 1..8 threads).
- The PCF command line itself (desi_amd/cli/pcf_main.cpp, ASan + UBSan)
  through its argument parsing, help, error messages and catalog reading, up
  to the device set-up (this container has no GPU, so the run stops at
  pcf_create with its error code).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CATS = os.path.join(ROOT, "tests", "golden", "catalogs")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _gxx(args):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread"] + args, check=True,
                   cwd=ROOT)


def _clean(r):
    text = r.stdout + r.stderr
    for marker in ("AddressSanitizer", "runtime error:", "ThreadSanitizer",
                   "UndefinedBehaviorSanitizer"):
        assert marker not in text, text[-4000:]


def test_oracle_and_io_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_check")
    _gxx(SAN + ["-fopenmp", "-ffp-contract=off", "-o", exe,
                os.path.join(ROOT, "tests", "sanitize_check.cpp"),
                os.path.join(ROOT, "oracle", "pcf_oracle.cpp")])
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=ENV,
                       timeout=300)
    _clean(r)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize_check ok" in r.stdout


def test_parallel_reader_tsan(tmp_path):
    exe = str(tmp_path / "reader_check_tsan")
    _gxx(["-fsanitize=thread", "-o", exe, os.path.join(ROOT, "tests", "reader_check.cpp")])
    cat = os.path.join(CATS, "data_5K.dat")
    for threads in (1, 2, 5, 8):
        out = str(tmp_path / f"r{threads}.f32")
        r = subprocess.run([exe, cat, "5000", str(threads), out], capture_output=True, text=True,
                           env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"), timeout=300)
        _clean(r)
        assert r.returncode == 0, r.stdout + r.stderr


def test_command_line_asan_ubsan(tmp_path):
    lib = os.path.join(ROOT, "desi_amd", "libpcf_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libpcf_hip.so not built")
    exe = str(tmp_path / "PCF")
    _gxx(SAN + ["-Wall", "-o", exe, os.path.join(ROOT, "desi_amd", "cli", "pcf_main.cpp"),
                "-I" + os.path.join(ROOT, "include"), "-L" + os.path.dirname(lib), "-lpcf_hip",
                "-Wl,-rpath," + os.path.dirname(lib)])
    data = tmp_path / "data"
    res = tmp_path / "results"
    data.mkdir()
    res.mkdir()
    for f in ("data_500.dat", "rand0_500.dat"):
        shutil.copy(os.path.join(CATS, f), data / f)

    def run(*args):
        r = subprocess.run([exe] + list(args), capture_output=True, text=True, env=ENV,
                           cwd=str(res), timeout=120)
        _clean(r)
        return r

    assert run("-h").returncode == 0
    assert run("--help").returncode == 0
    # Too few arguments, unknown type, missing -n / -b / -d, missing randoms.
    assert run("2iso", "-n", "10").returncode != 0
    assert run("4iso", "-n", "500", "-f", "data_500.dat", "-b", "10", "-d", "50").returncode != 0
    assert run("2iso", "-f", "data_500.dat", "-r", "rand0_500.dat", "-b", "10", "-d",
               "50").returncode != 0
    assert run("3iso", "-n", "500", "-f", "data_500.dat", "-b", "10", "-d", "50", "-p",
               "8").returncode != 0
    # Missing catalog file.
    assert run("2iso", "-n", "500", "-f", "nope.dat", "-r", "rand0_500.dat", "-b", "10", "-d",
               "50").returncode != 0
    # A complete command: reads both catalogs, then stops at the device set-up
    # here (no GPU in this container) or runs to the end where there is one.
    r = run("2ani", "-n", "500", "-f", "data_500.dat", "-r", "rand0_500.dat", "-b", "10", "-d",
            "50", "-bpc", "-s", "250")
    assert r.returncode == 0 or "pcf_create" in (r.stdout + r.stderr), r.stdout + r.stderr
