"""The device prefix sum behind the grid build, the work lists and the triplet
pair lists: totals of 2^31 or more are flagged (-1) rather than wrapped, so a
caller never sizes a buffer from a wrapped total."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

I32 = 2 ** 31 - 1


def test_scan_total_flags_overflow(engine):
    assert engine.selftest_scan([I32, I32, I32]) == -1     # 3 (2^31 - 1): > 2^32
    assert engine.selftest_scan([I32, 1]) == -1             # exactly 2^31
    assert engine.selftest_scan([I32 - 5, 5]) == I32        # the largest valid total
    assert engine.selftest_scan([I32, 0, 0]) == I32


def test_scan_totals_match_numpy(engine):
    rng = np.random.default_rng(3)
    for n in (0, 1, 4095, 4096, 4097, 100000, 3 * 4096 * 1024 + 7):
        c = rng.integers(0, 50, n).astype(np.int32)
        assert engine.selftest_scan(c) == int(c.sum()), n
    # spread over many tiles, overflowing only in the sum of tile sums
    c = np.full(70000, 40000, np.int32)                     # 2.8e9 in total
    assert engine.selftest_scan(c) == -1
