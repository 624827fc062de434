"""The device applies C3 (plp_common.hpp:304-312) as the integer test
2 * s >= w * d instead of the reference's FP64 `(double)s / (double)(w*d) < 0.5`
rejection.  The two verdicts agree for 0 <= s <= w*d <= 2^53; these
tests check that against IEEE FP64 division (numpy float64, correctly rounded)
exhaustively for small areas and on the boundary 2s = A +- few for areas up to
the engine's range (w*d < 2^31) and beyond it to 2^53."""
import numpy as np


def _fp64_pass(s, a):
    return ~((s.astype(np.float64) / a.astype(np.float64)) < 0.5)


def _int_pass(s, a):
    return 2 * s >= a


def test_exhaustive_small_areas():
    a = np.repeat(np.arange(1, 1201, dtype=np.int64), 1201)
    s = np.tile(np.arange(0, 1201, dtype=np.int64), 1200)
    keep = s <= a
    a, s = a[keep], s[keep]
    np.testing.assert_array_equal(_fp64_pass(s, a), _int_pass(s, a))


def _boundary(lo_bits, hi_bits, n, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(1 << lo_bits, 1 << hi_bits, size=n, dtype=np.int64)
    a = np.concatenate([a, np.array([(1 << hi_bits) - 1, (1 << hi_bits) - 2], np.int64)])
    s = np.concatenate([a // 2 + k for k in (-2, -1, 0, 1, 2)])
    a = np.tile(a, 5)
    keep = (s >= 0) & (s <= a)
    return s[keep], a[keep]


def test_boundary_engine_range():
    s, a = _boundary(20, 31, 200_000, 7)
    np.testing.assert_array_equal(_fp64_pass(s, a), _int_pass(s, a))


def test_boundary_up_to_2_53():
    s, a = _boundary(31, 53, 200_000, 11)
    np.testing.assert_array_equal(_fp64_pass(s, a), _int_pass(s, a))


def test_equivalence_breaks_past_2_53():
    # why the bound matters: at A = 2^55 + 1 neither operand is exact in FP64
    # and the quotient rounds to 0.5 although 2s < A
    a = np.array([(1 << 55) + 1], np.int64)
    s = (a - 1) // 2
    assert not _int_pass(s, a)[0]
    assert _fp64_pass(s, a)[0]
