"""The shared-divisor FP64 division the gradient and source kernels use
(hp_check_division) against the hardware quotient a / b, bit for bit, over
random bit patterns (every exponent, denormals, infinities, NaNs) and over
the operands where the shared form takes its special routes: zero and
exact numerators, and magnitudes at the edges of its [2^-300, 2^300] range."""
import ctypes

import numpy as np
import pytest

from hydraphase import _lib

pytestmark = pytest.mark.gpu


def mismatches(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    L = _lib.lib()
    n = ctypes.c_size_t()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.hp_check_division(a.ctypes.data_as(dp), b.ctypes.data_as(dp), a.size, ctypes.byref(n))
    assert rc == 0, L.hp_last_error(None).decode()
    return n.value


def scaled(rng, n, lo, hi):
    """random signs and mantissas, exponents uniform in [lo, hi]"""
    m = rng.uniform(1.0, 2.0, n) * rng.choice([-1.0, 1.0], n)
    return np.ldexp(m, rng.integers(lo, hi + 1, n))


def test_random_bit_patterns():
    rng = np.random.default_rng(11)
    n = 1 << 22
    with np.errstate(all="ignore"):
        a = rng.integers(0, 2**63, n, dtype=np.int64).view(np.float64) * rng.choice([-1.0, 1.0], n)
        b = rng.integers(0, 2**63, n, dtype=np.int64).view(np.float64)
    assert mismatches(a, b) == 0


def test_kernel_ranges_and_edges():
    rng = np.random.default_rng(12)
    n = 1 << 22
    # the magnitudes the kernels divide (geometry 1e-6..1, states 1e-12..1e7)
    a, b = scaled(rng, n, -60, 60), np.abs(scaled(rng, n, -60, 60))
    assert mismatches(a, b) == 0
    assert mismatches(a, -b) == 0
    # the edges of the shared form's range and beyond
    for lo, hi in ((-310, -290), (290, 310), (-330, 330), (-1074, -1000), (1000, 1023)):
        assert mismatches(scaled(rng, n // 8, lo, hi), scaled(rng, n // 8, -330, 330)) == 0
        assert mismatches(scaled(rng, n // 8, -330, 330), scaled(rng, n // 8, lo, hi)) == 0


def test_zero_and_exact_numerators():
    rng = np.random.default_rng(13)
    n = 1 << 20
    b = scaled(rng, n, -60, 60)
    zeros = np.where(rng.random(n) < 0.5, 0.0, -0.0)
    assert mismatches(zeros, b) == 0
    assert mismatches(zeros, -b) == 0
    # a = k * b exactly: b with a short mantissa, k a small integer
    bs = np.ldexp(rng.integers(1, 1 << 20, n).astype(np.float64), rng.integers(-80, 40, n))
    k = rng.integers(-1000, 1000, n).astype(np.float64)
    assert mismatches(k * bs, bs) == 0
    # infinities and NaNs on either side
    specials = np.array([np.inf, -np.inf, np.nan, 0.0, -0.0, 1.0, -1.0, 5e-324, 1.7976931348623157e308])
    aa, bb = np.meshgrid(specials, specials)
    with np.errstate(all="ignore"):
        assert mismatches(aa.ravel(), bb.ravel()) == 0
