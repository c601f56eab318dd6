"""The shared-denominator quotient (recip/qdiv in storm_amd/csrc/fvt_math.h)
against the plain IEEE division on the device, bit for bit: random operands
over the whole exponent range, operands whose quotient is near overflow or
underflow (where v_div_scale scales the denominator and qdiv must fall back),
subnormals, zeros, infinities and NaNs.  Also the device's pow_one_third
(the implicit friction's h**(1/3)): correctly rounded against 60-digit
arithmetic over the depths the step meets, and equal to the Python
restatement of its arithmetic in tests/test_pow_one_third.py."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqdiv_probe.so")


def _operands():
    rng = np.random.default_rng(2024)
    n = 1 << 20
    # random signs, mantissas and exponents over the normal range
    def rnd(m, emin=-1020, emax=1020):
        e = rng.integers(emin, emax, size=m)
        return np.ldexp(rng.uniform(1.0, 2.0, size=m), e) * rng.choice([-1.0, 1.0], size=m)
    den = rnd(n)
    num = np.stack([rnd(n), rnd(n), rnd(n)], axis=1)
    # the physical range of the FVT step (lengths, depths, determinants)
    k = n // 4
    den[:k] = rnd(k, -40, 40)
    num[:k] = np.stack([rnd(k, -40, 40) for _ in range(3)], axis=1)
    # quotients near overflow / underflow: large exponent gaps in one row
    m = 4096
    den[k:k + m] = rnd(m, -1022, -1000)
    num[k:k + m, 1] = rnd(m, 1000, 1023)
    den[k + m:k + 2 * m] = rnd(m, 1000, 1023)
    num[k + m:k + 2 * m, 2] = rnd(m, -1022, -1000)
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.2e-308,
                         1.7e308, 1.0, -1.0, 3.0], dtype=np.float64)
    s = specials.shape[0]
    j = k + 2 * m
    for a in range(s):
        for b in range(s):
            den[j] = specials[a]
            num[j] = [specials[b], specials[(b + 1) % s], 1.0]
            j += 1
    return num, den


def test_qdiv_is_bitwise_the_plain_division():
    lib = C.CDLL(LIB)
    lib.qdiv_probe.restype = C.c_int
    num, den = _operands()
    num = np.ascontiguousarray(num)
    n = den.shape[0]
    out = np.zeros((n, 3))
    ref = np.zeros((n, 3))
    dp = C.POINTER(C.c_double)
    rc = lib.qdiv_probe(n, num.ctypes.data_as(dp), den.ctypes.data_as(dp),
                        out.ctypes.data_as(dp), ref.ctypes.data_as(dp))
    assert rc == 0
    a, b = out.view(np.uint64), ref.view(np.uint64)
    same = (a == b) | (np.isnan(out) & np.isnan(ref))
    assert same.all(), "%d of %d quotients differ" % ((~same).sum(), same.size)
    # and the plain device division is IEEE's (the host's, as numpy divides)
    with np.errstate(all="ignore"):
        host = num / den[:, None]
    hb = host.view(np.uint64)
    agree = (b == hb) | (np.isnan(ref) & np.isnan(host))
    assert agree.all(), "%d device quotients differ from the host's" % (~agree).sum()


def test_device_pow_one_third_is_correctly_rounded():
    import math
    from decimal import Decimal, getcontext
    from test_pow_one_third import pow_one_third as restated
    lib = C.CDLL(LIB)
    lib.pow13_probe.restype = C.c_int
    rng = np.random.default_rng(9)
    h = np.exp(rng.uniform(math.log(1e-4), math.log(200.0), 4000))
    h = np.concatenate([h, [1e-30, 1e30, 1.0, 8.0, 27.0, 0.001, 1e-3 * (1 + 2 ** -52)]])
    out = np.zeros_like(h)
    dp = C.POINTER(C.c_double)
    assert lib.pow13_probe(h.shape[0], h.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    getcontext().prec = 60
    third = Decimal(1.0 / 3.0)
    for x, y in zip(h, out):
        want = float((Decimal(float(x)).ln() * third).exp())
        assert y == want, (x, y, want)
        assert y == restated(float(x)), x
