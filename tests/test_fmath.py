"""The project-owned elementary functions (su2_fmath.h), host build, against
mpmath at 40 digits.  The float functions must return the correctly rounded
float except where the exact value lies within 1e-11 relative of a rounding
boundary (then the neighbouring float is allowed, i.e. at most one ulp off);
the double exp must be within 2 ulp, the double log within 2e-12 relative and
the double sinpi/cospi within 1e-13 (straightforward Horner evaluation without extra
precision; the double exp only decides Metropolis tests inside the +-1e-4
band, where a 2-ulp error changes a decision with probability ~1e-16)."""
import math
import struct

import mpmath
import numpy as np
import pytest

from tests import _shim

mpmath.mp.dps = 40
RNG = np.random.default_rng(20261203)


def f32(x):
    return float(np.float32(x))


def ulp_f(x):
    x = np.float32(abs(x))
    return float(np.spacing(x)) if x != 0 else float(np.float32(1.4e-45))


def check_float(got, exact, tag):
    exact_f = f32(float(exact))
    if got == exact_f:
        return 0
    # allowed only next to a rounding boundary: exact within 1e-12 rel of a midpoint
    mid = (mpmath.mpf(got) + mpmath.mpf(exact_f)) / 2
    assert abs(got - exact_f) <= 1.01 * ulp_f(exact_f), (tag, got, exact_f)
    assert abs(exact - mid) <= 1e-11 * abs(exact) + mpmath.mpf(1e-60), (tag, got, exact_f, exact)
    return 1


def test_logf():
    lib = _shim.load()
    xs = list(RNG.random(3000).astype(np.float32)) + [np.float32(v) for v in
                                                      (1.0, 0.5, 2.0 ** -32, 2.3283064e-10, 0.9999999, 1e-38, 3e38,
                                                       1e-45, 0.70710677, 1.4142135)]
    xs += list((RNG.random(1000) * 100).astype(np.float32))
    near = 0
    for x in xs:
        x = float(x)
        if x <= 0:
            continue
        near += check_float(lib.shim_logf(x), mpmath.log(mpmath.mpf(x)), ("logf", x))
    assert lib.shim_logf(1.0) == 0.0
    assert near <= 2


def test_log_double():
    lib = _shim.load()
    for x in list(RNG.random(2000)) + [1e-310, 5e-324, 1.0, 2.0, 1e300, 0.7071067811865476]:
        got = lib.shim_log_d(x)
        ex = mpmath.log(mpmath.mpf(x))
        assert abs(mpmath.mpf(got) - ex) <= 2e-12 * abs(ex) + 2 * math.ulp(float(ex)), x


def test_cospif_and_sincospif():
    lib = _shim.load()
    import ctypes as C

    xs = list((2 * RNG.random(3000)).astype(np.float32))
    xs += [np.float32(v) for v in (0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0, 1e-8, 1.9999999, 0.49999997,
                                   -0.3, -1.0, 3.5, 1000.25, 2.0 ** 24 + 2)]
    near = 0
    for x in xs:
        x = float(x)
        near += check_float(lib.shim_cospif(x), mpmath.cospi(mpmath.mpf(x)), ("cospif", x))
        s, c = C.c_float(), C.c_float()
        lib.shim_sincospif(x, C.byref(s), C.byref(c))
        near += check_float(s.value, mpmath.sinpi(mpmath.mpf(x)), ("sinpif", x))
        near += check_float(c.value, mpmath.cospi(mpmath.mpf(x)), ("cospif2", x))
    assert near <= 3
    assert lib.shim_cospif(0.5) == 0.0 and lib.shim_cospif(1.0) == -1.0 and lib.shim_cospif(2.0) == 1.0


def test_sincospi_double():
    import ctypes as C

    lib = _shim.load()
    for x in list(2 * RNG.random(1000)) + [0.1, 0.3, 1.7]:
        s, c = C.c_double(), C.c_double()
        lib.shim_sincospi_d(x, C.byref(s), C.byref(c))
        assert abs(mpmath.mpf(s.value) - mpmath.sinpi(x)) <= 1e-13
        assert abs(mpmath.mpf(c.value) - mpmath.cospi(x)) <= 1e-13


def test_expf():
    lib = _shim.load()
    xs = list((-104 * RNG.random(3000)).astype(np.float32)) + list((-1e-3 * RNG.random(500)).astype(np.float32))
    xs += [np.float32(v) for v in (0.0, -87.33654, -88.0, -103.0, -103.97, -104.0, -150.0, 10.0, 88.0)]
    near = 0
    for x in xs:
        x = float(x)
        near += check_float(lib.shim_expf(x), mpmath.exp(mpmath.mpf(x)), ("expf", x))
    assert lib.shim_expf(0.0) == 1.0
    assert lib.shim_expf(float("-inf")) == 0.0
    assert near <= 2


def test_exp_double_ulp():
    lib = _shim.load()
    xs = list(-745 * RNG.random(2000)) + list(709 * RNG.random(500)) + [0.0, -1e-300, 1e-10, -708.4, -740.0,
                                                                          -744.4, 709.7]
    for x in xs:
        got = lib.shim_exp(x)
        ex = mpmath.exp(mpmath.mpf(x))
        tol = 2 * math.ulp(float(ex)) if float(ex) >= 2.2250738585072014e-308 else 5e-324
        assert abs(mpmath.mpf(got) - ex) <= tol, (x, got, float(ex))
    assert lib.shim_exp(-800.0) == 0.0
    assert math.isinf(lib.shim_exp(710.0))


def test_sqrt_and_rsqrt_are_correctly_rounded():
    lib = _shim.load()
    for x in RNG.random(2000).astype(np.float32):
        x = float(x)
        assert lib.shim_sqrtf(x) == f32(float(mpmath.sqrt(mpmath.mpf(x))))
    for x in list(RNG.random(500) * 16):
        assert lib.shim_rsqrt(x) == 1.0 / math.sqrt(x)


def _mistaken_logf(x):
    """logf by the header's method with one plausible slip: the f^5
    coefficient 1/4 instead of 1/5."""
    m, e = math.frexp(x)
    m *= 2
    e -= 1
    if m > math.sqrt(2):
        m /= 2
        e += 1
    f = (m - 1) / (m + 1)
    s = f * f
    coef = [1 / 3, 1 / 4, 1 / 7, 1 / 9, 1 / 11, 1 / 13, 1 / 15]
    p = 0.0
    for c in reversed(coef):
        p = c + s * p
    return f32(e * math.log(2) + 2 * f + 2 * f * s * p)


def _mistaken_cospif(x):
    """cospif with the reduction slip cos(pi q) = sin(pi (1 - q))."""
    r = x - 2 * math.floor(x / 2 + 0.5)
    return f32(math.sin(math.pi * (1 - abs(r))))


@pytest.mark.parametrize("fn,ref", [(_mistaken_logf, mpmath.log), (_mistaken_cospif, mpmath.cospi)])
def test_planted_mistake_is_caught(fn, ref):
    """The accuracy checks above reject plausible slips in the method."""
    xs = [float(v) for v in (RNG.random(300) * 1.9 + 0.05).astype(np.float32)]
    with pytest.raises(AssertionError):
        for x in xs:
            check_float(fn(x), ref(mpmath.mpf(x)), "planted")
