"""The project-owned elementary functions (su2_fmath.h), host build.

The heat-bath's float functions (logf, cospif, sincospif) are within one
float ulp of the exact value, the accuracy class of the library functions
the reference calls: checked on every float of the heat-bath's argument
ranges against libm's double-precision values (exact argument reductions,
about 1e-16 relative), and on random samples against mpmath at 40 digits.
expf and sqrtf are correctly rounded except where the exact value lies
within 1e-11 relative of a rounding boundary; the double exp must be within
2 ulp, the double log within 2e-12 relative, the double sinpi/cospi within
1e-13 (the double exp only decides Metropolis tests inside the +-1e-4 band,
where a 2-ulp error changes a decision with probability ~1e-16), and the
double rsqrt within 1 ulp."""
import ctypes as C
import math
import os
import re
import struct
import sys

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


def check_ulp(got, exact, tag, bound=1.0):
    """|got - exact| <= bound ulp, the ulp of the float binade of the exact
    value: 2^(floor(log2 |exact|) - 23), at least 2^-149."""
    e = abs(mpmath.mpf(got) - exact)
    if exact == 0:
        u = mpmath.mpf(2) ** -149
    else:
        u = max(mpmath.mpf(2) ** (mpmath.floor(mpmath.log(abs(exact), 2)) - 23), mpmath.mpf(2) ** -149)
    assert e <= bound * u, (tag, got, float(exact), float(e / u))
    return float(e / u)


def test_logf():
    """logf on random floats of the heat-bath's range (0, 1] and beyond,
    and special points, against mpmath: within one ulp."""
    lib = _shim.load()
    xs = list(RNG.random(3000).astype(np.float32)) + [np.float32(v) for v in
                                                      (1.0, 0.5, 2.0 ** -32, 2.3283064e-10, 0.9999999, 2.0 ** -126, 3e38,
                                                       0.70710677, 1.4142135, 0.70710671, 1.4142134)]
    xs += list((RNG.random(1000) * 100).astype(np.float32))
    for x in xs:
        x = float(x)
        if x <= 0:
            continue
        check_ulp(lib.shim_logf(x), mpmath.log(mpmath.mpf(x)), ("logf", x))
    assert lib.shim_logf(1.0) == 0.0


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def sweep(which, lo, hi, planted=0):
    lib = _shim.load()
    w, n = C.c_uint32(), C.c_longlong()
    err = lib.shim_fmath_max_ulp(which, f32_bits(lo), f32_bits(hi) + 1, planted, C.byref(w), C.byref(n))
    return err, struct.unpack("<f", struct.pack("<I", w.value))[0], n.value


def test_logf_every_float_of_the_heatbath_range():
    """Every float in [2^-33, 2] (285 M; the heat-bath's draws lie in
    [2^-32, 1]): within one ulp, in fact below 0.6 ulp, and correctly
    rounded for all but a small fraction."""
    err, worst, not_cr = sweep(0, 2.0 ** -33, 2.0)
    assert err < 0.6, (err, worst)
    assert not_cr < 1e-3 * (33 * 2 ** 23), not_cr


@pytest.mark.parametrize("which,name", [(1, "cospif"), (2, "sin of sincospif"), (3, "cos of sincospif")])
def test_sinpi_cospi_every_float_of_the_heatbath_range(which, name):
    """Every float in [2^-40, 2] (the heat-bath calls them on 2u, u a draw
    in (0, 1]): within one ulp, in fact below 0.501 ulp."""
    err, worst, not_cr = sweep(which, 2.0 ** -40, 2.0)
    assert err < 0.501, (name, err, worst)
    assert not_cr < 1e-3 * (41 * 2 ** 23), (name, not_cr)


@pytest.mark.parametrize("planted,which", [(1, 0), (2, 1)])
def test_planted_mistake_fails_the_sweep(planted, which):
    """The sweeps reject plausible slips: the log tail's f^2 coefficient
    off by 1 %, and cospif's reduction written cos(pi q) = sin(pi (1 - q))."""
    err, _, _ = sweep(which, 0.25, 2.0, planted=planted)
    assert err > 1.0, err


def test_coefficients_come_from_the_generator():
    """The polynomial coefficients in su2_fmath.h are exactly those that
    gen_fmath_coeffs.py derives (Chebyshev interpolation, 60 digits)."""
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qcd_2d_gpu_amd", "csrc")
    sys.path.insert(0, csrc)
    try:
        import gen_fmath_coeffs as G
    finally:
        sys.path.remove(csrc)
    text = open(os.path.join(csrc, "su2_fmath.h")).read()
    for name, cs in G.coefficients().items():
        prefix = G.PREFIXES[name]
        found = dict(re.findall(rf"#define {prefix}(\d+) (\S+)", text))
        assert len(found) == len(cs), (name, found)
        for k, c in enumerate(cs):
            assert float.fromhex(found[str(k)].rstrip("f")) == c, (name, k, found[str(k)], c.hex())


def test_log_double():
    lib = _shim.load()
    for x in list(RNG.random(2000)) + [1e-310, 5e-324, 1.0, 2.0, 1e300, 0.7071067811865476]:
        got = lib.shim_log_d(x)
        ex = mpmath.log(mpmath.mpf(x))
        assert abs(mpmath.mpf(got) - ex) <= 2e-12 * abs(ex) + 2 * math.ulp(float(ex)), x


def test_cospif_and_sincospif():
    lib = _shim.load()
    xs = list((2 * RNG.random(3000)).astype(np.float32))
    xs += [np.float32(v) for v in (0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0, 1e-8, 1.9999999, 0.49999997,
                                   -0.3, -1.0, 3.5, 1000.25, 2.0 ** 24 + 2, 1e-30, 1e-45)]
    for x in xs:
        x = float(x)
        check_ulp(lib.shim_cospif(x), mpmath.cospi(mpmath.mpf(x)), ("cospif", x))
        s, c = C.c_float(), C.c_float()
        lib.shim_sincospif(x, C.byref(s), C.byref(c))
        check_ulp(s.value, mpmath.sinpi(mpmath.mpf(x)), ("sinpif", x))
        check_ulp(c.value, mpmath.cospi(mpmath.mpf(x)), ("cospif2", x))
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

def test_rsqrt_within_one_ulp():
    """The Newton rsqrt against mpmath: within one ulp over link and staple
    norms and the whole normal range; rsqrt(0) = inf as 1/sqrt(0)."""
    lib = _shim.load()
    xs = list(RNG.random(4000) * 16) + list(1 + 1e-3 * (RNG.random(4000) - 0.5)) + list(10 ** RNG.uniform(-300, 300, 2000))
    xs += [1.0, 2.0, 4.0, 0.5, 3.0, 1e-300, 1e300, 2.2250738585072014e-308, 1.7976931348623157e308]
    for x in xs:
        got = lib.shim_rsqrt(float(x))
        ex = 1 / mpmath.sqrt(mpmath.mpf(float(x)))
        assert abs(mpmath.mpf(got) - ex) <= mpmath.mpf(math.ulp(float(ex))), (x, got)
    assert math.isinf(lib.shim_rsqrt(0.0))
