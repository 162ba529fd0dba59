"""Coefficients of the polynomials in su2_fmath.h, by Chebyshev interpolation.

Method (textbook numerical analysis; e.g. Trefethen, "Approximation Theory
and Approximation Practice", ch. 4 and 8): interpolate g on [lo, hi] at the
n + 1 Chebyshev points of the first kind, convert the interpolant to the
monomial basis in 60-digit arithmetic (mpmath), round each coefficient to
the nearest double.  The interpolant's error is within a factor
1 + (2/pi) log(n + 1) of the best (minimax) polynomial of that degree.

Functions (what the float elementary functions of the heat-bath need; each
result is rounded to float once):
  U(f) = (log1p(f) / f - 1 + f / 2) / f^2
                 on f in [0x3f3504f3 as float - 1, 0x3fb504f2 as float - 1]
                 (the mantissa range of su2_logf's argument reduction), so
                 that log1p(f) = f (1 - f / 2 + f^2 U(f)); degree 8, the
                 coefficients rounded to floats (su2_logf evaluates U in
                 float arithmetic, the rest in double);
  sin(pi d) / d  as a polynomial in w = d^2, d in [-1/2, 1/2]; degree 5,
                 double coefficients.

Run as a script to print the coefficient lines of su2_fmath.h and the
relative error of each polynomial (its coefficients as doubles, evaluated
exactly) against the function on a dense grid.
"""
from __future__ import annotations

import struct
import sys

import mpmath

mpmath.mp.dps = 60


def f32_bits(b: int) -> float:
    return struct.unpack("<f", struct.pack("<I", b))[0]


LOG_LO = f32_bits(0x3F3504F3) - 1.0  # m = 0.70710677 ...
LOG_HI = f32_bits(0x3FB504F2) - 1.0  # ... up to the float below 2 * 0.70710677
LOG_DEGREE = 8  # of U
SIN_DEGREE = 5  # in w = d^2: sin(pi d) through d^11


def log_g(f):
    """log1p(f) / f"""
    f = mpmath.mpf(f)
    return mpmath.log1p(f) / f if f != 0 else mpmath.mpf(1)


def log_tail_g(f):
    """U(f) = (log1p(f) / f - 1 + f / 2) / f^2 (its limit 1/3 at f = 0)."""
    f = mpmath.mpf(f)
    if abs(f) < mpmath.mpf("1e-15"):
        return mpmath.mpf(1) / 3 - f / 4
    return (log_g(f) - 1 + f / 2) / f ** 2


def sin_g(w):
    w = mpmath.mpf(w)
    if w == 0:
        return mpmath.pi
    d = mpmath.sqrt(w)
    return mpmath.sin(mpmath.pi * d) / d


def to_single(c):
    return struct.unpack("<f", struct.pack("<f", c))[0]


def cheb_monomial(g, lo, hi, n, single=False):
    """Monomial coefficients c[0..n] (c[k] multiplies x^k) of the degree-n
    Chebyshev interpolant of g on [lo, hi], rounded to doubles (to floats
    with single=True)."""
    lo, hi = mpmath.mpf(lo), mpmath.mpf(hi)
    c, r = (lo + hi) / 2, (hi - lo) / 2
    xs = [c + r * mpmath.cos(mpmath.pi * (k + mpmath.mpf(1) / 2) / (n + 1)) for k in range(n + 1)]
    A = mpmath.matrix([[x ** j for j in range(n + 1)] for x in xs])
    b = mpmath.matrix([g(x) for x in xs])
    sol = mpmath.lu_solve(A, b)
    out = [float(sol[j]) for j in range(n + 1)]
    if single:
        out = [to_single(c) for c in out]
    return out


def max_rel_error(g, coeffs, lo, hi, pts=4000):
    worst = mpmath.mpf(0)
    for i in range(pts + 1):
        x = mpmath.mpf(lo) + (mpmath.mpf(hi) - mpmath.mpf(lo)) * i / pts
        p = mpmath.mpf(0)
        for c in reversed(coeffs):
            p = p * x + mpmath.mpf(c)
        worst = max(worst, abs(p / g(x) - 1))
    return worst


# macro prefixes of the coefficients in su2_fmath.h
PREFIXES = {"log_tail": "SU2_FM_LOGU_C", "sinpi_over_d": "SU2_FM_SIN_C"}


def coefficients():
    """{name: list of floats}: the coefficients su2_fmath.h must hold."""
    return {
        "log_tail": cheb_monomial(log_tail_g, LOG_LO, LOG_HI, LOG_DEGREE, single=True),
        "sinpi_over_d": cheb_monomial(sin_g, 0.0, 0.25, SIN_DEGREE),
    }


def log_error_in_P(coeffs, pts=4000):
    """Largest |f^2 (U_poly(f) - U(f))| / (log1p(f) / f): the polynomial's
    share of the relative error of log1p(f)."""
    worst = mpmath.mpf(0)
    for i in range(pts + 1):
        f = mpmath.mpf(LOG_LO) + (mpmath.mpf(LOG_HI) - mpmath.mpf(LOG_LO)) * i / pts
        p = mpmath.mpf(0)
        for c in reversed(coeffs):
            p = p * f + mpmath.mpf(c)
        worst = max(worst, abs((p - log_tail_g(f)) * f ** 2 / log_g(f)))
    return worst


def main():
    co = coefficients()
    for name, cs in co.items():
        print(f"/* {name}: degree {len(cs) - 1} */")
        for k, c in enumerate(cs):
            single = name == "log_tail"
            print(f"#define {PREFIXES[name]}{k} {c.hex()}{'f' if single else ''}")
    if "--errors" in sys.argv:
        print("log1p: tail polynomial's relative error share", mpmath.nstr(log_error_in_P(co["log_tail"]), 5))
        print("sin(pi d)/d max rel error", mpmath.nstr(max_rel_error(sin_g, co["sinpi_over_d"], 0.0, 0.25), 5))


if __name__ == "__main__":
    main()
