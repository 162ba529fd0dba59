/*
 * su2_fmath.h - project-owned elementary functions for the SU(2) update
 * bodies, identical bit for bit on the x86-64 host (gcc, clang) and on
 * gfx950 (hipcc).
 *
 * The heat-bath and Metropolis updates of the reference evaluate logf,
 * cospif, sincospif, sqrtf, expf, exp and rsqrt
 * (functions/mc_updates.h:9,14,48-49,53,55-56,58; functions/su2.h:5,17).
 * A one-ulp difference in any of them can flip a Kennedy-Pendleton accept
 * test and shift every later draw of that site's random stream, so the
 * device kernels and the CPU checker must evaluate them identically.  Vendor
 * libraries (glibc libm, the device math library) do not promise that, so
 * this header builds each function from IEEE-754 basic operations only
 * (+ - * /, fused multiply-add and sqrt, all correctly rounded on both
 * targets) with a fixed evaluation order; every polynomial step is one fused
 * multiply-add.  Every file that includes it must be compiled with automatic
 * floating-point contraction disabled (-ffp-contract=off), so that the only
 * fused operations are the written ones; the pragma below enforces that for
 * clang/hipcc.
 *
 * Methods (all from the mathematics; no vendor library internals):
 *  - log: x = 2^e * m, m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(f),
 *    f = (m-1)/(m+1), |f| < 0.172, evaluated by its Taylor series through
 *    f^15 (truncation < 1.3e-12 relative); e*ln2 with a two-part (Cody-Waite)
 *    ln2 so that e*ln2_hi is exact.
 *  - exp: x = k ln2 + r, |r| <= ln2/2 (Cody-Waite), exp r by its Taylor
 *    series through r^13/13! (truncation < 5e-18), then scaled by 2^k with
 *    exact exponent arithmetic (two steps for subnormal results).
 *  - sinpi/cospi: exact reduction of the argument modulo 2 to r in [-1, 1),
 *    q = |r|; then sin(pi q) = sin(pi min(q, 1-q)) and
 *    cos(pi q) = sin(pi (1/2 - q)), so both come from one odd Taylor
 *    polynomial of sin t on |t| <= pi/2 (through t^17/17!, truncation
 *    < 5e-14) with no data-dependent branch: on the GPU every lane of a
 *    wavefront runs the same instructions.
 *  - float versions (logf, cospif, sincospif: the heat-bath's) evaluate in
 *    double arithmetic on the exactly converted argument, with polynomials
 *    just accurate enough for a float result, and round once to float:
 *      logf: x = 2^e m with m in [0x3f3504f3, 0x3fb504f2] (as floats, about
 *        [sqrt(2)/2, sqrt(2))), split off the float's bits with integer
 *        operations; log x = e ln2 + f (1 - f/2 + f^2 U(f)), f = m - 1 exact,
 *        U the degree-8 Chebyshev interpolant of (log1p(f)/f - 1 + f/2)/f^2
 *        with float coefficients, evaluated in float (each step one fused
 *        multiply-add with a 32-bit literal); the rest in double.  U's term
 *        is at most 7 % of log1p(f)/f, so its float rounding costs at most
 *        about 0.3 ulp of the result; no division;
 *      sinpi/cospi: the exact reduction above, then sin(pi d) = d R(d^2),
 *        R the degree-5 Chebyshev interpolant of sin(pi d)/d in d^2 on
 *        |d| <= 1/2 (relative error 2.7e-11).
 *    The coefficients come from gen_fmath_coeffs.py (Chebyshev
 *    interpolation in 60-digit arithmetic).  Every result is within one
 *    float ulp of the exact value (the accuracy class of the library
 *    functions the reference calls); the CPU tests check every float of the
 *    heat-bath's argument ranges.  Double arithmetic issues at the float
 *    rate on gfx950 (packed FP32 is not used), but a double constant that is
 *    not 0, 0.5, 1, 2 or 4 costs two register moves per use where a float
 *    constant is a free literal: hence float for the log's long tail.
 *  - sqrtf is the IEEE (correctly rounded) float square root on both
 *    targets.  rsqrt(x) = 1/sqrt(x) by four Newton steps
 *    y <- y + y (1/2 - (x/2) y^2) from the classic bit-level seed (the
 *    exponent halved and negated by an integer subtraction, relative error
 *    < 3.5 %); quadratic convergence leaves the rounding of the last step,
 *    a few ulp: three operations per step instead of a correctly rounded
 *    square root and a division.
 *
 * The header is valid C99 (for the CPU checker) and HIP C++.
 */
#ifndef SU2_FMATH_H
#define SU2_FMATH_H

#include <stdint.h>

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

#if defined(__HIP__) || defined(__HIPCC__)
#define SU2_FM_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define SU2_FM_FN static inline
#endif

/* Cody-Waite split of ln 2: LN2_HI has 32 significant bits, so k*LN2_HI is
 * exact for |k| < 2^21. */
#define SU2_LN2_HI 0x1.62e42fee00000p-1
#define SU2_LN2_LO 0x1.a39ef35793c76p-33
#define SU2_INV_LN2 0x1.71547652b82fep+0
#define SU2_PI 0x1.921fb54442d18p+1
#define SU2_SQRT2 0x1.6a09e667f3bcdp+0

SU2_FM_FN uint64_t su2_fm_bits(double x)
{
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}

SU2_FM_FN double su2_fm_from_bits(uint64_t u)
{
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

/* 2^k for -1022 <= k <= 1023, exact. */
SU2_FM_FN double su2_fm_pow2(int k)
{
    return su2_fm_from_bits((uint64_t)(k + 1023) << 52);
}

SU2_FM_FN double su2_sqrt_rn(double x) { return __builtin_sqrt(x); }

/* a * b + c with one rounding */
SU2_FM_FN double su2_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
SU2_FM_FN float su2_fmaf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

/* rsqrt of the reference (functions/su2.h:5,17; mc_updates.h:58) for x > 0
 * finite and normal (link norms and staple norms): four Newton steps from
 * the bit-level seed; rsqrt(0) = +inf as 1/sqrt(0), so a zero staple still
 * gives the NaN link it gives with exact arithmetic. */
SU2_FM_FN double su2_rsqrt(double x)
{
    double y = su2_fm_from_bits(0x5fe6eb50c7b537a9ULL - (su2_fm_bits(x) >> 1));
    const double h = 0.5 * x;
    /* y (3/2 - h y^2) as y + y (1/2 - h y^2) */
    y = su2_fma(y, su2_fma(-h * y, y, 0.5), y);
    y = su2_fma(y, su2_fma(-h * y, y, 0.5), y);
    y = su2_fma(y, su2_fma(-h * y, y, 0.5), y);
    y = su2_fma(y, su2_fma(-h * y, y, 0.5), y);
    return x == 0.0 ? __builtin_inf() : y;
}

/* sqrtf (functions/mc_updates.h:56), correctly rounded. */
SU2_FM_FN float su2_sqrtf(float x) { return __builtin_sqrtf(x); }

SU2_FM_FN float su2_fmaxf(float a, float b)
{
    /* fmaxf semantics for the non-NaN operands that occur here */
    return (a > b) ? a : ((b > a) ? b : (a == 0.0f ? (a + b) : a));
}

/* Natural logarithm of a normal, finite double x > 0. */
SU2_FM_FN double su2_log_normal_d(double x)
{
    uint64_t u = su2_fm_bits(x);
    int e = (int)((u >> 52) & 0x7ff) - 1023;
    double m = su2_fm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > SU2_SQRT2) {
        m = m * 0.5;
        e = e + 1;
    }
    double f = (m - 1.0) / (m + 1.0);
    double s = f * f;
    double p = 1.0 / 15.0;
    p = su2_fma(s, p, 1.0 / 13.0);
    p = su2_fma(s, p, 1.0 / 11.0);
    p = su2_fma(s, p, 1.0 / 9.0);
    p = su2_fma(s, p, 1.0 / 7.0);
    p = su2_fma(s, p, 1.0 / 5.0);
    p = su2_fma(s, p, 1.0 / 3.0);
    double logm = su2_fma(2.0 * f, s * p, 2.0 * f);
    double de = (double)e;
    return su2_fma(de, SU2_LN2_HI, su2_fma(de, SU2_LN2_LO, logm));
}

/* Natural logarithm for finite x > 0 (subnormals scaled into range). */
SU2_FM_FN double su2_log_d(double x)
{
    if (su2_fm_bits(x) < 0x0010000000000000ULL) return su2_log_normal_d(x * 0x1p54) - 54.0 * SU2_LN2_HI - 54.0 * SU2_LN2_LO;
    return su2_log_normal_d(x);
}

SU2_FM_FN uint32_t su2_fm_fbits(float x)
{
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}

SU2_FM_FN float su2_fm_ffrom_bits(uint32_t u)
{
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

/* Chebyshev coefficients (floats) of U(f) = (log1p(f) / f - 1 + f / 2) / f^2
 * on the mantissa range below (gen_fmath_coeffs.py, degree 8). */
#define SU2_FM_LOGU_C0 0x1.5555540000000p-2f
#define SU2_FM_LOGU_C1 -0x1.fffffc0000000p-3f
#define SU2_FM_LOGU_C2 0x1.999d5a0000000p-3f
#define SU2_FM_LOGU_C3 -0x1.555b4a0000000p-3f
#define SU2_FM_LOGU_C4 0x1.23d21a0000000p-3f
#define SU2_FM_LOGU_C5 -0x1.fcf4c60000000p-4f
#define SU2_FM_LOGU_C6 0x1.dea2800000000p-4f
#define SU2_FM_LOGU_C7 -0x1.d635c00000000p-4f
#define SU2_FM_LOGU_C8 0x1.1d8eb00000000p-4f
#define SU2_LN2 0x1.62e42fefa39efp-1

/* logf of the reference (functions/mc_updates.h:49) for positive normal
 * floats (the heat-bath calls it on draws in [2^-32, 1]).  x = 2^e m with
 * m in [0x3f3504f3, 0x3fb504f2] by integer operations on the float's bits:
 * subtracting the bits of 0.70710677 leaves e (signed) above the mantissa
 * field, and adding them back to the mantissa field alone gives m. */
SU2_FM_FN float su2_logf(float x)
{
    const uint32_t ix = su2_fm_fbits(x) - 0x3f3504f3u;
    const int e = (int)(ix >> 23) - (int)((ix >> 31) << 9); /* arithmetic ix >> 23 */
    const float ff = su2_fm_ffrom_bits((ix & 0x007fffffu) + 0x3f3504f3u) - 1.0f; /* exact */
    float u = SU2_FM_LOGU_C8;
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C7);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C6);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C5);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C4);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C3);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C2);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C1);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C0);
    const double f = (double)ff;
    const double p = su2_fma(f, su2_fma(f, (double)u, -0.5), 1.0); /* log1p(f) / f */
    return (float)su2_fma((double)e, SU2_LN2, f * p);
}

/* Exponential in double. */
SU2_FM_FN double su2_exp_d(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.1332191019412) return 0.0;
    double kd = __builtin_floor(su2_fma(x, SU2_INV_LN2, 0.5));
    int k = (int)kd;
    double r = su2_fma(-kd, SU2_LN2_LO, su2_fma(-kd, SU2_LN2_HI, x));
    double p = 1.0 / 6227020800.0;              /* 1/13! */
    p = su2_fma(r, p, 1.0 / 479001600.0);       /* 1/12! */
    p = su2_fma(r, p, 1.0 / 39916800.0);        /* 1/11! */
    p = su2_fma(r, p, 1.0 / 3628800.0);         /* 1/10! */
    p = su2_fma(r, p, 1.0 / 362880.0);          /* 1/9! */
    p = su2_fma(r, p, 1.0 / 40320.0);           /* 1/8! */
    p = su2_fma(r, p, 1.0 / 5040.0);            /* 1/7! */
    p = su2_fma(r, p, 1.0 / 720.0);             /* 1/6! */
    p = su2_fma(r, p, 1.0 / 120.0);             /* 1/5! */
    p = su2_fma(r, p, 1.0 / 24.0);              /* 1/4! */
    p = su2_fma(r, p, 1.0 / 6.0);               /* 1/3! */
    p = su2_fma(r, p, 0.5);                     /* 1/2! */
    p = su2_fma(r, p, 1.0);
    p = su2_fma(r, p, 1.0);
    if (k > 1023) return (p * 2.0) * su2_fm_pow2(k - 1);
    if (k < -1021) return (p * su2_fm_pow2(k + 600)) * 0x1p-600;
    return p * su2_fm_pow2(k);
}

/* expf of the reference (functions/mc_updates.h:9). */
SU2_FM_FN float su2_expf(float x) { return (float)su2_exp_d((double)x); }

/* exp of the reference (functions/mc_updates.h:14). */
SU2_FM_FN double su2_exp(double x) { return su2_exp_d(x); }

/* sin(pi d) for |d| <= 1/2: t = pi d, sin t = t - t z S(z), z = t^2,
 * S(z) = 1/3! - z/5! + ... + z^7/17!; truncation at t^19/19! < 5e-14. */
SU2_FM_FN double su2_fm_sinpi_k(double d)
{
    double t = d * SU2_PI;
    double z = t * t;
    double p = 1.0 / 355687428096000.0;         /*  1/17! */
    p = su2_fma(z, p, -1.0 / 1307674368000.0);  /* -1/15! */
    p = su2_fma(z, p, 1.0 / 6227020800.0);      /*  1/13! */
    p = su2_fma(z, p, -1.0 / 39916800.0);       /* -1/11! */
    p = su2_fma(z, p, 1.0 / 362880.0);          /*  1/9!  */
    p = su2_fma(z, p, -1.0 / 5040.0);           /* -1/7!  */
    p = su2_fma(z, p, 1.0 / 120.0);             /*  1/5!  */
    p = su2_fma(z, p, -1.0 / 6.0);              /* -1/3!  */
    return su2_fma(t * z, p, t);
}

/* Exact reduction: r = a - 2 round(a/2) in [-1, 1); returns q = |r|. */
SU2_FM_FN double su2_fm_reduce2(double a, int *neg)
{
    double r = a - 2.0 * __builtin_floor(a * 0.5 + 0.5);
    *neg = r < 0.0;
    return *neg ? -r : r;
}

/* cos(pi a) = sin(pi (1/2 - q)) in double for finite a. */
SU2_FM_FN double su2_cospi_d(double a)
{
    int neg;
    double q = su2_fm_reduce2(a, &neg);
    return su2_fm_sinpi_k(0.5 - q);
}

/* sin(pi a) and cos(pi a) in double for finite a. */
SU2_FM_FN void su2_sincospi_d(double a, double *sp, double *cp)
{
    int neg;
    double q = su2_fm_reduce2(a, &neg);
    double s = su2_fm_sinpi_k(q < 1.0 - q ? q : 1.0 - q);
    *sp = neg ? -s : s;
    *cp = su2_fm_sinpi_k(0.5 - q);
}

/* Chebyshev coefficients of sin(pi d) / d as a polynomial in d^2, |d| <= 1/2
 * (gen_fmath_coeffs.py, degree 5). */
#define SU2_FM_SIN_C0 0x1.921fb54425516p+1
#define SU2_FM_SIN_C1 -0x1.4abbce51c31dap+2
#define SU2_FM_SIN_C2 0x1.466bc06900900p+1
#define SU2_FM_SIN_C3 -0x1.32d12ebe6b560p-1
#define SU2_FM_SIN_C4 0x1.5014072bd9ca2p-4
#define SU2_FM_SIN_C5 -0x1.cc9b5f3456b49p-8

/* sin(pi d) for |d| <= 1/2, accurate enough for a float result. */
SU2_FM_FN double su2_fm_sinpi_f(double d)
{
    const double w = d * d;
    double p = SU2_FM_SIN_C5;
    p = su2_fma(w, p, SU2_FM_SIN_C4);
    p = su2_fma(w, p, SU2_FM_SIN_C3);
    p = su2_fma(w, p, SU2_FM_SIN_C2);
    p = su2_fma(w, p, SU2_FM_SIN_C1);
    p = su2_fma(w, p, SU2_FM_SIN_C0);
    return d * p;
}

/* cospif of the reference (functions/mc_updates.h:48) for finite x:
 * cos(pi q) = sin(pi (1/2 - q)), q = |x - 2 round(x/2)|, all exact. */
SU2_FM_FN float su2_cospif(float x)
{
    int neg;
    const double q = su2_fm_reduce2((double)x, &neg);
    return (float)su2_fm_sinpi_f(0.5 - q);
}

/* sincospif of the reference (functions/mc_updates.h:55) for finite x. */
SU2_FM_FN void su2_sincospif(float x, float *sp, float *cp)
{
    int neg;
    const double q = su2_fm_reduce2((double)x, &neg);
    const double s = su2_fm_sinpi_f(q < 1.0 - q ? q : 1.0 - q);
    *sp = (float)(neg ? -s : s);
    *cp = (float)su2_fm_sinpi_f(0.5 - q);
}

#endif /* SU2_FMATH_H */
