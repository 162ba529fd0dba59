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
 *  - float versions evaluate the double function on the exactly converted
 *    argument and round once to float; with the truncations above they are
 *    within one float ulp (correctly rounded unless the exact value lies
 *    within ~1e-12 relative of a rounding boundary), the accuracy class of
 *    the library functions the reference calls.
 *  - sqrtf = (float)sqrt((double)x) is the correctly rounded float square
 *    root (53 >= 2*24+2, so the double rounding is innocuous); rsqrt(x) is
 *    1.0/sqrt(x), two correctly rounded operations.
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

/* rsqrt of the reference (functions/su2.h:5,17; mc_updates.h:58) as two
 * correctly rounded operations. */
SU2_FM_FN double su2_rsqrt(double x) { return 1.0 / __builtin_sqrt(x); }

/* sqrtf (functions/mc_updates.h:56), correctly rounded. */
SU2_FM_FN float su2_sqrtf(float x) { return (float)__builtin_sqrt((double)x); }

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

/* logf of the reference (functions/mc_updates.h:49) for finite x > 0.
 * Every positive float, subnormals included, is a normal double, so no
 * range scaling is needed; the heat-bath calls it on draws in (0, 1]. */
SU2_FM_FN float su2_logf(float x) { return (float)su2_log_normal_d((double)x); }

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

/* cospif of the reference (functions/mc_updates.h:48). */
SU2_FM_FN float su2_cospif(float x) { return (float)su2_cospi_d((double)x); }

/* sincospif of the reference (functions/mc_updates.h:55). */
SU2_FM_FN void su2_sincospif(float x, float *sp, float *cp)
{
    double s, c;
    su2_sincospi_d((double)x, &s, &c);
    *sp = (float)s;
    *cp = (float)c;
}

#endif /* SU2_FMATH_H */
