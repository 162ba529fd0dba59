/*
 * su2_shim.c - host build of the product's numerics headers (the same
 * headers the gfx950 kernels include), exported for the CPU tests: the
 * elementary functions, the MRG32k3a generator and the single-link update
 * bodies.  Test infrastructure only.
 */
#include <math.h>
#include <string.h>

#include "su2_fmath.h"
#include "su2_mrg32k3a.h"
#include "su2_quat.h"
#include "su2_update.h"

float shim_logf(float x) { return su2_logf(x); }
double shim_log_d(double x) { return su2_log_d(x); }
float shim_cospif(float x) { return su2_cospif(x); }
void shim_sincospif(float x, float *s, float *c) { su2_sincospif(x, s, c); }
void shim_sincospi_d(double x, double *s, double *c) { su2_sincospi_d(x, s, c); }
double shim_cospi_d(double x) { return su2_cospi_d(x); }
float shim_expf(float x) { return su2_expf(x); }
double shim_exp(double x) { return su2_exp(x); }
float shim_sqrtf(float x) { return su2_sqrtf(x); }
double shim_rsqrt(double x) { return su2_rsqrt(x); }

void shim_mrg_seed(uint64_t seed, uint32_t *s) { su2_mrg_seed(seed, s); }
void shim_mrg_jump(uint32_t *s, uint32_t sub) { su2_mrg_jump(s, sub, su2_mrg_jump1_host, su2_mrg_jump2_host); }
uint32_t shim_mrg_next(uint32_t *s) { return su2_mrg_next(s); }
float shim_uniform(uint32_t *s) { return su2_mrg_uniform(s); }
double shim_uniform_double(uint32_t *s) { return su2_mrg_uniform_double(s); }
const uint32_t *shim_jump_table(int comp) { return comp == 1 ? su2_mrg_jump1_host : su2_mrg_jump2_host; }

static su2_q ld(const double *a) { return su2_q_make(a[0], a[1], a[2], a[3]); }
static void st(double *a, su2_q q)
{
    a[0] = q.a0;
    a[1] = q.a1;
    a[2] = q.a2;
    a[3] = q.a3;
}

void shim_heat_bath(const double *A, double beta, uint32_t *s, double *out) { st(out, su2_heat_bath(ld(A), beta, s)); }
int shim_metropolis(const double *A, double *U, double beta, double eps, int hits, uint32_t *s)
{
    su2_q u = ld(U);
    int acc = su2_metropolis(ld(A), &u, beta, eps, hits, s);
    st(U, u);
    return acc;
}
int shim_metropolis_accept(double s, double u) { return su2_metropolis_accept(s, u); }
void shim_over_relax(const double *A, const double *U, double *out) { st(out, su2_over_relax(ld(A), ld(U))); }
void shim_rand_link(uint32_t *s, double *out) { st(out, su2_rand_link(s)); }
void shim_haar_link(uint32_t *s, double *out) { st(out, su2_haar_link(s)); }
void shim_q_mul(const double *A, const double *B, double *C, int kind)
{
    su2_q a = ld(A), b = ld(B), c;
    switch (kind) {
    case 0: c = su2_q_mul(a, b); break;
    case 1: c = su2_q_mul_nd(a, b); break;
    case 2: c = su2_q_mul_dn(a, b); break;
    default: c = su2_q_mul_dd(a, b); break;
    }
    st(C, c);
}

/* float paths of the FP32-link staple; kind 0: A B, 1: A B^dag, 3: A^dag B^dag */
void shim_qf_mul(const float *A, const float *B, float *C, int kind)
{
    su2_qf a = su2_qf_make(A[0], A[1], A[2], A[3]), b = su2_qf_make(B[0], B[1], B[2], B[3]), c;
    switch (kind) {
    case 0: c = su2_qf_mul(a, b); break;
    case 1: c = su2_qf_mul_nd(a, b); break;
    default: c = su2_qf_mul_dd(a, b); break;
    }
    C[0] = c.a0;
    C[1] = c.a1;
    C[2] = c.a2;
    C[3] = c.a3;
}

/* kind 0: det A, 1: 1/2 Re tr(A B), 2: first component of A B^dag */
double shim_q_scalar(const double *A, const double *B, int kind)
{
    su2_q a = ld(A), b = ld(B);
    switch (kind) {
    case 0: return su2_q_det(a);
    case 1: return su2_q_retr(a, b);
    default: return su2_q_mul_nd_re(a, b);
    }
}

/* ---- exhaustive accuracy sweeps of the float elementary functions ----
 *
 * Error of su2_logf (which = 0), su2_cospif (1), and the sine (2) and
 * cosine (3) of su2_sincospif over every float whose bit pattern lies in
 * [lo, hi), in ulps of the exact result: |got - ref| / ulp(ref), ulp(y) =
 * 2^(floor(log2 |y|) - 23) (subnormal spacing below 2^-126).  ref is libm's
 * double-precision value, about 1e-16 relative; the trigonometric ones
 * after an exact reduction, sin(pi d) with |d| <= 1/2, so that ref keeps its
 * relative accuracy next to the zeros.  Returns the largest error; *worst
 * = the bit pattern where it occurs; *not_cr = the number of results that
 * differ from the correctly rounded float of ref.  `planted` (tests only)
 * evaluates a deliberately wrong variant instead of the product function:
 * 1 = su2_logf with its U polynomial's f^2 coefficient off by 1 %, 2 =
 * cospif with the reduction slip cos(pi q) = sin(pi (1 - q)). */
#define SHIM_PI 3.14159265358979323846

static double ulp_of(double y)
{
    y = fabs(y);
    if (y < 0x1p-126) return 0x1p-149;
    int e;
    frexp(y, &e); /* y = m 2^e, m in [0.5, 1) */
    return ldexp(1.0, e - 1 - 23);
}

static float planted_logf(float x)
{
    const uint32_t ix = su2_fm_fbits(x) - 0x3f3504f3u;
    const int e = (int)(ix >> 23) - (int)((ix >> 31) << 9);
    const float ff = su2_fm_ffrom_bits((ix & 0x007fffffu) + 0x3f3504f3u) - 1.0f;
    float u = SU2_FM_LOGU_C8;
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C7);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C6);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C5);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C4);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C3);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C2 * 1.01f);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C1);
    u = su2_fmaf(ff, u, SU2_FM_LOGU_C0);
    const double f = (double)ff;
    const double p = su2_fma(f, su2_fma(f, (double)u, -0.5), 1.0);
    return (float)su2_fma((double)e, SU2_LN2, f * p);
}

static float planted_cospif(float x)
{
    int neg;
    const double q = su2_fm_reduce2((double)x, &neg);
    return (float)su2_fm_sinpi_f(1.0 - q);
}

static double ref_value(int which, float x)
{
    if (which == 0) return log((double)x);
    const double a = (double)x;
    const double r = a - 2.0 * nearbyint(0.5 * a); /* exact, in [-1, 1] */
    const double q = fabs(r);
    if (which == 2) {
        const double d = q < 1.0 - q ? q : 1.0 - q;
        const double s = sin(SHIM_PI * d);
        return r < 0 ? -s : s;
    }
    return sin(SHIM_PI * (0.5 - q));
}

double shim_fmath_max_ulp(int which, uint32_t lo, uint32_t hi, int planted, uint32_t *worst, long long *not_cr)
{
    double best = 0.0;
    uint32_t wb = lo;
    long long ncr = 0;
#pragma omp parallel
    {
        double lbest = 0.0;
        uint32_t lwb = lo;
        long long lncr = 0;
#pragma omp for schedule(static)
        for (long long b = lo; b < (long long)hi; ++b) {
            const float x = su2_fm_ffrom_bits((uint32_t)b);
            float got;
            if (which == 0) {
                got = planted == 1 ? planted_logf(x) : su2_logf(x);
            } else if (which == 1) {
                got = planted == 2 ? planted_cospif(x) : su2_cospif(x);
            } else {
                float s, c;
                su2_sincospif(x, &s, &c);
                got = which == 2 ? s : c;
            }
            const double ref = ref_value(which, x);
            const double err = fabs((double)got - ref) / ulp_of(ref);
            if (got != (float)ref) ++lncr;
            if (err > lbest || err != err) {
                lbest = err != err ? 1e300 : err;
                lwb = (uint32_t)b;
            }
        }
#pragma omp critical
        {
            if (lbest > best) {
                best = lbest;
                wb = lwb;
            }
            ncr += lncr;
        }
    }
    if (worst) *worst = wb;
    if (not_cr) *not_cr = ncr;
    return best;
}
