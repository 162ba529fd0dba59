/*
 * su2_shim.c - host build of the product's numerics headers (the same
 * headers the gfx950 kernels include), exported for the CPU tests: the
 * elementary functions, the MRG32k3a generator and the single-link update
 * bodies.  Test infrastructure only.
 */
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
