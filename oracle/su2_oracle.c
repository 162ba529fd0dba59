/*
 * su2_oracle.c - CPU restatement of fezquerro/QCD_2d_GPU's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path of
 * libsu2mc is compared against; it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and by nothing
 * in the product.
 *
 * It restates, in plain C + OpenMP and in the reference's own
 * array-of-structures layout and index arithmetic, so that each line can be
 * set beside the reference:
 *   - the SU(2) algebra of functions/su2.h:1-97;
 *   - heat_bath / metropolis / metropolis_accept / over_relaxation /
 *     link_update of functions/mc_updates.h:7-128;
 *   - the MC_update_{even,odd}_{y,x,t} kernels and mc_up_per
 *     (functions/mc_updates.h:129-475), their PCC variants
 *     (functions/mc_updates_pcc.h:4-224) and PMC variants
 *     (functions/mc_updates_pmc.h:4-257), one OpenMP loop per launch;
 *   - setup_kernel, rand_init2, id_init2, rand_init2_pcc, normalize
 *     (functions/lattice_setup.h:15-81,202-207);
 *   - plaquette_yx/yt/xt and en_av<PMC> (functions/observables.h:1-21,
 *     99-141), with the block tree sum of the reference and the block sums
 *     added in block order.
 * The random generator is restated independently here from the published
 * MRG32k3a recurrence and rocRAND's (seed, subsequence, offset) rules (the
 * build's RNG contract); the jump-ahead matrices are computed at run time
 * by exact modular matrix squaring.  The one piece shared with the product
 * is su2_fmath.h (project-owned log/exp/sinpi/cospi built from IEEE basic
 * operations), which is pinned separately against high-precision values.
 *
 * Rounding: every product that feeds a sum is fused, in one fixed pattern:
 * the first product of a sum rounded on its own, each later term added by
 * fma()/fmaf() in the reference's term order (the contraction a CUDA
 * compiler applies to these expressions by default).  No other contraction
 * happens (-ffp-contract=off); -mfma makes the fused operations single
 * instructions, and without it the C library's correctly rounded fma()
 * gives the same values.
 *
 * FP32-link mode (an extension of the build): every store into the field is
 * rounded to float, and the three-link paths of each staple are multiplied
 * in float (path_up/path_dn); the staple sum, the updates, reunitarisation
 * and the observables stay in double.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../qcd_2d_gpu_amd/csrc/su2_fmath.h"

/* struct lattice (functions/lattice_setup.h:2-14); passed as the half
 * lattice `l2` (y and V halved, kernel.cu:59) unless noted. */
typedef struct orc_lattice {
    int y, x, t, V, d, dg, dt;
    float beta;
    float epsilon;
    int updates;
} orc_lattice;

/* ------------------------------------------------------------------------
 * MRG32k3a
 * ---------------------------------------------------------------------- */
typedef struct orc_rng {
    uint32_t g1[3];
    uint32_t g2[3];
} orc_rng;

#define M1 4294967087ull
#define M2 4294944443ull

static uint64_t red1(uint64_t p) /* one fold of 2^32 = m1 + 209, one subtraction */
{
    p = 209ull * (p >> 32) + (p & 0xffffffffull);
    if (p >= M1) p -= M1;
    return p;
}

static uint64_t red2(uint64_t p) /* two folds of 2^32 = m2 + 22853, one subtraction */
{
    p = 22853ull * (p >> 32) + (p & 0xffffffffull);
    p = 22853ull * (p >> 32) + (p & 0xffffffffull);
    if (p >= M2) p -= M2;
    return p;
}

/* x1[n] = 1403580 x1[n-2] - 810728 x1[n-3] (mod m1),
 * x2[n] = 527612 x2[n-1] - 1370589 x2[n-3] (mod m2), output x1 - x2 in
 * [1, m1]; the subtractions are done as additions of (m - x). */
uint32_t orc_mrg_next(orc_rng *r)
{
    uint32_t p1 = (uint32_t)red1(1403580ull * r->g1[1] + 810728ull * (uint32_t)((uint32_t)M1 - r->g1[0]));
    r->g1[0] = r->g1[1];
    r->g1[1] = r->g1[2];
    r->g1[2] = p1;
    uint32_t p2 = (uint32_t)red2(527612ull * r->g2[2] + 1370589ull * (uint32_t)((uint32_t)M2 - r->g2[0]));
    r->g2[0] = r->g2[1];
    r->g2[1] = r->g2[2];
    r->g2[2] = p2;
    return p1 > p2 ? p1 - p2 : p1 - p2 + (uint32_t)M1;
}

float orc_uniform(orc_rng *r) /* curand_uniform: (0, 1] float */
{
    return (float)((double)orc_mrg_next(r) * (1.0 / 4294967087.0));
}

double orc_uniform_double(orc_rng *r) /* curand_uniform_double: (0, 1] double */
{
    return (double)orc_mrg_next(r) * (1.0 / 4294967087.0);
}

static uint64_t seedmul(uint32_t i, uint64_t j, uint64_t (*red)(uint64_t))
{
    /* i * j split at 2^17, products wrapping modulo 2^64 */
    uint64_t hi = i >> 17, lo = i & 131071u;
    return red(red(hi * j) * 131072ull + red(lo * j));
}

typedef uint64_t mat3[9];

static void matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, uint64_t m)
{
    uint64_t t[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            uint64_t s = 0;
            for (int j = 0; j < 3; ++j) s += (uint64_t)(((unsigned __int128)a[3 * i + j] * b[3 * j + k]) % m);
            t[3 * i + k] = s % m;
        }
    memcpy(c, t, sizeof t);
}

static void matvec(const uint64_t *a, uint32_t *s, uint64_t m)
{
    uint64_t v[3];
    for (int i = 0; i < 3; ++i) {
        uint64_t acc = 0;
        for (int j = 0; j < 3; ++j) acc += (uint64_t)(((unsigned __int128)a[3 * i + j] * s[j]) % m);
        v[i] = acc % m;
    }
    for (int i = 0; i < 3; ++i) s[i] = (uint32_t)v[i];
}

/* A^(2^(76+i)) and A^(2^i), i < 64, for both components */
static mat3 J76_1[64], J76_2[64], J0_1[64], J0_2[64];
static int jumps_ready = 0;

static void build_jumps(void)
{
    if (jumps_ready) return;
    /* built once, before any parallel use, by orc_setup_kernel or under the
     * critical section of orc_mrg_init */
    mat3 a1 = {0, 1, 0, 0, 0, 1, M1 - 810728ull, 1403580ull, 0};
    mat3 a2 = {0, 1, 0, 0, 0, 1, M2 - 1370589ull, 0, 527612ull};
    mat3 p1, p2;
    memcpy(p1, a1, sizeof p1);
    memcpy(p2, a2, sizeof p2);
    for (int i = 0; i < 140; ++i) {
        if (i < 64) {
            memcpy(J0_1[i], p1, sizeof p1);
            memcpy(J0_2[i], p2, sizeof p2);
        }
        if (i >= 76) {
            memcpy(J76_1[i - 76], p1, sizeof p1);
            memcpy(J76_2[i - 76], p2, sizeof p2);
        }
        matmul(p1, p1, p1, M1);
        matmul(p2, p2, p2, M2);
    }
    jumps_ready = 1;
}

/* curand_init(seed, subsequence, offset) with rocRAND's semantics. */
void orc_mrg_init(uint64_t seed, uint64_t subsequence, uint64_t offset, orc_rng *r)
{
    if (!jumps_ready) {
#pragma omp critical(orc_jumps)
        build_jumps();
    }
    if (seed == 0) seed = 12345ull;
    uint32_t x = (uint32_t)seed ^ 0x55555555u;
    uint32_t y = (uint32_t)(seed >> 32) ^ 0xAAAAAAAAu;
    r->g1[0] = (uint32_t)seedmul(x, seed, red1);
    r->g1[1] = (uint32_t)seedmul(y, seed, red1);
    r->g1[2] = (uint32_t)seedmul(x, seed, red1);
    r->g2[0] = (uint32_t)seedmul(y, seed, red2);
    r->g2[1] = (uint32_t)seedmul(x, seed, red2);
    r->g2[2] = (uint32_t)seedmul(y, seed, red2);
    for (int i = 0; subsequence; ++i, subsequence >>= 1)
        if (subsequence & 1) {
            matvec(J76_1[i], r->g1, M1);
            matvec(J76_2[i], r->g2, M2);
        }
    for (int i = 0; offset; ++i, offset >>= 1)
        if (offset & 1) {
            matvec(J0_1[i], r->g1, M1);
            matvec(J0_2[i], r->g2, M2);
        }
}

/* setup_kernel (functions/lattice_setup.h:15-21) for all V/2 streams */
void orc_setup_kernel(orc_rng *state, int nstates, uint64_t seed)
{
    build_jumps();
#pragma omp parallel for schedule(static)
    for (int id = 0; id < nstates; ++id) orc_mrg_init(seed, (uint64_t)id, 0, state + id);
}

/* ------------------------------------------------------------------------
 * SU(2) algebra, functions/su2.h
 * ---------------------------------------------------------------------- */
double su2_det(const double *A) /* su2.h:1-3 */
{
    return fma(A[3], A[3], fma(A[2], A[2], fma(A[1], A[1], A[0] * A[0])));
}

void su2_normalize(double *A) /* su2.h:4-8 */
{
    double inv = su2_rsqrt(su2_det(A));
    for (int i = 0; i < 4; i++) A[i] = A[i] * inv;
}

double su2_normalize_det(double *A) /* su2.h:16-21 */
{
    double det = su2_det(A), inv = su2_rsqrt(det);
    for (int i = 0; i < 4; i++) A[i] = A[i] * inv;
    return det * inv;
}

void su2_rand(double *A, orc_rng *state) /* su2.h:29-34 */
{
    for (int i = 0; i < 4; i++) A[i] = orc_uniform_double(state) - 0.5;
    su2_normalize(A);
}

void su2_id(double *A) /* su2.h:35-39 */
{
    A[0] = 1;
    for (int i = 1; i < 4; i++) A[i] = 0;
}

void su2_zero(double *A)
{
    for (int i = 0; i < 4; i++) A[i] = 0;
}

void su2_product(const double *A, const double *B, double *C) /* su2.h:50-55: C = A B */
{
    double c0 = fma(-A[3], B[3], fma(-A[2], B[2], fma(-A[1], B[1], A[0] * B[0])));
    double c1 = fma(-A[3], B[2], fma(A[2], B[3], fma(A[0], B[1], A[1] * B[0])));
    double c2 = fma(-A[1], B[3], fma(A[0], B[2], fma(A[3], B[1], A[2] * B[0])));
    double c3 = fma(A[1], B[2], fma(A[0], B[3], fma(-A[2], B[1], A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

void su2_product_tc(const double *A, const double *B, double *C) /* su2.h:56-61: C = A B^dag */
{
    double c0 = fma(A[3], B[3], fma(A[2], B[2], fma(A[1], B[1], A[0] * B[0])));
    double c1 = fma(A[3], B[2], fma(-A[2], B[3], fma(-A[0], B[1], A[1] * B[0])));
    double c2 = fma(A[1], B[3], fma(-A[0], B[2], fma(-A[3], B[1], A[2] * B[0])));
    double c3 = fma(-A[1], B[2], fma(-A[0], B[3], fma(A[2], B[1], A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

void su2_product_tc1(const double *A, const double *B, double *C) /* su2.h:62-67: C = A^dag B */
{
    double c0 = fma(A[3], B[3], fma(A[2], B[2], fma(A[1], B[1], A[0] * B[0])));
    double c1 = fma(A[3], B[2], fma(-A[2], B[3], fma(A[0], B[1], -A[1] * B[0])));
    double c2 = fma(A[1], B[3], fma(A[0], B[2], fma(-A[3], B[1], -A[2] * B[0])));
    double c3 = fma(-A[1], B[2], fma(A[0], B[3], fma(A[2], B[1], -A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

void su2_product_tc2(const double *A, const double *B, double *C) /* su2.h:68-73: C = A^dag B^dag */
{
    double c0 = fma(-A[3], B[3], fma(-A[2], B[2], fma(-A[1], B[1], A[0] * B[0])));
    double c1 = fma(-A[3], B[2], fma(A[2], B[3], fma(-A[0], B[1], -A[1] * B[0])));
    double c2 = fma(-A[1], B[3], fma(-A[0], B[2], fma(A[3], B[1], -A[2] * B[0])));
    double c3 = fma(A[1], B[2], fma(-A[0], B[3], fma(-A[2], B[1], -A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

double su2_prod_trace(const double *A, const double *B) /* su2.h:74-76 */
{
    return fma(-A[3], B[3], fma(-A[2], B[2], fma(-A[1], B[1], A[0] * B[0])));
}

void su2_sum(double *A, const double *B) /* su2.h:77-80 */
{
    for (int i = 0; i < 4; i++) A[i] = A[i] + B[i];
}

void su2_sub(const double *A, const double *B, double *C) /* su2.h:81-84 */
{
    for (int i = 0; i < 4; i++) C[i] = A[i] - B[i];
}

void su2_update(const double *A, double *B, double ep, orc_rng *state) /* su2.h:87-97 */
{
    double r[4];
    for (int i = 1; i < 4; i++) r[i] = orc_uniform_double(state) - 0.5;
    double n = sqrt(fma(r[3], r[3], fma(r[2], r[2], r[1] * r[1])));
    for (int i = 0; i < 3; i++) r[i + 1] = ep * r[i + 1] / n;
    r[0] = sqrt(fma(-ep, ep, 1.0));
    su2_product(r, A, B);
}

/* ------------------------------------------------------------------------
 * Link updates, functions/mc_updates.h:7-128
 * ---------------------------------------------------------------------- */
int orc_metropolis_accept(double s, double u) /* mc_updates.h:7-15 */
{
    const double tol = 1e-4;
    double ef = su2_expf(-(float)s);
    if (u < ef * (1 - tol)) return 1;
    if (u > ef * (1 + tol)) return 0;
    return su2_exp(-s) > u;
}

int orc_metropolis(const double *A, double *phi, double beta, double epsilon, orc_rng *state) /* mc_updates.h:16-36 */
{
    double Unew[4], deltaU[4];
    su2_update(phi, Unew, epsilon, state);
    su2_sub(Unew, phi, deltaU);
    double s = -beta * su2_prod_trace(deltaU, A);
    if (s < 0) {
        memcpy(phi, Unew, sizeof Unew);
        return 1;
    }
    if (orc_metropolis_accept(s, orc_uniform_double(state))) {
        memcpy(phi, Unew, sizeof Unew);
        return 1;
    }
    return 0;
}

int orc_heat_bath(double *A, double *phi, double beta, orc_rng *state) /* mc_updates.h:43-64 */
{
    double a = su2_normalize_det(A), X[4];
    float inv = (float)(0.5 / (a * beta)), lf, aux;
    do {
        float r0 = orc_uniform(state), r1 = orc_uniform(state), r2 = orc_uniform(state);
        float c = su2_cospif(2 * r1);
        lf = -fmaf(c * c, su2_logf(r2), su2_logf(r0)) * inv;
        aux = 1 - orc_uniform(state);
        /* a zero or non-finite staple (inv infinite or NaN) would be rejected
         * for ever, as in the reference; the loop ends after one trial */
    } while (aux * aux > 1 - lf && inv <= FLT_MAX);
    X[0] = 1 - 2 * (double)lf;
    double n = sqrt(fma(-X[0], X[0], 1.0));
    float ct = 2 * orc_uniform(state) - 1, sp, cp;
    su2_sincospif(2 * orc_uniform(state), &sp, &cp);
    float st = su2_sqrtf(su2_fmaxf(0.f, fmaf(-ct, ct, 1.f)));
    double v1 = ct, v2 = (double)st * cp, v3 = (double)st * sp;
    double s = n * su2_rsqrt(fma(v3, v3, fma(v2, v2, v1 * v1)));
    X[1] = v1 * s;
    X[2] = v2 * s;
    X[3] = v3 * s;
    su2_product_tc(X, A, phi);
    return 1;
}

int orc_over_relaxation(double *A, double *phi) /* mc_updates.h:109-115 */
{
    double B[4];
    (void)su2_normalize_det(A);
    su2_product_tc2(phi, A, B);
    su2_product_tc1(A, B, phi);
    return 1;
}

/* One hot-start link drawn from the Haar measure (extension; no
 * counterpart in the reference): normalised vector of four standard normals
 * from two Box-Muller pairs. */
static void su2_haar(double *A, orc_rng *state)
{
    double u1 = orc_uniform_double(state), u2 = orc_uniform_double(state);
    double u3 = orc_uniform_double(state), u4 = orc_uniform_double(state);
    double r1 = sqrt(-2.0 * su2_log_d(u1)), r2 = sqrt(-2.0 * su2_log_d(u3));
    double s1, c1, s2, c2;
    su2_sincospi_d(2.0 * u2, &s1, &c1);
    su2_sincospi_d(2.0 * u4, &s2, &c2);
    A[0] = r1 * c1;
    A[1] = r1 * s1;
    A[2] = r2 * c2;
    A[3] = r2 * s2;
    if (su2_det(A) == 0.0)
        su2_id(A);
    else
        su2_normalize(A);
}

/* Store rounding of the FP32-link extension. */
static void round_link(double *a, int f32)
{
    if (f32)
        for (int i = 0; i < 4; ++i) a[i] = (double)(float)a[i];
}

/* link_update (mc_updates.h:119-128) + the `update` switch of the kernels
 * (mc_updates.h:178-181).  algo: 0 Metropolis, 1 heat-bath. */
static int link_update(double *A, double *phi, const orc_lattice *l, orc_rng *state, int algo, int update, int f32)
{
    int acc = 0;
    if (update == 1) {
        acc = orc_over_relaxation(A, phi);
    } else if (algo == 1) {
        acc = orc_heat_bath(A, phi, l->beta, state);
    } else {
        double u[4];
        memcpy(u, phi, sizeof u);
        for (int i = 0; i < l->updates; i++) acc += orc_metropolis(A, u, l->beta, l->epsilon, state);
        memcpy(phi, u, sizeof u);
    }
    round_link(phi, f32);
    return acc;
}

/* ------------------------------------------------------------------------
 * The six kernels per boundary condition.
 *
 * For an even launch the thread index id is the even half-site: idr =
 * id * dt is the site, pos = (id + V) * dt the odd half at the same index
 * (mc_updates.h:130-136).  For an odd launch the thread index is pos: posr =
 * pos * dt is the even half (the other parity), id = (pos + V) * dt the odd
 * site (mc_updates.h:186-192).  Below, `self` and `other` stand for
 * (idr, pos) and (id, posr) respectively, so one staple routine per
 * direction restates both kernels of that direction.
 * ---------------------------------------------------------------------- */
typedef struct offs {
    int yp, ym, xp, xm, tp, tm;
} offs;

/* Neighbour offsets (mc_updates.h:139-153 even, 195-210 odd; the y-minus
 * ones at 252-254 even, 309-311 odd). */
static offs make_offs(const orc_lattice *l, int parity, int pos_y, int pos_x, int pos_t, int par_aux)
{
    offs o;
    if (parity == 0) {
        o.yp = par_aux * l->dt;
        if (pos_y == (l->y - 1) && par_aux == 1) o.yp = (-l->y + 1) * l->dt;
        o.ym = (par_aux - 1) * l->dt;
        if (pos_y == 0 && par_aux == 0) o.ym = (l->y - 1) * l->dt;
    } else {
        o.yp = (1 - par_aux) * l->dt;
        if (pos_y == (l->y - 1) && par_aux == 0) o.yp = (-l->y + 1) * l->dt;
        o.ym = -par_aux * l->dt;
        if (pos_y == 0 && par_aux == 1) o.ym = (l->y - 1) * l->dt;
    }
    o.xp = l->y * l->dt;
    if (pos_x == (l->x - 1)) o.xp = -l->y * (l->x - 1) * l->dt;
    o.xm = -l->y * l->dt;
    if (pos_x == 0) o.xm = l->y * (l->x - 1) * l->dt;
    o.tp = l->y * l->x * l->dt;
    if (pos_t == (l->t - 1)) o.tp = -l->y * l->x * (l->t - 1) * l->dt;
    o.tm = -l->y * l->x * l->dt;
    if (pos_t == 0) o.tm = l->y * l->x * (l->t - 1) * l->dt;
    return o;
}

/* FP32-link extension: the three-link paths of a staple are multiplied in
 * float (fmaf in the fused pattern of su2_product and its variants) and each
 * path is added to the double staple sum.  The field holds floats in
 * doubles, so the narrowing below is exact. */
static void qf_load(const double *a, float *f)
{
    for (int i = 0; i < 4; i++) f[i] = (float)a[i];
}

void su2_productf(const float *A, const float *B, float *C) /* C = A B */
{
    float c0 = fmaf(-A[3], B[3], fmaf(-A[2], B[2], fmaf(-A[1], B[1], A[0] * B[0])));
    float c1 = fmaf(-A[3], B[2], fmaf(A[2], B[3], fmaf(A[0], B[1], A[1] * B[0])));
    float c2 = fmaf(-A[1], B[3], fmaf(A[0], B[2], fmaf(A[3], B[1], A[2] * B[0])));
    float c3 = fmaf(A[1], B[2], fmaf(A[0], B[3], fmaf(-A[2], B[1], A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

void su2_product_tcf(const float *A, const float *B, float *C) /* C = A B^dag */
{
    float c0 = fmaf(A[3], B[3], fmaf(A[2], B[2], fmaf(A[1], B[1], A[0] * B[0])));
    float c1 = fmaf(A[3], B[2], fmaf(-A[2], B[3], fmaf(-A[0], B[1], A[1] * B[0])));
    float c2 = fmaf(A[1], B[3], fmaf(-A[0], B[2], fmaf(-A[3], B[1], A[2] * B[0])));
    float c3 = fmaf(-A[1], B[2], fmaf(-A[0], B[3], fmaf(A[2], B[1], A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

void su2_product_tc2f(const float *A, const float *B, float *C) /* C = A^dag B^dag */
{
    float c0 = fmaf(-A[3], B[3], fmaf(-A[2], B[2], fmaf(-A[1], B[1], A[0] * B[0])));
    float c1 = fmaf(-A[3], B[2], fmaf(A[2], B[3], fmaf(-A[0], B[1], -A[1] * B[0])));
    float c2 = fmaf(-A[1], B[3], fmaf(-A[0], B[2], fmaf(A[3], B[1], -A[2] * B[0])));
    float c3 = fmaf(A[1], B[2], fmaf(-A[0], B[3], fmaf(-A[2], B[1], -A[3] * B[0])));
    C[0] = c0;
    C[1] = c1;
    C[2] = c2;
    C[3] = c3;
}

/* A += [B C^dag] D^dag, the "positive" path of a staple */
static void path_up(const double *B, const double *C, const double *D, int f32, double *A)
{
    if (f32) {
        float b[4], c[4], d[4], aux[4], aux2[4];
        qf_load(B, b);
        qf_load(C, c);
        qf_load(D, d);
        su2_product_tcf(b, c, aux);
        su2_product_tcf(aux, d, aux2);
        for (int i = 0; i < 4; i++) A[i] = A[i] + (double)aux2[i];
    } else {
        double aux[4], aux2[4];
        su2_product_tc(B, C, aux);
        su2_product_tc(aux, D, aux2);
        su2_sum(A, aux2);
    }
}

/* A += [B^dag C^dag] D, the "negative" path of a staple */
static void path_dn(const double *B, const double *C, const double *D, int f32, double *A)
{
    if (f32) {
        float b[4], c[4], d[4], aux[4], aux2[4];
        qf_load(B, b);
        qf_load(C, c);
        qf_load(D, d);
        su2_product_tc2f(b, c, aux);
        su2_productf(aux, d, aux2);
        for (int i = 0; i < 4; i++) A[i] = A[i] + (double)aux2[i];
    } else {
        double aux[4], aux2[4];
        su2_product_tc2(B, C, aux);
        su2_product(aux, D, aux2);
        su2_sum(A, aux2);
    }
}

/* y links: x and t plaquettes (mc_updates.h:155-176, 212-232) */
static void staple_y(const double *phi, int self, int other, const offs *o, int dg, int f32, double *A)
{
    su2_zero(A);
    path_up(phi + other + o->yp + dg, phi + other + o->xp, phi + self + dg, f32, A);
    path_dn(phi + self + o->yp + dg + o->xm, phi + other + o->xm, phi + other + o->xm + dg, f32, A);
    path_up(phi + other + o->yp + 2 * dg, phi + other + o->tp, phi + self + 2 * dg, f32, A);
    path_dn(phi + self + o->yp + 2 * dg + o->tm, phi + other + o->tm, phi + other + o->tm + 2 * dg, f32, A);
}

/* x links: y and t plaquettes (mc_updates.h:265-285, 323-343; PMC skips as
 * mc_updates_pmc.h:33-45, 96-108) */
static void staple_x(const double *phi, int self, int other, const offs *o, int dg, int skip_yup, int skip_ydn,
                     int f32, double *A)
{
    su2_zero(A);
    if (!skip_yup) path_up(phi + other + o->xp, phi + other + o->yp + dg, phi + self, f32, A);
    if (!skip_ydn) path_dn(phi + self + o->xp + o->ym, phi + other + o->ym + dg, phi + other + o->ym, f32, A);
    path_up(phi + other + o->xp + 2 * dg, phi + other + o->tp + dg, phi + self + 2 * dg, f32, A);
    path_dn(phi + self + o->xp + 2 * dg + o->tm, phi + other + o->tm + dg, phi + other + o->tm + 2 * dg, f32, A);
}

/* t links: y and x plaquettes (mc_updates.h:377-397, 437-457; PMC skips as
 * mc_updates_pmc.h:156-167, 219-230) */
static void staple_t(const double *phi, int self, int other, const offs *o, int dg, int skip_yup, int skip_ydn,
                     int f32, double *A)
{
    su2_zero(A);
    if (!skip_yup) path_up(phi + other + o->tp, phi + other + o->yp + 2 * dg, phi + self, f32, A);
    if (!skip_ydn) path_dn(phi + self + o->tp + o->ym, phi + other + o->ym + 2 * dg, phi + other + o->ym, f32, A);
    path_up(phi + other + o->tp + dg, phi + other + o->xp + 2 * dg, phi + self + dg, f32, A);
    path_dn(phi + self + o->tp + dg + o->xm, phi + other + o->xm + 2 * dg, phi + other + o->xm + dg, f32, A);
}

/* One (direction mu, parity) launch: MC_update_{even,odd}_{y,x,t} with the
 * boundary variant bc (0 periodic, 1 PCC, 2 PMC).  l is the half lattice. */
void orc_mc_update(double *phi, const orc_lattice *l, int *accept, orc_rng *state, int bc, int algo, int update,
                   int mu, int parity, int f32)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < l->V; ++k) {
        int pos_y = k % l->y;
        int pos_x = (k % (l->x * l->y)) / (l->y);
        int pos_t = k / (l->x * l->y);
        int par_aux = (pos_t + pos_x) % 2;
        /* even: id = k, self = idr, other = pos; odd: pos = k, self = id, other = posr */
        int self = parity == 0 ? k * l->dt : (k + l->V) * l->dt;
        int other = parity == 0 ? (k + l->V) * l->dt : k * l->dt;
        /* the site's y-coordinate is 0 / Ly-1 at these (pos_y, par_aux) */
        int at_y0 = pos_y == 0 && par_aux == (parity == 0 ? 0 : 1);
        int at_ylast = pos_y == l->y - 1 && par_aux == (parity == 0 ? 1 : 0);
        offs o = make_offs(l, parity, pos_y, pos_x, pos_t, par_aux);
        double A[4];
        if (mu != 0 && bc == 1 && at_y0) continue; /* mc_updates_pcc.h:29,83,134,189 */
        int skip_up = (bc == 2) && at_ylast, skip_dn = (bc == 2) && at_y0;
        if (mu == 0)
            staple_y(phi, self, other, &o, l->dg, f32, A);
        else if (mu == 1)
            staple_x(phi, self, other, &o, l->dg, skip_up, skip_dn, f32, A);
        else
            staple_t(phi, self, other, &o, l->dg, skip_up, skip_dn, f32, A);
        accept[self / l->dt] += link_update(A, phi + self + mu * l->dg, l, state + k, algo, update, f32);
    }
}

/* mc_up_per / mc_up_pcc / mc_up_pmc (mc_updates.h:467-475,
 * mc_updates_pcc.h:216-224, mc_updates_pmc.h:249-257) */
void orc_mc_up(double *phi, const orc_lattice *l, int *accept, orc_rng *state, int bc, int algo, int update, int f32)
{
    for (int mu = 0; mu < 3; ++mu)
        for (int parity = 0; parity < 2; ++parity)
            orc_mc_update(phi, l, accept, state, bc, algo, update, mu, parity, f32);
}

/* ------------------------------------------------------------------------
 * Starts and normalisation, functions/lattice_setup.h
 * start: 0 rand_init2, 1 id_init2, 2 rand_init2_pcc, 3 Haar, 4 Haar + PCC
 * ---------------------------------------------------------------------- */
void orc_init(double *phi, orc_rng *state, const orc_lattice *l, int start, int f32)
{
    const int D = l->d * l->dg;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < l->V; ++id) {
        int pos_y = id % l->y;
        int pos_x = (id % (l->x * l->y)) / l->y;
        int pos_t = id / (l->x * l->y);
        int par_aux = (pos_t + pos_x) % 2;
        double *e = phi + id * D, *od = phi + (id + l->V) * D;
        if (start == 1) { /* lattice_setup.h:38-51 */
            for (int m = 0; m < 3; ++m) {
                su2_id(e + m * l->dg);
                su2_id(od + m * l->dg);
            }
            continue;
        }
        int pcc = (start == 2 || start == 4), haar = (start >= 3);
        /* lattice_setup.h:24-37 and 52-81: even y, x, t then odd y, x, t */
        for (int p = 0; p < 2; ++p) {
            double *s = p == 0 ? e : od;
            int y0 = pos_y == 0 && par_aux == p;
            for (int m = 0; m < 3; ++m) {
                if (pcc && m != 0 && y0)
                    su2_id(s + m * l->dg);
                else if (haar)
                    su2_haar(s + m * l->dg, state + id);
                else
                    su2_rand(s + m * l->dg, state + id);
                round_link(s + m * l->dg, f32);
            }
        }
    }
}

void orc_normalize(double *phi, long long nlinks, int f32) /* lattice_setup.h:202-207 */
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nlinks; ++i) {
        su2_normalize(phi + 4 * i);
        round_link(phi + 4 * i, f32);
    }
}

/* ------------------------------------------------------------------------
 * Observables, functions/observables.h
 * ---------------------------------------------------------------------- */
static double plaquette_yx(const double *phi, int tp, int xp, int yp, int dg, int sub) /* observables.h:1-7 */
{
    double A[4], B[4];
    su2_product(phi, phi + yp + dg + sub, A);
    su2_product_tc(A, phi + xp + sub, B);
    su2_product_tc(B, phi + dg, A);
    (void)tp;
    return 1 - A[0];
}

static double plaquette_yt(const double *phi, int tp, int xp, int yp, int dg, int sub) /* observables.h:8-14 */
{
    double A[4], B[4];
    su2_product(phi, phi + yp + 2 * dg + sub, A);
    su2_product_tc(A, phi + tp + sub, B);
    su2_product_tc(B, phi + 2 * dg, A);
    (void)xp;
    return 1 - A[0];
}

static double plaquette_xt(const double *phi, int tp, int xp, int yp, int dg, int sub) /* observables.h:15-21 */
{
    double A[4], B[4];
    su2_product(phi + dg, phi + xp + 2 * dg + sub, A);
    su2_product_tc(A, phi + tp + dg + sub, B);
    su2_product_tc(B, phi + 2 * dg, A);
    (void)yp;
    return 1 - A[0];
}

/* Per-site terms of en_av (observables.h:102-128): e = -c, p = (3-a-b-c)/d. */
static void en_av_site(const double *phi, const orc_lattice *l, int pmc, int posf, double *e, double *p)
{
    int parity = posf / l->V;
    int pos = posf % l->V;
    int pos_y = pos % l->y;
    int pos_x = (pos % (l->x * l->y)) / l->y;
    int pos_t = pos / (l->x * l->y);
    int par_aux = (pos_t + pos_x + parity) % 2;
    int yp, xp, tp, aux = (1 - 2 * parity) * l->V * l->dt, posfr = posf * l->dt;
    yp = par_aux * l->dt;
    if (pos_y == (l->y - 1) && par_aux == 1) yp = (-l->y + 1) * l->dt;
    xp = l->y * l->dt;
    if (pos_x == (l->x - 1)) xp = -l->y * (l->x - 1) * l->dt;
    tp = l->y * l->x * l->dt;
    if (pos_t == (l->t - 1)) tp = -l->y * l->x * (l->t - 1) * l->dt;
    double a = plaquette_yx(phi + posfr, tp, xp, yp, l->dg, aux);
    double b = plaquette_yt(phi + posfr, tp, xp, yp, l->dg, aux);
    double c = plaquette_xt(phi + posfr, tp, xp, yp, l->dg, aux);
    if (pmc && pos_y == l->y - 1 && par_aux == 1) {
        a = 0;
        b = 0;
    }
    *e = -c;
    *p = (3 - a - b - c) / l->d;
}

/* en_av<PMC> (observables.h:99-141) with the reference's launch shape: nt =
 * largest block <= 256 dividing V/2 (kernel.cu:66-70), a shared-memory tree
 * per block, and block sums / V added to *res, *res2 in block order. */
void orc_en_av(const double *phi, const orc_lattice *l, int pmc, double *res, double *res2)
{
    int nt = 256;
    while (l->V % nt != 0) nt--;
    int nb = 2 * l->V / nt;
    double *bs = (double *)malloc(sizeof(double) * 2 * (size_t)nb);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b) {
        double sh[2][256];
        for (int i = 0; i < nt; ++i) en_av_site(phi, l, pmc, b * nt + i, &sh[0][i], &sh[1][i]);
        for (int s = 128; s > 0; s >>= 1)
            for (int i = 0; i < s; ++i)
                if (i + s < nt) {
                    sh[0][i] += sh[0][i + s];
                    sh[1][i] += sh[1][i + s];
                }
        bs[2 * b] = sh[0][0];
        bs[2 * b + 1] = sh[1][0];
    }
    for (int b = 0; b < nb; ++b) {
        *res += bs[2 * b] / (2 * l->V);
        *res2 += bs[2 * b + 1] / (2 * l->V);
    }
    free(bs);
}

/* ------------------------------------------------------------------------
 * Timed sweep loop for the CPU baseline: n sweeps of the given kind
 * (update == 0: link update with algo; `or_per` over-relaxation sweeps
 * after each).  Returns wall seconds.
 * ---------------------------------------------------------------------- */
double orc_time_sweeps(double *phi, const orc_lattice *l, int *accept, orc_rng *state, int bc, int algo, int or_per,
                       int nsweeps, int f32)
{
#ifdef _OPENMP
    double t0 = omp_get_wtime();
#endif
    for (int i = 0; i < nsweeps; ++i) {
        orc_mc_up(phi, l, accept, state, bc, algo, 0, f32);
        for (int k = 0; k < or_per; ++k) orc_mc_up(phi, l, accept, state, bc, algo, 1, f32);
    }
#ifdef _OPENMP
    return omp_get_wtime() - t0;
#else
    return 0.0;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
