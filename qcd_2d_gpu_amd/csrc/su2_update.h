/*
 * su2_update.h - the single-link update bodies: Kennedy-Pendleton heat-bath,
 * Metropolis with the single-then-double accept test, over-relaxation.
 *
 * Each takes the staple sum A of the link being updated (the sum of the
 * four "U_nu U_mu^dag U_nu^dag" paths around it) and the link's MRG32k3a
 * state s[6]; the local weight is exp(beta * 1/2 Re tr(U A)).  The random
 * draws happen in exactly the order of the reference, because the stream of
 * a half-site is shared by the even and odd site and by all six launches of
 * a sweep: any change of order or count would desynchronise it.
 */
#ifndef SU2_UPDATE_H
#define SU2_UPDATE_H

#include <float.h>

#include "su2_fmath.h"
#include "su2_mrg32k3a.h"
#include "su2_quat.h"

#if defined(__HIP__) || defined(__HIPCC__)
#define SU2_U_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define SU2_U_FN static inline
#endif

/*
 * Every product that feeds a sum below is fused (one fma, written out), as
 * in the quaternion algebra: the contraction a CUDA compiler applies to the
 * reference's expressions by default, fixed here so that the host and the
 * device round alike.
 *
 * Kennedy-Pendleton heat-bath (functions/mc_updates.h:43-64).
 * A is normalised to A^ with norm a; X0 = 1 - 2 lambda is drawn from
 * sqrt(1 - X0^2) exp(beta a X0) by the rejection loop, evaluated in float as
 * the reference does; the direction (cos theta, phi) is drawn in float and
 * renormalised in double; the new link is X A^^dagger.
 * Split in three so a sweep kernel can run the trials of sites rejected at
 * the first trial elsewhere (su2_sweep_site.h): su2_kp_setup normalises A
 * and returns the float scale 0.5 / (a beta), su2_kp_trial is one pass of
 * the loop, su2_kp_finish draws the direction.
 */
SU2_U_FN float su2_kp_setup(su2_q *A, double beta)
{
    double a;
    *A = su2_q_normalize_det(*A, &a);
    return (float)(0.5 / (a * beta));
}

/* One trial of the rejection loop: four draws; returns whether it ends the
 * loop, with lambda in *lf.  A zero or non-finite staple makes the scale inv
 * infinite or NaN, and every trial would be rejected for ever (the
 * reference hangs there); the loop then ends after its first trial, with a
 * NaN link, instead of hanging the device.  For every finite scale this is
 * the reference's test. */
SU2_U_FN int su2_kp_trial(uint32_t *s, float inv, float *lf)
{
    float r0 = su2_mrg_uniform(s);
    float r1 = su2_mrg_uniform(s);
    float r2 = su2_mrg_uniform(s);
    float c = su2_cospif(2.0f * r1);
    float l = -su2_fmaf(c * c, su2_logf(r2), su2_logf(r0)) * inv;
    float aux = 1.0f - su2_mrg_uniform(s);
    *lf = l;
    return !(aux * aux > 1.0f - l) || !(inv <= FLT_MAX);
}

/* The new link from an accepted lambda: two more draws for the direction. */
SU2_U_FN su2_q su2_kp_finish(su2_q A, float lf, uint32_t *s)
{
    double x0 = 1.0 - 2.0 * (double)lf;
    double n = su2_sqrt_rn(su2_fma(-x0, x0, 1.0));
    float ct = 2.0f * su2_mrg_uniform(s) - 1.0f;
    float sp, cp;
    su2_sincospif(2.0f * su2_mrg_uniform(s), &sp, &cp);
    float st = su2_sqrtf(su2_fmaxf(0.0f, su2_fmaf(-ct, ct, 1.0f)));
    double v1 = ct, v2 = (double)st * cp, v3 = (double)st * sp;
    double sc = n * su2_rsqrt(su2_fma(v3, v3, su2_fma(v2, v2, v1 * v1)));
    su2_q X = su2_q_make(x0, v1 * sc, v2 * sc, v3 * sc);
    return su2_q_mul_nd(X, A);
}

SU2_U_FN su2_q su2_heat_bath(su2_q A, double beta, uint32_t *s)
{
    float inv = su2_kp_setup(&A, beta);
    float lf;
    while (!su2_kp_trial(s, inv, &lf)) {
    }
    return su2_kp_finish(A, lf, s);
}

/*
 * Metropolis decision exp(-s) > u (functions/mc_updates.h:7-15): a float
 * estimate decides outside a +-1e-4 relative band, the double exponential
 * inside it.
 */
SU2_U_FN int su2_metropolis_accept(double s, double u)
{
    const double tol = 1e-4;
    double ef = (double)su2_expf(-(float)s);
    if (u < ef * (1.0 - tol)) return 1;
    if (u > ef * (1.0 + tol)) return 0;
    return su2_exp(-s) > u;
}

/*
 * `hits` Metropolis hits with a fixed staple (functions/mc_updates.h:16-36,
 * 119-128; proposal functions/su2.h:87-97).  The proposal is R U with
 * R = (sqrt(1 - eps^2), eps r^) and r^ the normalised vector of three
 * (u - 0.5) draws.  A hit with dS < 0 is accepted without a fourth draw.
 * Returns the number of accepted hits.
 */
SU2_U_FN int su2_metropolis(su2_q A, su2_q *U, double beta, double eps, int hits, uint32_t *s)
{
    int acc = 0;
    for (int i = 0; i < hits; ++i) {
        double r1 = su2_mrg_uniform_double(s) - 0.5;
        double r2 = su2_mrg_uniform_double(s) - 0.5;
        double r3 = su2_mrg_uniform_double(s) - 0.5;
        double n = su2_sqrt_rn(su2_fma(r3, r3, su2_fma(r2, r2, r1 * r1)));
        r1 = eps * r1 / n;
        r2 = eps * r2 / n;
        r3 = eps * r3 / n;
        double r0 = su2_sqrt_rn(su2_fma(-eps, eps, 1.0));
        su2_q Unew = su2_q_mul(su2_q_make(r0, r1, r2, r3), *U);
        su2_q dU = su2_q_sub(Unew, *U);
        double dS = -beta * su2_q_retr(dU, A);
        if (dS < 0.0) {
            *U = Unew;
            ++acc;
            continue;
        }
        if (su2_metropolis_accept(dS, su2_mrg_uniform_double(s))) {
            *U = Unew;
            ++acc;
        }
    }
    return acc;
}

/*
 * Over-relaxation (functions/mc_updates.h:109-115): U <- A^^dag U^dag A^^dag,
 * which keeps 1/2 Re tr(U A) and draws nothing.
 */
SU2_U_FN su2_q su2_over_relax(su2_q A, su2_q U)
{
    double a;
    A = su2_q_normalize_det(A, &a);
    (void)a;
    su2_q B = su2_q_mul_dd(U, A);
    return su2_q_mul_dn(A, B);
}

/* One hot-start link (functions/su2.h:29-34): four (u - 0.5) double draws,
 * normalised. */
SU2_U_FN su2_q su2_rand_link(uint32_t *s)
{
    double a0 = su2_mrg_uniform_double(s) - 0.5;
    double a1 = su2_mrg_uniform_double(s) - 0.5;
    double a2 = su2_mrg_uniform_double(s) - 0.5;
    double a3 = su2_mrg_uniform_double(s) - 0.5;
    return su2_q_normalize(su2_q_make(a0, a1, a2, a3));
}

/* One Haar-distributed link (extension, no counterpart in the reference):
 * four standard normals from two Box-Muller pairs (four double draws),
 * normalised; the uniform measure on S^3 is the Haar measure of SU(2). */
SU2_U_FN su2_q su2_haar_link(uint32_t *s)
{
    double u1 = su2_mrg_uniform_double(s);
    double u2 = su2_mrg_uniform_double(s);
    double u3 = su2_mrg_uniform_double(s);
    double u4 = su2_mrg_uniform_double(s);
    double r1 = su2_sqrt_rn(-2.0 * su2_log_d(u1));
    double r2 = su2_sqrt_rn(-2.0 * su2_log_d(u3));
    double s1, c1, s2, c2;
    su2_sincospi_d(2.0 * u2, &s1, &c1);
    su2_sincospi_d(2.0 * u4, &s2, &c2);
    su2_q g = su2_q_make(r1 * c1, r1 * s1, r2 * c2, r2 * s2);
    double det = su2_q_det(g);
    if (det == 0.0) return su2_q_id();
    return su2_q_normalize(g);
}

#endif /* SU2_UPDATE_H */
