/*
 * su2_quat.h - SU(2) link algebra in the quaternion ("sigma") basis.
 *
 * A link is four reals a0..a3 standing for the matrix
 *     [[a0 + i a1,  a2 + i a3], [-a2 + i a3,  a0 - i a1]]
 * (functions/su2.h:98-100).  The products below are the 2x2 matrix products
 * in that basis.  Their term order follows functions/su2.h:50-76 exactly,
 * and each four-term sum is contracted in one fixed pattern: the first
 * product rounded on its own, every later term added by a fused
 * multiply-add, t = fma(+-a_k, b_k, t) (the contraction a CUDA compiler
 * applies to these expressions by default; fma is a correctly rounded
 * IEEE-754 operation on the host and on gfx950 alike).  With automatic
 * contraction disabled everywhere else, that fixes every rounding, which is
 * what lets the device kernels agree bit for bit with the CPU checker.
 * Fusing cuts each product from 28 to 16 FP64 instructions.
 */
#ifndef SU2_QUAT_H
#define SU2_QUAT_H

#include "su2_fmath.h"

#if defined(__HIP__) || defined(__HIPCC__)
#define SU2_Q_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define SU2_Q_FN static inline
#endif

typedef struct su2_q {
    double a0, a1, a2, a3;
} su2_q;

/* fused multiply-add, one rounding */
#define SU2_FMA(a, b, c) __builtin_fma((a), (b), (c))

SU2_Q_FN su2_q su2_q_make(double a0, double a1, double a2, double a3)
{
    su2_q r;
    r.a0 = a0;
    r.a1 = a1;
    r.a2 = a2;
    r.a3 = a3;
    return r;
}

SU2_Q_FN su2_q su2_q_zero(void) { return su2_q_make(0.0, 0.0, 0.0, 0.0); }
SU2_Q_FN su2_q su2_q_id(void) { return su2_q_make(1.0, 0.0, 0.0, 0.0); }

/* A B  (functions/su2.h:50-55) */
SU2_Q_FN su2_q su2_q_mul(su2_q A, su2_q B)
{
    return su2_q_make(SU2_FMA(-A.a3, B.a3, SU2_FMA(-A.a2, B.a2, SU2_FMA(-A.a1, B.a1, A.a0 * B.a0))),
                      SU2_FMA(-A.a3, B.a2, SU2_FMA(A.a2, B.a3, SU2_FMA(A.a0, B.a1, A.a1 * B.a0))),
                      SU2_FMA(-A.a1, B.a3, SU2_FMA(A.a0, B.a2, SU2_FMA(A.a3, B.a1, A.a2 * B.a0))),
                      SU2_FMA(A.a1, B.a2, SU2_FMA(A.a0, B.a3, SU2_FMA(-A.a2, B.a1, A.a3 * B.a0))));
}

/* A B^dagger  (functions/su2.h:56-61) */
SU2_Q_FN su2_q su2_q_mul_nd(su2_q A, su2_q B)
{
    return su2_q_make(SU2_FMA(A.a3, B.a3, SU2_FMA(A.a2, B.a2, SU2_FMA(A.a1, B.a1, A.a0 * B.a0))),
                      SU2_FMA(A.a3, B.a2, SU2_FMA(-A.a2, B.a3, SU2_FMA(-A.a0, B.a1, A.a1 * B.a0))),
                      SU2_FMA(A.a1, B.a3, SU2_FMA(-A.a0, B.a2, SU2_FMA(-A.a3, B.a1, A.a2 * B.a0))),
                      SU2_FMA(-A.a1, B.a2, SU2_FMA(-A.a0, B.a3, SU2_FMA(A.a2, B.a1, A.a3 * B.a0))));
}

/* A^dagger B  (functions/su2.h:62-67) */
SU2_Q_FN su2_q su2_q_mul_dn(su2_q A, su2_q B)
{
    return su2_q_make(SU2_FMA(A.a3, B.a3, SU2_FMA(A.a2, B.a2, SU2_FMA(A.a1, B.a1, A.a0 * B.a0))),
                      SU2_FMA(A.a3, B.a2, SU2_FMA(-A.a2, B.a3, SU2_FMA(A.a0, B.a1, -A.a1 * B.a0))),
                      SU2_FMA(A.a1, B.a3, SU2_FMA(A.a0, B.a2, SU2_FMA(-A.a3, B.a1, -A.a2 * B.a0))),
                      SU2_FMA(-A.a1, B.a2, SU2_FMA(A.a0, B.a3, SU2_FMA(A.a2, B.a1, -A.a3 * B.a0))));
}

/* A^dagger B^dagger  (functions/su2.h:68-73) */
SU2_Q_FN su2_q su2_q_mul_dd(su2_q A, su2_q B)
{
    return su2_q_make(SU2_FMA(-A.a3, B.a3, SU2_FMA(-A.a2, B.a2, SU2_FMA(-A.a1, B.a1, A.a0 * B.a0))),
                      SU2_FMA(-A.a3, B.a2, SU2_FMA(A.a2, B.a3, SU2_FMA(-A.a0, B.a1, -A.a1 * B.a0))),
                      SU2_FMA(-A.a1, B.a3, SU2_FMA(-A.a0, B.a2, SU2_FMA(A.a3, B.a1, -A.a2 * B.a0))),
                      SU2_FMA(A.a1, B.a2, SU2_FMA(-A.a0, B.a3, SU2_FMA(-A.a2, B.a1, -A.a3 * B.a0))));
}

/* First component of A B^dagger: 1/2 Re tr(A B^dagger). */
SU2_Q_FN double su2_q_mul_nd_re(su2_q A, su2_q B)
{
    return SU2_FMA(A.a3, B.a3, SU2_FMA(A.a2, B.a2, SU2_FMA(A.a1, B.a1, A.a0 * B.a0)));
}

/* 1/2 Re tr(A B)  (functions/su2.h:74-76) */
SU2_Q_FN double su2_q_retr(su2_q A, su2_q B)
{
    return SU2_FMA(-A.a3, B.a3, SU2_FMA(-A.a2, B.a2, SU2_FMA(-A.a1, B.a1, A.a0 * B.a0)));
}

/* A + B, component-wise (functions/su2.h:77-80) */
SU2_Q_FN su2_q su2_q_add(su2_q A, su2_q B)
{
    return su2_q_make(A.a0 + B.a0, A.a1 + B.a1, A.a2 + B.a2, A.a3 + B.a3);
}

/* A - B (functions/su2.h:81-84) */
SU2_Q_FN su2_q su2_q_sub(su2_q A, su2_q B)
{
    return su2_q_make(A.a0 - B.a0, A.a1 - B.a1, A.a2 - B.a2, A.a3 - B.a3);
}

SU2_Q_FN su2_q su2_q_scale(su2_q A, double s)
{
    return su2_q_make(A.a0 * s, A.a1 * s, A.a2 * s, A.a3 * s);
}

/* |a|^2 = det  (functions/su2.h:1-3) */
SU2_Q_FN double su2_q_det(su2_q A)
{
    return SU2_FMA(A.a3, A.a3, SU2_FMA(A.a2, A.a2, SU2_FMA(A.a1, A.a1, A.a0 * A.a0)));
}

/* a <- a / |a|  (functions/su2.h:4-8) */
SU2_Q_FN su2_q su2_q_normalize(su2_q A)
{
    return su2_q_scale(A, su2_rsqrt(su2_q_det(A)));
}

/* a <- a / |a|, returning |a| as det * (1/sqrt det)  (functions/su2.h:16-21) */
SU2_Q_FN su2_q su2_q_normalize_det(su2_q A, double *norm)
{
    double det = su2_q_det(A);
    double inv = su2_rsqrt(det);
    *norm = det * inv;
    return su2_q_scale(A, inv);
}

/*
 * FP32-link staple paths.  With links stored in float, each three-link path
 * of a staple is multiplied in float, in the fused pattern above with fmaf,
 * and the path enters the staple sum as a double: the sum of the paths, the
 * accept tests, reunitarisation and the observables stay double.  The float
 * operands are the stored links themselves (exact), so the products round
 * like the FP64 ones with a 24-bit significand.
 */
typedef struct su2_qf {
    float a0, a1, a2, a3;
} su2_qf;

#define SU2_FMAF(a, b, c) __builtin_fmaf((a), (b), (c))

SU2_Q_FN su2_qf su2_qf_make(float a0, float a1, float a2, float a3)
{
    su2_qf r;
    r.a0 = a0;
    r.a1 = a1;
    r.a2 = a2;
    r.a3 = a3;
    return r;
}

/* A B in float (the pattern of su2_q_mul) */
SU2_Q_FN su2_qf su2_qf_mul(su2_qf A, su2_qf B)
{
    return su2_qf_make(SU2_FMAF(-A.a3, B.a3, SU2_FMAF(-A.a2, B.a2, SU2_FMAF(-A.a1, B.a1, A.a0 * B.a0))),
                       SU2_FMAF(-A.a3, B.a2, SU2_FMAF(A.a2, B.a3, SU2_FMAF(A.a0, B.a1, A.a1 * B.a0))),
                       SU2_FMAF(-A.a1, B.a3, SU2_FMAF(A.a0, B.a2, SU2_FMAF(A.a3, B.a1, A.a2 * B.a0))),
                       SU2_FMAF(A.a1, B.a2, SU2_FMAF(A.a0, B.a3, SU2_FMAF(-A.a2, B.a1, A.a3 * B.a0))));
}

/* A B^dagger in float (the pattern of su2_q_mul_nd) */
SU2_Q_FN su2_qf su2_qf_mul_nd(su2_qf A, su2_qf B)
{
    return su2_qf_make(SU2_FMAF(A.a3, B.a3, SU2_FMAF(A.a2, B.a2, SU2_FMAF(A.a1, B.a1, A.a0 * B.a0))),
                       SU2_FMAF(A.a3, B.a2, SU2_FMAF(-A.a2, B.a3, SU2_FMAF(-A.a0, B.a1, A.a1 * B.a0))),
                       SU2_FMAF(A.a1, B.a3, SU2_FMAF(-A.a0, B.a2, SU2_FMAF(-A.a3, B.a1, A.a2 * B.a0))),
                       SU2_FMAF(-A.a1, B.a2, SU2_FMAF(-A.a0, B.a3, SU2_FMAF(A.a2, B.a1, A.a3 * B.a0))));
}

/* A^dagger B^dagger in float (the pattern of su2_q_mul_dd) */
SU2_Q_FN su2_qf su2_qf_mul_dd(su2_qf A, su2_qf B)
{
    return su2_qf_make(SU2_FMAF(-A.a3, B.a3, SU2_FMAF(-A.a2, B.a2, SU2_FMAF(-A.a1, B.a1, A.a0 * B.a0))),
                       SU2_FMAF(-A.a3, B.a2, SU2_FMAF(A.a2, B.a3, SU2_FMAF(-A.a0, B.a1, -A.a1 * B.a0))),
                       SU2_FMAF(-A.a1, B.a3, SU2_FMAF(-A.a0, B.a2, SU2_FMAF(A.a3, B.a1, -A.a2 * B.a0))),
                       SU2_FMAF(A.a1, B.a2, SU2_FMAF(-A.a0, B.a3, SU2_FMAF(-A.a2, B.a1, -A.a3 * B.a0))));
}

/* widening to double, exact */
SU2_Q_FN su2_q su2_q_of_qf(su2_qf A) { return su2_q_make((double)A.a0, (double)A.a1, (double)A.a2, (double)A.a3); }

#endif /* SU2_QUAT_H */
