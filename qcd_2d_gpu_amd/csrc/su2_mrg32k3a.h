/*
 * su2_mrg32k3a.h - MRG32k3a random streams, one per half-site.
 *
 * The reference keeps one curandStateMRG32k3a_t per half-site h, seeded with
 * (seed, subsequence = h, offset = 0) (functions/lattice_setup.h:15-21,
 * kernel.cu:77-79), and draws from it with curand_uniform (float, (0,1]) and
 * curand_uniform_double (double, (0,1]) (functions/su2.h:31,90;
 * functions/mc_updates.h:30,47,50,54-55).
 *
 * This file implements the generator with the published (seed, subsequence,
 * offset) semantics of AMD's MIT-licensed rocRAND MRG32k3a engine, which the
 * build treats as the contract (cuRAND's seeding is NVIDIA library code and
 * stays out of this project):
 *   - seeding: x = lo32(seed) ^ 0x55555555, y = hi32(seed) ^ 0xAAAAAAAA,
 *     g1 = (x*seed, y*seed, x*seed) and g2 = (y*seed, x*seed, y*seed) with
 *     the engine's own modular-multiply steps (exact for seeds below 2^32,
 *     reproduced step for step above); seed 0 means 12345;
 *   - a subsequence is 2^76 draws; subsequence k multiplies the state by
 *     A^(2^76 k) (table from gen_mrg32k3a_jump.py);
 *   - next() returns an integer in [1, m1]; the float draw is
 *     (float)(v * 2.3283065498378288e-10) and the double draw
 *     v * 2.3283065498378288e-10, one step each.
 * Fixtures written by rocRAND's own host code pin all of the above.
 *
 * Device state layout: six uint32 planes g1[0..2], g2[0..2], each of length
 * nstreams (structure of arrays: one coalesced dword load per plane).
 */
#ifndef SU2_MRG32K3A_H
#define SU2_MRG32K3A_H

#include <stdint.h>

#include "su2_mrg32k3a_jump.h"

#if defined(__HIP__) || defined(__HIPCC__)
#define SU2_MRG_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define SU2_MRG_FN static inline
#endif

#define SU2_MRG_M1 4294967087u
#define SU2_MRG_M2 4294944443u
#define SU2_MRG_M1C 209u
#define SU2_MRG_M2C 22853u
#define SU2_MRG_A12 1403580u
#define SU2_MRG_A13N 810728u
#define SU2_MRG_A21 527612u
#define SU2_MRG_A23N 1370589u
#define SU2_MRG_NORM 2.3283065498378288e-10 /* 1/m1 rounded to double */

/* One fold of the 2^32 = m1 + 209 identity plus one conditional subtraction
 * (the engine's reduction; exact for inputs below about 2^53). */
SU2_MRG_FN uint64_t su2_mrg_fold1(uint64_t p)
{
    p = (uint64_t)SU2_MRG_M1C * (uint32_t)(p >> 32) + (uint32_t)p;
    if (p >= SU2_MRG_M1) p -= SU2_MRG_M1;
    return p;
}

/* Two folds of 2^32 = m2 + 22853 plus one conditional subtraction. */
SU2_MRG_FN uint64_t su2_mrg_fold2(uint64_t p)
{
    p = (uint64_t)SU2_MRG_M2C * (uint32_t)(p >> 32) + (uint32_t)p;
    p = (uint64_t)SU2_MRG_M2C * (uint32_t)(p >> 32) + (uint32_t)p;
    if (p >= SU2_MRG_M2) p -= SU2_MRG_M2;
    return p;
}

/* The last fold and conditional subtraction of the recurrence step, in
 * 32-bit arithmetic: v = c * hi + lo reduced once by m = 2^32 - c.  With
 * t = lo + c * (hi + 1) mod 2^32, v >= m exactly when that addition
 * carries, and then v - m = t; otherwise v = t - c.  Equal to
 * su2_mrg_fold1's (c = 209) and su2_mrg_fold2's last (c = 22853) fold and
 * subtraction for every input the step produces.  hi_mask covers every
 * value hi takes there (below 2^22 and 2^5); it only tells the compiler
 * that c * (hi + 1) fits a 24-bit multiply with a 32-bit result. */
SU2_MRG_FN uint32_t su2_mrg_tail(uint32_t hi, uint32_t lo, uint32_t c, uint32_t hi_mask)
{
    const uint32_t t = lo + ((hi & hi_mask) * c + c);
    return (t < lo) ? t : t - c;
}

/* One step of both recurrences; s = {g1[0], g1[1], g1[2], g2[0], g2[1],
 * g2[2]}.  Returns the combined output in [1, m1].  The products are below
 * 2^54, so the high words entering su2_mrg_tail are below 2^22 (component
 * 1) and, after component 2's first 64-bit fold, below 2^5. */
SU2_MRG_FN uint32_t su2_mrg_next(uint32_t *s)
{
    const uint64_t a = (uint64_t)SU2_MRG_A12 * s[1] + (uint64_t)SU2_MRG_A13N * (uint32_t)(SU2_MRG_M1 - s[0]);
    const uint32_t p1 = su2_mrg_tail((uint32_t)(a >> 32), (uint32_t)a, SU2_MRG_M1C, 0x3fffffu);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = p1;
    uint64_t b = (uint64_t)SU2_MRG_A21 * s[5] + (uint64_t)SU2_MRG_A23N * (uint32_t)(SU2_MRG_M2 - s[3]);
    b = (uint64_t)SU2_MRG_M2C * (uint32_t)(b >> 32) + (uint32_t)b;
    const uint32_t p2 = su2_mrg_tail((uint32_t)(b >> 32), (uint32_t)b, SU2_MRG_M2C, 0x1fu);
    s[3] = s[4];
    s[4] = s[5];
    s[5] = p2;
    return (p1 - p2) + (p1 <= p2 ? SU2_MRG_M1 : 0u);
}

/* curand_uniform semantics: float in (0, 1]. */
SU2_MRG_FN float su2_mrg_uniform(uint32_t *s)
{
    return (float)((double)su2_mrg_next(s) * SU2_MRG_NORM);
}

/* curand_uniform_double semantics: double in (0, 1]. */
SU2_MRG_FN double su2_mrg_uniform_double(uint32_t *s)
{
    return (double)su2_mrg_next(s) * SU2_MRG_NORM;
}

/* The engine's seeding multiply i*j "mod m" for a 32-bit i and 64-bit j,
 * split as i = hi*2^17 + lo; the 64-bit products wrap as unsigned. */
SU2_MRG_FN uint64_t su2_mrg_seed_mul1(uint32_t i, uint64_t j)
{
    uint64_t hi = i / 131072u;
    uint64_t lo = i - hi * 131072u;
    uint64_t t1 = su2_mrg_fold1(hi * j) * 131072u;
    uint64_t t2 = su2_mrg_fold1(lo * j);
    return su2_mrg_fold1(t1 + t2);
}

SU2_MRG_FN uint64_t su2_mrg_seed_mul2(uint32_t i, uint64_t j)
{
    uint64_t hi = i / 131072u;
    uint64_t lo = i - hi * 131072u;
    uint64_t t1 = su2_mrg_fold2(hi * j) * 131072u;
    uint64_t t2 = su2_mrg_fold2(lo * j);
    return su2_mrg_fold2(t1 + t2);
}

/* State at (seed, subsequence 0, offset 0). */
SU2_MRG_FN void su2_mrg_seed(uint64_t seed, uint32_t *s)
{
    if (seed == 0) seed = 12345ull;
    uint32_t x = (uint32_t)seed ^ 0x55555555u;
    uint32_t y = (uint32_t)(seed >> 32) ^ 0xAAAAAAAAu;
    s[0] = (uint32_t)su2_mrg_seed_mul1(x, seed);
    s[1] = (uint32_t)su2_mrg_seed_mul1(y, seed);
    s[2] = (uint32_t)su2_mrg_seed_mul1(x, seed);
    s[3] = (uint32_t)su2_mrg_seed_mul2(y, seed);
    s[4] = (uint32_t)su2_mrg_seed_mul2(x, seed);
    s[5] = (uint32_t)su2_mrg_seed_mul2(y, seed);
}

/* s[0..2] <- M s[0..2] with the engine's reduction order. */
SU2_MRG_FN void su2_mrg_matvec1(const uint32_t *M, uint32_t *s)
{
    uint64_t x0 = s[0], x1 = s[1], x2 = s[2];
    s[0] = (uint32_t)su2_mrg_fold1(su2_mrg_fold1(M[0] * x0) + su2_mrg_fold1(M[1] * x1) + su2_mrg_fold1(M[2] * x2));
    s[1] = (uint32_t)su2_mrg_fold1(su2_mrg_fold1(M[3] * x0) + su2_mrg_fold1(M[4] * x1) + su2_mrg_fold1(M[5] * x2));
    s[2] = (uint32_t)su2_mrg_fold1(su2_mrg_fold1(M[6] * x0) + su2_mrg_fold1(M[7] * x1) + su2_mrg_fold1(M[8] * x2));
}

SU2_MRG_FN void su2_mrg_matvec2(const uint32_t *M, uint32_t *s)
{
    uint64_t x0 = s[0], x1 = s[1], x2 = s[2];
    s[0] = (uint32_t)su2_mrg_fold2(su2_mrg_fold2(M[0] * x0) + su2_mrg_fold2(M[1] * x1) + su2_mrg_fold2(M[2] * x2));
    s[1] = (uint32_t)su2_mrg_fold2(su2_mrg_fold2(M[3] * x0) + su2_mrg_fold2(M[4] * x1) + su2_mrg_fold2(M[5] * x2));
    s[2] = (uint32_t)su2_mrg_fold2(su2_mrg_fold2(M[6] * x0) + su2_mrg_fold2(M[7] * x1) + su2_mrg_fold2(M[8] * x2));
}

/* Jump s ahead by `sub` subsequences using tables J1, J2 of
 * A^(2^(76+i)) (9 entries each, SU2_MRG_NJUMP of them). */
SU2_MRG_FN void su2_mrg_jump(uint32_t *s, uint32_t sub, const uint32_t *J1, const uint32_t *J2)
{
    int i = 0;
    while (sub) {
        if (sub & 1u) {
            su2_mrg_matvec1(J1 + 9 * i, s);
            su2_mrg_matvec2(J2 + 9 * i, s + 3);
        }
        sub >>= 1;
        ++i;
    }
}

#endif /* SU2_MRG32K3A_H */
