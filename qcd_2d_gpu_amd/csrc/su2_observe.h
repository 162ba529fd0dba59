/*
 * su2_observe.h - the plaquette terms of en_av and the deterministic block
 * sums, shared by the single-lattice en_av, the sweep launch that carries a
 * measurement, and the batched reduction (so a lattice's energy and
 * plaquette come out bit for bit the same in all of them).
 */
#ifndef SU2_OBSERVE_H
#define SU2_OBSERVE_H

#include "su2_internal.h"
#include "su2_quat.h"

/*
 * en_av (functions/observables.h:1-21, 99-141) in owner form.  With
 *   a = 1 - 1/2 tr P_yx,  b = 1 - 1/2 tr P_yt,  c = 1 - 1/2 tr P_xt,
 *   P_yx(n) = U_y(n) U_x(n+y) U_y(n+x)^dag U_x(n)^dag  (P_yt, P_xt alike),
 * the reference sums E_site = -c and P_site = (3 - a - b - c) / 3 over the
 * sites (with PMC, a = b = 0 at y = Ly-1).  Here the thread of parity-0
 * half-site h, site m, owns six plaquettes:
 *   a(m), b(m), c(m), a(m - x), c(m - x), b(m - t),
 * and contributes e = -(c(m) + c(m - x)) and
 * p = (6 - a(m) - b(m) - c(m) - a(m - x) - c(m - x) - b(m - t)) / 3.
 * m -> m - x and m -> m - t map the parity-0 sites one to one onto the
 * parity-1 sites (Lx and Lt are even), so every site's three plaquettes are
 * counted once, and the sums equal the reference's up to rounding.  m, m - x
 * and m - t share y, so the PMC rule is one test per thread.  The links
 * these six plaquettes read are the twelve of the U_y staple at m, U_y(m)
 * itself and four more (U_t(m+x), U_x(m+t), U_x(m+t-x), U_t(m-x)); of them
 * only U_y(m) is written by the sweep launch (y, parity 0), and only by this
 * thread, so that launch can carry the measurement of the field it starts
 * from (su2_site_staple).
 */
struct su2_owner_plaq {
    double a, b, c, a1, c1, b2; /* a(m) b(m) c(m) a(m-x) c(m-x) b(m-t) */
    int edge;                   /* y(m) = Ly - 1 */
};

/* 1 - 1/2 tr (A B C^dag D^dag) */
__host__ __device__ static inline double su2_plaq_term(su2_q A, su2_q B, su2_q C, su2_q D)
{
    return 1.0 - su2_q_mul_nd_re(su2_q_mul_nd(su2_q_mul(A, B), C), D);
}

/* The thread's contribution to the energy and plaquette sums. */
__host__ __device__ static inline void su2_owner_finish(su2_owner_plaq q, int pmc, double *e, double *pl)
{
    if (pmc && q.edge) {
        q.a = 0.0;
        q.b = 0.0;
        q.a1 = 0.0;
        q.b2 = 0.0;
    }
    *e = -(q.c + q.c1);
    *pl = (6.0 - q.a - q.b - q.c - q.a1 - q.c1 - q.b2) / 3.0;
}

/* Workgroup size of every launch that makes en_av partials (single-lattice,
 * the measuring sweep launch, batched): logical block b, the owners of
 * parity-0 half-sites [b * SU2_EN_BLOCK, (b + 1) * SU2_EN_BLOCK), writes the
 * partial pair 2b, 2b + 1, which fixes the partials and their summation
 * order.  Blocks past the lattice write zeros. */
#define SU2_EN_BLOCK 256u
/* Workgroup size of every en_av pass-2 launch. */
#define SU2_EN_FINAL_BLOCK 1024u

/* Logical blocks (partial pairs) of a lattice's en_av: the half-sites in
 * blocks of SU2_EN_BLOCK, padded to a multiple of 8 (the XCD count). */
static inline uint32_t su2_en_av_blocks(const su2_kgeom &g) { return su2_grid_blocks(g.Vh, SU2_EN_BLOCK); }

/* Deterministic block sum of two doubles over a workgroup of nthreads (the
 * launch's blockDim.x, a multiple of 64 up to 1024): wavefront shuffles in a
 * fixed butterfly, then the wavefront totals in order (result in thread 0). */
__device__ static inline void su2_block_sum2(double &u, double &v, uint32_t nthreads)
{
    __shared__ double sh[2][16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        u += __shfl_xor(u, off, 64);
        v += __shfl_xor(v, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (int)((nthreads + 63u) >> 6);
    if (lane == 0) {
        sh[0][wave] = u;
        sh[1][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double su = 0.0, sv = 0.0;
        for (int w = 0; w < nw; ++w) {
            su += sh[0][w];
            sv += sh[1][w];
        }
        u = su;
        v = sv;
    }
}

/* The partial sums of pass 2 for one lattice, per thread of a
 * SU2_EN_FINAL_BLOCK workgroup: thread t adds partial t + k B + 4 m B (B the
 * workgroup size) into accumulator k, m ascending, k = 0..3, and returns
 * (s0 + s1) + (s2 + s3).  Eight partials are loaded per trip, so a lattice
 * of up to 8 B partials takes one memory round trip; a slot past n reads
 * partial 0 and is not added. */
__device__ static inline __attribute__((always_inline)) void su2_en_final_partials_at(const double *__restrict__ work,
                                                                                   uint32_t n, uint32_t t,
                                                                                   double *te, double *tp)
{
    const uint32_t B = SU2_EN_FINAL_BLOCK;
    double se[4] = {0.0, 0.0, 0.0, 0.0}, sp[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = t; i < n; i += 8 * B) {
        double ve[8], vp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t j = i + (uint32_t)u * B, jj = j < n ? j : 0u;
            ve[u] = work[2 * jj];
            vp[u] = work[2 * jj + 1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (i + (uint32_t)u * B < n) {
                se[u & 3] += ve[u];
                sp[u & 3] += vp[u];
            }
        }
    }
    *te = (se[0] + se[1]) + (se[2] + se[3]);
    *tp = (sp[0] + sp[1]) + (sp[2] + sp[3]);
}

__device__ static inline __attribute__((always_inline)) void su2_en_final_partials(const double *__restrict__ work,
                                                                                uint32_t n, double *te, double *tp)
{
    su2_en_final_partials_at(work, n, threadIdx.x, te, tp);
}

/* Pass 2 of en_av by a workgroup of SU2_EN_BLOCK threads with the sums and
 * the order of a SU2_EN_FINAL_BLOCK workgroup (su2_en_final_partials, then
 * su2_block_sum2), so the bits are the same: thread r stands for threads
 * r + SU2_EN_BLOCK j, j = 0..3, of the larger workgroup.  Their wavefront
 * is (r / 64) + 4 j and their lane r % 64, the lane of r in its own
 * wavefront, so one butterfly per j over the wavefront's lanes forms that
 * wavefront's total; thread 0 adds the 16 totals in wavefront order and
 * accumulates sum * invV into res, res2. */
__device__ static inline void su2_en_final_small(const double *__restrict__ work, uint32_t n, double invV,
                                                 double *__restrict__ res, double *__restrict__ res2)
{
    static_assert(SU2_EN_FINAL_BLOCK == 4 * SU2_EN_BLOCK, "four threads of the larger workgroup per thread");
    __shared__ double sh[2][16];
    const uint32_t r = threadIdx.x, lane = r & 63u, w = r >> 6;
    /* one j at a time: unrolled, the loads of all four would hold 128 VGPRs
     * and set the register count of the sweep kernel that carries this */
#pragma nounroll
    for (uint32_t j = 0; j < 4; ++j) {
        double u, v;
        su2_en_final_partials_at(work, n, r + SU2_EN_BLOCK * j, &u, &v);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            u += __shfl_xor(u, off, 64);
            v += __shfl_xor(v, off, 64);
        }
        if (lane == 0) {
            sh[0][w + 4 * j] = u;
            sh[1][w + 4 * j] = v;
        }
    }
    __syncthreads();
    if (r == 0) {
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < 16; ++k) {
            su += sh[0][k];
            sv += sh[1][k];
        }
        *res += su * invV;
        *res2 += sv * invV;
    }
}

#endif /* SU2_OBSERVE_H */
