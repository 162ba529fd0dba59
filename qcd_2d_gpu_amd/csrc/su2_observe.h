/*
 * su2_observe.h - the per-site plaquette terms of en_av and the
 * deterministic block sum, shared by the single-lattice and the batched
 * reductions (so a lattice's energy and plaquette come out bit for bit the
 * same in both).
 */
#ifndef SU2_OBSERVE_H
#define SU2_OBSERVE_H

#include "su2_sweep_site.h"

/* The three plaquettes at site n of parity p (functions/observables.h:1-21),
 * each returned as 1 - 1/2 tr P:
 *   P_yx = U_y(n) U_x(n+y) U_y(n+x)^dag U_x(n)^dag
 *   P_yt = U_y(n) U_t(n+y) U_y(n+t)^dag U_t(n)^dag
 *   P_xt = U_x(n) U_t(n+x) U_x(n+t)^dag U_t(n)^dag
 * The reduction of each site (functions/observables.h:99-141) gives
 * E_site = -(1 - P_xt) and P_site = (3 - a - b - c) / 3; with PMC the y-x and
 * y-t plaquettes of y = Ly-1 count as 1 (a = b = 0).  (t, x, y') is the
 * decoded half-site shared by both parities. */
template <int PMC, typename LinkT>
__device__ static inline void su2_site_plaquettes(const LinkT *__restrict__ U, const su2_kgeom &g, uint32_t p, uint32_t h,
                                              uint32_t t, uint32_t x, uint32_t yh, double *e, double *pl)
{
    const size_t Vh = g.Vh;
    const uint32_t y = 2u * yh + (((t + x) & 1u) ^ p);
    const uint32_t yp = (y + 1u == g.Ly) ? 0u : y + 1u;
    const uint32_t xp = (x + 1u == g.Lx) ? 0u : x + 1u;
    const uint32_t tp = (t + 1u == g.Lt) ? 0u : t + 1u;
    const uint32_t h_py = (yp >> 1) + x * g.Lyh + t * g.plane;
    const uint32_t h_px = yh + xp * g.Lyh + t * g.plane;
    const uint32_t h_pt = yh + x * g.Lyh + tp * g.plane;
    const LinkT *Us = U + (size_t)(p * 3) * Vh;
    const LinkT *Uo = U + (size_t)((1u - p) * 3) * Vh;

    const su2_q uy = su2_link_io<LinkT>::load(Us + 0 * Vh + h);
    const su2_q ux = su2_link_io<LinkT>::load(Us + 1 * Vh + h);
    const su2_q ut = su2_link_io<LinkT>::load(Us + 2 * Vh + h);
    const su2_q ux_py = su2_link_io<LinkT>::load(Uo + 1 * Vh + h_py);
    const su2_q ut_py = su2_link_io<LinkT>::load(Uo + 2 * Vh + h_py);
    const su2_q uy_px = su2_link_io<LinkT>::load(Uo + 0 * Vh + h_px);
    const su2_q ut_px = su2_link_io<LinkT>::load(Uo + 2 * Vh + h_px);
    const su2_q uy_pt = su2_link_io<LinkT>::load(Uo + 0 * Vh + h_pt);
    const su2_q ux_pt = su2_link_io<LinkT>::load(Uo + 1 * Vh + h_pt);

    double a = 1.0 - su2_q_mul_nd_re(su2_q_mul_nd(su2_q_mul(uy, ux_py), uy_px), ux);
    double b = 1.0 - su2_q_mul_nd_re(su2_q_mul_nd(su2_q_mul(uy, ut_py), uy_pt), ut);
    double c = 1.0 - su2_q_mul_nd_re(su2_q_mul_nd(su2_q_mul(ux, ut_px), ux_pt), ut);
    if (PMC && y == g.Ly - 1u) {
        a = 0.0;
        b = 0.0;
    }
    *e = -c;
    *pl = (3.0 - a - b - c) / 3.0;
}

/* Workgroup size of every en_av pass-1 launch (single-lattice and batched):
 * logical block b covers half-sites [(b >> 1) * SU2_EN_BLOCK, ...) of parity
 * b & 1, which fixes the partials and their summation order. */
#define SU2_EN_BLOCK 256u
/* Workgroup size of every en_av pass-2 launch. */
#define SU2_EN_FINAL_BLOCK 1024u

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
__device__ static inline __attribute__((always_inline)) void su2_en_final_partials(const double *__restrict__ work,
                                                                                uint32_t n, double *te, double *tp)
{
    const uint32_t B = SU2_EN_FINAL_BLOCK;
    double se[4] = {0.0, 0.0, 0.0, 0.0}, sp[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < n; i += 8 * B) {
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

#endif /* SU2_OBSERVE_H */
