/*
 * su2_sweep_site.h - the update of one link U_mu(n), n = half-site h of
 * parity p, shared by the single-lattice launch (su2_sweep.hip) and the
 * batched launch over many lattices (su2_batch.hip), so a lattice's links
 * and random streams evolve bit for bit the same in both.
 *
 * Staple of U_mu(n) (functions/mc_updates.h:155-176 and the other five
 * kernels), summed over nu != mu in the reference's order (nu ascending in
 * y, x, t; the "positive" path before the "negative" one):
 *   A += [U_nu(n+mu) U_mu(n+nu)^dag] U_nu(n)^dag
 *   A += [U_nu(n+mu-nu)^dag U_mu(n-nu)^dag] U_nu(n-nu)
 * Parities: n+mu, n+nu, n-nu lie on 1-p; n and n+mu-nu on p.
 *
 * Boundary conditions in y:
 *   PCC (mc_updates_pcc.h:29,83,134,189): x and t links at y = 0 are not
 *       updated (no draws, no acceptance count).
 *   PMC (mc_updates_pmc.h:33,40,96,103,156,163,219,225): for the x and t
 *       links the nu = y positive path is dropped at y = Ly-1 and the
 *       negative path at y = 0.  y links use the periodic staple in both.
 */
#ifndef SU2_SWEEP_SITE_H
#define SU2_SWEEP_SITE_H

#include "su2_internal.h"
#include "su2_observe.h"
#include "su2_update.h"

template <typename LinkT> struct su2_link_io;

template <> struct su2_link_io<su2_l64> {
    __device__ static inline __attribute__((always_inline)) su2_q load(const su2_l64 *p)
    {
        su2_l64 v = *p;
        return su2_q_make(v.a0, v.a1, v.a2, v.a3);
    }
    __device__ static inline __attribute__((always_inline)) void store(su2_l64 *p, su2_q q)
    {
        su2_l64 v;
        v.a0 = q.a0;
        v.a1 = q.a1;
        v.a2 = q.a2;
        v.a3 = q.a3;
        *p = v;
    }
};

template <> struct su2_link_io<su2_l32> {
    __device__ static inline __attribute__((always_inline)) su2_qf loadf(const su2_l32 *p)
    {
        su2_l32 v = *p;
        return su2_qf_make(v.a0, v.a1, v.a2, v.a3);
    }
    __device__ static inline __attribute__((always_inline)) su2_q load(const su2_l32 *p)
    {
        su2_l32 v = *p;
        return su2_q_make((double)v.a0, (double)v.a1, (double)v.a2, (double)v.a3);
    }
    __device__ static inline __attribute__((always_inline)) void store(su2_l32 *p, su2_q q)
    {
        su2_l32 v;
        v.a0 = (float)q.a0;
        v.a1 = (float)q.a1;
        v.a2 = (float)q.a2;
        v.a3 = (float)q.a3;
        *p = v;
    }
};

/* Workgroup -> block of 256 half-sites.  The hardware deals workgroups
 * round-robin to the 8 XCDs, each with its own 4 MB L2; the block order
 * decides which part of the lattice an XCD walks, and in which order.
 *
 * ORDER_SLAB: XCD k takes the k-th contiguous eighth of the lattice (a slab
 * of t-planes) and walks it in memory order.
 *
 * ORDER_STRIPS (the host sets it only when every plane is a whole number
 * of blocks, the slabs are whole planes and the plane's bpp blocks split
 * into strips of `strip` blocks): the XCD walks its slab strip by strip,
 * all its t-planes for one strip of x-rows before the next, so the
 * t-neighbour rows of a strip are re-read from L2 a few hundred KB later
 * instead of two full planes (2.6 MB at 128^3) later.  Measured with FP64
 * links at 128^3: over-relaxation launch -10 %, heat-bath -5 %.  At 256^3
 * (128 blocks per plane) the L2-miss bytes of a heat-bath launch drop from
 * 1.55 to 1.10 times the compulsory bytes, the t-neighbour planes no longer
 * re-fetched through the Infinity Cache, but the launch gains only 1.7 %:
 * there the time is set by the DRAM access pattern itself (a copy of the
 * kernel with its arithmetic removed takes 99 % of its time).  FP32 links
 * (half the footprint) gain nothing from it and keep the slab order.
 * Strips are 32 blocks wide on planes of 64 blocks or more, else 16 where
 * the plane's block count allows it, else 8: at 128^3 (32 blocks per plane)
 * 16 against 8 gave heat-bath -4 %, step -2 %, and 4 and 32 were slower
 * than 8; at 256^3 (128 blocks per plane) 8 was 5 % slower than 16, and 32
 * against 16 gave over-relaxation -3.3 % (380 -> 368 us) with the heat-bath
 * unchanged (alternating runs on one box; 64 and 128 were 3-4 % slower).
 *
 * ORDER_COLUMNS (thin lattices, Lt <= 32, with bpp a multiple of 8): a slab
 * of Lt/8 <= 4 planes would send half or more of an XCD's t-neighbour reads
 * to another XCD's planes.  Instead XCD k takes the k-th eighth of every
 * plane's blocks (a column of x-rows through all t) and walks it t-major;
 * only the x-neighbour rows at the column's two edges belong to other
 * XCDs.  Measured with FP64 links against ORDER_SLAB: 32x256x256
 * heat-bath -10 %, over-relaxation -30 %; 16x256x256 -13 % and -25 %
 * (one GPU's share of 256^3 slab-decomposed over 8 or 16 GPUs); on 256^3
 * itself it was 2.5 % slower. */
enum { ORDER_SLAB = 0, ORDER_STRIPS = 1, ORDER_COLUMNS = 2 };


/* Workgroup b of nblk (nblk a multiple of 8) -> block (su2_block_of for a
 * launch whose first workgroups do other work). */
__device__ static inline __attribute__((always_inline)) uint32_t su2_block_at(uint32_t b, uint32_t nblk,
                                                                              uint32_t order, uint32_t strip,
                                                                              uint32_t bpp)
{
    const uint32_t xcd = b & 7u, j = b >> 3, per = nblk >> 3;
    if (order == ORDER_COLUMNS) {
        const uint32_t W = bpp >> 3; /* blocks of a plane in this XCD's column */
        const uint32_t tl = j / W;
        return tl * bpp + xcd * W + (j - tl * W);
    }
    if (order == ORDER_SLAB) return xcd * per + j;
    const uint32_t T = per / bpp; /* t-planes of this XCD's slab */
    const uint32_t s = j / (T * strip), r = j - s * (T * strip);
    const uint32_t tl = r / strip, c = r - tl * strip;
    return (xcd * T + tl) * bpp + s * strip + c;
}

__device__ static inline __attribute__((always_inline)) uint32_t su2_block_of(uint32_t order, uint32_t strip,
                                                                              uint32_t bpp)
{
    return su2_block_at(blockIdx.x, gridDim.x, order, strip, bpp);
}

/* The block order of a launch over the half-sites in blocks of `block`
 * (see above): columns or strips for FP64 links (f64) when the lattice tiles
 * evenly, else the slab order.  bpp = blocks per t-plane. */
static inline void su2_pick_block_order(const su2_kgeom &g, uint32_t block, bool f64, uint32_t *order,
                                        uint32_t *strip, uint32_t *bpp)
{
    const uint32_t nb = su2_grid_blocks(g.Vh, block);
    *order = ORDER_SLAB;
    *strip = 0;
    *bpp = g.plane / block;
    const uint32_t w = (*bpp >= 64u && *bpp % 32u == 0) ? 32u : (*bpp % 16u == 0) ? 16u : 8u;
    const bool whole = g.plane % block == 0 && nb * block == g.Vh;
    if (f64 && whole && g.Lt <= 32u && *bpp % 8u == 0) {
        *order = ORDER_COLUMNS;
    } else if (f64 && whole && *bpp % w == 0 && *bpp <= 128 && (g.Vh / block) % (8u * *bpp) == 0) {
        *order = ORDER_STRIPS;
        *strip = w;
    }
}

/* Half-site index of (t, x, y). */
__device__ static inline __attribute__((always_inline)) uint32_t su2_hidx(const su2_kgeom &g, uint32_t t, uint32_t x,
                                                                      uint32_t y)
{
    return (y >> 1) + x * g.Lyh + t * g.plane;
}

/* The staple A of U_mu at half-site h of parity p; returns 0 (and leaves A
 * alone) for a link the boundary condition does not update (PCC: x and t
 * links at y = 0).  BC = SU2_K_BC_ANY takes the boundary from bc.
 *
 * MEAS (MU = y, parity 0 only): also the six owned plaquettes of the
 * thread's site m (su2_observe.h) into *meas, from the field as it is before
 * this launch writes: they reuse the staple's twelve links and read U_y(m)
 * and four more; the writes of the launch touch U_y at parity 0 only, and
 * U_y(m) only in this thread, after this returns. */
template <int MU, int BC, typename LinkT, bool MEAS = false>
__device__ static inline __attribute__((always_inline)) int su2_site_staple(const LinkT *__restrict__ U,
                                                                            const su2_kgeom &g, int parity,
                                                                            uint32_t h, su2_q *Aout, int bc = BC,
                                                                            su2_owner_plaq *meas = nullptr)
{
    static_assert(!MEAS || MU == SU2_DIR_Y, "a measuring launch is a y launch");
    typedef su2_link_io<LinkT> io;
    const bool pcc = BC == SU2_K_BC_PCC || (BC == SU2_K_BC_ANY && bc == SU2_K_BC_PCC);
    const bool pmc = BC == SU2_K_BC_PMC || (BC == SU2_K_BC_ANY && bc == SU2_K_BC_PMC);

    const uint32_t q = su2_div(h, g.div_lyh);
    const uint32_t yh = h - q * g.Lyh;
    const uint32_t t = su2_div(q, g.div_lx);
    const uint32_t x = q - t * g.Lx;
    const uint32_t y = 2u * yh + (((t + x) & 1u) ^ (uint32_t)parity);

    if (pcc && MU != SU2_DIR_Y && y == 0u) return 0;

    /* wrapped neighbour coordinates */
    const uint32_t yp = (y + 1u == g.Ly) ? 0u : y + 1u;
    const uint32_t ym = (y == 0u) ? g.Ly - 1u : y - 1u;
    const uint32_t xp = (x + 1u == g.Lx) ? 0u : x + 1u;
    const uint32_t xm = (x == 0u) ? g.Lx - 1u : x - 1u;
    const uint32_t tp = (t + 1u == g.Lt) ? 0u : t + 1u;
    const uint32_t tm = (t == 0u) ? g.Lt - 1u : t - 1u;

    const LinkT *Us = U + (size_t)(parity * 3) * g.Vh;       /* parity p */
    const LinkT *Uo = U + (size_t)((1 - parity) * 3) * g.Vh; /* parity 1-p */
    const size_t Vh = g.Vh;

    /* coordinates of n + mu */
    const uint32_t tmu = (MU == SU2_DIR_T) ? tp : t;
    const uint32_t xmu = (MU == SU2_DIR_X) ? xp : x;
    const uint32_t ymu = (MU == SU2_DIR_Y) ? yp : y;
    const uint32_t h_pmu = su2_hidx(g, tmu, xmu, ymu);

    /* MEAS: U_y(m) before this launch overwrites it, and the links the x-t
     * plaquettes share with the staple */
    su2_q uy_m, ux_m, ux_mx, ut_m;
    if constexpr (MEAS) uy_m = io::load(Us + SU2_DIR_Y * Vh + h);

    su2_q A = su2_q_zero();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int NU = (k == 0) ? (MU == SU2_DIR_Y ? SU2_DIR_X : SU2_DIR_Y) : (MU == SU2_DIR_T ? SU2_DIR_X : SU2_DIR_T);
        /* n + nu, n - nu, n + mu - nu */
        const uint32_t h_pnu = su2_hidx(g, NU == SU2_DIR_T ? tp : t, NU == SU2_DIR_X ? xp : x, NU == SU2_DIR_Y ? yp : y);
        const uint32_t h_mnu = su2_hidx(g, NU == SU2_DIR_T ? tm : t, NU == SU2_DIR_X ? xm : x, NU == SU2_DIR_Y ? ym : y);
        const uint32_t h_pmu_mnu =
            su2_hidx(g, NU == SU2_DIR_T ? tm : tmu, NU == SU2_DIR_X ? xm : xmu, NU == SU2_DIR_Y ? ym : ymu);

        su2_q up, dn;
        /* the six links in double, for a measuring launch */
        su2_q l0, l1, l2, l3, l4, l5;
        if constexpr (sizeof(LinkT) == sizeof(su2_l32)) {
            /* FP32 links: the paths multiplied in float (su2_quat.h) */
            const su2_qf u_nu_pmu = io::loadf(Uo + NU * Vh + h_pmu);
            const su2_qf u_mu_pnu = io::loadf(Uo + MU * Vh + h_pnu);
            const su2_qf u_nu_n = io::loadf(Us + NU * Vh + h);
            const su2_qf u_nu_pmu_mnu = io::loadf(Us + NU * Vh + h_pmu_mnu);
            const su2_qf u_mu_mnu = io::loadf(Uo + MU * Vh + h_mnu);
            const su2_qf u_nu_mnu = io::loadf(Uo + NU * Vh + h_mnu);
            up = su2_q_of_qf(su2_qf_mul_nd(su2_qf_mul_nd(u_nu_pmu, u_mu_pnu), u_nu_n));
            dn = su2_q_of_qf(su2_qf_mul(su2_qf_mul_dd(u_nu_pmu_mnu, u_mu_mnu), u_nu_mnu));
            if constexpr (MEAS) {
                l0 = su2_q_of_qf(u_nu_pmu);
                l1 = su2_q_of_qf(u_mu_pnu);
                l2 = su2_q_of_qf(u_nu_n);
                l3 = su2_q_of_qf(u_nu_pmu_mnu);
                l4 = su2_q_of_qf(u_mu_mnu);
                l5 = su2_q_of_qf(u_nu_mnu);
            }
        } else {
            const su2_q u_nu_pmu = io::load(Uo + NU * Vh + h_pmu);
            const su2_q u_mu_pnu = io::load(Uo + MU * Vh + h_pnu);
            const su2_q u_nu_n = io::load(Us + NU * Vh + h);
            const su2_q u_nu_pmu_mnu = io::load(Us + NU * Vh + h_pmu_mnu);
            const su2_q u_mu_mnu = io::load(Uo + MU * Vh + h_mnu);
            const su2_q u_nu_mnu = io::load(Uo + NU * Vh + h_mnu);
            up = su2_q_mul_nd(su2_q_mul_nd(u_nu_pmu, u_mu_pnu), u_nu_n);
            dn = su2_q_mul(su2_q_mul_dd(u_nu_pmu_mnu, u_mu_mnu), u_nu_mnu);
            if constexpr (MEAS) {
                l0 = u_nu_pmu;
                l1 = u_mu_pnu;
                l2 = u_nu_n;
                l3 = u_nu_pmu_mnu;
                l4 = u_mu_mnu;
                l5 = u_nu_mnu;
            }
        }
        if constexpr (MEAS) {
            /* k = 0 (nu = x): l0..l5 = U_x(m+y), U_y(m+x), U_x(m), U_x(m+y-x),
             * U_y(m-x), U_x(m-x): a(m) and a(m-x); k = 1 (nu = t) likewise
             * b(m) and b(m-t) */
            const double fwd = su2_plaq_term(uy_m, l0, l1, l2);
            const double bwd = su2_plaq_term(l4, l3, uy_m, l5);
            if (k == 0) {
                meas->a = fwd;
                meas->a1 = bwd;
                ux_m = l2;
                ux_mx = l5;
            } else {
                meas->b = fwd;
                meas->b2 = bwd;
                ut_m = l2;
            }
        }
        if (pmc && MU != SU2_DIR_Y && NU == SU2_DIR_Y) {
            /* A is never -0 here, so adding +0 leaves it bit-identical. */
            if (y == g.Ly - 1u) up = su2_q_zero();
            if (y == 0u) dn = su2_q_zero();
        }
        A = su2_q_add(A, up);
        A = su2_q_add(A, dn);
    }
    if constexpr (MEAS) {
        /* c(m) = P_xt(m) and c(m-x) with U_t(m+x), U_x(m+t), U_x(m+t-x), U_t(m-x) */
        const uint32_t h_px = su2_hidx(g, t, xp, y), h_pt = su2_hidx(g, tp, x, y);
        const uint32_t h_mx = su2_hidx(g, t, xm, y), h_ptmx = su2_hidx(g, tp, xm, y);
        const su2_q ut_px = io::load(Uo + SU2_DIR_T * Vh + h_px);
        const su2_q ux_pt = io::load(Uo + SU2_DIR_X * Vh + h_pt);
        const su2_q ux_ptmx = io::load(Us + SU2_DIR_X * Vh + h_ptmx);
        const su2_q ut_mx = io::load(Uo + SU2_DIR_T * Vh + h_mx);
        meas->c = su2_plaq_term(ux_m, ut_px, ux_pt, ut_m);
        meas->c1 = su2_plaq_term(ux_mx, ut_m, ux_ptmx, ut_mx);
        meas->edge = y == g.Ly - 1u;
    }
    *Aout = A;
    return 1;
}

/* The owned plaquettes of parity-0 half-site h (su2_observe.h), for an
 * en_av launch of its own: the same links as a measuring y launch reads,
 * loaded as doubles.  PART 0: a(m), a(m-x), b(m) (ten links); PART 1:
 * b(m-t), c(m), c(m-x) (eleven links); PART 2: all six.  Two threads that
 * split a site between them (parts 0 and 1) hold between them exactly the
 * values one thread computes alone. */
template <int PART, typename LinkT>
__device__ static inline __attribute__((always_inline)) void su2_owner_terms(const LinkT *__restrict__ U,
                                                                             const su2_kgeom &g, uint32_t h,
                                                                             su2_owner_plaq *meas)
{
    typedef su2_link_io<LinkT> io;
    const uint32_t q = su2_div(h, g.div_lyh);
    const uint32_t yh = h - q * g.Lyh;
    const uint32_t t = su2_div(q, g.div_lx);
    const uint32_t x = q - t * g.Lx;
    const uint32_t y = 2u * yh + ((t + x) & 1u); /* parity 0 */
    const uint32_t yp = (y + 1u == g.Ly) ? 0u : y + 1u;
    const uint32_t xp = (x + 1u == g.Lx) ? 0u : x + 1u;
    const uint32_t xm = (x == 0u) ? g.Lx - 1u : x - 1u;
    const uint32_t tp = (t + 1u == g.Lt) ? 0u : t + 1u;
    const uint32_t tm = (t == 0u) ? g.Lt - 1u : t - 1u;
    const size_t Vh = g.Vh;
    const LinkT *U0 = U, *U1 = U + 3 * Vh; /* parity 0, parity 1 */
    const uint32_t h_py = su2_hidx(g, t, x, yp), h_px = su2_hidx(g, t, xp, y), h_pt = su2_hidx(g, tp, x, y);
    const uint32_t h_mx = su2_hidx(g, t, xm, y), h_mt = su2_hidx(g, tm, x, y);
    const su2_q uy = io::load(U0 + SU2_DIR_Y * Vh + h);
    const su2_q ut = io::load(U0 + SU2_DIR_T * Vh + h);
    if constexpr (PART != 1) {
        /* y-x: a(m), a(m-x); y-t: b(m) */
        const uint32_t h_pymx = su2_hidx(g, t, xm, yp);
        const su2_q ux = io::load(U0 + SU2_DIR_X * Vh + h);
        const su2_q ux_py = io::load(U1 + SU2_DIR_X * Vh + h_py), uy_px = io::load(U1 + SU2_DIR_Y * Vh + h_px);
        const su2_q ux_pymx = io::load(U0 + SU2_DIR_X * Vh + h_pymx), uy_mx = io::load(U1 + SU2_DIR_Y * Vh + h_mx);
        const su2_q ux_mx = io::load(U1 + SU2_DIR_X * Vh + h_mx);
        const su2_q ut_py = io::load(U1 + SU2_DIR_T * Vh + h_py), uy_pt = io::load(U1 + SU2_DIR_Y * Vh + h_pt);
        meas->a = su2_plaq_term(uy, ux_py, uy_px, ux);
        meas->a1 = su2_plaq_term(uy_mx, ux_pymx, uy, ux_mx);
        meas->b = su2_plaq_term(uy, ut_py, uy_pt, ut);
    }
    if constexpr (PART != 0) {
        /* y-t: b(m-t); x-t: c(m), c(m-x) */
        const uint32_t h_pymt = su2_hidx(g, tm, x, yp), h_ptmx = su2_hidx(g, tp, xm, y);
        const su2_q ux = io::load(U0 + SU2_DIR_X * Vh + h);
        const su2_q ut_pymt = io::load(U0 + SU2_DIR_T * Vh + h_pymt), uy_mt = io::load(U1 + SU2_DIR_Y * Vh + h_mt);
        const su2_q ut_mt = io::load(U1 + SU2_DIR_T * Vh + h_mt);
        meas->b2 = su2_plaq_term(uy_mt, ut_pymt, uy, ut_mt);
        const su2_q ut_px = io::load(U1 + SU2_DIR_T * Vh + h_px), ux_pt = io::load(U1 + SU2_DIR_X * Vh + h_pt);
        const su2_q ux_mx = io::load(U1 + SU2_DIR_X * Vh + h_mx);
        const su2_q ux_ptmx = io::load(U0 + SU2_DIR_X * Vh + h_ptmx), ut_mx = io::load(U1 + SU2_DIR_T * Vh + h_mx);
        meas->c = su2_plaq_term(ux, ut_px, ux_pt, ut);
        meas->c1 = su2_plaq_term(ux_mx, ut, ux_ptmx, ut_mx);
    }
    meas->edge = y == g.Ly - 1u;
}

/* Workgroup size of the en_av partial launches (single-lattice and
 * batched): two threads per owned site with FP64 links, one with FP32
 * links (su2_en_av_owner_block). */
static inline uint32_t su2_en_av_threads(int f32) { return f32 ? SU2_EN_BLOCK : 2 * SU2_EN_BLOCK; }

/* One logical block of en_av pass 1 (owners of parity-0 half-sites
 * [b * SU2_EN_BLOCK, ...)): partials 2b, 2b + 1 into work, the same bits as a
 * measuring sweep launch writes.  FP64 links: each owned site takes two
 * threads of a 2 * SU2_EN_BLOCK workgroup, the first half computing a(m),
 * a(m-x), b(m), the second b(m-t), c(m), c(m-x) (ten and eleven links
 * instead of seventeen in one thread); the second half hands its three
 * values over in LDS and the first half forms and reduces the owners' sums
 * exactly as one thread per owner does.  128^3 FP64: 40 us against 50 us
 * with one thread per owner; FP32 links: 26 us against 24 us, so they keep
 * one thread per owner.  Every thread of the workgroup calls this. */
template <typename LinkT>
__device__ static inline __attribute__((always_inline)) void su2_en_av_owner_block(const LinkT *__restrict__ U,
                                                                                   const su2_kgeom &g, int pmc,
                                                                                   uint32_t b,
                                                                                   double *__restrict__ work)
{
    constexpr bool split = sizeof(LinkT) == sizeof(su2_l64);
    const uint32_t o = threadIdx.x % SU2_EN_BLOCK, second = threadIdx.x / SU2_EN_BLOCK;
    const uint32_t h = b * SU2_EN_BLOCK + o;
    su2_owner_plaq m = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    double e = 0.0, pl = 0.0;
    if constexpr (split) {
        __shared__ double xch[3][SU2_EN_BLOCK];
        if (h < g.Vh) {
            if (second)
                su2_owner_terms<1, LinkT>(U, g, h, &m);
            else
                su2_owner_terms<0, LinkT>(U, g, h, &m);
        }
        if (second) {
            xch[0][o] = m.b2;
            xch[1][o] = m.c;
            xch[2][o] = m.c1;
        }
        __syncthreads();
        if (!second && h < g.Vh) {
            m.b2 = xch[0][o];
            m.c = xch[1][o];
            m.c1 = xch[2][o];
            su2_owner_finish(m, pmc, &e, &pl);
        }
    } else {
        if (h < g.Vh) {
            su2_owner_terms<2, LinkT>(U, g, h, &m);
            su2_owner_finish(m, pmc, &e, &pl);
        }
    }
    /* the owners' sums (a second half adds zeros to wavefront totals that
     * are not read) */
    su2_block_sum2(e, pl, SU2_EN_BLOCK);
    if (threadIdx.x == 0) {
        work[2 * b] = e;
        work[2 * b + 1] = pl;
    }
}

/* MEAS: the launch also measures the field it starts from
 * (su2_site_staple): *meas_e and *meas_p receive the thread's energy and
 * plaquette sums (su2_owner_finish). */
template <int MU, int BC, int UPD, typename LinkT, bool MEAS = false>
__device__ static inline __attribute__((always_inline)) void su2_sweep_site(LinkT *__restrict__ U,
                                                                            uint32_t *__restrict__ rng,
                                                                            int *__restrict__ accept,
                                                                            const su2_kgeom &g, const su2_kphys &ph,
                                                                            int parity, uint32_t h, int bc = BC,
                                                                            int acc_mode = SU2_K_ACC_LAUNCH,
                                                                            double *meas_e = nullptr,
                                                                            double *meas_p = nullptr)
{
    typedef su2_link_io<LinkT> io;
    static_assert(!MEAS || MU == SU2_DIR_Y, "a measuring launch is a y launch");
    su2_owner_plaq m = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    su2_owner_plaq *meas = MEAS ? &m : nullptr;

    /* the RNG state is loaded ahead of the staple, so that its latency
     * overlaps the link loads (FP32-link heat-bath at 128^3: -0.8 %,
     * alternating on one box; Metropolis unchanged) */
    uint32_t s[6];
    if (UPD != SU2_K_OR) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = rng[(size_t)i * g.Vh + h];
    }
    su2_q A;
    if (!su2_site_staple<MU, BC, LinkT, MEAS>(U, g, parity, h, &A, bc, meas)) {
        /* a link the boundary does not update (PCC x/t at y = 0): its
         * half-site's sweep count is the y link's update alone */
        if (acc_mode == SU2_K_ACC_SWEEP) atomicAdd(&accept[(size_t)parity * g.Vh + h], 1);
        return;
    }
    if constexpr (MEAS) {
        /* The two sums are formed here, before the update, and passed
         * through an empty asm statement that the compiler cannot move
         * past the Kennedy-Pendleton loop.  Left to itself it computes the
         * six plaquettes after the loop, where the sums are used, and keeps
         * their links live across it: FP64 links 214 against 166 VGPRs, two
         * waves per SIMD against three; the 128^3 step's measurement
         * 10.2-11.3 against 8.1-8.7 us, alternating on two boxes.  The
         * same values, the same bits. */
        const bool pmc = BC == SU2_K_BC_PMC || (BC == SU2_K_BC_ANY && bc == SU2_K_BC_PMC);
        double e, pl;
        su2_owner_finish(m, pmc, &e, &pl);
        asm volatile("" : "+v"(e), "+v"(pl));
        *meas_e = e;
        *meas_p = pl;
    }
    const size_t Vh = g.Vh;
    LinkT *dst = U + (size_t)(parity * 3 + MU) * g.Vh + h;
    int acc;
    if (UPD == SU2_K_OR) {
        su2_q u = io::load(dst);
        io::store(dst, su2_over_relax(A, u));
        acc = 1;
    } else {
        if (UPD == SU2_K_HB) {
            io::store(dst, su2_heat_bath(A, ph.beta, s));
            acc = 1;
        } else {
            su2_q u = io::load(dst);
            acc = su2_metropolis(A, &u, ph.beta, ph.eps, ph.hits, s);
            io::store(dst, u);
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) rng[(size_t)i * Vh + h] = s[i];
    }
    /* A no-return atomic add: the wavefront does not wait for a load of the
     * counter before it ends (each counter belongs to one thread per
     * launch, so the sum is the same; heat-bath -1.6 %, over-relaxation
     * -1 % to -3 % per launch). */
    if (acc_mode == SU2_K_ACC_LAUNCH)
        atomicAdd(&accept[(size_t)parity * Vh + h], acc);
    else if (acc_mode == SU2_K_ACC_SWEEP)
        atomicAdd(&accept[(size_t)parity * Vh + h], 3);
}

/*
 * Heat-bath with the rejection loop shared by the workgroup.  A wavefront
 * runs the Kennedy-Pendleton loop until its slowest lane accepts: with
 * about 2 % of the sites rejecting their first trial, nearly every
 * wavefront pays for a second trial (1.8 trials per wavefront against 1.02
 * per site, 128^3 at beta = 5).  Here every site runs its first trial in
 * place; the sites that reject it move their RNG state and scale into an
 * LDS queue, the first lanes of the workgroup run those sites' remaining
 * trials to acceptance, and each site takes its state and lambda back to
 * finish its own link.  A site's stream is consumed in the same order as
 * in su2_heat_bath, so the results are identical bit for bit.
 *
 * Every thread of the workgroup must call this the same number of times
 * (it holds three barriers): threads without a site pass in_range = 0.
 */
struct su2_kp_queue {
    uint32_t n;
    uint32_t s[6][256];
    float inv[256];
    float lf[256];
};

template <int MU, int BC, typename LinkT, bool EARLY_RNG = true, bool STREAM_RNG = false, bool MEAS = false>
__device__ static inline __attribute__((always_inline)) void
su2_sweep_site_hb_wg(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept,
                     const su2_kgeom &g, const su2_kphys &ph, int parity, uint32_t h, int in_range,
                     su2_kp_queue &q, int bc = BC, int acc_mode = SU2_K_ACC_LAUNCH, su2_owner_plaq *meas = nullptr)
{
    typedef su2_link_io<LinkT> io;

    /* EARLY_RNG: the stream's state is loaded before the staple, so that
     * its latency overlaps the neighbour loads; loaded after it, each
     * wavefront waits for one more memory round trip.  It keeps the six
     * words live across the staple.  Measured, alternating on one box: the
     * FP64-link single-lattice sweep (capped at 4 waves per SIMD anyway)
     * gains 2-4 % per launch at 128^3; the uncapped FP32-link one is within
     * 1 % either way and keeps the late load; the batched sweep gains 4.5 %
     * with FP32 links and loses 1.6-2 % with FP64 links.  The callers choose
     * accordingly.  (Issuing all twelve neighbour loads before the first
     * product gained nothing more and cost FP32-link over-relaxation 11 %.)
     *
     * STREAM_RNG: the stream's state is loaded and stored with nontemporal
     * hints, for lattices larger than the Infinity Cache, whose launches
     * run from DRAM: the state is read once and written once per launch,
     * and the hints keep it from displacing the neighbour links in the
     * caches.  Measured with FP64 links, alternating on one box: 256^3
     * heat-bath launch 410 -> 398 us; at 128^3 (resident in the Infinity
     * Cache, where the next launch re-reads the state from it) the same
     * hints cost 4.5 %, so they are not used there. */
    uint32_t s[6];
    if constexpr (EARLY_RNG) {
        if (in_range) {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                if constexpr (STREAM_RNG)
                    s[i] = __builtin_nontemporal_load(&rng[(size_t)i * g.Vh + h]);
                else
                    s[i] = rng[(size_t)i * g.Vh + h];
            }
        }
    }
    su2_q A = su2_q_zero();
    const int active = in_range ? su2_site_staple<MU, BC, LinkT, MEAS>(U, g, parity, h, &A, bc, meas) : 0;
    const size_t Vh = g.Vh;
    float inv = 0.0f, lf = 0.0f;
    int done = 1;
    if (active) {
        if constexpr (!EARLY_RNG) {
#pragma unroll
            for (int i = 0; i < 6; ++i) s[i] = rng[(size_t)i * Vh + h];
        }
        inv = su2_kp_setup(&A, ph.beta);
        done = su2_kp_trial(s, inv, &lf);
    }
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    uint32_t slot = 0;
    if (!done) {
        slot = atomicAdd(&q.n, 1u);
#pragma unroll
        for (int i = 0; i < 6; ++i) q.s[i][slot] = s[i];
        q.inv[slot] = inv;
    }
    __syncthreads();
    const uint32_t n = q.n;
    if (threadIdx.x < n) {
        uint32_t w[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) w[i] = q.s[i][threadIdx.x];
        const float iw = q.inv[threadIdx.x];
        float l;
        while (!su2_kp_trial(w, iw, &l)) {
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) q.s[i][threadIdx.x] = w[i];
        q.lf[threadIdx.x] = l;
    }
    __syncthreads();
    if (!done) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = q.s[i][slot];
        lf = q.lf[slot];
    }
    if (active) {
        LinkT *dst = U + (size_t)(parity * 3 + MU) * g.Vh + h;
        io::store(dst, su2_kp_finish(A, lf, s));
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if constexpr (STREAM_RNG)
                __builtin_nontemporal_store(s[i], &rng[(size_t)i * Vh + h]);
            else
                rng[(size_t)i * Vh + h] = s[i];
        }
        /* no-return adds, see su2_sweep_site */
        if (acc_mode == SU2_K_ACC_LAUNCH)
            atomicAdd(&accept[(size_t)parity * Vh + h], 1);
        else if (acc_mode == SU2_K_ACC_SWEEP)
            atomicAdd(&accept[(size_t)parity * Vh + h], 3);
    } else if (in_range && acc_mode == SU2_K_ACC_SWEEP) {
        atomicAdd(&accept[(size_t)parity * Vh + h], 1); /* PCC x/t at y = 0: the y link's update */
    }
}

#endif /* SU2_SWEEP_SITE_H */
