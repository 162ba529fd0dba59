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

/* Half-site index of (t, x, y). */
__device__ static inline __attribute__((always_inline)) uint32_t su2_hidx(const su2_kgeom &g, uint32_t t, uint32_t x,
                                                                      uint32_t y)
{
    return (y >> 1) + x * g.Lyh + t * g.plane;
}

/* The staple A of U_mu at half-site h of parity p; returns 0 (and leaves A
 * alone) for a link the boundary condition does not update (PCC: x and t
 * links at y = 0).  BC = SU2_K_BC_ANY takes the boundary from bc. */
template <int MU, int BC, typename LinkT>
__device__ static inline __attribute__((always_inline)) int su2_site_staple(const LinkT *__restrict__ U,
                                                                            const su2_kgeom &g, int parity,
                                                                            uint32_t h, su2_q *Aout, int bc = BC)
{
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
        } else {
            const su2_q u_nu_pmu = io::load(Uo + NU * Vh + h_pmu);
            const su2_q u_mu_pnu = io::load(Uo + MU * Vh + h_pnu);
            const su2_q u_nu_n = io::load(Us + NU * Vh + h);
            const su2_q u_nu_pmu_mnu = io::load(Us + NU * Vh + h_pmu_mnu);
            const su2_q u_mu_mnu = io::load(Uo + MU * Vh + h_mnu);
            const su2_q u_nu_mnu = io::load(Uo + NU * Vh + h_mnu);
            up = su2_q_mul_nd(su2_q_mul_nd(u_nu_pmu, u_mu_pnu), u_nu_n);
            dn = su2_q_mul(su2_q_mul_dd(u_nu_pmu_mnu, u_mu_mnu), u_nu_mnu);
        }
        if (pmc && MU != SU2_DIR_Y && NU == SU2_DIR_Y) {
            /* A is never -0 here, so adding +0 leaves it bit-identical. */
            if (y == g.Ly - 1u) up = su2_q_zero();
            if (y == 0u) dn = su2_q_zero();
        }
        A = su2_q_add(A, up);
        A = su2_q_add(A, dn);
    }
    *Aout = A;
    return 1;
}

template <int MU, int BC, int UPD, typename LinkT>
__device__ static inline __attribute__((always_inline)) void su2_sweep_site(LinkT *__restrict__ U,
                                                                            uint32_t *__restrict__ rng,
                                                                            int *__restrict__ accept,
                                                                            const su2_kgeom &g, const su2_kphys &ph,
                                                                            int parity, uint32_t h, int bc = BC,
                                                                            int acc_mode = SU2_K_ACC_LAUNCH)
{
    typedef su2_link_io<LinkT> io;

    su2_q A;
    if (!su2_site_staple<MU, BC, LinkT>(U, g, parity, h, &A, bc)) {
        /* a link the boundary does not update (PCC x/t at y = 0): its
         * half-site's sweep count is the y link's update alone */
        if (acc_mode == SU2_K_ACC_SWEEP) atomicAdd(&accept[(size_t)parity * g.Vh + h], 1);
        return;
    }
    const size_t Vh = g.Vh;
    LinkT *dst = U + (size_t)(parity * 3 + MU) * g.Vh + h;
    int acc;
    if (UPD == SU2_K_OR) {
        su2_q u = io::load(dst);
        io::store(dst, su2_over_relax(A, u));
        acc = 1;
    } else {
        uint32_t s[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = rng[(size_t)i * Vh + h];
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

template <int MU, int BC, typename LinkT, bool EARLY_RNG = true, bool STREAM_RNG = false>
__device__ static inline __attribute__((always_inline)) void
su2_sweep_site_hb_wg(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept,
                     const su2_kgeom &g, const su2_kphys &ph, int parity, uint32_t h, int in_range,
                     su2_kp_queue &q, int bc = BC, int acc_mode = SU2_K_ACC_LAUNCH)
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
    const int active = in_range ? su2_site_staple<MU, BC, LinkT>(U, g, parity, h, &A, bc) : 0;
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
