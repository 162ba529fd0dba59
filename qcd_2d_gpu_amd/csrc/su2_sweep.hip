/*
 * su2_sweep.hip - the checkerboard link update: one launch per
 * (direction mu, parity p), as the reference's six MC_update_{even,odd}_
 * {y,x,t}[_pcc|_pmc] kernels (functions/mc_updates.h:129-465,
 * functions/mc_updates_pcc.h:4-214, functions/mc_updates_pmc.h:4-247).
 *
 * One thread updates U_mu at one half-site h of parity p.  Within the launch
 * no link that is written (U_mu at parity p) is read by another thread: the
 * staple reads U_mu only at parity 1-p and the other two directions at both
 * parities, so the launch is race-free without synchronisation.
 *
 * Staple of U_mu(n), summed over nu != mu in the reference's order (nu
 * ascending in y, x, t; the "positive" path before the "negative" one):
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
 *
 * Roofline: HBM.  Per heat-bath update the compulsory traffic is 5 link
 * reads (U_mu at 1-p and U_nu, U_rho at both parities, each array once),
 * one link write and 48 B of RNG state, i.e. 240 B in FP64 storage; the
 * twelve neighbour reads of a wavefront are contiguous runs that the other
 * wavefronts of the same slab re-read from L2.
 */
#include "su2_internal.h"
#include "su2_update.h"

namespace {

template <typename LinkT> struct link_io;

template <> struct link_io<su2_l64> {
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

template <> struct link_io<su2_l32> {
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
__device__ static inline __attribute__((always_inline)) uint32_t hidx(const su2_kgeom &g, uint32_t t, uint32_t x,
                                                                      uint32_t y)
{
    return (y >> 1) + x * g.Lyh + t * g.plane;
}

template <int MU, int BC, int UPD, typename LinkT>
__global__ __launch_bounds__(256) void k_sweep(LinkT *__restrict__ U, uint32_t *__restrict__ rng,
                                               int *__restrict__ accept, su2_kgeom g, su2_kphys ph, int parity)
{
    typedef link_io<LinkT> io;
    const uint32_t h = su2_xcd_block() * blockDim.x + threadIdx.x;
    if (h >= g.Vh) return;

    const uint32_t q = su2_div(h, g.div_lyh);
    const uint32_t yh = h - q * g.Lyh;
    const uint32_t t = su2_div(q, g.div_lx);
    const uint32_t x = q - t * g.Lx;
    const uint32_t y = 2u * yh + (((t + x) & 1u) ^ (uint32_t)parity);

    if (BC == SU2_K_BC_PCC && MU != SU2_DIR_Y && y == 0u) return;

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
    const uint32_t h_pmu = hidx(g, tmu, xmu, ymu);

    su2_q A = su2_q_zero();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int NU = (k == 0) ? (MU == SU2_DIR_Y ? SU2_DIR_X : SU2_DIR_Y) : (MU == SU2_DIR_T ? SU2_DIR_X : SU2_DIR_T);
        /* n + nu, n - nu, n + mu - nu */
        const uint32_t h_pnu = hidx(g, NU == SU2_DIR_T ? tp : t, NU == SU2_DIR_X ? xp : x, NU == SU2_DIR_Y ? yp : y);
        const uint32_t h_mnu = hidx(g, NU == SU2_DIR_T ? tm : t, NU == SU2_DIR_X ? xm : x, NU == SU2_DIR_Y ? ym : y);
        const uint32_t h_pmu_mnu =
            hidx(g, NU == SU2_DIR_T ? tm : tmu, NU == SU2_DIR_X ? xm : xmu, NU == SU2_DIR_Y ? ym : ymu);

        const su2_q u_nu_pmu = io::load(Uo + NU * Vh + h_pmu);
        const su2_q u_mu_pnu = io::load(Uo + MU * Vh + h_pnu);
        const su2_q u_nu_n = io::load(Us + NU * Vh + h);
        const su2_q u_nu_pmu_mnu = io::load(Us + NU * Vh + h_pmu_mnu);
        const su2_q u_mu_mnu = io::load(Uo + MU * Vh + h_mnu);
        const su2_q u_nu_mnu = io::load(Uo + NU * Vh + h_mnu);

        su2_q up = su2_q_mul_nd(su2_q_mul_nd(u_nu_pmu, u_mu_pnu), u_nu_n);
        su2_q dn = su2_q_mul(su2_q_mul_dd(u_nu_pmu_mnu, u_mu_mnu), u_nu_mnu);
        if (BC == SU2_K_BC_PMC && MU != SU2_DIR_Y && NU == SU2_DIR_Y) {
            /* A is never -0 here, so adding +0 leaves it bit-identical. */
            if (y == g.Ly - 1u) up = su2_q_zero();
            if (y == 0u) dn = su2_q_zero();
        }
        A = su2_q_add(A, up);
        A = su2_q_add(A, dn);
    }

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
    accept[(size_t)parity * Vh + h] += acc;
}

template <int MU, int BC, int UPD, typename LinkT>
hipError_t launch_one(void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                      int block, hipStream_t st)
{
    const uint32_t nb = su2_grid_blocks(g.Vh, (uint32_t)block);
    hipLaunchKernelGGL((k_sweep<MU, BC, UPD, LinkT>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng, accept, g, p,
                       parity);
    return hipGetLastError();
}

template <int BC, int UPD, typename LinkT>
hipError_t launch_mu(int mu, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                     int block, hipStream_t st)
{
    switch (mu) {
    case SU2_DIR_Y: return launch_one<SU2_DIR_Y, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, st);
    case SU2_DIR_X: return launch_one<SU2_DIR_X, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, st);
    case SU2_DIR_T: return launch_one<SU2_DIR_T, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, st);
    }
    return hipErrorInvalidValue;
}

template <int BC, typename LinkT>
hipError_t launch_upd(int mu, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p,
                      int parity, int block, hipStream_t st)
{
    switch (upd) {
    case SU2_K_METRO: return launch_mu<BC, SU2_K_METRO, LinkT>(mu, U, rng, accept, g, p, parity, block, st);
    case SU2_K_HB: return launch_mu<BC, SU2_K_HB, LinkT>(mu, U, rng, accept, g, p, parity, block, st);
    case SU2_K_OR: return launch_mu<BC, SU2_K_OR, LinkT>(mu, U, rng, accept, g, p, parity, block, st);
    }
    return hipErrorInvalidValue;
}

template <typename LinkT>
hipError_t launch_bc(int mu, int bc, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g,
                     const su2_kphys &p, int parity, int block, hipStream_t st)
{
    switch (bc) {
    case SU2_K_BC_PER: return launch_upd<SU2_K_BC_PER, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, st);
    case SU2_K_BC_PCC: return launch_upd<SU2_K_BC_PCC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, st);
    case SU2_K_BC_PMC: return launch_upd<SU2_K_BC_PMC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, st);
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t su2_launch_sweep(int mu, int bc, int upd, int f32, void *U, uint32_t *rng, int *accept,
                            const su2_kgeom &g, const su2_kphys &p, int parity, int block, hipStream_t st)
{
    if (f32) return launch_bc<su2_l32>(mu, bc, upd, U, rng, accept, g, p, parity, block, st);
    return launch_bc<su2_l64>(mu, bc, upd, U, rng, accept, g, p, parity, block, st);
}
