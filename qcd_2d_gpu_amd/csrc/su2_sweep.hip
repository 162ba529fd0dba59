/*
 * su2_sweep.hip - the checkerboard link update of one lattice: one launch
 * per (direction mu, parity p), as the reference's six
 * MC_update_{even,odd}_{y,x,t}[_pcc|_pmc] kernels (functions/mc_updates.h:
 * 129-465, functions/mc_updates_pcc.h:4-214, functions/mc_updates_pmc.h:
 * 4-247).
 *
 * One thread updates U_mu at one half-site h of parity p (su2_sweep_site.h).
 * Within the launch no link that is written (U_mu at parity p) is read by
 * another thread: the staple reads U_mu only at parity 1-p and the other two
 * directions at both parities, so the launch is race-free without
 * synchronisation.
 *
 * Roofline: HBM.  Per heat-bath update the compulsory traffic is 5 link
 * reads (U_mu at 1-p and U_nu, U_rho at both parities, each array once),
 * one link write and 48 B of RNG state, i.e. 240 B in FP64 storage; the
 * twelve neighbour reads of a wavefront are contiguous runs that the other
 * wavefronts of the same slab re-read from L2.
 */
#include "su2_sweep_site.h"

namespace {

/* MEAS (the y launch of parity 0, 256-thread workgroups): the launch also
 * measures the field it starts from and writes the en_av partial pair of its
 * block (su2_observe.h) to meas_work, in the layout of su2_launch_en_av. */
/* FIN: the launch has 8 more workgroups in front of the sweep's; the first
 * runs en_av pass 2 of the measuring launch before it (su2_en_final_small),
 * the other seven end at once.  In the sweep that replaces a launch of its
 * own: the kernel (4.7 us) and its dispatch gap, 8 us per 128^3 step. */
template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG = false, bool MEAS = false, bool FIN = false>
__device__ static inline __attribute__((always_inline)) void sweep_body(LinkT *__restrict__ U,
                                                                        uint32_t *__restrict__ rng,
                                                                        int *__restrict__ accept, const su2_kgeom &g,
                                                                        const su2_kphys &ph, int parity,
                                                                        uint32_t order, uint32_t strip,
                                                                        uint32_t bpp, uint32_t nthr, int acc_mode,
                                                                        double *__restrict__ meas_work,
                                                                        const su2_en_final_args &fin)
{
    /* nthr = blockDim.x, passed as an argument: blockDim.x itself is read
     * from the dispatch packet with a vector load whose wait would sit in
     * front of every wavefront's first link and RNG loads */
    uint32_t wg = blockIdx.x, nwg = gridDim.x;
    if constexpr (FIN) {
        if (wg < 8u) {
            if (wg == 0u) su2_en_final_small(fin.work, fin.n, fin.invV, fin.res, fin.res2);
            return;
        }
        wg -= 8u;
        nwg -= 8u;
    }
    const uint32_t lb = su2_block_at(wg, nwg, order, strip, bpp);
    const uint32_t h = lb * nthr + threadIdx.x;
    if constexpr (MEAS) {
        double e = 0.0, pl = 0.0;
        if constexpr (UPD == SU2_K_HB && STREAM_RNG) {
            __shared__ su2_kp_queue q;
            su2_owner_plaq m = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
            su2_sweep_site_hb_wg<MU, BC, LinkT, true, STREAM_RNG, true>(U, rng, accept, g, ph, parity, h, h < g.Vh,
                                                                       q, BC, acc_mode, &m);
            if (h < g.Vh) su2_owner_finish(m, BC == SU2_K_BC_PMC, &e, &pl);
        } else {
            if (h < g.Vh)
                su2_sweep_site<MU, BC, UPD, LinkT, true>(U, rng, accept, g, ph, parity, h, BC, acc_mode, &e, &pl);
        }
        su2_block_sum2(e, pl, SU2_EN_BLOCK);
        if (threadIdx.x == 0) {
            meas_work[2 * lb] = e;
            meas_work[2 * lb + 1] = pl;
        }
        return;
    }
    /* The heat-bath runs its Kennedy-Pendleton loop in each lane
     * (su2_sweep_site), except the streaming heat-bath of a lattice beyond
     * the Infinity Cache, which shares the rejection loop across the
     * workgroup (su2_sweep_site_hb_wg).  Measured with FP64 links,
     * alternating on one box, per-lane against shared: 16^3 and 32^3 -10 %,
     * 64^3 -5 %, 128^3 -1 %, 256^3 +1 %; FP32 links: 128^3 -3 % (28.7
     * against 29.6 us).  The wavefront repeats its trial for its slowest lane
     * (1.8 trials per wavefront against 1.02 per site at 128^3, beta = 5),
     * but saves the queue's three barriers and LDS round trips. */
    if constexpr (UPD == SU2_K_HB && STREAM_RNG) {
        __shared__ su2_kp_queue q;
        su2_sweep_site_hb_wg<MU, BC, LinkT, true, STREAM_RNG>(U, rng, accept, g, ph, parity, h, h < g.Vh, q, BC,
                                                             acc_mode);
    } else {
        if (h >= g.Vh) return;
        su2_sweep_site<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, h, BC, acc_mode);
    }
}

template <int MU, int BC, int UPD, typename LinkT, bool MEAS = false, bool FIN = false>
__global__ __launch_bounds__(256) void k_sweep(LinkT *__restrict__ U, uint32_t *__restrict__ rng,
                                               int *__restrict__ accept, su2_kgeom g, su2_kphys ph, int parity,
                                               uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nthr,
                                               int acc_mode, double *__restrict__ meas_work, su2_en_final_args fin)
{
    sweep_body<MU, BC, UPD, LinkT, false, MEAS, FIN>(U, rng, accept, g, ph, parity, order, strip, bpp, nthr,
                                                     acc_mode, meas_work, fin);
}

/* FP64-link sweeps run at most 4 waves per SIMD.  More workgroups in flight
 * spread each XCD's working set over more t-planes.  Measured on one box
 * when the kernels needed few enough VGPRs for 6-7 waves, 4 waves against
 * uncapped: heat-bath 128^3 -4 %, 256^3 -9 % (one plane of the five link
 * arrays is already larger than an XCD's 4 MB L2 there); over-relaxation
 * -6 % and -5 %.  With the fused arithmetic (heat-bath 106 VGPRs,
 * over-relaxation 94) a cap of 5 (100 VGPRs, no spills) timed the same as
 * 4 at 128^3.  FP32 links (half the footprint) are as fast or faster
 * uncapped.  Over-relaxation, and the heat-bath of a lattice beyond the
 * Infinity Cache, run at most 3 (k_sweep_w3): alternating on one box,
 * over-relaxation -1 % at 128^3 and -0.4 % at 256^3, heat-bath -1.5 % at
 * 256^3 but +5 % at 128^3 (so the cache-resident heat-bath keeps 4; 2 waves
 * took the 256^3 heat-bath +8 %). */
template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG, bool MEAS = false, bool FIN = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void
k_sweep_w4(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept, su2_kgeom g, su2_kphys ph,
           int parity, uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nthr, int acc_mode,
           double *__restrict__ meas_work, su2_en_final_args fin)
{
    sweep_body<MU, BC, UPD, LinkT, STREAM_RNG, MEAS, FIN>(U, rng, accept, g, ph, parity, order, strip, bpp, nthr,
                                                         acc_mode, meas_work, fin);
}

template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG, bool MEAS = false, bool FIN = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void
k_sweep_w3(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept, su2_kgeom g, su2_kphys ph,
           int parity, uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nthr, int acc_mode,
           double *__restrict__ meas_work, su2_en_final_args fin)
{
    sweep_body<MU, BC, UPD, LinkT, STREAM_RNG, MEAS, FIN>(U, rng, accept, g, ph, parity, order, strip, bpp, nthr,
                                                         acc_mode, meas_work, fin);
}

/* The kernel of one launch: FP32 links uncapped; FP64 links at most 3 waves
 * per SIMD for over-relaxation and the streaming heat-bath, else 4. */
template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG, bool MEAS, bool FIN>
hipError_t launch_kernel(dim3 G, dim3 B, hipStream_t st, LinkT *u, uint32_t *rng, int *accept, const su2_kgeom &g,
                         const su2_kphys &p, int parity, uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nt,
                         int acc_mode, double *meas_work, const su2_en_final_args &fin)
{
    if constexpr (sizeof(LinkT) != sizeof(su2_l64)) {
        static_assert(!STREAM_RNG, "streaming RNG hints are an FP64-link heat-bath form");
        hipLaunchKernelGGL((k_sweep<MU, BC, UPD, LinkT, MEAS, FIN>), G, B, 0, st, u, rng, accept, g, p, parity, order,
                           strip, bpp, nt, acc_mode, meas_work, fin);
    } else if constexpr (UPD == SU2_K_OR || STREAM_RNG) {
        hipLaunchKernelGGL((k_sweep_w3<MU, BC, UPD, LinkT, STREAM_RNG, MEAS, FIN>), G, B, 0, st, u, rng, accept, g, p,
                           parity, order, strip, bpp, nt, acc_mode, meas_work, fin);
    } else {
        hipLaunchKernelGGL((k_sweep_w4<MU, BC, UPD, LinkT, STREAM_RNG, MEAS, FIN>), G, B, 0, st, u, rng, accept, g, p,
                           parity, order, strip, bpp, nt, acc_mode, meas_work, fin);
    }
    return hipGetLastError();
}

/* Bytes of links and RNG state above which a lattice's launches run from
 * DRAM rather than from the 256 MB Infinity Cache (128^3 FP64: 226 MB). */
#define SU2_CACHE_RESIDENT_BYTES (256ull << 20)

template <int MU, int BC, int UPD, typename LinkT>
hipError_t launch_one(void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                      int block, int acc_mode, double *meas_work, const su2_en_final_args *fin, hipStream_t st)
{
    const uint32_t nb = su2_grid_blocks(g.Vh, (uint32_t)block);
    const bool f64 = sizeof(LinkT) == sizeof(su2_l64);
    /* block order for FP64 links when the lattice tiles evenly (see
     * su2_block_of); FP32 links keep the slab order */
    uint32_t order, strip, bpp;
    su2_pick_block_order(g, (uint32_t)block, f64, &order, &strip, &bpp);
    /* heat-bath of a lattice beyond the Infinity Cache: streaming hints
     * on the RNG state (su2_sweep_site_hb_wg) */
    const bool dram = 6ull * g.Vh * sizeof(LinkT) + 24ull * g.Vh > SU2_CACHE_RESIDENT_BYTES;
    const dim3 G(nb), B(block);
    const uint32_t nt = (uint32_t)block;
    LinkT *u = (LinkT *)U;
    const su2_en_final_args nofin = {nullptr, 0u, 0.0, nullptr, nullptr};
    /* the streaming kernel: heat-bath, FP64 links, beyond the Infinity Cache */
    constexpr bool srng_ok = sizeof(LinkT) == sizeof(su2_l64) && UPD == SU2_K_HB;
    if constexpr (MU == SU2_DIR_Y) {
        if (meas_work) {
            /* the measuring launch: the host guarantees parity 0 and
             * SU2_EN_BLOCK threads */
            if constexpr (srng_ok)
                if (dram)
                    return launch_kernel<MU, BC, UPD, LinkT, true, true, false>(G, B, st, u, rng, accept, g, p, parity,
                                                                               order, strip, bpp, nt, acc_mode,
                                                                               meas_work, nofin);
            return launch_kernel<MU, BC, UPD, LinkT, false, true, false>(G, B, st, u, rng, accept, g, p, parity, order,
                                                                        strip, bpp, nt, acc_mode, meas_work, nofin);
        }
        if (fin) {
            /* pass 2 of the measuring launch in 8 leading workgroups: the
             * host guarantees parity 1 and SU2_EN_BLOCK threads */
            const dim3 GF(nb + 8u);
            if constexpr (srng_ok)
                if (dram)
                    return launch_kernel<MU, BC, UPD, LinkT, true, false, true>(GF, B, st, u, rng, accept, g, p,
                                                                               parity, order, strip, bpp, nt,
                                                                               acc_mode, nullptr, *fin);
            return launch_kernel<MU, BC, UPD, LinkT, false, false, true>(GF, B, st, u, rng, accept, g, p, parity,
                                                                        order, strip, bpp, nt, acc_mode, nullptr,
                                                                        *fin);
        }
    }
    if constexpr (srng_ok)
        if (dram)
            return launch_kernel<MU, BC, UPD, LinkT, true, false, false>(G, B, st, u, rng, accept, g, p, parity, order,
                                                                        strip, bpp, nt, acc_mode, nullptr, nofin);
    return launch_kernel<MU, BC, UPD, LinkT, false, false, false>(G, B, st, u, rng, accept, g, p, parity, order, strip,
                                                                 bpp, nt, acc_mode, nullptr, nofin);
}

template <int BC, int UPD, typename LinkT>
hipError_t launch_mu(int mu, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                     int block, int acc_mode, double *mw, const su2_en_final_args *fin, hipStream_t st)
{
    switch (mu) {
    case SU2_DIR_Y: return launch_one<SU2_DIR_Y, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_DIR_X: return launch_one<SU2_DIR_X, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_DIR_T: return launch_one<SU2_DIR_T, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    }
    return hipErrorInvalidValue;
}

template <int BC, typename LinkT>
hipError_t launch_upd(int mu, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p,
                      int parity, int block, int acc_mode, double *mw, const su2_en_final_args *fin,
                      hipStream_t st)
{
    switch (upd) {
    case SU2_K_METRO: return launch_mu<BC, SU2_K_METRO, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_K_HB: return launch_mu<BC, SU2_K_HB, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_K_OR: return launch_mu<BC, SU2_K_OR, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    }
    return hipErrorInvalidValue;
}

template <typename LinkT>
hipError_t launch_bc(int mu, int bc, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g,
                     const su2_kphys &p, int parity, int block, int acc_mode, double *mw,
                     const su2_en_final_args *fin, hipStream_t st)
{
    switch (bc) {
    case SU2_K_BC_PER:
        return launch_upd<SU2_K_BC_PER, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_K_BC_PCC:
        return launch_upd<SU2_K_BC_PCC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    case SU2_K_BC_PMC:
        return launch_upd<SU2_K_BC_PMC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, mw, fin, st);
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t su2_launch_sweep(int mu, int bc, int upd, int f32, void *U, uint32_t *rng, int *accept,
                            const su2_kgeom &g, const su2_kphys &p, int parity, int block, int acc_mode,
                            hipStream_t st, double *meas_work, const su2_en_final_args *fin)
{
    if (meas_work && (mu != SU2_DIR_Y || parity != 0 || block != (int)SU2_EN_BLOCK)) return hipErrorInvalidValue;
    if (fin && (mu != SU2_DIR_Y || parity != 1 || block != (int)SU2_EN_BLOCK || !fin->work)) return hipErrorInvalidValue;
    if (f32) return launch_bc<su2_l32>(mu, bc, upd, U, rng, accept, g, p, parity, block, acc_mode, meas_work, fin, st);
    return launch_bc<su2_l64>(mu, bc, upd, U, rng, accept, g, p, parity, block, acc_mode, meas_work, fin, st);
}
