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
 * Strips are 16 blocks wide where the plane's block count allows it, else
 * 8: at 128^3 (32 blocks per plane) 16 against 8 gave heat-bath -4 %, step
 * -2 %; 4 and 32 were slower than 8, and at 256^3 8 was 5 % slower than
 * 16.
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

__device__ static inline __attribute__((always_inline)) uint32_t sweep_block(uint32_t order, uint32_t strip,
                                                                             uint32_t bpp)
{
    const uint32_t b = blockIdx.x, xcd = b & 7u, j = b >> 3, per = gridDim.x >> 3;
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

template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG = false>
__device__ static inline __attribute__((always_inline)) void sweep_body(LinkT *__restrict__ U,
                                                                        uint32_t *__restrict__ rng,
                                                                        int *__restrict__ accept, const su2_kgeom &g,
                                                                        const su2_kphys &ph, int parity,
                                                                        uint32_t order, uint32_t strip,
                                                                        uint32_t bpp, uint32_t nthr, int acc_mode)
{
    /* nthr = blockDim.x, passed as an argument: blockDim.x itself is read
     * from the dispatch packet with a vector load whose wait would sit in
     * front of every wavefront's first link and RNG loads */
    const uint32_t h = sweep_block(order, strip, bpp) * nthr + threadIdx.x;
    /* Heat-bath with FP64 links shares the rejection loop across the
     * workgroup (su2_sweep_site_hb_wg, early RNG load); with FP32 links each
     * lane runs its own loop (su2_sweep_site), measured 3 % faster per
     * launch at 128^3 (28.7 against 29.6 us) with FP64 links unchanged. */
    if constexpr (UPD == SU2_K_HB && sizeof(LinkT) == sizeof(su2_l64)) {
        __shared__ su2_kp_queue q;
        su2_sweep_site_hb_wg<MU, BC, LinkT, true, STREAM_RNG>(U, rng, accept, g, ph, parity, h, h < g.Vh, q, BC,
                                                             acc_mode);
    } else {
        if (h >= g.Vh) return;
        su2_sweep_site<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, h, BC, acc_mode);
    }
}

template <int MU, int BC, int UPD, typename LinkT>
__global__ __launch_bounds__(256) void k_sweep(LinkT *__restrict__ U, uint32_t *__restrict__ rng,
                                               int *__restrict__ accept, su2_kgeom g, su2_kphys ph, int parity,
                                               uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nthr,
                                               int acc_mode)
{
    sweep_body<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, order, strip, bpp, nthr, acc_mode);
}

/* FP64-link sweeps run at most 4 waves per SIMD.  More workgroups in flight
 * spread each XCD's working set over more t-planes.  Measured on one box
 * when the kernels needed few enough VGPRs for 6-7 waves, 4 waves against
 * uncapped: heat-bath 128^3 -4 %, 256^3 -9 % (one plane of the five link
 * arrays is already larger than an XCD's 4 MB L2 there); over-relaxation
 * -6 % and -5 %.  With the fused arithmetic (heat-bath 106 VGPRs,
 * over-relaxation 94) a cap of 5 (100 VGPRs, no spills) timed the same as
 * 4 at 128^3.  FP32 links (half the footprint) are as fast or faster
 * uncapped. */
template <int MU, int BC, int UPD, typename LinkT, bool STREAM_RNG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void
k_sweep_w4(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept, su2_kgeom g, su2_kphys ph,
           int parity, uint32_t order, uint32_t strip, uint32_t bpp, uint32_t nthr, int acc_mode)
{
    sweep_body<MU, BC, UPD, LinkT, STREAM_RNG>(U, rng, accept, g, ph, parity, order, strip, bpp, nthr, acc_mode);
}

/* Bytes of links and RNG state above which a lattice's launches run from
 * DRAM rather than from the 256 MB Infinity Cache (128^3 FP64: 226 MB). */
#define SU2_CACHE_RESIDENT_BYTES (256ull << 20)

template <int MU, int BC, int UPD, typename LinkT>
hipError_t launch_one(void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                      int block, int acc_mode, hipStream_t st)
{
    const uint32_t nb = su2_grid_blocks(g.Vh, (uint32_t)block);
    /* block order for FP64 links when the lattice tiles evenly (see
     * sweep_block); FP32 links keep the slab order */
    uint32_t order = ORDER_SLAB, strip = 0, bpp = g.plane / (uint32_t)block;
    const uint32_t w = (bpp % 16u == 0) ? 16u : 8u;
    const bool f64 = sizeof(LinkT) == sizeof(su2_l64);
    const bool whole = g.plane % (uint32_t)block == 0 && nb * (uint32_t)block == g.Vh;
    if (f64 && whole && g.Lt <= 32u && bpp % 8u == 0) {
        order = ORDER_COLUMNS;
    } else if (f64 && whole && bpp % w == 0 && bpp <= 128 && (g.Vh / (uint32_t)block) % (8u * bpp) == 0) {
        order = ORDER_STRIPS;
        strip = w;
    }
    /* heat-bath of a lattice beyond the Infinity Cache: streaming hints
     * on the RNG state (su2_sweep_site_hb_wg) */
    const bool dram = 6ull * g.Vh * sizeof(LinkT) + 24ull * g.Vh > SU2_CACHE_RESIDENT_BYTES;
    if constexpr (sizeof(LinkT) == sizeof(su2_l64) && UPD == SU2_K_HB) {
        if (dram)
            hipLaunchKernelGGL((k_sweep_w4<MU, BC, UPD, LinkT, true>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng,
                               accept, g, p, parity, order, strip, bpp, (uint32_t)block, acc_mode);
        else
            hipLaunchKernelGGL((k_sweep_w4<MU, BC, UPD, LinkT, false>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng,
                               accept, g, p, parity, order, strip, bpp, (uint32_t)block, acc_mode);
    } else if constexpr (sizeof(LinkT) == sizeof(su2_l64))
        hipLaunchKernelGGL((k_sweep_w4<MU, BC, UPD, LinkT, false>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng,
                           accept, g, p, parity, order, strip, bpp, (uint32_t)block, acc_mode);
    else
        hipLaunchKernelGGL((k_sweep<MU, BC, UPD, LinkT>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng, accept, g,
                           p, parity, order, strip, bpp, (uint32_t)block, acc_mode);
    return hipGetLastError();
}

template <int BC, int UPD, typename LinkT>
hipError_t launch_mu(int mu, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                     int block, int acc_mode, hipStream_t st)
{
    switch (mu) {
    case SU2_DIR_Y: return launch_one<SU2_DIR_Y, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_DIR_X: return launch_one<SU2_DIR_X, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_DIR_T: return launch_one<SU2_DIR_T, BC, UPD, LinkT>(U, rng, accept, g, p, parity, block, acc_mode, st);
    }
    return hipErrorInvalidValue;
}

template <int BC, typename LinkT>
hipError_t launch_upd(int mu, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p,
                      int parity, int block, int acc_mode, hipStream_t st)
{
    switch (upd) {
    case SU2_K_METRO: return launch_mu<BC, SU2_K_METRO, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_K_HB: return launch_mu<BC, SU2_K_HB, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_K_OR: return launch_mu<BC, SU2_K_OR, LinkT>(mu, U, rng, accept, g, p, parity, block, acc_mode, st);
    }
    return hipErrorInvalidValue;
}

template <typename LinkT>
hipError_t launch_bc(int mu, int bc, int upd, void *U, uint32_t *rng, int *accept, const su2_kgeom &g,
                     const su2_kphys &p, int parity, int block, int acc_mode, hipStream_t st)
{
    switch (bc) {
    case SU2_K_BC_PER:
        return launch_upd<SU2_K_BC_PER, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_K_BC_PCC:
        return launch_upd<SU2_K_BC_PCC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, st);
    case SU2_K_BC_PMC:
        return launch_upd<SU2_K_BC_PMC, LinkT>(mu, upd, U, rng, accept, g, p, parity, block, acc_mode, st);
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t su2_launch_sweep(int mu, int bc, int upd, int f32, void *U, uint32_t *rng, int *accept,
                            const su2_kgeom &g, const su2_kphys &p, int parity, int block, int acc_mode,
                            hipStream_t st)
{
    if (f32) return launch_bc<su2_l32>(mu, bc, upd, U, rng, accept, g, p, parity, block, acc_mode, st);
    return launch_bc<su2_l64>(mu, bc, upd, U, rng, accept, g, p, parity, block, acc_mode, st);
}
