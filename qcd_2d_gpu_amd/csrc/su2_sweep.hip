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
 * round-robin to the 8 XCDs; XCD k is given the k-th contiguous eighth of
 * the lattice (a slab of t-planes).  With strip > 0 (the host sets it only
 * when every plane is a whole number of blocks, the slabs are whole planes
 * and the plane's bpp blocks split into strips of `strip` blocks), the XCD
 * walks its slab strip by strip, all its t-planes for one strip of x-rows
 * before the next: the t-neighbour rows of a strip are then re-read from
 * L2 a few hundred KB later instead of two full planes (2.6 MB at 128^3)
 * later.  Measured with FP64 links at 128^3: over-relaxation launch -10 %,
 * heat-bath -5 %; FP32 links (half the footprint) and planes of more than
 * 64 blocks (256^3, DRAM-bound) gain nothing from it and keep the plain
 * slab order. */
__device__ static inline __attribute__((always_inline)) uint32_t sweep_block(uint32_t strip, uint32_t bpp)
{
    const uint32_t b = blockIdx.x, xcd = b & 7u, j = b >> 3, per = gridDim.x >> 3;
    if (strip == 0u) return xcd * per + j;
    const uint32_t T = per / bpp; /* t-planes of this XCD's slab */
    const uint32_t s = j / (T * strip), r = j - s * (T * strip);
    const uint32_t tl = r / strip, c = r - tl * strip;
    return (xcd * T + tl) * bpp + s * strip + c;
}

template <int MU, int BC, int UPD, typename LinkT>
__device__ static inline __attribute__((always_inline)) void sweep_body(LinkT *__restrict__ U,
                                                                        uint32_t *__restrict__ rng,
                                                                        int *__restrict__ accept, const su2_kgeom &g,
                                                                        const su2_kphys &ph, int parity,
                                                                        uint32_t strip, uint32_t bpp)
{
    const uint32_t h = sweep_block(strip, bpp) * blockDim.x + threadIdx.x;
    if constexpr (UPD == SU2_K_HB) {
        __shared__ su2_kp_queue q;
        su2_sweep_site_hb_wg<MU, BC, LinkT>(U, rng, accept, g, ph, parity, h, h < g.Vh, q);
    } else {
        if (h >= g.Vh) return;
        su2_sweep_site<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, h);
    }
}

template <int MU, int BC, int UPD, typename LinkT>
__global__ __launch_bounds__(256) void k_sweep(LinkT *__restrict__ U, uint32_t *__restrict__ rng,
                                               int *__restrict__ accept, su2_kgeom g, su2_kphys ph, int parity,
                                               uint32_t strip, uint32_t bpp)
{
    sweep_body<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, strip, bpp);
}

/* FP64-link sweeps run at most 4 waves per SIMD.  The heat-bath with the
 * shared rejection loop needs only 74-108 VGPRs and over-relaxation 66, so
 * the compiler would allow 5-7 waves, but more workgroups in flight spread
 * each XCD's working set over more t-planes.  Measured on one box, 4 waves
 * against uncapped: heat-bath 128^3 -4 %, 256^3 -9 % (one plane of the five
 * link arrays is already larger than an XCD's 4 MB L2 there);
 * over-relaxation -6 % and -5 %.  FP32 links (half the footprint) are as
 * fast or faster uncapped. */
template <int MU, int BC, int UPD, typename LinkT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void
k_sweep_w4(LinkT *__restrict__ U, uint32_t *__restrict__ rng, int *__restrict__ accept, su2_kgeom g, su2_kphys ph,
           int parity, uint32_t strip, uint32_t bpp)
{
    sweep_body<MU, BC, UPD, LinkT>(U, rng, accept, g, ph, parity, strip, bpp);
}

template <int MU, int BC, int UPD, typename LinkT>
hipError_t launch_one(void *U, uint32_t *rng, int *accept, const su2_kgeom &g, const su2_kphys &p, int parity,
                      int block, hipStream_t st)
{
    const uint32_t nb = su2_grid_blocks(g.Vh, (uint32_t)block);
    /* strip order for FP64 links when the lattice tiles evenly */
    uint32_t strip = 0, bpp = g.plane / (uint32_t)block;
    const uint32_t w = 8;
    if (sizeof(LinkT) == sizeof(su2_l64) && g.plane % (uint32_t)block == 0 && bpp % w == 0 && bpp <= 64 &&
        (g.Vh / (uint32_t)block) % (8u * bpp) == 0 && nb * (uint32_t)block == g.Vh)
        strip = w;
    if constexpr (sizeof(LinkT) == sizeof(su2_l64))
        hipLaunchKernelGGL((k_sweep_w4<MU, BC, UPD, LinkT>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng, accept, g,
                           p, parity, strip, bpp);
    else
        hipLaunchKernelGGL((k_sweep<MU, BC, UPD, LinkT>), dim3(nb), dim3(block), 0, st, (LinkT *)U, rng, accept, g,
                           p, parity, strip, bpp);
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
