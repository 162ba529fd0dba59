/*
 * su2_internal.h - device data layout, geometry arguments and the launcher
 * interface shared by the HIP translation units of libsu2mc.
 *
 * Field layout in HBM (structure of arrays, checkerboarded):
 *     link(parity p, direction mu, half-site h) = U[(p * 3 + mu) * Vh + h]
 * with one link = 4 doubles (32 B, FP64 storage) or 4 floats (16 B, FP32
 * storage).  Directions are 0 = y, 1 = x, 2 = t as in the reference.
 * The half-site index is h = y' + (Ly/2) (x + Lx t) with y' = y >> 1, the
 * reference's even/odd numbering (functions/mc_updates.h:130-134): parity 0
 * holds the sites with (t + x + y) even, and within a parity the site at
 * (t, x, y') has y = 2 y' + ((t + x + p) & 1).  A wavefront of 64
 * consecutive h therefore reads 64 consecutive links of every array it
 * touches, for the site itself and for every neighbour.
 *
 * RNG state: six uint32 planes (g1[0..2], g2[0..2]) of Vh entries; stream h
 * is shared by the even and the odd site h (functions/mc_updates.h:179,235).
 * Acceptance counters: int accept[2 * Vh], indexed p * Vh + h
 * (functions/mc_updates.h:179,235).
 */
#ifndef SU2_INTERNAL_H
#define SU2_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Largest lattice, in sites.  Link indices (p * 3 + mu) * Vh + h < 3V are
 * 32-bit in the kernels, and every launch's thread count padded to whole
 * blocks (at most 3V + 2047) must stay below 2^32, the dispatch limit:
 * 3V <= 2^32 - 2^16.  (1.43 G sites = 160 GB of FP64 links, RNG state and
 * acceptance counters, in 288 GB of HBM.) */
#define SU2_MAX_SITES 1431633920ll /* = SU2_MAX_LATTICE_SITES of su2mc.h */

enum { SU2_DIR_Y = 0, SU2_DIR_X = 1, SU2_DIR_T = 2 };
enum { SU2_K_BC_PER = 0, SU2_K_BC_PCC = 1, SU2_K_BC_PMC = 2 };
/* Template value for a boundary condition chosen at run time (the batched
 * launches, whose virtual blocks belong to lattices of any boundary): the
 * site code then tests the bc argument instead of a constant. */
enum { SU2_K_BC_ANY = -1 };
enum { SU2_K_METRO = 0, SU2_K_HB = 1, SU2_K_OR = 2 };
/* Acceptance counting of a sweep launch (argument acc_mode):
 *   SU2_K_ACC_LAUNCH  each updated link adds its accepted count (the
 *                     reference's per-kernel accept[site] += ...);
 *   SU2_K_ACC_NONE    no counting;
 *   SU2_K_ACC_SWEEP   each half-site adds the whole sweep's count of its
 *                     parity: 3 updated links, 1 at PCC's y = 0 (only its y
 *                     link is updated).  For heat-bath and over-relaxation,
 *                     which accept every update, the six launches of a sweep
 *                     run (y, x) with NONE and t with SWEEP: after the sweep
 *                     every counter holds what per-launch counting gives, with
 *                     one read-modify-write per site instead of three. */
enum { SU2_K_ACC_LAUNCH = 0, SU2_K_ACC_NONE = 1, SU2_K_ACC_SWEEP = 2 };

struct __attribute__((aligned(32))) su2_l64 {
    double a0, a1, a2, a3;
};
struct __attribute__((aligned(16))) su2_l32 {
    float a0, a1, a2, a3;
};

/* n / d for n < 2^32 with M = ceil(2^64 / d): q = floor(n M / 2^64), exact
 * because n (M - 2^64/d) / 2^64 < 1/d.  d == 1 is handled separately. */
struct su2_fastdiv {
    uint32_t d, mlo, mhi, one;
};

static inline su2_fastdiv su2_make_fastdiv(uint32_t d)
{
    su2_fastdiv f;
    f.d = d;
    f.one = (d == 1u);
    if (d <= 1u) {
        f.mlo = f.mhi = 0u;
    } else {
        unsigned __int128 m = ((((unsigned __int128)1) << 64) + d - 1) / d;
        f.mlo = (uint32_t)m;
        f.mhi = (uint32_t)(m >> 32);
    }
    return f;
}

__device__ static inline __attribute__((always_inline)) uint32_t su2_div(uint32_t n, su2_fastdiv f)
{
    if (f.one) return n;
    uint64_t t = (uint64_t)n * f.mhi + __umulhi(n, f.mlo);
    return (uint32_t)(t >> 32);
}

/* Lattice geometry as the kernels see it (full extents, all even). */
struct su2_kgeom {
    uint32_t Lt, Lx, Ly;
    uint32_t Lyh;   /* Ly / 2: half-sites per x-row */
    uint32_t plane; /* Lx * Ly / 2: half-sites per t-plane */
    uint32_t Vh;    /* V / 2 */
    su2_fastdiv div_lyh, div_lx;
};

static inline su2_kgeom su2_make_kgeom(int Lt, int Lx, int Ly)
{
    su2_kgeom g;
    g.Lt = (uint32_t)Lt;
    g.Lx = (uint32_t)Lx;
    g.Ly = (uint32_t)Ly;
    g.Lyh = (uint32_t)(Ly / 2);
    g.plane = (uint32_t)Lx * g.Lyh;
    g.Vh = g.plane * (uint32_t)Lt;
    g.div_lyh = su2_make_fastdiv(g.Lyh);
    g.div_lx = su2_make_fastdiv(g.Lx);
    return g;
}

/* Smallest heat-bath coupling accepted on the host (= SU2_HB_BETA_MIN of
 * su2mc.h).  A Kennedy-Pendleton trial accepts with probability about
 * 1.26 k^1.5 at small k = beta |A| (|A| <= 4, the staple norm): some 800
 * trials per update at k = 0.01, 25 000 at 0.001, 10^9 at 1e-6, where a
 * workgroup would loop for minutes.  beta >= 0.01 keeps k small only where
 * the staple itself nearly vanishes. */
#define SU2_HB_BETA_MIN 0.01f

/* Physics parameters of one launch.  beta and eps arrive as the float values
 * of the reference's struct lattice (lattice_setup.h:10-11) widened to
 * double, as the reference's heat_bath(double beta) does. */
struct su2_kphys {
    double beta, eps;
    int hits;
};

/* Blocks are dealt round-robin to the 8 XCDs of the MI355X, each with its
 * own L2.  Remap so that XCD k gets the k-th contiguous eighth of the
 * lattice (a slab of t-planes) and neighbouring planes share an L2.  The
 * host pads the grid to a multiple of 8. */
__device__ static inline __attribute__((always_inline)) uint32_t su2_xcd_block(void)
{
    const uint32_t nb = gridDim.x;
    const uint32_t b = blockIdx.x;
    return (b & 7u) * (nb >> 3) + (b >> 3);
}

static inline uint32_t su2_grid_blocks(uint64_t nthreads, uint32_t block)
{
    uint64_t nb = (nthreads + block - 1) / block;
    nb = (nb + 7) & ~(uint64_t)7;
    return (uint32_t)nb;
}

/* en_av pass 2 carried by a sweep launch: the partials, their count (logical
 * blocks), 1 / V and the result slots it adds into. */
struct su2_en_final_args {
    const double *work;
    uint32_t n;
    double invV;
    double *res, *res2;
};

/* ---- launchers (host), defined in the .hip files ---- */
/* meas_work non-null: the launch (mu = y, parity 0, block = SU2_EN_BLOCK
 * threads, otherwise hipErrorInvalidValue) also writes the en_av partials of
 * the field it starts from there.  fin non-null: the launch (mu = y, parity
 * 1, block = SU2_EN_BLOCK) also runs en_av pass 2 over fin->work, in one
 * workgroup beside its own (the bits of su2_launch_en_av_final). */
hipError_t su2_launch_sweep(int mu, int bc, int upd, int f32, void *U, uint32_t *rng, int *accept,
                            const su2_kgeom &g, const su2_kphys &p, int parity, int block, int acc_mode,
                            hipStream_t st, double *meas_work = nullptr, const su2_en_final_args *fin = nullptr);
hipError_t su2_launch_setup_rng(uint32_t *rng, const su2_kgeom &g, uint64_t seed, hipStream_t st);
/* start: 0 reference hot start, 1 cold, 2 reference PCC hot start, 3 Haar,
 * 4 Haar with PCC identity links at y = 0 */
hipError_t su2_launch_init(int start, int f32, void *U, uint32_t *rng, const su2_kgeom &g, hipStream_t st);
hipError_t su2_launch_normalize(int f32, void *U, const su2_kgeom &g, hipStream_t st);
hipError_t su2_launch_en_av(int pmc, int f32, const void *U, const su2_kgeom &g, double *res, double *res2,
                            double *work, hipStream_t st);
size_t su2_en_av_work_doubles(const su2_kgeom &g);
/* pass 2 of en_av over partials written by a measuring sweep launch */
hipError_t su2_launch_en_av_final(const su2_kgeom &g, const double *work, double *res, double *res2, hipStream_t st);
/* en_av with its partial sums in the library's own static device array (no
 * allocation); lattices of at most SU2_EN_AV_NOALLOC_MAX_SITES sites (=
 * SU2_EN_AV_NOALLOC_MAX_SITES of su2mc.h). */
#define SU2_EN_AV_NOALLOC_MAX_SITES (1ll << 26)
hipError_t su2_launch_en_av_noalloc(int pmc, int f32, const void *U, const su2_kgeom &g, double *res, double *res2,
                                    hipStream_t st);
hipError_t su2_launch_to_ref(int f32, const void *U, double *phi_ref, const su2_kgeom &g, hipStream_t st);
hipError_t su2_launch_from_ref(int f32, const double *phi_ref, void *U, const su2_kgeom &g, hipStream_t st);
hipError_t su2_launch_fill_i32(int *p, int value, size_t n, hipStream_t st);
hipError_t su2_launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
hipError_t su2_launch_stream_probe(void *dst, const void *src, size_t bytes, int reads, int unroll, int nt,
                                   uint32_t blocks, hipStream_t st);
/* su2_phase_probe: synchronous; launches on st and waits for it;
 * hipErrorUnknown when an XCD-local barrier ran out of spins or read stale data */
hipError_t su2_phase_probe_run(double *buf, uint32_t blocks, int words, int phases, hipStream_t st, double *launch_us,
                               double *barrier_us, double *xcd_barrier_us);

/* su2_set_error (defined in su2_capi.cpp), shared with the C host driver */
#include "su2_host.h"

#endif /* SU2_INTERNAL_H */
