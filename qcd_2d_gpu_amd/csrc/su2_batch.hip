/*
 * su2_batch.hip - batched launches over many lattices of different shapes,
 * couplings and boundary conditions: one persistent launch per
 * (direction, parity) for all of them, plus batched reunitarisation and
 * energy/plaquette reduction.
 *
 * The work list lives in device memory (su2_bhead + a compact array of
 * su2_bdesc, one per resident lattice).  Every lattice is cut into virtual
 * blocks; block ranges are consecutive in the descriptor order, and a
 * launch's workgroups walk the virtual blocks with a grid stride, finding
 * their lattice by binary search on the range starts (compact arrays of
 * them, staged in LDS).  A virtual block never spans two lattices, so the
 * boundary condition and algorithm are uniform in a workgroup.  Because the
 * launch geometry does not depend on the work list, a captured hipGraph
 * stays valid when lattices join or leave: only the descriptor table is
 * rewritten between graph launches.
 *
 * Each lattice advances by its own schedule, the reference's run loop
 * (kernel.cu:161-208): with s = sweeps done before this tick,
 *   thermalisation  s < n_therm: normalise first if (s+1) % nevery == 0;
 *                   after the last one (s = n_therm-1) normalise once more;
 *   measurement     k = s - n_therm; after the last of n_cor sweeps of
 *                   measurement i = k / n_cor + 1: normalise if
 *                   i % nevery == 0, then en_av into slot i-1.
 * The per-site code is the single-lattice code (su2_sweep_site.h,
 * su2_observe.h) and the en_av partial layout and summation order are the
 * single-lattice ones, so every lattice gives the same bits alone, in a
 * batch, or in a served mix.
 */
#include "su2_batch.h"
#include "su2_observe.h"
#include "su2_sweep_site.h"

namespace {

/* The virtual blocks [*vb0, *vend) with stride *step that this workgroup
 * walks.  Workgroups are dispatched to the 8 XCDs round-robin by blockIdx,
 * so each XCD gets one contiguous eighth of the virtual blocks: tiles of a
 * lattice that share neighbour links meet in the same L2. */
__device__ static inline void xcd_walk_at(uint32_t wg, uint32_t G, uint32_t nblk, uint32_t *vb0, uint32_t *vend,
                                          uint32_t *step)
{
    if (G & 7u) {
        *vb0 = wg;
        *vend = nblk;
        *step = G;
        return;
    }
    const uint32_t per = (nblk + 7u) >> 3, x = wg & 7u;
    *vb0 = x * per + (wg >> 3);
    *vend = min(nblk, (x + 1u) * per);
    *step = G >> 3;
}

__device__ static inline void xcd_walk(uint32_t nblk, uint32_t *vb0, uint32_t *vend, uint32_t *step)
{
    xcd_walk_at(blockIdx.x, gridDim.x, nblk, vb0, vend, step);
}

/* Stages one launch kind's range starts in LDS when they fit (the search
 * then costs no dependent global loads); returns whether they did.  A lone
 * lattice needs no search (find_desc returns 0 without a load), so nothing
 * is staged for it and the workgroup skips the load and the barrier (n is
 * uniform over the workgroup). */
__device__ static inline bool stage_begins(const uint32_t *__restrict__ Bk, uint32_t n, uint32_t *s_begin)
{
    if (n <= 1u) return false;
    const bool lds = n <= SU2_BATCH_LDS_DESC;
    if (lds)
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_begin[i] = Bk[i];
    __syncthreads();
    return lds;
}

/* Index of the descriptor whose range [begin, next begin) holds vb.  vb is
 * uniform in the workgroup, so the result is too; readfirstlane says so to
 * the compiler, which then reads the descriptor with scalar loads into
 * SGPRs instead of one copy per lane in VGPRs. */
__device__ static inline uint32_t find_desc(const uint32_t *__restrict__ Bk, const uint32_t *s_begin, bool lds,
                                            uint32_t n, uint32_t vb)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t b = lds ? s_begin[mid] : Bk[mid];
        if (b <= vb)
            lo = mid;
        else
            hi = mid;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

__device__ static inline bool measure_now(const su2_bdesc &d, uint32_t s)
{
    if (s < (uint32_t)d.n_therm) return false;
    const uint32_t k = s - (uint32_t)d.n_therm;
    return (k % (uint32_t)d.n_cor) == (uint32_t)d.n_cor - 1u;
}

/* The measurement of sweep s is deferred to the lattice's next tick
 * (deferred: its final pass sets the lattice's pending flag) when the sweep
 * launches run workgroups of SU2_EN_BLOCK threads (carry_on) and the
 * lattice has a next tick.  There the first update launch forms the en_av
 * partials of the field it starts from into the carried partials (partc),
 * or, if that tick's launches do not run SU2_EN_BLOCK threads, an en_av
 * pass before the sweep does; no launch changes the field in between,
 * since a reunitarisation before the sweep happens only during
 * thermalisation.  The tick's final pass adds them and clears the flag.
 * Every other measurement runs as en_av at the end of its tick. */
__device__ static inline uint32_t total_sweeps(const su2_bdesc &d)
{
    return (uint32_t)d.n_therm + (uint32_t)d.n_meas * (uint32_t)d.n_cor;
}

__device__ static inline bool deferred(const su2_bdesc &d, uint32_t s, int carry_on)
{
    return carry_on && s + 1u < total_sweeps(d);
}

__device__ static inline bool normalize_pre(const su2_bdesc &d, uint32_t s)
{
    return s < (uint32_t)d.n_therm && ((s + 1u) % (uint32_t)d.nevery) == 0u;
}

__device__ static inline bool normalize_post(const su2_bdesc &d, uint32_t s)
{
    if (s + 1u == (uint32_t)d.n_therm) return true;
    if (!measure_now(d, s)) return false;
    const uint32_t i = (s - (uint32_t)d.n_therm) / (uint32_t)d.n_cor + 1u;
    return (i % (uint32_t)d.nevery) == 0u;
}

/* Acceptance of heat-bath and over-relaxation, which accept every update,
 * is counted once per sweep, in the t launches (SU2_K_ACC_SWEEP); Metropolis
 * counts in every launch. */
template <int MU>
__device__ static inline constexpr int accept_once_per_sweep()
{
    return MU == SU2_DIR_T ? SU2_K_ACC_SWEEP : SU2_K_ACC_NONE;
}

/* One link update of lattice d at half-site h (in_range = 0 past the end of
 * the lattice); the algorithm is uniform over a virtual block.  The
 * heat-bath runs its Kennedy-Pendleton loop in each lane, as the
 * single-lattice heat-bath of a cache-resident lattice does: measured
 * against the workgroup-shared rejection loop, alternating on one box, the
 * 64-lattice ensemble -5 %, a lone job's tick -8 to -11 % at 32^3-64^3,
 * the served p50 -10 %; the results are the same bits. */
template <int MU, int ALGO, typename LinkT>
__device__ static inline __attribute__((always_inline)) void site_link(const su2_bdesc &d, int parity, uint32_t h,
                                                                       int in_range)
{
    if (!in_range) return;
    if (ALGO == SU2_BALGO_HB || (ALGO == SU2_BALGO_MIXED && d.algo == 1))
        su2_sweep_site<MU, SU2_K_BC_ANY, SU2_K_HB, LinkT>((LinkT *)d.U, d.rng, d.accept, d.g, d.ph, parity, h, d.bc,
                                                         accept_once_per_sweep<MU>());
    else
        su2_sweep_site<MU, SU2_K_BC_ANY, SU2_K_METRO, LinkT>((LinkT *)d.U, d.rng, d.accept, d.g, d.ph, parity, h,
                                                            d.bc);
}

/* site_link that also measures the field it starts from (MU = y, parity 0,
 * workgroups of SU2_EN_BLOCK threads): logical en_av block lb of the lattice
 * goes to its carried partials, the bits su2_en_av_owner_block writes. */
template <int MU, int ALGO, typename LinkT>
__device__ static inline __attribute__((always_inline)) void site_link_measured(const su2_bdesc &d, int parity,
                                                                                uint32_t h, int in_range, uint32_t lb)
{
    double e = 0.0, pl = 0.0;
    if (in_range) {
        if (ALGO == SU2_BALGO_HB || (ALGO == SU2_BALGO_MIXED && d.algo == 1))
            su2_sweep_site<MU, SU2_K_BC_ANY, SU2_K_HB, LinkT, true>((LinkT *)d.U, d.rng, d.accept, d.g, d.ph, parity,
                                                                   h, d.bc, accept_once_per_sweep<MU>(), &e, &pl);
        else
            su2_sweep_site<MU, SU2_K_BC_ANY, SU2_K_METRO, LinkT, true>((LinkT *)d.U, d.rng, d.accept, d.g, d.ph,
                                                                      parity, h, d.bc, SU2_K_ACC_LAUNCH, &e, &pl);
    }
    su2_block_sum2(e, pl, SU2_EN_BLOCK);
    if (threadIdx.x == 0) {
        d.partc[2 * lb] = e;
        d.partc[2 * lb + 1] = pl;
    }
    __syncthreads(); /* su2_block_sum2's shared array is reused by the next block */
}

/* The end of a tick for lattice d by a workgroup of SU2_EN_BLOCK threads
 * (the leading workgroups of a tick's last sweep launch): en_av pass 2 of
 * its pending measurement, whose partials this tick's first update launch
 * formed, with the sums and order of k_ben_av_final's 1024-thread pass
 * (su2_en_final_small), then its sweep counter and pending flag.  The host
 * ends a tick this way only when no lattice measures or reunitarises after
 * the sweep (no end-of-tick en_av or normalise launch follows it), so the
 * end-of-tick measurement of k_ben_av_final never arises here.  No sweep
 * launch reads the counters or, after the first, the pending flags. */
__device__ static inline void tick_end_small(const su2_bdesc &d, uint32_t *__restrict__ ctr,
                                             uint32_t *__restrict__ pend, int carry_on)
{
    const uint32_t s = ctr[d.slot], p = pend[d.slot];
    __syncthreads(); /* every wave has read them before lane 0 sets them */
    if (p) {
        const uint32_t m = (s - 1u - (uint32_t)d.n_therm) / (uint32_t)d.n_cor;
        su2_en_final_small(d.partc, d.en_nb, 1.0 / (2.0 * (double)d.g.Vh), d.resE + m, d.resP + m);
    }
    const bool later = measure_now(d, s) && deferred(d, s, carry_on);
    if (threadIdx.x == 0) {
        ctr[d.slot] = s + 1u;
        pend[d.slot] = later ? 1u : 0u;
    }
    __syncthreads(); /* su2_en_final_small's shared array is reused by the next call */
}

/* OR = 0: link update; ALGO is the algorithm class of the resident lattices
 * (all heat-bath, all Metropolis, or mixed), so the common homogeneous case
 * carries one update body.  OR = 1: over-relaxation pass number or_pass
 * (lattices with fewer over-relaxation sweeps per tick skip it). */
/* END: the tick's last sweep launch also ends the tick: its first `lead`
 * workgroups (a multiple of 8, so the others keep their XCD) run
 * tick_end_small for the lattices j = wg, wg + lead, ...; the workgroups
 * after them sweep as usual. */
template <int MU, int OR, int ALGO, typename LinkT, bool MEAS = false, bool END = false>
__global__ __launch_bounds__(256) void k_bsweep(const su2_bhead *__restrict__ H, const su2_bdesc *__restrict__ D,
                                                const uint32_t *__restrict__ B, int parity, int or_pass,
                                                const uint32_t *__restrict__ pend, uint32_t *__restrict__ ctr,
                                                uint32_t *__restrict__ pend_end, uint32_t lead, int carry_on)
{
    __shared__ uint32_t s_begin[SU2_BATCH_LDS_DESC];
    const uint32_t n = H->n_active;
    uint32_t wg = blockIdx.x, G = gridDim.x;
    if constexpr (END) {
        if (wg < lead) {
            for (uint32_t j = wg; j < n; j += lead) tick_end_small(D[j], ctr, pend_end, carry_on);
            return;
        }
        wg -= lead;
        G -= lead;
    }
    uint32_t vb0, vend, step;
    xcd_walk_at(wg, G, H->sw_blocks, &vb0, &vend, &step);
    if (vb0 >= vend) return;
    const bool lds = stage_begins(B, n, s_begin);
    if constexpr (OR) {
        for (uint32_t vb = vb0; vb < vend; vb += step) {
            const su2_bdesc &d = D[find_desc(B, s_begin, lds, n, vb)];
            if (or_pass >= d.or_sweeps) continue;
            const uint32_t h = (vb - d.sw_begin) * blockDim.x + threadIdx.x;
            if (h >= d.g.Vh) continue;
            /* one body for every boundary: the boundary is a run-time test
             * (uniform in the workgroup), which keeps the register count of
             * a single update instead of the largest of three inlined ones */
            su2_sweep_site<MU, SU2_K_BC_ANY, SU2_K_OR, LinkT>((LinkT *)d.U, d.rng, d.accept, d.g, d.ph, parity, h,
                                                             d.bc, accept_once_per_sweep<MU>());
        }
    } else {
        for (uint32_t vb = vb0; vb < vend; vb += step) {
            /* everything up to the update is uniform in the workgroup */
            const su2_bdesc &d = D[find_desc(B, s_begin, lds, n, vb)];
            const uint32_t lb = vb - d.sw_begin;
            const uint32_t h = lb * blockDim.x + threadIdx.x;
            const int in_range = h < d.g.Vh;
            if constexpr (MEAS) {
                if (pend[d.slot]) {
                    site_link_measured<MU, ALGO, LinkT>(d, parity, h, in_range, lb);
                    continue;
                }
            }
            site_link<MU, ALGO, LinkT>(d, parity, h, in_range);
        }
    }
}

/* Reunitarisation of the lattices whose schedule asks for it this tick
 * (post = 0: before the sweep, post = 1: after it); the post pass at the end
 * of thermalisation also clears the lattice's acceptance counters. */
template <typename LinkT>
__global__ __launch_bounds__(256) void k_bnormalize(const su2_bhead *__restrict__ H, const su2_bdesc *__restrict__ D,
                                                    const uint32_t *__restrict__ B, uint32_t max_desc,
                                                    const uint32_t *__restrict__ ctr, int post)
{
    __shared__ uint32_t s_begin[SU2_BATCH_LDS_DESC];
    const uint32_t n = H->n_active;
    uint32_t vb0, vend, step;
    xcd_walk(H->nm_blocks, &vb0, &vend, &step);
    if (vb0 >= vend) return;
    const uint32_t *Bk = B + 2u * max_desc;
    const bool lds = stage_begins(Bk, n, s_begin);
    for (uint32_t vb = vb0; vb < vend; vb += step) {
        const su2_bdesc &d = D[find_desc(Bk, s_begin, lds, n, vb)];
        const uint32_t s = ctr[d.slot];
        if (!(post ? normalize_post(d, s) : normalize_pre(d, s))) continue;
        const uint32_t i = (vb - d.nm_begin) * blockDim.x + threadIdx.x;
        /* acceptance restarts with the measurement (kernel.cu:190) */
        if (post && s + 1u == (uint32_t)d.n_therm && i < 2u * d.g.Vh) d.accept[i] = 0;
        if (i >= 6u * d.g.Vh) continue;
        LinkT *U = (LinkT *)d.U + i;
        su2_link_io<LinkT>::store(U, su2_q_normalize(su2_link_io<LinkT>::load(U)));
    }
}

/* en_av pass 1; logical block b of a lattice has the single-lattice layout
 * and bits (su2_en_av_owner_block).  pre = 0: at the end of the tick, for
 * the lattices that measure there, into part; pre = 1: before the sweeps,
 * for the lattices with a pending measurement, into partc. */
template <typename LinkT>
__device__ static inline __attribute__((always_inline)) void ben_av_partial(const su2_bhead *__restrict__ H,
                                                                            const su2_bdesc *__restrict__ D,
                                                                            const uint32_t *__restrict__ B,
                                                                            uint32_t max_desc,
                                                                            const uint32_t *__restrict__ ctr,
                                                                            const uint32_t *__restrict__ pend,
                                                                            int carry_on, int pre)
{
    __shared__ uint32_t s_begin[SU2_BATCH_LDS_DESC];
    const uint32_t n = H->n_active;
    uint32_t vb0, vend, step;
    xcd_walk(H->en_blocks, &vb0, &vend, &step);
    if (vb0 >= vend) return;
    const uint32_t *Bk = B + max_desc;
    const bool lds = stage_begins(Bk, n, s_begin);
    for (uint32_t vb = vb0; vb < vend; vb += step) {
        const su2_bdesc &d = D[find_desc(Bk, s_begin, lds, n, vb)];
        if (pre) {
            if (!pend[d.slot]) continue;
        } else {
            const uint32_t s = ctr[d.slot];
            if (!measure_now(d, s) || deferred(d, s, carry_on)) continue;
        }
        su2_en_av_owner_block<LinkT>((const LinkT *)d.U, d.g, d.bc == SU2_K_BC_PMC, vb - d.en_begin,
                                     pre ? d.partc : d.part);
        __syncthreads(); /* su2_block_sum2's shared array is reused by the next block */
    }
}

/* FP64 links, two threads per owner: held to 80 VGPRs (6 waves per SIMD, 3
 * workgroups per CU instead of 2 at 85; no spill), 2 % off the 64-lattice
 * ensemble.  FP32 links, one thread per owner, would spill under that cap. */
__global__ __launch_bounds__(2 * SU2_EN_BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_ben_av_partial64(
    const su2_bhead *__restrict__ H, const su2_bdesc *__restrict__ D, const uint32_t *__restrict__ B,
    uint32_t max_desc, const uint32_t *__restrict__ ctr, const uint32_t *__restrict__ pend, int carry_on, int pre)
{
    ben_av_partial<su2_l64>(H, D, B, max_desc, ctr, pend, carry_on, pre);
}

__global__ __launch_bounds__(SU2_EN_BLOCK) void k_ben_av_partial32(const su2_bhead *__restrict__ H,
                                                                   const su2_bdesc *__restrict__ D,
                                                                   const uint32_t *__restrict__ B, uint32_t max_desc,
                                                                   const uint32_t *__restrict__ ctr,
                                                                   const uint32_t *__restrict__ pend, int carry_on,
                                                                   int pre)
{
    ben_av_partial<su2_l32>(H, D, B, max_desc, ctr, pend, carry_on, pre);
}

/* Adds one measurement of lattice d, pass 2 over its partials part, into
 * measurement slot m: the single-lattice order and bits. */
__device__ static inline void add_measurement(const su2_bdesc &d, const double *part, uint32_t m)
{
    double te, tp;
    su2_en_final_partials(part, d.en_nb, &te, &tp);
    su2_block_sum2(te, tp, SU2_EN_FINAL_BLOCK);
    if (threadIdx.x == 0) {
        const double invV = 1.0 / (2.0 * (double)d.g.Vh);
        d.resE[m] += te * invV;
        d.resP[m] += tp * invV;
    }
    __syncthreads(); /* su2_block_sum2's shared array is reused by the next call */
}

/* en_av pass 2: workgroup j (1024 threads) adds lattice j's pending
 * measurement (of the sweep before this tick, from its carried partials)
 * and its end-of-tick one, each in the single-lattice order, into their
 * measurement slots, or defers the end-of-tick one to the next tick.  It
 * is the last launch of a measuring tick, so it also ends the tick: every
 * lattice's sweep counter and pending flag are read once and then set. */
__global__ __launch_bounds__(1024) void k_ben_av_final(const su2_bhead *__restrict__ H,
                                                       const su2_bdesc *__restrict__ D,
                                                       uint32_t *__restrict__ ctr, uint32_t *__restrict__ pend,
                                                       int carry_on)
{
    const uint32_t n = H->n_active;
    for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
        const su2_bdesc &d = D[j];
        const uint32_t s = ctr[d.slot], p = pend[d.slot];
        __syncthreads(); /* every wave has read them before lane 0 sets them */
        if (p) add_measurement(d, d.partc, (s - 1u - (uint32_t)d.n_therm) / (uint32_t)d.n_cor);
        const bool now = measure_now(d, s), later = now && deferred(d, s, carry_on);
        if (now && !later) add_measurement(d, d.part, (s - (uint32_t)d.n_therm) / (uint32_t)d.n_cor);
        if (threadIdx.x == 0) {
            ctr[d.slot] = s + 1u;
            pend[d.slot] = later ? 1u : 0u;
        }
    }
}

/* End of a tick without measurement: every resident lattice has done one
 * more sweep. */
__global__ __launch_bounds__(256) void k_btick_end(const su2_bhead *__restrict__ H, const su2_bdesc *__restrict__ D,
                                                   uint32_t *__restrict__ ctr)
{
    const uint32_t n = H->n_active;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) ctr[D[j].slot] += 1u;
}

template <int MU, int OR, int ALGO, typename LinkT>
hipError_t launch_bsweep(const su2_bhead *H, const su2_bdesc *D, const uint32_t *B, int parity, int or_pass,
                         uint32_t grid, uint32_t block, hipStream_t st, const uint32_t *pend, int measure,
                         const su2_btick_end *end)
{
    if constexpr (MU == SU2_DIR_Y && !OR) {
        if (measure) {
            hipLaunchKernelGGL((k_bsweep<MU, OR, ALGO, LinkT, true>), dim3(grid), dim3(block), 0, st, H, D, B, parity,
                               or_pass, pend, nullptr, nullptr, 0u, 0);
            return hipGetLastError();
        }
    }
    if constexpr (MU == SU2_DIR_T) {
        if (end) {
            hipLaunchKernelGGL((k_bsweep<MU, OR, ALGO, LinkT, false, true>), dim3(grid + end->lead), dim3(block), 0,
                               st, H, D, B, parity, or_pass, pend, end->ctr, end->pend, end->lead, end->carry_on);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_bsweep<MU, OR, ALGO, LinkT>), dim3(grid), dim3(block), 0, st, H, D, B, parity, or_pass,
                       pend, nullptr, nullptr, 0u, 0);
    return hipGetLastError();
}

template <int OR, int ALGO, typename LinkT>
hipError_t launch_bsweep_mu(int mu, const su2_bhead *H, const su2_bdesc *D, const uint32_t *B, int parity,
                            int or_pass, uint32_t grid, uint32_t block, hipStream_t st, const uint32_t *pend,
                            int measure, const su2_btick_end *end)
{
    switch (mu) {
    case 0: return launch_bsweep<0, OR, ALGO, LinkT>(H, D, B, parity, or_pass, grid, block, st, pend, measure, nullptr);
    case 1: return launch_bsweep<1, OR, ALGO, LinkT>(H, D, B, parity, or_pass, grid, block, st, pend, 0, nullptr);
    default: return launch_bsweep<2, OR, ALGO, LinkT>(H, D, B, parity, or_pass, grid, block, st, pend, 0, end);
    }
}

template <typename LinkT>
hipError_t launch_bsweep_t(int mu, int over_relax, int algo_class, const su2_bhead *H, const su2_bdesc *D,
                           const uint32_t *B, int parity, int or_pass, uint32_t grid, uint32_t block, hipStream_t st,
                           const uint32_t *pend, int measure, const su2_btick_end *end)
{
    if (over_relax)
        return launch_bsweep_mu<1, SU2_BALGO_MIXED, LinkT>(mu, H, D, B, parity, or_pass, grid, block, st, pend, 0,
                                                          end);
    switch (algo_class) {
    case SU2_BALGO_HB:
        return launch_bsweep_mu<0, SU2_BALGO_HB, LinkT>(mu, H, D, B, parity, 0, grid, block, st, pend, measure, end);
    case SU2_BALGO_METRO:
        return launch_bsweep_mu<0, SU2_BALGO_METRO, LinkT>(mu, H, D, B, parity, 0, grid, block, st, pend, measure,
                                                          end);
    default:
        return launch_bsweep_mu<0, SU2_BALGO_MIXED, LinkT>(mu, H, D, B, parity, 0, grid, block, st, pend, measure,
                                                          end);
    }
}

} // namespace

hipError_t su2_launch_bsweep(int f32, int mu, int over_relax, int algo_class, const su2_bhead *H,
                             const su2_bdesc *D, const uint32_t *B, uint32_t max_desc, int parity, int or_pass,
                             uint32_t grid, uint32_t block, hipStream_t st, const uint32_t *pend, int measure,
                             const su2_btick_end *end)
{
    (void)max_desc; /* the sweep range starts are the first array of B */
    if (measure && (mu != SU2_DIR_Y || parity != 0 || over_relax || block != SU2_EN_BLOCK || !pend))
        return hipErrorInvalidValue;
    if (end && (mu != SU2_DIR_T || parity != 1 || block != SU2_EN_BLOCK || !end->ctr || !end->pend ||
                end->lead == 0 || (end->lead & 7u)))
        return hipErrorInvalidValue;
    if (f32)
        return launch_bsweep_t<su2_l32>(mu, over_relax, algo_class, H, D, B, parity, or_pass, grid, block, st, pend,
                                        measure, end);
    return launch_bsweep_t<su2_l64>(mu, over_relax, algo_class, H, D, B, parity, or_pass, grid, block, st, pend,
                                    measure, end);
}

hipError_t su2_launch_bnormalize(int f32, const su2_bhead *H, const su2_bdesc *D, const uint32_t *B,
                                 uint32_t max_desc, const uint32_t *ctr, int post, uint32_t grid, hipStream_t st)
{
    if (f32)
        hipLaunchKernelGGL(k_bnormalize<su2_l32>, dim3(grid), dim3(256), 0, st, H, D, B, max_desc, ctr, post);
    else
        hipLaunchKernelGGL(k_bnormalize<su2_l64>, dim3(grid), dim3(256), 0, st, H, D, B, max_desc, ctr, post);
    return hipGetLastError();
}

hipError_t su2_launch_ben_av_partial(int f32, const su2_bhead *H, const su2_bdesc *D, const uint32_t *B,
                                     uint32_t max_desc, const uint32_t *ctr, const uint32_t *pend, uint32_t grid,
                                     int carry_on, int pre, hipStream_t st)
{
    if (f32)
        hipLaunchKernelGGL(k_ben_av_partial32, dim3(grid), dim3(su2_en_av_threads(1)), 0, st, H, D, B, max_desc, ctr,
                           pend, carry_on, pre);
    else
        hipLaunchKernelGGL(k_ben_av_partial64, dim3(grid), dim3(su2_en_av_threads(0)), 0, st, H, D, B, max_desc, ctr,
                           pend, carry_on, pre);
    return hipGetLastError();
}

hipError_t su2_launch_ben_av_final(const su2_bhead *H, const su2_bdesc *D, uint32_t *ctr, uint32_t *pend,
                                   uint32_t final_grid, int carry_on, hipStream_t st)
{
    hipLaunchKernelGGL(k_ben_av_final, dim3(final_grid), dim3(SU2_EN_FINAL_BLOCK), 0, st, H, D, ctr, pend, carry_on);
    return hipGetLastError();
}

hipError_t su2_launch_btick_end(const su2_bhead *H, const su2_bdesc *D, uint32_t *ctr, hipStream_t st)
{
    hipLaunchKernelGGL(k_btick_end, dim3(1), dim3(256), 0, st, H, D, ctr);
    return hipGetLastError();
}

uint32_t su2_en_av_logical_blocks(const su2_kgeom &g) { return su2_en_av_blocks(g); }
