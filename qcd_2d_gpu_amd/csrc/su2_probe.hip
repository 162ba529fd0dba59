/*
 * su2_probe.hip - the phase-separation probe (su2_phase_probe): what it
 * costs on this device to separate two dependent phases of a sweep by a
 * kernel boundary, by a grid-wide barrier inside one cooperative launch, and
 * by a barrier among only the workgroups that share one XCD's L2.  It bounds
 * how fast a small lattice can be swept, whatever carries its six
 * (direction, parity) phases.
 */
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "su2_internal.h"

namespace {

/* Phase-separation probe (su2_phase_probe).  One phase: every thread adds
 * 1 to `words` doubles of its workgroup's region (WRITE) and, after the
 * separation, reads as many of the next workgroup's region (READ), the
 * dependence of one (direction, parity) launch of a sweep on the one
 * before.  k_phase_half runs one half per launch; k_phase_coop runs every
 * phase in one cooperative launch with a grid-wide barrier after each half. */
__global__ __launch_bounds__(256) void k_phase_half(double *__restrict__ buf, int words, int read)
{
    if (read) {
        const size_t nb = (size_t)((blockIdx.x + 1u) % gridDim.x) * 256u * (size_t)words;
        double acc = 0.0;
        for (int w = 0; w < words; ++w) acc += buf[nb + (size_t)w * 256u + threadIdx.x];
        if (acc < -1.0) buf[0] = acc; /* never: every value is a count >= 0; keeps the loads */
    } else {
        const size_t base = (size_t)blockIdx.x * 256u * (size_t)words;
        for (int w = 0; w < words; ++w) buf[base + (size_t)w * 256u + threadIdx.x] += 1.0;
    }
}

__global__ __launch_bounds__(256) void k_phase_coop(double *__restrict__ buf, int words, int phases)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const size_t base = (size_t)blockIdx.x * 256u * (size_t)words;
    const size_t nb = (size_t)((blockIdx.x + 1u) % gridDim.x) * 256u * (size_t)words;
    for (int s = 0; s < phases; ++s) {
        for (int w = 0; w < words; ++w) buf[base + (size_t)w * 256u + threadIdx.x] += 1.0;
        grid.sync();
        double acc = 0.0;
        for (int w = 0; w < words; ++w) acc += buf[nb + (size_t)w * 256u + threadIdx.x];
        if (acc < -1.0) buf[0] = acc;
        grid.sync();
    }
}

/* Control words of the XCD-local barrier: per XCD the workgroups that found
 * themselves on it (reg), the order they registered in (who), and the
 * barrier's arrival counter (bar); how many workgroups registered at all
 * (checked); spins that ran out and neighbour reads that saw stale data. */
constexpr uint32_t XCD_SLOTS = 16, XCD_MAX_BLOCKS = 1024;
struct xcd_ctl {
    uint32_t reg[XCD_SLOTS], bar[XCD_SLOTS], checked, spin_out, stale;
    uint32_t who[XCD_SLOTS * XCD_MAX_BLOCKS];
};
/* a bound on every wait (each try is one L2 round trip, well under 1 us):
 * a barrier that cannot complete ends the kernel with spin_out set, it does
 * not hang it; with every workgroup resident (cooperative launch) it never
 * runs out */
constexpr uint32_t XCD_SPIN_MAX = 1u << 22;

__device__ inline uint32_t ld_agent(uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

/* Thread 0 arrives and waits for the group's n workgroups, the other
 * threads at the workgroup barrier.  Before arriving, the workgroup's stores
 * are drained (s_waitcnt 0: each store acknowledged by the L2, which every
 * member of the group shares); after it, the acquire invalidates this CU's
 * L1 so that the reads that follow go to that L2.  No L2 write-back: that
 * is what an agent-scope release (HIP's grid barrier) pays for, so that
 * other XCDs see the data.  false: a wait ran out. */
__device__ inline bool xcd_barrier(xcd_ctl *c, uint32_t xcd, uint32_t target, uint32_t *abort_flag)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(&c->bar[xcd], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spin = 0;
        while (ld_agent(&c->bar[xcd]) < target)
            if (++spin > XCD_SPIN_MAX) {
                *abort_flag = 1u;
                atomicAdd(&c->spin_out, 1u);
                break;
            }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return *abort_flag == 0u;
}

/* k_phase_coop with the barrier among the workgroups of one XCD only: every
 * workgroup reads its XCD from the XCC_ID hardware register (hwreg 20, bits
 * 3:0), registers in that XCD's group, and works with the next workgroup of
 * its own group, so the groups run side by side, one per XCD, whatever the
 * dispatcher's placement.  Every read is checked against the value its
 * writer stored: a stale read (an L1 or L2 copy the barrier failed to
 * expose) is counted. */
__global__ __launch_bounds__(256) void k_phase_xcd(double *__restrict__ buf, xcd_ctl *c, int words, int phases)
{
    __shared__ uint32_t s_xcd, s_tk, s_n, s_nb, s_abort;
    if (threadIdx.x == 0) {
        const uint32_t xcd = __builtin_amdgcn_s_getreg((3 << 11) | 20) & (XCD_SLOTS - 1u);
        const uint32_t tk = __hip_atomic_fetch_add(&c->reg[xcd], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&c->who[xcd * XCD_MAX_BLOCKS + tk], blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&c->checked, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spin = 0, ab = 0;
        while (ld_agent(&c->checked) < gridDim.x)
            if (++spin > XCD_SPIN_MAX) {
                ab = 1u;
                atomicAdd(&c->spin_out, 1u);
                break;
            }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const uint32_t n = ld_agent(&c->reg[xcd]);
        s_xcd = xcd;
        s_tk = tk;
        s_n = n;
        s_nb = ab ? blockIdx.x : ld_agent(&c->who[xcd * XCD_MAX_BLOCKS + (tk + 1u) % n]);
        s_abort = ab;
    }
    __syncthreads();
    if (s_abort) return;
    const uint32_t xcd = s_xcd, n = s_n;
    const size_t base = (size_t)blockIdx.x * 256u * (size_t)words;
    const size_t nb = (size_t)s_nb * 256u * (size_t)words;
    uint32_t stale = 0;
    for (int s = 0; s < phases; ++s) {
        double mine = 0.0;
        for (int w = 0; w < words; ++w) mine = (buf[base + (size_t)w * 256u + threadIdx.x] += 1.0);
        if (!xcd_barrier(c, xcd, n * (uint32_t)(2 * s + 1), &s_abort)) break;
        /* every element holds the same count (each phase adds 1 to all of
         * them), so the neighbour's last element must equal this thread's */
        double acc = 0.0;
        for (int w = 0; w < words; ++w) acc = buf[nb + (size_t)w * 256u + threadIdx.x];
        stale += acc != mine;
        if (!xcd_barrier(c, xcd, n * (uint32_t)(2 * s + 2), &s_abort)) break;
    }
    if (stale) atomicAdd(&c->stale, stale);
}

} // namespace

hipError_t su2_phase_probe_run(double *buf, uint32_t blocks, int words, int phases, hipStream_t st, double *launch_us,
                               double *barrier_us, double *xcd_barrier_us)
{
    int per_cu = 0, per_cu_x = 0, ncu = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phase_coop, 256, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_x, k_phase_xcd, 256, 0);
    if (e != hipSuccess) return e;
    /* every workgroup of a cooperative launch must be resident at once */
    if ((uint64_t)blocks > (uint64_t)std::min(per_cu, per_cu_x) * (uint64_t)ncu || blocks > XCD_MAX_BLOCKS)
        return hipErrorCooperativeLaunchTooLarge;
    hipEvent_t a = nullptr, b = nullptr;
    xcd_ctl *ctl = nullptr;
    e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess && xcd_barrier_us) e = hipMalloc((void **)&ctl, sizeof(xcd_ctl));
    float ms_launch = 0.0f, ms_coop = 0.0f, ms_xcd = 0.0f;
    uint32_t bad[2] = {0, 0};
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) { /* the first round warms up */
        e = hipEventRecord(a, st);
        for (int s = 0; s < phases && e == hipSuccess; ++s)
            for (int half = 0; half < 2 && e == hipSuccess; ++half) {
                hipLaunchKernelGGL(k_phase_half, dim3(blocks), dim3(256), 0, st, buf, words, half);
                e = hipGetLastError();
            }
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms_launch, a, b);
        void *args[] = {&buf, &words, &phases};
        if (e == hipSuccess) e = hipEventRecord(a, st);
        if (e == hipSuccess) e = hipLaunchCooperativeKernel((const void *)k_phase_coop, dim3(blocks), dim3(256), args, 0, st);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms_coop, a, b);
        if (!ctl) continue;
        /* a kernel clears the control words (no memset node on the stream) */
        void *xargs[] = {&buf, &ctl, &words, &phases};
        if (e == hipSuccess) e = su2_launch_fill_i32((int *)ctl, 0, sizeof(xcd_ctl) / 4, st);
        if (e == hipSuccess) e = hipEventRecord(a, st);
        if (e == hipSuccess) e = hipLaunchCooperativeKernel((const void *)k_phase_xcd, dim3(blocks), dim3(256), xargs, 0, st);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms_xcd, a, b);
        if (e == hipSuccess) e = hipMemcpy(bad, &ctl->spin_out, sizeof(bad), hipMemcpyDeviceToHost);
        if (e == hipSuccess && (bad[0] || bad[1])) e = hipErrorUnknown; /* a wait ran out or a read was stale */
    }
    if (ctl) hipFree(ctl);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (e != hipSuccess) return e;
    *launch_us = 1e3 * (double)ms_launch / (2.0 * phases);
    *barrier_us = 1e3 * (double)ms_coop / (2.0 * phases);
    if (xcd_barrier_us) *xcd_barrier_us = 1e3 * (double)ms_xcd / (2.0 * phases);
    return hipSuccess;
}
