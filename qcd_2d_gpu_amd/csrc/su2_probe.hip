/*
 * su2_probe.hip - the phase-separation probe (su2_phase_probe): what it
 * costs on this device to separate two dependent phases of a sweep by a
 * kernel boundary, and by a grid-wide barrier inside one cooperative
 * launch.  It bounds how fast a small lattice can be swept, whatever
 * carries its six (direction, parity) phases.
 */
#include <hip/hip_cooperative_groups.h>

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

} // namespace

hipError_t su2_phase_probe_run(double *buf, uint32_t blocks, int words, int phases, hipStream_t st, double *launch_us,
                               double *barrier_us)
{
    int per_cu = 0, ncu = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phase_coop, 256, 0);
    if (e != hipSuccess) return e;
    /* every workgroup of the cooperative launch must be resident at once */
    if ((uint64_t)blocks > (uint64_t)per_cu * (uint64_t)ncu) return hipErrorCooperativeLaunchTooLarge;
    hipEvent_t a = nullptr, b = nullptr;
    e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    float ms_launch = 0.0f, ms_coop = 0.0f;
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
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (e != hipSuccess) return e;
    *launch_us = 1e3 * (double)ms_launch / (2.0 * phases);
    *barrier_us = 1e3 * (double)ms_coop / (2.0 * phases);
    return hipSuccess;
}

