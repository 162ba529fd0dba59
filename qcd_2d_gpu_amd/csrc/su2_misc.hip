/*
 * su2_misc.hip - the non-sweep kernels of the path: RNG set-up, starts,
 * reunitarisation, the fused energy/plaquette reduction and the conversion
 * to and from the reference's on-disk field layout.
 */
#include <mutex>

#include "su2_observe.h"
#include "su2_sweep_site.h"

namespace {

__constant__ uint32_t c_jump1[SU2_MRG_NJUMP * 9];
__constant__ uint32_t c_jump2[SU2_MRG_NJUMP * 9];

__device__ static inline uint32_t gtid(void) { return blockIdx.x * blockDim.x + threadIdx.x; }

/* setup_kernel (functions/lattice_setup.h:15-21): stream h starts at
 * subsequence h of the seeded generator, offset 0. */
__global__ __launch_bounds__(256) void k_setup_rng(uint32_t *__restrict__ rng, uint32_t Vh, uint32_t s0, uint32_t s1,
                                                   uint32_t s2, uint32_t s3, uint32_t s4, uint32_t s5)
{
    const uint32_t h = gtid();
    if (h >= Vh) return;
    uint32_t s[6] = {s0, s1, s2, s3, s4, s5};
    su2_mrg_jump(s, h, c_jump1, c_jump2);
#pragma unroll
    for (int i = 0; i < 6; ++i) rng[(size_t)i * Vh + h] = s[i];
}

/* Starts.  Draw order per half-site h: parity 0 links y, x, t, then parity 1
 * links y, x, t (functions/lattice_setup.h:24-37).  PCC variant
 * (lattice_setup.h:52-81): the x and t links of the site at y = 0 are the
 * identity and consume no draws. */
template <int START, typename LinkT>
__global__ __launch_bounds__(256) void k_init(LinkT *__restrict__ U, uint32_t *__restrict__ rng, su2_kgeom g)
{
    const uint32_t h = gtid();
    if (h >= g.Vh) return;
    const size_t Vh = g.Vh;
    if (START == 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) su2_link_io<LinkT>::store(U + k * Vh + h, su2_q_id());
        return;
    }
    const uint32_t q = su2_div(h, g.div_lyh);
    const uint32_t yh = h - q * g.Lyh;
    const uint32_t t = su2_div(q, g.div_lx);
    const uint32_t x = q - t * g.Lx;
    const uint32_t par = (t + x) & 1u;
    uint32_t s[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) s[i] = rng[(size_t)i * Vh + h];
    for (int p = 0; p < 2; ++p) {
        const uint32_t y = 2u * yh + (par ^ (uint32_t)p);
        for (int mu = 0; mu < 3; ++mu) {
            su2_q l;
            const bool pcc = (START == 2 || START == 4);
            if (pcc && mu != SU2_DIR_Y && y == 0u)
                l = su2_q_id();
            else if (START == 3 || START == 4)
                l = su2_haar_link(s);
            else
                l = su2_rand_link(s);
            su2_link_io<LinkT>::store(U + (size_t)(p * 3 + mu) * Vh + h, l);
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) rng[(size_t)i * Vh + h] = s[i];
}

/* normalize (functions/lattice_setup.h:202-207): every link a <- a / |a|. */
template <typename LinkT> __global__ __launch_bounds__(256) void k_normalize(LinkT *__restrict__ U, uint32_t nlinks)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nlinks; i += (size_t)gridDim.x * blockDim.x)
        su2_link_io<LinkT>::store(U + i, su2_q_normalize(su2_link_io<LinkT>::load(U + i)));
}

/* Pass 1 of en_av in owner form (su2_observe.h): logical block b owns
 * parity-0 half-sites [b * SU2_EN_BLOCK, ...) and writes partials 2b,
 * 2b + 1 (su2_en_av_owner_block).  The blocks run in the sweeps' XCD-aware
 * order (su2_block_of: the t-neighbour rows a block reads stay in its XCD's
 * L2). */
template <int PMC, typename LinkT>
__global__ __launch_bounds__(2 * SU2_EN_BLOCK) void k_en_av_partial(const LinkT *__restrict__ U, su2_kgeom g,
                                                                    double *__restrict__ work, uint32_t order,
                                                                    uint32_t strip, uint32_t bpp)
{
    const uint32_t b = su2_block_of(order, strip, bpp);
    su2_en_av_owner_block<LinkT>(U, g, PMC, b, work);
}

/* Pass 2: one block of 1024 threads adds the partials in a fixed order
 * (su2_en_final_partials, then the block sum) and accumulates sum / V into
 * res, res2 (the reference adds into zeroed slots with atomicAdd,
 * functions/observables.h:137-140). */
__global__ __launch_bounds__(1024) void k_en_av_final(const double *__restrict__ work, uint32_t n, double invV,
                                                      double *__restrict__ res, double *__restrict__ res2)
{
    double te, tp;
    su2_en_final_partials(work, n, &te, &tp);
    su2_block_sum2(te, tp, SU2_EN_FINAL_BLOCK);
    if (threadIdx.x == 0) {
        *res += te * invV;
        *res2 += tp * invV;
    }
}

/* Reference layout (functions/lattice_setup.h:27-32, output.h:85):
 * phi[((p * Vh + h) * 3 + mu) * 4 + c]. */
template <typename LinkT>
__global__ __launch_bounds__(256) void k_to_ref(const LinkT *__restrict__ U, double *__restrict__ phi, uint32_t Vh)
{
    const size_t n = 6 * (size_t)Vh;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)j; /* < 3V < 2^32 */
        const uint32_t pm = i / Vh, h = i - pm * Vh, p = pm / 3u, mu = pm - 3u * p;
        su2_q l = su2_link_io<LinkT>::load(U + i);
        double *d = phi + ((size_t)(p * Vh + h) * 3 + mu) * 4;
        d[0] = l.a0;
        d[1] = l.a1;
        d[2] = l.a2;
        d[3] = l.a3;
    }
}

template <typename LinkT>
__global__ __launch_bounds__(256) void k_from_ref(const double *__restrict__ phi, LinkT *__restrict__ U, uint32_t Vh)
{
    const size_t n = 6 * (size_t)Vh;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)j; /* < 3V < 2^32 */
        const uint32_t pm = i / Vh, h = i - pm * Vh, p = pm / 3u, mu = pm - 3u * p;
        const double *d = phi + ((size_t)(p * Vh + h) * 3 + mu) * 4;
        su2_link_io<LinkT>::store(U + i, su2_q_make(d[0], d[1], d[2], d[3]));
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int *__restrict__ p, int v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* Plain streaming copy, 16 B per thread per iteration. */
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

/* Bandwidth probe: dst[i] = src_0[i] ^ ... ^ src_{READS-1}[i] over READS
 * source regions of n 16-byte items each, laid out one after the other in
 * src (READS = 1 is a copy).  Each thread keeps UNROLL items of every
 * source in flight (all loads issued before the first XOR); a wavefront's
 * load of one item index is one contiguous 1 KB run.  NT marks loads and
 * stores nontemporal.  This is the ceiling for a kernel that streams READS
 * arrays in and one out with nothing else to do. */
typedef unsigned int su2_v4u __attribute__((ext_vector_type(4)));

template <int NT> __device__ static inline __attribute__((always_inline)) su2_v4u probe_load(const su2_v4u *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <int NT> __device__ static inline __attribute__((always_inline)) void probe_store(su2_v4u *p, su2_v4u v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <int READS, int UNROLL, int NT>
__global__ __launch_bounds__(256) void k_stream_probe(su2_v4u *__restrict__ dst, const su2_v4u *__restrict__ src,
                                                      size_t n)
{
    const size_t step = (size_t)gridDim.x * (256 * UNROLL);
    for (size_t base = (size_t)blockIdx.x * (256 * UNROLL) + threadIdx.x; base < n; base += step) {
        su2_v4u v[READS][UNROLL];
#pragma unroll
        for (int r = 0; r < READS; ++r)
#pragma unroll
            for (int k = 0; k < UNROLL; ++k) {
                const size_t i = base + (size_t)k * 256;
                v[r][k] = i < n ? probe_load<NT>(src + (size_t)r * n + i) : su2_v4u{0u, 0u, 0u, 0u};
            }
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
            su2_v4u a = v[0][k];
#pragma unroll
            for (int r = 1; r < READS; ++r) a ^= v[r][k];
            const size_t i = base + (size_t)k * 256;
            if (i < n) probe_store<NT>(dst + i, a);
        }
    }
}

/* Partial sums of su2_en_av_noalloc: two doubles per pass-1 block of a
 * lattice of at most SU2_EN_AV_NOALLOC_MAX_SITES sites (2^25 half-sites,
 * 2^17 blocks of 256): 2 MiB, held by each device the library runs on until
 * the process ends.  Every call on a device writes
 * every partial it reads, so nothing in it needs clearing; calls that
 * overlap on one device (other streams, other host threads) race on it. */
#define SU2_EN_AV_SCRATCH_DOUBLES ((size_t)2 * (size_t)(SU2_EN_AV_NOALLOC_MAX_SITES / 2 / SU2_EN_BLOCK))
static __device__ __align__(256) double scratch[SU2_EN_AV_SCRATCH_DOUBLES];

/* The device address of `scratch`, looked up once per device with that
 * device current (host code never passes the array's name to a kernel: in
 * host code it is a stand-in, not a device address). */
std::mutex g_scratch_mu;
double *g_scratch_ptr[64];

hipError_t scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

/* The jump tables go to each device's __constant__ memory once; set-up may
 * be called from several host threads (a server thread and its caller), so
 * the check and the upload are serialised. */
std::mutex g_jump_mu;
bool g_jump_uploaded[64];

hipError_t ensure_jump_tables(void)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_jump_mu);
    if (dev >= 0 && dev < 64 && g_jump_uploaded[dev]) return hipSuccess;
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_jump1), su2_mrg_jump1_host, sizeof(su2_mrg_jump1_host));
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_jump2), su2_mrg_jump2_host, sizeof(su2_mrg_jump2_host));
    if (e != hipSuccess) return e;
    if (dev >= 0 && dev < 64) g_jump_uploaded[dev] = true;
    return hipSuccess;
}

} // namespace

hipError_t su2_launch_setup_rng(uint32_t *rng, const su2_kgeom &g, uint64_t seed, hipStream_t st)
{
    hipError_t e = ensure_jump_tables();
    if (e != hipSuccess) return e;
    uint32_t s[6];
    su2_mrg_seed(seed, s);
    const uint32_t nb = (g.Vh + 255) / 256;
    hipLaunchKernelGGL(k_setup_rng, dim3(nb), dim3(256), 0, st, rng, g.Vh, s[0], s[1], s[2], s[3], s[4], s[5]);
    return hipGetLastError();
}

template <typename LinkT>
static hipError_t launch_init_t(int start, void *U, uint32_t *rng, const su2_kgeom &g, hipStream_t st)
{
    const uint32_t nb = (g.Vh + 255) / 256;
    LinkT *u = (LinkT *)U;
    switch (start) {
    case 0: hipLaunchKernelGGL((k_init<0, LinkT>), dim3(nb), dim3(256), 0, st, u, rng, g); break;
    case 1: hipLaunchKernelGGL((k_init<1, LinkT>), dim3(nb), dim3(256), 0, st, u, rng, g); break;
    case 2: hipLaunchKernelGGL((k_init<2, LinkT>), dim3(nb), dim3(256), 0, st, u, rng, g); break;
    case 3: hipLaunchKernelGGL((k_init<3, LinkT>), dim3(nb), dim3(256), 0, st, u, rng, g); break;
    case 4: hipLaunchKernelGGL((k_init<4, LinkT>), dim3(nb), dim3(256), 0, st, u, rng, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t su2_launch_init(int start, int f32, void *U, uint32_t *rng, const su2_kgeom &g, hipStream_t st)
{
    return f32 ? launch_init_t<su2_l32>(start, U, rng, g, st) : launch_init_t<su2_l64>(start, U, rng, g, st);
}

/* Blocks of 256 for a grid-stride walk over n items: one item per thread
 * up to 2^20 blocks (about 1 100 per CU), then each thread takes several. */
static uint32_t stride_blocks(size_t n)
{
    const size_t nb = (n + 255) / 256;
    return (uint32_t)(nb < ((size_t)1 << 20) ? nb : ((size_t)1 << 20));
}

hipError_t su2_launch_normalize(int f32, void *U, const su2_kgeom &g, hipStream_t st)
{
    const uint32_t n = 6u * g.Vh; /* 3V < 2^32 */
    const uint32_t nb = stride_blocks(n);
    if (f32)
        hipLaunchKernelGGL(k_normalize<su2_l32>, dim3(nb), dim3(256), 0, st, (su2_l32 *)U, n);
    else
        hipLaunchKernelGGL(k_normalize<su2_l64>, dim3(nb), dim3(256), 0, st, (su2_l64 *)U, n);
    return hipGetLastError();
}

/* partial sums, 2 per logical block */
size_t su2_en_av_work_doubles(const su2_kgeom &g) { return 2 * (size_t)su2_en_av_blocks(g); }

hipError_t su2_launch_en_av_final(const su2_kgeom &g, const double *work, double *res, double *res2, hipStream_t st)
{
    const double invV = 1.0 / (2.0 * (double)g.Vh);
    hipLaunchKernelGGL(k_en_av_final, dim3(1), dim3(SU2_EN_FINAL_BLOCK), 0, st, work, su2_en_av_blocks(g), invV, res,
                       res2);
    return hipGetLastError();
}

hipError_t su2_launch_en_av(int pmc, int f32, const void *U, const su2_kgeom &g, double *res, double *res2,
                            double *work, hipStream_t st)
{
    const uint32_t nb = su2_en_av_blocks(g);
    uint32_t order, strip, bpp;
    su2_pick_block_order(g, SU2_EN_BLOCK, !f32, &order, &strip, &bpp);
    const dim3 G(nb), B(su2_en_av_threads(f32));
    if (f32) {
        if (pmc)
            hipLaunchKernelGGL((k_en_av_partial<1, su2_l32>), G, B, 0, st, (const su2_l32 *)U, g, work, order, strip, bpp);
        else
            hipLaunchKernelGGL((k_en_av_partial<0, su2_l32>), G, B, 0, st, (const su2_l32 *)U, g, work, order, strip, bpp);
    } else {
        if (pmc)
            hipLaunchKernelGGL((k_en_av_partial<1, su2_l64>), G, B, 0, st, (const su2_l64 *)U, g, work, order, strip, bpp);
        else
            hipLaunchKernelGGL((k_en_av_partial<0, su2_l64>), G, B, 0, st, (const su2_l64 *)U, g, work, order, strip, bpp);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return su2_launch_en_av_final(g, work, res, res2, st);
}

hipError_t su2_launch_en_av_noalloc(int pmc, int f32, const void *U, const su2_kgeom &g, double *res, double *res2,
                                    hipStream_t st)
{
    if (su2_en_av_work_doubles(g) > SU2_EN_AV_SCRATCH_DOUBLES) return hipErrorInvalidValue;
    double *work = nullptr;
    hipError_t e = scratch_address(&work);
    if (e != hipSuccess) return e;
    return su2_launch_en_av(pmc, f32, U, g, res, res2, work, st);
}

hipError_t su2_launch_to_ref(int f32, const void *U, double *phi_ref, const su2_kgeom &g, hipStream_t st)
{
    const uint32_t nb = stride_blocks(6 * (size_t)g.Vh);
    if (f32)
        hipLaunchKernelGGL(k_to_ref<su2_l32>, dim3(nb), dim3(256), 0, st, (const su2_l32 *)U, phi_ref, g.Vh);
    else
        hipLaunchKernelGGL(k_to_ref<su2_l64>, dim3(nb), dim3(256), 0, st, (const su2_l64 *)U, phi_ref, g.Vh);
    return hipGetLastError();
}

hipError_t su2_launch_from_ref(int f32, const double *phi_ref, void *U, const su2_kgeom &g, hipStream_t st)
{
    const uint32_t nb = stride_blocks(6 * (size_t)g.Vh);
    if (f32)
        hipLaunchKernelGGL(k_from_ref<su2_l32>, dim3(nb), dim3(256), 0, st, phi_ref, (su2_l32 *)U, g.Vh);
    else
        hipLaunchKernelGGL(k_from_ref<su2_l64>, dim3(nb), dim3(256), 0, st, phi_ref, (su2_l64 *)U, g.Vh);
    return hipGetLastError();
}

hipError_t su2_launch_fill_i32(int *p, int v, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_i32, dim3((uint32_t)nb), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t su2_launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    size_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_copy16, dim3((uint32_t)nb), dim3(256), 0, st, (uint4 *)dst, (const uint4 *)src, n);
    return hipGetLastError();
}

template <int READS, int UNROLL>
static hipError_t launch_probe_t(void *dst, const void *src, size_t n, int nt, uint32_t blocks, hipStream_t st)
{
    if (nt)
        hipLaunchKernelGGL((k_stream_probe<READS, UNROLL, 1>), dim3(blocks), dim3(256), 0, st, (su2_v4u *)dst,
                           (const su2_v4u *)src, n);
    else
        hipLaunchKernelGGL((k_stream_probe<READS, UNROLL, 0>), dim3(blocks), dim3(256), 0, st, (su2_v4u *)dst,
                           (const su2_v4u *)src, n);
    return hipGetLastError();
}

template <int READS>
static hipError_t launch_probe_r(void *dst, const void *src, size_t n, int unroll, int nt, uint32_t blocks,
                                 hipStream_t st)
{
    switch (unroll) {
    case 1: return launch_probe_t<READS, 1>(dst, src, n, nt, blocks, st);
    case 2: return launch_probe_t<READS, 2>(dst, src, n, nt, blocks, st);
    case 4: return launch_probe_t<READS, 4>(dst, src, n, nt, blocks, st);
    case 8: return launch_probe_t<READS, 8>(dst, src, n, nt, blocks, st);
    }
    return hipErrorInvalidValue;
}

hipError_t su2_launch_stream_probe(void *dst, const void *src, size_t bytes, int reads, int unroll, int nt,
                                   uint32_t blocks, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    if (blocks == 0) {
        const size_t nb = (n + 256 * (size_t)unroll - 1) / (256 * (size_t)unroll);
        blocks = (uint32_t)(nb < ((size_t)1 << 20) ? nb : ((size_t)1 << 20));
    }
    switch (reads) {
    case 1: return launch_probe_r<1>(dst, src, n, unroll, nt, blocks, st);
    case 2: return launch_probe_r<2>(dst, src, n, unroll, nt, blocks, st);
    case 4: return launch_probe_r<4>(dst, src, n, unroll, nt, blocks, st);
    case 5: return launch_probe_r<5>(dst, src, n, unroll, nt, blocks, st);
    }
    return hipErrorInvalidValue;
}
