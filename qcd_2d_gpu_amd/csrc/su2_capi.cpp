/*
 * su2_capi.cpp - argument checking and dispatch of the C ABI declared in
 * include/su2mc.h.  Every entry point validates its arguments on the host
 * before anything is launched, so a bad size or pointer becomes an error
 * code instead of an out-of-bounds kernel.
 */
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "su2_internal.h"
#include "su2_observe.h"
#include "su2mc.h"

namespace {

thread_local char g_err[512];
std::atomic<int> g_sweep_block{256}; /* set by su2_set_sweep_block, read by every launch */

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(SU2_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define SU2_HIP(call)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call);                                                              \
    } while (0)

/* Validate the struct: even extents >= 2, consistent derived members, and
 * sizes whose indices fit the kernels' 32-bit arithmetic (6 * V/2 links). */
int check_lattice(const su2_lattice *l)
{
    if (!l) return fail(SU2_EINVAL, "lattice is NULL");
    if (l->t < 2 || l->x < 2 || l->y < 2 || (l->t & 1) || (l->x & 1) || (l->y & 1))
        return fail(SU2_EODD_EXTENT, "Lt, Lx and Ly must be even and >= 2 (got %d %d %d)", l->t, l->x, l->y);
    long long V = (long long)l->t * l->x * l->y;
    if (V > SU2_MAX_SITES) return fail(SU2_EINVAL, "lattice too large (V = %lld > %lld)", V, SU2_MAX_SITES);
    if (l->V != V) return fail(SU2_EINVAL, "lattice V = %d does not equal Lt*Lx*Ly = %lld", l->V, V);
    return SU2_OK;
}

int check_prec(int p)
{
    if (p != SU2_PREC_F64 && p != SU2_PREC_F32) return fail(SU2_EINVAL, "precision must be 0 (F64) or 1 (F32)");
    return SU2_OK;
}

su2_kgeom geom(const su2_lattice *l) { return su2_make_kgeom(l->t, l->x, l->y); }

hipStream_t S(void *s) { return (hipStream_t)s; }

/* Coupling checks of a link update (over-relaxation uses neither).  The
 * heat-bath's Kennedy-Pendleton loop accepts with probability ~ (beta |A|)^1.5
 * for small beta and never at beta <= 0 (its scale 0.5 / (|A| beta) is
 * infinite or negative), so beta must be at least SU2_HB_BETA_MIN; a
 * Metropolis proposal (sqrt(1 - eps^2), eps r) needs 0 <= eps <= 1. */
int check_couplings(const su2_lattice *l, int algo, int update)
{
    if (update != SU2_UPDATE_LINK) return SU2_OK;
    if (!std::isfinite(l->beta)) return fail(SU2_EINVAL, "beta is not finite");
    if (algo == SU2_ALGO_HEATBATH && !(l->beta >= SU2_HB_BETA_MIN))
        return fail(SU2_EINVAL, "heat-bath needs beta >= %g (got %g)", (double)SU2_HB_BETA_MIN, (double)l->beta);
    if (algo == SU2_ALGO_METROPOLIS) {
        if (!(l->epsilon >= 0.0f && l->epsilon <= 1.0f))
            return fail(SU2_EINVAL, "Metropolis needs 0 <= epsilon <= 1 (got %g)", (double)l->epsilon);
        if (l->updates < 0) return fail(SU2_EINVAL, "negative Metropolis hits");
    }
    return SU2_OK;
}

/* One sweep; with res non-null it also measures the field it starts from
 * (en_av into res, res2 with work), on its first launch when the sweep block
 * is SU2_EN_BLOCK threads, else with su2_launch_en_av before it (the same
 * partials, so the same bits either way). */
int sweep(const su2_lattice *l, void *phi, int prec, int *accept, uint32_t *states, int bc, int algo, int update,
          void *stream, double *res = nullptr, double *res2 = nullptr, void *work = nullptr)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(prec))) return rc;
    if (!phi || !accept) return fail(SU2_EINVAL, "field or accept pointer is NULL");
    if (bc < SU2_BC_PERIODIC || bc > SU2_BC_PMC) return fail(SU2_EINVAL, "unknown boundary %d", bc);
    if (update != SU2_UPDATE_LINK && update != SU2_UPDATE_OVERRELAX)
        return fail(SU2_EINVAL, "unknown update %d", update);
    if (algo != SU2_ALGO_HEATBATH && algo != SU2_ALGO_METROPOLIS) return fail(SU2_EINVAL, "unknown algorithm %d", algo);
    if ((rc = check_couplings(l, algo, update))) return rc;
    if (update == SU2_UPDATE_LINK && !states) return fail(SU2_EINVAL, "RNG states pointer is NULL");
    const su2_kgeom g = geom(l);
    su2_kphys p;
    p.beta = (double)l->beta;
    p.eps = (double)l->epsilon;
    p.hits = l->updates;
    const int upd = update == SU2_UPDATE_OVERRELAX ? SU2_K_OR : (algo == SU2_ALGO_HEATBATH ? SU2_K_HB : SU2_K_METRO);
    /* (y,e) (y,o) (x,e) (x,o) (t,e) (t,o): functions/mc_updates.h:467-474.
     * Heat-bath and over-relaxation accept every update they make, so the
     * t launches add each half-site's count for the whole sweep and the y
     * and x launches count nothing; Metropolis counts in every launch. */
    const int block = g_sweep_block;
    /* a measurement rides on the sweep with workgroups of SU2_EN_BLOCK
     * threads: pass 1 in the (y, even) launch, pass 2 in the (y, odd) one */
    double *mw = nullptr;
    su2_en_final_args fin = {nullptr, 0u, 0.0, nullptr, nullptr};
    if (res) {
        if (block == (int)SU2_EN_BLOCK) {
            mw = (double *)work;
            fin = {mw, su2_en_av_blocks(g), 1.0 / (2.0 * (double)g.Vh), res, res2};
        } else {
            hipError_t e = su2_launch_en_av(bc == SU2_BC_PMC, prec, phi, g, res, res2, (double *)work, S(stream));
            if (e != hipSuccess) return hip_fail(e, "en_av launch");
        }
    }
    for (int mu = 0; mu < 3; ++mu)
        for (int par = 0; par < 2; ++par) {
            const int acc_mode = upd == SU2_K_METRO ? SU2_K_ACC_LAUNCH
                                                    : (mu == SU2_DIR_T ? SU2_K_ACC_SWEEP : SU2_K_ACC_NONE);
            const bool y = mu == SU2_DIR_Y && mw;
            hipError_t e = su2_launch_sweep(mu, bc, upd, prec, phi, states, accept, g, p, par, block, acc_mode,
                                            S(stream), y && par == 0 ? mw : nullptr,
                                            y && par == 1 ? &fin : nullptr);
            if (e != hipSuccess) return hip_fail(e, "sweep launch");
        }
    return SU2_OK;
}

} // namespace

int su2_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" {

int su2_abi_version(void) { return SU2MC_ABI_VERSION; }

const char *su2_last_error(void) { return g_err; }

const char *su2_strerror(int status)
{
    switch (status) {
    case SU2_OK: return "ok";
    case SU2_EINVAL: return "invalid argument";
    case SU2_EODD_EXTENT: return "lattice extents must be even";
    case SU2_ENOMEM: return "out of memory";
    case SU2_EHIP: return "HIP runtime error";
    case SU2_EIO: return "I/O error";
    case SU2_ENODEV: return "no usable device";
    case SU2_EBUDGET: return "memory budget exceeded";
    case SU2_ECANCELLED: return "cancelled";
    }
    return "unknown status";
}

int su2_lattice_init(su2_lattice *l, int Lt, int Lx, int Ly, float beta, float epsilon, int updates)
{
    if (!l) return fail(SU2_EINVAL, "lattice is NULL");
    std::memset(l, 0, sizeof *l);
    l->y = Ly;
    l->x = Lx;
    l->t = Lt;
    long long V = (long long)Lt * Lx * Ly;
    l->V = (V > 0 && V < (1ll << 31)) ? (int)V : 0;
    l->d = 3;
    l->dg = 4;
    l->dt = 12;
    l->beta = beta;
    l->epsilon = epsilon;
    l->updates = updates;
    if (Lt < 2 || Lx < 2 || Ly < 2 || (Lt & 1) || (Lx & 1) || (Ly & 1))
        return fail(SU2_EODD_EXTENT, "Lt, Lx and Ly must be even and >= 2 (got %d %d %d)", Lt, Lx, Ly);
    return check_lattice(l);
}

size_t su2_field_bytes(const su2_lattice *l, int precision)
{
    if (check_lattice(l) || check_prec(precision)) return 0;
    return (size_t)l->V * 12 * (precision == SU2_PREC_F32 ? sizeof(float) : sizeof(double));
}

size_t su2_rng_bytes(const su2_lattice *l)
{
    if (check_lattice(l)) return 0;
    return (size_t)(l->V / 2) * 6 * sizeof(uint32_t);
}

size_t su2_accept_bytes(const su2_lattice *l)
{
    if (check_lattice(l)) return 0;
    return (size_t)l->V * sizeof(int);
}

size_t su2_en_av_work_bytes(const su2_lattice *l)
{
    if (check_lattice(l)) return 0;
    return su2_en_av_work_doubles(geom(l)) * sizeof(double);
}

int su2_setup_kernel(const su2_lattice *l, uint32_t *states, uint64_t seed, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if (!states) return fail(SU2_EINVAL, "RNG states pointer is NULL");
    hipError_t e = su2_launch_setup_rng(states, geom(l), seed, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "setup_kernel launch");
}

static int init_common(const su2_lattice *l, void *phi, int prec, uint32_t *states, int start, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(prec))) return rc;
    if (!phi) return fail(SU2_EINVAL, "field pointer is NULL");
    if (start != 1 && !states) return fail(SU2_EINVAL, "RNG states pointer is NULL");
    hipError_t e = su2_launch_init(start, prec, phi, states, geom(l), S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "init launch");
}

int su2_id_init2(const su2_lattice *l, void *phi, int precision, void *stream)
{
    return init_common(l, phi, precision, nullptr, 1, stream);
}

int su2_rand_init2(const su2_lattice *l, void *phi, int precision, uint32_t *states, void *stream)
{
    return init_common(l, phi, precision, states, 0, stream);
}

int su2_rand_init2_pcc(const su2_lattice *l, void *phi, int precision, uint32_t *states, void *stream)
{
    return init_common(l, phi, precision, states, 2, stream);
}

int su2_haar_init(const su2_lattice *l, void *phi, int precision, uint32_t *states, int pcc, void *stream)
{
    return init_common(l, phi, precision, states, pcc ? 4 : 3, stream);
}

int su2_mc_up_per(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream)
{
    return sweep(l, phi, precision, accept, states, SU2_BC_PERIODIC, algo, update, stream);
}

int su2_mc_up_pcc(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream)
{
    return sweep(l, phi, precision, accept, states, SU2_BC_PCC, algo, update, stream);
}

int su2_mc_up_pmc(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream)
{
    return sweep(l, phi, precision, accept, states, SU2_BC_PMC, algo, update, stream);
}

int su2_mc_up_measured(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int bc,
                       int algo, int update, double *res, double *res2, void *work, void *stream)
{
    if (!res || !res2 || !work) return fail(SU2_EINVAL, "NULL pointer passed to mc_up_measured");
    return sweep(l, phi, precision, accept, states, bc, algo, update, stream, res, res2, work);
}

int su2_mc_update(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int bc, int algo,
                  int update, int mu, int parity, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi || !accept) return fail(SU2_EINVAL, "field or accept pointer is NULL");
    if (mu < 0 || mu > 2 || parity < 0 || parity > 1) return fail(SU2_EINVAL, "mu must be 0..2 and parity 0..1");
    if (bc < SU2_BC_PERIODIC || bc > SU2_BC_PMC) return fail(SU2_EINVAL, "unknown boundary %d", bc);
    if (update != SU2_UPDATE_LINK && update != SU2_UPDATE_OVERRELAX)
        return fail(SU2_EINVAL, "unknown update %d", update);
    if (algo != SU2_ALGO_HEATBATH && algo != SU2_ALGO_METROPOLIS) return fail(SU2_EINVAL, "unknown algorithm %d", algo);
    if ((rc = check_couplings(l, algo, update))) return rc;
    if (update == SU2_UPDATE_LINK && !states) return fail(SU2_EINVAL, "RNG states pointer is NULL");
    su2_kphys p;
    p.beta = (double)l->beta;
    p.eps = (double)l->epsilon;
    p.hits = l->updates;
    const int upd = update == SU2_UPDATE_OVERRELAX ? SU2_K_OR : (algo == SU2_ALGO_HEATBATH ? SU2_K_HB : SU2_K_METRO);
    hipError_t e =
        su2_launch_sweep(mu, bc, upd, precision, phi, states, accept, geom(l), p, parity, g_sweep_block,
                         SU2_K_ACC_LAUNCH, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "sweep launch");
}

int su2_normalize(const su2_lattice *l, void *phi, int precision, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi) return fail(SU2_EINVAL, "field pointer is NULL");
    hipError_t e = su2_launch_normalize(precision, phi, geom(l), S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "normalize launch");
}

int su2_en_av(const su2_lattice *l, const void *phi, int precision, int pmc, double *res, double *res2, void *work,
              void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi || !res || !res2 || !work) return fail(SU2_EINVAL, "NULL pointer passed to en_av");
    hipError_t e = su2_launch_en_av(pmc ? 1 : 0, precision, phi, geom(l), res, res2, (double *)work, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "en_av launch");
}

int su2_en_av_noalloc(const su2_lattice *l, const void *phi, int precision, int pmc, double *res, double *res2,
                      void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi || !res || !res2) return fail(SU2_EINVAL, "NULL pointer passed to en_av_noalloc");
    if ((long long)l->V > SU2_EN_AV_NOALLOC_MAX_SITES)
        return fail(SU2_EINVAL, "en_av_noalloc: %d sites exceed its scratch limit of %lld sites (use su2_en_av)", l->V,
                    (long long)SU2_EN_AV_NOALLOC_MAX_SITES);
    hipError_t e = su2_launch_en_av_noalloc(pmc ? 1 : 0, precision, phi, geom(l), res, res2, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "en_av_noalloc launch");
}

int su2_accept_clear(const su2_lattice *l, int *accept, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if (!accept) return fail(SU2_EINVAL, "accept pointer is NULL");
    hipError_t e = su2_launch_fill_i32(accept, 0, (size_t)l->V, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "accept fill launch");
}

int su2_mean_lattice(const su2_lattice *l, const int *accept, long long n, double *mean, double *err, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if (!accept || !mean) return fail(SU2_EINVAL, "NULL pointer passed to mean_lattice");
    std::vector<int> h((size_t)l->V);
    SU2_HIP(hipStreamSynchronize(S(stream)));
    SU2_HIP(hipMemcpy(h.data(), accept, (size_t)l->V * sizeof(int), hipMemcpyDeviceToHost));
    /* statistics.h:17-28: mean = sum / V / n; error from the site variance */
    double m = 0.0, v = 0.0;
    for (int i = 0; i < l->V; ++i) {
        m += h[(size_t)i];
        v += (double)h[(size_t)i] * h[(size_t)i];
    }
    m = m / l->V / (double)n;
    double e = std::sqrt((v / l->V / (double)n - m * m) / l->V / (double)(n - 1));
    *mean = m;
    if (err) *err = e;
    return SU2_OK;
}

int su2_field_to_ref(const su2_lattice *l, const void *phi, int precision, double *phi_ref, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi || !phi_ref) return fail(SU2_EINVAL, "NULL field pointer");
    hipError_t e = su2_launch_to_ref(precision, phi, phi_ref, geom(l), S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "to_ref launch");
}

int su2_field_from_ref(const su2_lattice *l, const double *phi_ref, void *phi, int precision, void *stream)
{
    int rc = check_lattice(l);
    if (rc) return rc;
    if ((rc = check_prec(precision))) return rc;
    if (!phi || !phi_ref) return fail(SU2_EINVAL, "NULL field pointer");
    hipError_t e = su2_launch_from_ref(precision, phi_ref, phi, geom(l), S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "from_ref launch");
}

int su2_device_count(int *n)
{
    if (!n) return fail(SU2_EINVAL, "NULL");
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        return fail(SU2_ENODEV, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return SU2_OK;
}

int su2_set_device(int device)
{
    SU2_HIP(hipSetDevice(device));
    return SU2_OK;
}

int su2_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(SU2_EINVAL, "NULL");
    *ptr = nullptr;
    if (bytes == 0) return SU2_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return fail(SU2_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return SU2_OK;
}

int su2_free(void *ptr)
{
    if (ptr) SU2_HIP(hipFree(ptr));
    return SU2_OK;
}

int su2_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return SU2_OK;
    if (!dst || !src) return fail(SU2_EINVAL, "NULL");
    SU2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    SU2_HIP(hipStreamSynchronize(S(stream)));
    return SU2_OK;
}

int su2_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return SU2_OK;
    if (!dst || !src) return fail(SU2_EINVAL, "NULL");
    SU2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
    SU2_HIP(hipStreamSynchronize(S(stream)));
    return SU2_OK;
}

int su2_memset_zero(void *dst, size_t bytes, void *stream)
{
    if (bytes == 0) return SU2_OK;
    if (!dst) return fail(SU2_EINVAL, "NULL");
    SU2_HIP(hipMemsetAsync(dst, 0, bytes, S(stream)));
    return SU2_OK;
}

int su2_device_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return SU2_OK;
    if (!dst || !src || (bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(SU2_EINVAL, "device_copy needs 16-byte aligned pointers and a size multiple of 16");
    hipError_t e = su2_launch_copy(dst, src, bytes, S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "copy launch");
}

int su2_stream_probe(void *dst, const void *src, size_t bytes, int reads, int unroll, int nontemporal, int blocks,
                     void *stream)
{
    if (bytes == 0) return SU2_OK;
    if (!dst || !src || (bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(SU2_EINVAL, "stream_probe needs 16-byte aligned pointers and a size multiple of 16");
    if (!(reads == 1 || reads == 2 || reads == 4 || reads == 5) ||
        !(unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8) || blocks < 0)
        return fail(SU2_EINVAL, "stream_probe: reads 1, 2, 4 or 5; unroll 1, 2, 4 or 8; blocks >= 0");
    hipError_t e = su2_launch_stream_probe(dst, src, bytes, reads, unroll, nontemporal != 0, (uint32_t)blocks,
                                           S(stream));
    return e == hipSuccess ? SU2_OK : hip_fail(e, "stream probe launch");
}

int su2_phase_probe(double *buf, int blocks, int words, int phases, void *stream, double *launch_us,
                    double *barrier_us, double *xcd_barrier_us)
{
    if (!buf || !launch_us || !barrier_us || blocks < 1 || blocks > 1024 || words < 1 || words > 64 || phases < 1 ||
        phases > 100000)
        return fail(SU2_EINVAL, "phase_probe: buf, outputs, 1 <= blocks <= 1024, 1 <= words <= 64, 1 <= phases <= 100000");
    hipError_t e =
        su2_phase_probe_run(buf, (uint32_t)blocks, words, phases, S(stream), launch_us, barrier_us, xcd_barrier_us);
    if (e == hipErrorUnknown)
        return fail(SU2_EHIP, "phase probe: an XCD-local barrier ran out of spins or a read after it was stale");
    return e == hipSuccess ? SU2_OK : hip_fail(e, "phase probe");
}

int su2_stream_create(void **stream)
{
    if (!stream) return fail(SU2_EINVAL, "NULL");
    hipStream_t s;
    SU2_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return SU2_OK;
}

int su2_stream_destroy(void *stream)
{
    if (stream) SU2_HIP(hipStreamDestroy(S(stream)));
    return SU2_OK;
}

int su2_stream_synchronize(void *stream)
{
    SU2_HIP(hipStreamSynchronize(S(stream)));
    return SU2_OK;
}

int su2_event_create(void **event)
{
    if (!event) return fail(SU2_EINVAL, "NULL");
    hipEvent_t ev;
    SU2_HIP(hipEventCreate(&ev));
    *event = (void *)ev;
    return SU2_OK;
}

int su2_event_destroy(void *event)
{
    if (event) SU2_HIP(hipEventDestroy((hipEvent_t)event));
    return SU2_OK;
}

int su2_event_record(void *event, void *stream)
{
    if (!event) return fail(SU2_EINVAL, "NULL");
    SU2_HIP(hipEventRecord((hipEvent_t)event, S(stream)));
    return SU2_OK;
}

int su2_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!start || !stop || !ms) return fail(SU2_EINVAL, "NULL");
    SU2_HIP(hipEventSynchronize((hipEvent_t)stop));
    SU2_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SU2_OK;
}

int su2_set_sweep_block(int threads)
{
    if (threads == 0) {
        g_sweep_block = 256;
        return SU2_OK;
    }
    if (threads < 64 || threads > 256 || (threads % 64))
        return fail(SU2_EINVAL, "sweep block must be 64, 128, 192 or 256 threads");
    g_sweep_block = threads;
    return SU2_OK;
}

} /* extern "C" */
