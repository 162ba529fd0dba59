/*
 * su2_handle.cpp - the handle API of su2mc.h: a context (device + stream)
 * and lattices that own their buffers, built on the raw-pointer entry
 * points of su2_capi.cpp.  This is what kernel.cu:main does to its one
 * lattice (kernel.cu:58-236) with the allocations folded into the handle.
 */
#include <hip/hip_runtime.h>

#include <vector>

#include "su2_internal.h"
#include "su2mc.h"

struct su2_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    int lattices = 0;
};

struct su2_lat {
    su2_ctx *ctx = nullptr;
    su2_lattice l{};
    int bc = SU2_BC_PERIODIC, precision = SU2_PREC_F64;
    void *phi = nullptr;
    uint32_t *states = nullptr;
    int *accept = nullptr;
    void *work = nullptr;
    double *res = nullptr; /* two device slots: energy, plaquette */
    int64_t attempts_per_site = 0;
};

namespace {

#define H_HIP(call)                                                                                                    \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return su2_set_error(SU2_EHIP, "%s: %s", #call, hipGetErrorString(e_));                \
    } while (0)

#define H_RC(call)                                                                                                     \
    do {                                                                                                               \
        int rc_ = (call);                                                                                              \
        if (rc_ != SU2_OK) return rc_;                                                                                 \
    } while (0)

int enter(const su2_lat *h)
{
    if (!h || !h->ctx) return su2_set_error(SU2_EINVAL, "NULL lattice handle");
    H_HIP(hipSetDevice(h->ctx->device));
    return SU2_OK;
}

void release(su2_lat *h)
{
    if (h->res) hipFree(h->res);
    if (h->work) hipFree(h->work);
    if (h->accept) hipFree(h->accept);
    if (h->states) hipFree(h->states);
    if (h->phi) hipFree(h->phi);
}

} // namespace

extern "C" {

int su2_ctx_create(int backend, int device, su2_ctx **out)
{
    if (!out) return su2_set_error(SU2_EINVAL, "NULL output");
    *out = nullptr;
    if (backend == SU2_BACKEND_CPU)
        return su2_set_error(SU2_EINVAL, "libsu2mc has no CPU backend; SU2_BACKEND_HIP is the only one");
    if (backend != SU2_BACKEND_HIP) return su2_set_error(SU2_EINVAL, "unknown backend %d", backend);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return su2_set_error(SU2_ENODEV, "no HIP device");
    if (device < 0 || device >= n) return su2_set_error(SU2_EINVAL, "device %d of %d", device, n);
    H_HIP(hipSetDevice(device));
    su2_ctx *c = new su2_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return su2_set_error(SU2_EHIP, "stream create: %s", hipGetErrorString(e));
    }
    *out = c;
    return SU2_OK;
}

int su2_ctx_destroy(su2_ctx *c)
{
    if (!c) return SU2_OK;
    if (c->lattices) return su2_set_error(SU2_EINVAL, "context still has %d lattices", c->lattices);
    hipSetDevice(c->device);
    hipStreamSynchronize(c->st);
    hipStreamDestroy(c->st);
    delete c;
    return SU2_OK;
}

int su2_ctx_stream(su2_ctx *c, void **stream)
{
    if (!c || !stream) return su2_set_error(SU2_EINVAL, "NULL");
    *stream = (void *)c->st;
    return SU2_OK;
}

int su2_ctx_synchronize(su2_ctx *c)
{
    if (!c) return su2_set_error(SU2_EINVAL, "NULL context");
    H_HIP(hipSetDevice(c->device));
    H_HIP(hipStreamSynchronize(c->st));
    return SU2_OK;
}

int su2_lattice_create(su2_ctx *c, int Lt, int Lx, int Ly, int bc, int precision, su2_lat **out)
{
    if (!c || !out) return su2_set_error(SU2_EINVAL, "NULL");
    *out = nullptr;
    if (bc < SU2_BC_PERIODIC || bc > SU2_BC_PMC) return su2_set_error(SU2_EINVAL, "boundary %d", bc);
    if (precision != SU2_PREC_F64 && precision != SU2_PREC_F32)
        return su2_set_error(SU2_EINVAL, "precision %d", precision);
    su2_lattice l;
    H_RC(su2_lattice_init(&l, Lt, Lx, Ly, 0.0f, 0.0f, 1));
    H_HIP(hipSetDevice(c->device));
    su2_lat *h = new su2_lat();
    h->ctx = c;
    h->l = l;
    h->bc = bc;
    h->precision = precision;
    if (hipMalloc(&h->phi, su2_field_bytes(&l, precision)) != hipSuccess ||
        hipMalloc((void **)&h->states, su2_rng_bytes(&l)) != hipSuccess ||
        hipMalloc((void **)&h->accept, su2_accept_bytes(&l)) != hipSuccess ||
        hipMalloc(&h->work, su2_en_av_work_bytes(&l)) != hipSuccess ||
        hipMalloc((void **)&h->res, 2 * sizeof(double)) != hipSuccess) {
        release(h);
        delete h;
        return su2_set_error(SU2_ENOMEM, "lattice %dx%dx%d buffers", Lt, Lx, Ly);
    }
    /* acceptance counters start at zero, whatever start the caller uses */
    int rc = su2_accept_clear(&h->l, h->accept, c->st);
    if (rc != SU2_OK) {
        release(h);
        delete h;
        return rc;
    }
    ++c->lattices;
    *out = h;
    return SU2_OK;
}

int su2_lattice_destroy(su2_lat *h)
{
    if (!h) return SU2_OK;
    if (h->ctx) {
        hipSetDevice(h->ctx->device);
        hipStreamSynchronize(h->ctx->st);
        --h->ctx->lattices;
    }
    release(h);
    delete h;
    return SU2_OK;
}

int su2_lattice_get(const su2_lat *h, su2_lattice *l, int *bc, int *precision)
{
    if (!h) return su2_set_error(SU2_EINVAL, "NULL lattice handle");
    if (l) *l = h->l;
    if (bc) *bc = h->bc;
    if (precision) *precision = h->precision;
    return SU2_OK;
}

int su2_lattice_buffers(su2_lat *h, void **phi, uint32_t **states, int **accept)
{
    if (!h) return su2_set_error(SU2_EINVAL, "NULL lattice handle");
    if (phi) *phi = h->phi;
    if (states) *states = h->states;
    if (accept) *accept = h->accept;
    return SU2_OK;
}

int su2_rng_init(su2_lat *h, uint64_t seed)
{
    H_RC(enter(h));
    return su2_setup_kernel(&h->l, h->states, seed, h->ctx->st);
}

int su2_init(su2_lat *h, int start)
{
    H_RC(enter(h));
    void *st = h->ctx->st;
    switch (start) {
    case SU2_START_COLD: H_RC(su2_id_init2(&h->l, h->phi, h->precision, st)); break;
    case SU2_START_HOT:
        if (h->bc == SU2_BC_PCC)
            H_RC(su2_rand_init2_pcc(&h->l, h->phi, h->precision, h->states, st));
        else
            H_RC(su2_rand_init2(&h->l, h->phi, h->precision, h->states, st));
        break;
    case SU2_START_HAAR:
        H_RC(su2_haar_init(&h->l, h->phi, h->precision, h->states, h->bc == SU2_BC_PCC, st));
        break;
    default: return su2_set_error(SU2_EINVAL, "start %d", start);
    }
    return su2_accept_reset(h);
}

int su2_sweep(su2_lat *h, int algo, int update, float beta, float eps, int hits)
{
    H_RC(enter(h));
    if (algo == SU2_ALGO_METROPOLIS && update == SU2_UPDATE_LINK && hits < 1)
        return su2_set_error(SU2_EINVAL, "Metropolis hits %d < 1", hits);
    h->l.beta = beta;
    h->l.epsilon = eps;
    h->l.updates = hits < 1 ? 1 : hits;
    void *st = h->ctx->st;
    int rc;
    switch (h->bc) {
    case SU2_BC_PCC: rc = su2_mc_up_pcc(&h->l, h->phi, h->precision, h->accept, h->states, algo, update, st); break;
    case SU2_BC_PMC: rc = su2_mc_up_pmc(&h->l, h->phi, h->precision, h->accept, h->states, algo, update, st); break;
    default: rc = su2_mc_up_per(&h->l, h->phi, h->precision, h->accept, h->states, algo, update, st); break;
    }
    if (rc != SU2_OK) return rc;
    const int per = update == SU2_UPDATE_OVERRELAX ? 1 : (algo == SU2_ALGO_METROPOLIS ? h->l.updates : 1);
    h->attempts_per_site += 3 * per;
    return SU2_OK;
}

int su2_renormalize(su2_lat *h)
{
    H_RC(enter(h));
    return su2_normalize(&h->l, h->phi, h->precision, h->ctx->st);
}

int su2_measure(su2_lat *h, double *E, double *P)
{
    H_RC(enter(h));
    if (!E || !P) return su2_set_error(SU2_EINVAL, "NULL output");
    hipStream_t st = h->ctx->st;
    H_RC(su2_launch_fill_i32((int *)h->res, 0, 4, st) == hipSuccess ? SU2_OK
                                                                        : su2_set_error(SU2_EHIP, "slot clear"));
    H_RC(su2_en_av(&h->l, h->phi, h->precision, h->bc == SU2_BC_PMC, h->res, h->res + 1, h->work, st));
    double r[2];
    H_HIP(hipMemcpyAsync(r, h->res, sizeof r, hipMemcpyDeviceToHost, st));
    H_HIP(hipStreamSynchronize(st));
    *E = r[0];
    *P = r[1];
    return SU2_OK;
}

int su2_accept_read(su2_lat *h, int64_t *accepted, int64_t *attempts)
{
    H_RC(enter(h));
    hipStream_t st = h->ctx->st;
    std::vector<int> a((size_t)h->l.V);
    H_HIP(hipMemcpyAsync(a.data(), h->accept, a.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    H_HIP(hipStreamSynchronize(st));
    int64_t s = 0;
    for (int v : a) s += v;
    if (accepted) *accepted = s;
    if (attempts) *attempts = (int64_t)h->l.V * h->attempts_per_site;
    return SU2_OK;
}

int su2_accept_reset(su2_lat *h)
{
    H_RC(enter(h));
    h->attempts_per_site = 0;
    return su2_accept_clear(&h->l, h->accept, h->ctx->st);
}

int su2_field_download(su2_lat *h, double *host_phi_ref)
{
    H_RC(enter(h));
    if (!host_phi_ref) return su2_set_error(SU2_EINVAL, "NULL output");
    const size_t n = (size_t)12 * h->l.V * sizeof(double);
    hipStream_t st = h->ctx->st;
    double *tmp = nullptr;
    H_HIP(hipMalloc((void **)&tmp, n));
    int rc = su2_field_to_ref(&h->l, h->phi, h->precision, tmp, st);
    if (rc == SU2_OK) {
        hipError_t e = hipMemcpyAsync(host_phi_ref, tmp, n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = su2_set_error(SU2_EHIP, "field download: %s", hipGetErrorString(e));
    }
    hipFree(tmp);
    return rc;
}

int su2_field_upload(su2_lat *h, const double *host_phi_ref)
{
    H_RC(enter(h));
    if (!host_phi_ref) return su2_set_error(SU2_EINVAL, "NULL input");
    const size_t n = (size_t)12 * h->l.V * sizeof(double);
    hipStream_t st = h->ctx->st;
    double *tmp = nullptr;
    H_HIP(hipMalloc((void **)&tmp, n));
    hipError_t e = hipMemcpyAsync(tmp, host_phi_ref, n, hipMemcpyHostToDevice, st);
    int rc = e == hipSuccess ? su2_field_from_ref(&h->l, tmp, h->phi, h->precision, st)
                             : su2_set_error(SU2_EHIP, "field upload: %s", hipGetErrorString(e));
    e = hipStreamSynchronize(st);
    if (rc == SU2_OK && e != hipSuccess) rc = su2_set_error(SU2_EHIP, "field upload: %s", hipGetErrorString(e));
    hipFree(tmp);
    return rc;
}

int su2_rng_download(su2_lat *h, uint32_t *host_states)
{
    H_RC(enter(h));
    if (!host_states) return su2_set_error(SU2_EINVAL, "NULL output");
    hipStream_t st = h->ctx->st;
    H_HIP(hipMemcpyAsync(host_states, h->states, su2_rng_bytes(&h->l), hipMemcpyDeviceToHost, st));
    H_HIP(hipStreamSynchronize(st));
    return SU2_OK;
}

int su2_rng_upload(su2_lat *h, const uint32_t *host_states)
{
    H_RC(enter(h));
    if (!host_states) return su2_set_error(SU2_EINVAL, "NULL input");
    hipStream_t st = h->ctx->st;
    H_HIP(hipMemcpyAsync(h->states, host_states, su2_rng_bytes(&h->l), hipMemcpyHostToDevice, st));
    H_HIP(hipStreamSynchronize(st));
    return SU2_OK;
}

} /* extern "C" */
