/*
 * su2_engine.cpp - the batched engine (many resident lattices advanced by
 * shared launches, su2_batch.hip) and the serving loop on top of it.
 *
 * Engine: one device arena (first-fit allocator, 256-byte aligned regions)
 * holds each job's field, RNG planes, acceptance counters, en_av partials
 * and measurement series; a compact descriptor table (su2_bdesc) of the
 * resident, unfinished jobs is uploaded from pinned staging whenever the
 * membership changes.  A tick is a sequence of launches whose grids do not
 * depend on the membership, so it can be captured as a hipGraph and
 * replayed.  Per-job schedules (thermalisation, reunitarisation,
 * measurement) are evaluated on the device from per-slot sweep counters; the
 * host keeps identical counters, to know when a job is done and to leave out
 * of a tick the launches that no resident job needs in it (reunitarisation,
 * measurement, over-relaxation passes beyond the largest count, the mixed
 * heat-bath/Metropolis update body when all jobs use one algorithm).  Each
 * such tick plan is captured once, on first use, and its graph kept.
 *
 * Server: a thread owning an engine.  Submitted jobs queue; each loop
 * iteration admits queued jobs that fit (arrival order, first that fits),
 * enqueues one tick, waits for the previous tick, and retires the jobs it
 * finished (results copied with plain copies that do not wait for the
 * engine's non-blocking stream, then the slot freed).  One tick is always
 * queued ahead of the host, so the device does not idle while the host does
 * bookkeeping.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "su2_batch.h"
#include "su2_observe.h"
#include "su2mc.h"

namespace {

#define ENG_HIP(call)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return su2_set_error(SU2_EHIP, "%s: %s", #call, hipGetErrorString(e_));                \
    } while (0)

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

struct job_layout {
    size_t U, rng, acc, part, partc, resE, resP, total;
    uint32_t en_nb;
};

job_layout layout_of(const su2_job &j, int precision)
{
    job_layout L;
    const size_t V = (size_t)j.Lt * j.Lx * j.Ly, Vh = V / 2;
    const su2_kgeom g = su2_make_kgeom(j.Lt, j.Lx, j.Ly);
    L.en_nb = su2_en_av_logical_blocks(g);
    size_t o = 0;
    L.U = o;
    o += align256(V * 12 * (precision == SU2_PREC_F32 ? 4 : 8));
    L.rng = o;
    o += align256(Vh * 6 * 4);
    L.acc = o;
    o += align256(V * 4);
    L.part = o;
    o += align256((size_t)L.en_nb * 16);
    L.partc = o;
    o += align256((size_t)L.en_nb * 16);
    L.resE = o;
    o += align256((size_t)j.n_meas * 8);
    L.resP = o;
    o += align256((size_t)j.n_meas * 8);
    L.total = o;
    return L;
}

int check_job(const su2_job *j, int max_or)
{
    if (!j) return su2_set_error(SU2_EINVAL, "job is NULL");
    if (j->Lt < 2 || j->Lx < 2 || j->Ly < 2 || (j->Lt & 1) || (j->Lx & 1) || (j->Ly & 1))
        return su2_set_error(SU2_EODD_EXTENT, "job extents must be even and >= 2");
    if ((long long)j->Lt * j->Lx * j->Ly > SU2_MAX_SITES) return su2_set_error(SU2_EINVAL, "job too large");
    if (j->n_therm < 1 || j->n_meas < 1 || j->n_cor < 1 || j->or_sweeps < 0 || j->or_sweeps > max_or ||
        j->normalize_every < 0)
        return su2_set_error(SU2_EINVAL, "job schedule out of range (n_therm, n_meas, n_cor >= 1; or_sweeps <= %d)",
                             max_or);
    if (j->bc < SU2_BC_PERIODIC || j->bc > SU2_BC_PMC) return su2_set_error(SU2_EINVAL, "job boundary");
    if (j->algo != SU2_ALGO_HEATBATH && j->algo != SU2_ALGO_METROPOLIS) return su2_set_error(SU2_EINVAL, "job algo");
    if (j->algo == SU2_ALGO_METROPOLIS && j->hits < 1) return su2_set_error(SU2_EINVAL, "Metropolis hits < 1");
    if (!std::isfinite(j->beta)) return su2_set_error(SU2_EINVAL, "beta is not finite");
    /* the Kennedy-Pendleton loop never accepts at beta <= 0 and crawls
     * below SU2_HB_BETA_MIN */
    if (j->algo == SU2_ALGO_HEATBATH && !(j->beta >= SU2_HB_BETA_MIN))
        return su2_set_error(SU2_EINVAL, "heat-bath needs beta >= %g", (double)SU2_HB_BETA_MIN);
    if (j->algo == SU2_ALGO_METROPOLIS && !(j->epsilon >= 0.0f && j->epsilon <= 1.0f))
        return su2_set_error(SU2_EINVAL, "Metropolis needs 0 <= epsilon <= 1");
    if (j->start < SU2_START_HOT || j->start > SU2_START_HAAR) return su2_set_error(SU2_EINVAL, "job start");
    if (j->fp64_twin != 0 && j->fp64_twin != 1) return su2_set_error(SU2_EINVAL, "job fp64_twin must be 0 or 1");
    return SU2_OK;
}

struct slot_t {
    bool used = false, done = false;
    su2_job job;
    job_layout lay;
    size_t off = 0;
    su2_kgeom g;
    int total = 0, swept = 0;
    bool pend = false; /* a measurement deferred to the next tick */
};

struct tune_key {
    int prec, log2vh, max_or;
    bool operator<(const tune_key &o) const
    {
        return prec != o.prec ? prec < o.prec : (log2vh != o.log2vh ? log2vh < o.log2vh : max_or < o.max_or);
    }
};
std::mutex g_tune_mu;
std::map<tune_key, std::pair<int, int>> g_tune_cache;

/* Host copies of the device schedule tests (su2_batch.hip) on a job with s
 * sweeps done. */
int nevery_of(const su2_job &j) { return j.normalize_every ? j.normalize_every : 100; }

bool h_measure_now(const su2_job &j, int s)
{
    if (s < j.n_therm) return false;
    return (s - j.n_therm) % j.n_cor == j.n_cor - 1;
}

bool h_normalize_pre(const su2_job &j, int s) { return s < j.n_therm && (s + 1) % nevery_of(j) == 0; }

/* a measurement after sweep s is deferred to the next tick */
bool h_deferred(int s, int total, bool carry_on) { return carry_on && s + 1 < total; }

bool h_normalize_post(const su2_job &j, int s)
{
    if (s + 1 == j.n_therm) return true;
    if (!h_measure_now(j, s)) return false;
    return ((s - j.n_therm) / j.n_cor + 1) % nevery_of(j) == 0;
}

/* which launches a tick needs, and their grids */
struct tick_plan {
    /* meas: en_av pass 1 at the end of the tick; carry: the first update
     * launch forms the partials of the pending measurements, carry_pre: a
     * pass 1 before the sweeps does (a tick that cannot carry them); fin:
     * en_av pass 2 ends the tick; carry_on: the sweep launches run
     * SU2_EN_BLOCK threads, so measurements may be deferred */
    bool pre = false, post = false, meas = false, carry = false, carry_pre = false, fin = false, carry_on = false;
    /* end: no launch follows the sweeps (no end-of-tick reunitarisation or
     * en_av), so the last sweep launch ends the tick in leading workgroups
     * instead of a launch of its own (a function of carry_on, post and
     * meas, which the key holds) */
    bool end = false;
    int n_or = 0;
    int algo = SU2_BALGO_MIXED;
    uint32_t g_sw = 0, g_nm = 0, g_en = 0, g_fin = 0; /* workgroups per launch kind */
    uint64_t key() const
    {
        auto lg = [](uint32_t v) {
            uint64_t l = 0;
            while ((1u << l) < v) ++l;
            return l;
        };
        const uint64_t flags =
            ((((((((uint64_t)(n_or * 3 + algo) * 2 + pre) * 2 + post) * 2 + meas) * 2 + carry) * 2 + carry_pre) * 2 +
              fin) *
                 2 +
             carry_on);
        return (flags << 24) | (lg(g_sw) << 18) | (lg(g_nm) << 12) | (lg(g_en) << 6) | lg(g_fin);
    }
};

/* The grid of a persistent launch over nblk virtual blocks: the power of two
 * (at least 8, one workgroup per XCD) that covers them, capped at `cap`.  A
 * light load then dispatches tens of workgroups instead of thousands that
 * would exit at once; the grid class is part of the plan, so each class has
 * its own captured graph. */
uint32_t fit_grid(uint32_t nblk, uint32_t cap)
{
    uint32_t g = 8;
    while (g < nblk && g < cap) g <<= 1;
    return std::min(g, cap);
}

} // namespace

struct su2_engine {
    su2_engine_config cfg{};
    hipStream_t st = nullptr;
    char *arena = nullptr;
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> free_;
    std::vector<slot_t> slots;
    su2_bhead *dH = nullptr;
    su2_bdesc *dD = nullptr;
    uint32_t *dctr = nullptr;
    uint32_t *dpend = nullptr; /* per slot: a measurement deferred to the next tick */
    uint32_t *dB = nullptr; /* range starts: sweep, en_av, normalise; max_jobs each */
    su2_bhead *hH = nullptr;
    su2_bdesc *hD = nullptr;
    uint32_t *hB = nullptr;
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false, dirty = true;
    std::map<uint64_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs; /* by tick_plan::key */
    int block = 256, grid = 1024, aux_grid = 1024, en_grid = 2048, ncu = 256;
    int n_active = 0;

    bool alloc(size_t n, size_t *off)
    {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const size_t o = it->first, rem = it->second - n;
                free_.erase(it);
                if (rem) free_[o + n] = rem;
                used += n;
                *off = o;
                return true;
            }
        }
        return false;
    }

    void release(size_t off, size_t n)
    {
        n = align256(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

    int rebuild()
    {
        if (upload_pending) {
            ENG_HIP(hipEventSynchronize(upload_done));
            upload_pending = false;
        }
        uint32_t n = 0, sw = 0, en = 0, nm = 0;
        for (size_t i = 0; i < slots.size(); ++i) {
            slot_t &s = slots[i];
            if (!s.used || s.done) continue;
            su2_bdesc &d = hD[n++];
            std::memset(&d, 0, sizeof d);
            d.U = arena + s.off + s.lay.U;
            d.rng = (uint32_t *)(arena + s.off + s.lay.rng);
            d.accept = (int *)(arena + s.off + s.lay.acc);
            d.part = (double *)(arena + s.off + s.lay.part);
            d.partc = (double *)(arena + s.off + s.lay.partc);
            d.resE = (double *)(arena + s.off + s.lay.resE);
            d.resP = (double *)(arena + s.off + s.lay.resP);
            d.g = s.g;
            d.ph.beta = (double)s.job.beta;
            d.ph.eps = (double)s.job.epsilon;
            d.ph.hits = s.job.hits;
            d.bc = s.job.bc;
            d.algo = s.job.algo;
            d.slot = (int)i;
            d.n_therm = s.job.n_therm;
            d.n_cor = s.job.n_cor;
            d.n_meas = s.job.n_meas;
            d.nevery = s.job.normalize_every ? s.job.normalize_every : 100;
            d.or_sweeps = s.job.or_sweeps;
            d.sw_begin = sw;
            sw += (s.g.Vh + block - 1) / block;
            d.en_begin = en;
            d.en_nb = s.lay.en_nb;
            en += s.lay.en_nb;
            d.nm_begin = nm;
            nm += (6u * s.g.Vh + 255) / 256;
            const size_t mj = (size_t)cfg.max_jobs;
            hB[n - 1] = d.sw_begin;
            hB[mj + n - 1] = d.en_begin;
            hB[2 * mj + n - 1] = d.nm_begin;
        }
        hH->n_active = n;
        hH->sw_blocks = sw;
        hH->en_blocks = en;
        hH->nm_blocks = nm;
        n_active = (int)n;
        ENG_HIP(hipMemcpyAsync(dH, hH, sizeof(su2_bhead), hipMemcpyHostToDevice, st));
        if (n) {
            ENG_HIP(hipMemcpyAsync(dD, hD, n * sizeof(su2_bdesc), hipMemcpyHostToDevice, st));
            ENG_HIP(hipMemcpyAsync(dB, hB, 3 * sizeof(uint32_t) * cfg.max_jobs, hipMemcpyHostToDevice, st));
        }
        ENG_HIP(hipEventRecord(upload_done, st));
        upload_pending = true;
        dirty = false;
        return SU2_OK;
    }

    /* the launches the next tick needs, from the host sweep counters */
    tick_plan plan() const
    {
        tick_plan p;
        bool hb = false, metro = false;
        p.carry_on = block == (int)SU2_EN_BLOCK;
        bool pend = false;
        for (const auto &s : slots) {
            if (!s.used || s.done) continue;
            p.pre |= h_normalize_pre(s.job, s.swept);
            p.post |= h_normalize_post(s.job, s.swept);
            const bool now = h_measure_now(s.job, s.swept);
            p.meas |= now && !h_deferred(s.swept, s.total, p.carry_on);
            p.fin |= now || s.pend;
            pend |= s.pend;
            p.n_or = std::max(p.n_or, s.job.or_sweeps);
            (s.job.algo == SU2_ALGO_HEATBATH ? hb : metro) = true;
        }
        p.carry = pend && p.carry_on;
        p.carry_pre = pend && !p.carry_on;
        p.end = p.carry_on && !p.post && !p.meas;
        p.algo = hb && metro ? SU2_BALGO_MIXED : (hb ? SU2_BALGO_HB : SU2_BALGO_METRO);
        p.g_sw = fit_grid(hH->sw_blocks, (uint32_t)grid);
        p.g_nm = fit_grid(hH->nm_blocks, (uint32_t)aux_grid);
        p.g_en = fit_grid(hH->en_blocks, (uint32_t)en_grid);
        p.g_fin = fit_grid(hH->n_active, (uint32_t)cfg.max_jobs);
        return p;
    }

    int enqueue_tick(const tick_plan &p)
    {
        const int f32 = cfg.precision == SU2_PREC_F32;
        const uint32_t mj = (uint32_t)cfg.max_jobs;
        hipError_t e = hipSuccess;
        /* the pending measurements, of the field as the last tick left it */
        if (p.carry_pre) e = su2_launch_ben_av_partial(f32, dH, dD, dB, mj, dctr, dpend, p.g_en, p.carry_on, 1, st);
        if (p.pre && e == hipSuccess) e = su2_launch_bnormalize(f32, dH, dD, dB, mj, dctr, 0, p.g_nm, st);
        /* the tick's last sweep launch, (t, parity 1) of the last pass, ends
         * the tick when nothing follows it (p.end) */
        const su2_btick_end te = {dctr, dpend, (p.g_fin + 7u) & ~7u, p.carry_on};
        for (int mu = 0; mu < 3 && e == hipSuccess; ++mu)
            for (int par = 0; par < 2 && e == hipSuccess; ++par) {
                const bool last = p.end && p.n_or == 0 && mu == SU2_DIR_T && par == 1;
                e = su2_launch_bsweep(f32, mu, 0, p.algo, dH, dD, dB, mj, par, 0, p.g_sw, block, st, dpend,
                                      p.carry && mu == 0 && par == 0, last ? &te : nullptr);
            }
        for (int pass = 0; pass < p.n_or && e == hipSuccess; ++pass)
            for (int mu = 0; mu < 3 && e == hipSuccess; ++mu)
                for (int par = 0; par < 2 && e == hipSuccess; ++par) {
                    const bool last = p.end && pass == p.n_or - 1 && mu == SU2_DIR_T && par == 1;
                    e = su2_launch_bsweep(f32, mu, 1, p.algo, dH, dD, dB, mj, par, pass, p.g_sw, block, st, nullptr,
                                          0, last ? &te : nullptr);
                }
        if (p.post && e == hipSuccess) e = su2_launch_bnormalize(f32, dH, dD, dB, mj, dctr, 1, p.g_nm, st);
        if (p.meas && e == hipSuccess)
            e = su2_launch_ben_av_partial(f32, dH, dD, dB, mj, dctr, dpend, p.g_en, p.carry_on, 0, st);
        if (!p.end && p.fin && e == hipSuccess)
            e = su2_launch_ben_av_final(dH, dD, dctr, dpend, p.g_fin, p.carry_on, st);
        if (!p.end && !p.fin && e == hipSuccess) e = su2_launch_btick_end(dH, dD, dctr, st);
        if (e != hipSuccess) return su2_set_error(SU2_EHIP, "tick launch: %s", hipGetErrorString(e));
        return SU2_OK;
    }

    void drop_graph()
    {
        for (auto &kv : graphs) {
            if (kv.second.second) hipGraphExecDestroy(kv.second.second);
            if (kv.second.first) hipGraphDestroy(kv.second.first);
        }
        graphs.clear();
    }

    int capture(const tick_plan &p, hipGraphExec_t *out)
    {
        ENG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_tick(p);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc != SU2_OK) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return su2_set_error(SU2_EHIP, "graph capture: %s", hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            return su2_set_error(SU2_EHIP, "graph instantiate: %s", hipGetErrorString(e));
        }
        graphs[p.key()] = {g, x};
        *out = x;
        return SU2_OK;
    }

    /* one tick; host counters advance, finished jobs drop out of the list */
    int tick_once(bool *any)
    {
        if (dirty) {
            int rc = rebuild();
            if (rc) return rc;
        }
        *any = n_active > 0;
        if (!n_active) return SU2_OK;
        const tick_plan p = plan();
        if (cfg.use_graph) {
            hipGraphExec_t x = nullptr;
            auto it = graphs.find(p.key());
            if (it != graphs.end()) {
                x = it->second.second;
            } else {
                if (graphs.size() >= 64) {
                    /* bound the cache: retire every graph once the queued ones ran */
                    ENG_HIP(hipStreamSynchronize(st));
                    drop_graph();
                }
                int rc = capture(p, &x);
                if (rc) return rc;
            }
            ENG_HIP(hipGraphLaunch(x, st));
        } else {
            int rc = enqueue_tick(p);
            if (rc) return rc;
        }
        for (auto &s : slots) {
            if (!s.used || s.done) continue;
            s.pend = h_measure_now(s.job, s.swept) && h_deferred(s.swept, s.total, p.carry_on);
            if (++s.swept == s.total) {
                s.done = true;
                dirty = true;
            }
        }
        return SU2_OK;
    }
};

extern "C" {

size_t su2_job_bytes(const su2_job *job, int precision)
{
    if (!job || check_job(job, 1 << 20) != SU2_OK) return 0;
    return layout_of(*job, precision).total;
}

int su2_engine_create(const su2_engine_config *cfg, su2_engine **out)
{
    if (!cfg || !out) return su2_set_error(SU2_EINVAL, "NULL");
    *out = nullptr;
    if (cfg->precision != SU2_PREC_F64 && cfg->precision != SU2_PREC_F32)
        return su2_set_error(SU2_EINVAL, "precision");
    if (cfg->block && (cfg->block < 64 || cfg->block > 256 || cfg->block % 64))
        return su2_set_error(SU2_EINVAL, "block must be 64, 128, 192 or 256");
    if (cfg->max_jobs < 0 || cfg->max_jobs > 65536 || cfg->max_or < 0 || cfg->grid < 0)
        return su2_set_error(SU2_EINVAL, "engine limits out of range");
    ENG_HIP(hipSetDevice(cfg->device));
    su2_engine *e = new su2_engine();
    e->cfg = *cfg;
    if (!e->cfg.max_jobs) e->cfg.max_jobs = 256;
    if (!e->cfg.max_or) e->cfg.max_or = 1;
    if (!e->cfg.budget_bytes) e->cfg.budget_bytes = (size_t)1 << 30;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0)
        e->ncu = prop.multiProcessorCount;
    e->block = cfg->block ? cfg->block : 256;
    /* the sweep grid is a cap: a tick launches the power of two that covers
     * its virtual blocks, up to this (served burst of 200 jobs: 16 per CU
     * +2 % over 4 per CU, light loads unchanged) */
    e->grid = cfg->grid ? cfg->grid : 16 * e->ncu;
    e->aux_grid = 4 * e->ncu;
    /* en_av: up to one workgroup per virtual block (fit_grid takes the power
     * of two that covers them); a persistent walk of several blocks per
     * workgroup waits on each block's descriptor and counter loads in turn
     * (measured on the 64-lattice ensemble: FP64 +2 % against 8 per CU) */
    e->en_grid = 64 * e->ncu;
    int rc = SU2_OK;
    auto bail = [&](int code) {
        su2_engine_destroy(e);
        return code;
    };
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) return bail(SU2_EHIP);
    /* the arena is the budget, or 90 % of the free device memory if that is
     * smaller */
    size_t freeb = 0, totalb = 0, cap = e->cfg.budget_bytes;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && freeb > 0) cap = std::min(cap, freeb / 10 * 9);
    cap &= ~(size_t)255;
    if (hipMalloc((void **)&e->arena, cap) != hipSuccess)
        return bail(su2_set_error(SU2_ENOMEM, "arena of %zu bytes", cap));
    e->cap = cap;
    e->free_[0] = e->cap;
    e->slots.resize((size_t)e->cfg.max_jobs);
    if (hipMalloc((void **)&e->dH, sizeof(su2_bhead)) != hipSuccess ||
        hipMalloc((void **)&e->dD, sizeof(su2_bdesc) * e->cfg.max_jobs) != hipSuccess ||
        hipMalloc((void **)&e->dctr, sizeof(uint32_t) * e->cfg.max_jobs) != hipSuccess ||
        hipMalloc((void **)&e->dpend, sizeof(uint32_t) * e->cfg.max_jobs) != hipSuccess ||
        hipMalloc((void **)&e->dB, 3 * sizeof(uint32_t) * e->cfg.max_jobs) != hipSuccess ||
        hipHostMalloc((void **)&e->hH, sizeof(su2_bhead), 0) != hipSuccess ||
        hipHostMalloc((void **)&e->hD, sizeof(su2_bdesc) * e->cfg.max_jobs, 0) != hipSuccess ||
        hipHostMalloc((void **)&e->hB, 3 * sizeof(uint32_t) * e->cfg.max_jobs, 0) != hipSuccess ||
        hipEventCreateWithFlags(&e->upload_done, hipEventDisableTiming) != hipSuccess)
        return bail(su2_set_error(SU2_ENOMEM, "engine tables"));
    if (hipMemsetAsync(e->dctr, 0, sizeof(uint32_t) * e->cfg.max_jobs, e->st) != hipSuccess ||
        hipMemsetAsync(e->dpend, 0, sizeof(uint32_t) * e->cfg.max_jobs, e->st) != hipSuccess)
        return bail(SU2_EHIP);
    rc = e->rebuild();
    if (rc) return bail(rc);
    *out = e;
    return SU2_OK;
}

int su2_engine_destroy(su2_engine *e)
{
    if (!e) return SU2_OK;
    if (e->st) hipStreamSynchronize(e->st);
    e->drop_graph();
    if (e->upload_done) hipEventDestroy(e->upload_done);
    if (e->hB) hipHostFree(e->hB);
    if (e->dB) hipFree(e->dB);
    if (e->hD) hipHostFree(e->hD);
    if (e->hH) hipHostFree(e->hH);
    if (e->dctr) hipFree(e->dctr);
    if (e->dpend) hipFree(e->dpend);
    if (e->dD) hipFree(e->dD);
    if (e->dH) hipFree(e->dH);
    if (e->arena) hipFree(e->arena);
    if (e->st) hipStreamDestroy(e->st);
    delete e;
    return SU2_OK;
}

int su2_engine_add(su2_engine *e, const su2_job *job, int *slot)
{
    if (!e || !slot) return su2_set_error(SU2_EINVAL, "NULL");
    int rc = check_job(job, e->cfg.max_or);
    if (rc) return rc;
    int id = -1;
    for (size_t i = 0; i < e->slots.size(); ++i)
        if (!e->slots[i].used) {
            id = (int)i;
            break;
        }
    if (id < 0) return su2_set_error(SU2_EBUDGET, "all %d job slots are in use", e->cfg.max_jobs);
    job_layout lay = layout_of(*job, e->cfg.precision);
    size_t off;
    if (!e->alloc(lay.total, &off))
        return su2_set_error(SU2_EBUDGET, "job needs %zu bytes; %zu of %zu in use", lay.total, e->used, e->cap);
    slot_t &s = e->slots[(size_t)id];
    s.used = true;
    s.done = false;
    s.job = *job;
    s.lay = lay;
    s.off = off;
    s.g = su2_make_kgeom(job->Lt, job->Lx, job->Ly);
    s.total = job->n_therm + job->n_meas * job->n_cor;
    s.swept = 0;
    s.pend = false;
    char *base = e->arena + off;
    const int f32 = e->cfg.precision == SU2_PREC_F32;
    /* setup_kernel, the start of kernel.cu:85-136, zeroed counters/series */
    hipError_t he = su2_launch_setup_rng((uint32_t *)(base + lay.rng), s.g, job->seed, e->st);
    int start = job->start == SU2_START_COLD ? 1
                : job->start == SU2_START_HAAR ? (job->bc == SU2_BC_PCC ? 4 : 3)
                                                : (job->bc == SU2_BC_PCC ? 2 : 0);
    if (he == hipSuccess) he = su2_launch_init(start, f32, base + lay.U, (uint32_t *)(base + lay.rng), s.g, e->st);
    if (he == hipSuccess) he = su2_launch_fill_i32((int *)(base + lay.acc), 0, 2 * (size_t)s.g.Vh, e->st);
    /* the carried partials past the last sweep block stay zero (the sweep
     * launches write only blocks that hold half-sites) */
    if (he == hipSuccess) he = hipMemsetAsync(base + lay.partc, 0, lay.total - lay.partc, e->st);
    if (he == hipSuccess) he = hipMemsetAsync(e->dctr + id, 0, sizeof(uint32_t), e->st);
    if (he == hipSuccess) he = hipMemsetAsync(e->dpend + id, 0, sizeof(uint32_t), e->st);
    if (he != hipSuccess) {
        e->release(off, lay.total);
        s.used = false;
        return su2_set_error(SU2_EHIP, "job set-up: %s", hipGetErrorString(he));
    }
    e->dirty = true;
    *slot = id;
    return SU2_OK;
}

int su2_engine_tick(su2_engine *e, int n)
{
    if (!e || n < 0) return su2_set_error(SU2_EINVAL, "bad tick request");
    for (int i = 0; i < n; ++i) {
        bool any;
        int rc = e->tick_once(&any);
        if (rc) return rc;
        if (!any) break;
    }
    return SU2_OK;
}

int su2_engine_run(su2_engine *e)
{
    if (!e) return su2_set_error(SU2_EINVAL, "NULL");
    for (;;) {
        bool any;
        int rc = e->tick_once(&any);
        if (rc) return rc;
        if (!any) break;
    }
    ENG_HIP(hipStreamSynchronize(e->st));
    return SU2_OK;
}

int su2_engine_sync(su2_engine *e)
{
    if (!e) return su2_set_error(SU2_EINVAL, "NULL");
    ENG_HIP(hipStreamSynchronize(e->st));
    return SU2_OK;
}

int su2_engine_job_status(su2_engine *e, int slot, int *done_sweeps, int *total_sweeps)
{
    if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[(size_t)slot].used)
        return su2_set_error(SU2_EINVAL, "no such job slot");
    const slot_t &s = e->slots[(size_t)slot];
    if (done_sweeps) *done_sweeps = s.swept;
    if (total_sweeps) *total_sweeps = s.total;
    return SU2_OK;
}

} /* extern "C" */

/* Copies a job's results.  With sync = false the caller guarantees that the
 * tick that finished the job has completed (later ticks never touch a
 * finished job), so the engine stream is not waited on. */
static int engine_read_impl(su2_engine *e, int slot, double *E, double *P, double *accept, double *field_ref,
                            uint32_t *rng, bool sync)
{
    if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[(size_t)slot].used)
        return su2_set_error(SU2_EINVAL, "no such job slot");
    const slot_t &s = e->slots[(size_t)slot];
    char *base = e->arena + s.off;
    if (sync || field_ref) ENG_HIP(hipStreamSynchronize(e->st));
    const size_t nm = (size_t)s.job.n_meas;
    if (E) ENG_HIP(hipMemcpy(E, base + s.lay.resE, nm * 8, hipMemcpyDeviceToHost));
    if (P) ENG_HIP(hipMemcpy(P, base + s.lay.resP, nm * 8, hipMemcpyDeviceToHost));
    if (accept) {
        std::vector<int> a(2 * (size_t)s.g.Vh);
        ENG_HIP(hipMemcpy(a.data(), base + s.lay.acc, a.size() * 4, hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (int v : a) sum += v;
        const int meas_sweeps = std::max(0, s.swept - s.job.n_therm);
        const int hits = s.job.algo == SU2_ALGO_METROPOLIS ? s.job.hits : 1;
        const double n = (double)meas_sweeps * 3.0 * (hits + s.job.or_sweeps);
        double m = n > 0 ? sum / (double)a.size() / n : 0.0;
        if (s.job.bc == SU2_BC_PCC) m = m / ((double)(s.g.Ly - 1) / s.g.Ly + 1 / (double)s.g.Ly / 3); /* kernel.cu:226-228 */
        *accept = m;
    }
    if (field_ref) {
        double *tmp = nullptr;
        ENG_HIP(hipMalloc((void **)&tmp, (size_t)s.g.Vh * 24 * 8));
        hipError_t he = su2_launch_to_ref(e->cfg.precision == SU2_PREC_F32, base + s.lay.U, tmp, s.g, e->st);
        if (he == hipSuccess) he = hipStreamSynchronize(e->st);
        if (he == hipSuccess) he = hipMemcpy(field_ref, tmp, (size_t)s.g.Vh * 24 * 8, hipMemcpyDeviceToHost);
        hipFree(tmp);
        if (he != hipSuccess) return su2_set_error(SU2_EHIP, "field read: %s", hipGetErrorString(he));
    }
    if (rng) ENG_HIP(hipMemcpy(rng, base + s.lay.rng, (size_t)s.g.Vh * 24, hipMemcpyDeviceToHost));
    return SU2_OK;
}

extern "C" {

int su2_engine_read(su2_engine *e, int slot, double *E, double *P, double *accept, double *field_ref, uint32_t *rng)
{
    return engine_read_impl(e, slot, E, P, accept, field_ref, rng, true);
}

int su2_engine_remove(su2_engine *e, int slot)
{
    if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[(size_t)slot].used)
        return su2_set_error(SU2_EINVAL, "no such job slot");
    slot_t &s = e->slots[(size_t)slot];
    if (!s.done) e->dirty = true; /* an unfinished job leaves the work list */
    e->release(s.off, s.lay.total);
    s = slot_t();
    return SU2_OK;
}

int su2_engine_memory(su2_engine *e, size_t *used, size_t *capacity)
{
    if (!e) return su2_set_error(SU2_EINVAL, "NULL");
    if (used) *used = e->used;
    if (capacity) *capacity = e->cap;
    return SU2_OK;
}

int su2_engine_set_shape(su2_engine *e, int block, int grid)
{
    if (!e) return su2_set_error(SU2_EINVAL, "NULL");
    if (block < 64 || block > 256 || block % 64) return su2_set_error(SU2_EINVAL, "block must be 64, 128, 192 or 256");
    if (grid < 0) return su2_set_error(SU2_EINVAL, "grid must be >= 0");
    /* ticks enqueued before may still run their graphs */
    ENG_HIP(hipStreamSynchronize(e->st));
    e->block = block;
    if (grid) e->grid = grid;
    e->dirty = true;
    e->drop_graph();
    return SU2_OK;
}

int su2_engine_autotune(su2_engine *e, int *block, int *grid)
{
    if (!e) return su2_set_error(SU2_EINVAL, "NULL");
    if (e->dirty) {
        int rc = e->rebuild();
        if (rc) return rc;
    }
    uint64_t vh = 0;
    for (auto &s : e->slots)
        if (s.used && !s.done) vh += s.g.Vh;
    if (!vh) return su2_set_error(SU2_EINVAL, "autotune needs resident jobs");
    int lg = 0;
    while ((1ull << (lg + 1)) <= vh) ++lg;
    tune_key key{e->cfg.precision, lg, e->cfg.max_or};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            /* ticks enqueued before may still run their graphs */
            ENG_HIP(hipStreamSynchronize(e->st));
            e->block = it->second.first;
            e->grid = it->second.second;
            e->dirty = true;
            e->drop_graph();
            if (block) *block = e->block;
            if (grid) *grid = e->grid;
            return SU2_OK;
        }
    }
    /* Trials run real update sweeps, so the resident state is saved first and
     * restored afterwards, also when a trial fails: tuning leaves every job's
     * results unchanged. */
    size_t hi = 0;
    for (auto &s : e->slots)
        if (s.used) hi = std::max(hi, s.off + s.lay.total);
    char *backup = nullptr;
    ENG_HIP(hipMalloc((void **)&backup, hi));
    hipEvent_t a = nullptr, b = nullptr;
    const int keep_block = e->block, keep_grid = e->grid;
    int bb = keep_block, bg = keep_grid;
    hipError_t he = hipMemcpyAsync(backup, e->arena, hi, hipMemcpyDeviceToDevice, e->st);
    if (he == hipSuccess) he = hipEventCreate(&a);
    if (he == hipSuccess) he = hipEventCreate(&b);
    int rc = he == hipSuccess ? SU2_OK : su2_set_error(SU2_EHIP, "autotune set-up: %s", hipGetErrorString(he));
    const int f32 = e->cfg.precision == SU2_PREC_F32;
    const int algo = e->plan().algo;
    const int blocks[3] = {64, 128, 256};
    const int grids[6] = {2 * e->ncu, 4 * e->ncu, 8 * e->ncu, 16 * e->ncu, 32 * e->ncu, 64 * e->ncu};
    float best = 1e30f;
    for (int bi = 0; bi < 3 && rc == SU2_OK; ++bi)
        for (int gi = 0; gi < 6 && rc == SU2_OK; ++gi) {
            e->block = blocks[bi];
            e->grid = grids[gi];
            e->dirty = true;
            rc = e->rebuild();
            float tot = 0.0f;
            for (int rep = 0; rep < 3 && rc == SU2_OK; ++rep) {
                he = hipEventRecord(a, e->st);
                for (int mu = 0; mu < 3 && he == hipSuccess; ++mu)
                    for (int par = 0; par < 2 && he == hipSuccess; ++par)
                        he = su2_launch_bsweep(f32, mu, 0, algo, e->dH, e->dD, e->dB, (uint32_t)e->cfg.max_jobs, par,
                                               0, (uint32_t)e->grid, (uint32_t)e->block, e->st);
                if (he == hipSuccess) he = hipEventRecord(b, e->st);
                if (he == hipSuccess) he = hipEventSynchronize(b);
                float ms = 0.0f;
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
                if (he != hipSuccess) rc = su2_set_error(SU2_EHIP, "autotune trial: %s", hipGetErrorString(he));
                if (rep) tot += ms;
            }
            if (rc == SU2_OK && tot < best) {
                best = tot;
                bb = e->block;
                bg = e->grid;
            }
        }
    /* restore the resident state whatever happened above */
    he = hipMemcpyAsync(e->arena, backup, hi, hipMemcpyDeviceToDevice, e->st);
    if (he == hipSuccess) he = hipStreamSynchronize(e->st);
    hipFree(backup);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (he != hipSuccess && rc == SU2_OK)
        rc = su2_set_error(SU2_EHIP, "autotune restore: %s", hipGetErrorString(he));
    if (rc != SU2_OK) {
        e->block = keep_block;
        e->grid = keep_grid;
        e->dirty = true;
        return rc;
    }
    e->block = bb;
    e->grid = bg;
    e->dirty = true;
    e->drop_graph();
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {bb, bg};
    }
    if (block) *block = bb;
    if (grid) *grid = bg;
    return SU2_OK;
}

} /* extern "C" */

/* ------------------------------------------------------------------------
 * Server
 * ---------------------------------------------------------------------- */
namespace {

struct pending_job {
    long long id;
    su2_job job;
    double t_submit;
};

struct running_job {
    long long id;
    int slot;
    int slot64; /* the FP64 twin's slot in the server's FP64 engine, or -1 */
    double t_submit, t_admit;
    su2_job job;
};

struct finished_job {
    su2_job_result res;
    std::vector<double> E, P;
};

double mean_err(const std::vector<double> &m, double *err)
{
    /* mean_variance (functions/statistics.h:30-39) */
    const size_t n = m.size();
    double s = 0.0, s2 = 0.0;
    for (double v : m) {
        s += v;
        s2 += v * v;
    }
    const double mean = n ? s / n : 0.0;
    const double var = n ? s2 / n - mean * mean : 0.0;
    *err = n > 1 ? std::sqrt(std::max(var, 0.0) / (n - 1)) : 0.0;
    return mean;
}

} // namespace

struct su2_server {
    su2_engine *e = nullptr;
    su2_engine *e64 = nullptr; /* FP64 twins of an FP32-link server; made on the first twin job */
    hipEvent_t ev64[2] = {nullptr, nullptr};
    su2_engine_config cfg{};
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, idle_cv;
    int created = 0; /* 0 pending, 1 engine ready, -1 creation failed; guarded by mu */
    int create_rc = SU2_OK;
    std::string create_msg;
    /* the engine's limits, copied at creation: submit checks jobs with them
     * and never touches the engine, which the server thread owns */
    int max_or = 1, precision = SU2_PREC_F64;
    size_t cap = 0;
    std::deque<pending_job> queue;        /* guarded by mu */
    std::map<long long, finished_job> finished; /* guarded by mu */
    std::vector<running_job> running;     /* server thread only */
    int inflight = 0;                     /* submitted and not yet retired; guarded by mu */
    long long next_id = 1;
    bool stop = false;
    int error = SU2_OK;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    void publish(finished_job &&f)
    {
        std::lock_guard<std::mutex> lk(mu);
        finished[f.res.id] = std::move(f);
        --inflight;
        idle_cv.notify_all();
    }

    /* the tick that finished r has completed */
    void retire(const running_job &r)
    {
        finished_job f;
        f.res = su2_job_result();
        f.res.id = r.id;
        f.res.t_submit = r.t_submit;
        f.res.t_admit = r.t_admit;
        f.E.resize((size_t)r.job.n_meas);
        f.P.resize((size_t)r.job.n_meas);
        double acc = 0.0;
        int rc = engine_read_impl(e, r.slot, f.E.data(), f.P.data(), &acc, nullptr, nullptr, false);
        su2_engine_remove(e, r.slot);
        if (r.slot64 >= 0) {
            std::vector<double> P64((size_t)r.job.n_meas);
            int rc64 = engine_read_impl(e64, r.slot64, nullptr, P64.data(), nullptr, nullptr, nullptr, false);
            su2_engine_remove(e64, r.slot64);
            const int n = r.job.n_meas, nb = n < 20 ? n : 20;
            if (rc64 == SU2_OK && nb >= 2) {
                double m32 = 0.0;
                rc64 = su2_binned_mean_error(f.P.data(), n, nb, &m32, &f.res.P_err_binned);
                if (rc64 == SU2_OK)
                    rc64 = su2_binned_mean_error(P64.data(), n, nb, &f.res.P_fp64_twin, &f.res.P_fp64_twin_err);
                if (rc64 == SU2_OK)
                    rc64 = su2_series_compare(f.P.data(), P64.data(), n, nb, &f.res.P_diff, &f.res.P_diff_sigma,
                                              nullptr);
            }
            if (rc == SU2_OK) rc = rc64;
        }
        f.res.status = rc;
        f.res.E_mean = mean_err(f.E, &f.res.E_err);
        f.res.P_mean = mean_err(f.P, &f.res.P_err);
        f.res.accept = acc;
        f.res.link_updates = 3ll * r.job.Lt * r.job.Lx * r.job.Ly *
                             ((long long)r.job.n_therm + (long long)r.job.n_meas * r.job.n_cor) * (1 + r.job.or_sweeps);
        f.res.t_done = now();
        publish(std::move(f));
    }

    /* records a fatal error: the loop stops and drain/poll report it */
    void fail(int rc)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (error == SU2_OK) error = rc;
        stop = true;
        idle_cv.notify_all();
    }

    /* The server thread creates the engine (its arena, tables and stream),
     * uses it and, after the loop, destroys it: every HIP object of the
     * engine is made and used by this one thread. */
    void run()
    {
        su2_engine *eng = nullptr;
        const int rc = su2_engine_create(&cfg, &eng);
        {
            std::lock_guard<std::mutex> lk(mu);
            e = eng;
            create_rc = rc;
            if (rc != SU2_OK) {
                create_msg = su2_last_error();
            } else {
                max_or = eng->cfg.max_or;
                precision = eng->cfg.precision;
                cap = eng->cap;
            }
            created = rc == SU2_OK ? 1 : -1;
        }
        idle_cv.notify_all();
        if (rc != SU2_OK) return;
        loop();
        if (e64) {
            su2_engine_destroy(e64);
            for (hipEvent_t &x : ev64)
                if (x) hipEventDestroy(x);
        }
        su2_engine_destroy(e);
    }

    /* the FP64 engine of the twins, created on first use by this thread */
    int twin_engine()
    {
        if (e64) return SU2_OK;
        su2_engine_config c = cfg;
        c.precision = SU2_PREC_F64;
        su2_engine *x = nullptr;
        int rc = su2_engine_create(&c, &x);
        if (rc) return rc;
        hipError_t he = hipEventCreateWithFlags(&ev64[0], hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&ev64[1], hipEventDisableTiming);
        if (he != hipSuccess) {
            for (hipEvent_t &v : ev64)
                if (v) hipEventDestroy(v);
            ev64[0] = ev64[1] = nullptr;
            su2_engine_destroy(x);
            return su2_set_error(SU2_EHIP, "twin events: %s", hipGetErrorString(he));
        }
        e64 = x;
        return SU2_OK;
    }

    /* admits a job and, for an FP64 twin, its twin: both or neither */
    int admit_job(const su2_job &job, int *slot, int *slot64)
    {
        *slot64 = -1;
        const bool twin = job.fp64_twin && cfg.precision == SU2_PREC_F32;
        if (twin) {
            int rc = twin_engine();
            if (rc) return rc;
        }
        int rc = su2_engine_add(e, &job, slot);
        if (rc || !twin) return rc;
        rc = su2_engine_add(e64, &job, slot64);
        if (rc) {
            su2_engine_remove(e, *slot);
            *slot64 = -1;
        }
        return rc;
    }

    void loop()
    {
        hipEvent_t ev[2] = {nullptr, nullptr};
        hipError_t he = hipSetDevice(e->cfg.device);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
        if (he != hipSuccess) {
            fail(su2_set_error(SU2_EHIP, "server set-up: %s", hipGetErrorString(he)));
            return;
        }
        int cur = 0;
        bool prev_valid = false;
        std::vector<running_job> prev_done;
        for (;;) {
            std::vector<pending_job> admit, reject;
            std::vector<int> reject_rc;
            {
                std::unique_lock<std::mutex> lk(mu);
                /* admit in arrival order, first that fits */
                for (auto it = queue.begin(); it != queue.end();) {
                    int slot = -1, slot64 = -1;
                    int rc = admit_job(it->job, &slot, &slot64);
                    if (rc == SU2_OK) {
                        running.push_back({it->id, slot, slot64, it->t_submit, now(), it->job});
                        it = queue.erase(it);
                    } else if (rc == SU2_EBUDGET) {
                        ++it;
                    } else {
                        reject.push_back(*it);
                        reject_rc.push_back(rc);
                        it = queue.erase(it);
                    }
                }
                if (running.empty() && !prev_valid && reject.empty()) {
                    idle_cv.notify_all();
                    if (stop && queue.empty()) break;
                    cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
            for (size_t i = 0; i < reject.size(); ++i) {
                finished_job f;
                f.res = su2_job_result();
                f.res.id = reject[i].id;
                f.res.status = reject_rc[i];
                f.res.t_submit = reject[i].t_submit;
                f.res.t_admit = f.res.t_done = now();
                publish(std::move(f));
            }
            /* one tick for everything resident */
            std::vector<running_job> done_now;
            bool ticked = false;
            if (!running.empty()) {
                int rc = su2_engine_tick(e, 1);
                if (rc == SU2_OK && (he = hipEventRecord(ev[cur], e->st)) != hipSuccess)
                    rc = su2_set_error(SU2_EHIP, "server tick event: %s", hipGetErrorString(he));
                /* the FP64 twins advance in the same tick, on their own stream */
                if (rc == SU2_OK && e64) rc = su2_engine_tick(e64, 1);
                if (rc == SU2_OK && e64 && (he = hipEventRecord(ev64[cur], e64->st)) != hipSuccess)
                    rc = su2_set_error(SU2_EHIP, "server twin tick event: %s", hipGetErrorString(he));
                if (rc != SU2_OK) {
                    fail(rc);
                    break;
                }
                ticked = true;
                for (auto it = running.begin(); it != running.end();) {
                    int d = 0, t = 0, d64 = 0, t64 = 0;
                    su2_engine_job_status(e, it->slot, &d, &t);
                    if (it->slot64 >= 0) su2_engine_job_status(e64, it->slot64, &d64, &t64);
                    if (d == t && d64 == t64) {
                        done_now.push_back(*it);
                        it = running.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            /* retire the jobs the previous tick finished while this one runs */
            if (prev_valid) {
                if ((he = hipEventSynchronize(ev[cur ^ 1])) != hipSuccess ||
                    (e64 && (he = hipEventSynchronize(ev64[cur ^ 1])) != hipSuccess)) {
                    fail(su2_set_error(SU2_EHIP, "server tick wait: %s", hipGetErrorString(he)));
                    break;
                }
                for (auto &r : prev_done) retire(r);
                prev_done.clear();
                prev_valid = false;
            }
            if (ticked) {
                prev_done = std::move(done_now);
                prev_valid = true;
                cur ^= 1;
            }
        }
        if (prev_valid && error == SU2_OK) {
            if ((he = hipEventSynchronize(ev[cur ^ 1])) == hipSuccess &&
                (!e64 || (he = hipEventSynchronize(ev64[cur ^ 1])) == hipSuccess))
                for (auto &r : prev_done) retire(r);
            else
                fail(su2_set_error(SU2_EHIP, "server tick wait: %s", hipGetErrorString(he)));
        }
        hipEventDestroy(ev[0]);
        hipEventDestroy(ev[1]);
        std::lock_guard<std::mutex> lk(mu);
        idle_cv.notify_all();
    }
};

extern "C" {

int su2_server_create(const su2_engine_config *cfg, su2_server **out)
{
    if (!cfg || !out) return su2_set_error(SU2_EINVAL, "NULL");
    *out = nullptr;
    su2_server *s = new su2_server();
    s->cfg = *cfg;
    s->th = std::thread([s] { s->run(); });
    int rc;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->idle_cv.wait(lk, [s] { return s->created != 0; });
        rc = s->create_rc;
    }
    if (rc != SU2_OK) {
        s->th.join();
        const std::string msg = s->create_msg; /* recorded on the server thread */
        delete s;
        return su2_set_error(rc, "%s", msg.c_str());
    }
    *out = s;
    return SU2_OK;
}

int su2_server_destroy(su2_server *s)
{
    if (!s) return SU2_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join(); /* the thread destroys its engine */
    int rc = s->error;
    delete s;
    return rc;
}

int su2_server_submit(su2_server *s, const su2_job *job, long long *id)
{
    if (!s || !job || !id) return su2_set_error(SU2_EINVAL, "NULL");
    int rc = check_job(job, s->max_or);
    if (rc) return rc;
    if (layout_of(*job, s->precision).total > s->cap)
        return su2_set_error(SU2_EBUDGET, "job larger than the whole memory budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return su2_set_error(SU2_ECANCELLED, "server is stopping");
    *id = s->next_id++;
    ++s->inflight;
    s->queue.push_back({*id, *job, s->now()});
    s->cv.notify_all();
    return SU2_OK;
}

int su2_server_poll(su2_server *s, su2_job_result *out, int max, int *n)
{
    if (!s || !n || (max > 0 && !out)) return su2_set_error(SU2_EINVAL, "NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    int k = 0;
    for (auto it = s->finished.begin(); it != s->finished.end() && k < max;) {
        out[k++] = it->second.res;
        it = s->finished.erase(it);
    }
    *n = k;
    return s->error;
}

int su2_server_series(su2_server *s, long long id, double *E, double *P, int n)
{
    if (!s) return su2_set_error(SU2_EINVAL, "NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return su2_set_error(SU2_EINVAL, "job %lld is not finished or already polled", id);
    const int m = std::min(n, (int)it->second.E.size());
    if (E) std::memcpy(E, it->second.E.data(), (size_t)m * 8);
    if (P) std::memcpy(P, it->second.P.data(), (size_t)m * 8);
    return SU2_OK;
}

int su2_server_drain(su2_server *s)
{
    if (!s) return su2_set_error(SU2_EINVAL, "NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    s->idle_cv.wait(lk, [s] { return s->inflight == 0 || s->error != SU2_OK; });
    return s->error;
}

double su2_server_now(su2_server *s) { return s ? s->now() : 0.0; }

} /* extern "C" */
