/*
 * su2mc.h - C ABI of libsu2mc, the MI355X (gfx950) SU(2) Yang-Mills Monte
 * Carlo hot path: the Wilson-action checkerboard sweep with Kennedy-Pendleton
 * heat-bath, Metropolis and over-relaxation, periodic / PCC / PMC boundaries
 * in y, and the fused energy/plaquette reduction.
 *
 * Drop-in boundary.  fezquerro/QCD_2d_GPU has no FFI: kernel.cu:main calls
 * the header functions of functions/ directly (kernel.cu:15-22).  Each entry
 * point below replaces one of those calls; the reference function it
 * replaces is named beside it.  Conventions:
 *   - The caller allocates every device buffer (as kernel.cu:78,82,145,150
 *     does) with the sizes given by su2_*_bytes and passes raw device
 *     pointers.  su2_malloc/su2_free are provided for callers without their
 *     own HIP allocator.
 *   - All entry points take the FULL-lattice struct (the reference's `l`,
 *     kernel.cu:58), also where the reference passes the half-lattice `l2`.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the default
 *     stream) and is asynchronous unless stated otherwise; the reference
 *     uses the legacy default stream with implicit ordering.
 *   - Every entry point returns an int status: SU2_OK (0) or a negative
 *     SU2_E* code; su2_last_error() gives a message for the calling thread.
 *     (The reference returns void and checks nothing.)
 *   - Device field layout: structure of arrays, link(p, mu, h) at index
 *     (p * 3 + mu) * (V/2) + h, 4 doubles (SU2_PREC_F64) or 4 floats
 *     (SU2_PREC_F32) per link.  su2_field_to_ref/su2_field_from_ref convert
 *     to and from the reference's array-of-structures layout, which is also
 *     its Config_ file format (functions/output.h:85).
 *   - RNG state: six uint32 planes of V/2 entries, 24 B per half-site
 *     (replaces curandStateMRG32k3a_t[V/2], kernel.cu:77-78).
 */
#ifndef SU2MC_H
#define SU2MC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SU2MC_ABI_VERSION 5

/* ---- status codes ---- */
#define SU2_OK 0
#define SU2_EINVAL (-1)      /* bad argument */
#define SU2_EODD_EXTENT (-2) /* an extent is odd or < 2 (kernel.cu:61-64) */
#define SU2_ENOMEM (-3)      /* device or host allocation failed */
#define SU2_EHIP (-4)        /* a HIP runtime call failed */
#define SU2_EIO (-5)         /* file I/O failed (the reference exit(1)s) */
#define SU2_ENODEV (-6)      /* no usable gfx950 device */
#define SU2_EBUDGET (-7)     /* request exceeds the memory budget */
#define SU2_ECANCELLED (-8)  /* job was cancelled */

/* ---- enumerations (values follow kernel.cu:11-13,28,31) ---- */
#define SU2_ALGO_METROPOLIS 0 /* kernel.cu:11 METROPOLIS */
#define SU2_ALGO_HEATBATH 1   /* kernel.cu:12 HEATBATH */
#define SU2_UPDATE_LINK 0     /* mc_updates.h:178 update == 0: link_update */
#define SU2_UPDATE_OVERRELAX 1 /* mc_updates.h:180 update == 1: over_relaxation */
#define SU2_BC_PERIODIC 0     /* kernel.cu:28 Boundary 0 */
#define SU2_BC_PCC 1          /* Boundary 1: perfect colour conductor */
#define SU2_BC_PMC 2          /* Boundary 2: perfect magnetic conductor */
#define SU2_START_HOT 0       /* kernel.cu:31 init 0: rand_init2 / rand_init2_pcc */
#define SU2_START_COLD 1      /* init 1: id_init2 */
#define SU2_START_HAAR 2      /* extension: Haar-random links */
#define SU2_PREC_F64 0        /* links stored as 4 doubles */
#define SU2_PREC_F32 1        /* links stored as 4 floats; all arithmetic FP64 */

/*
 * Lattice parameters.  Same members, types and order as the reference's
 * struct lattice (functions/lattice_setup.h:2-14), holding the FULL lattice:
 * y, x, t = Ly, Lx, Lt; V = Lt*Lx*Ly; d = 3; dg = 4; dt = 12; beta and
 * epsilon are float as in the reference; updates = Metropolis hits per link.
 */
typedef struct su2_lattice {
    int y;
    int x;
    int t;
    int V;
    int d;
    int dg;
    int dt;
    float beta;
    float epsilon;
    int updates;
} su2_lattice;

/* Fills *l for an Lt x Lx x Ly lattice (kernel.cu:58).  Returns
 * SU2_EODD_EXTENT when an extent is odd or below 2 (kernel.cu:61-64) and
 * SU2_EINVAL above SU2_MAX_LATTICE_SITES sites (3V + padding must fit the
 * kernels' 32-bit link indices and a dispatch's 2^32 work-items; this is
 * 160 GB of FP64 device state). */
#define SU2_MAX_LATTICE_SITES 1431633920ll

int su2_lattice_init(su2_lattice *l, int Lt, int Lx, int Ly, float beta, float epsilon, int updates);

/* Buffer sizes in bytes. */
size_t su2_field_bytes(const su2_lattice *l, int precision); /* kernel.cu:82  12 V reals */
size_t su2_rng_bytes(const su2_lattice *l);                  /* kernel.cu:78  V/2 streams */
size_t su2_accept_bytes(const su2_lattice *l);               /* kernel.cu:145 int[V] */
size_t su2_en_av_work_bytes(const su2_lattice *l);           /* scratch of su2_en_av */

/* ---- set-up and starts ---- */

/* setup_kernel (functions/lattice_setup.h:15-21): stream h = subsequence h
 * of MRG32k3a(seed), offset 0, for h < V/2. */
int su2_setup_kernel(const su2_lattice *l, uint32_t *states, uint64_t seed, void *stream);

/* id_init2 (functions/lattice_setup.h:38-51): every link the identity. */
int su2_id_init2(const su2_lattice *l, void *phi, int precision, void *stream);

/* rand_init2 (functions/lattice_setup.h:24-37): hot start, each link the
 * normalised vector of four (u - 0.5) draws (functions/su2.h:29-34). */
int su2_rand_init2(const su2_lattice *l, void *phi, int precision, uint32_t *states, void *stream);

/* rand_init2_pcc (functions/lattice_setup.h:52-81): as rand_init2 but the x
 * and t links at y = 0 are the identity and consume no draws. */
int su2_rand_init2_pcc(const su2_lattice *l, void *phi, int precision, uint32_t *states, void *stream);

/* Extension (no counterpart in the reference): Haar-random hot start; with
 * pcc != 0 the x and t links at y = 0 are the identity. */
int su2_haar_init(const su2_lattice *l, void *phi, int precision, uint32_t *states, int pcc, void *stream);

/* ---- the sweep ---- */

/* mc_up_per (functions/mc_updates.h:467-475): one sweep, six launches in
 * the order (y,even) (y,odd) (x,even) (x,odd) (t,even) (t,odd).
 * algo = SU2_ALGO_HEATBATH or SU2_ALGO_METROPOLIS (the reference's compile-
 * time ALGO, kernel.cu:13); update = SU2_UPDATE_LINK or SU2_UPDATE_OVERRELAX
 * (the reference's `int update`).  accept[site] += accepted hits, as the
 * six reference kernels leave it after the sweep (the heat-bath and
 * over-relaxation, which accept every update, add a site's count for the
 * whole sweep in its t launch; in between, the counters are not defined).
 * Couplings (from *l): the heat-bath needs a finite beta >=
 * SU2_HB_BETA_MIN (its Kennedy-Pendleton trial accepts with probability
 * about 1.26 (beta |A|)^1.5 for small beta |A|: some 800 trials per update at
 * 0.01, never at beta <= 0); Metropolis needs 0 <= epsilon <= 1; otherwise
 * SU2_EINVAL. */
#define SU2_HB_BETA_MIN 0.01f
int su2_mc_up_per(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream);

/* mc_up_pcc (functions/mc_updates_pcc.h:216-224): y links periodic; x and t
 * links at y = 0 not updated.  update = SU2_UPDATE_OVERRELAX is an extension
 * (the reference has no PCC over-relaxation). */
int su2_mc_up_pcc(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream);

/* mc_up_pmc (functions/mc_updates_pmc.h:249-257): y links periodic; for x
 * and t links the y-plaquettes between y = Ly-1 and y = 0 are dropped from
 * the staple.  update = SU2_UPDATE_OVERRELAX is an extension. */
int su2_mc_up_pmc(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int algo,
                  int update, void *stream);

/* One (direction, parity) launch of the above: mu 0 = y, 1 = x, 2 = t;
 * parity 0 = even, 1 = odd (one MC_update_{even,odd}_{y,x,t}[_pcc|_pmc]
 * kernel, functions/mc_updates.h:129-465); counts its own accepted updates
 * into accept. */
int su2_mc_update(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int bc, int algo,
                  int update, int mu, int parity, void *stream);

/* normalize (functions/lattice_setup.h:202-207): reunitarise all 3V links. */
int su2_normalize(const su2_lattice *l, void *phi, int precision, void *stream);

/* en_av<PMC> (functions/observables.h:99-141): *res += mean over sites of
 * -(1 - P_xt) (energy), *res2 += mean over sites and planes of 1/2 tr P
 * (average plaquette); res, res2 are device pointers, as in the reference
 * (end + i - 1, end2 + i - 1 at kernel.cu:208).  work is device scratch of
 * su2_en_av_work_bytes(l) (block partial sums).  The sum is deterministic
 * (partials added in a fixed order, no atomics), so repeated calls give
 * identical bits.  Calls that share one work buffer must not overlap. */
int su2_en_av(const su2_lattice *l, const void *phi, int precision, int pmc, double *res, double *res2, void *work,
              void *stream);

/* en_av of the field as it is (kernel.cu:208), then one sweep
 * (kernel.cu:192-202, mc_up_per / mc_up_pcc / mc_up_pmc by bc): the same
 * results, bit for bit, as su2_en_av(l, phi, precision, bc == SU2_BC_PMC,
 * res, res2, work, stream) followed by the sweep, in fewer passes over the
 * field.  The reference's run loop measures after the last sweep of a
 * measurement and sweeps again right after; a driver that holds a
 * measurement back until its next sweep makes the pair one call.  The
 * measurement rides on the sweep's first launch (y, parity 0), which reads
 * nearly every link the measurement needs before it writes, and its sum on
 * the second; with a sweep block size other than 256 threads
 * (su2_set_sweep_block) the two run separately.  work as for su2_en_av. */
int su2_mc_up_measured(const su2_lattice *l, void *phi, int precision, int *accept, uint32_t *states, int bc,
                       int algo, int update, double *res, double *res2, void *work, void *stream);

/* No counterpart in the reference: su2_en_av for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels and summation
 * order as su2_en_av, so the results are identical bit for bit, with the
 * block partial sums in the library's own static device array (2 MiB per
 * device, held until the process ends) instead of a work buffer; no
 * allocation of any kind.  Lattices of at most SU2_EN_AV_NOALLOC_MAX_SITES
 * sites, else SU2_EINVAL.  All calls on one device share that array: calls
 * that overlap on one device (several streams or host threads) race and
 * must be serialised by the caller. */
#define SU2_EN_AV_NOALLOC_MAX_SITES (1ll << 26)
int su2_en_av_noalloc(const su2_lattice *l, const void *phi, int precision, int pmc, double *res, double *res2,
                      void *stream);

/* cudaMemset(accept, 0, V * sizeof(int)) of kernel.cu:145,190, done with a
 * fill kernel. */
int su2_accept_clear(const su2_lattice *l, int *accept, void *stream);

/* mean_lattice(int*) (functions/statistics.h:17-28): copies accept to the
 * host (synchronising the stream) and returns the mean acceptance
 * sum / V / n and its error; n = sweeps * 3 * hits (kernel.cu:225,246). */
int su2_mean_lattice(const su2_lattice *l, const int *accept, long long n, double *mean, double *err, void *stream);

/* Conversion to/from the reference array-of-structures field layout
 * double[12 V] (functions/lattice_setup.h:27-32; the Config_ file format,
 * functions/output.h:85).  Both pointers are device pointers. */
int su2_field_to_ref(const su2_lattice *l, const void *phi, int precision, double *phi_ref, void *stream);
int su2_field_from_ref(const su2_lattice *l, const double *phi_ref, void *phi, int precision, void *stream);

/* ---- device plumbing ---- */
int su2_device_count(int *n);
int su2_set_device(int device);
int su2_malloc(void **ptr, size_t bytes);
int su2_free(void *ptr);
int su2_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream); /* synchronous */
int su2_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream); /* synchronous */
int su2_memset_zero(void *dst, size_t bytes, void *stream);
/* Device-to-device copy by a plain streaming kernel (16 B per thread); also
 * the reference kernel for calibrating hardware byte counters. */
int su2_device_copy(void *dst, const void *src, size_t bytes, void *stream);
/* Extension (no counterpart in the reference): streaming-bandwidth probe,
 * the measured ceiling a sweep's HBM rate is compared with.  dst[i] = XOR of
 * src[r * bytes + i] over r < reads (reads 1, 2, 4 or 5; reads = 1 is a
 * copy), 16-byte items; src holds reads regions of `bytes` each.  Every
 * thread keeps `unroll` (1, 2, 4, 8) items of each source in flight;
 * nontemporal != 0 marks loads and stores nontemporal; blocks = workgroups
 * of 256 (0: one item per thread per unroll, at most 2^20). */
int su2_stream_probe(void *dst, const void *src, size_t bytes, int reads, int unroll, int nontemporal, int blocks,
                     void *stream);
/* Extension (no counterpart in the reference): what separating two
 * dependent phases of a sweep costs on this device.  A phase is every thread
 * of `blocks` workgroups of 256 adding 1 to `words` doubles of its own and
 * then reading `words` doubles of the next workgroup, as a (direction,
 * parity) launch depends on the one before it.  The halves are separated
 * once by kernel boundaries (back-to-back launches on `stream`), once by a
 * grid-wide barrier inside one cooperative launch, and, unless
 * xcd_barrier_us is NULL, once by a barrier among only the workgroups on one
 * XCD (each finds its XCD from the hardware's XCC_ID register and pairs
 * with the next workgroup of its own XCD; the groups run side by side).
 * *launch_us, *barrier_us and *xcd_barrier_us are the mean time per
 * separation over `phases` phases.  buf holds at least blocks * 256 * words
 * doubles; blocks must all fit on the device at once.  SU2_EHIP when the
 * XCD-local barrier read stale data or one of its bounded waits ran out.
 * Synchronous. */
int su2_phase_probe(double *buf, int blocks, int words, int phases, void *stream, double *launch_us,
                    double *barrier_us, double *xcd_barrier_us);
int su2_stream_create(void **stream);
int su2_stream_destroy(void *stream);
int su2_stream_synchronize(void *stream);
int su2_event_create(void **event);
int su2_event_destroy(void *event);
int su2_event_record(void *event, void *stream);
int su2_event_elapsed_ms(void *start, void *stop, float *ms); /* synchronises on stop */
const char *su2_strerror(int status);
const char *su2_last_error(void);
int su2_abi_version(void);

/* Launch block size of the sweep kernels (threads per block, a multiple of
 * 64, at most 256); 0 restores the built-in default. */
int su2_set_sweep_block(int threads);

/* ---- handle API: one lattice with its own buffers ----
 * The same operations as above on an opaque lattice that owns its field,
 * RNG state, acceptance counters and reduction scratch, all enqueued on its
 * context's stream.  These are the calls kernel.cu:main makes on its one
 * lattice (kernel.cu:77-236) with the buffers and the cudaMalloc calls
 * (kernel.cu:78,82,145,150) folded into the handle.  A context is one
 * device and one stream; lattices of one context are stream-ordered, and a
 * context must be used by one host thread at a time.
 */
#define SU2_BACKEND_CPU 0 /* not provided: the library has no CPU fallback */
#define SU2_BACKEND_HIP 1

typedef struct su2_ctx su2_ctx;
typedef struct su2_lat su2_lat;

/* backend must be SU2_BACKEND_HIP (SU2_BACKEND_CPU returns SU2_EINVAL).
 * Creates a non-blocking stream on `device`. */
int su2_ctx_create(int backend, int device, su2_ctx **out);
int su2_ctx_destroy(su2_ctx *ctx); /* synchronises; the context's lattices must be destroyed first */
int su2_ctx_stream(su2_ctx *ctx, void **stream);
int su2_ctx_synchronize(su2_ctx *ctx);

/* Allocates an Lt x Lx x Ly lattice (kernel.cu:58,77-82,145); SU2_EODD_EXTENT
 * for odd extents (kernel.cu:61-64).  bc = SU2_BC_*, precision = SU2_PREC_*.
 * The field and RNG are uninitialised until su2_rng_init / su2_init (or
 * the uploads); the acceptance counters start at zero. */
int su2_lattice_create(su2_ctx *ctx, int Lt, int Lx, int Ly, int bc, int precision, su2_lat **out);
int su2_lattice_destroy(su2_lat *h);
/* The lattice's su2_lattice struct (beta, epsilon, updates as last swept),
 * boundary and precision; NULL skips an output. */
int su2_lattice_get(const su2_lat *h, su2_lattice *l, int *bc, int *precision);
/* Raw device buffers of the lattice, for the entry points above; NULL
 * skips an output. */
int su2_lattice_buffers(su2_lat *h, void **phi, uint32_t **states, int **accept);

/* setup_kernel (functions/lattice_setup.h:15-21, kernel.cu:79). */
int su2_rng_init(su2_lat *h, uint64_t seed);
/* The start of kernel.cu:85-136: SU2_START_COLD id_init2; SU2_START_HOT
 * rand_init2 (rand_init2_pcc under PCC); SU2_START_HAAR the Haar start.
 * Also clears the acceptance counters. */
int su2_init(su2_lat *h, int start);
/* One sweep (mc_up_per / mc_up_pcc / mc_up_pmc by the lattice's boundary,
 * kernel.cu:167,195): algo SU2_ALGO_*, update SU2_UPDATE_*; beta, eps and
 * hits as the reference's betal, epsilonl and upd. */
int su2_sweep(su2_lat *h, int algo, int update, float beta, float eps, int hits);
/* normalize (functions/lattice_setup.h:202-207) on the handle's field
 * (su2_normalize is the raw-pointer form). */
int su2_renormalize(su2_lat *h);
/* en_av<PMC> (functions/observables.h:99-141) with PMC set by the lattice's
 * boundary; synchronous: *E and *P receive the energy and the average
 * plaquette (host pointers).  Deterministic, like su2_en_av. */
int su2_measure(su2_lat *h, double *E, double *P);
/* Acceptance counters (kernel.cu:144-145,190,225): *accepted = sum over
 * sites, *attempts = V x (3 x hits for each link sweep, 3 for each
 * over-relaxation sweep) since su2_init or su2_accept_reset, the
 * denominator of kernel.cu:225,246 (no PCC correction); synchronous. */
int su2_accept_read(su2_lat *h, int64_t *accepted, int64_t *attempts);
int su2_accept_reset(su2_lat *h); /* cudaMemset of kernel.cu:190 */
/* Field in the reference layout double[12 V] on the host (the Config_ file
 * layout, functions/output.h:85, functions/lattice_setup.h:83-100);
 * synchronous. */
int su2_field_download(su2_lat *h, double *host_phi_ref);
int su2_field_upload(su2_lat *h, const double *host_phi_ref);
/* RNG state uint32[6 * V/2] on the host (six planes of V/2); synchronous.
 * Extension: the reference cannot save its RNG state. */
int su2_rng_download(su2_lat *h, uint32_t *host_states);
int su2_rng_upload(su2_lat *h, const uint32_t *host_states);

/* ---- batched lattices: one launch per (direction, parity) for many ----
 *
 * Extension (no counterpart in the reference, whose driver runs one lattice
 * per process): an engine holds many lattices of different shapes,
 * couplings, boundary conditions and seeds in one device arena and advances
 * all of them with one persistent launch per (direction, parity).  Each job
 * follows the reference's run loop (kernel.cu:161-208) on its own schedule
 * and gives bit for bit the results it gives alone.  A tick advances every
 * resident, unfinished job by one update sweep (+ its over-relaxation
 * sweeps), with its reunitarisation and measurement when due.  Ticks can be
 * replayed from captured hipGraphs (use_graph); jobs may join or leave
 * between ticks.  Calls on one engine must come from one host thread at a
 * time (the engine of a server belongs to the server's thread).
 */
typedef struct su2_job {
    int Lt, Lx, Ly;
    float beta, epsilon;
    int hits;            /* Metropolis hits per link */
    int bc;              /* SU2_BC_* */
    int algo;            /* SU2_ALGO_* */
    int start;           /* SU2_START_* */
    uint64_t seed;
    int n_therm;         /* thermalisation sweeps (>= 1) */
    int n_meas;          /* measurements */
    int n_cor;           /* sweeps per measurement (>= 1) */
    int or_sweeps;       /* over-relaxation sweeps after each update sweep (<= engine max_or) */
    int normalize_every; /* reunitarisation cadence (0: 100, the reference's) */
    int fp64_twin;       /* server of SU2_PREC_F32 only (engines ignore it): also run the job with FP64 links and
                            report the plaquette difference in its su2_job_result */
} su2_job;

typedef struct su2_engine_config {
    int precision;       /* SU2_PREC_* for every job of the engine */
    int device;
    int max_jobs;        /* resident jobs (0: 256) */
    int max_or;          /* over-relaxation passes per tick (0: 1) */
    size_t budget_bytes; /* device memory for job data (0: 1 GiB; at most 90 % of the free memory) */
    int use_graph;       /* replay ticks from a captured hipGraph (measured 4-11 us per tick slower than the
                            launches enqueued one by one for lone 16^3-64^3 jobs, 1-4 % on a 64-lattice
                            ensemble, on MI355X) */
    int block;           /* sweep workgroup size: 64, 128, 192 or 256 (0: 256; su2_engine_autotune may change it) */
    int grid;            /* most workgroups of a sweep launch (0: 16 per CU; su2_engine_autotune may change it) */
} su2_engine_config;

typedef struct su2_engine su2_engine;

/* Device bytes a job occupies in the arena. */
size_t su2_job_bytes(const su2_job *job, int precision);
int su2_engine_create(const su2_engine_config *cfg, su2_engine **out);
int su2_engine_destroy(su2_engine *e);
/* Admits a job (allocates, seeds setup_kernel-style, starts).  Returns
 * SU2_EBUDGET when its data does not fit in the free arena now and
 * SU2_EINVAL for a bad job; *slot identifies it afterwards. */
int su2_engine_add(su2_engine *e, const su2_job *job, int *slot);
/* Enqueues n ticks (asynchronous). */
int su2_engine_tick(su2_engine *e, int n);
/* Ticks until every resident job is finished; synchronises. */
int su2_engine_run(su2_engine *e);
int su2_engine_sync(su2_engine *e);
/* sweeps done and total sweeps of a job (host counters) */
int su2_engine_job_status(su2_engine *e, int slot, int *done_sweeps, int *total_sweeps);
/* Copies a job's results (synchronising): energies and plaquettes of its
 * n_meas measurements, mean acceptance (kernel.cu:246-249 normalisation,
 * PCC-corrected), and optionally its field in the reference layout
 * (double[12 V]) and RNG state (uint32[6 * V/2]).  NULL skips an output. */
int su2_engine_read(su2_engine *e, int slot, double *E, double *P, double *accept, double *field_ref,
                    uint32_t *rng);
/* Frees a job's slot and memory. */
int su2_engine_remove(su2_engine *e, int slot);
/* Times the candidate (block, grid) shapes on the resident jobs and keeps
 * the fastest; the choice is cached per shape class for the process. */
int su2_engine_autotune(su2_engine *e, int *block, int *grid);
/* Sets the sweep launches' shape: block 64, 128, 192 or 256 threads, grid
 * the cap on their workgroups (0 keeps it).  Ticks already enqueued finish
 * first.  Results are the same bits for every shape; a measurement is
 * carried by the next tick's first launch only with 256 threads. */
int su2_engine_set_shape(su2_engine *e, int block, int grid);
/* Device memory in use / free in the arena. */
int su2_engine_memory(su2_engine *e, size_t *used, size_t *capacity);

/* ---- serving: continuous batching of run requests ----
 *
 * Extension.  A server thread owns an engine: submitted jobs queue, are
 * admitted as soon as a slot and arena memory are free (first that fits,
 * in arrival order), run in the shared ticks with everything else resident,
 * and are retired with their results when done.  Times are seconds on the
 * server's steady clock.  On an FP32-link server a job with fp64_twin also
 * runs, tick for tick, in a second engine with FP64 links that the server
 * thread creates on the first such job (an arena of the same budget): it is
 * admitted when both halves fit, and its result states the plaquette
 * difference.
 */
typedef struct su2_server su2_server;

typedef struct su2_job_result {
    long long id;
    int status;          /* SU2_OK or an error code */
    double t_submit, t_admit, t_done;
    double E_mean, E_err, P_mean, P_err, accept;
    long long link_updates;
    /* jobs with fp64_twin on an FP32-link server, else 0: the FP64 twin's
     * average plaquette and the errors from min(20, n_meas) bins; P_diff =
     * P_mean - P_fp64_twin, P_diff_sigma = the combined binned error */
    double P_fp64_twin, P_fp64_twin_err, P_err_binned, P_diff, P_diff_sigma;
} su2_job_result;

int su2_server_create(const su2_engine_config *cfg, su2_server **out);
/* Stops the server thread after the jobs already submitted are done. */
int su2_server_destroy(su2_server *s);
int su2_server_submit(su2_server *s, const su2_job *job, long long *id);
/* Moves up to max finished results into out; *n = number returned. */
int su2_server_poll(su2_server *s, su2_job_result *out, int max, int *n);
/* Blocks until no job is queued or running. */
int su2_server_drain(su2_server *s);
/* Series of a finished job (before it is polled): NULL skips. */
int su2_server_series(su2_server *s, long long id, double *E, double *P, int n);
double su2_server_now(su2_server *s);

/* ---- host driver (kernel.cu:49-268) ---- */

/*
 * Run configuration: the compile-time #defines of kernel.cu:10-46 as run-
 * time fields (README.md:37-50).  Zero-initialise, then set what you need;
 * su2_run_config_default gives the reference's own defaults where they are
 * meaningful.
 */
typedef struct su2_run_config {
    int Lt, Lx, Ly;     /* d_t, d_x, d_y */
    float beta;         /* betal */
    int boundary;       /* Boundary: SU2_BC_* */
    int algo;           /* ALGO: SU2_ALGO_* */
    int readconf;       /* 0 new start, 1 read Config_ then thermalise, 2 read without thermalisation;
                           the RNG state is restored from Config_*_rng.bin when present (extension),
                           else reseeded from seed as the reference does */
    int init;           /* SU2_START_* (init: 0 hot, 1 cold; 2 Haar extension) */
    int sconf;          /* 0: save the configuration (and, as an extension, the RNG state) after thermalisation */
    int save;           /* 0: write the result files */
    int N;              /* measurements */
    int N_term;         /* thermalisation sweeps */
    int N_cor;          /* sweeps between measurements */
    int Nsave;          /* times the results are saved during the run */
    float epsilon;      /* epsilonl */
    int upd;            /* upd: Metropolis hits per link */
    int W;              /* autocorrelation lags written */
    int or_sweeps;      /* extension: over-relaxation sweeps after each update sweep */
    int precision;      /* SU2_PREC_* */
    int normalize_every; /* reunitarisation cadence in sweeps (reference: 100) */
    uint64_t seed;      /* replaces time(NULL) at kernel.cu:79 */
    const char *outdir; /* directory of the output files (NULL = ".") */
    const char *nsave;  /* run counter in result file names (reference "1") */
    const char *nconfr; /* counter of the Config_ file read */
    const char *nconfs; /* counter of the Config_ file written */
    int verbose;        /* print progress like the reference */
    int device;         /* HIP device ordinal */
    int N_block;        /* progress lines during the measurement (reference 20) */
    int Nblock_term;    /* progress lines during the thermalisation (reference 10) */
    int fp64_twin;      /* extension: with precision SU2_PREC_F32, also run the same configuration with FP64
                           links (same seed, start and schedule, writing nothing) and state the plaquette
                           difference in su2_run_result and the Energy_av_ file */
} su2_run_config;

typedef struct su2_run_result {
    double *energy;      /* caller buffer of N doubles, or NULL */
    double *plaquette;   /* caller buffer of N doubles, or NULL */
    double E_mean, E_err; /* mean_variance of the energy series (statistics.h:30-39) */
    double P_mean, P_err; /* same for the plaquette */
    double accept;       /* mean acceptance (kernel.cu:246-249, PCC-corrected) */
    double det_mean;     /* mean link determinant after the run (print_det) */
    double seconds_therm; /* wall time of the thermalisation loop */
    double seconds_meas; /* wall time of the measurement loop */
    long long link_updates; /* link updates performed (update + over-relaxation) */
    int Nterms;          /* total thermalisation sweeps of the configuration */
    /* with fp64_twin (FP32 links), else 0: the FP64 twin's average plaquette
     * and the errors from 20 bins of the measurements (fewer measurements:
     * one per bin); P_diff_fp32_fp64 = P_mean - P_fp64_twin, P_diff_sigma =
     * the combined binned error (twin runs are the twin's own link updates and
     * are not counted in link_updates or the timings) */
    double P_fp64_twin, P_fp64_twin_err, P_err_binned;
    double P_diff_fp32_fp64, P_diff_sigma;
} su2_run_result;

/* Mean and standard error of a series from nbins equal blocks (trailing
 * values that do not fill a block are dropped; nbins >= 2, n >= nbins):
 * the binning README.md:74 of the reference recommends for final errors. */
int su2_binned_mean_error(const double *m, int n, int nbins, double *mean, double *err);
/* Two series (FP32 links against FP64, or any pair): the difference of
 * their binned means and its combined binned error sqrt(e_a^2 + e_b^2);
 * NULL skips an output. */
int su2_series_compare(const double *a, const double *b, int n, int nbins, double *diff, double *sigma,
                       double *diff_over_sigma);

void su2_run_config_default(su2_run_config *cfg);
int su2_run(const su2_run_config *cfg, su2_run_result *res);

#ifdef __cplusplus
}
#endif

#endif /* SU2MC_H */
