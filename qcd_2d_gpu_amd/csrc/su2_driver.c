/*
 * su2_driver.c - the host driver su2_run, in C on top of the C ABI: the run
 * loop of kernel.cu:49-268 (start or resume, thermalisation, measurement,
 * reunitarisation cadence, periodic statistics and result files).
 *
 * Result files follow functions/output.h:1-94 and functions/lattice_setup.h:
 * 83-100 byte for byte in their text formats, so downstream scripts written
 * for the reference read them unchanged.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "su2_host.h"
#include "su2mc.h"

#define PATH_LEN 4096

static const char *bound_name(int bc)
{
    return bc == SU2_BC_PERIODIC ? "per" : (bc == SU2_BC_PCC ? "pcc" : "pmc");
}

/* SU2_EIO with the file and what failed (the reference exit(1)s) */
static int io_fail(const char *what, const char *f) { return su2_set_error(SU2_EIO, "%s %s", what, f); }

/* "<outdir>/<prefix><Lt>_<Lx>_<Ly>_B_<beta>_<bc>_<counter><suffix>", the
 * reference's file names (output.h:3,17,32,47,63,81,88) */
static int file_name(const su2_run_config *c, const su2_lattice *l, const char *bound, const char *counter,
                     const char *prefix, const char *suffix, char *out)
{
    const char *d = (c->outdir && *c->outdir) ? c->outdir : ".";
    int n = snprintf(out, PATH_LEN, "%s/%s%d_%d_%d_B_%.2lf_%s_%s%s", d, prefix, l->t, l->x, l->y, (double)l->beta,
                     bound, counter ? counter : "1", suffix);
    if (n < 0 || n >= PATH_LEN) return su2_set_error(SU2_EINVAL, "output path longer than %d bytes", PATH_LEN - 1);
    return SU2_OK;
}

/* mean_variance (functions/statistics.h:30-39): population variance. */
static void mean_variance(const double *m, int n, double *res)
{
    double s = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) {
        s += m[i];
        s2 += m[i] * m[i];
    }
    res[0] = s / n;
    res[1] = s2 / n - res[0] * res[0];
}

/* Mean and standard error from nbins equal blocks of the series (trailing
 * values that do not fill a block are dropped): the binning the reference's
 * README recommends for final errors (README.md:74). */
int su2_binned_mean_error(const double *m, int n, int nbins, double *mean, double *err)
{
    if (!m || !mean || !err || nbins < 2 || n < nbins)
        return su2_set_error(SU2_EINVAL, "binned error needs nbins >= 2 and at least nbins values");
    const int b = n / nbins;
    double s = 0.0;
    for (int k = 0; k < nbins; ++k) {
        double bm = 0.0;
        for (int i = 0; i < b; ++i) bm += m[k * b + i];
        s += bm / b;
    }
    const double mu = s / nbins;
    double s2 = 0.0; /* second pass: squared deviations of the bin means */
    for (int k = 0; k < nbins; ++k) {
        double bm = 0.0;
        for (int i = 0; i < b; ++i) bm += m[k * b + i];
        const double d = bm / b - mu;
        s2 += d * d;
    }
    *mean = mu;
    *err = sqrt(s2 / (nbins - 1) / nbins); /* unbiased variance of the bin means / nbins */
    return SU2_OK;
}

int su2_series_compare(const double *a, const double *b, int n, int nbins, double *diff, double *sigma,
                       double *diff_over_sigma)
{
    double ma, ea, mb, eb;
    int rc = su2_binned_mean_error(a, n, nbins, &ma, &ea);
    if (rc) return rc;
    if ((rc = su2_binned_mean_error(b, n, nbins, &mb, &eb))) return rc;
    const double sg = sqrt(ea * ea + eb * eb);
    if (diff) *diff = ma - mb;
    if (sigma) *sigma = sg;
    if (diff_over_sigma) *diff_over_sigma = sg > 0.0 ? (ma - mb) / sg : 0.0;
    return SU2_OK;
}

/* autocorrelation (functions/statistics.h:41-51), normalised by the
 * variance; lags with no pairs (i >= n) are reported as 0. */
static void autocorrelation(const double *m, int n, int W, double mean, double var, double *out)
{
    for (int i = 1; i <= W; ++i) {
        double s = 0.0;
        for (int j = 0; j < n - i; ++j) s += (m[j] - mean) * (m[j + i] - mean);
        out[i - 1] = (n - i > 0 && var != 0.0) ? s / (n - i) / var : 0.0;
    }
}

static int write_autocorrelation(const su2_run_config *c, const char *prefix, const double *r, int W, int N,
                                 int N_cor, int N_term, const su2_lattice *l, const char *bound)
{
    /* save_autocorrelation / save_autocorrelation_av (output.h:1-14,45-58) */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nsave, prefix, ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalizacion %d N %d N_correlation %d  D_t %d D_x %d D_ y %d beta %.2lf deltaphi %.2lf\n", N_term,
            N, N_cor, l->t, l->x, l->y, (double)l->beta, (double)l->epsilon);
    fprintf(fp, "N Autocorrelation\n");
    for (int i = 0; i < W; ++i) fprintf(fp, "%d %.10lf\n", i + 1, r[i]);
    fclose(fp);
    return SU2_OK;
}

static int write_energy(const su2_run_config *c, double mean, double var, int N_term, int N, int N_cor, double acc,
                        const su2_lattice *l, const char *bound)
{
    /* save_energy (output.h:15-28) */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nsave, "Energy_", ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalizacion %d N %d N_correlation %d  D_t %d D_x %d D_ y %d deltaphi %.2lf accept %.2lf\n", N_term,
            N, N_cor, l->t, l->x, l->y, (double)l->epsilon, acc);
    fprintf(fp, "M Energy Error\n");
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l->beta, mean, sqrt(var / (N - 1)));
    fclose(fp);
    return SU2_OK;
}

/* The FP32-link run against its FP64 twin (same seed, start and schedule):
 * the average plaquettes with 20-bin errors and their difference in units of
 * the combined error. */
typedef struct twin_cmp {
    double P64, P64_err, P32_err, diff, sigma;
} twin_cmp;

static int write_energy_av(const su2_run_config *c, double m1, double v1, double m2, double v2, int N_term, int N,
                           int N_cor, double acc, const su2_lattice *l, const char *bound, const twin_cmp *tw)
{
    /* save_energy_av (output.h:30-44) */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nsave, "Energy_av_", ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp,
            "N_termalizacion %d N %d N_correlation %d N_updates %d  D_t %d D_x %d D_ y %d deltaphi %.2lf accept "
            "%.2lf\n",
            N_term, N, N_cor, l->updates, l->t, l->x, l->y, (double)l->epsilon, acc);
    fprintf(fp, "M Energy/Avplaquette Error\n");
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l->beta, m1, sqrt(v1 / (N - 1)));
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l->beta, m2, sqrt(v2 / (N - 1)));
    if (tw) /* extension: only in FP32-link runs with an FP64 twin */
        fprintf(fp, "FP64_twin_plaquette %.10lf %.10lf FP32_minus_FP64 %.10lf sigma %.10lf diff_over_sigma %.4lf\n",
                tw->P64, tw->P64_err, tw->diff, tw->sigma, tw->sigma > 0.0 ? tw->diff / tw->sigma : 0.0);
    fclose(fp);
    return SU2_OK;
}

static int write_en_seq(const su2_run_config *c, const double *plaq, const double *en, int N_term, int N, int N_cor,
                        double acc, const su2_lattice *l, const char *bound)
{
    /* save_en_seq (output.h:61-75): "i E P" per measurement */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nsave, "Check_", ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp,
            "N_termalizacion %d N %d N_correlation %d N_updates %d  D_t %d D_x %d D_y %d deltaphi %.2lf accept "
            "%.2lf\n",
            N_term, N, N_cor, l->updates, l->t, l->x, l->y, (double)l->epsilon, acc);
    fprintf(fp, "I Energy Avpl\n");
    for (int i = 0; i < N; ++i) fprintf(fp, "%d %.10lf %.10lf\n", i, en[i], plaq[i]);
    fclose(fp);
    return SU2_OK;
}

static int write_conf(const su2_run_config *c, const double *phi, size_t n, const su2_lattice *l, int N_term, int init,
                      const char *bound)
{
    /* save_conf (output.h:77-94): raw native-endian double[12 V] + _data.txt */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nconfs, "Config_", ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "wb");
    if (!fp) return io_fail("cannot open", f);
    size_t w = fwrite(phi, sizeof(double), n, fp);
    fclose(fp);
    if (w != n) return io_fail("short write to", f);
    if ((rc = file_name(c, l, bound, c->nconfs, "Config_", "_data.txt", f))) return rc;
    fp = fopen(f, "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalization %d Type_termalization %d", N_term, init);
    fclose(fp);
    return SU2_OK;
}

static int read_conf(const su2_run_config *c, double *phi, size_t n, const su2_lattice *l, int *Nterm, int *init,
                     const char *bound)
{
    /* read_conf (lattice_setup.h:83-100) */
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nconfr, "Config_", ".txt", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "rb");
    if (!fp) return io_fail("cannot open", f);
    size_t r = fread(phi, sizeof(double), n, fp);
    fclose(fp);
    if (r != n) return io_fail("short read from", f);
    if ((rc = file_name(c, l, bound, c->nconfr, "Config_", "_data.txt", f))) return rc;
    fp = fopen(f, "r");
    if (!fp) return io_fail("cannot open", f);
    int got = fscanf(fp, "N_termalization %d Type_termalization %d", Nterm, init);
    fclose(fp);
    return got == 2 ? SU2_OK : io_fail("malformed", f);
}

/* RNG checkpoint (extension: the reference reseeds on resume, so its
 * resumed runs are not reproducible): Config_<tail>_rng.bin holds the raw
 * uint32 state planes g1[0..2], g2[0..2] of the V/2 streams. */
static int write_rng(const su2_run_config *c, const uint32_t *st, size_t n, const su2_lattice *l, const char *bound)
{
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nconfs, "Config_", "_rng.bin", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "wb");
    if (!fp) return io_fail("cannot open", f);
    size_t w = fwrite(st, sizeof(uint32_t), n, fp);
    fclose(fp);
    return w == n ? SU2_OK : io_fail("short write to", f);
}

/* 1: restored, 0: no checkpoint file, < 0: error */
static int read_rng(const su2_run_config *c, uint32_t *st, size_t n, const su2_lattice *l, const char *bound)
{
    char f[PATH_LEN];
    int rc = file_name(c, l, bound, c->nconfr, "Config_", "_rng.bin", f);
    if (rc) return rc;
    FILE *fp = fopen(f, "rb");
    if (!fp) return 0;
    size_t r = fread(st, sizeof(uint32_t), n, fp);
    fclose(fp);
    return r == n ? 1 : io_fail("short read from", f);
}

static double mean_det(const double *phi, size_t n)
{
    /* print_det (lattice_setup.h:176-183) */
    double det = 0.0;
    const size_t nl = n / 4;
    for (size_t i = 0; i < nl; ++i) {
        const double *a = phi + 4 * i;
        det += a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    }
    return det / (double)nl;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void su2_run_config_default(su2_run_config *c)
{
    memset(c, 0, sizeof *c);
    c->Lt = 16;
    c->Lx = 16;
    c->Ly = 16;
    c->beta = 3.0f;
    c->boundary = SU2_BC_PERIODIC;
    c->algo = SU2_ALGO_HEATBATH;
    c->readconf = 0;
    c->init = SU2_START_COLD;
    c->sconf = 1;
    c->save = 1;
    c->N = 100;
    c->N_term = 100;
    c->N_cor = 1;
    c->Nsave = 1;
    c->epsilon = 0.1f;
    c->upd = 1;
    c->W = 80;
    c->or_sweeps = 0;
    c->precision = SU2_PREC_F64;
    c->normalize_every = 100;
    c->seed = 1234;
    c->outdir = ".";
    c->nsave = "1";
    c->nconfr = "1";
    c->nconfs = "1";
    c->verbose = 0;
    c->device = 0;
    c->N_block = 20;
    c->Nblock_term = 10;
}

/* the state of one run */
typedef struct run_ctx {
    const su2_run_config *c;
    su2_lattice l;
    int prec;
    void *phi, *st, *acc, *end, *end2, *work, *ref;
    void *stream;
    long long updates;
    long long per_meas; /* acceptance denominator per measurement, per site */
} run_ctx;

/* one mc_up_per / mc_up_pcc / mc_up_pmc (kernel.cu:166-170,196-200) */
static int sweep_kind(run_ctx *r, int update)
{
    const su2_run_config *c = r->c;
    switch (c->boundary) {
    case SU2_BC_PERIODIC:
        return su2_mc_up_per(&r->l, r->phi, r->prec, (int *)r->acc, (uint32_t *)r->st, c->algo, update, r->stream);
    case SU2_BC_PCC:
        return su2_mc_up_pcc(&r->l, r->phi, r->prec, (int *)r->acc, (uint32_t *)r->st, c->algo, update, r->stream);
    default:
        return su2_mc_up_pmc(&r->l, r->phi, r->prec, (int *)r->acc, (uint32_t *)r->st, c->algo, update, r->stream);
    }
}

/* en_av of measurement i (kernel.cu:208): into slots end[i-1], end2[i-1] */
static int measure(run_ctx *r, int i)
{
    return su2_en_av(&r->l, r->phi, r->prec, r->c->boundary == SU2_BC_PMC, (double *)r->end + i - 1,
                     (double *)r->end2 + i - 1, r->work, r->stream);
}

/* one update sweep, then the over-relaxation sweeps (extension).  With
 * *pending > 0, the en_av of measurement *pending, held back since the end
 * of its sweeps, runs first, on the update sweep's first launch
 * (su2_mc_up_measured: the same bits as su2_en_av before the sweep). */
static int one_sweep(run_ctx *r, int *pending)
{
    int rc;
    if (*pending > 0) {
        const int i = *pending;
        rc = su2_mc_up_measured(&r->l, r->phi, r->prec, (int *)r->acc, (uint32_t *)r->st, r->c->boundary,
                                r->c->algo, SU2_UPDATE_LINK, (double *)r->end + i - 1, (double *)r->end2 + i - 1,
                                r->work, r->stream);
        *pending = 0;
    } else {
        rc = sweep_kind(r, SU2_UPDATE_LINK);
    }
    if (rc) return rc;
    r->updates += 3ll * r->l.V;
    for (int k = 0; k < r->c->or_sweeps; ++k) {
        if ((rc = sweep_kind(r, SU2_UPDATE_OVERRELAX))) return rc;
        r->updates += 3ll * r->l.V;
    }
    return SU2_OK;
}

/* mean acceptance after i measurements (kernel.cu:225-228,246-249) */
static int acceptance(run_ctx *r, int i, double *out)
{
    double m = 0.0, e = 0.0;
    int rc = su2_mean_lattice(&r->l, (const int *)r->acc, (long long)i * r->per_meas, &m, &e, r->stream);
    if (rc) return rc;
    if (r->c->boundary == SU2_BC_PCC) m = m / ((double)(r->l.y - 1) / r->l.y + 1 / (double)r->l.y / 3);
    *out = m;
    return SU2_OK;
}

#define CK(x)                                                                                                          \
    do {                                                                                                               \
        rc = (x);                                                                                                      \
        if (rc != SU2_OK) goto done;                                                                                   \
    } while (0)

#define CK_ALLOC(p, n)                                                                                                 \
    do {                                                                                                               \
        if (!((p) = malloc(n))) {                                                                                      \
            rc = su2_set_error(SU2_ENOMEM, "host buffer of %zu bytes", (size_t)(n));                                   \
            goto done;                                                                                                 \
        }                                                                                                              \
    } while (0)

int su2_run(const su2_run_config *c, su2_run_result *res)
{
    if (!c || !res) return su2_set_error(SU2_EINVAL, "NULL configuration or result");
    if (c->N < 1 || c->N_term < 0 || c->N_cor < 1 || c->Nsave < 1 || c->W < 0 || c->or_sweeps < 0 ||
        c->normalize_every < 1)
        return su2_set_error(SU2_EINVAL, "run lengths out of range (N, N_cor, Nsave, normalize_every >= 1; "
                                         "N_term, W, or >= 0)");
    if (c->boundary < SU2_BC_PERIODIC || c->boundary > SU2_BC_PMC)
        return su2_set_error(SU2_EINVAL, "boundary %d", c->boundary);
    if (c->algo != SU2_ALGO_HEATBATH && c->algo != SU2_ALGO_METROPOLIS)
        return su2_set_error(SU2_EINVAL, "algo %d", c->algo);
    if (c->readconf < 0 || c->readconf > 2) return su2_set_error(SU2_EINVAL, "readconf %d", c->readconf);
    if (c->init < SU2_START_HOT || c->init > SU2_START_HAAR) return su2_set_error(SU2_EINVAL, "init %d", c->init);

    run_ctx r;
    memset(&r, 0, sizeof r);
    r.c = c;
    double *h_phi = NULL, *en = NULL, *en2 = NULL, *autoc = NULL, *autoc2 = NULL, *p64 = NULL;
    const int twin = c->precision == SU2_PREC_F32 && c->fp64_twin;
    twin_cmp tw;
    memset(&tw, 0, sizeof tw);
    uint32_t *hst = NULL;
    int rc = su2_lattice_init(&r.l, c->Lt, c->Lx, c->Ly, c->beta, c->epsilon, c->upd);
    if (rc) return rc;
    if ((rc = su2_set_device(c->device))) return rc;
    if (twin) {
        /* the FP64 twin: the same run with FP64 links, writing nothing */
        su2_run_config c64 = *c;
        c64.precision = SU2_PREC_F64;
        c64.fp64_twin = 0;
        c64.save = 1;
        c64.sconf = 1;
        c64.verbose = 0;
        su2_run_result r64;
        memset(&r64, 0, sizeof r64);
        if (!(p64 = malloc((size_t)c->N * sizeof(double))))
            return su2_set_error(SU2_ENOMEM, "host buffer of %zu bytes", (size_t)c->N * sizeof(double));
        r64.plaquette = p64;
        if ((rc = su2_run(&c64, &r64))) {
            free(p64);
            return rc;
        }
    }
    su2_lattice *l = &r.l;
    const int prec = r.prec = c->precision;
    const char *bound = bound_name(c->boundary);
    const int hits = (c->algo == SU2_ALGO_METROPOLIS) ? l->updates : 1; /* kernel.cu:74 */
    const int N = c->N;
    const size_t nphi = (size_t)l->V * 12, nst = (size_t)l->V / 2 * 6;
    const size_t W = c->W > 0 ? (size_t)c->W : 1;
    void *stream = r.stream = NULL;

    CK(su2_malloc(&r.st, su2_rng_bytes(l)));
    CK(su2_malloc(&r.phi, su2_field_bytes(l, prec)));
    CK(su2_malloc(&r.acc, su2_accept_bytes(l)));
    CK(su2_malloc(&r.end, (size_t)N * sizeof(double)));
    CK(su2_malloc(&r.end2, (size_t)N * sizeof(double)));
    CK(su2_malloc(&r.work, su2_en_av_work_bytes(l)));
    CK(su2_memset_zero(r.work, su2_en_av_work_bytes(l), stream));
    CK(su2_malloc(&r.ref, nphi * sizeof(double)));
    CK(su2_memset_zero(r.end, (size_t)N * sizeof(double), stream));
    CK(su2_memset_zero(r.end2, (size_t)N * sizeof(double), stream));
    CK(su2_setup_kernel(l, (uint32_t *)r.st, c->seed, stream)); /* kernel.cu:79 */
    CK_ALLOC(h_phi, nphi * sizeof(double));
    CK_ALLOC(hst, nst * sizeof(uint32_t));
    CK_ALLOC(en, (size_t)N * sizeof(double));
    CK_ALLOC(en2, (size_t)N * sizeof(double));
    CK_ALLOC(autoc, W * sizeof(double));
    CK_ALLOC(autoc2, W * sizeof(double));

    int Nterms = 0, inits = c->init;
    if (c->readconf == 0) { /* kernel.cu:85-136 */
        if (c->init == SU2_START_COLD)
            CK(su2_id_init2(l, r.phi, prec, stream));
        else if (c->init == SU2_START_HAAR)
            CK(su2_haar_init(l, r.phi, prec, (uint32_t *)r.st, c->boundary == SU2_BC_PCC, stream));
        else if (c->boundary == SU2_BC_PCC)
            CK(su2_rand_init2_pcc(l, r.phi, prec, (uint32_t *)r.st, stream));
        else
            CK(su2_rand_init2(l, r.phi, prec, (uint32_t *)r.st, stream));
    } else {
        CK(read_conf(c, h_phi, nphi, l, &Nterms, &inits, bound));
        CK(su2_memcpy_h2d(r.ref, h_phi, nphi * sizeof(double), stream));
        CK(su2_field_from_ref(l, (const double *)r.ref, r.phi, prec, stream));
        int got = read_rng(c, hst, nst, l, bound);
        if (got < 0) {
            rc = got;
            goto done;
        }
        if (got == 1) CK(su2_memcpy_h2d(r.st, hst, nst * sizeof(uint32_t), stream));
    }
    CK(su2_field_to_ref(l, r.phi, prec, (double *)r.ref, stream));
    CK(su2_memcpy_d2h(h_phi, r.ref, nphi * sizeof(double), stream));
    if (c->verbose) printf("Av det %.16lf\n", mean_det(h_phi, nphi));
    CK(su2_accept_clear(l, (int *)r.acc, stream));

    /* progress lines (kernel.cu:156,172-176,212-216): every N_term/Nblock_term
     * and N/N_block sweeps; a zero block prints nothing (the reference's
     * modulus by zero) */
    const int block_t = c->Nblock_term > 0 ? c->N_term / c->Nblock_term : 0;
    const int block_m = c->N_block > 0 ? N / c->N_block : 0;
    double t0 = now_s(), t_n = t0;
    int pending = 0; /* a measurement whose en_av waits for the next sweep */
    if (c->readconf < 2) { /* kernel.cu:161-187 */
        for (int i = 1; i < c->N_term + 1; ++i) {
            if (i % c->normalize_every == 0) CK(su2_normalize(l, r.phi, prec, stream));
            CK(one_sweep(&r, &pending));
            if (c->verbose && block_t > 0 && i % block_t == 0) {
                CK(su2_stream_synchronize(stream));
                const double t = now_s();
                printf("%lf Time %d Left(min)%lf\n", (double)i / c->N_term, (int)(t - t_n),
                       (t - t0) * (c->N_term - i) / i / 60);
                fflush(stdout);
                t_n = t;
            }
        }
        CK(su2_normalize(l, r.phi, prec, stream));
        Nterms = c->N_term + Nterms;
        if (c->sconf == 0) {
            CK(su2_field_to_ref(l, r.phi, prec, (double *)r.ref, stream));
            CK(su2_memcpy_d2h(h_phi, r.ref, nphi * sizeof(double), stream));
            CK(write_conf(c, h_phi, nphi, l, Nterms, inits, bound));
            CK(su2_memcpy_d2h(hst, r.st, nst * sizeof(uint32_t), stream));
            CK(write_rng(c, hst, nst, l, bound));
        }
    }
    CK(su2_stream_synchronize(stream));
    double t1 = now_s();
    t_n = t1;

    CK(su2_accept_clear(l, (int *)r.acc, stream)); /* kernel.cu:190 */
    double energy[2] = {0, 0}, energy2[2] = {0, 0}, saccept = 0.0;
    r.per_meas = (long long)c->N_cor * l->d * (hits + c->or_sweeps);
    const int block_s = N / c->Nsave > 0 ? N / c->Nsave : N;
    for (int i = 1; i < N + 1; ++i) { /* kernel.cu:192-236 */
        for (int j = 0; j < c->N_cor; ++j) CK(one_sweep(&r, &pending));
        if (i % c->normalize_every == 0) CK(su2_normalize(l, r.phi, prec, stream));
        /* en_av of measurement i (kernel.cu:208) on the next sweep's first
         * launch, or here when the series is read before that */
        pending = i;
        if (i == N || (i % block_s == 0 && i != N)) {
            CK(measure(&r, i));
            pending = 0;
        }
        if (c->verbose && block_m > 0 && i % block_m == 0) {
            CK(su2_stream_synchronize(stream));
            const double t = now_s();
            printf("%lf Time %d Left(min)%lf\n", (double)i / N, (int)(t - t_n), (t - t1) * (N - i) / i / 60);
            fflush(stdout);
            t_n = t;
        }
        if (i % block_s == 0 && i != N) {
            CK(su2_memcpy_d2h(en, r.end, (size_t)i * sizeof(double), stream));
            CK(su2_memcpy_d2h(en2, r.end2, (size_t)i * sizeof(double), stream));
            mean_variance(en, i, energy);
            autocorrelation(en, i, c->W, energy[0], energy[1], autoc);
            mean_variance(en2, i, energy2);
            autocorrelation(en2, i, c->W, energy2[0], energy2[1], autoc2);
            if (c->verbose)
                printf("Energy %.10lf +- %.10lf\nPlaquette %.10lf +- %.10lf\n", energy[0], sqrt(energy[1] / (i - 1)),
                       energy2[0], sqrt(energy2[1] / (i - 1)));
            if (c->save == 0) {
                CK(acceptance(&r, i, &saccept));
                CK(write_energy(c, energy[0], energy[1], Nterms, N, c->N_cor, saccept, l, bound));
                CK(write_energy_av(c, energy[0], energy[1], energy2[0], energy2[1], Nterms, N, c->N_cor, saccept, l,
                                   bound, NULL));
                CK(write_autocorrelation(c, "Autocorrelation_", autoc, c->W, i, c->N_cor, Nterms, l, bound));
                CK(write_autocorrelation(c, "Autocorrelation_av_", autoc2, c->W, i, c->N_cor, Nterms, l, bound));
                CK(write_en_seq(c, en2, en, Nterms, i, c->N_cor, 1.0, l, bound));
            }
        }
    }
    CK(su2_stream_synchronize(stream));
    double t2 = now_s();

    CK(su2_memcpy_d2h(en, r.end, (size_t)N * sizeof(double), stream));
    CK(su2_memcpy_d2h(en2, r.end2, (size_t)N * sizeof(double), stream));
    CK(acceptance(&r, N, &saccept));
    mean_variance(en, N, energy);
    autocorrelation(en, N, c->W, energy[0], energy[1], autoc);
    mean_variance(en2, N, energy2);
    autocorrelation(en2, N, c->W, energy2[0], energy2[1], autoc2);
    if (twin) {
        double z;
        CK(su2_series_compare(en2, p64, N, N < 20 ? N : 20, &tw.diff, &tw.sigma, &z));
        double m32;
        CK(su2_binned_mean_error(en2, N, N < 20 ? N : 20, &m32, &tw.P32_err));
        CK(su2_binned_mean_error(p64, N, N < 20 ? N : 20, &tw.P64, &tw.P64_err));
    }
    if (c->save == 0) { /* kernel.cu:254-260 */
        CK(write_energy(c, energy[0], energy[1], Nterms, N, c->N_cor, saccept, l, bound));
        CK(write_energy_av(c, energy[0], energy[1], energy2[0], energy2[1], Nterms, N, c->N_cor, saccept, l, bound,
                           twin ? &tw : NULL));
        CK(write_en_seq(c, en2, en, Nterms, N, c->N_cor, 1.0, l, bound));
        CK(write_autocorrelation(c, "Autocorrelation_", autoc, c->W, N, c->N_cor, Nterms, l, bound));
        CK(write_autocorrelation(c, "Autocorrelation_av_", autoc2, c->W, N, c->N_cor, Nterms, l, bound));
    }
    if (c->verbose) {
        printf("accept %lf\n\n", saccept);
        printf("Energy %.10lf +- %.10lf\n", energy[0], N > 1 ? sqrt(energy[1] / (N - 1)) : 0.0);
        printf("Plaquette %.10lf +- %.10lf\n", energy2[0], N > 1 ? sqrt(energy2[1] / (N - 1)) : 0.0);
    }
    CK(su2_field_to_ref(l, r.phi, prec, (double *)r.ref, stream));
    CK(su2_memcpy_d2h(h_phi, r.ref, nphi * sizeof(double), stream));

    if (res->energy) memcpy(res->energy, en, (size_t)N * sizeof(double));
    if (res->plaquette) memcpy(res->plaquette, en2, (size_t)N * sizeof(double));
    res->E_mean = energy[0];
    res->E_err = N > 1 ? sqrt(energy[1] / (N - 1)) : 0.0;
    res->P_mean = energy2[0];
    res->P_err = N > 1 ? sqrt(energy2[1] / (N - 1)) : 0.0;
    res->accept = saccept;
    res->det_mean = mean_det(h_phi, nphi);
    res->seconds_therm = t1 - t0;
    res->seconds_meas = t2 - t1;
    res->link_updates = r.updates;
    res->Nterms = Nterms;
    res->P_fp64_twin = twin ? tw.P64 : 0.0;
    res->P_fp64_twin_err = twin ? tw.P64_err : 0.0;
    res->P_err_binned = twin ? tw.P32_err : 0.0;
    res->P_diff_fp32_fp64 = twin ? tw.diff : 0.0;
    res->P_diff_sigma = twin ? tw.sigma : 0.0;
    rc = SU2_OK;

done:
    free(p64);
    free(h_phi);
    free(hst);
    free(en);
    free(en2);
    free(autoc);
    free(autoc2);
    void *bufs[] = {r.ref, r.work, r.end2, r.end, r.acc, r.phi, r.st};
    for (size_t k = 0; k < sizeof bufs / sizeof bufs[0]; ++k)
        if (bufs[k]) su2_free(bufs[k]);
    return rc;
}
