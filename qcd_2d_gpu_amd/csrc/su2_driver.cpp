/*
 * su2_driver.cpp - the host driver su2_run: the run loop of kernel.cu:49-268
 * (start or resume, thermalisation, measurement, reunitarisation cadence,
 * periodic statistics and result files) on top of the C ABI.
 *
 * Result files follow functions/output.h:1-94 and functions/lattice_setup.h:
 * 83-100 byte for byte in their text formats, so downstream scripts written
 * for the reference read them unchanged.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "su2_internal.h"
#include "su2mc.h"

namespace {

struct dev_buf {
    void *p = nullptr;
    ~dev_buf()
    {
        if (p) su2_free(p);
    }
};

const char *bound_name(int bc)
{
    return bc == SU2_BC_PERIODIC ? "per" : (bc == SU2_BC_PCC ? "pcc" : "pmc");
}

/* SU2_EIO with the file and what failed (the reference exit(1)s) */
int io_fail(const char *what, const std::string &f) { return su2_set_error(SU2_EIO, "%s %s", what, f.c_str()); }

std::string path_of(const su2_run_config *c, const char *name)
{
    std::string d = (c->outdir && *c->outdir) ? c->outdir : ".";
    return d + "/" + name;
}

/* mean_variance (functions/statistics.h:30-39): population variance. */
void mean_variance(const double *m, int n, double *res)
{
    double s = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) {
        s += m[i];
        s2 += m[i] * m[i];
    }
    res[0] = s / n;
    res[1] = s2 / n - res[0] * res[0];
}

/* autocorrelation (functions/statistics.h:41-51), normalised by the
 * variance; lags with no pairs (i >= n) are reported as 0. */
void autocorrelation(const double *m, int n, int W, double mean, double var, double *out)
{
    for (int i = 1; i <= W; ++i) {
        double s = 0.0;
        for (int j = 0; j < n - i; ++j) s += (m[j] - mean) * (m[j + i] - mean);
        out[i - 1] = (n - i > 0 && var != 0.0) ? s / (n - i) / var : 0.0;
    }
}

struct names {
    char tail[160]; /* "%d_%d_%d_B_%.2lf_%s_%s" */
};

names make_tail(const su2_lattice &l, const char *bound, const char *counter)
{
    names n;
    snprintf(n.tail, sizeof n.tail, "%d_%d_%d_B_%.2lf_%s_%s", l.t, l.x, l.y, (double)l.beta, bound, counter);
    return n;
}

int write_autocorrelation(const su2_run_config *c, const char *prefix, const double *r, int W, int N, int N_cor,
                          int N_term, const su2_lattice &l, const char *bound)
{
    /* save_autocorrelation / save_autocorrelation_av (output.h:1-14,45-58) */
    std::string f = path_of(c, (std::string(prefix) + make_tail(l, bound, c->nsave).tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalizacion %d N %d N_correlation %d  D_t %d D_x %d D_ y %d beta %.2lf deltaphi %.2lf\n", N_term,
            N, N_cor, l.t, l.x, l.y, (double)l.beta, (double)l.epsilon);
    fprintf(fp, "N Autocorrelation\n");
    for (int i = 0; i < W; ++i) fprintf(fp, "%d %.10lf\n", i + 1, r[i]);
    fclose(fp);
    return SU2_OK;
}

int write_energy(const su2_run_config *c, double mean, double var, int N_term, int N, int N_cor, double acc,
                 const su2_lattice &l, const char *bound)
{
    /* save_energy (output.h:15-28) */
    std::string f = path_of(c, (std::string("Energy_") + make_tail(l, bound, c->nsave).tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalizacion %d N %d N_correlation %d  D_t %d D_x %d D_ y %d deltaphi %.2lf accept %.2lf\n", N_term,
            N, N_cor, l.t, l.x, l.y, (double)l.epsilon, acc);
    fprintf(fp, "M Energy Error\n");
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l.beta, mean, std::sqrt(var / (N - 1)));
    fclose(fp);
    return SU2_OK;
}

int write_energy_av(const su2_run_config *c, double m1, double v1, double m2, double v2, int N_term, int N, int N_cor,
                    double acc, const su2_lattice &l, const char *bound)
{
    /* save_energy_av (output.h:30-44) */
    std::string f = path_of(c, (std::string("Energy_av_") + make_tail(l, bound, c->nsave).tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp,
            "N_termalizacion %d N %d N_correlation %d N_updates %d  D_t %d D_x %d D_ y %d deltaphi %.2lf accept "
            "%.2lf\n",
            N_term, N, N_cor, l.updates, l.t, l.x, l.y, (double)l.epsilon, acc);
    fprintf(fp, "M Energy/Avplaquette Error\n");
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l.beta, m1, std::sqrt(v1 / (N - 1)));
    fprintf(fp, "%f %.10lf %.10lf\n", (double)l.beta, m2, std::sqrt(v2 / (N - 1)));
    fclose(fp);
    return SU2_OK;
}

int write_en_seq(const su2_run_config *c, const double *plaq, const double *en, int N_term, int N, int N_cor,
                 double acc, const su2_lattice &l, const char *bound)
{
    /* save_en_seq (output.h:61-75): "i E P" per measurement */
    std::string f = path_of(c, (std::string("Check_") + make_tail(l, bound, c->nsave).tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp,
            "N_termalizacion %d N %d N_correlation %d N_updates %d  D_t %d D_x %d D_y %d deltaphi %.2lf accept "
            "%.2lf\n",
            N_term, N, N_cor, l.updates, l.t, l.x, l.y, (double)l.epsilon, acc);
    fprintf(fp, "I Energy Avpl\n");
    for (int i = 0; i < N; ++i) fprintf(fp, "%d %.10lf %.10lf\n", i, en[i], plaq[i]);
    fclose(fp);
    return SU2_OK;
}

int write_conf(const su2_run_config *c, const std::vector<double> &phi, const su2_lattice &l, int N_term, int init,
               const char *bound)
{
    /* save_conf (output.h:77-94): raw native-endian double[12 V] + _data.txt */
    names n = make_tail(l, bound, c->nconfs);
    std::string f = path_of(c, (std::string("Config_") + n.tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "wb");
    if (!fp) return io_fail("cannot open", f);
    size_t w = fwrite(phi.data(), sizeof(double), phi.size(), fp);
    fclose(fp);
    if (w != phi.size()) return io_fail("short write to", f);
    f = path_of(c, (std::string("Config_") + n.tail + "_data.txt").c_str());
    fp = fopen(f.c_str(), "w");
    if (!fp) return io_fail("cannot open", f);
    fprintf(fp, "N_termalization %d Type_termalization %d", N_term, init);
    fclose(fp);
    return SU2_OK;
}

int read_conf(const su2_run_config *c, std::vector<double> &phi, const su2_lattice &l, int *Nterm, int *init,
              const char *bound)
{
    /* read_conf (lattice_setup.h:83-100) */
    names n = make_tail(l, bound, c->nconfr);
    std::string f = path_of(c, (std::string("Config_") + n.tail + ".txt").c_str());
    FILE *fp = fopen(f.c_str(), "rb");
    if (!fp) return io_fail("cannot open", f);
    size_t r = fread(phi.data(), sizeof(double), phi.size(), fp);
    fclose(fp);
    if (r != phi.size()) return io_fail("short read from", f);
    f = path_of(c, (std::string("Config_") + n.tail + "_data.txt").c_str());
    fp = fopen(f.c_str(), "r");
    if (!fp) return io_fail("cannot open", f);
    int got = fscanf(fp, "N_termalization %d Type_termalization %d", Nterm, init);
    fclose(fp);
    return got == 2 ? SU2_OK : io_fail("malformed", f);
}

/* RNG checkpoint (extension: the reference reseeds on resume, so its
 * resumed runs are not reproducible): Config_<tail>_rng.bin holds the raw
 * uint32 state planes g1[0..2], g2[0..2] of the V/2 streams. */
int write_rng(const su2_run_config *c, const std::vector<uint32_t> &st, const su2_lattice &l, const char *bound)
{
    names n = make_tail(l, bound, c->nconfs);
    std::string f = path_of(c, (std::string("Config_") + n.tail + "_rng.bin").c_str());
    FILE *fp = fopen(f.c_str(), "wb");
    if (!fp) return io_fail("cannot open", f);
    size_t w = fwrite(st.data(), sizeof(uint32_t), st.size(), fp);
    fclose(fp);
    return w == st.size() ? SU2_OK : io_fail("short write to", f);
}

/* 1: restored, 0: no checkpoint file, < 0: error */
int read_rng(const su2_run_config *c, std::vector<uint32_t> &st, const su2_lattice &l, const char *bound)
{
    names n = make_tail(l, bound, c->nconfr);
    std::string f = path_of(c, (std::string("Config_") + n.tail + "_rng.bin").c_str());
    FILE *fp = fopen(f.c_str(), "rb");
    if (!fp) return 0;
    size_t r = fread(st.data(), sizeof(uint32_t), st.size(), fp);
    fclose(fp);
    return r == st.size() ? 1 : io_fail("short read from", f);
}

double mean_det(const std::vector<double> &phi)
{
    /* print_det (lattice_setup.h:176-183) */
    double det = 0.0;
    const size_t nl = phi.size() / 4;
    for (size_t i = 0; i < nl; ++i) {
        const double *a = &phi[4 * i];
        det += a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    }
    return det / (double)nl;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

extern "C" {

void su2_run_config_default(su2_run_config *c)
{
    std::memset(c, 0, sizeof *c);
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
}

#define RUN_CHECK(x)                                                                                                   \
    do {                                                                                                               \
        int rc_ = (x);                                                                                                 \
        if (rc_ != SU2_OK) return rc_;                                                                                 \
    } while (0)

int su2_run(const su2_run_config *c, su2_run_result *res)
{
    if (!c || !res) return SU2_EINVAL;
    if (c->N < 1 || c->N_term < 0 || c->N_cor < 1 || c->Nsave < 1 || c->W < 0 || c->or_sweeps < 0 ||
        c->normalize_every < 1)
        return SU2_EINVAL;
    if (c->boundary < SU2_BC_PERIODIC || c->boundary > SU2_BC_PMC) return SU2_EINVAL;
    if (c->algo != SU2_ALGO_HEATBATH && c->algo != SU2_ALGO_METROPOLIS) return SU2_EINVAL;
    if (c->readconf < 0 || c->readconf > 2) return SU2_EINVAL;
    if (c->init < SU2_START_HOT || c->init > SU2_START_HAAR) return SU2_EINVAL;

    su2_lattice l;
    RUN_CHECK(su2_lattice_init(&l, c->Lt, c->Lx, c->Ly, c->beta, c->epsilon, c->upd));
    RUN_CHECK(su2_set_device(c->device));
    const int prec = c->precision;
    const char *bound = bound_name(c->boundary);
    const int hits = (c->algo == SU2_ALGO_METROPOLIS) ? l.updates : 1; /* kernel.cu:74 */
    const int N = c->N;

    dev_buf phi, st, acc, end, end2, work, ref;
    void *stream = nullptr;
    RUN_CHECK(su2_malloc(&st.p, su2_rng_bytes(&l)));
    RUN_CHECK(su2_malloc(&phi.p, su2_field_bytes(&l, prec)));
    RUN_CHECK(su2_malloc(&acc.p, su2_accept_bytes(&l)));
    RUN_CHECK(su2_malloc(&end.p, (size_t)N * sizeof(double)));
    RUN_CHECK(su2_malloc(&end2.p, (size_t)N * sizeof(double)));
    RUN_CHECK(su2_malloc(&work.p, su2_en_av_work_bytes(&l)));
    RUN_CHECK(su2_memset_zero(work.p, su2_en_av_work_bytes(&l), stream));
    RUN_CHECK(su2_malloc(&ref.p, (size_t)l.V * 12 * sizeof(double)));
    RUN_CHECK(su2_memset_zero(end.p, (size_t)N * sizeof(double), stream));
    RUN_CHECK(su2_memset_zero(end2.p, (size_t)N * sizeof(double), stream));
    RUN_CHECK(su2_setup_kernel(&l, (uint32_t *)st.p, c->seed, stream)); /* kernel.cu:79 */

    std::vector<double> h_phi((size_t)l.V * 12);
    int Nterms = 0, inits = c->init;
    if (c->readconf == 0) { /* kernel.cu:85-136 */
        if (c->init == SU2_START_COLD)
            RUN_CHECK(su2_id_init2(&l, phi.p, prec, stream));
        else if (c->init == SU2_START_HAAR)
            RUN_CHECK(su2_haar_init(&l, phi.p, prec, (uint32_t *)st.p, c->boundary == SU2_BC_PCC, stream));
        else if (c->boundary == SU2_BC_PCC)
            RUN_CHECK(su2_rand_init2_pcc(&l, phi.p, prec, (uint32_t *)st.p, stream));
        else
            RUN_CHECK(su2_rand_init2(&l, phi.p, prec, (uint32_t *)st.p, stream));
    } else {
        RUN_CHECK(read_conf(c, h_phi, l, &Nterms, &inits, bound));
        RUN_CHECK(su2_memcpy_h2d(ref.p, h_phi.data(), h_phi.size() * sizeof(double), stream));
        RUN_CHECK(su2_field_from_ref(&l, (const double *)ref.p, phi.p, prec, stream));
        std::vector<uint32_t> hst((size_t)l.V / 2 * 6);
        int got = read_rng(c, hst, l, bound);
        if (got < 0) return got;
        if (got == 1) RUN_CHECK(su2_memcpy_h2d(st.p, hst.data(), hst.size() * sizeof(uint32_t), stream));
    }
    RUN_CHECK(su2_field_to_ref(&l, phi.p, prec, (double *)ref.p, stream));
    RUN_CHECK(su2_memcpy_d2h(h_phi.data(), ref.p, h_phi.size() * sizeof(double), stream));
    if (c->verbose) printf("Av det %.16lf\n", mean_det(h_phi));
    RUN_CHECK(su2_accept_clear(&l, (int *)acc.p, stream));

    long long updates = 0;
    auto one_sweep = [&]() -> int {
        int rc = SU2_OK;
        if (c->boundary == SU2_BC_PERIODIC)
            rc = su2_mc_up_per(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_LINK, stream);
        else if (c->boundary == SU2_BC_PCC)
            rc = su2_mc_up_pcc(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_LINK, stream);
        else
            rc = su2_mc_up_pmc(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_LINK, stream);
        if (rc) return rc;
        updates += 3ll * l.V;
        for (int k = 0; k < c->or_sweeps; ++k) {
            if (c->boundary == SU2_BC_PERIODIC)
                rc = su2_mc_up_per(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_OVERRELAX,
                                   stream);
            else if (c->boundary == SU2_BC_PCC)
                rc = su2_mc_up_pcc(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_OVERRELAX,
                                   stream);
            else
                rc = su2_mc_up_pmc(&l, phi.p, prec, (int *)acc.p, (uint32_t *)st.p, c->algo, SU2_UPDATE_OVERRELAX,
                                   stream);
            if (rc) return rc;
            updates += 3ll * l.V;
        }
        return SU2_OK;
    };

    double t0 = now_s();
    if (c->readconf < 2) { /* kernel.cu:161-187 */
        for (int i = 1; i < c->N_term + 1; ++i) {
            if (i % c->normalize_every == 0) RUN_CHECK(su2_normalize(&l, phi.p, prec, stream));
            RUN_CHECK(one_sweep());
            if (c->verbose && c->N_term >= 10 && i % (c->N_term / 10) == 0) {
                RUN_CHECK(su2_stream_synchronize(stream));
                printf("%lf Time %.3f\n", (double)i / c->N_term, now_s() - t0);
                fflush(stdout);
            }
        }
        RUN_CHECK(su2_normalize(&l, phi.p, prec, stream));
        Nterms = c->N_term + Nterms;
        if (c->sconf == 0) {
            RUN_CHECK(su2_field_to_ref(&l, phi.p, prec, (double *)ref.p, stream));
            RUN_CHECK(su2_memcpy_d2h(h_phi.data(), ref.p, h_phi.size() * sizeof(double), stream));
            RUN_CHECK(write_conf(c, h_phi, l, Nterms, inits, bound));
            std::vector<uint32_t> hst((size_t)l.V / 2 * 6);
            RUN_CHECK(su2_memcpy_d2h(hst.data(), st.p, hst.size() * sizeof(uint32_t), stream));
            RUN_CHECK(write_rng(c, hst, l, bound));
        }
    }
    RUN_CHECK(su2_stream_synchronize(stream));
    double t1 = now_s();

    RUN_CHECK(su2_accept_clear(&l, (int *)acc.p, stream)); /* kernel.cu:190 */
    std::vector<double> en((size_t)N), en2((size_t)N);
    std::vector<double> autoc((size_t)c->W > 0 ? (size_t)c->W : 1), autoc2(autoc.size());
    double energy[2] = {0, 0}, energy2[2] = {0, 0}, saccept = 0.0;
    const long long per_meas = (long long)c->N_cor * l.d * (hits + c->or_sweeps);
    const int block_s = N / c->Nsave > 0 ? N / c->Nsave : N;
    auto acceptance = [&](int i, double *out) -> int {
        double m = 0.0, e = 0.0;
        RUN_CHECK(su2_mean_lattice(&l, (const int *)acc.p, (long long)i * per_meas, &m, &e, stream));
        if (c->boundary == SU2_BC_PCC) m = m / ((double)(l.y - 1) / l.y + 1 / (double)l.y / 3); /* kernel.cu:226-228 */
        *out = m;
        return SU2_OK;
    };
    for (int i = 1; i < N + 1; ++i) { /* kernel.cu:192-236 */
        for (int j = 0; j < c->N_cor; ++j) RUN_CHECK(one_sweep());
        if (i % c->normalize_every == 0) RUN_CHECK(su2_normalize(&l, phi.p, prec, stream));
        RUN_CHECK(su2_en_av(&l, phi.p, prec, c->boundary == SU2_BC_PMC, (double *)end.p + i - 1,
                            (double *)end2.p + i - 1, work.p, stream));
        if (i % block_s == 0 && i != N) {
            RUN_CHECK(su2_memcpy_d2h(en.data(), end.p, (size_t)i * sizeof(double), stream));
            RUN_CHECK(su2_memcpy_d2h(en2.data(), end2.p, (size_t)i * sizeof(double), stream));
            mean_variance(en.data(), i, energy);
            autocorrelation(en.data(), i, c->W, energy[0], energy[1], autoc.data());
            mean_variance(en2.data(), i, energy2);
            autocorrelation(en2.data(), i, c->W, energy2[0], energy2[1], autoc2.data());
            if (c->verbose)
                printf("Energy %.10lf +- %.10lf\nPlaquette %.10lf +- %.10lf\n", energy[0],
                       std::sqrt(energy[1] / (i - 1)), energy2[0], std::sqrt(energy2[1] / (i - 1)));
            if (c->save == 0) {
                RUN_CHECK(acceptance(i, &saccept));
                RUN_CHECK(write_energy(c, energy[0], energy[1], Nterms, N, c->N_cor, saccept, l, bound));
                RUN_CHECK(write_energy_av(c, energy[0], energy[1], energy2[0], energy2[1], Nterms, N, c->N_cor,
                                          saccept, l, bound));
                RUN_CHECK(write_autocorrelation(c, "Autocorrelation_", autoc.data(), c->W, i, c->N_cor, Nterms, l,
                                                bound));
                RUN_CHECK(write_autocorrelation(c, "Autocorrelation_av_", autoc2.data(), c->W, i, c->N_cor, Nterms,
                                                l, bound));
                RUN_CHECK(write_en_seq(c, en2.data(), en.data(), Nterms, i, c->N_cor, 1.0, l, bound));
            }
        }
    }
    RUN_CHECK(su2_stream_synchronize(stream));
    double t2 = now_s();

    RUN_CHECK(su2_memcpy_d2h(en.data(), end.p, (size_t)N * sizeof(double), stream));
    RUN_CHECK(su2_memcpy_d2h(en2.data(), end2.p, (size_t)N * sizeof(double), stream));
    RUN_CHECK(acceptance(N, &saccept));
    mean_variance(en.data(), N, energy);
    autocorrelation(en.data(), N, c->W, energy[0], energy[1], autoc.data());
    mean_variance(en2.data(), N, energy2);
    autocorrelation(en2.data(), N, c->W, energy2[0], energy2[1], autoc2.data());
    if (c->save == 0) { /* kernel.cu:254-260 */
        RUN_CHECK(write_energy(c, energy[0], energy[1], Nterms, N, c->N_cor, saccept, l, bound));
        RUN_CHECK(write_energy_av(c, energy[0], energy[1], energy2[0], energy2[1], Nterms, N, c->N_cor, saccept, l,
                                  bound));
        RUN_CHECK(write_en_seq(c, en2.data(), en.data(), Nterms, N, c->N_cor, 1.0, l, bound));
        RUN_CHECK(write_autocorrelation(c, "Autocorrelation_", autoc.data(), c->W, N, c->N_cor, Nterms, l, bound));
        RUN_CHECK(write_autocorrelation(c, "Autocorrelation_av_", autoc2.data(), c->W, N, c->N_cor, Nterms, l, bound));
    }
    if (c->verbose) {
        printf("accept %lf\n\n", saccept);
        printf("Energy %.10lf +- %.10lf\n", energy[0], N > 1 ? std::sqrt(energy[1] / (N - 1)) : 0.0);
        printf("Plaquette %.10lf +- %.10lf\n", energy2[0], N > 1 ? std::sqrt(energy2[1] / (N - 1)) : 0.0);
    }
    RUN_CHECK(su2_field_to_ref(&l, phi.p, prec, (double *)ref.p, stream));
    RUN_CHECK(su2_memcpy_d2h(h_phi.data(), ref.p, h_phi.size() * sizeof(double), stream));

    if (res->energy) std::memcpy(res->energy, en.data(), (size_t)N * sizeof(double));
    if (res->plaquette) std::memcpy(res->plaquette, en2.data(), (size_t)N * sizeof(double));
    res->E_mean = energy[0];
    res->E_err = N > 1 ? std::sqrt(energy[1] / (N - 1)) : 0.0;
    res->P_mean = energy2[0];
    res->P_err = N > 1 ? std::sqrt(energy2[1] / (N - 1)) : 0.0;
    res->accept = saccept;
    res->det_mean = mean_det(h_phi);
    res->seconds_therm = t1 - t0;
    res->seconds_meas = t2 - t1;
    res->link_updates = updates;
    res->Nterms = Nterms;
    return SU2_OK;
}

} /* extern "C" */
