/*
 * su2mc_main.c - command-line front end of su2_run (C, on the C ABI): the reference's
 * compile-time #defines (kernel.cu:10-46, README.md:37-50) as flags, e.g.
 *   su2mc --Lt 16 --Lx 16 --Ly 16 --beta 3.0 --boundary 0 --algo 1 --init 1 \
 *         --N 100 --N_term 100 --seed 1234 --save 0
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "su2mc.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: su2mc [--Lt n] [--Lx n] [--Ly n] [--beta f] [--boundary 0|1|2] [--algo 0|1]\n"
            "             [--readconf 0|1|2] [--init 0|1|2] [--sconf 0|1] [--save 0|1] [--N n] [--N_term n]\n"
            "             [--N_cor n] [--Nsave n] [--epsilon f] [--upd n] [--W n] [--or n] [--precision 0|1]\n"
            "             [--normalize_every n] [--seed n] [--outdir dir] [--nsave s] [--nconfr s]\n"
            "             [--nconfs s] [--device n] [--N_block n] [--Nblock_term n] [--fp64_twin 0|1] [--quiet]\n");
}

int main(int argc, char **argv)
{
    su2_run_config c;
    su2_run_config_default(&c);
    c.verbose = 1;
    for (int i = 1; i < argc; ++i) {
        const char *k = argv[i];
        if (!strcmp(k, "--quiet")) {
            c.verbose = 0;
            continue;
        }
        if (!strcmp(k, "--help") || !strcmp(k, "-h")) {
            usage();
            return 0;
        }
        if (i + 1 >= argc) {
            usage();
            return 2;
        }
        const char *v = argv[++i];
#define INT_FLAG(name, field)                                                                                          \
    if (!strcmp(k, "--" name)) {                                                                                       \
        c.field = atoi(v);                                                                                             \
        continue;                                                                                                      \
    }
        INT_FLAG("Lt", Lt)
        INT_FLAG("Lx", Lx)
        INT_FLAG("Ly", Ly)
        INT_FLAG("boundary", boundary)
        INT_FLAG("algo", algo)
        INT_FLAG("readconf", readconf)
        INT_FLAG("init", init)
        INT_FLAG("sconf", sconf)
        INT_FLAG("save", save)
        INT_FLAG("N", N)
        INT_FLAG("N_term", N_term)
        INT_FLAG("N_cor", N_cor)
        INT_FLAG("Nsave", Nsave)
        INT_FLAG("upd", upd)
        INT_FLAG("W", W)
        INT_FLAG("or", or_sweeps)
        INT_FLAG("precision", precision)
        INT_FLAG("normalize_every", normalize_every)
        INT_FLAG("device", device)
        INT_FLAG("N_block", N_block)
        INT_FLAG("Nblock_term", Nblock_term)
        INT_FLAG("fp64_twin", fp64_twin)
#undef INT_FLAG
        if (!strcmp(k, "--beta")) c.beta = (float)atof(v);
        else if (!strcmp(k, "--epsilon")) c.epsilon = (float)atof(v);
        else if (!strcmp(k, "--seed")) c.seed = strtoull(v, NULL, 10);
        else if (!strcmp(k, "--outdir")) c.outdir = v;
        else if (!strcmp(k, "--nsave")) c.nsave = v;
        else if (!strcmp(k, "--nconfr")) c.nconfr = v;
        else if (!strcmp(k, "--nconfs")) c.nconfs = v;
        else {
            usage();
            return 2;
        }
    }
    printf("N %d Lt %d L %d Ly %d V %d\n", c.N, c.Lt, c.Lx, c.Ly, c.Lt * c.Lx * c.Ly);
    su2_run_result r;
    memset(&r, 0, sizeof r);
    int rc = su2_run(&c, &r);
    if (rc != SU2_OK) {
        fprintf(stderr, "su2mc: %s (%s)\n", su2_strerror(rc), su2_last_error());
        return 1;
    }
    printf("Time therm %.3f s meas %.3f s, %.3e link updates/s\n", r.seconds_therm, r.seconds_meas,
           (double)r.link_updates / (r.seconds_therm + r.seconds_meas));
    if (c.precision == SU2_PREC_F32 && c.fp64_twin)
        printf("FP32 links: plaquette %.10f +- %.10f, FP64 twin %.10f +- %.10f, difference %.3f sigma\n", r.P_mean,
               r.P_err_binned, r.P_fp64_twin, r.P_fp64_twin_err,
               r.P_diff_sigma > 0 ? r.P_diff_fp32_fp64 / r.P_diff_sigma : 0.0);
    return 0;
}
