/*
 * The CPU checker (oracle/su2_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: every start, boundary, algorithm, update kind
 * and storage precision on small and ragged lattices (extents down to 2,
 * Ly/2 odd), plus normalize, en_av and the timed sweep loop.  Built by the
 * Makefile beside it with gcc (host code only); any report aborts the run.
 */
#include "../../oracle/su2_oracle.c"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static int run(int Lt, int Lx, int Ly, int f32)
{
    const int V = Lt * Lx * Ly, Vh = V / 2;
    orc_lattice l2 = {Ly / 2, Lx, Lt, Vh, 3, 4, 12, 2.5f, 0.4f, 3};
    double *phi = (double *)malloc(sizeof(double) * 12 * (size_t)V);
    orc_rng *st = (orc_rng *)malloc(sizeof(orc_rng) * (size_t)Vh);
    int *acc = (int *)calloc((size_t)V, sizeof(int));
    if (!phi || !st || !acc) return 1;
    for (int start = 0; start < 5; ++start) {
        orc_setup_kernel(st, Vh, 99u + (uint64_t)start);
        orc_init(phi, st, &l2, start, f32);
        for (int bc = 0; bc < 3; ++bc)
            for (int algo = 0; algo < 2; ++algo)
                for (int update = 0; update < 2; ++update) {
                    orc_mc_up(phi, &l2, acc, st, bc, algo, update, f32);
                    for (int mu = 0; mu < 3; ++mu)
                        orc_mc_update(phi, &l2, acc, st, bc, algo, update, mu, (mu + start) & 1, f32);
                }
        orc_normalize(phi, 3ll * V, f32);
        for (int pmc = 0; pmc < 2; ++pmc) {
            double e = 0.0, p = 0.0;
            orc_en_av(phi, &l2, pmc, &e, &p);
            if (!isfinite(e) || !isfinite(p) || p <= -1.0 || p > 1.0) {
                fprintf(stderr, "bad observables %d %d %d: %g %g\n", Lt, Lx, Ly, e, p);
                return 1;
            }
        }
    }
    orc_time_sweeps(phi, &l2, acc, st, 2, 0, 1, 2, f32);
    free(phi);
    free(st);
    free(acc);
    return 0;
}

int main(void)
{
    const int shapes[][3] = {{2, 2, 2}, {4, 6, 2}, {4, 4, 10}, {6, 2, 4}, {2, 8, 6}};
    for (size_t i = 0; i < sizeof shapes / sizeof shapes[0]; ++i)
        for (int f32 = 0; f32 < 2; ++f32)
            if (run(shapes[i][0], shapes[i][1], shapes[i][2], f32)) return 1;
    printf("sanitized checker ok\n");
    return 0;
}
