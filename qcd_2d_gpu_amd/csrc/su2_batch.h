/*
 * su2_batch.h - device work list of the batched launches (su2_batch.hip)
 * and their host launchers.
 */
#ifndef SU2_BATCH_H
#define SU2_BATCH_H

#include "su2_internal.h"

/* One resident lattice.  All pointers are device pointers into the engine's
 * arena. */
struct su2_bdesc {
    void *U;        /* field, SoA layout of su2_internal.h */
    uint32_t *rng;  /* 6 planes of Vh */
    int *accept;    /* 2 Vh */
    double *part;   /* en_av partials, 2 per logical block */
    double *partc;  /* the same, of the measurement carried by the first update launch of a tick */
    double *resE;   /* n_meas energies */
    double *resP;   /* n_meas plaquettes */
    su2_kgeom g;
    su2_kphys ph;
    int bc, algo, slot;
    int n_therm, n_cor, n_meas, nevery, or_sweeps;
    uint32_t sw_begin; /* first virtual block in the sweep launches */
    uint32_t en_begin; /* first virtual block in the en_av launch */
    uint32_t en_nb;    /* en_av logical blocks of this lattice */
    uint32_t nm_begin; /* first virtual block in the normalise launch */
};

struct su2_bhead {
    uint32_t n_active;  /* descriptors in use */
    uint32_t sw_blocks; /* total virtual blocks of a sweep launch */
    uint32_t en_blocks;
    uint32_t nm_blocks;
};

/* The first virtual block of every descriptor, per launch kind, as compact
 * arrays (sw, en, nm; max_desc entries each) that workgroups stage in LDS
 * for the block -> lattice search. */
#define SU2_BATCH_LDS_DESC 1024

/* algorithm classes of the resident jobs for the update launches */
enum { SU2_BALGO_METRO = 0, SU2_BALGO_HB = 1, SU2_BALGO_MIXED = 2 };

/* The end of a tick carried by its last sweep launch: lead leading
 * workgroups (a multiple of 8) add the pending measurements and advance the
 * sweep counters and pending flags, the work of su2_launch_ben_av_final when
 * no lattice measures at the end of the tick. */
struct su2_btick_end {
    uint32_t *ctr, *pend;
    uint32_t lead;
    int carry_on;
};
/* measure: the y launch of parity 0 also forms the carried partials of the
 * lattices with a pending measurement (pend[slot], the previous sweep's),
 * for workgroups of SU2_EN_BLOCK threads only.  end: the launch (mu = t,
 * parity 1, SU2_EN_BLOCK threads, the tick's last sweep launch) also ends
 * the tick. */
hipError_t su2_launch_bsweep(int f32, int mu, int over_relax, int algo_class, const su2_bhead *H,
                             const su2_bdesc *D, const uint32_t *B, uint32_t max_desc, int parity, int or_pass,
                             uint32_t grid, uint32_t block, hipStream_t st, const uint32_t *pend = nullptr,
                             int measure = 0, const su2_btick_end *end = nullptr);
hipError_t su2_launch_bnormalize(int f32, const su2_bhead *H, const su2_bdesc *D, const uint32_t *B,
                                 uint32_t max_desc, const uint32_t *ctr, int post, uint32_t grid, hipStream_t st);
/* en_av pass 1: pre = 0 at the end of a tick, for the lattices that
 * measure there (not deferred: carry_on is whether the tick's sweep
 * launches run workgroups of SU2_EN_BLOCK threads), into their partials;
 * pre = 1 before a tick's sweeps, for the lattices with a pending
 * measurement, into their carried partials (a tick whose sweep launches
 * cannot carry them). */
hipError_t su2_launch_ben_av_partial(int f32, const su2_bhead *H, const su2_bdesc *D, const uint32_t *B,
                                     uint32_t max_desc, const uint32_t *ctr, const uint32_t *pend, uint32_t grid,
                                     int carry_on, int pre, hipStream_t st);
/* en_av pass 2 of the pending and the end-of-tick measurements; it sets
 * the pending flags of the deferred ones and advances every sweep counter,
 * so a measuring tick needs no su2_launch_btick_end. */
hipError_t su2_launch_ben_av_final(const su2_bhead *H, const su2_bdesc *D, uint32_t *ctr, uint32_t *pend,
                                   uint32_t final_grid, int carry_on, hipStream_t st);
hipError_t su2_launch_btick_end(const su2_bhead *H, const su2_bdesc *D, uint32_t *ctr, hipStream_t st);
uint32_t su2_en_av_logical_blocks(const su2_kgeom &g);

#endif /* SU2_BATCH_H */
