/*
 * nav_ba_debug.h — inspection hooks of the BA back-end, for parity tests.
 *
 * No counterpart in the reference (g2o keeps its linear system internal).
 * They expose intermediate results of the GPU path so that each stage can be
 * compared with the FP64 CPU oracle on its own.
 */
#ifndef NAV_BA_DEBUG_H
#define NAV_BA_DEBUG_H

#include <stdint.h>
#include "nav_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* After nav_ba_prepare: linearise at the initial state and form the reduced
 * camera system for damping `lambda`: S (n x n, both triangles) and g (n),
 * n = 6 * free cameras with observations.  *n_out receives n; nothing is
 * written when n > max_n. */
int32_t nav_ba_debug_reduced_system(nav_ba_handle* h, double lambda, double* S, double* g,
                                    int32_t max_n, int32_t* n_out);

/* Stages of the LM solve, for per-stage device timing. */
enum nav_ba_stage {
    NAV_BA_ST_LIN_PT = 0,      /* per-observation residual/Jacobian, Hll/bl per point */
    NAV_BA_ST_LIN_CAM,         /* Hpp/bp per camera (+ chunk reduction) */
    NAV_BA_ST_LAMBDA,          /* lambda initialisation */
    NAV_BA_ST_HINV,            /* unused: (Hll + lambda I)^-1 is formed inside the Schur rows and back-substitution */
    NAV_BA_ST_SCHUR_ROW,       /* reduced camera system rows */
    NAV_BA_ST_SCHUR_REDUCE,    /* partial sums -> dense S, g */
    NAV_BA_ST_CHOLESKY,        /* dense factorisation and substitutions */
    NAV_BA_ST_BACKSUB,         /* point back-substitution and candidate points */
    NAV_BA_ST_UPDATE_CAM,      /* SE3 exp-map update */
    NAV_BA_ST_COST,            /* robust chi2 */
    NAV_BA_ST_ACCEPT,          /* LM decision, commit, status */
    NAV_BA_ST_FLAGS,           /* final outlier flags and FP64 cost */
    NAV_BA_N_STAGES
};

typedef struct nav_ba_timing {
    double  ms[16];        /* accumulated device time per stage */
    int32_t launches[16];  /* timed launches per stage */
} nav_ba_timing;

/* enable != 0: later nav_ba_run calls launch the stages one by one (no
 * graphs) with HIP events on the handle's stream around each stage, and
 * accumulate the times; enable == 0 returns to graph replay.  Enabling
 * clears the accumulators. */
int32_t nav_ba_set_timing(nav_ba_handle* h, int32_t enable);
int32_t nav_ba_get_timing(const nav_ba_handle* h, nav_ba_timing* out);

/* Factorisation of a single problem's reduced system, for the next
 * nav_ba_prepare on this handle: 0 = automatic (3 up to 30 free cameras,
 * 1 above), 1 = 64x64 tiles over many workgroups, 2 = 6x6 camera blocks in
 * one workgroup with the system in global memory (falls back to tiles above
 * 128 free cameras), 3 = 6x6 camera blocks in the LDS of one workgroup,
 * summed there from the Schur partials (the batched windows' kernel; falls
 * back to tiles above 30 free cameras), 4 = block-Jacobi preconditioned
 * conjugate gradients on the dense reduced system (stop at a relative
 * residual of 1e-10 or 2000 iterations; chosen automatically above 16384
 * unknowns, up to 49152).  The factorisations give the same step up to
 * rounding, conjugate gradients up to its tolerance; the switch lets tests
 * cover each on the same problem. */
int32_t nav_ba_debug_set_solver(nav_ba_handle* h, int32_t mode);

/* The conjugate-gradient solve of the last trial of the last run (solver mode
 * 4, or a system above 16384 unknowns): its iteration count and relative
 * residual ||r|| / ||g||; 0 and 0 when the problem is factored instead. */
int32_t nav_ba_debug_pcg_stats(const nav_ba_handle* h, int32_t* iterations, double* rel_residual);

/* Observations per Schur work item for the next nav_ba_prepare on this
 * handle: 0 = automatic (problems of at least 131072 observations time the
 * candidates 128, 256, ..., 4096 with each allowed Schur-row kernel on their
 * initial linearisation at prepare and keep the fastest, remembered per
 * device, precision, kernel setting and problem shape for the life of the
 * process; smaller problems use 256), or a multiple of 64 up to 4096 to force
 * it.  Any size gives the same reduced system up to the order of FP64
 * additions. */
int32_t nav_ba_debug_set_schur_chunk(nav_ba_handle* h, int32_t chunk);
/* The chunk size the last nav_ba_prepare used (0 before a prepare, or when
 * the problem had nothing to optimise). */
int32_t nav_ba_debug_get_schur_chunk(const nav_ba_handle* h, int32_t* chunk);

/* Whether the last nav_ba_prepare (or the prepare inside nav_ba_solve) found
 * the same structure as the one before it on this handle (sizes, observation
 * indices, fixed flags, camera, options and debug settings all equal) and so
 * only uploaded the new poses, points and measurements: *cached = 1, else 0. */
int32_t nav_ba_debug_prepare_cached(const nav_ba_handle* h, int32_t* cached);

/* Schur-row kernel for the next nav_ba_prepare on this handle: 0 = automatic
 * (timed with the chunk sizes above 131072 observations; below, producer/
 * consumer for batches and rounds for one problem), 1 = rounds of 256
 * observations with the lanes bound to column blocks per round, 2 =
 * producer/consumer (two waves linearise rounds of 128 observations while two
 * waves, bound to column blocks for the whole work item, sum the previous
 * round's pairs).  Both give the same reduced system up to the order of FP64
 * additions; each run of either gives the same bits. */
int32_t nav_ba_debug_set_schur_kernel(nav_ba_handle* h, int32_t kernel);
/* The kernel the last nav_ba_prepare used: 1 or 2 (0 before a prepare, or
 * when the problem had nothing to optimise). */
int32_t nav_ba_debug_get_schur_kernel(const nav_ba_handle* h, int32_t* kernel);

/* After a prepare (one problem or a batch): linearise every window at its
 * initial state, damping lambda, and time the Schur-row kernel over all
 * windows: one warm-up launch, then `reps` launches back to back between two
 * HIP events on the handle's stream; *avg_us = their average.  For roofline
 * figures of that kernel with every window running. */
int32_t nav_ba_debug_time_schur_rows(nav_ba_handle* h, double lambda, int32_t reps, double* avg_us);

/* The same for one stage's launches (NAV_BA_ST_LIN_PT, NAV_BA_ST_LIN_CAM with
 * its chunk reduction, NAV_BA_ST_COST at the initial state, or
 * NAV_BA_ST_SCHUR_ROW): every window running and doing the stage's work, so
 * the average is the time of a launch that does it (a solve's stage timing
 * also counts the launches that a window in its trial phase skips). */
int32_t nav_ba_debug_time_stage(nav_ba_handle* h, int32_t stage, double lambda, int32_t reps, double* avg_us);

/* Host-only check of the structure a prepare builds (no device needed): the
 * flat graph arrays and the Schur work items of `kernel` (1 or 2, as
 * nav_ba_debug_set_schur_kernel) at `chunk` observations per item
 * (0 = the default), every index against the buffer it addresses and every
 * Schur round's pairs against a direct enumeration of the row's pairs.
 * Returns NAV_BA_OK when consistent (pair entries and rounds in the outputs,
 * either may be NULL), -1 with the first violation in msg, or a status for
 * invalid input. */
int32_t nav_ba_debug_check_layout(const nav_ba_problem* problems, int32_t n_windows, const nav_ba_options* o,
                                  int32_t batched, int32_t chunk, int32_t kernel, int64_t* n_pairs,
                                  int32_t* n_rounds, char* msg, int32_t msg_len);

#ifdef __cplusplus
}
#endif
#endif
