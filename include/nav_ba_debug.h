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

#ifdef __cplusplus
}
#endif
#endif
