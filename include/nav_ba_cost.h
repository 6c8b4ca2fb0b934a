/*
 * nav_ba_cost.h — differentiable robust reprojection cost on the device.
 *
 * The total robust cost of a batch of BA problems that share one camera,
 *     C_p = sum_{obs of problem p} rho(|z - pi_d(T_cw X)|^2),
 * and its gradient with respect to every pose (6-vector tangent, left
 * perturbation T <- exp(xi) T, xi = (omega, upsilon): the update and column
 * order of the reference edge, OptimizableTypes.cpp:140-161 and
 * OptimizableTypes.hpp:221-227), every point, and the shared intrinsics
 * (fx, fy, cx, cy) and radial-tangential distortion (k1, k2, p1, p2).
 * pi_d is the pinhole projection of Pinhole.cpp:30-36 applied after the
 * Brown-Conrady distortion of the normalised coordinates,
 *     x_d = x (1 + k1 r^2 + k2 r^4) + 2 p1 x y + p2 (r^2 + 2 x^2),
 *     y_d = y (1 + k1 r^2 + k2 r^4) + p1 (r^2 + 2 y^2) + 2 p2 x y,
 * the model the reference's rad-tan camera removes from key points before BA
 * (PinholeRadTan.cpp:11-24).  rho is the Huber kernel of the BA edge
 * (BE_GraphOptim.cpp:191-195) or the identity.
 *
 * No counterpart exists in the reference (it never differentiates the cost
 * with respect to intrinsics or distortion); the gradient with respect to
 * poses and points is -2 rho' J^T e with the edge Jacobians J of the BA edge
 * (OptimizableTypes.cpp:140-161) when the distortion is zero.
 *
 * The structure (which camera and point each observation links, and the
 * measurements) is uploaded once; evaluations read the state from and write
 * the results to device memory (e.g. torch tensors) on a caller-given HIP
 * stream, so nothing crosses PCIe per evaluation.  Sums are in fixed order:
 * results are deterministic.
 */
#ifndef NAV_BA_COST_H
#define NAV_BA_COST_H

#include <stdint.h>
#include "nav_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nav_ba_cost_handle nav_ba_cost_handle;

enum nav_ba_cost_dtype { NAV_BA_COST_F64 = 0, NAV_BA_COST_F32 = 1 };

int32_t nav_ba_cost_create(nav_ba_cost_handle** out, int32_t device);
void    nav_ba_cost_destroy(nav_ba_cost_handle* h);
const char* nav_ba_cost_last_error(const nav_ba_cost_handle* h);

/* Structure of a batch of problems laid end to end: problem p's cameras and
 * points follow those of problems 0..p-1 in the state arrays.  Only n_cam,
 * n_pt, n_obs, obs_cam, obs_pt and obs_uv of each nav_ba_problem are read
 * (host arrays); poses, points and intrinsics come with each evaluation. */
int32_t nav_ba_cost_prepare(nav_ba_cost_handle* h, const nav_ba_problem* problems, int32_t n_problems,
                            int32_t robust, double huber_delta);

/* One evaluation on `stream` (a hipStream_t, NULL = the null stream), all
 * arrays in device memory of type dtype (double or float):
 *   T_cw    [n_cam][12]  R row-major, then t (all problems' cameras)
 *   X       [n_pt][3]
 *   intr    [8]          fx fy cx cy k1 k2 p1 p2
 *   cost    [n_problems] total robust cost per problem (always written)
 *   g_pose  [n_cam][6]   d cost / d xi (omega, upsilon), or NULL
 *   g_X     [n_pt][3]    d cost / d X (zero for points without observations), or NULL
 *   g_intr  [8]          d (sum of costs) / d intr, or NULL
 * Gradients are computed when any of the three is non-NULL. */
int32_t nav_ba_cost_eval(nav_ba_cost_handle* h, int32_t dtype, const void* T_cw, const void* X, const void* intr,
                         void* cost, void* g_pose, void* g_X, void* g_intr, void* stream);

/* Sizes of the prepared batch. */
int32_t nav_ba_cost_sizes(const nav_ba_cost_handle* h, int32_t* n_cam, int32_t* n_pt, int32_t* n_obs,
                          int32_t* n_problems);

#ifdef __cplusplus
}
#endif
#endif /* NAV_BA_COST_H */
