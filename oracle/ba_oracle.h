/*
 * ba_oracle.h — FP64 CPU restatement of nav24's bundle-adjustment LM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).
 * Nothing in the product library links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   graph build / write-back        core/backEnd/BE_GraphOptim.cpp:31-244
 *   reprojection edge + Jacobians   core/dataTypes/optimization/OptimizableTypes.hpp:102-113,
 *                                   OptimizableTypes.cpp:140-161
 *   pinhole projection + Jacobian   core/sensor/camera/models/Pinhole.cpp:30-36,70-80
 *   fisheye projection + Jacobian   core/sensor/camera/models/KannalaBrandt8.cpp:80-98,195-227
 *   FP32 boundary roundings         BE_GraphOptim.cpp:147,239-240; DataConversion.cpp:122-128
 * and the semantics of the un-vendored g2o library the reference calls
 * (BE_GraphOptim.cpp:40-49,66-68): Levenberg-Marquardt schedule, BlockSolver
 * <6,3> Schur complement, SE3Quat exp-map update, Huber robust kernel.
 *
 * Parity status per function is recorded in the header comment of each
 * function in ba_oracle.cpp; no reference test or golden vector exists for
 * this path, so every pin is a closed form, an invariant or a brute force.
 */
#ifndef NAV24_BA_ORACLE_H
#define NAV24_BA_ORACLE_H

#include <stdint.h>
#include "../include/nav_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Order of the oracle's long sums for the calls that follow: 0 as written
 * (edges and points ascending, the default), 1 reversed (the cost over edges
 * and the Schur complement over points descending).  The reversed order is a
 * test twin: the same algorithm with other roundings. */
void oracle_set_sum_order(int32_t order);

/* Full solve with g2o semantics.  Same contract as nav_ba_solve. */
int32_t oracle_ba_solve(const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r,
                        double* lambda_history /* [n_iter] lambda at start of each iteration, or NULL */,
                        int32_t* trials_history /* [n_iter] or NULL */);

/* The same solve, recording every LM trial: trace[k][7] = {iteration, trial
 * within the iteration, lambda used, nu at the trial, rho, chi2 before the
 * trial, chi2 of the candidate}, at most max_trace trials; *n_trace = count. */
int32_t oracle_ba_solve_trace(const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r,
                              double* trace, int32_t max_trace, int32_t* n_trace);

/* Robust chi2 at the state (T_cw, X) given in caller layout; optional per-observation
 * chi2 (unweighted e'e) and outlier flags. */
double oracle_cost(const nav_ba_problem* p, const nav_ba_options* o,
                   const double* T_cw, const double* X,
                   double* chi2_per_obs, uint8_t* outlier);

/* Linearisation blocks at the problem's initial state, caller indexing; rows of
 * fixed or inactive vertices are zero.  Hpp [n_cam][36], bp [n_cam][6],
 * Hll [n_pt][9], bl [n_pt][3].  Returns robust chi2. */
double oracle_linear_blocks(const nav_ba_problem* p, const nav_ba_options* o,
                            double* Hpp, double* bp, double* Hll, double* bl);

/* Reduced camera system at the initial state for a given lambda:
 * S [n][n] (n = 6 * number of free active cameras, camera order), g [n].
 * Returns n; writes at most max_n*max_n entries. */
int32_t oracle_reduced_system(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                              double* S, double* g, int32_t max_n);

/* One LM step at the initial state through the Schur path (dx = [poses; points]
 * in index order).  Returns the number of unknowns, or -1 if the factorisation failed. */
int32_t oracle_schur_step(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                          double* dx, int32_t max_n);

/* The same step by brute force: the full (6 Nc + 3 Np) system with lambda on the
 * diagonal, solved densely in long double.  Pins the Schur path. */
int32_t oracle_dense_step(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                          double* dx, int32_t max_n);

/* Math kernels exposed for unit tests. */
void oracle_quat_from_R(const double R[9], double q[4] /* x y z w */);
void oracle_R_from_quat(const double q[4], double R[9]);
void oracle_se3_exp(const double xi[6] /* omega, upsilon */, double q[4], double t[3]);
/* T_out = exp(xi) * T_in, poses as (q xyzw, t). */
void oracle_se3_oplus(const double xi[6], const double q_in[4], const double t_in[3],
                      double q_out[4], double t_out[3]);
void oracle_huber(double s, double delta, double rho[3]);
void oracle_inv3(const double A[9], double Ainv[9]);
/* Edge at pose (q, t), point X: residual e (2), Jacobians J_X (2x3 row-major) and
 * J_xi (2x6 row-major, columns omega then upsilon), camera point Xc (3). */
void oracle_edge(const double q[4], const double t[3], const double X[3],
                 double fx, double fy, double cx, double cy, const double uv[2],
                 double e[2], double J_X[6], double J_xi[12], double Xc[3]);
/* The same edge with the Kannala-Brandt camera, K8 = fx fy cx cy k1 k2 k3 k4. */
void oracle_edge_kb8(const double q[4], const double t[3], const double X[3], const double K8[8],
                     const double uv[2], double e[2], double J_X[6], double J_xi[12], double Xc[3]);

#ifdef __cplusplus
}
#endif
#endif
