/*
 * nav_ba.h — C ABI of the MI355X bundle-adjustment back-end.
 *
 * This library replaces the reference back-end solver
 *   NAV24::BE::GraphOptim::solve(const ProblemPtr&)      core/backEnd/BE_GraphOptim.cpp:31-77
 * behind the interface
 *   NAV24::BE::BackEnd::solve(const ProblemPtr&)          core/backEnd/BackEnd.hpp:15
 * with a flat, caller-owned structure-of-arrays problem instead of the
 * reference's shared-pointer object graph (PR_VBA, core/backEnd/Problem.hpp:41-47).
 *
 * Plain C: no torch, HIP or C++ types in any signature.  Every entry point
 * returns an nav_ba_status; nothing throws across the ABI.  A handle owns one
 * HIP stream and its device buffers; handles are independent, so one handle
 * per host thread is safe (the reference solve is synchronous on the caller's
 * thread, FE_SlamMonoV.cpp:164-165).
 *
 * ABI version 2 added the camera model and k1..k4 at the end of
 * nav_ba_problem; a zero-initialised problem is the version-1 pinhole.
 */
#ifndef NAV_BA_H
#define NAV_BA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAV_BA_ABI_VERSION 2

typedef enum nav_ba_status {
    NAV_BA_OK = 0,
    NAV_BA_EMPTY = 1,            /* no pose or no point in the graph (BE_GraphOptim.cpp:59-63) */
    NAV_BA_BAD_ARG = 2,          /* null pointer, negative size, index out of range */
    NAV_BA_NOT_PD_TERMINATED = 3,/* LM stopped after 10 rejected trials (factorisation failures) */
    NAV_BA_HIP_ERROR = 4,        /* a HIP runtime call failed; see nav_ba_last_error */
    NAV_BA_NO_DEVICE = 5,        /* no usable gfx950 device */
    NAV_BA_TOO_LARGE = 6,        /* problem exceeds a documented size limit */
    NAV_BA_NOT_PREPARED = 7      /* nav_ba_run / nav_ba_fetch before nav_ba_prepare */
} nav_ba_status;

typedef enum nav_ba_precision {
    NAV_BA_FP64 = 0,   /* every stage in FP64 */
    NAV_BA_MIXED = 1   /* FP32 residuals/Jacobians/Hessian blocks; FP64 Hll^-1, reduced system, update */
} nav_ba_precision;

/* Camera model of the reprojection edge: the model the reference's
 * Calibration loads (Calibration.cpp:78-87) and the edge reaches through
 * Calibration::project / projectJac (Calibration.cpp:235-247). */
typedef enum nav_ba_camera {
    NAV_BA_CAM_PINHOLE = 0,  /* Pinhole::project / projectJac (Pinhole.cpp:30-36,70-80); also the
                                radial-tangential camera, whose edge sees undistorted key points */
    NAV_BA_CAM_KB8 = 1       /* KannalaBrandt8::project / projectJac (KannalaBrandt8.cpp:80-98,195-227):
                                theta = atan2f of float arguments in the residual, double in the
                                Jacobian; undefined on the optical axis (x = y = 0), as there */
} nav_ba_camera;

typedef enum nav_ba_stop {
    NAV_BA_STOP_MAX_ITER = 0,    /* ran all requested outer iterations */
    NAV_BA_STOP_TERMINATE = 1,   /* LM returned Terminate: 10 trials or rho == 0 */
    NAV_BA_STOP_EMPTY = 2
} nav_ba_stop;

/*
 * One BA problem, caller-owned, read only during the call.
 *
 * Camera poses are T_cw (world -> camera), the pose a Frame carries
 * (core/trajectory/pose/Pose.hpp:95,104); the reprojection edge maps world
 * points with it (OptimizableTypes.hpp:102-107).  Observations are the
 * undistorted key-point coordinates (Point2D.hpp:23) linking one camera and
 * one point — the weak observation links of WorldObject.hpp:26.
 */
typedef struct nav_ba_problem {
    int32_t n_cam;
    int32_t n_pt;
    int32_t n_obs;
    const double*  T_cw;      /* [n_cam][12]: R row-major (9) then t (3) */
    const uint8_t* cam_fixed; /* [n_cam] SmartObject::isOptFixed (BE_GraphOptim.cpp:95); may be NULL = none */
    const double*  X;         /* [n_pt][3] world points (Point3D.hpp:28) */
    const uint8_t* pt_fixed;  /* [n_pt] (BE_GraphOptim.cpp:150-153); may be NULL = none */
    const int32_t* obs_cam;   /* [n_obs] camera index of each observation */
    const int32_t* obs_pt;    /* [n_obs] point index of each observation */
    const double*  obs_uv;    /* [n_obs][2] the key point the edge measures (Point2D.hpp:23) */
    double fx, fy, cx, cy;    /* intrinsics (Pinhole.cpp:30-36; mvParameters[0..3] of either model) */
    int32_t camera_model;     /* nav_ba_camera; 0 (zero-initialised) = pinhole */
    double k1, k2, k3, k4;    /* NAV_BA_CAM_KB8: mvParameters[4..7]; ignored for the pinhole */
} nav_ba_problem;

/* Solver settings.  The reference hard-codes these (its defaults:
 * nIterations Problem.hpp:18,36, mbRobust Problem.hpp:46, Huber delta
 * BE_GraphOptim.cpp:125); here they are explicit. */
typedef struct nav_ba_options {
    int32_t n_iter;         /* outer LM iterations, SparseOptimizer::optimize(n) (BE_GraphOptim.cpp:68) */
    int32_t robust;         /* Huber kernel on every edge (PR_VBA::mbRobust) */
    double  huber_delta;    /* Huber threshold delta (reference: (float)sqrt(5.99)) */
    double  lambda_init;    /* > 0: user initial lambda; <= 0: 1e-5 * max|diag H| */
    int32_t precision;      /* nav_ba_precision */
    int32_t compat_fp32_io; /* reproduce the reference's FP32 boundary roundings:
                               X in/out (BE_GraphOptim.cpp:147,239), uv (Point2D.hpp:33),
                               intrinsics (GeometricCamera.h:117) */
    double  outlier_chi2;   /* per-observation flag threshold on chi2 (5.991) */
    int32_t max_trials;     /* LM trials per iteration before Terminate (g2o: 10); <= 0 -> 10 */
} nav_ba_options;

/* Results and statistics.  Output arrays are caller-owned; any may be NULL. */
typedef struct nav_ba_result {
    double*  T_cw;          /* [n_cam][12] refined poses (BE_GraphOptim.cpp:213-229) */
    double*  X;             /* [n_pt][3] refined points (BE_GraphOptim.cpp:231-244); points
                               without observations are copied unchanged */
    uint8_t* outlier;       /* [n_obs] 1 if chi2 > outlier_chi2 or depth <= 0 at the final
                               state, in the caller's observation order */
    double*  cost_history;  /* [n_iter + 1] chi2 before iteration 0 and after each iteration */
    int32_t  iterations;    /* outer iterations run */
    int32_t  trials;        /* LM trials in total */
    int32_t  stop_reason;   /* nav_ba_stop */
    int32_t  n_outliers;
    double   initial_cost;  /* robust chi2 at the initial state */
    double   final_cost;    /* robust chi2 at the final state (FP64) */
    double   final_lambda;
    double   solve_ms;      /* device time of the LM solve */
} nav_ba_result;

typedef struct nav_ba_handle nav_ba_handle;

/* Library version string; never NULL. */
const char* nav_ba_version(void);

/* Create a handle bound to HIP device `device` (owns one HIP stream). */
int32_t nav_ba_create(nav_ba_handle** out, int32_t device);
void    nav_ba_destroy(nav_ba_handle* h);
/* Text of the last error on this handle ("" if none); valid until the next call. */
const char* nav_ba_last_error(const nav_ba_handle* h);

/* Fill `o` with the reference's defaults: 5 iterations, robust off,
 * delta (float)sqrt(5.99), lambda from 1e-5*max diag, FP64, compat on,
 * flag threshold 5.991, 10 trials. */
void nav_ba_default_options(nav_ba_options* o);

/* One call = GraphOptim::solve: upload, build structure, run the LM solve,
 * write results back.  Replaces BE_GraphOptim.cpp:31-77. */
int32_t nav_ba_solve(nav_ba_handle* h, const nav_ba_problem* p,
                     const nav_ba_options* o, nav_ba_result* r);

/* The same solve split in three, so that the problem can stay resident in
 * HBM across repeated solves:  prepare = the graph build of
 * BE_GraphOptim.cpp:51-63 (upload, ordering, block structure);  run = the
 * optimize() call of :66-68 from the uploaded initial state;  fetch = the
 * write-back of :71-76. */
int32_t nav_ba_prepare(nav_ba_handle* h, const nav_ba_problem* p, const nav_ba_options* o);
int32_t nav_ba_run(nav_ba_handle* h);
int32_t nav_ba_fetch(nav_ba_handle* h, nav_ba_result* r);

/* A batch of independent local windows solved together: the loop of
 * per-window GraphOptim::solve calls the front end makes on each local map
 * (FE_SlamMonoV.cpp:161-165, BE_GraphOptim.cpp:31-77 once per window).  Every
 * window gets its own LM schedule (lambda, nu, trials, termination) exactly as
 * if it were solved alone; all windows share the options and the intrinsics.
 * A window may hold at most 30 free cameras (NAV_BA_TOO_LARGE otherwise).
 * `results[w]` receives window w's results, with the
 * same meaning as for nav_ba_fetch; solve_ms is the time of the whole batch.
 * nav_ba_run runs a prepared batch; an empty window reports NAV_BA_STOP_EMPTY
 * and fetch returns NAV_BA_EMPTY if any window was empty (the others are
 * still written). */
int32_t nav_ba_prepare_batched(nav_ba_handle* h, const nav_ba_problem* windows, int32_t n_windows,
                               const nav_ba_options* o);
int32_t nav_ba_fetch_batched(nav_ba_handle* h, nav_ba_result* results, int32_t n_windows);
int32_t nav_ba_solve_batched(nav_ba_handle* h, const nav_ba_problem* windows, int32_t n_windows,
                             const nav_ba_options* o, nav_ba_result* results);

#ifdef __cplusplus
}
#endif
#endif /* NAV_BA_H */
