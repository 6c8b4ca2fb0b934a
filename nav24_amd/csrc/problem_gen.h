/*
 * problem_gen.h — seeded synthetic bundle-adjustment problems.
 *
 * Produces the workloads the solver is measured and tested on: circle of
 * keyframes looking at a cube of points, random-walk trajectory in a box,
 * 200 m loop, batched uneven windows, and distorted-camera problems for the
 * differentiable cost.  Deterministic across platforms: std::mt19937_64 (fully
 * specified by the C++ standard) with this file's own uniform and Box-Muller
 * normal samplers (std::normal_distribution is implementation-defined).
 */
#ifndef NAV24_PROBLEM_GEN_H
#define NAV24_PROBLEM_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { NAV_GEN_CIRCLE = 0, NAV_GEN_RANDOM_WALK = 1, NAV_GEN_LOOP = 2 };

typedef struct nav_gen_params {
    int32_t  geometry;        /* NAV_GEN_* */
    int32_t  n_cam;
    int32_t  n_pt;
    double   noise_px;        /* pixel noise sigma */
    double   outlier_rate;    /* fraction of observations replaced by uniform pixels */
    double   rot_sigma;       /* initial pose rotation perturbation (rad, per axis) */
    double   trans_sigma;     /* initial pose translation perturbation (m, per axis) */
    double   point_sigma;     /* initial point perturbation (m, per axis) */
    int32_t  n_fixed;         /* first n_fixed cameras fixed (and left unperturbed) */
    int32_t  obs_min;         /* 0: every visible camera observes; else per point */
    int32_t  obs_max;         /*    U[obs_min, min(obs_max, visible)] cameras */
    int32_t  distortion;      /* 1: radial-tangential k1 k2 p1 p2 ~ U(-0.1, 0.1) */
    uint64_t seed;
} nav_gen_params;

typedef struct nav_gen_problem {
    int32_t  n_cam, n_pt, n_obs;
    double*  T_cw;            /* [n_cam][12] initial (perturbed) poses, R row-major + t */
    uint8_t* cam_fixed;       /* [n_cam] */
    double*  X;               /* [n_pt][3] initial (perturbed) points, FP32-representable */
    uint8_t* pt_fixed;        /* [n_pt] */
    int32_t* obs_cam;         /* [n_obs] camera-major order */
    int32_t* obs_pt;
    double*  obs_uv;          /* [n_obs][2], FP32-representable */
    uint8_t* obs_outlier;     /* [n_obs] 1 where the measurement was replaced */
    double*  T_gt;            /* [n_cam][12] ground truth */
    double*  X_gt;            /* [n_pt][3] ground truth */
    double   fx, fy, cx, cy;
    double   dist[4];         /* k1 k2 p1 p2 (zero unless distortion) */
    int32_t  width, height;
} nav_gen_problem;

/* Parameters of the named workloads:
 *   0  circle r=2 m, 10 cams, 1500 pts, all visible obs, 5% outliers, 2 fixed
 *   1  random walk, 60 cams, 20000 pts in 40x40x10 m, 5% outliers, 2 fixed
 *   2  loop r=30 m, 300 cams, 150000 pts within 15 m, 3% outliers, 2 fixed
 *   3  one batched window: cams U[5,30], pts U[300,5000], obs/pt U[2,min(8,cams)]
 *      (sizes drawn from `seed`; window w uses seed = base + w)
 *   4  circle geometry, 20 cams, 5000 pts, radial-tangential distortion
 * Returns 0, or 2 for an unknown workload. */
int32_t nav_gen_params_for(int32_t workload, uint64_t seed, nav_gen_params* out);

/* Generate; *out is owned by the library (nav_gen_free).  Returns 0 or 2. */
int32_t nav_gen_make(const nav_gen_params* p, nav_gen_problem** out);
void    nav_gen_free(nav_gen_problem* p);

/* Known-answer helpers for the samplers: k-th raw draw of mt19937_64(seed),
 * and the first n normals of the generator's Box-Muller stream. */
uint64_t nav_gen_mt_draw(uint64_t seed, int32_t k);
void     nav_gen_normals(uint64_t seed, int32_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
