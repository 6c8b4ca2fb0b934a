// ba_kernels.h — launch interface between the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ba_device.hpp"

namespace navba {

constexpr int kBlock = 256;        // threads per workgroup for the per-observation kernels
constexpr int kTile = 256;         // observations per point-aligned linearisation tile
constexpr int kChunk = 1024;       // camera-major observations per Schur / Hpp work item
constexpr int kMaxSlots = 160;     // S-row column blocks per work item (LDS budget:
                                   // 160*36*8 B + 2*kMaxFreeCams B < 64 KiB)
constexpr int kNB = 64;            // dense Cholesky tile
constexpr int kMaxFreeCams = 8192; // slot map in LDS is int16
constexpr int kCamLdsMax = 512;    // Schur rows keep the camera table in LDS up to this many cameras

struct alignas(32) D4 { double x, y, z, w; };

// Scalars of the LM schedule, resident on the device.
struct LMDev {
    double lambda, nu, chi, temp_chi, scale, rho;
    int fail, accepted, trials, pad;
};

// Problem structure built once per prepare (device pointers).
struct DevProblem {
    int n_cam, n_pt, n_obs, n_fc;
    int robust;
    double delta;
    double fx, fy, cx, cy;
    // observations, point-major (by point, then camera, then input order)
    const int* o_cam;
    const int* o_pt;
    const double* o_uv;      // 2 per observation
    const int* o_orig;       // caller index
    const int* pt_ptr;       // [n_pt + 1]
    const uint8_t* pt_free;  // [n_pt]
    const int* cam_free;     // [n_cam] -> free index or -1
    const int* free_cam;     // [n_fc] -> camera
    // camera-major lists of free cameras: observation indices ascending
    const int* cam_ptr;      // [n_fc + 1]
    const int* cam_obs;
    const int* cam_obs_b0;   // per cam_obs entry: first observation of its point on a camera >= this one
    // linearisation tiles (whole points, <= kTile observations or one point)
    const int* tile_obs;     // [n_tiles + 1]
    const int* tile_pt;      // [n_tiles + 1]
    int n_tiles;
    // Hpp work items (free camera, chunk)
    const int* hitem_f;
    const int* hitem_a0;
    const int* hitem_a1;
    const int* cam_hitem_ptr;  // [n_fc + 1]
    int n_hitems;
    // Schur work items (free camera, column group, chunk), group-major per camera
    const int* sitem_f;
    const int* sitem_a0;
    const int* sitem_a1;
    const int* sitem_g;
    const int* sitem_nslots;
    const long long* sitem_pofs;
    const int* row_sitem_ptr;  // [n_fc + 1]
    const int* row_nchunks;    // [n_fc]
    int n_sitems;
    int max_ns;                // largest slot count of any Schur work item
    // column blocks of each S row: free indices >= row, first is the row itself
    const int* colptr;         // [n_fc + 1]
    const int* colcam;
    // dense reduced system
    int n, n_pad, nt;
};

struct DevState {
    PoseQT* pose;  // [n_cam]
    CamRT* cam;    // [n_cam]
    D4* X;         // [n_pt]
};

// Per-iteration / per-trial work buffers.
struct DevWork {
    void* w;            // [n_obs] T: robust weights
    void* Hl;           // [n_pt * 9] T: Hll (6 packed) + bl (3)
    void* hpart;        // [n_hitems * 27] T
    double* Hc;         // [n_fc * 28]: Hpp packed (21) + bp (6)
    double* ptsys;      // [n_pt * 12]: Dinv packed (6) + Dinv*bl (3)
    double* spart;      // Schur row partials (FP64 for both compute types)
    double* S;          // [n_pad * n_pad]
    double* g;          // [n_pad]
    double* y;          // [n_pad]
    double* dp;         // [n_pad]
    double* Lm;         // [n_pad * n_pad] Cholesky factor (lower tiles)
    double* red_cost;   // [n_cost_blocks]
    double* red_scale;  // [1 + n_pt_blocks]
    double* red_aux;    // [n_cost_blocks] outlier counts
    uint8_t* outlier;   // [n_obs] caller order
    LMDev* lm;
    int n_cost_blocks;
    int n_pt_blocks;
};

__device__ __forceinline__ int sym6(int i, int j) {
    if (i > j) { const int t = i; i = j; j = t; }
    return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

template <typename T>
__device__ __forceinline__ void load_cam(const CamRT* cam, int c, T R[9], T t[3]) {
    const double* p = reinterpret_cast<const double*>(cam + c);
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = (T)p[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = (T)p[9 + k];
}

// Launchers.  T = double (FP64 path) or float (mixed path).
template <typename T> int prepare_schur_row(const DevProblem& P);  // LDS attribute; < 0 if too large
void launch_cam_rt(const DevProblem& P, const DevState& s, hipStream_t st);
void launch_lm_reset(LMDev* lm, double lambda0, hipStream_t st);
template <typename T> void launch_lin_pt(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
template <typename T> void launch_lin_cam(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
template <typename T> void launch_lambda_init(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_hinv(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_schur_row(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
template <typename T> void launch_schur_reduce(const DevProblem& P, const DevWork& W, hipStream_t st);
void launch_cholesky_solve(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_backsub(const DevProblem& P, const DevState& cur, const DevState& cand,
                                          const DevWork& W, hipStream_t st);
void launch_update_cam(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st);
template <typename T> void launch_cost(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
void launch_lm_accept(const DevWork& W, hipStream_t st);
void launch_lm_set_chi(const DevWork& W, hipStream_t st);
void launch_commit(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st);
void launch_flags(const DevProblem& P, const DevState& s, const DevWork& W, double thr, hipStream_t st);
void launch_final(const DevWork& W, double* out2, hipStream_t st);

}  // namespace navba
