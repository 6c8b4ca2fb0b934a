// ba_kernels.h — launch interface between the host runtime and the gfx950 kernels.
//
// A prepared problem is one or more independent BA windows laid end to end:
// cameras, points and observations of window w occupy contiguous ranges, and
// every kernel covers all windows in one launch.  A single problem is the
// one-window case.  Each window carries its own LM state on the device, and
// the accept kernel runs g2o's trial loop per window, so one replayed graph
// (linearise if needed, one trial, decide) advances every window at once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ba_device.hpp"

namespace navba {

constexpr int kBlock = 256;        // threads per workgroup for the per-observation kernels
constexpr int kTile = 256;         // observations per point-aligned linearisation tile
constexpr int kChunk = 1024;       // camera-major observations per Hpp work item
constexpr int kSchurChunk = 256;   // camera-major observations per Schur work item (default; tuned per problem)
constexpr int kMaxSlots = 112;     // S-row column blocks per work item: 78 KiB of LDS, two workgroups per CU
constexpr int kPcRound = 128;      // producer/consumer Schur rows: observations per round (two producer waves)
constexpr int kPcCons = 128;       // ... and consumer lanes (two waves), each bound to one slot per item
constexpr int kNB = 64;            // tiled dense Cholesky tile
constexpr int kSmallMaxN = 180;    // per-window in-LDS Cholesky (every window of a batch, and a single
                                   // problem by default): at most 30 free cameras
constexpr int kBlocksMaxFc = 128;  // one problem: the single-workgroup block Cholesky in global memory is
                                   // possible up to this size (chosen only on request)
constexpr int kCostObs = 1024;     // observations per cost / flag block
constexpr int kXcds = 8;           // XCDs of one MI355X: workgroup i of a launch runs on XCD i % 8
constexpr int kBurstWindows = 64;  // up to this many windows the host reads their LM state between bursts
constexpr int kCommitInAcceptCams = 1024;  // windows up to this many cameras commit them inside the LM decision

// LM phase of one window
enum { kPhaseLin = 0, kPhaseTrial = 1, kPhaseDone = 2 };

struct alignas(32) D4 { double x, y, z, w; };

// Scalars of the LM schedule of one window, resident on the device.
struct LMDev {
    double lambda, nu, chi, temp_chi, scale, rho;
    int fail, accepted, trials, phase;
    int iter, q, stop;
    int xsel;  // the point buffer (DevProblem xb0 / xb1) that holds the window's current points
};

// State of the conjugate-gradient solve of one problem's reduced system
struct PcgState {
    double rz;     // r.z of the current iterate
    double rr0;    // ||g||^2
    double rr;     // ||r||^2 of the current iterate
    double alpha, beta;
    int iter, done, fail, pad;
};

struct DevProblem {
    int n_cam, n_pt, n_obs, n_fc;
    int robust;
    double delta;
    double fx, fy, cx, cy;
    int cam_model;             // CameraModel: pinhole, or Kannala-Brandt with k1..k4
    double k1, k2, k3, k4;
    // LM settings (same for every window)
    int n_iter, max_trials, auto_lambda;
    double lambda0;
    // observations, point-major (by point, then camera, then input order)
    const int* o_cam;
    const int* o_pt;
    const double* o_uv;      // 2 per observation
    const int* o_orig;       // caller index (batched: index in the concatenated caller order)
    const int* pt_ptr;       // [n_pt + 1]
    const uint8_t* pt_free;  // [n_pt]
    const int* cam_free;     // [n_cam] -> free index or -1
    const int* free_cam;     // [n_fc] -> camera
    // camera-major lists of free cameras: observation indices ascending
    const int* cam_ptr;      // [n_fc + 1]
    const int* cam_obs;
    const int* cam_obs_b0;   // per cam_obs entry: first observation of its point on a camera >= this one
    const int* c_pt;         // per cam_obs entry: its point (camera-major copy of o_pt)
    const double* c_uv;      // per cam_obs entry: its measurement, 2 values (camera-major copy of o_uv)
    // linearisation tiles (whole points of one window, <= kTile observations or one point)
    const int* tile_obs;     // [n_tiles + 1]
    const int* tile_pt;      // [n_tiles + 1]
    const int* tile_win;     // [n_tiles]
    int n_tiles;
    // Hpp work items (free camera, chunk)
    const int* hitem_f;
    const int* hitem_a0;
    const int* hitem_a1;
    const int* cam_hitem_ptr;  // [n_fc + 1]
    int n_hitems;
    // Schur work items (free camera, column group, chunk), group-major per camera;
    // each item is taken in rounds of kBlock observations
    const int* sitem_f;
    const int* sitem_a0;
    const int* sitem_a1;
    const int* sitem_g;
    const int* sitem_nslots;
    const long long* sitem_pofs;
    const int* sitem_rnd0;     // [n_sitems] first round of the item
    const int* row_sitem_ptr;  // [n_fc + 1]
    const int* row_nchunks;    // [n_fc]
    const long long* row_pofs0; // [n_fc] partial-sum offset of each row's first item (the items of a row
                                //   of one column group lie back to back)
    const int* sitem_order;    // [n_sitems] dispatch order: estimated work, heaviest first
    int schur_pc;              // 0: rounds of kBlock, lanes bound per round; 1: producer/consumer rounds
    const uint8_t* rnd_lane_slot;  // slot of each lane (slots' lanes contiguous), 255 idle: [rounds][kBlock],
                                   // or (producer/consumer) [items][kPcCons] for the consumer lanes
    const int* rnd_pair0;      // [rounds] first pair entry of the round
    const int4* rnd_rows;      // [rounds] pair rows of each wave (64 entries a row); producer/consumer: .x, .y
                               // are the two consumer waves
    const uint8_t* rnd_group;  // [rounds] lanes per group (2 or 4): a slot's lanes are whole aligned groups
    const int2* pairs;         // {observation b, row of observation a in the round}; {-1, 0} pads a lane's list
    int n_sitems;
    int max_ns;                // largest slot count of any Schur work item
    // column blocks of each S row: free indices >= row (same window), first is the row itself
    const int* colptr;         // [n_fc + 1]
    const int* colcam;
    const int* zcolptr;        // [n_fc + 1] per S row: the blocks g > f with no co-visibility
    const int* zcolcam;        //   (written as zeros by the reduction every trial)
    // windows
    int n_win;
    int max_win_fc;            // free cameras of the largest window
    int batched;               // 1: per-window dense systems factored in LDS
    const int* cam_win;        // [n_cam]
    const int* pt_win;         // [n_pt]
    const int* win_cam0;       // [n_win + 1]
    const int* win_fc0;        // [n_win + 1]
    const int* win_pt0;        // [n_win + 1]
    const long long* win_S;    // [n_win] offset of the window's dense system in S
    const int* win_lda;        // [n_win] leading dimension of the window's dense system
    const int* cblk_o0;        // [n_cblk + 1] cost blocks: window-aligned observation ranges
    const int* cblk_win;       // [n_cblk]
    const int* win_cblk0;      // [n_win + 1]
    int n_cblk;
    const int* win_tile0;      // [n_win + 1] first linearisation tile of each window
    const int* win_chol_order; // [n_win] windows by free-camera count, largest first (batched factorisation)
    // dense reduced system of one problem: tiled (n_pad = multiple of 64) or 6x6 blocks
    int n, n_pad, nt;
    int use_blocks;            // 1: single-workgroup block Cholesky in global memory
    int use_win;               // 1: per-window in-LDS Cholesky summed from the Schur partials (every batch)
    int commit_in_accept;      // 1: k_lm_accept commits accepted trials' cameras (small windows), no k_commit
    int use_pcg;               // 1: block-Jacobi conjugate gradients on the reduced system (one problem)
    // the points' two state buffers: window w's current points are in xb1
    // when its LM state's xsel is set, else in xb0, and its candidate (a
    // trial's back-substitution) in the other; an accepted trial flips xsel
    // instead of copying its points
    D4* xb0;
    D4* xb1;
};

// Window win's current and candidate points.
__device__ __forceinline__ const D4* x_cur(const DevProblem& P, const LMDev* lm, int win) {
    return lm[win].xsel ? P.xb1 : P.xb0;
}
__device__ __forceinline__ D4* x_cand(const DevProblem& P, const LMDev* lm, int win) {
    return lm[win].xsel ? P.xb0 : P.xb1;
}

struct DevState {
    PoseQT* pose;  // [n_cam]
    CamRT* cam;    // [n_cam]
    D4* X;         // [n_pt]: state buffer 0 (cur) or 1 (cand); which one holds a window's
                   // current points is its LM state's xsel
};

// Work buffers.
struct DevWork {
    void* w;            // [n_obs] T: robust weights
    double* Hl;         // [n_pt * 9]: Hll (6 packed) + bl (3)
    double* hpart;      // [n_hitems * 27]: Hpp (21 packed) + bp (6) per chunk
    double* Hc;         // [n_fc * 28]: Hpp packed (21) + bp (6)
    double* spart;      // Schur row partials (FP64 for both compute types)
    double* S;          // dense reduced systems
    double* g;          // reduced right-hand sides (free-camera order, padded for the tiled path)
    double* y;
    double* dp;
    double* Lm;         // one problem: Cholesky factor (tiled: n_pad^2; blocks: packed 6x6 blocks)
    double* Li;         // tiled path: inverses of the diagonal factor tiles [nt][64][64]
    double* fpart;      // tiled path: forward-step products L_ij y_j [nt][nt][64]
    double* red_cost;   // [n_cblk] the cost at the current state (start of a solve, flags)
    double* red_tcost;  // [n_tiles] the trial cost at the candidate state, per linearisation tile
    double* red_scale;  // [n_win + n_tiles]: per-window camera part, then per-tile point parts
    double* red_aux;    // [n_cblk] outlier counts
    uint8_t* outlier;   // [n_obs] caller order
    double* hist;       // [n_win * (n_iter + 1)] chi2 before / after each iteration
    double* final;      // [n_win * 2] FP64 final cost, outlier count
    int* done_count;    // windows finished
    LMDev* lm;          // [n_win]
    double* pcg;        // conjugate gradients (one problem): 36 per camera block, x r z p q, partial sums
    PcgState* pcg_st;
};

__device__ __forceinline__ bool needs_lin(const LMDev* lm, int w) { return lm[w].phase == kPhaseLin; }
__device__ __forceinline__ bool running(const LMDev* lm, int w) { return lm[w].phase != kPhaseDone; }

__device__ __forceinline__ int sym6(int i, int j) {
    if (i > j) { const int t = i; i = j; j = t; }
    return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

// Huber weight rho'(s) computed in FP64 and rounded to T: the value stored per
// observation and every recomputation of it agree bit for bit.
template <typename T>
__device__ __forceinline__ T huber_weight(double s, const DevProblem& P) {
    double w;
    robustify(s, P.robust, P.delta, w);
    return (T)w;
}

template <typename T>
__device__ __forceinline__ void load_cam(const CamRT* cam, int c, T R[9], T t[3]) {
    const double* p = reinterpret_cast<const double*>(cam + c);
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = (T)p[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = (T)p[9 + k];
}

// The inputs of camera c that do not depend on the step: its pose, free index
// and b_p (zero for a fixed camera); loaded ahead of the step by some callers.
struct CamIn {
    PoseQT pose;
    double bp[6];
    int f;
};

__device__ __forceinline__ CamIn camera_inputs(const DevProblem& P, const PoseQT* __restrict__ pose,
                                              const double* __restrict__ Hc, int c) {
    CamIn ci;
    ci.f = P.cam_free[c];
    ci.pose = pose[c];
    const double* bp = Hc + 28 * (size_t)(ci.f >= 0 ? ci.f : 0) + 21;
#pragma unroll
    for (int i = 0; i < 6; ++i) ci.bp[i] = ci.f >= 0 ? bp[i] : 0.0;
    return ci;
}

// Candidate pose of camera c, T <- exp(xi) T with xi = step(f, 0..5) for a
// free camera f (a copy for a fixed camera or a failed factorisation), written
// to pn / cn; returns the camera's term of x'(lambda x + b).
template <class Step>
__device__ __forceinline__ double camera_candidate(const CamIn& ci, double lambda, int fail, int c, Step step,
                                                   PoseQT* __restrict__ pn, CamRT* __restrict__ cn) {
    const int f = ci.f;
    PoseQT o = ci.pose;
    double sc = 0.0;
    if (f >= 0 && !fail) {
        double xi[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) xi[i] = step(f, i);
        se3_oplus(xi, ci.pose, o);
#pragma unroll
        for (int i = 0; i < 6; ++i) sc += xi[i] * (lambda * xi[i] + ci.bp[i]);
    }
    pn[c] = o;
    CamRT r;
    quat_to_R(o.q, r.R);
    r.t[0] = o.t[0]; r.t[1] = o.t[1]; r.t[2] = o.t[2];
    cn[c] = r;
    return sc;
}

// Launchers.  T = double (FP64 path) or float (mixed path).
template <typename T> int prepare_schur_row(const DevProblem& P);  // LDS attribute; < 0 if too large
void launch_cam_rt(const DevProblem& P, const DevState& s, hipStream_t st);
void launch_lm_reset(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_lin_pt(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
template <typename T> void launch_lin_cam(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
template <typename T> void launch_lambda_init(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_schur_row(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st);
void launch_fill_system(const DevProblem& P, const DevWork& W, hipStream_t st);  // once per prepare
template <typename T> void launch_schur_reduce(const DevProblem& P, const DevWork& W, hipStream_t st);
// The per-window factorisation and the tiled path's one-workgroup backward launch
// also form the candidate poses (cholesky_updates_cameras); otherwise launch_update_cam does.
void launch_cholesky_solve(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                           hipStream_t st);
constexpr int kBackSingleMax = 10;  // tiles up to which the backward step runs in one workgroup
inline bool cholesky_updates_cameras(const DevProblem& P) {
    return P.n_fc > 0 && (P.use_win || (!P.batched && !P.use_blocks && !P.use_pcg && P.nt <= kBackSingleMax));
}
// Every window of a batch, or a single problem of at most 30 free cameras (use_win):
// the reduced system is summed from the Schur partials inside the factorisation,
// which also forms the candidate poses.
void launch_cholesky_small(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                           hipStream_t st);
int prepare_cholesky_small(const DevProblem& P);  // LDS attribute of the per-window factorisation
int prepare_cholesky(const DevProblem& P);  // LDS attributes of the one-problem factorisation kernels
// Block-Jacobi conjugate gradients on one problem's reduced system (use_pcg):
// the start, `count` iterations (each stops at once after the stop test), and
// the step into dp.  The host reads W.pcg_st between chunks of iterations.
size_t pcg_work_doubles(int n);
void launch_pcg_start(const DevProblem& P, const DevWork& W, hipStream_t st);
void launch_pcg_iterations(const DevProblem& P, const DevWork& W, int count, double tol, int max_iter,
                           hipStream_t st);
void launch_pcg_finish(const DevProblem& P, const DevWork& W, hipStream_t st);
template <typename T> void launch_backsub(const DevProblem& P, const DevState& cur, const DevState& cand,
                                          const DevWork& W, hipStream_t st);
void launch_update_cam(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st);
// The cost at the current state; trial = 1 leaves out finished and failed windows
template <typename T> void launch_cost(const DevProblem& P, const DevState& s, const DevWork& W, int trial,
                                       hipStream_t st);
// The LM decision (an accepted trial flips the window's point buffer); with
// commit_in_accept (every window at most kCommitInAcceptCams cameras) it also
// commits the accepted cameras, and launch_commit does nothing.
void launch_lm_accept(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                      hipStream_t st);
void launch_lm_set_chi(const DevProblem& P, const DevWork& W, hipStream_t st);
void launch_commit(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st);
void launch_flags(const DevProblem& P, const DevState& s, const DevWork& W, double thr, hipStream_t st);
void launch_final(const DevProblem& P, const DevWork& W, hipStream_t st);

}  // namespace navba
