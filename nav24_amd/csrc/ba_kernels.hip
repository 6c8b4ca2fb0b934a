// ba_kernels.hip — gfx950 kernels of the BA Levenberg-Marquardt solve.
//
// One replay of the step graph advances every window by one LM trial (g2o
// OptimizationAlgorithmLevenberg::solve, the optimizer nav24 runs at
// core/backEnd/BE_GraphOptim.cpp:66-68):
//   lin_pt     windows starting an iteration: one lane per observation in
//              point-aligned tiles; residual, Huber weight, Hll / bl per point
//              (segment sums in LDS)                                     [HBM-bound]
//   lin_cam    Hpp / bp per free camera from camera-major chunks          [HBM-bound]
//   lambda     first iteration without a user lambda: 1e-5 max |diag H|
//   schur_row  reduced camera system rows (schur_kernels.hip); each forms
//              the damped point inverse (Hll + lambda I)^-1 itself, FP64
//   schur_red  partials -> dense S (FP64) and g per window
//   cholesky   tiled (one large window) or in-LDS per window (chol_kernels.hip)
//   update_cam T <- exp(dxi) T (FP64 SE3 exp-map), candidate cameras
//   backsub    dX = (Hll + lambda I)^-1 (bl - Hpl^T dxi), candidate points into
//              the window's other state buffer, x'(lambda x + b), and the
//              robust chi2 at the candidate per tile
//   accept     rho, lambda / nu schedule, accept or reject, next phase; an
//              accepted trial flips the window's point buffer (no copy)
//   commit     candidate cameras -> current for windows whose trial was accepted
//   cost       robust chi2 at the current state (a solve's start)
// Every reduction has a fixed order: results are deterministic.
#include <float.h>

#include "ba_kernels.h"

namespace navba {


// ---------------------------------------------------------------- cameras
__global__ void k_cam_rt(int n, const PoseQT* __restrict__ pose, CamRT* __restrict__ cam) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const PoseQT p = pose[c];
    CamRT o;
    quat_to_R(p.q, o.R);
    o.t[0] = p.t[0]; o.t[1] = p.t[1]; o.t[2] = p.t[2];
    cam[c] = o;
}

void launch_cam_rt(const DevProblem& P, const DevState& s, hipStream_t st) {
    hipLaunchKernelGGL(k_cam_rt, dim3((P.n_cam + 255) / 256), dim3(256), 0, st, P.n_cam, s.pose, s.cam);
}

__global__ void k_lm_reset(int n_win, int n_iter, double lambda0, LMDev* lm, int* done_count) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w == 0) *done_count = n_iter > 0 ? 0 : n_win;
    if (w >= n_win) return;
    LMDev L{};
    L.lambda = lambda0;
    L.nu = 2.0;
    L.phase = n_iter > 0 ? kPhaseLin : kPhaseDone;
    lm[w] = L;
}

void launch_lm_reset(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_reset, dim3((P.n_win + 255) / 256), dim3(256), 0, st, P.n_win, P.n_iter,
                       P.auto_lambda ? 0.0 : P.lambda0, W.lm, W.done_count);
}

// ---------------------------------------------------------------- lin_pt
// One workgroup per tile of whole points.  Per observation: residual and
// Jacobian in T, Huber weight (stored as T), and the point's Hll / bl terms
// formed and summed in FP64 (fixed order from LDS).  A point with more than
// kTile observations has a tile of its own and is summed by a block reduction.
// Registers capped at 72 for the pinhole (7 waves per SIMD instead of 6): the
// kernel waits on its gathers, and the extra wave gave 6 % at 3 M
// observations.  The fisheye model's transcendental functions need more.
template <typename T, int CM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CM == kCamPinhole ? 7 : 1, 8))) void k_lin_pt(DevProblem P, const CamRT* __restrict__ cam,
                                                   const LMDev* lm,
                                                   T* __restrict__ w_out, double* __restrict__ Hl) {
    // rows padded by one double: the 9 sums of one point read one column, which
    // an unpadded 2 KB row stride would put in a single LDS bank
    __shared__ double sh[9][kTile + 1];
    __shared__ double red[kBlock / kWave][9];
    const int t = blockIdx.x;
    const D4* __restrict__ X = x_cur(P, lm, P.tile_win[t]);  // its selector loaded beside the LM state
    if (!needs_lin(lm, P.tile_win[t])) return;
    const int o0 = P.tile_obs[t], o1 = P.tile_obs[t + 1];
    const int p0 = P.tile_pt[t], p1 = P.tile_pt[t + 1];
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = 0.0;
    for (int a = o0 + (int)threadIdx.x; a < o1; a += kBlock) {
        const int p = P.o_pt[a], ci = P.o_cam[a];
        double R[9], tt[3];
        load_cam(cam, ci, R, tt);
        const D4 xp = X[p];
        EdgeLin<T> L;
        const double s = edge_linearize_m<CM>(P, R, tt, xp.x, xp.y, xp.z, P.o_uv[2 * (size_t)a],
                                              P.o_uv[2 * (size_t)a + 1], L);
        const T wt = huber_weight<T>(s, P);
        w_out[a] = wt;
        if (P.pt_free[p]) {
            const double w = (double)wt;
            const double r0 = -L.e0 * w, r1 = -L.e1 * w;
            double A[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) A[k] = (double)L.JX[k];
            const double w0 = A[0] * w, w1 = A[1] * w, w2 = A[2] * w;
            const double w3 = A[3] * w, w4 = A[4] * w, w5 = A[5] * w;
            c[0] = fma(w3, A[3], fma(w0, A[0], c[0]));
            c[1] = fma(w3, A[4], fma(w0, A[1], c[1]));
            c[2] = fma(w3, A[5], fma(w0, A[2], c[2]));
            c[3] = fma(w4, A[4], fma(w1, A[1], c[3]));
            c[4] = fma(w4, A[5], fma(w1, A[2], c[4]));
            c[5] = fma(w5, A[5], fma(w2, A[2], c[5]));
            c[6] = fma(A[3], r1, fma(A[0], r0, c[6]));
            c[7] = fma(A[4], r1, fma(A[1], r0, c[7]));
            c[8] = fma(A[5], r1, fma(A[2], r0, c[8]));
        }
    }
    if (o1 - o0 <= kTile) {
#pragma unroll
        for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = c[k];
        __syncthreads();
        const int n = (p1 - p0) * 9;
        for (int idx = threadIdx.x; idx < n; idx += kBlock) {
            const int p = p0 + idx / 9, k = idx % 9;
            const int b0 = P.pt_ptr[p] - o0, b1 = P.pt_ptr[p + 1] - o0;
            // four interleaved partial sums: four LDS reads in flight per step
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int a = b0;
            for (; a + 4 <= b1; a += 4) {
                s0 += sh[k][a];
                s1 += sh[k][a + 1];
                s2 += sh[k][a + 2];
                s3 += sh[k][a + 3];
            }
            if (a < b1) s0 += sh[k][a];
            if (a + 1 < b1) s1 += sh[k][a + 1];
            if (a + 2 < b1) s2 += sh[k][a + 2];
            Hl[9 * (size_t)p + k] = (s0 + s1) + (s2 + s3);
        }
    } else {
        const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double v = wave_sum(c[k]);
            if (lane == 0) red[wid][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < 9) {
            double s = 0.0;
            for (int i = 0; i < kBlock / kWave; ++i) s += red[i][threadIdx.x];
            Hl[9 * (size_t)p0 + threadIdx.x] = s;
        }
    }
}

template <typename T>
void launch_lin_pt(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_tiles == 0) return;
    const auto fn = P.cam_model == kCamKB8 ? k_lin_pt<T, kCamKB8> : k_lin_pt<T, kCamPinhole>;
    hipLaunchKernelGGL(fn, dim3(P.n_tiles),
                       dim3(kBlock), 0, st, P, s.cam, (const LMDev*)W.lm, (T*)W.w, W.Hl);
}

// ---------------------------------------------------------------- lin_cam
// One workgroup per (free camera, chunk of its observations): Hpp (21 packed)
// and bp (6) from T Jacobians, formed in FP64 and reduced per wave by a
// reduce-scatter of the 27 sums, then over the waves in order; then one
// workgroup per camera sums the chunks.  Inputs are read in camera-major
// order (c_pt, c_uv), so the per-observation reads are contiguous.
template <typename T, int CM>
__global__ __launch_bounds__(kBlock) void k_lin_cam(DevProblem P, const CamRT* __restrict__ cam,
                                                    const LMDev* lm, double* __restrict__ hpart) {
    __shared__ double red[kBlock / kWave][27];
    const int it = blockIdx.x;
    const int f = P.hitem_f[it], a0 = P.hitem_a0[it], a1 = P.hitem_a1[it];
    const int cf = P.free_cam[f];
    const D4* __restrict__ X = x_cur(P, lm, P.cam_win[cf]);  // its selector loaded beside the LM state
    if (!needs_lin(lm, P.cam_win[cf])) return;
    double R[9], tt[3];
    load_cam(cam, cf, R, tt);
    double acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = 0.0;
    // camera-major copies of the point index and measurement: contiguous
    // reads, all issued before the first point gather; the Huber weight is
    // recomputed from the residual (the same arithmetic as k_lin_pt)
    constexpr int kU = kChunk / kBlock;
    int pk[kU];
    double uk[kU], vk[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        const int k = min(a0 + (int)threadIdx.x + kBlock * u, a1 - 1);  // a1 > a0: items are never empty
        pk[u] = P.c_pt[k];
        uk[u] = P.c_uv[2 * (size_t)k];
        vk[u] = P.c_uv[2 * (size_t)k + 1];
    }
    D4 xs[kU];  // the point gathers, all in flight before the first is used
#pragma unroll
    for (int u = 0; u < kU; ++u) xs[u] = X[pk[u]];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        if (a0 + (int)threadIdx.x + kBlock * u >= a1) break;
        const D4 xp = xs[u];
        EdgeLin<T> L;
        const double s = edge_linearize_m<CM>(P, R, tt, xp.x, xp.y, xp.z, uk[u], vk[u], L);
        const double wa = (double)huber_weight<T>(s, P);
        const double r0 = -L.e0 * wa, r1 = -L.e1 * wa;
        double J[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) J[i] = (double)L.Jxi[i];
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double A0 = J[i] * wa, A1 = J[6 + i] * wa;
#pragma unroll
            for (int j = i; j < 6; ++j, ++q) {  // a pinhole J_xi is zero at (0, 4) and (1, 3)
                constexpr bool zx = CamTraits<CM>::xi_zeros;
                const bool z0 = zx && (i == 4 || j == 4), z1 = zx && (i == 3 || j == 3);
                if (!z0 && !z1) acc[q] = fma(A1, J[6 + j], fma(A0, J[j], acc[q]));
                else if (!z0) acc[q] = fma(A0, J[j], acc[q]);
                else if (!z1) acc[q] = fma(A1, J[6 + j], acc[q]);
            }
        }
#pragma unroll
        for (int i = 0; i < 6; ++i)
            acc[21 + i] = (CamTraits<CM>::xi_zeros && i == 4) ? fma(J[6 + i], r1, acc[21 + i])
                          : (CamTraits<CM>::xi_zeros && i == 3) ? fma(J[i], r0, acc[21 + i])
                                                                 : fma(J[6 + i], r1, fma(J[i], r0, acc[21 + i]));
    }
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    int k;
    const double v = wave_reduce_scatter27(acc, lane, k);
    if ((lane & 1) == 0 && k >= 0) red[wid][k] = v;
    __syncthreads();
    if (threadIdx.x < 27) {
        double s = 0.0;
        for (int i = 0; i < kBlock / kWave; ++i) s += red[i][threadIdx.x];
        hpart[27 * (size_t)it + threadIdx.x] = s;
    }
}

__global__ void k_cam_reduce(DevProblem P, const double* __restrict__ hpart, const LMDev* lm, double* __restrict__ Hc) {
    const int f = blockIdx.x;
    if (!needs_lin(lm, P.cam_win[P.free_cam[f]])) return;
    if (threadIdx.x < 28) {
        // chunks in a fixed order, four interleaved partial sums (loads in flight)
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (threadIdx.x < 27) {
            const int i1 = P.cam_hitem_ptr[f + 1];
            int it = P.cam_hitem_ptr[f];
            const double* hp = hpart + threadIdx.x;
            for (; it + 4 <= i1; it += 4) {
                s0 += hp[27 * (size_t)it];
                s1 += hp[27 * (size_t)(it + 1)];
                s2 += hp[27 * (size_t)(it + 2)];
                s3 += hp[27 * (size_t)(it + 3)];
            }
            for (; it < i1; ++it) s0 += hp[27 * (size_t)it];
        }
        Hc[28 * (size_t)f + threadIdx.x] = (s0 + s1) + (s2 + s3);
    }
}

template <typename T>
void launch_lin_cam(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_fc == 0) return;
    if (P.n_hitems > 0) {
        const auto fn = P.cam_model == kCamKB8 ? k_lin_cam<T, kCamKB8> : k_lin_cam<T, kCamPinhole>;
        hipLaunchKernelGGL(fn, dim3(P.n_hitems), dim3(kBlock), 0, st, P, s.cam, (const LMDev*)W.lm, W.hpart);
    }
    hipLaunchKernelGGL(k_cam_reduce, dim3(P.n_fc), dim3(64), 0, st, P, (const double*)W.hpart, (const LMDev*)W.lm, W.Hc);
}

// ---------------------------------------------------------------- lambda init
// g2o computeLambdaInit with no user lambda: tau * max |H_jj| over the free
// vertices of the window, tau = 1e-5.  Launched in the first step only, where
// every window is at iteration 0 and its lambda slot holds 0 (k_lm_reset).
// k_diag_max: one thread per free vertex (free cameras, then points); the
// maximum goes into the window's lambda slot by an integer atomic max on the
// bit pattern, which orders non-negative doubles as their values (NaN
// diagonals are skipped, as fmax skips them).  One atomic per wave when the
// wave's vertices share a window.
__global__ __launch_bounds__(256) void k_diag_max(DevProblem P, const double* __restrict__ Hc,
                                                  const double* __restrict__ Hl, LMDev* lm) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double m = 0.0;
    int win = -1;
    if (i < P.n_fc) {
        win = P.cam_win[P.free_cam[i]];
        const double* h = Hc + 28 * (size_t)i;
#pragma unroll
        for (int d = 0; d < 6; ++d) m = fmax(m, fabs(h[sym6(d, d)]));
    } else if (i < P.n_fc + P.n_pt) {
        const int q = i - P.n_fc;
        if (P.pt_free[q]) {
            win = P.pt_win[q];
            const double* h = Hl + 9 * (size_t)q;  // packed (00 01 02 11 12 22)
            m = fmax(fmax(fabs(h[0]), fabs(h[3])), fabs(h[5]));
        }
    }
    if (win >= 0 && (!needs_lin(lm, win) || lm[win].iter != 0)) win = -1;
    const int w0 = __builtin_amdgcn_readfirstlane(win);
    if (__all(win == w0 || win < 0)) {
        int wl = win;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            m = fmax(m, __shfl_xor(m, off, 64));
            wl = max(wl, __shfl_xor(wl, off, 64));
        }
        if ((threadIdx.x & 63) == 0 && wl >= 0)
            atomicMax((unsigned long long*)&lm[wl].lambda, (unsigned long long)__double_as_longlong(m));
    } else if (win >= 0) {
        atomicMax((unsigned long long*)&lm[win].lambda, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void k_lambda_scale(int n_win, LMDev* lm) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_win || !needs_lin(lm, w) || lm[w].iter != 0) return;
    lm[w].lambda *= 1e-5;
    lm[w].nu = 2.0;
}

template <typename T>
void launch_lambda_init(const DevProblem& P, const DevWork& W, hipStream_t st) {
    if (!P.auto_lambda) return;
    const int n = P.n_fc + P.n_pt;
    if (n > 0)
        hipLaunchKernelGGL(k_diag_max, dim3((n + 255) / 256), dim3(256), 0, st, P, W.Hc, (const double*)W.Hl, W.lm);
    hipLaunchKernelGGL(k_lambda_scale, dim3((P.n_win + 255) / 256), dim3(256), 0, st, P.n_win, W.lm);
}

// ---------------------------------------------------------------- schur reduce
// Clear the dense systems (padding rows of the tiled path get a unit
// diagonal) and their right-hand sides; run once per prepare.  The reduction
// rewrites every lower-triangle block of a window each trial (zeros where no
// point links two cameras), and the padding is never changed.
__global__ void k_fill_system(DevProblem P, double* __restrict__ S, double* __restrict__ g) {
    const int win = blockIdx.y;
    const int lda = P.win_lda[win];
    const int n = 6 * (P.win_fc0[win + 1] - P.win_fc0[win]);
    double* Sw = S + P.win_S[win];
    const size_t tot = (size_t)lda * lda;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / lda, c = i % lda;
        Sw[i] = (r == c && (int)r >= n) ? 1.0 : 0.0;
    }
    if (blockIdx.x == 0) {
        const int g0 = 6 * P.win_fc0[win];
        const int gn = P.batched ? n : lda;
        for (int i = threadIdx.x; i < gn; i += blockDim.x) g[g0 + i] = 0.0;
    }
}

// Workgroups (f, part) for S row block f, each taking a quarter of the row's
// entries: sum the chunk partials (FP64, fixed order), form
// S_ff = Hpp + lambda I - sum, S_gf = -sum^T (lower triangle), and write
// zeros to the row's blocks without co-visibility (the factorisation of the
// previous trial left values there); part 0 also forms g = bp - coef.
constexpr int kReduceParts = 8;
constexpr int kReduceStage = 2048;  // item offsets of one S row staged in LDS

__global__ __launch_bounds__(kBlock) void k_schur_reduce(DevProblem P, const double* __restrict__ spart,
                                                         const double* __restrict__ Hc, const LMDev* lm,
                                                         double* __restrict__ S, double* __restrict__ gv) {
    __shared__ double cpart[4][6];
    __shared__ long long pofs_s[kReduceStage];
    const int f = blockIdx.x;
    const int win = P.cam_win[P.free_cam[f]];
    if (!running(lm, win)) return;
    if (!P.batched && !P.use_blocks) {
        // the padding rows n..n_pad-1 of the tiled system (identity): the
        // factorisation reads them, and a failed trial can leave non-finite
        // values there (a zero row times an overflowed factor entry)
        const int pr = P.n_pad - P.n;
        const long long tot = (long long)pr * P.n_pad;
        for (long long q = ((long long)blockIdx.x * gridDim.y + blockIdx.y) * kBlock + threadIdx.x; q < tot;
             q += (long long)gridDim.x * gridDim.y * kBlock) {
            const int r = P.n + (int)(q / P.n_pad), c = (int)(q % P.n_pad);
            if (c <= r) S[(size_t)r * P.n_pad + c] = (r == c) ? 1.0 : 0.0;
        }
    }
    const int fc0 = P.win_fc0[win];
    const int lf = f - fc0;
    double* Sw = S + P.win_S[win];
    const int lda = P.win_lda[win];
    const int cbase = P.colptr[f];
    const int ncols = P.colptr[f + 1] - cbase;
    const int nch = P.row_nchunks[f];
    const int base = P.row_sitem_ptr[f];
    const double lambda = lm[win].lambda;
    const int nent = ncols * 36;
    const int per = (nent + kReduceParts - 1) / kReduceParts;
    const int e0 = blockIdx.y * per, e1 = min(nent, e0 + per);
    // the offsets of the row's item partials, staged in LDS once (the sums
    // below then have no dependent global load per chunk)
    const int nitems = P.row_sitem_ptr[f + 1] - base;
    const bool staged = nitems <= kReduceStage;
    if (staged)
        for (int q = threadIdx.x; q < nitems; q += kBlock) pofs_s[q] = P.sitem_pofs[base + q];
    __syncthreads();
    for (int idx = e0 + threadIdx.x; idx < e1; idx += kBlock) {
        const int s = idx / 36, e = idx % 36;
        const int gr = s / kMaxSlots, sl = s % kMaxSlots;
        // chunks summed in a fixed order by eight interleaved partial sums (loads in flight)
        const int q0 = gr * nch;
        const double* sp = spart + 36 * sl + e;
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = 0.0;
        int ch = 0;
        if (staged) {
            for (; ch + 8 <= nch; ch += 8) {
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] += sp[pofs_s[q0 + ch + u]];
            }
            for (; ch < nch; ++ch) a[0] += sp[pofs_s[q0 + ch]];
        } else {
            for (; ch + 8 <= nch; ch += 8) {
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] += sp[P.sitem_pofs[base + q0 + ch + u]];
            }
            for (; ch < nch; ++ch) a[0] += sp[P.sitem_pofs[base + q0 + ch]];
        }
        const double sum = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        const int i = e / 6, c = e % 6;
        if (s == 0) {
            Sw[(size_t)(6 * lf + i) * lda + 6 * lf + c] = Hc[28 * (size_t)f + sym6(i, c)] + (i == c ? lambda : 0.0) - sum;
        } else {
            const int lg = P.colcam[cbase + s] - fc0;
            Sw[(size_t)(6 * lg + c) * lda + 6 * lf + i] = -sum;
            if (P.use_pcg) Sw[(size_t)(6 * lf + i) * lda + 6 * lg + c] = -sum;  // both triangles for the matvec
        }
    }
    {
        const int z0 = P.zcolptr[f], nz = (P.zcolptr[f + 1] - z0) * 36;
        const int zper = (nz + kReduceParts - 1) / kReduceParts;
        const int q0 = blockIdx.y * zper, q1 = min(nz, q0 + zper);
        for (int idx = q0 + threadIdx.x; idx < q1; idx += kBlock) {
            const int lg = P.zcolcam[z0 + idx / 36] - fc0, e = idx % 36;
            Sw[(size_t)(6 * lg + e % 6) * lda + 6 * lf + e / 6] = 0.0;
            if (P.use_pcg) Sw[(size_t)(6 * lf + e / 6) * lda + 6 * lg + e % 6] = 0.0;
        }
    }
    if (blockIdx.y == 0) {
        // coefficient of group 0's chunks: 24 threads, entry t % 6, chunks = t / 6 (mod 4)
        if (threadIdx.x < 24) {
            const int e = threadIdx.x % 6, q = threadIdx.x / 6;
            double c = 0.0;
            for (int ch = q; ch < nch; ch += 4) {
                const int it = base + ch;
                c += spart[P.sitem_pofs[it] + 36 * P.sitem_nslots[it] + e];
            }
            cpart[q][e] = c;
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const double coef = (cpart[0][threadIdx.x] + cpart[1][threadIdx.x]) +
                                (cpart[2][threadIdx.x] + cpart[3][threadIdx.x]);
            gv[6 * f + threadIdx.x] = Hc[28 * (size_t)f + 21 + threadIdx.x] - coef;
        }
    }
}

// (Batches have no dense system in memory: the per-window factorisation sums
// the partials into LDS itself.)
void launch_fill_system(const DevProblem& P, const DevWork& W, hipStream_t st) {
    if (P.batched) return;
    const size_t tiles = ((size_t)P.n_pad * P.n_pad + 255) / 256;
    const int chunks = P.batched ? 4 : (int)(tiles < 65535 ? (tiles > 0 ? tiles : 1) : 65535);
    hipLaunchKernelGGL(k_fill_system, dim3(chunks, P.n_win), dim3(256), 0, st, P, W.S, W.g);
}

template <typename T>
void launch_schur_reduce(const DevProblem& P, const DevWork& W, hipStream_t st) {
    if (P.n_fc == 0 || P.batched) return;  // batches: summed by the per-window factorisation
    hipLaunchKernelGGL(k_schur_reduce, dim3(P.n_fc, kReduceParts), dim3(kBlock), 0, st, P, (const double*)W.spart,
                       W.Hc, (const LMDev*)W.lm, W.S, W.g);
}

// ---------------------------------------------------------------- back-substitution
// One workgroup per linearisation tile (whole points of one window).  Per
// observation b on a free camera: c_b = J_X,b^T w_b J_xi,b dxi_b (one lane per
// observation); per point, in fixed order from LDS: cl = bl - sum_b c_b,
// dX = Dinv cl, X' = X + dX, and the point part of x'(lambda x + b), summed
// per tile.  A point with more than kTile observations has a tile of its own
// and is summed by a block reduction.  The same workgroup then forms the
// trial cost of the tile's observations at the candidate state (the candidate
// cameras are formed before it), summed per tile in a fixed order, so that the
// trial needs no separate cost pass over the observations.
// c_b of one observation with its camera, weight and camera step d (6) loaded.
template <typename T, int CM>
__device__ __forceinline__ void backsub_obs_at(const DevProblem& P, const double R[9], const double tt[3], const D4& xp,
                                               double wb, const double* __restrict__ d, double c[3]) {
    EdgeLin<T> L;
    edge_linearize_m<CM>(P, R, tt, xp.x, xp.y, xp.z, 0.0, 0.0, L);
    double t0 = 0.0, t1 = 0.0;
    constexpr bool zx = CamTraits<CM>::xi_zeros;  // a pinhole J_xi is zero at (0, 4) and (1, 3)
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (!zx || i != 4) t0 += (double)L.Jxi[i] * d[i];
        if (!zx || i != 3) t1 += (double)L.Jxi[6 + i] * d[i];
    }
    t0 *= wb; t1 *= wb;
    c[0] = (double)L.JX[0] * t0 + (double)L.JX[3] * t1;
    c[1] = (double)L.JX[1] * t0 + (double)L.JX[4] * t1;
    c[2] = (double)L.JX[2] * t0 + (double)L.JX[5] * t1;
}

template <typename T, int CM>
__device__ __forceinline__ void backsub_obs(const DevProblem& P, const CamRT* __restrict__ cam, const D4& xp,
                                            int ob, const T* __restrict__ w, const double* __restrict__ dp,
                                            double c[3]) {
    c[0] = c[1] = c[2] = 0.0;
    const int cb = P.o_cam[ob];
    const int fb = P.cam_free[cb];
    if (fb < 0) return;
    double R[9], tt[3];
    load_cam(cam, cb, R, tt);
    backsub_obs_at<T, CM>(P, R, tt, xp, (double)w[ob], dp + 6 * (size_t)fb, c);
}

// Point p's update from bl (3) and Dinv (6, packed) already loaded.
__device__ __forceinline__ double backsub_point_at(int p, double lambda, const double bl[3], const double D[6],
                                                   const double s[3], const D4& xp, D4* __restrict__ Xn, D4& xn) {
    const double bl0 = bl[0], bl1 = bl[1], bl2 = bl[2];
    const double c0 = bl0 - s[0], c1 = bl1 - s[1], c2 = bl2 - s[2];
    const double d0 = D[0] * c0 + D[1] * c1 + D[2] * c2;
    const double d1 = D[1] * c0 + D[3] * c1 + D[4] * c2;
    const double d2 = D[2] * c0 + D[4] * c1 + D[5] * c2;
    xn = D4{xp.x + d0, xp.y + d1, xp.z + d2, xp.w};
    Xn[p] = xn;
    return d0 * (lambda * d0 + bl0) + d1 * (lambda * d1 + bl1) + d2 * (lambda * d2 + bl2);
}

__device__ __forceinline__ double backsub_point(const DevProblem& P, int p, double lambda, const double* __restrict__ Hl,
                                                const double s[3], const D4& xp, D4* __restrict__ Xn, D4& xn) {
    const double* h = Hl + 9 * (size_t)p;
    double D[6];
    point_dinv(h, lambda, D);
    return backsub_point_at(p, lambda, h + 6, D, s, xp, Xn, xn);
}

// Robust cost of observation a at the candidate state (its camera's candidate
// pose and its point's candidate position xn), as k_cost forms it.
template <int CM>
__device__ __forceinline__ double trial_term(const DevProblem& P, const CamRT* __restrict__ cn, int cb, const D4& xn,
                                             int a) {
    double R[9], tt[3];
    load_cam(cn, cb, R, tt);
    double e0, e1, zc, w;
    const double s = edge_residual_m<CM>(P, R, tt, xn.x, xn.y, xn.z, P.o_uv[2 * (size_t)a],
                                         P.o_uv[2 * (size_t)a + 1], e0, e1, zc);
    return robustify(s, P.robust, P.delta, w);
}

// Tile t of window win: X the window's current points, Xn the buffer of its
// candidate (restrict parameters: the compiler may keep point loads ahead of
// the candidate stores).
template <typename T, int CM>
__device__ __forceinline__ void backsub_tile(const DevProblem& P, const CamRT* __restrict__ cam,
                                             const D4* __restrict__ X, const T* __restrict__ w,
                                             const double* __restrict__ Hl, const double* __restrict__ dp,
                                             const LMDev* lm, D4* __restrict__ Xn, double* __restrict__ red_scale,
                                             const CamRT* __restrict__ cn, double* __restrict__ red_tcost, int t,
                                             int win) {
    __shared__ double sh[3][kTile];
    __shared__ double red[kBlock / kWave][3];
    __shared__ D4 xs_new[kTile];  // the tile's candidate points (a point has an observation: at most kTile)
    const int o0 = P.tile_obs[t], o1 = P.tile_obs[t + 1];
    const int p0 = P.tile_pt[t], p1 = P.tile_pt[t + 1];
    const bool fail = lm[win].fail != 0;
    const double lambda = lm[win].lambda;
    double sc = 0.0, cost = 0.0;
    if (o1 - o0 <= kTile) {
        // loads in three waves, each issued together: the observation's indices
        // and weight with this thread's first point (X, pt_free, bl, Dinv); the
        // gathers they index (point, camera, free index); the camera step
        const int a = o0 + (int)threadIdx.x;
        const bool va = a < o1;  // a tile of unobserved points has no observations
        int p = 0, cb = 0;
        double wb = 0.0;
        if (va) {
            p = P.o_pt[a];
            cb = P.o_cam[a];
            wb = (double)w[a];
        }
        const int q = p0 + (int)threadIdx.x;
        D4 xq = {0.0, 0.0, 0.0, 0.0};
        bool fq = false;
        double hq[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (q < p1) {
            xq = X[q];
            fq = P.pt_free[q] != 0;
#pragma unroll
            for (int k = 0; k < 9; ++k) hq[k] = Hl[9 * (size_t)q + k];
        }
        bool fp = false;
        D4 xp = {0.0, 0.0, 0.0, 0.0};
        int fb = -1;
        double R[9], tt[3];
        if (va) {
            fp = P.pt_free[p] != 0;
            xp = X[p];
            fb = P.cam_free[cb];
            load_cam(cam, cb, R, tt);
        }
        double c[3] = {0.0, 0.0, 0.0};
        if (va && !fail && fp && fb >= 0) backsub_obs_at<T, CM>(P, R, tt, xp, wb, dp + 6 * (size_t)fb, c);
        sh[0][threadIdx.x] = c[0]; sh[1][threadIdx.x] = c[1]; sh[2][threadIdx.x] = c[2];
        __syncthreads();
        for (int pp = q; pp < p1; pp += kBlock) {
            const bool first = pp == q;
            const D4 xpp = first ? xq : X[pp];
            if ((first ? fq : P.pt_free[pp] != 0) && !fail) {
                // the point's observations in a fixed order, two interleaved sums per component
                const int b1 = P.pt_ptr[pp + 1] - o0;
                int b = P.pt_ptr[pp] - o0;
                double s[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
                for (; b + 2 <= b1; b += 2) {
                    s[0] += sh[0][b]; s[1] += sh[1][b]; s[2] += sh[2][b];
                    t[0] += sh[0][b + 1]; t[1] += sh[1][b + 1]; t[2] += sh[2][b + 1];
                }
                for (; b < b1; ++b) { s[0] += sh[0][b]; s[1] += sh[1][b]; s[2] += sh[2][b]; }
                s[0] += t[0]; s[1] += t[1]; s[2] += t[2];
                D4 xn;
                if (first) {
                    double D[6];
                    point_dinv(hq, lambda, D);
                    sc += backsub_point_at(pp, lambda, hq + 6, D, s, xpp, Xn, xn);
                } else {
                    sc += backsub_point(P, pp, lambda, Hl, s, xpp, Xn, xn);
                }
                if (pp - p0 < kTile) xs_new[pp - p0] = xn;
            } else {
                Xn[pp] = xpp;
                if (pp - p0 < kTile) xs_new[pp - p0] = xpp;
            }
        }
        __syncthreads();
        if (va && !fail) cost = trial_term<CM>(P, cn, cb, p - p0 < kTile ? xs_new[p - p0] : Xn[p], a);
    } else {
        // one point (p0) with more than kTile observations
        const D4 xp = X[p0];
        const bool active = P.pt_free[p0] && !fail;
        double c[3] = {0.0, 0.0, 0.0};
        if (active)
            for (int a = o0 + (int)threadIdx.x; a < o1; a += kBlock) {
                double ca[3];
                backsub_obs<T, CM>(P, cam, xp, a, w, dp, ca);
                c[0] += ca[0]; c[1] += ca[1]; c[2] += ca[2];
            }
        const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double v = wave_sum(c[k]);
            if (lane == 0) red[wid][k] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (active) {
                double s[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    s[k] = 0.0;
                    for (int i = 0; i < kBlock / kWave; ++i) s[k] += red[i][k];
                }
                D4 xn;
                sc = backsub_point(P, p0, lambda, Hl, s, xp, Xn, xn);
                xs_new[0] = xn;
            } else {
                Xn[p0] = xp;
                xs_new[0] = xp;
            }
        }
        __syncthreads();
        if (!fail) {
            const D4 xn = xs_new[0];
            for (int a = o0 + (int)threadIdx.x; a < o1; a += kBlock) cost += trial_term<CM>(P, cn, P.o_cam[a], xn, a);
        }
    }
    // both tile sums in one pass: wave sums, then the waves in order (the
    // additions of block_sum, with one barrier for the two; shs is used
    // nowhere else)
    __shared__ double shs[2][kBlock / kWave];
    const double ws = wave_sum(sc), wc = wave_sum(cost);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) {
        shs[0][wid] = ws;
        shs[1][wid] = wc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, cs = 0.0;
#pragma unroll
        for (int i = 0; i < kBlock / kWave; ++i) {
            s += shs[0][i];
            cs += shs[1][i];
        }
        red_scale[P.n_win + t] = s;
        red_tcost[t] = cs;
    }
}

template <typename T, int CM>
__global__ __launch_bounds__(kBlock) void k_backsub(DevProblem P, const CamRT* __restrict__ cam,
                                                    const T* __restrict__ w, const double* __restrict__ Hl,
                                                    const double* __restrict__ dp, const LMDev* lm,
                                                    double* __restrict__ red_scale,
                                                    const CamRT* __restrict__ cn, double* __restrict__ red_tcost) {
    const int t = blockIdx.x;
    const int win = P.tile_win[t];
    if (!running(lm, win)) return;
    backsub_tile<T, CM>(P, cam, x_cur(P, lm, win), w, Hl, dp, lm, x_cand(P, lm, win), red_scale, cn, red_tcost, t, win);
}

template <typename T>
void launch_backsub(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    if (P.n_tiles == 0) return;
    const auto fn = P.cam_model == kCamKB8 ? k_backsub<T, kCamKB8> : k_backsub<T, kCamPinhole>;
    hipLaunchKernelGGL(fn, dim3(P.n_tiles),
                       dim3(kBlock), 0, st, P, cur.cam, (const T*)W.w, (const double*)W.Hl, W.dp,
                       (const LMDev*)W.lm, W.red_scale, (const CamRT*)cand.cam, W.red_tcost);
}

// ---------------------------------------------------------------- camera update
// One workgroup per window: T <- exp(dxi) T for its free cameras and the camera
// part of x'(lambda x + b).
__global__ __launch_bounds__(256) void k_update_cam(DevProblem P, const PoseQT* __restrict__ pose,
                                                    const double* __restrict__ Hc, const double* __restrict__ dp,
                                                    const LMDev* lm, PoseQT* __restrict__ pn, CamRT* __restrict__ cn,
                                                    double* __restrict__ red_scale) {
    __shared__ double sh[4];
    const int win = blockIdx.x;
    if (!running(lm, win)) return;
    const double lambda = lm[win].lambda;
    const int fail = lm[win].fail;
    double sc = 0.0;
    for (int c = P.win_cam0[win] + threadIdx.x; c < P.win_cam0[win + 1]; c += 256)
        sc += camera_candidate(camera_inputs(P, pose, Hc, c), lambda, fail, c,
                               [&](int f, int i) { return dp[6 * (size_t)f + i]; }, pn, cn);
    const double s = block_sum<double, 256>(sc, sh);
    if (threadIdx.x == 0) red_scale[win] = s;
}

void launch_update_cam(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_update_cam, dim3(P.n_win), dim3(256), 0, st, P, (const PoseQT*)cur.pose, (const double*)W.Hc,
                       (const double*)W.dp, (const LMDev*)W.lm, cand.pose, cand.cam, W.red_scale);
}

// ---------------------------------------------------------------- cost
template <typename T, int CM>
__global__ __launch_bounds__(256) void k_cost(DevProblem P, const CamRT* __restrict__ cam,
                                              const LMDev* lm, int trial, double* __restrict__ red_cost) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    const int win = P.cblk_win[b];
    const D4* __restrict__ X = x_cur(P, lm, win);  // the window's current points
    if (trial && (!running(lm, win) || lm[win].fail)) {
        if (threadIdx.x == 0) red_cost[b] = 0.0;
        return;
    }
    // kCostObs / 256 observations per thread; every index and measurement
    // load of the block is issued before the first gather (the gathers of one
    // observation depend on its indices)
    constexpr int kU = kCostObs / 256;
    const int a0 = P.cblk_o0[b], a1 = P.cblk_o0[b + 1];
    int ci[kU], pi[kU];
    double uu[kU], vv[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) {
        const int a = min(a0 + (int)threadIdx.x + 256 * q, a1 - 1);  // a1 > a0: blocks are never empty
        ci[q] = P.o_cam[a];
        pi[q] = P.o_pt[a];
        uu[q] = P.o_uv[2 * (size_t)a];
        vv[q] = P.o_uv[2 * (size_t)a + 1];
    }
    D4 xs[kU];  // the point gathers, all in flight before the first is used
#pragma unroll
    for (int q = 0; q < kU; ++q) xs[q] = X[pi[q]];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < kU; ++q) {
        double R[9], tt[3];
        load_cam(cam, ci[q], R, tt);
        const D4 xp = xs[q];
        double e0, e1, zc, w;
        const double s = edge_residual_m<CM>(P, R, tt, xp.x, xp.y, xp.z, uu[q], vv[q], e0, e1, zc);
        const double rho = robustify(s, P.robust, P.delta, w);
        if (a0 + (int)threadIdx.x + 256 * q < a1) acc += rho;
    }
    const double s = block_sum<double, 256>(acc, sh);
    if (threadIdx.x == 0) red_cost[b] = s;
}

template <typename T>
void launch_cost(const DevProblem& P, const DevState& s, const DevWork& W, int trial, hipStream_t st) {
    if (P.n_cblk == 0) return;
    const auto fn = P.cam_model == kCamKB8 ? k_cost<T, kCamKB8> : k_cost<T, kCamPinhole>;
    hipLaunchKernelGGL(fn, dim3(P.n_cblk), dim3(256),
                       0, st, P, s.cam, (const LMDev*)W.lm, trial, W.red_cost);
}

// ---------------------------------------------------------------- LM decision
// This thread's part of the sum of v[i0..i1) over 256 threads: four
// interleaved partial sums (four loads in flight), in a fixed order.
__device__ __forceinline__ double range_part(const double* v, int i0, int i1) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = i0 + threadIdx.x;
    for (; i + 768 < i1; i += 1024) {
        s0 += v[i];
        s1 += v[i + 256];
        s2 += v[i + 512];
        s3 += v[i + 768];
    }
    for (; i < i1; i += 256) s0 += v[i];
    return (s0 + s1) + (s2 + s3);
}

__device__ double sum_range(const double* v, int i0, int i1, double* sh) {
    return block_sum<double, 256>(range_part(v, i0, i1), sh);
}

// g2o OptimizationAlgorithmLevenberg trial decision, one workgroup per window:
//   rho = (chi - chi') / (x'(lambda x + b) + 1e-3)
//   good: lambda *= max(1/3, min(2/3, 1 - (2 rho - 1)^3)), nu = 2, chi = chi'
//   bad:  lambda *= nu, nu *= 2
// then the trial loop of solve(): another trial while rho < 0 and fewer than
// max_trials trials; otherwise the iteration ends, and the window stops when
// the trial limit was reached, rho == 0, or the iteration count is done.
__global__ __launch_bounds__(256) void k_lm_accept(DevProblem P, const double* __restrict__ red_tcost,
                                                   const double* __restrict__ red_scale, LMDev* lm,
                                                   double* __restrict__ hist, int* done_count,
                                                   PoseQT* __restrict__ pose, const PoseQT* __restrict__ pn,
                                                   CamRT* __restrict__ cam, const CamRT* __restrict__ cn) {
    __shared__ double sh[4];
    __shared__ int acc;
    const int win = blockIdx.x;
    if (!running(lm, win)) {
        if (threadIdx.x == 0) lm[win].accepted = 0;
        return;
    }
    // both ranges' loads in flight together (the trial cost per linearisation tile, from k_backsub)
    const double pc = range_part(red_tcost, P.win_tile0[win], P.win_tile0[win + 1]);
    const double ps = range_part(red_scale, P.n_win + P.win_tile0[win], P.n_win + P.win_tile0[win + 1]);
    const double cost = block_sum<double, 256>(pc, sh);
    __syncthreads();
    const double scale = red_scale[win] + block_sum<double, 256>(ps, sh);
    if (threadIdx.x == 0) {
        LMDev L = lm[win];
        const double temp = L.fail ? DBL_MAX : cost;
        const double sc = scale + 1e-3;
        const double rho = (L.chi - temp) / sc;
        if (rho > 0 && isfinite(temp)) {
            double alpha = 1. - pow((2 * rho - 1), 3.0);
            alpha = fmin(alpha, 2. / 3.);
            L.lambda *= fmax(1. / 3., alpha);
            L.nu = 2;
            L.chi = temp;
            L.accepted = 1;
            // the trial's points become current: the back-substitution wrote
            // them to the other state buffer
            L.xsel ^= 1;
        } else {
            L.lambda *= L.nu;
            L.nu *= 2;
            L.accepted = 0;
        }
        L.temp_chi = temp;
        L.scale = sc;
        L.rho = rho;
        L.trials += 1;
        L.q += 1;
        if (rho < 0 && L.q < P.max_trials) {
            L.phase = kPhaseTrial;
        } else {
            L.iter += 1;
            hist[(size_t)win * (P.n_iter + 1) + L.iter] = L.chi;
            if (L.q == P.max_trials || rho == 0) {
                L.phase = kPhaseDone;
                L.stop = 1;
            } else if (L.iter >= P.n_iter) {
                L.phase = kPhaseDone;
                L.stop = 0;
            } else {
                L.phase = kPhaseLin;
            }
            L.q = 0;
            if (L.phase == kPhaseDone) atomicAdd(done_count, 1);
        }
        L.fail = 0;
        lm[win] = L;
        acc = L.accepted;
    }
    if (!P.commit_in_accept) return;
    // small windows: the commit of an accepted trial's cameras (k_commit) by the same workgroup
    __syncthreads();
    if (!acc) return;
    for (int c = P.win_cam0[win] + threadIdx.x; c < P.win_cam0[win + 1]; c += 256) {
        pose[c] = pn[c];
        cam[c] = cn[c];
    }
}

void launch_lm_accept(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_lm_accept, dim3(P.n_win), dim3(256), 0, st, P, (const double*)W.red_tcost,
                       (const double*)W.red_scale, W.lm, W.hist, W.done_count, cur.pose,
                       (const PoseQT*)cand.pose, cur.cam, (const CamRT*)cand.cam);
}

__global__ __launch_bounds__(256) void k_lm_set_chi(DevProblem P, const double* __restrict__ red_cost, LMDev* lm,
                                                    double* __restrict__ hist) {
    __shared__ double sh[4];
    const int win = blockIdx.x;
    const double cost = sum_range(red_cost, P.win_cblk0[win], P.win_cblk0[win + 1], sh);
    if (threadIdx.x == 0) {
        lm[win].chi = cost;
        lm[win].temp_chi = cost;
        hist[(size_t)win * (P.n_iter + 1)] = cost;
    }
}

void launch_lm_set_chi(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_set_chi, dim3(P.n_win), dim3(256), 0, st, P, (const double*)W.red_cost, W.lm, W.hist);
}

__global__ void k_commit(DevProblem P, const LMDev* lm, PoseQT* __restrict__ pose, const PoseQT* __restrict__ pn,
                         CamRT* __restrict__ cam, const CamRT* __restrict__ cn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n_cam && lm[P.cam_win[i]].accepted) { pose[i] = pn[i]; cam[i] = cn[i]; }
}

void launch_commit(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    if (P.commit_in_accept || P.n_cam == 0) return;  // done by k_lm_accept
    hipLaunchKernelGGL(k_commit, dim3((P.n_cam + 255) / 256), dim3(256), 0, st, P, (const LMDev*)W.lm, cur.pose,
                       (const PoseQT*)cand.pose, cur.cam, (const CamRT*)cand.cam);
}

// ---------------------------------------------------------------- outlier flags
// FP64 at the given state: flag = chi2 > thr or depth <= 0; written in the
// caller's observation order.  Also the FP64 robust cost and the flag count.
template <int CM>
__global__ __launch_bounds__(256) void k_flags(DevProblem P, const CamRT* __restrict__ cam, const LMDev* lm,
                                               double thr, uint8_t* __restrict__ out, double* __restrict__ red_cost,
                                               double* __restrict__ red_cnt) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    const D4* __restrict__ X = x_cur(P, lm, P.cblk_win[b]);  // the window's final points
    double acc = 0.0, cnt = 0.0;
    for (int a = P.cblk_o0[b] + threadIdx.x; a < P.cblk_o0[b + 1]; a += 256) {
        double R[9], tt[3];
        load_cam(cam, P.o_cam[a], R, tt);
        const D4 xp = X[P.o_pt[a]];
        double e0, e1, zc, w;
        const double s = edge_residual_m<CM>(P, R, tt, xp.x, xp.y, xp.z, P.o_uv[2 * (size_t)a],
                                             P.o_uv[2 * (size_t)a + 1], e0, e1, zc);
        acc += robustify(s, P.robust, P.delta, w);
        const uint8_t f = (s > thr || !(zc > 0.0)) ? 1 : 0;
        out[P.o_orig[a]] = f;
        cnt += f;
    }
    const double s = block_sum<double, 256>(acc, sh);
    __syncthreads();
    const double c = block_sum<double, 256>(cnt, sh);
    if (threadIdx.x == 0) { red_cost[b] = s; red_cnt[b] = c; }
}

void launch_flags(const DevProblem& P, const DevState& s, const DevWork& W, double thr, hipStream_t st) {
    if (P.n_cblk == 0) return;
    const auto fn = P.cam_model == kCamKB8 ? k_flags<kCamKB8> : k_flags<kCamPinhole>;
    hipLaunchKernelGGL(fn, dim3(P.n_cblk), dim3(256), 0,
                       st, P, s.cam, (const LMDev*)W.lm, thr, W.outlier, W.red_cost, W.red_aux);
}

__global__ __launch_bounds__(256) void k_final(DevProblem P, const double* __restrict__ red_cost,
                                               const double* __restrict__ red_cnt, double* __restrict__ out) {
    __shared__ double sh[4];
    const int win = blockIdx.x;
    const double c = sum_range(red_cost, P.win_cblk0[win], P.win_cblk0[win + 1], sh);
    __syncthreads();
    const double n = sum_range(red_cnt, P.win_cblk0[win], P.win_cblk0[win + 1], sh);
    if (threadIdx.x == 0) { out[2 * win] = c; out[2 * win + 1] = n; }
}

void launch_final(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_final, dim3(P.n_win), dim3(256), 0, st, P, (const double*)W.red_cost,
                       (const double*)W.red_aux, W.final);
}

// explicit instantiations
#define NAVBA_INST(T)                                                                                         \
    template void launch_lin_pt<T>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);         \
    template void launch_lin_cam<T>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);        \
    template void launch_lambda_init<T>(const DevProblem&, const DevWork&, hipStream_t);                     \
    template void launch_schur_reduce<T>(const DevProblem&, const DevWork&, hipStream_t);                    \
    template void launch_backsub<T>(const DevProblem&, const DevState&, const DevState&, const DevWork&,     \
                                    hipStream_t);                                                            \
    template void launch_cost<T>(const DevProblem&, const DevState&, const DevWork&, int, hipStream_t);
NAVBA_INST(double)
NAVBA_INST(float)
#undef NAVBA_INST

}  // namespace navba
