// ba_kernels.hip — gfx950 kernels of the BA Levenberg-Marquardt solve.
//
// Data flow of one LM iteration (g2o OptimizationAlgorithmLevenberg::solve, the
// optimizer nav24 runs at core/backEnd/BE_GraphOptim.cpp:66-68):
//   lin_pt     one lane per observation, point-aligned tiles: residual, Huber
//              weight, Hll / bl per point (segment sums in LDS)           [HBM-bound]
//   lin_cam    camera-major chunks: Hpp / bp per free camera (block sums)  [HBM-bound]
//   per trial:
//   hinv       (Hll + lambda I)^-1 and Hll^-1 bl per point, FP64
//   schur_row  per (camera row, chunk): S_row = Hpp - sum Hpl Hll^-1 Hpl^T accumulated
//              in LDS (one lane per observation, recomputed Jacobians)
//   schur_red  partials -> dense reduced camera system S (FP64) and g
//   potrf/trsm/syrk/trsv   blocked dense Cholesky of S, FP64
//   backsub    dX = Hll^-1 (bl - Hpl^T dxi), candidate points, x'(lambda x + b)
//   update_cam T <- exp(dxi) T (FP64 SE3 exp-map), candidate cameras
//   cost       robust chi2 at the candidate (per-block partial sums)
//   accept     rho, lambda / nu schedule, accept or reject (single workgroup)
//   commit     copy candidate -> current when accepted
// Reductions use fixed trees (deterministic), except the off-diagonal S-row
// blocks which use LDS atomics.
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

__global__ void k_lm_reset(LMDev* lm, double lambda0) {
    LMDev L{};
    L.lambda = lambda0;
    L.nu = 2.0;
    *lm = L;
}

void launch_lm_reset(LMDev* lm, double lambda0, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_reset, dim3(1), dim3(1), 0, st, lm, lambda0);
}

// ---------------------------------------------------------------- lin_pt
// One workgroup per tile of whole points.  Per observation: residual, Huber
// weight (stored), and the point's Hll / bl terms; per point the terms are
// summed in fixed order from LDS.  A point with more than kTile observations
// has a tile of its own and is summed by a block reduction.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lin_pt(DevProblem P, const CamRT* __restrict__ cam,
                                                   const D4* __restrict__ X, T* __restrict__ w_out,
                                                   T* __restrict__ Hl) {
    __shared__ T sh[9][kTile];
    __shared__ T red[kBlock / kWave][9];
    const int t = blockIdx.x;
    const int o0 = P.tile_obs[t], o1 = P.tile_obs[t + 1];
    const int p0 = P.tile_pt[t], p1 = P.tile_pt[t + 1];
    const T fx = (T)P.fx, fy = (T)P.fy, cx = (T)P.cx, cy = (T)P.cy, delta = (T)P.delta;
    T c[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = T(0);
    for (int a = o0 + (int)threadIdx.x; a < o1; a += kBlock) {
        const int p = P.o_pt[a], ci = P.o_cam[a];
        T R[9], tt[3];
        load_cam(cam, ci, R, tt);
        const D4 xp = X[p];
        EdgeLin<T> L;
        const T s = edge_linearize(R, tt, (T)xp.x, (T)xp.y, (T)xp.z, fx, fy, cx, cy,
                                   (T)P.o_uv[2 * (size_t)a], (T)P.o_uv[2 * (size_t)a + 1], L);
        T w;
        robustify(s, P.robust, delta, w);
        w_out[a] = w;
        if (P.pt_free[p]) {
            const T r0 = -L.e0 * w, r1 = -L.e1 * w;
            const T* A = L.JX;
            c[0] += (A[0] * w) * A[0] + (A[3] * w) * A[3];
            c[1] += (A[0] * w) * A[1] + (A[3] * w) * A[4];
            c[2] += (A[0] * w) * A[2] + (A[3] * w) * A[5];
            c[3] += (A[1] * w) * A[1] + (A[4] * w) * A[4];
            c[4] += (A[1] * w) * A[2] + (A[4] * w) * A[5];
            c[5] += (A[2] * w) * A[2] + (A[5] * w) * A[5];
            c[6] += A[0] * r0 + A[3] * r1;
            c[7] += A[1] * r0 + A[4] * r1;
            c[8] += A[2] * r0 + A[5] * r1;
        }
    }
    if (o1 - o0 <= kTile) {
#pragma unroll
        for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = c[k];
        __syncthreads();
        const int n = (p1 - p0) * 9;
        for (int idx = threadIdx.x; idx < n; idx += kBlock) {
            const int p = p0 + idx / 9, k = idx % 9;
            T s = T(0);
            for (int a = P.pt_ptr[p]; a < P.pt_ptr[p + 1]; ++a) s += sh[k][a - o0];
            Hl[9 * (size_t)p + k] = s;
        }
    } else {
        const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const T v = wave_sum(c[k]);
            if (lane == 0) red[wid][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < 9) {
            T s = T(0);
            for (int i = 0; i < kBlock / kWave; ++i) s += red[i][threadIdx.x];
            Hl[9 * (size_t)p0 + threadIdx.x] = s;
        }
    }
}

template <typename T>
void launch_lin_pt(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_tiles == 0) return;
    hipLaunchKernelGGL(k_lin_pt<T>, dim3(P.n_tiles), dim3(kBlock), 0, st, P, s.cam, s.X,
                       (T*)W.w, (T*)W.Hl);
}

// ---------------------------------------------------------------- lin_cam
// One workgroup per (free camera, chunk of its observations): Hpp (21 packed)
// and bp (6), block-reduced; then one workgroup per camera sums the chunks.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lin_cam(DevProblem P, const CamRT* __restrict__ cam,
                                                    const D4* __restrict__ X, const T* __restrict__ w,
                                                    T* __restrict__ hpart) {
    __shared__ T red[kBlock / kWave][27];
    const int it = blockIdx.x;
    const int f = P.hitem_f[it], a0 = P.hitem_a0[it], a1 = P.hitem_a1[it];
    T R[9], tt[3];
    load_cam(cam, P.free_cam[f], R, tt);
    const T fx = (T)P.fx, fy = (T)P.fy, cx = (T)P.cx, cy = (T)P.cy;
    T acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = T(0);
    for (int k = a0 + (int)threadIdx.x; k < a1; k += kBlock) {
        const int a = P.cam_obs[k];
        const D4 xp = X[P.o_pt[a]];
        EdgeLin<T> L;
        edge_linearize(R, tt, (T)xp.x, (T)xp.y, (T)xp.z, fx, fy, cx, cy,
                       (T)P.o_uv[2 * (size_t)a], (T)P.o_uv[2 * (size_t)a + 1], L);
        const T wa = w[a];
        const T r0 = -L.e0 * wa, r1 = -L.e1 * wa;
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const T A0 = L.Jxi[i] * wa, A1 = L.Jxi[6 + i] * wa;
#pragma unroll
            for (int j = i; j < 6; ++j) acc[q++] += A0 * L.Jxi[j] + A1 * L.Jxi[6 + j];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) acc[21 + i] += L.Jxi[i] * r0 + L.Jxi[6 + i] * r1;
    }
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const T v = wave_sum(acc[k]);
        if (lane == 0) red[wid][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 27) {
        T s = T(0);
        for (int i = 0; i < kBlock / kWave; ++i) s += red[i][threadIdx.x];
        hpart[27 * (size_t)it + threadIdx.x] = s;
    }
}

template <typename T>
__global__ void k_cam_reduce(DevProblem P, const T* __restrict__ hpart, double* __restrict__ Hc) {
    const int f = blockIdx.x;
    if (threadIdx.x < 28) {
        double s = 0.0;
        if (threadIdx.x < 27)
            for (int it = P.cam_hitem_ptr[f]; it < P.cam_hitem_ptr[f + 1]; ++it)
                s += (double)hpart[27 * (size_t)it + threadIdx.x];
        Hc[28 * (size_t)f + threadIdx.x] = s;
    }
}

template <typename T>
void launch_lin_cam(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_fc == 0) return;
    if (P.n_hitems > 0)
        hipLaunchKernelGGL(k_lin_cam<T>, dim3(P.n_hitems), dim3(kBlock), 0, st, P, s.cam, s.X,
                           (const T*)W.w, (T*)W.hpart);
    hipLaunchKernelGGL(k_cam_reduce<T>, dim3(P.n_fc), dim3(64), 0, st, P, (const T*)W.hpart, W.Hc);
}

// ---------------------------------------------------------------- lambda init
// g2o computeLambdaInit with no user lambda: tau * max |H_jj| over the free
// vertices, tau = 1e-5.
template <typename T>
__global__ __launch_bounds__(256) void k_lambda_init(DevProblem P, const double* __restrict__ Hc,
                                                     const T* __restrict__ Hl, LMDev* lm) {
    __shared__ double sh[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < P.n_fc * 6; i += 256) {
        const int d = i % 6;
        m = fmax(m, fabs(Hc[28 * (size_t)(i / 6) + sym6(d, d)]));
    }
    for (int i = threadIdx.x; i < P.n_pt * 3; i += 256) {
        const int d = i % 3;  // packed (00 01 02 11 12 22): diagonal at 0, 3, 5
        if (P.pt_free[i / 3]) m = fmax(m, fabs((double)Hl[9 * (size_t)(i / 3) + d * (7 - d) / 2]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        lm->lambda = 1e-5 * m;
        lm->nu = 2.0;
    }
}

template <typename T>
void launch_lambda_init(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_lambda_init<T>, dim3(1), dim3(256), 0, st, P, W.Hc, (const T*)W.Hl, W.lm);
}

// ---------------------------------------------------------------- hinv
// Per free point: Dinv = (Hll + lambda I)^-1 (FP64 adjugate) and Dinv * bl.
template <typename T>
__global__ void k_hinv(DevProblem P, const T* __restrict__ Hl, LMDev* lm, double* __restrict__ ptsys) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) lm->fail = 0;
    if (p >= P.n_pt || !P.pt_free[p]) return;
    const double lambda = lm->lambda;
    const T* h = Hl + 9 * (size_t)p;
    double a[6] = {(double)h[0] + lambda, (double)h[1], (double)h[2],
                   (double)h[3] + lambda, (double)h[4], (double)h[5] + lambda};
    double d[6];
    inv3_sym(a, d);
    const double b0 = h[6], b1 = h[7], b2 = h[8];
    double* o = ptsys + 12 * (size_t)p;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = d[k];
    o[6] = d[0] * b0 + d[1] * b1 + d[2] * b2;
    o[7] = d[1] * b0 + d[3] * b1 + d[4] * b2;
    o[8] = d[2] * b0 + d[4] * b1 + d[5] * b2;
}

template <typename T>
void launch_hinv(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_hinv<T>, dim3((P.n_pt + 255) / 256), dim3(256), 0, st, P, (const T*)W.Hl, W.lm, W.ptsys);
}

// ---------------------------------------------------------------- schur reduce
__global__ void k_fill_system(int n, int n_pad, double* __restrict__ S, double* __restrict__ g) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)n_pad * n_pad;
    if (i < tot) {
        const size_t r = i / n_pad, c = i % n_pad;
        S[i] = (r == c && (int)r >= n) ? 1.0 : 0.0;
    }
    if (i < (size_t)n_pad) g[i] = 0.0;
}

// One workgroup per S row block f: sum the chunk partials (FP64, fixed order),
// form S_ff = Hpp + lambda I - sum, S_gf = -sum^T (lower triangle), g = bp - coef.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_schur_reduce(DevProblem P, const double* __restrict__ spart,
                                                         const double* __restrict__ Hc, const LMDev* lm,
                                                         double* __restrict__ S, double* __restrict__ gv) {
    const int f = blockIdx.x;
    const int cbase = P.colptr[f];
    const int ncols = P.colptr[f + 1] - cbase;
    const int nch = P.row_nchunks[f];
    const int base = P.row_sitem_ptr[f];
    const double lambda = lm->lambda;
    const int lda = P.n_pad;
    for (int idx = threadIdx.x; idx < ncols * 36; idx += kBlock) {
        const int s = idx / 36, e = idx % 36;
        const int gr = s / kMaxSlots, sl = s % kMaxSlots;
        double sum = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int it = base + gr * nch + ch;
            sum += (double)spart[P.sitem_pofs[it] + 36 * sl + e];
        }
        const int i = e / 6, c = e % 6;
        if (s == 0) {
            S[(size_t)(6 * f + i) * lda + 6 * f + c] = Hc[28 * (size_t)f + sym6(i, c)] + (i == c ? lambda : 0.0) - sum;
        } else {
            const int gc = P.colcam[cbase + s];
            S[(size_t)(6 * gc + c) * lda + 6 * f + i] = -sum;
        }
    }
    if (threadIdx.x < 6) {
        double coef = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const int it = base + ch;
            coef += (double)spart[P.sitem_pofs[it] + 36 * P.sitem_nslots[it] + threadIdx.x];
        }
        gv[6 * f + threadIdx.x] = Hc[28 * (size_t)f + 21 + threadIdx.x] - coef;
    }
}

template <typename T>
void launch_schur_reduce(const DevProblem& P, const DevWork& W, hipStream_t st) {
    const size_t tot = (size_t)P.n_pad * P.n_pad;
    hipLaunchKernelGGL(k_fill_system, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, P.n, P.n_pad, W.S, W.g);
    if (P.n_fc == 0) return;
    hipLaunchKernelGGL(k_schur_reduce<T>, dim3(P.n_fc), dim3(kBlock), 0, st, P, (const double*)W.spart, W.Hc, W.lm, W.S, W.g);
}

// ---------------------------------------------------------------- back-substitution
// Per point: cl = bl - sum_b Hpl_b^T dxi_b, dX = Dinv cl, X' = X + dX and the
// point part of x'(lambda x + b).
template <typename T>
__global__ __launch_bounds__(256) void k_backsub(DevProblem P, const CamRT* __restrict__ cam,
                                                 const D4* __restrict__ X, const T* __restrict__ w,
                                                 const T* __restrict__ Hl, const double* __restrict__ ptsys,
                                                 const double* __restrict__ dp, const LMDev* lm,
                                                 D4* __restrict__ Xn, double* __restrict__ red_scale) {
    __shared__ double sh[4];
    const int p = blockIdx.x * 256 + threadIdx.x;
    double sc = 0.0;
    if (p < P.n_pt) {
        const D4 xp = X[p];
        D4 o = xp;
        if (P.pt_free[p] && !lm->fail) {
            const double lambda = lm->lambda;
            const T* h = Hl + 9 * (size_t)p;
            const double bl0 = (double)h[6], bl1 = (double)h[7], bl2 = (double)h[8];
            double c0 = bl0, c1 = bl1, c2 = bl2;
            const T x = (T)xp.x, y = (T)xp.y, z = (T)xp.z;
            const T fx = (T)P.fx, fy = (T)P.fy, cx = (T)P.cx, cy = (T)P.cy;
            for (int b = P.pt_ptr[p]; b < P.pt_ptr[p + 1]; ++b) {
                const int cb = P.o_cam[b];
                const int fb = P.cam_free[cb];
                if (fb < 0) continue;
                T R[9], tt[3];
                load_cam(cam, cb, R, tt);
                EdgeLin<T> L;
                edge_linearize(R, tt, x, y, z, fx, fy, cx, cy, T(0), T(0), L);
                const double wb = (double)w[b];
                const double* d = dp + 6 * (size_t)fb;
                double t0 = 0.0, t1 = 0.0;
#pragma unroll
                for (int i = 0; i < 6; ++i) { t0 += (double)L.Jxi[i] * d[i]; t1 += (double)L.Jxi[6 + i] * d[i]; }
                t0 *= wb; t1 *= wb;
                c0 -= (double)L.JX[0] * t0 + (double)L.JX[3] * t1;
                c1 -= (double)L.JX[1] * t0 + (double)L.JX[4] * t1;
                c2 -= (double)L.JX[2] * t0 + (double)L.JX[5] * t1;
            }
            const double* D = ptsys + 12 * (size_t)p;
            const double d0 = D[0] * c0 + D[1] * c1 + D[2] * c2;
            const double d1 = D[1] * c0 + D[3] * c1 + D[4] * c2;
            const double d2 = D[2] * c0 + D[4] * c1 + D[5] * c2;
            o.x += d0; o.y += d1; o.z += d2;
            sc = d0 * (lambda * d0 + bl0) + d1 * (lambda * d1 + bl1) + d2 * (lambda * d2 + bl2);
        }
        Xn[p] = o;
    }
    const double s = block_sum<double, 256>(sc, sh);
    if (threadIdx.x == 0) red_scale[1 + blockIdx.x] = s;
}

template <typename T>
void launch_backsub(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_backsub<T>, dim3(W.n_pt_blocks), dim3(256), 0, st, P, cur.cam, cur.X, (const T*)W.w,
                       (const T*)W.Hl, W.ptsys, W.dp, (const LMDev*)W.lm, cand.X, W.red_scale);
}

// ---------------------------------------------------------------- camera update
__global__ __launch_bounds__(256) void k_update_cam(DevProblem P, const PoseQT* __restrict__ pose,
                                                    const double* __restrict__ Hc, const double* __restrict__ dp,
                                                    const LMDev* lm, PoseQT* __restrict__ pn, CamRT* __restrict__ cn,
                                                    double* __restrict__ red_scale) {
    __shared__ double sh[4];
    const double lambda = lm->lambda;
    const int fail = lm->fail;
    double sc = 0.0;
    for (int c = threadIdx.x; c < P.n_cam; c += 256) {
        const int f = P.cam_free[c];
        const PoseQT in = pose[c];
        PoseQT o = in;
        if (f >= 0 && !fail) {
            double xi[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) xi[i] = dp[6 * f + i];
            se3_oplus(xi, in, o);
            const double* bp = Hc + 28 * (size_t)f + 21;
#pragma unroll
            for (int i = 0; i < 6; ++i) sc += xi[i] * (lambda * xi[i] + bp[i]);
        }
        pn[c] = o;
        CamRT r;
        quat_to_R(o.q, r.R);
        r.t[0] = o.t[0]; r.t[1] = o.t[1]; r.t[2] = o.t[2];
        cn[c] = r;
    }
    const double s = block_sum<double, 256>(sc, sh);
    if (threadIdx.x == 0) red_scale[0] = s;
}

void launch_update_cam(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_update_cam, dim3(1), dim3(256), 0, st, P, (const PoseQT*)cur.pose, (const double*)W.Hc,
                       (const double*)W.dp, (const LMDev*)W.lm, cand.pose, cand.cam, W.red_scale);
}

// ---------------------------------------------------------------- cost
constexpr int kCostPerThread = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_cost(DevProblem P, const CamRT* __restrict__ cam, const D4* __restrict__ X,
                                              const LMDev* lm, int skip_on_fail, double* __restrict__ red_cost) {
    __shared__ double sh[4];
    double acc = 0.0;
    if (!(skip_on_fail && lm->fail)) {
        const T fx = (T)P.fx, fy = (T)P.fy, cx = (T)P.cx, cy = (T)P.cy, delta = (T)P.delta;
        const int base = blockIdx.x * 256 * kCostPerThread + threadIdx.x;
#pragma unroll
        for (int i = 0; i < kCostPerThread; ++i) {
            const int a = base + 256 * i;
            if (a < P.n_obs) {
                T R[9], tt[3];
                load_cam(cam, P.o_cam[a], R, tt);
                const D4 xp = X[P.o_pt[a]];
                T e0, e1, zc, w;
                const T s = edge_residual(R, tt, (T)xp.x, (T)xp.y, (T)xp.z, fx, fy, cx, cy,
                                          (T)P.o_uv[2 * (size_t)a], (T)P.o_uv[2 * (size_t)a + 1], e0, e1, zc);
                acc += (double)robustify(s, P.robust, delta, w);
            }
        }
    }
    const double s = block_sum<double, 256>(acc, sh);
    if (threadIdx.x == 0) red_cost[blockIdx.x] = s;
}

template <typename T>
void launch_cost(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_cost<T>, dim3(W.n_cost_blocks), dim3(256), 0, st, P, s.cam, s.X, (const LMDev*)W.lm, 1, W.red_cost);
}

// ---------------------------------------------------------------- LM decision
__device__ double sum_partials(const double* v, int n, double* sh) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i];
    return block_sum<double, 256>(s, sh);
}

// g2o OptimizationAlgorithmLevenberg trial decision:
//   rho = (chi - chi') / (x'(lambda x + b) + 1e-3)
//   good: lambda *= max(1/3, min(2/3, 1 - (2 rho - 1)^3)), nu = 2, chi = chi'
//   bad:  lambda *= nu, nu *= 2
__global__ __launch_bounds__(256) void k_lm_accept(const double* __restrict__ red_cost, int nc,
                                                   const double* __restrict__ red_scale, int ns, LMDev* lm) {
    __shared__ double sh[4];
    const double cost = sum_partials(red_cost, nc, sh);
    __syncthreads();
    const double scale = sum_partials(red_scale, ns, sh);
    if (threadIdx.x == 0) {
        LMDev L = *lm;
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
        } else {
            L.lambda *= L.nu;
            L.nu *= 2;
            L.accepted = 0;
        }
        L.temp_chi = temp;
        L.scale = sc;
        L.rho = rho;
        L.trials += 1;
        *lm = L;
    }
}

void launch_lm_accept(const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_accept, dim3(1), dim3(256), 0, st, (const double*)W.red_cost, W.n_cost_blocks,
                       (const double*)W.red_scale, 1 + W.n_pt_blocks, W.lm);
}

__global__ __launch_bounds__(256) void k_lm_set_chi(const double* __restrict__ red_cost, int nc, LMDev* lm) {
    __shared__ double sh[4];
    const double cost = sum_partials(red_cost, nc, sh);
    if (threadIdx.x == 0) { lm->chi = cost; lm->temp_chi = cost; }
}

void launch_lm_set_chi(const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_set_chi, dim3(1), dim3(256), 0, st, (const double*)W.red_cost, W.n_cost_blocks, W.lm);
}

__global__ void k_commit(int n_pt, int n_cam, const LMDev* lm, D4* __restrict__ X, const D4* __restrict__ Xn,
                         PoseQT* __restrict__ pose, const PoseQT* __restrict__ pn, CamRT* __restrict__ cam,
                         const CamRT* __restrict__ cn) {
    if (!lm->accepted) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pt) X[i] = Xn[i];
    if (i < n_cam) { pose[i] = pn[i]; cam[i] = cn[i]; }
}

void launch_commit(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W, hipStream_t st) {
    const int n = P.n_pt > P.n_cam ? P.n_pt : P.n_cam;
    hipLaunchKernelGGL(k_commit, dim3((n + 255) / 256), dim3(256), 0, st, P.n_pt, P.n_cam, (const LMDev*)W.lm,
                       cur.X, (const D4*)cand.X, cur.pose, (const PoseQT*)cand.pose, cur.cam, (const CamRT*)cand.cam);
}

// ---------------------------------------------------------------- outlier flags
// FP64 at the given state: flag = chi2 > thr or depth <= 0; written in the
// caller's observation order.  Also the FP64 robust cost and the flag count.
__global__ __launch_bounds__(256) void k_flags(DevProblem P, const CamRT* __restrict__ cam, const D4* __restrict__ X,
                                               double thr, uint8_t* __restrict__ out, double* __restrict__ red_cost,
                                               double* __restrict__ red_cnt) {
    __shared__ double sh[4];
    double acc = 0.0, cnt = 0.0;
    const int base = blockIdx.x * 256 * kCostPerThread + threadIdx.x;
    for (int i = 0; i < kCostPerThread; ++i) {
        const int a = base + 256 * i;
        if (a < P.n_obs) {
            double R[9], tt[3];
            load_cam(cam, P.o_cam[a], R, tt);
            const D4 xp = X[P.o_pt[a]];
            double e0, e1, zc, w;
            const double s = edge_residual(R, tt, xp.x, xp.y, xp.z, P.fx, P.fy, P.cx, P.cy,
                                           P.o_uv[2 * (size_t)a], P.o_uv[2 * (size_t)a + 1], e0, e1, zc);
            acc += robustify(s, P.robust, P.delta, w);
            const uint8_t f = (s > thr || !(zc > 0.0)) ? 1 : 0;
            out[P.o_orig[a]] = f;
            cnt += f;
        }
    }
    const double s = block_sum<double, 256>(acc, sh);
    __syncthreads();
    const double c = block_sum<double, 256>(cnt, sh);
    if (threadIdx.x == 0) { red_cost[blockIdx.x] = s; red_cnt[blockIdx.x] = c; }
}

void launch_flags(const DevProblem& P, const DevState& s, const DevWork& W, double thr, hipStream_t st) {
    hipLaunchKernelGGL(k_flags, dim3(W.n_cost_blocks), dim3(256), 0, st, P, s.cam, s.X, thr, W.outlier,
                       W.red_cost, W.red_aux);
}

__global__ __launch_bounds__(256) void k_final(const double* __restrict__ red_cost, const double* __restrict__ red_cnt,
                                               int nc, double* __restrict__ out2) {
    __shared__ double sh[4];
    const double c = sum_partials(red_cost, nc, sh);
    __syncthreads();
    const double n = sum_partials(red_cnt, nc, sh);
    if (threadIdx.x == 0) { out2[0] = c; out2[1] = n; }
}

void launch_final(const DevWork& W, double* out2, hipStream_t st) {
    hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, (const double*)W.red_cost, (const double*)W.red_aux,
                       W.n_cost_blocks, out2);
}

// explicit instantiations
#define NAVBA_INST(T)                                                                                         \
    template void launch_lin_pt<T>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);         \
    template void launch_lin_cam<T>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);        \
    template void launch_lambda_init<T>(const DevProblem&, const DevWork&, hipStream_t);                     \
    template void launch_hinv<T>(const DevProblem&, const DevWork&, hipStream_t);                            \
    template void launch_schur_reduce<T>(const DevProblem&, const DevWork&, hipStream_t);                    \
    template void launch_backsub<T>(const DevProblem&, const DevState&, const DevState&, const DevWork&,     \
                                    hipStream_t);                                                            \
    template void launch_cost<T>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);
NAVBA_INST(double)
NAVBA_INST(float)
#undef NAVBA_INST

}  // namespace navba
