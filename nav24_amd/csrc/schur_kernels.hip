// schur_kernels.hip — reduced camera system rows (Schur complement onto the poses).
//
// g2o BlockSolver<6,3>::solve (the solver nav24 builds at
// core/backEnd/BE_GraphOptim.cpp:43-47) forms, per landmark p with blocks
// Hpl_ip (6x3) and D_p = Hll_p + lambda I:
//     S_ij -= Hpl_ip D_p^-1 Hpl_jp^T  (j >= i),   coef_i += Hpl_ip D_p^-1 bl_p.
// Here one workgroup owns one block row f of S (restricted to the cameras
// co-visible with f, and to one chunk of camera f's observations):
//   phase A  one lane per observation a of camera f on a free point p:
//            W_a = Hpl_a Dinv_p (FP64) and the point's coordinates go to the
//            lane's slot of a per-wave LDS buffer; the diagonal pair (a, a)
//            adds W_a Hpl_a^T (symmetric, 21 values) to the lane's registers;
//   phase B  the wave's other pairs (a, b), b a later observation of the same
//            point on a free camera of this column group, come from a pair
//            list built once at prepare time (observation b, owner lane,
//            column slot, camera: 8 bytes, read coalesced), packed as batches
//            of 64 runs: each run holds a few pairs of one slot.  A lane
//            takes one run, reads its camera once, forms each
//            W_a Hpl_b^T = (W_a J_X,b^T)(w_b J_xi,b) from recomputed
//            Jacobians, sums the 6x6 blocks in registers and adds the sum to
//            the slot in LDS once per batch (the lanes of one batch hold
//            different runs, mostly of different slots).
// Work items are dispatched heaviest first (a host-side permutation).
// T is the type of the Jacobian entries only.  Cameras, points, W, the pair
// products and the S-row slots are FP64 for both compute types: the point elimination
// cancels Hpp against Hpl Dinv Hpl^T, and with a few-view point's Dinv
// rounded to FP32 the final cost of an uneven window moved by 1e-3 relative
// in a CPU emulation, against 1e-6 for FP32 Jacobians alone.  (On this GPU
// ds_add_f64 is also the fast LDS atomic; ds_add_f32 measured 8x slower on
// the same access pattern.)
// LDS layouts avoid bank conflicts: the S-row accumulator is entry-major
// (acc[e][slot], lanes on different slots hit consecutive words) and the
// camera table uses an odd stride of 13 values.
#include "ba_kernels.h"

namespace navba {

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kCamStride = 13;  // odd: consecutive cameras fall in different LDS banks

template <typename T>
struct SchurLayout {
    static constexpr int kWbuf = 21;  // W (18) + point (3), FP64
    size_t acc, cam, wbuf, total;
    __host__ __device__ SchurLayout(int ns, int n_cam, bool cam_lds) {  // n_cam: cameras of the largest window
        size_t o = 0;
        acc = o;  o += sizeof(double) * 36 * (size_t)ns;
        cam = o;  o += cam_lds ? sizeof(double) * kCamStride * (size_t)n_cam : 0;
        wbuf = o; o += sizeof(double) * kWbuf * kBlock;
        total = (o + 15) & ~(size_t)15;
    }
};

// Camera c of the window (local index; cam0 = the window's first camera).
template <bool kCamLds>
__device__ __forceinline__ void cam_get(const double* camt, const CamRT* cam, int c, int cam0, double R[9],
                                        double t[3]) {
    if (kCamLds) {
        const double* p = camt + kCamStride * c;
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = p[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) t[k] = p[9 + k];
    } else {
        load_cam(cam, cam0 + c, R, t);
    }
}

template <typename T, bool kCamLds>
__global__ __launch_bounds__(kBlock) void k_schur_row(DevProblem P, const CamRT* __restrict__ cam,
                                                      const D4* __restrict__ X, const T* __restrict__ w,
                                                      const double* __restrict__ Hl, const LMDev* lm,
                                                      double* __restrict__ spart) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int it = P.sitem_order[blockIdx.x];
    const int f = P.sitem_f[it], a0 = P.sitem_a0[it], a1 = P.sitem_a1[it];
    const int g = P.sitem_g[it], ns = P.sitem_nslots[it];
    const int cf = P.free_cam[f];
    const int win = P.cam_win[cf];
    if (!running(lm, win)) return;
    const double lambda = lm[win].lambda;
    // cameras are addressed relative to the window's first camera
    const int cam0 = P.win_cam0[win], ncw = P.win_cam0[win + 1] - cam0;
    const bool g0 = (g == 0);
    const SchurLayout<T> L(ns, P.max_win_cams, kCamLds);
    double* acc = reinterpret_cast<double*>(smem + L.acc);
    double* camt = reinterpret_cast<double*>(smem + L.cam);
    double* wbuf_all = reinterpret_cast<double*>(smem + L.wbuf);

    const int tid = threadIdx.x, lane = tid % kWave;
    const int wid = __builtin_amdgcn_readfirstlane(tid / kWave);  // scalar: rounds and batches are wave-uniform
    for (int i = tid; i < ns * 36; i += kBlock) acc[i] = 0.0;
    if (kCamLds)
        for (int i = tid; i < ncw * 12; i += kBlock)
            camt[kCamStride * (i / 12) + i % 12] = reinterpret_cast<const double*>(cam + cam0)[i];
    __syncthreads();

    double* wbuf = wbuf_all + (size_t)wid * kWave * SchurLayout<T>::kWbuf;

    double coef[6], dsum[21];
#pragma unroll
    for (int k = 0; k < 6; ++k) coef[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) dsum[k] = 0.0;

    int round = 0;
    for (int kb = a0 + wid * kWave; kb < a1; kb += kBlock, ++round) {
        wave_fence();
        // ---- phase A: one lane per observation of camera f
        const int k = kb + lane;
        if (k < a1) {
            const int p = P.c_pt[k];
            if (P.pt_free[p]) {
                double Rf[9], tf[3];  // camera f, reloaded per round (fewer live registers in phase B)
                cam_get<kCamLds>(camt, cam, cf - cam0, cam0, Rf, tf);
                const D4 xp = X[p];
                EdgeLin<T> La;
                const double sa = edge_linearize(Rf, tf, xp.x, xp.y, xp.z, P.fx, P.fy, P.cx, P.cy,
                                                 P.c_uv[2 * (size_t)k], P.c_uv[2 * (size_t)k + 1], La);
                const double wa = (double)huber_weight<T>(sa, P);  // as k_lin_pt, without a gather of w
                double JXa[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) JXa[j] = (double)La.JX[j];
                double Ha[18];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    const double A0 = (double)La.Jxi[i] * wa, A1 = (double)La.Jxi[6 + i] * wa;
#pragma unroll
                    for (int j = 0; j < 3; ++j)  // J_xi is zero at (0, 4) and (1, 3)
                        Ha[3 * i + j] = i == 4 ? A1 * JXa[3 + j] : i == 3 ? A0 * JXa[j] : A0 * JXa[j] + A1 * JXa[3 + j];
                }
                const double* hp = Hl + 9 * (size_t)p;
                double Dp[6];
                point_dinv(hp, lambda, Dp);
                const double d00 = Dp[0], d01 = Dp[1], d02 = Dp[2], d11 = Dp[3], d12 = Dp[4], d22 = Dp[5];
                double* wb = wbuf + lane * SchurLayout<T>::kWbuf;
                double W[18];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    const double h0 = Ha[3 * i], h1 = Ha[3 * i + 1], h2 = Ha[3 * i + 2];
                    W[3 * i + 0] = h0 * d00 + h1 * d01 + h2 * d02;
                    W[3 * i + 1] = h0 * d01 + h1 * d11 + h2 * d12;
                    W[3 * i + 2] = h0 * d02 + h1 * d12 + h2 * d22;
                }
#pragma unroll
                for (int e = 0; e < 18; ++e) wb[e] = W[e];
                wb[18] = xp.x; wb[19] = xp.y; wb[20] = xp.z;
                if (g0) {
                    // Dinv bl of the point (the right-hand side's point term)
                    const double b0 = hp[6], b1 = hp[7], b2 = hp[8];
                    const double e0 = d00 * b0 + d01 * b1 + d02 * b2;
                    const double e1 = d01 * b0 + d11 * b1 + d12 * b2;
                    const double e2 = d02 * b0 + d12 * b1 + d22 * b2;
#pragma unroll
                    for (int i = 0; i < 6; ++i)
                        coef[i] = fma(Ha[3 * i + 2], e2, fma(Ha[3 * i + 1], e1, fma(Ha[3 * i], e0, coef[i])));
                    int q = 0;
#pragma unroll
                    for (int i = 0; i < 6; ++i)
#pragma unroll
                        for (int c = i; c < 6; ++c, ++q)
                            dsum[q] = fma(W[3 * i + 2], Ha[3 * c + 2],
                                          fma(W[3 * i + 1], Ha[3 * c + 1], fma(W[3 * i], Ha[3 * c], dsum[q])));
                }
            }
        }
        wave_fence();
        // ---- phase B: this wave's pair batches of this round
        const int seg = it * P.schur_segs + round * (kBlock / kWave) + wid;
        const int bt0 = P.seg_batch[seg], bt1 = P.seg_batch[seg + 1];
        // the pair entries of a group of kRunLen rows, then their weights, each
        // issued together (one round trip each per group, not per pair); the
        // next batch's first entries are loaded while a batch is formed
        int2 pe[kRunLen];
        if (bt0 < bt1) {
            const int e0 = P.batch_ptr[bt0], nrow = (P.batch_ptr[bt0 + 1] - e0) / kWave;
#pragma unroll
            for (int j = 0; j < kRunLen; ++j) pe[j] = j < nrow ? P.pairs[e0 + j * kWave + lane] : int2{-1, 0};
        }
        for (int bt = bt0; bt < bt1; ++bt) {
            const int e0 = P.batch_ptr[bt], nrow = (P.batch_ptr[bt + 1] - e0) / kWave;
            double pw[kRunLen];
            int2 pn[kRunLen];
            if (bt + 1 < bt1) {
                const int f0 = P.batch_ptr[bt + 1], nr1 = (P.batch_ptr[bt + 2] - f0) / kWave;
#pragma unroll
                for (int j = 0; j < kRunLen; ++j) pn[j] = j < nr1 ? P.pairs[f0 + j * kWave + lane] : int2{-1, 0};
            }
            const int2 p0 = pe[0];
            const bool has = p0.x >= 0;
            const int s = (p0.y >> 6) & 255, cb = has ? p0.y >> 14 : 0;
#pragma unroll
            for (int j = 0; j < kRunLen; ++j) pw[j] = pe[j].x >= 0 ? (double)w[pe[j].x] : 0.0;
            double Rb[9], tb[3];
            cam_get<kCamLds>(camt, cam, cb, cam0, Rb, tb);
            double sacc[36];
#pragma unroll
            for (int e = 0; e < 36; ++e) sacc[e] = 0.0;
            for (int j0 = 0; j0 < nrow; j0 += kRunLen) {
                if (j0 > 0) {  // runs longer than kRunLen (a forced run length)
#pragma unroll
                    for (int j = 0; j < kRunLen; ++j)
                        pe[j] = j0 + j < nrow ? P.pairs[e0 + (j0 + j) * kWave + lane] : int2{-1, 0};
#pragma unroll
                    for (int j = 0; j < kRunLen; ++j) pw[j] = pe[j].x >= 0 ? (double)w[pe[j].x] : 0.0;
                }
#pragma unroll
                for (int jj = 0; jj < kRunLen; ++jj) {
                const int2 pr = pe[jj];
                if (pr.x < 0) continue;
                const double* wb = wbuf + (pr.y & 63) * SchurLayout<T>::kWbuf;
                EdgeLin<T> Lb;
                edge_linearize(Rb, tb, wb[18], wb[19], wb[20], P.fx, P.fy, P.cx, P.cy, 0.0, 0.0, Lb);
                // the Huber weight of b folded into J_X,b: W_a (w_b J_X,b)^T J_xi,b
                const double wgt = pw[jj];
                double JXb[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) JXb[q] = wgt * (double)Lb.JX[q];
                double M0[6], M1[6];
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    const double w0 = wb[3 * r], w1 = wb[3 * r + 1], w2 = wb[3 * r + 2];
                    M0[r] = w0 * JXb[0] + w1 * JXb[1] + w2 * JXb[2];
                    M1[r] = w0 * JXb[3] + w1 * JXb[4] + w2 * JXb[5];
                }
                double B0[6], B1[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) { B0[c] = (double)Lb.Jxi[c]; B1[c] = (double)Lb.Jxi[6 + c]; }
                // J_xi,b is zero at (0, 4) and (1, 3) (edge_linearize: d e_u / d upsilon_y and
                // d e_v / d upsilon_x of a pinhole); those products are skipped
#pragma unroll
                for (int r = 0; r < 6; ++r)
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        double& a = sacc[6 * r + c];
                        if (c == 4) a = fma(M1[r], B1[c], a);
                        else if (c == 3) a = fma(M0[r], B0[c], a);
                        else a = fma(M1[r], B1[c], fma(M0[r], B0[c], a));
                    }
                }
            }
            if (bt + 1 < bt1) {
#pragma unroll
                for (int j = 0; j < kRunLen; ++j) pe[j] = pn[j];
            }
            if (has) {
#pragma unroll
                for (int e = 0; e < 36; ++e) atomicAdd(acc + e * ns + s, sacc[e]);
            }
        }
    }
    // ---- diagonal block (slot 0 of group 0): wave sums, one LDS add per wave;
    // coefficient: fixed-order block reduction; S-row slots out (slot-major)
    __shared__ double red[kBlock / kWave][6];
    if (g0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double v = wave_sum(coef[k]);
            if (lane == 0) red[wid][k] = v;
        }
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int c = i; c < 6; ++c, ++q) {
                const double v = wave_sum(dsum[q]);
                if (lane == 0) {
                    atomicAdd(acc + (6 * i + c) * ns, v);
                    if (c != i) atomicAdd(acc + (6 * c + i) * ns, v);
                }
            }
    }
    __syncthreads();
    double* out = spart + P.sitem_pofs[it];
    if (g0 && tid < 6) {
        double s = 0.0;
        for (int i = 0; i < kBlock / kWave; ++i) s += red[i][tid];
        out[36 * ns + tid] = s;
    }
    for (int i = tid; i < ns * 36; i += kBlock) out[i] = acc[(i % 36) * ns + i / 36];
}

template <typename T>
size_t schur_row_smem(const DevProblem& P) {
    return SchurLayout<T>(P.max_ns, P.max_win_cams, P.max_win_cams <= kCamLdsMax).total;
}

template <typename T>
int prepare_schur_row(const DevProblem& P) {
    const size_t bytes = schur_row_smem<T>(P);
    if (bytes > 160 * 1024) return -1;
    hipError_t e = P.max_win_cams <= kCamLdsMax
        ? hipFuncSetAttribute((const void*)k_schur_row<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
        : hipFuncSetAttribute((const void*)k_schur_row<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e == hipSuccess ? 0 : -2;
}

template <typename T>
void launch_schur_row(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_sitems == 0) return;
    const size_t smem = schur_row_smem<T>(P);
    if (P.max_win_cams <= kCamLdsMax)
        hipLaunchKernelGGL((k_schur_row<T, true>), dim3(P.n_sitems), dim3(kBlock), smem, st, P, s.cam, s.X,
                           (const T*)W.w, (const double*)W.Hl, (const LMDev*)W.lm, W.spart);
    else
        hipLaunchKernelGGL((k_schur_row<T, false>), dim3(P.n_sitems), dim3(kBlock), smem, st, P, s.cam, s.X,
                           (const T*)W.w, (const double*)W.Hl, (const LMDev*)W.lm, W.spart);
}

template void launch_schur_row<double>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);
template void launch_schur_row<float>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);
template int prepare_schur_row<double>(const DevProblem&);
template int prepare_schur_row<float>(const DevProblem&);

}  // namespace navba
