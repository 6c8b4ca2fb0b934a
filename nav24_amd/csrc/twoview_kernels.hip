// twoview_kernels.hip — scoring of two-view RANSAC hypotheses (include/nav_twoview.h):
// TwoViewReconstruction::CheckHomography / ::CheckFundamental
// (core/operators/mapInit/OP_2ViewReconstruction.cpp:447-528, :530-611) over
// every hypothesis of FindHomography / FindFundamental (:266-315, :317-366) at
// once, and the keep-the-best rule of those loops.
//
// k_tv_score: one workgroup of 1024 lanes per hypothesis.  The lanes evaluate
// the two chi-square terms of up to 2048 matches at a time (float, no fused
// multiply-adds: contraction is off in this file), drop the terms that are 0
// (matches above the threshold) and list the rest in LDS in match order; one
// lane then adds the list to the running score in that order, so the score is
// the reference's sequential float sum bit for bit.  That serial chain (one
// dependent add per kept term, up to 2 per match) bounds the kernel; the 24 B
// per match of key points and match indices come from L2 after the first
// hypothesis, so HBM is not the limit.
// k_tv_select: every workgroup picks the kept hypothesis from the scores
// (first strict maximum above 0: the reference's `currentScore > score` scan
// from score = 0) and writes the inlier flags of its share of the matches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_scratch.h"
#include "../../include/nav_twoview.h"

#pragma clang fp contract(off)

namespace navba {

constexpr int kTvThreads = 256;        // k_tv_select
constexpr int kTvScoreThreads = 1024;  // k_tv_score: up to one match per lane per pass
constexpr int kTvChunk = 2048;         // matches per LDS pass: 2 terms each, 16 KB (a multiple of 32)

struct TvArgs {
    float inv_s2;    // 1 / sigma^2, formed in float as the reference does
    float th;        // chi-square threshold of a term
    float th_score;  // what a term below the threshold adds: th_score - chi
};

// Homography: H = (H21, H12), both row-major.  x2in1 = H12 x2 against x1, then
// x1in2 = H21 x1 against x2 (OP_2ViewReconstruction.cpp:494-520).
__device__ __forceinline__ void tv_terms_h(const float* h21, const float* h12, float u1, float v1, float u2,
                                           float v2, const TvArgs& a, float& t1, float& t2, bool& in) {
    const float w2in1inv = 1.0f / (h12[6] * u2 + h12[7] * v2 + h12[8]);
    const float u2in1 = (h12[0] * u2 + h12[1] * v2 + h12[2]) * w2in1inv;
    const float v2in1 = (h12[3] * u2 + h12[4] * v2 + h12[5]) * w2in1inv;
    const float sq1 = (u1 - u2in1) * (u1 - u2in1) + (v1 - v2in1) * (v1 - v2in1);
    const float chi1 = sq1 * a.inv_s2;
    const float w1in2inv = 1.0f / (h21[6] * u1 + h21[7] * v1 + h21[8]);
    const float u1in2 = (h21[0] * u1 + h21[1] * v1 + h21[2]) * w1in2inv;
    const float v1in2 = (h21[3] * u1 + h21[4] * v1 + h21[5]) * w1in2inv;
    const float sq2 = (u2 - u1in2) * (u2 - u1in2) + (v2 - v1in2) * (v2 - v1in2);
    const float chi2 = sq2 * a.inv_s2;
    // `chi > th` is false for NaN: the reference then adds th - NaN
    const bool out1 = chi1 > a.th, out2 = chi2 > a.th;
    t1 = out1 ? 0.0f : a.th_score - chi1;
    t2 = out2 ? 0.0f : a.th_score - chi2;
    in = !(out1 || out2);
}

// Fundamental: l2 = F21 x1 = (a2, b2, c2) against x2, then l1 = F21^T x2
// against x1 (OP_2ViewReconstruction.cpp:571-603).
__device__ __forceinline__ void tv_terms_f(const float* f, float u1, float v1, float u2, float v2, const TvArgs& a,
                                           float& t1, float& t2, bool& in) {
    const float a2 = f[0] * u1 + f[1] * v1 + f[2];
    const float b2 = f[3] * u1 + f[4] * v1 + f[5];
    const float c2 = f[6] * u1 + f[7] * v1 + f[8];
    const float num2 = a2 * u2 + b2 * v2 + c2;
    const float sq1 = num2 * num2 / (a2 * a2 + b2 * b2);
    const float chi1 = sq1 * a.inv_s2;
    const float a1 = f[0] * u2 + f[3] * v2 + f[6];
    const float b1 = f[1] * u2 + f[4] * v2 + f[7];
    const float c1 = f[2] * u2 + f[5] * v2 + f[8];
    const float num1 = a1 * u1 + b1 * v1 + c1;
    const float sq2 = num1 * num1 / (a1 * a1 + b1 * b1);
    const float chi2 = sq2 * a.inv_s2;
    const bool out1 = chi1 > a.th, out2 = chi2 > a.th;
    t1 = out1 ? 0.0f : a.th_score - chi1;
    t2 = out2 ? 0.0f : a.th_score - chi2;
    in = !(out1 || out2);
}

// Both terms of match i under hypothesis `h` (kHom: H21 = m0 + 9 h,
// H12 = m1 + 9 h; otherwise F21 = m0 + 9 h).
template <bool kHom>
__device__ __forceinline__ void tv_match(const float2* __restrict__ kp1, const float2* __restrict__ kp2,
                                         const int2* __restrict__ mt, int i, const float* m0, const float* m1,
                                         const TvArgs& a, float& t1, float& t2, bool& in) {
    const int2 p = mt[i];
    const float2 x1 = kp1[p.x], x2 = kp2[p.y];
    if (kHom)
        tv_terms_h(m0, m1, x1.x, x1.y, x2.x, x2.y, a, t1, t2, in);
    else
        tv_terms_f(m0, x1.x, x1.y, x2.x, x2.y, a, t1, t2, in);
}

// Lane 0's in-order sum of terms[0, p): p a multiple of 64, read as float4
// in two alternating batches of 8 so that the next batch's LDS reads are in
// flight while the current one is added.
__device__ __forceinline__ float tv_serial_sum(const float* terms, int p, float score) {
    const float4* T = reinterpret_cast<const float4*>(terms);
    const int nb2 = p >> 6;  // pairs of 8-float4 batches
    float4 a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = T[k];
    for (int i = 0; i < nb2; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) b[k] = T[16 * i + 8 + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            score += a[k].x;
            score += a[k].y;
            score += a[k].z;
            score += a[k].w;
        }
        const int nx = 16 * min(i + 1, nb2 - 1);  // the last pass rereads a batch in range
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = T[nx + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            score += b[k].x;
            score += b[k].y;
            score += b[k].z;
            score += b[k].w;
        }
    }
    return score;
}

template <bool kHom>
__global__ __launch_bounds__(kTvScoreThreads) void k_tv_score(const float2* __restrict__ kp1,
                                                              const float2* __restrict__ kp2,
                                                              const int2* __restrict__ mt, int n,
                                                              const float* __restrict__ M0,
                                                              const float* __restrict__ M1, TvArgs a,
                                                              float* __restrict__ scores) {
    __shared__ __align__(16) float terms[2 * kTvChunk];
    const int h = blockIdx.x;
    float m0[9], m1[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        m0[c] = M0[9 * h + c];
        m1[c] = kHom ? M1[9 * h + c] : 0.0f;
    }
    __shared__ int wtot[kTvScoreThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    float score = 0.0f;  // lane 0 only
    for (int c0 = 0; c0 < n; c0 += kTvChunk) {
        const int cn = min(kTvChunk, n - c0);
        // The reference adds term 1 then term 2 of each match in turn.  A term
        // above the threshold is skipped there; here it is 0 and is left out
        // of the list, and so is any other +0 term: adding +0 to a sum that is
        // never -0 leaves it unchanged.  The kept terms go to terms[] in match
        // order (ballot counts within a wave, wave totals across the group).
        int fill = 0;
        for (int i0 = 0; i0 < cn; i0 += kTvScoreThreads) {
            const int i = i0 + threadIdx.x;
            float t1 = 0.0f, t2 = 0.0f;
            if (i < cn) {
                bool in;
                tv_match<kHom>(kp1, kp2, mt, c0 + i, m0, m1, a, t1, t2, in);
            }
            const bool k1 = t1 != 0.0f, k2 = t2 != 0.0f;  // NaN is kept
            const uint64_t b1 = __ballot(k1), b2 = __ballot(k2);
            if (lane == 0) wtot[w] = __popcll(b1) + __popcll(b2);
            __syncthreads();
            int pos = fill + __popcll(b1 & below) + __popcll(b2 & below), tot = 0;
#pragma unroll
            for (int v = 0; v < kTvScoreThreads / 64; ++v) {
                const int x = wtot[v];
                pos += v < w ? x : 0;
                tot += x;
            }
            if (k1) terms[pos++] = t1;
            if (k2) terms[pos] = t2;
            fill += tot;
            __syncthreads();  // wtot is rewritten by the next pass
        }
        const int p = (fill + 63) & ~63;  // zero padding to whole read batches
        for (int i = fill + threadIdx.x; i < p; i += kTvScoreThreads) terms[i] = 0.0f;
        __syncthreads();
        if (threadIdx.x == 0 && p > 0) score = tv_serial_sum(terms, p, score);
        __syncthreads();
    }
    if (threadIdx.x == 0) scores[h] = score;
}

// Kept hypothesis: highest score above 0, the first of equal ones; NaN never.
struct TvBest {
    float s;
    int i;
};

__device__ __forceinline__ TvBest tv_better(TvBest x, TvBest y) {
    if (y.i < 0) return x;
    if (x.i < 0) return y;
    if (y.s > x.s || (y.s == x.s && y.i < x.i)) return y;
    return x;
}

template <bool kHom>
__global__ __launch_bounds__(kTvThreads) void k_tv_select(const float2* __restrict__ kp1,
                                                          const float2* __restrict__ kp2,
                                                          const int2* __restrict__ mt, int n,
                                                          const float* __restrict__ M0,
                                                          const float* __restrict__ M1, TvArgs a,
                                                          const float* __restrict__ scores, int n_hyp,
                                                          int32_t* __restrict__ best, float* __restrict__ best_score,
                                                          uint8_t* __restrict__ inliers) {
    __shared__ TvBest red[kTvThreads / 64];
    // this lane's first match, loaded while the kept hypothesis is chosen
    const int i0 = blockIdx.x * kTvThreads + threadIdx.x, stride = gridDim.x * kTvThreads;
    float2 x1 = make_float2(0.0f, 0.0f), x2 = x1;
    if (i0 < n) {
        const int2 p = mt[i0];
        x1 = kp1[p.x];
        x2 = kp2[p.y];
    }
    TvBest b{0.0f, -1};
    for (int k = threadIdx.x; k < n_hyp; k += kTvThreads) {
        const float s = scores[k];
        if (s > 0.0f) b = tv_better(b, TvBest{s, k});
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const TvBest o{__shfl_xor(b.s, off), __shfl_xor(b.i, off)};
        b = tv_better(b, o);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    TvBest r = red[0];
#pragma unroll
    for (int w = 1; w < kTvThreads / 64; ++w) r = tv_better(r, red[w]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *best = r.i;
        *best_score = r.i < 0 ? 0.0f : r.s;
    }
    float m0[9], m1[9];
    const int h = max(r.i, 0);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        m0[c] = r.i < 0 ? 0.0f : M0[9 * h + c];
        m1[c] = (kHom && r.i >= 0) ? M1[9 * h + c] : 0.0f;
    }
    for (int i = i0; i < n; i += stride) {
        if (i != i0) {
            const int2 p = mt[i];
            x1 = kp1[p.x];
            x2 = kp2[p.y];
        }
        bool in = false;
        if (r.i >= 0) {
            float t1, t2;
            if (kHom)
                tv_terms_h(m0, m1, x1.x, x1.y, x2.x, x2.y, a, t1, t2, in);
            else
                tv_terms_f(m0, x1.x, x1.y, x2.x, x2.y, a, t1, t2, in);
        }
        inliers[i] = in ? 1 : 0;
    }
}

// ---- CheckRT (OP_2ViewReconstruction.cpp:1244-1352) with Triangulate (:1179-1194)
//
// k_tv_check_rt: one lane per (match, motion candidate).  Each inlier match is
// triangulated from the 4x4 system of Triangulate, its null vector taken by a
// one-sided Jacobi SVD in double and rounded to float; then CheckRT's float
// tests in its order: finite point, parallax cosine, depth in front of both
// cameras unless the parallax is tiny, reprojection error in both images.
// Matrix products, norms and dot products accumulate float products in
// double and round to float, as OpenCV's 32F gemm / norm / dot do.  A passing
// match writes its point and good flag at its image-1 key point, and its
// cosine and pass flag at its own index.
// k_tv_rt_select: one workgroup per candidate counts the passing matches and
// selects the cosine at index min(min_tri, nGood - 1) of their ascending
// order (the reference sorts and indexes) by a 4-pass radix select over
// order-preserving keys of the float bits; parallax = acos of it in degrees.

constexpr int kTvMaxCand = NAV_TV_MAX_CANDIDATES;

struct RtCand {
    float P2[12];  // K [R|t], row-major 3x4
    float O2[3];   // -R^T t
    float R[9];
    float t[3];
};

struct RtArgs {
    float fx, fy, cx, cy;
    float P1[12];  // K [I|0]
    float th2, min_cos;
    RtCand c[kTvMaxCand];
};

// Right singular vector of the smallest singular value of the 4x4 row-major a
// (overwritten): Jacobi rotations on column pairs (0,1) (0,2) (0,3) (1,2)
// (1,3) (2,3), up to 30 sweeps, a pair skipped once orthogonal to 1e-15 of
// its norms; the column of V with the smallest rotated norm (the first of equal).
__device__ void tv_null_vector4(double* a, double* x) {
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    al = al + a[4 * r + p] * a[4 * r + p];
                    be = be + a[4 * r + q] * a[4 * r + q];
                    ga = ga + a[4 * r + p] * a[4 * r + q];
                }
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double ap = a[4 * r + p], aq = a[4 * r + q];
                    a[4 * r + p] = c * ap - s * aq;
                    a[4 * r + q] = s * ap + c * aq;
                    const double vp = v[4 * r + p], vq = v[4 * r + q];
                    v[4 * r + p] = c * vp - s * vq;
                    v[4 * r + q] = s * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    int k = 0;
    double best = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double nrm = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) nrm = nrm + a[4 * r + c] * a[4 * r + c];
        if (c == 0 || nrm < best) {
            best = nrm;
            k = c;
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = k == 0 ? v[4 * r] : k == 1 ? v[4 * r + 1] : k == 2 ? v[4 * r + 2] : v[4 * r + 3];
}

// a . b over 3 floats (b with stride sb), products and sum in double, rounded to float
__device__ __forceinline__ float tv_dot3(const float* a, const float* b, int sb) {
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) r = r + (double)a[k] * (double)b[sb * k];
    return (float)r;
}

__device__ __forceinline__ float tv_norm3(const float* x) {
    return (float)sqrt(((0.0 + (double)x[0] * x[0]) + (double)x[1] * x[1]) + (double)x[2] * x[2]);
}

__global__ __launch_bounds__(256) void k_tv_rt_clear(float* __restrict__ p3d, uint8_t* __restrict__ good, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    p3d[3 * i] = 0.0f;
    p3d[3 * i + 1] = 0.0f;
    p3d[3 * i + 2] = 0.0f;
    good[i] = 0;
}

__global__ __launch_bounds__(256) void k_tv_check_rt(const float2* __restrict__ kp1, const float2* __restrict__ kp2,
                                                     const int2* __restrict__ mt, const uint8_t* __restrict__ inl,
                                                     int n, int n_kp1, RtArgs g, float* __restrict__ p3d,
                                                     uint8_t* __restrict__ good, float* __restrict__ cosp,
                                                     uint8_t* __restrict__ passed) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int c = blockIdx.y;
    const RtCand& C = g.c[c];
    float cp = 0.0f;
    uint8_t ok = 0;
    if (inl[i]) {
        const int2 m = mt[i];
        const float2 x1 = kp1[m.x], x2 = kp2[m.y];
        double A[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            A[q] = (double)(x1.x * g.P1[8 + q] - g.P1[q]);
            A[4 + q] = (double)(x1.y * g.P1[8 + q] - g.P1[4 + q]);
            A[8 + q] = (double)(x2.x * C.P2[8 + q] - C.P2[q]);
            A[12 + q] = (double)(x2.y * C.P2[8 + q] - C.P2[4 + q]);
        }
        double xh[4];
        tv_null_vector4(A, xh);
        const float w = (float)xh[3];
        const float X[3] = {(float)xh[0] / w, (float)xh[1] / w, (float)xh[2] / w};
        do {
            if (!isfinite(X[0]) || !isfinite(X[1]) || !isfinite(X[2])) break;
            const float n2[3] = {X[0] - C.O2[0], X[1] - C.O2[1], X[2] - C.O2[2]};
            const float d1 = tv_norm3(X), d2 = tv_norm3(n2);
            const double dt = ((0.0 + (double)X[0] * n2[0]) + (double)X[1] * n2[1]) + (double)X[2] * n2[2];
            const float cosP = (float)(dt / (double)(d1 * d2));
            if (X[2] <= 0 && cosP < g.min_cos) break;
            float X2[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) X2[r] = tv_dot3(C.R + 3 * r, X, 1) + C.t[r];
            if (X2[2] <= 0 && cosP < g.min_cos) break;
            const float iz1 = 1.0f / X[2];
            const float im1x = g.fx * X[0] * iz1 + g.cx, im1y = g.fy * X[1] * iz1 + g.cy;
            const float e1 = (im1x - x1.x) * (im1x - x1.x) + (im1y - x1.y) * (im1y - x1.y);
            if (e1 > g.th2) break;
            const float iz2 = 1.0f / X2[2];
            const float im2x = g.fx * X2[0] * iz2 + g.cx, im2y = g.fy * X2[1] * iz2 + g.cy;
            const float e2 = (im2x - x2.x) * (im2x - x2.x) + (im2y - x2.y) * (im2y - x2.y);
            if (e2 > g.th2) break;
            ok = 1;
            cp = cosP;
            float* P = p3d + 3 * ((size_t)n_kp1 * c + m.x);
            P[0] = X[0];
            P[1] = X[1];
            P[2] = X[2];
            good[(size_t)n_kp1 * c + m.x] = cosP < g.min_cos ? 1 : 0;
        } while (false);
    }
    cosp[(size_t)n * c + i] = cp;
    passed[(size_t)n * c + i] = ok;
}

// order-preserving unsigned key of a float (NaN above +inf)
__device__ __forceinline__ uint32_t tv_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float tv_unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

constexpr int kTvSelThreads = 1024;

__global__ __launch_bounds__(kTvSelThreads) void k_tv_rt_select(const float* __restrict__ cosp,
                                                                const uint8_t* __restrict__ passed, int n,
                                                                int min_tri, int32_t* __restrict__ n_good,
                                                                float* __restrict__ parallax,
                                                                float* __restrict__ cos_sel) {
    __shared__ int hist[256];
    __shared__ int wsum[kTvSelThreads / 64];
    __shared__ int sel[2];  // digit, rank left
    const int c = blockIdx.x;
    const float* cs = cosp + (size_t)n * c;
    const uint8_t* ps = passed + (size_t)n * c;
    int cnt = 0;
    for (int i = threadIdx.x; i < n; i += kTvSelThreads) cnt += ps[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    int ng = 0;
#pragma unroll
    for (int w = 0; w < kTvSelThreads / 64; ++w) ng += wsum[w];
    if (ng == 0) {
        if (threadIdx.x == 0) {
            n_good[c] = 0;
            parallax[c] = 0.0f;
            cos_sel[c] = __uint_as_float(0x7fc00000u);
        }
        return;
    }
    int rank = min(min_tri, ng - 1);
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += kTvSelThreads) hist[b] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kTvSelThreads) {
            if (!ps[i]) continue;
            const uint32_t k = tv_key(cs[i]);
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1);
        }
        __syncthreads();
        // the bucket holding the rank: inclusive prefix of the 256 counts over
        // the first 4 waves (shuffles within a wave, wave totals through LDS);
        // exactly one bucket has excl <= rank < incl
        int h = 0, incl = 0;
        if (threadIdx.x < 256) {
            h = hist[threadIdx.x];
            incl = h;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off);
                if ((threadIdx.x & 63) >= off) incl += y;
            }
            if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
        }
        __syncthreads();
        if (threadIdx.x < 256) {
#pragma unroll
            for (int w = 0; w < 4; ++w) incl += w < (threadIdx.x >> 6) ? wsum[w] : 0;
            const int excl = incl - h;
            if (excl <= rank && rank < incl) {
                sel[0] = threadIdx.x;
                sel[1] = rank - excl;
            }
        }
        __syncthreads();
        prefix |= (uint32_t)sel[0] << shift;
        mask |= 255u << shift;
        rank = sel[1];
        __syncthreads();  // hist and sel are rewritten by the next pass
    }
    if (threadIdx.x == 0) {
        const float cv = tv_unkey(prefix);
        n_good[c] = ng;
        cos_sel[c] = cv;
        parallax[c] = static_cast<float>(acosf(cv) * 180 / 3.14159265358979323846);
    }
}

// ---- hypotheses of FindHomography / FindFundamental (OP_2ViewReconstruction.cpp:266-366)
//
// k_tv_hyp: one lane per RANSAC set (8 matches) and model, both models in one
// launch.  ComputeH21 (:368-407) / ComputeF21 (:410-446) take the right
// singular vector of the smallest singular value of A from OpenCV's float
// SVD; here it is the eigenvector of the smallest eigenvalue of A^T A
// (accumulated in double, 45 entries mirrored), by inverse iteration on
// A^T A + mu I (mu = 1e-13 trace; Cholesky in place, the diagonal's
// reciprocals formed once; 40 iterations from (1, ..., 1)), unit length, sign
// with the largest entry positive, rounded to float.  F's rank-2 step is
// Fpre (I - w w^T) with w that eigenvector of Fpre^T Fpre.  H21 = T2^-1 Hn T1,
// H12 = H21^-1 and F21 = T2^T Fn T1 with 3x3 products summed in double and
// rounded to float, and the inverse by cofactors over the determinant in double.

struct HypArgs {
    float T1[9], T2inv[9], T2t[9];
};

template <int N>
__device__ void tv_small_eigvec(double* M, double* v) {  // M (N x N, row-major) is overwritten
    double tr = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) tr = tr + M[N * i + i];
    const double mu = 1e-13 * tr;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double a = M[N * i + j] + (i == j ? mu : 0.0);
#pragma unroll
            for (int k = 0; k < j; ++k) a = a - M[N * i + k] * M[N * j + k];
            if (i == j)
                M[N * i + i] = sqrt(a > 0.0 ? a : mu);
            else
                M[N * i + j] = a / M[N * j + j];
        }
    double x[N], rd[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        x[i] = 1.0;
        rd[i] = 1.0 / M[N * i + i];  // the diagonal's reciprocals, formed once
    }
    for (int it = 0; it < 40; ++it) {
        double y[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double a = x[i];
#pragma unroll
            for (int k = 0; k < i; ++k) a = a - M[N * i + k] * y[k];
            y[i] = a * rd[i];
        }
#pragma unroll
        for (int i = N - 1; i >= 0; --i) {
            double a = y[i];
#pragma unroll
            for (int k = i + 1; k < N; ++k) a = a - M[N * k + i] * x[k];
            x[i] = a * rd[i];
        }
        double nrm = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) nrm = nrm + x[i] * x[i];
        const double r = 1.0 / sqrt(nrm);
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = x[i] * r;
    }
    int imax = 0;
    double amax = fabs(x[0]);
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (fabs(x[i]) > amax) {
            amax = fabs(x[i]);
            imax = i;
        }
    double xm = x[0];
#pragma unroll
    for (int i = 1; i < N; ++i) xm = imax == i ? x[i] : xm;
    const double sg = xm < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = x[i] * sg;
}

__device__ __forceinline__ void tv_mul3(const float* A, const float* B, float* C) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) a = a + (double)A[3 * r + k] * (double)B[3 * k + c];
            C[3 * r + c] = (float)a;
        }
}

__host__ __device__ inline void tv_inv3f(const float* S, float* D) {
    const double d0 = (double)S[4] * S[8] - (double)S[5] * S[7];
    const double d1 = (double)S[3] * S[8] - (double)S[5] * S[6];
    const double d2 = (double)S[3] * S[7] - (double)S[4] * S[6];
    double det = (double)S[0] * d0 - (double)S[1] * d1 + (double)S[2] * d2;
    if (det == 0.0) {
        for (int k = 0; k < 9; ++k) D[k] = 0.0f;
        return;
    }
    det = 1.0 / det;
    D[0] = (float)(d0 * det);
    D[1] = (float)(((double)S[2] * S[7] - (double)S[1] * S[8]) * det);
    D[2] = (float)(((double)S[1] * S[5] - (double)S[2] * S[4]) * det);
    D[3] = (float)(-d1 * det);
    D[4] = (float)(((double)S[0] * S[8] - (double)S[2] * S[6]) * det);
    D[5] = (float)(((double)S[2] * S[3] - (double)S[0] * S[5]) * det);
    D[6] = (float)(d2 * det);
    D[7] = (float)(((double)S[1] * S[6] - (double)S[0] * S[7]) * det);
    D[8] = (float)(((double)S[0] * S[4] - (double)S[1] * S[3]) * det);
}

// A^T A of ComputeH21's (kHom) or ComputeF21's rows for 8 normalised pairs:
// the upper triangle accumulated pair by pair, mirrored to the lower.
template <bool kHom>
__device__ __forceinline__ void tv_hyp_normal(const float* p1, const float* p2, double* M) {
#pragma unroll
    for (int k = 0; k < 81; ++k) M[k] = 0.0;
    for (int i = 0; i < 8; ++i) {
        const float u1 = p1[2 * i], v1 = p1[2 * i + 1], u2 = p2[2 * i], v2 = p2[2 * i + 1];
        float r0[9], r1[9];
        if (kHom) {
            const float a[9] = {0.0f, 0.0f, 0.0f, -u1, -v1, -1.0f, v2 * u1, v2 * v1, v2};
            const float b[9] = {u1, v1, 1.0f, 0.0f, 0.0f, 0.0f, -u2 * u1, -u2 * v1, -u2};
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                r0[k] = a[k];
                r1[k] = b[k];
            }
        } else {
            const float a[9] = {u2 * u1, u2 * v1, u2, v2 * u1, v2 * v1, v2, u1, v1, 1.0f};
#pragma unroll
            for (int k = 0; k < 9; ++k) r0[k] = a[k];
        }
#pragma unroll
        for (int a = 0; a < 9; ++a)
#pragma unroll
            for (int b = a; b < 9; ++b) M[9 * a + b] = M[9 * a + b] + (double)r0[a] * (double)r0[b];
        if (kHom) {
#pragma unroll
            for (int a = 0; a < 9; ++a)
#pragma unroll
                for (int b = a; b < 9; ++b) M[9 * a + b] = M[9 * a + b] + (double)r1[a] * (double)r1[b];
        }
    }
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
        for (int b = 0; b < a; ++b) M[9 * a + b] = M[9 * b + a];
}

template <bool kHom>
__device__ __forceinline__ void tv_hyp_one(const float2* __restrict__ q1, const float2* __restrict__ q2,
                                           const int2* __restrict__ mt, const int* __restrict__ sets, int it,
                                           const HypArgs& g, float* __restrict__ M0, float* __restrict__ M1) {
    float p1[16], p2[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int2 m = mt[sets[8 * it + j]];
        const float2 a = q1[m.x], b = q2[m.y];
        p1[2 * j] = a.x;
        p1[2 * j + 1] = a.y;
        p2[2 * j] = b.x;
        p2[2 * j + 1] = b.y;
    }
    double M[81], v[9];
    tv_hyp_normal<kHom>(p1, p2, M);
    tv_small_eigvec<9>(M, v);
    float Xn[9], tmp[9], out[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Xn[k] = (float)v[k];
    if (kHom) {
        tv_mul3(g.T2inv, Xn, tmp);
        tv_mul3(tmp, g.T1, out);
#pragma unroll
        for (int k = 0; k < 9; ++k) M0[9 * it + k] = out[k];
        float inv[9];
        tv_inv3f(out, inv);
#pragma unroll
        for (int k = 0; k < 9; ++k) M1[9 * it + k] = inv[k];
    } else {
        double G[9], w[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc = acc + (double)Xn[3 * k + a] * (double)Xn[3 * k + b];
                G[3 * a + b] = acc;
            }
        tv_small_eigvec<3>(G, w);
        float Fn[9];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double fw = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) fw = fw + (double)Xn[3 * r + k] * w[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) Fn[3 * r + c] = (float)((double)Xn[3 * r + c] - fw * w[c]);
        }
        tv_mul3(g.T2t, Fn, tmp);
        tv_mul3(tmp, g.T1, out);
#pragma unroll
        for (int k = 0; k < 9; ++k) M0[9 * it + k] = out[k];
    }
}

// Both models in one launch, side by side: workgroups [0, nb) the homographies,
// [nb, 2 nb) the fundamental matrices (nb = ceil(n_sets / 64)).
__global__ __launch_bounds__(64) void k_tv_hyp(const float2* __restrict__ q1, const float2* __restrict__ q2,
                                               const int2* __restrict__ mt, const int* __restrict__ sets, int n_sets,
                                               HypArgs g, float* __restrict__ H21, float* __restrict__ H12,
                                               float* __restrict__ F21) {
    const int nb = (n_sets + 63) / 64;
    const bool hom = (int)blockIdx.x < nb;
    const int it = (hom ? (int)blockIdx.x : (int)blockIdx.x - nb) * 64 + threadIdx.x;
    if (it >= n_sets) return;
    if (hom)
        tv_hyp_one<true>(q1, q2, mt, sets, it, g, H21, H12);
    else
        tv_hyp_one<false>(q1, q2, mt, sets, it, g, F21, nullptr);
}

// Normalize (OP_2ViewReconstruction.cpp:1195-1240) on the host: float sums in input order.
static void tv_normalize_host(const float* kp, int32_t n, float* out, float* T) {
    float meanX = 0, meanY = 0;
    for (int32_t i = 0; i < n; ++i) {
        meanX += kp[2 * i];
        meanY += kp[2 * i + 1];
    }
    meanX = meanX / n;
    meanY = meanY / n;
    float meanDevX = 0, meanDevY = 0;
    for (int32_t i = 0; i < n; ++i) {
        out[2 * i] = kp[2 * i] - meanX;
        out[2 * i + 1] = kp[2 * i + 1] - meanY;
        meanDevX += std::fabs(out[2 * i]);
        meanDevY += std::fabs(out[2 * i + 1]);
    }
    meanDevX = meanDevX / n;
    meanDevY = meanDevY / n;
    const float sX = 1.0f / meanDevX, sY = 1.0f / meanDevY;
    for (int32_t i = 0; i < n; ++i) {
        out[2 * i] = out[2 * i] * sX;
        out[2 * i + 1] = out[2 * i + 1] * sY;
    }
    const float Tm[9] = {sX, 0.0f, -meanX * sX, 0.0f, sY, -meanY * sY, 0.0f, 0.0f, 1.0f};
    for (int k = 0; k < 9; ++k) T[k] = Tm[k];
}

static HypArgs hyp_args(const float* T1, const float* T2) {
    HypArgs g;
    for (int k = 0; k < 9; ++k) g.T1[k] = T1[k];
    tv_inv3f(T2, g.T2inv);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) g.T2t[3 * r + c] = T2[3 * c + r];
    return g;
}

static TvArgs tv_args(float sigma, float th, float th_score) {
    TvArgs a;
    a.inv_s2 = 1.0f / (sigma * sigma);
    a.th = th;
    a.th_score = th_score;
    return a;
}

template <bool kHom>
static int32_t tv_launch(const float* kp1, const float* kp2, const int32_t* matches, int32_t n, const float* M0,
                         const float* M1, int32_t n_hyp, TvArgs a, float* scores, int32_t* best,
                         float* best_score, uint8_t* inliers, hipStream_t st) {
    const float2* k1 = reinterpret_cast<const float2*>(kp1);
    const float2* k2 = reinterpret_cast<const float2*>(kp2);
    const int2* mt = reinterpret_cast<const int2*>(matches);
    if (n_hyp > 0)
        hipLaunchKernelGGL(k_tv_score<kHom>, dim3(n_hyp), dim3(kTvScoreThreads), 0, st, k1, k2, mt, n, M0, M1, a,
                           scores);
    const int blocks = std::max(1, std::min(1024, (n + kTvThreads - 1) / kTvThreads));
    hipLaunchKernelGGL(k_tv_select<kHom>, dim3(blocks), dim3(kTvThreads), 0, st, k1, k2, mt, n, M0, M1, a,
                       (const float*)scores, n_hyp, best, best_score, inliers);
    return hipGetLastError() == hipSuccess ? NAV_BA_OK : NAV_BA_HIP_ERROR;
}

static bool tv_sizes_ok(int32_t n, int32_t n_hyp) { return n >= 0 && n <= NAV_TV_MAX_MATCHES && n_hyp >= 0; }

static bool tv_gfx950(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Host arrays: one device buffer for inputs and outputs, copied in and out.
template <bool kHom>
static int32_t tv_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2, const int32_t* matches,
                       int32_t n, const float* M0, const float* M1, int32_t n_hyp, TvArgs a, float* scores,
                       int32_t* best, float* best_score, uint8_t* inliers, int32_t device) {
    if (!tv_gfx950(device)) return NAV_BA_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return NAV_BA_HIP_ERROR;
    const int nm = kHom ? 2 : 1;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_k1 = al(8 * (size_t)n1), b_k2 = al(8 * (size_t)n2), b_mt = al(8 * (size_t)n),
                 b_m = al(36 * (size_t)n_hyp), b_sc = al(4 * (size_t)n_hyp), b_out = al(8), b_in = al((size_t)n);
    const size_t total = b_k1 + b_k2 + b_mt + nm * b_m + b_sc + b_out + b_in;
    HostScratch& hs = host_scratch(device);
    std::lock_guard<std::mutex> lock(hs.mu);
    char* d = (char*)host_scratch_get(hs, total);
    if (!d) return NAV_BA_HIP_ERROR;
    char* p = d;
    float* d_k1 = (float*)p;
    p += b_k1;
    float* d_k2 = (float*)p;
    p += b_k2;
    int32_t* d_mt = (int32_t*)p;
    p += b_mt;
    float* d_m0 = (float*)p;
    p += b_m;
    float* d_m1 = kHom ? (float*)p : nullptr;
    if (kHom) p += b_m;
    float* d_sc = (float*)p;
    p += b_sc;
    int32_t* d_best = (int32_t*)p;
    float* d_bs = (float*)(p + 4);
    p += b_out;
    uint8_t* d_in = (uint8_t*)p;
    int32_t st = NAV_BA_OK;
    auto up = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && b > 0 && hipMemcpy(dst, src, b, hipMemcpyHostToDevice) != hipSuccess)
            st = NAV_BA_HIP_ERROR;
    };
    auto down = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && b > 0 && hipMemcpy(dst, src, b, hipMemcpyDeviceToHost) != hipSuccess)
            st = NAV_BA_HIP_ERROR;
    };
    up(d_k1, kp1, 8 * (size_t)n1);
    up(d_k2, kp2, 8 * (size_t)n2);
    up(d_mt, matches, 8 * (size_t)n);
    up(d_m0, M0, 36 * (size_t)n_hyp);
    if (kHom) up(d_m1, M1, 36 * (size_t)n_hyp);
    if (st == NAV_BA_OK)
        st = tv_launch<kHom>(d_k1, d_k2, d_mt, n, d_m0, d_m1, n_hyp, a, d_sc, d_best, d_bs, d_in, nullptr);
    down(scores, d_sc, 4 * (size_t)n_hyp);
    down(best, d_best, 4);
    down(best_score, d_bs, 4);
    down(inliers, d_in, (size_t)n);
    return st;
}

// Host-side check of every match against the key-point counts.
static bool tv_matches_ok(const int32_t* matches, int32_t n, int32_t n1, int32_t n2) {
    for (int32_t i = 0; i < n; ++i) {
        const int32_t a = matches[2 * i], b = matches[2 * i + 1];
        if (a < 0 || a >= n1 || b < 0 || b >= n2) return false;
    }
    return true;
}

}  // namespace navba

using namespace navba;

extern "C" int32_t nav_tv_score_homography(const float* kp1, const float* kp2, const int32_t* matches,
                                           int32_t n_matches, const float* H21, const float* H12, int32_t n_hyp,
                                           float sigma, float th_score, float* scores, int32_t* best,
                                           float* best_score, uint8_t* inliers, void* stream) {
    if (!tv_sizes_ok(n_matches, n_hyp) || !(sigma > 0.0f) || !best || !best_score) return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers)) return NAV_BA_BAD_ARG;
    if (n_hyp > 0 && (!H21 || !H12 || !scores)) return NAV_BA_BAD_ARG;
    return tv_launch<true>(kp1, kp2, matches, n_matches, H21, H12, n_hyp, tv_args(sigma, th_score, th_score), scores,
                           best, best_score, inliers, (hipStream_t)stream);
}

extern "C" int32_t nav_tv_score_fundamental(const float* kp1, const float* kp2, const int32_t* matches,
                                            int32_t n_matches, const float* F21, int32_t n_hyp, float sigma,
                                            float th_f, float th_score, float* scores, int32_t* best,
                                            float* best_score, uint8_t* inliers, void* stream) {
    if (!tv_sizes_ok(n_matches, n_hyp) || !(sigma > 0.0f) || !best || !best_score) return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers)) return NAV_BA_BAD_ARG;
    if (n_hyp > 0 && (!F21 || !scores)) return NAV_BA_BAD_ARG;
    return tv_launch<false>(kp1, kp2, matches, n_matches, F21, nullptr, n_hyp, tv_args(sigma, th_f, th_score),
                            scores, best, best_score, inliers, (hipStream_t)stream);
}

extern "C" int32_t nav_tv_score_homography_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2,
                                                const int32_t* matches, int32_t n_matches, const float* H21,
                                                const float* H12, int32_t n_hyp, float sigma, float th_score,
                                                float* scores, int32_t* best, float* best_score, uint8_t* inliers,
                                                int32_t device) {
    if (!tv_sizes_ok(n_matches, n_hyp) || n1 < 0 || n2 < 0 || !(sigma > 0.0f) || !best || !best_score)
        return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers)) return NAV_BA_BAD_ARG;
    if (n_hyp > 0 && (!H21 || !H12 || !scores)) return NAV_BA_BAD_ARG;
    if (!tv_matches_ok(matches, n_matches, n1, n2)) return NAV_BA_BAD_ARG;
    return tv_host<true>(kp1, n1, kp2, n2, matches, n_matches, H21, H12, n_hyp, tv_args(sigma, th_score, th_score),
                         scores, best, best_score, inliers, device);
}

extern "C" int32_t nav_tv_score_fundamental_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2,
                                                 const int32_t* matches, int32_t n_matches, const float* F21,
                                                 int32_t n_hyp, float sigma, float th_f, float th_score,
                                                 float* scores, int32_t* best, float* best_score, uint8_t* inliers,
                                                 int32_t device) {
    if (!tv_sizes_ok(n_matches, n_hyp) || n1 < 0 || n2 < 0 || !(sigma > 0.0f) || !best || !best_score)
        return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers)) return NAV_BA_BAD_ARG;
    if (n_hyp > 0 && (!F21 || !scores)) return NAV_BA_BAD_ARG;
    if (!tv_matches_ok(matches, n_matches, n1, n2)) return NAV_BA_BAD_ARG;
    return tv_host<false>(kp1, n1, kp2, n2, matches, n_matches, F21, nullptr, n_hyp, tv_args(sigma, th_f, th_score),
                          scores, best, best_score, inliers, device);
}

// ---- CheckRT entry points

static bool rt_args(const float* K, const float* R, const float* t, int32_t n_cand, float th2, float min_cos,
                    RtArgs& g) {
    if (!K || !R || !t || n_cand < 1 || n_cand > kTvMaxCand) return false;
    g = RtArgs{};
    g.fx = K[0];
    g.fy = K[4];
    g.cx = K[2];
    g.cy = K[5];
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 4; ++q) g.P1[4 * r + q] = q < 3 ? K[3 * r + q] : 0.0f;
    g.th2 = th2;
    g.min_cos = min_cos;
    for (int c = 0; c < n_cand; ++c) {
        RtCand& C = g.c[c];
        const float* Rc = R + 9 * c;
        const float* tc = t + 3 * c;
        std::memcpy(C.R, Rc, sizeof(C.R));
        std::memcpy(C.t, tc, sizeof(C.t));
        float Rt[12];
        for (int r = 0; r < 3; ++r) {
            for (int q = 0; q < 3; ++q) Rt[4 * r + q] = Rc[3 * r + q];
            Rt[4 * r + 3] = tc[r];
        }
        // P2 = K [R|t] and O2 = -R^T t, float products summed in double and rounded to float
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 4; ++q) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc = acc + (double)K[3 * r + k] * (double)Rt[4 * k + q];
                C.P2[4 * r + q] = (float)acc;
            }
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc = acc + (double)Rc[3 * k + r] * (double)tc[k];
            C.O2[r] = -(float)acc;
        }
    }
    return true;
}

extern "C" int32_t nav_tv_check_rt(const float* kp1, int32_t n_kp1, const float* kp2, const int32_t* matches,
                                   const uint8_t* inliers, int32_t n_matches, const float* K, const float* R,
                                   const float* t, int32_t n_cand, float th2, float min_parallax_cos,
                                   int32_t min_triangulated, int32_t* n_good, float* parallax, float* cos_sel,
                                   float* p3d, uint8_t* good, float* cos_parallax, uint8_t* passed, void* stream) {
    RtArgs g;
    if (n_matches < 0 || n_matches > NAV_TV_MAX_MATCHES || n_kp1 < 0 || min_triangulated < 0) return NAV_BA_BAD_ARG;
    if (!rt_args(K, R, t, n_cand, th2, min_parallax_cos, g)) return NAV_BA_BAD_ARG;
    if (!n_good || !parallax || !cos_sel) return NAV_BA_BAD_ARG;
    if (n_kp1 > 0 && (!p3d || !good)) return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers || !cos_parallax || !passed)) return NAV_BA_BAD_ARG;
    if ((size_t)n_kp1 * n_cand > (size_t)INT32_MAX) return NAV_BA_BAD_ARG;
    const hipStream_t st = (hipStream_t)stream;
    const int nk = n_kp1 * n_cand;
    if (nk > 0) hipLaunchKernelGGL(k_tv_rt_clear, dim3((nk + 255) / 256), dim3(256), 0, st, p3d, good, nk);
    if (n_matches > 0)
        hipLaunchKernelGGL(k_tv_check_rt, dim3((n_matches + 255) / 256, n_cand), dim3(256), 0, st,
                           reinterpret_cast<const float2*>(kp1), reinterpret_cast<const float2*>(kp2),
                           reinterpret_cast<const int2*>(matches), inliers, n_matches, n_kp1, g, p3d, good,
                           cos_parallax, passed);
    hipLaunchKernelGGL(k_tv_rt_select, dim3(n_cand), dim3(kTvSelThreads), 0, st, (const float*)cos_parallax,
                       (const uint8_t*)passed, n_matches, min_triangulated, n_good, parallax, cos_sel);
    return hipGetLastError() == hipSuccess ? NAV_BA_OK : NAV_BA_HIP_ERROR;
}

extern "C" int32_t nav_tv_check_rt_host(const float* kp1, int32_t n_kp1, const float* kp2, int32_t n_kp2,
                                        const int32_t* matches, const uint8_t* inliers, int32_t n_matches,
                                        const float* K, const float* R, const float* t, int32_t n_cand, float th2,
                                        float min_parallax_cos, int32_t min_triangulated, int32_t* n_good,
                                        float* parallax, float* cos_sel, float* p3d, uint8_t* good,
                                        int32_t device) {
    RtArgs g;
    if (n_matches < 0 || n_matches > NAV_TV_MAX_MATCHES || n_kp1 < 0 || n_kp2 < 0 || min_triangulated < 0)
        return NAV_BA_BAD_ARG;
    if (!rt_args(K, R, t, n_cand, th2, min_parallax_cos, g)) return NAV_BA_BAD_ARG;
    if (!n_good || !parallax || !cos_sel) return NAV_BA_BAD_ARG;
    if (n_kp1 > 0 && (!p3d || !good)) return NAV_BA_BAD_ARG;
    if (n_matches > 0 && (!kp1 || !kp2 || !matches || !inliers)) return NAV_BA_BAD_ARG;
    if (!tv_matches_ok(matches, n_matches, n_kp1, n_kp2)) return NAV_BA_BAD_ARG;
    if ((size_t)n_kp1 * n_cand > (size_t)INT32_MAX) return NAV_BA_BAD_ARG;
    if (!tv_gfx950(device)) return NAV_BA_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return NAV_BA_HIP_ERROR;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t nm = (size_t)n_matches, nk = (size_t)n_kp1, nc = (size_t)n_cand;
    const size_t b_k1 = al(8 * nk), b_k2 = al(8 * (size_t)n_kp2), b_mt = al(8 * nm), b_in = al(nm),
                 b_p3 = al(12 * nk * nc), b_gd = al(nk * nc), b_cp = al(4 * nm * nc), b_ps = al(nm * nc),
                 b_o = al(12 * nc);
    HostScratch& hs = host_scratch(device);
    std::lock_guard<std::mutex> lock(hs.mu);
    char* d = (char*)host_scratch_get(hs, b_k1 + b_k2 + b_mt + b_in + b_p3 + b_gd + b_cp + b_ps + b_o);
    if (!d) return NAV_BA_HIP_ERROR;
    char* p = d;
    float* d_k1 = (float*)p;
    p += b_k1;
    float* d_k2 = (float*)p;
    p += b_k2;
    int32_t* d_mt = (int32_t*)p;
    p += b_mt;
    uint8_t* d_in = (uint8_t*)p;
    p += b_in;
    float* d_p3 = (float*)p;
    p += b_p3;
    uint8_t* d_gd = (uint8_t*)p;
    p += b_gd;
    float* d_cp = (float*)p;
    p += b_cp;
    uint8_t* d_ps = (uint8_t*)p;
    p += b_ps;
    int32_t* d_ng = (int32_t*)p;
    float* d_par = (float*)(p + 4 * nc);
    float* d_cs = (float*)(p + 8 * nc);
    int32_t st = NAV_BA_OK;
    auto up = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && b > 0 && hipMemcpy(dst, src, b, hipMemcpyHostToDevice) != hipSuccess)
            st = NAV_BA_HIP_ERROR;
    };
    auto down = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && b > 0 && hipMemcpy(dst, src, b, hipMemcpyDeviceToHost) != hipSuccess)
            st = NAV_BA_HIP_ERROR;
    };
    up(d_k1, kp1, 8 * nk);
    up(d_k2, kp2, 8 * (size_t)n_kp2);
    up(d_mt, matches, 8 * nm);
    up(d_in, inliers, nm);
    if (st == NAV_BA_OK)
        st = nav_tv_check_rt(d_k1, n_kp1, d_k2, d_mt, d_in, n_matches, K, R, t, n_cand, th2, min_parallax_cos,
                             min_triangulated, d_ng, d_par, d_cs, d_p3, d_gd, d_cp, d_ps, nullptr);
    down(n_good, d_ng, 4 * nc);
    down(parallax, d_par, 4 * nc);
    down(cos_sel, d_cs, 4 * nc);
    down(p3d, d_p3, 12 * nk * nc);
    down(good, d_gd, nk * nc);
    return st;
}

// ---- hypothesis entry points

extern "C" int32_t nav_tv_normalize(const float* kp, int32_t n, float* out, float* T) {
    if (n < 1 || !kp || !out || !T) return NAV_BA_BAD_ARG;
    tv_normalize_host(kp, n, out, T);
    return NAV_BA_OK;
}

extern "C" int32_t nav_tv_hypotheses(const float* q1, const float* q2, const int32_t* matches,
                                     const int32_t* sets, int32_t n_sets, const float* T1, const float* T2,
                                     float* H21, float* H12, float* F21, void* stream) {
    if (n_sets < 0 || !T1 || !T2) return NAV_BA_BAD_ARG;
    if (n_sets == 0) return NAV_BA_OK;
    if (!q1 || !q2 || !matches || !sets || !H21 || !H12 || !F21) return NAV_BA_BAD_ARG;
    const HypArgs g = hyp_args(T1, T2);
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tv_hyp, dim3(2 * ((n_sets + 63) / 64)), dim3(64), 0, st, reinterpret_cast<const float2*>(q1),
                       reinterpret_cast<const float2*>(q2), reinterpret_cast<const int2*>(matches), sets, n_sets, g,
                       H21, H12, F21);
    return hipGetLastError() == hipSuccess ? NAV_BA_OK : NAV_BA_HIP_ERROR;
}

extern "C" int32_t nav_tv_hypotheses_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2,
                                          const int32_t* matches, int32_t n_matches, const int32_t* sets,
                                          int32_t n_sets, float* H21, float* H12, float* F21, int32_t device) {
    if (n1 < 1 || n2 < 1 || n_matches < 1 || n_sets < 0 || n_matches > NAV_TV_MAX_MATCHES) return NAV_BA_BAD_ARG;
    if (!kp1 || !kp2 || !matches || (n_sets > 0 && (!sets || !H21 || !H12 || !F21))) return NAV_BA_BAD_ARG;
    if (!tv_matches_ok(matches, n_matches, n1, n2)) return NAV_BA_BAD_ARG;
    for (int64_t i = 0; i < 8 * (int64_t)n_sets; ++i)
        if (sets[i] < 0 || sets[i] >= n_matches) return NAV_BA_BAD_ARG;
    if (!tv_gfx950(device)) return NAV_BA_NO_DEVICE;
    if (n_sets == 0) return NAV_BA_OK;
    if (hipSetDevice(device) != hipSuccess) return NAV_BA_HIP_ERROR;
    std::vector<float> q1(2 * (size_t)n1), q2(2 * (size_t)n2);
    float T1[9], T2[9];
    tv_normalize_host(kp1, n1, q1.data(), T1);
    tv_normalize_host(kp2, n2, q2.data(), T2);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t ns = (size_t)n_sets;
    const size_t b1 = al(8 * (size_t)n1), b2 = al(8 * (size_t)n2), bm = al(8 * (size_t)n_matches), bs = al(32 * ns),
                 bo = al(36 * ns);
    HostScratch& hs = host_scratch(device);
    std::lock_guard<std::mutex> lock(hs.mu);
    char* d = (char*)host_scratch_get(hs, b1 + b2 + bm + bs + 3 * bo);
    if (!d) return NAV_BA_HIP_ERROR;
    float* d_q1 = (float*)d;
    float* d_q2 = (float*)(d + b1);
    int32_t* d_m = (int32_t*)(d + b1 + b2);
    int32_t* d_s = (int32_t*)(d + b1 + b2 + bm);
    float* d_h21 = (float*)(d + b1 + b2 + bm + bs);
    float* d_h12 = (float*)(d + b1 + b2 + bm + bs + bo);
    float* d_f21 = (float*)(d + b1 + b2 + bm + bs + 2 * bo);
    int32_t st = NAV_BA_OK;
    auto up = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && hipMemcpy(dst, src, b, hipMemcpyHostToDevice) != hipSuccess) st = NAV_BA_HIP_ERROR;
    };
    auto down = [&](void* dst, const void* src, size_t b) {
        if (st == NAV_BA_OK && hipMemcpy(dst, src, b, hipMemcpyDeviceToHost) != hipSuccess) st = NAV_BA_HIP_ERROR;
    };
    up(d_q1, q1.data(), 8 * (size_t)n1);
    up(d_q2, q2.data(), 8 * (size_t)n2);
    up(d_m, matches, 8 * (size_t)n_matches);
    up(d_s, sets, 32 * ns);
    if (st == NAV_BA_OK) st = nav_tv_hypotheses(d_q1, d_q2, d_m, d_s, n_sets, T1, T2, d_h21, d_h12, d_f21, nullptr);
    down(H21, d_h21, 36 * ns);
    down(H12, d_h12, 36 * ns);
    down(F21, d_f21, 36 * ns);
    return st;
}
