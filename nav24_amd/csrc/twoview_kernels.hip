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
#include <cstdint>
#include <cstring>

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
    char* d = nullptr;
    if (hipMalloc(&d, total) != hipSuccess) return NAV_BA_HIP_ERROR;
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
    (void)hipFree(d);
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
