// cost_kernels.hip — differentiable robust reprojection cost (include/nav_ba_cost.h).
//
// Hand-written adjoint of C = sum rho(|z - pi_d(T X)|^2) for a batch of
// problems sharing one pinhole + radial-tangential camera.  The observation
// structure is uploaded once (point-major order, tiles of whole points that
// never span two problems, camera-major index lists); each evaluation runs
//   k_cg_obs    one lane per observation: cost term and its adjoint; the
//               point gradient is summed per point from LDS in fixed order,
//               cost and intrinsic gradients per tile, and the pose gradient
//               term of each observation goes to a buffer;
//   k_cg_cam    one workgroup per camera: its observations' pose terms in
//               camera-major order (fixed order);
//   k_cg_final  per-problem costs and the intrinsic gradient over the tiles.
// The forward (cost only) skips the adjoint and the second kernel.
// T = double or float: the projection and residual are FP64 either way, the
// adjoint runs in T; tile partials and the final sums are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nav_ba_cost.h"
#include "ba_device.hpp"

namespace navcost {

using navba::block_sum;
using navba::kWave;
using navba::wave_sum;

constexpr int kTileObs = 256;  // observations per tile (whole points; one larger point has a tile of its own)
constexpr int kThreads = 256;

struct CostProblem {
    int n_obs, n_tiles, n_cam, n_pt, n_prob;
    int robust;
    double delta;
    const int* o_cam;      // point-major, caller (global) camera index
    const int* o_pt;       // point-major, caller (global) point index
    const double* uv_d;    // 2 per observation (FP64 for both compute types: the residual is FP64)
    const int* pt_list;    // unique points with observations, ascending
    const int* pt_ptr;     // [n_pts_used + 1]
    const int* tile_obs;   // [n_tiles + 1]
    const int* tile_pt;    // [n_tiles + 1] range in pt_list
    const int* tile_prob;  // [n_tiles]
    const int* prob_tile0; // [n_prob + 1]
    const int* cam_ptr;    // [n_cam + 1] camera-major lists
    const int* cam_obs;    // positions in point-major order
};

template <typename T>
struct ObsGrad {
    double cost;  // the robust term, from the FP64 residual
    T gX[3];
    T gpose[6];
    T gintr[8];
};

template <typename T>
__device__ __forceinline__ T sqrt_t(T v) { return sqrt(v); }

// Cost term of one observation and, if kGrad, its adjoint.  The projection
// and residual are formed in FP64 whatever T is: the residual is a difference
// of two pixel coordinates of a few hundred pixels, and in FP32 that
// cancellation alone left the cost of an FP32 problem 6.6e-4 off its FP64
// value (FP64 issues at the FP32 rate on gfx950 without packed math, so the
// forward costs the same).  The adjoint then runs in T.
template <typename T, bool kGrad>
__device__ __forceinline__ void obs_term(const T* __restrict__ Tcw, const T* __restrict__ X, const T in[8], double u,
                                         double v, int robust, double delta, ObsGrad<T>& g) {
    const double R0 = Tcw[0], R1 = Tcw[1], R2 = Tcw[2], R3 = Tcw[3], R4 = Tcw[4], R5 = Tcw[5];
    const double R6 = Tcw[6], R7 = Tcw[7], R8 = Tcw[8];
    const double x0 = X[0], x1 = X[1], x2 = X[2];
    const double Xc0 = R0 * x0 + R1 * x1 + R2 * x2 + (double)Tcw[9];
    const double Xc1 = R3 * x0 + R4 * x1 + R5 * x2 + (double)Tcw[10];
    const double Xc2 = R6 * x0 + R7 * x1 + R8 * x2 + (double)Tcw[11];
    const double iz = 1.0 / Xc2;
    const double x = Xc0 * iz, y = Xc1 * iz;
    const double fx = in[0], fy = in[1], cx = in[2], cy = in[3], k1 = in[4], k2 = in[5], p1 = in[6], p2 = in[7];
    const double r2 = x * x + y * y;
    const double rad = 1.0 + k1 * r2 + k2 * r2 * r2;
    const double xy = x * y;
    const double xd = x * rad + 2.0 * p1 * xy + p2 * (r2 + 2.0 * x * x);
    const double yd = y * rad + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * xy;
    const double e0 = u - (fx * xd + cx);
    const double e1 = v - (fy * yd + cy);
    const double s = e0 * e0 + e1 * e1;
    double w;
    const double d2 = delta * delta;
    if (!robust || s <= d2) {
        g.cost = s;
        w = 1.0;
    } else {
        const double r = sqrt(s);
        g.cost = 2.0 * delta * r - d2;
        w = delta / r;
    }
    if (!kGrad) return;
    // the adjoint in T from the FP64 forward values
    const T tx = (T)x, ty = (T)y, tr2 = (T)r2, trad = (T)rad, txy = (T)xy, tiz = (T)iz;
    const T tk1 = in[4], tk2 = in[5], tp1 = in[6], tp2 = in[7], tfx = in[0], tfy = in[1];
    const T gpu = (T)(-2.0 * w * e0), gpv = (T)(-2.0 * w * e1);  // d C / d (projected u, v)
    g.gintr[0] = gpu * (T)xd;
    g.gintr[1] = gpv * (T)yd;
    g.gintr[2] = gpu;
    g.gintr[3] = gpv;
    const T gxd = gpu * tfx, gyd = gpv * tfy;
    g.gintr[4] = (gxd * tx + gyd * ty) * tr2;
    g.gintr[5] = (gxd * tx + gyd * ty) * tr2 * tr2;
    g.gintr[6] = gxd * T(2) * txy + gyd * (tr2 + T(2) * ty * ty);
    g.gintr[7] = gxd * (tr2 + T(2) * tx * tx) + gyd * T(2) * txy;
    const T kr = tk1 + T(2) * tk2 * tr2;
    const T dxdx = trad + T(2) * tx * tx * kr + T(2) * tp1 * ty + T(6) * tp2 * tx;
    const T dxdy = T(2) * txy * kr + T(2) * tp1 * tx + T(2) * tp2 * ty;
    const T dydx = dxdy;
    const T dydy = trad + T(2) * ty * ty * kr + T(6) * tp1 * ty + T(2) * tp2 * tx;
    const T gx = gxd * dxdx + gyd * dydx;
    const T gy = gxd * dxdy + gyd * dydy;
    const T gc0 = gx * tiz, gc1 = gy * tiz, gc2 = -(gx * tx + gy * ty) * tiz;  // d C / d Xc
    // world point: R^T gc
    g.gX[0] = Tcw[0] * gc0 + Tcw[3] * gc1 + Tcw[6] * gc2;
    g.gX[1] = Tcw[1] * gc0 + Tcw[4] * gc1 + Tcw[7] * gc2;
    g.gX[2] = Tcw[2] * gc0 + Tcw[5] * gc1 + Tcw[8] * gc2;
    // pose, left perturbation: d Xc / d omega = -[Xc]x, d Xc / d upsilon = I
    const T tc0 = (T)Xc0, tc1 = (T)Xc1, tc2 = (T)Xc2;
    g.gpose[0] = tc1 * gc2 - tc2 * gc1;
    g.gpose[1] = tc2 * gc0 - tc0 * gc2;
    g.gpose[2] = tc0 * gc1 - tc1 * gc0;
    g.gpose[3] = gc0;
    g.gpose[4] = gc1;
    g.gpose[5] = gc2;
}

// One workgroup per tile.  part[t] = {cost, d/d intrinsics (8)} of the tile (FP64).
template <typename T, bool kGrad>
__global__ __launch_bounds__(kThreads) void k_cg_obs(CostProblem P, const T* __restrict__ Tcw,
                                                     const T* __restrict__ X, const T* __restrict__ intr,
                                                     T* __restrict__ gpo, T* __restrict__ gX,
                                                     double* __restrict__ part) {
    __shared__ T sh[3][kTileObs];
    __shared__ double red[kThreads / kWave];
    __shared__ T redp[kThreads / kWave][3];
    const int t = blockIdx.x;
    const int o0 = P.tile_obs[t], o1 = P.tile_obs[t + 1];
    const int q0 = P.tile_pt[t], q1 = P.tile_pt[t + 1];
    T in[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) in[k] = intr[k];
    double cost = 0.0, gi[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    T gp[3] = {T(0), T(0), T(0)};
    for (int a = o0 + (int)threadIdx.x; a < o1; a += kThreads) {
        const int c = P.o_cam[a], q = P.o_pt[a];
        const double* uv = P.uv_d + 2 * (size_t)a;
        ObsGrad<T> g;
        obs_term<T, kGrad>(Tcw + 12 * (size_t)c, X + 3 * (size_t)q, in, uv[0], uv[1], P.robust, P.delta, g);
        cost += g.cost;
        if (kGrad) {
#pragma unroll
            for (int k = 0; k < 8; ++k) gi[k] += (double)g.gintr[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) gpo[6 * (size_t)a + k] = g.gpose[k];
            gp[0] += g.gX[0]; gp[1] += g.gX[1]; gp[2] += g.gX[2];
        }
    }
    if (kGrad && gX) {  // uniform: gX is a kernel argument
        if (o1 - o0 <= kTileObs) {
            // one observation per lane: per-point sums in fixed order from LDS
            sh[0][threadIdx.x] = gp[0]; sh[1][threadIdx.x] = gp[1]; sh[2][threadIdx.x] = gp[2];
            __syncthreads();
            for (int i = q0 + (int)threadIdx.x; i < q1; i += kThreads) {
                T s0 = T(0), s1 = T(0), s2 = T(0);
                for (int a = P.pt_ptr[i]; a < P.pt_ptr[i + 1]; ++a) {
                    s0 += sh[0][a - o0]; s1 += sh[1][a - o0]; s2 += sh[2][a - o0];
                }
                const int q = P.pt_list[i];
                gX[3 * (size_t)q] = s0; gX[3 * (size_t)q + 1] = s1; gX[3 * (size_t)q + 2] = s2;
            }
        } else {
            // one point with more than kTileObs observations
            const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T v = wave_sum(gp[k]);
                if (lane == 0) redp[wid][k] = v;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                T s = T(0);
                for (int i = 0; i < kThreads / kWave; ++i) s += redp[i][threadIdx.x];
                gX[3 * (size_t)P.pt_list[q0] + threadIdx.x] = s;
            }
        }
    }
    const double cs = block_sum<double, kThreads>(cost, red);
    if (threadIdx.x == 0) part[9 * (size_t)t] = cs;
    if (kGrad) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            __syncthreads();
            const double v = block_sum<double, kThreads>(gi[k], red);
            if (threadIdx.x == 0) part[9 * (size_t)t + 1 + k] = v;
        }
    }
}

// One workgroup per camera: d C / d xi in fixed order over its observations.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_cam(CostProblem P, const T* __restrict__ gpo, T* __restrict__ gpose) {
    __shared__ double red[kThreads / kWave];
    const int c = blockIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = P.cam_ptr[c] + (int)threadIdx.x; k < P.cam_ptr[c + 1]; k += kThreads) {
        const int a = P.cam_obs[k];
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[j] += (double)gpo[6 * (size_t)a + j];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        __syncthreads();
        const double v = block_sum<double, kThreads>(acc[j], red);
        if (threadIdx.x == 0) gpose[6 * (size_t)c + j] = (T)v;
    }
}

// Workgroup p < n_prob: cost of problem p over its tiles; workgroup n_prob:
// the intrinsic gradient over all tiles (fixed order).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_final(CostProblem P, const double* __restrict__ part, int grad,
                                                       T* __restrict__ cost, T* __restrict__ gintr) {
    __shared__ double red[kThreads / kWave];
    const int pr = blockIdx.x;
    if (pr < P.n_prob) {
        double s = 0.0;
        for (int t = P.prob_tile0[pr] + (int)threadIdx.x; t < P.prob_tile0[pr + 1]; t += kThreads)
            s += part[9 * (size_t)t];
        const double v = block_sum<double, kThreads>(s, red);
        if (threadIdx.x == 0) cost[pr] = (T)v;
        return;
    }
    if (!grad || !gintr) return;
    for (int k = 0; k < 8; ++k) {
        double s = 0.0;
        for (int t = threadIdx.x; t < P.n_tiles; t += kThreads) s += part[9 * (size_t)t + 1 + k];
        const double v = block_sum<double, kThreads>(s, red);
        if (threadIdx.x == 0) gintr[k] = (T)v;
    }
}

template <typename T>
__global__ void k_zero(T* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = T(0);
}

}  // namespace navcost

using namespace navcost;

struct nav_ba_cost_handle {
    int device = 0;
    std::string err;
    bool prepared = false;
    CostProblem P{};
    std::vector<void*> bufs;
    void* gpo = nullptr;
    size_t gpo_bytes = 0;
    double* part = nullptr;
};

namespace {

void free_all(nav_ba_cost_handle* h) {
    for (void* p : h->bufs) (void)hipFree(p);
    h->bufs.clear();
    if (h->gpo) (void)hipFree(h->gpo);
    h->gpo = nullptr;
    h->gpo_bytes = 0;
    h->part = nullptr;
}

template <class V>
int up(nav_ba_cost_handle* h, const V& v, const typename V::value_type** out) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(sizeof(typename V::value_type) * v.size(), 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) { h->err = std::string("hipMalloc: ") + hipGetErrorString(e); return NAV_BA_HIP_ERROR; }
    h->bufs.push_back(p);
    if (!v.empty() && hipMemcpy(p, v.data(), sizeof(typename V::value_type) * v.size(), hipMemcpyHostToDevice) != hipSuccess) {
        h->err = "hipMemcpy failed";
        return NAV_BA_HIP_ERROR;
    }
    *out = (const typename V::value_type*)p;
    return NAV_BA_OK;
}

template <typename T>
int32_t eval_t(nav_ba_cost_handle* h, const T* Tcw, const T* X, const T* intr, T* cost, T* gpose, T* gX,
                      T* gintr, hipStream_t s) {
    const CostProblem& P = h->P;
    const bool grad = gpose || gX || gintr;
    T* gpo = (T*)h->gpo;
    if (grad && gX) hipLaunchKernelGGL(k_zero<T>, dim3(256), dim3(256), 0, s, gX, (size_t)3 * P.n_pt);
    if (P.n_tiles > 0) {
        if (grad)  // gX may be null: the per-point sums are then skipped
            hipLaunchKernelGGL((k_cg_obs<T, true>), dim3(P.n_tiles), dim3(kThreads), 0, s, P, Tcw, X, intr, gpo, gX,
                               h->part);
        else
            hipLaunchKernelGGL((k_cg_obs<T, false>), dim3(P.n_tiles), dim3(kThreads), 0, s, P, Tcw, X, intr,
                               (T*)nullptr, (T*)nullptr, h->part);
    }
    if (grad && gpose && P.n_cam > 0)
        hipLaunchKernelGGL(k_cg_cam<T>, dim3(P.n_cam), dim3(kThreads), 0, s, P, (const T*)gpo, gpose);
    hipLaunchKernelGGL(k_cg_final<T>, dim3(P.n_prob + 1), dim3(kThreads), 0, s, P, (const double*)h->part,
                       grad ? 1 : 0, cost, gintr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { h->err = std::string("launch: ") + hipGetErrorString(e); return NAV_BA_HIP_ERROR; }
    return NAV_BA_OK;
}

}  // namespace

extern "C" {

int32_t nav_ba_cost_create(nav_ba_cost_handle** out, int32_t device) {
    if (!out) return NAV_BA_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NAV_BA_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return NAV_BA_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return NAV_BA_NO_DEVICE;
    nav_ba_cost_handle* h = new nav_ba_cost_handle();
    h->device = device;
    *out = h;
    return NAV_BA_OK;
}

void nav_ba_cost_destroy(nav_ba_cost_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
}

const char* nav_ba_cost_last_error(const nav_ba_cost_handle* h) { return h ? h->err.c_str() : "null handle"; }

int32_t nav_ba_cost_prepare(nav_ba_cost_handle* h, const nav_ba_problem* probs, int32_t n_probs, int32_t robust,
                            double huber_delta) {
    if (!h) return NAV_BA_BAD_ARG;
    h->err.clear();
    h->prepared = false;
    if (!probs || n_probs <= 0) { h->err = "no problems"; return NAV_BA_BAD_ARG; }
    if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice"; return NAV_BA_HIP_ERROR; }
    (void)hipDeviceSynchronize();
    free_all(h);
    std::vector<int> cam_off(n_probs + 1, 0), pt_off(n_probs + 1, 0);
    size_t no = 0;
    for (int p = 0; p < n_probs; ++p) {
        const nav_ba_problem& q = probs[p];
        if (q.n_cam < 0 || q.n_pt < 0 || q.n_obs < 0 || (q.n_obs > 0 && (!q.obs_cam || !q.obs_pt || !q.obs_uv))) {
            h->err = "invalid problem " + std::to_string(p);
            return NAV_BA_BAD_ARG;
        }
        for (int i = 0; i < q.n_obs; ++i)
            if (q.obs_cam[i] < 0 || q.obs_cam[i] >= q.n_cam || q.obs_pt[i] < 0 || q.obs_pt[i] >= q.n_pt) {
                h->err = "observation index out of range in problem " + std::to_string(p);
                return NAV_BA_BAD_ARG;
            }
        cam_off[p + 1] = cam_off[p] + q.n_cam;
        pt_off[p + 1] = pt_off[p] + q.n_pt;
        no += (size_t)q.n_obs;
    }
    if (no > (size_t)0x7fffffff) { h->err = "too many observations"; return NAV_BA_TOO_LARGE; }
    const int nc = cam_off[n_probs], np = pt_off[n_probs];
    std::vector<int> gcam(no), gpt(no);
    std::vector<double> guv(2 * no);
    {
        size_t k = 0;
        for (int p = 0; p < n_probs; ++p)
            for (int i = 0; i < probs[p].n_obs; ++i, ++k) {
                gcam[k] = cam_off[p] + probs[p].obs_cam[i];
                gpt[k] = pt_off[p] + probs[p].obs_pt[i];
                guv[2 * k] = probs[p].obs_uv[2 * (size_t)i];
                guv[2 * k + 1] = probs[p].obs_uv[2 * (size_t)i + 1];
            }
    }
    // point-major order (point, camera, input order): two stable counting sorts
    std::vector<int> byc(no), ord(no);
    {
        std::vector<int> off(nc + 1, 0);
        for (size_t i = 0; i < no; ++i) off[gcam[i] + 1]++;
        for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
        for (size_t i = 0; i < no; ++i) byc[off[gcam[i]]++] = (int)i;
        std::vector<int> poff(np + 1, 0);
        for (size_t i = 0; i < no; ++i) poff[gpt[i] + 1]++;
        for (int q = 0; q < np; ++q) poff[q + 1] += poff[q];
        for (size_t k = 0; k < no; ++k) { const int i = byc[k]; ord[poff[gpt[i]]++] = i; }
    }
    std::vector<int> o_cam(no), o_pt(no);
    std::vector<double> uv_d(2 * no);
    for (size_t k = 0; k < no; ++k) {
        const int i = ord[k];
        o_cam[k] = gcam[i]; o_pt[k] = gpt[i];
        uv_d[2 * k] = guv[2 * (size_t)i]; uv_d[2 * k + 1] = guv[2 * (size_t)i + 1];
    }
    // unique points, tiles of whole points per problem
    std::vector<int> pt_list, pt_ptr{0}, tile_obs{0}, tile_pt{0}, tile_prob, prob_tile0{0};
    {
        size_t k = 0;
        for (int p = 0; p < n_probs; ++p) {  // a problem's points, hence observations, are contiguous
            while (k < no && o_pt[k] < pt_off[p + 1]) {
                const int q = o_pt[k];
                size_t e = k;
                while (e < no && o_pt[e] == q) ++e;
                pt_list.push_back(q);
                pt_ptr.push_back((int)e);
                k = e;
            }
        }
        int i = 0;
        const int nu = (int)pt_list.size();
        for (int p = 0; p < n_probs; ++p) {
            while (i < nu && pt_list[i] < pt_off[p + 1]) {
                int j = i;
                while (j < nu && pt_list[j] < pt_off[p + 1] && pt_ptr[j + 1] - pt_ptr[i] <= kTileObs) ++j;
                if (j == i) j = i + 1;
                tile_obs.push_back(pt_ptr[j]);
                tile_pt.push_back(j);
                tile_prob.push_back(p);
                i = j;
            }
            prob_tile0.push_back((int)tile_prob.size());
        }
    }
    std::vector<int> cam_ptr(nc + 1, 0), cam_obs(no);
    for (size_t k = 0; k < no; ++k) cam_ptr[o_cam[k] + 1]++;
    for (int c = 0; c < nc; ++c) cam_ptr[c + 1] += cam_ptr[c];
    {
        std::vector<int> pos(cam_ptr.begin(), cam_ptr.end() - 1);
        for (size_t k = 0; k < no; ++k) cam_obs[pos[o_cam[k]]++] = (int)k;
    }
    CostProblem& P = h->P;
    P = CostProblem{};
    P.n_obs = (int)no; P.n_tiles = (int)tile_prob.size(); P.n_cam = nc; P.n_pt = np; P.n_prob = n_probs;
    P.robust = robust ? 1 : 0;
    P.delta = huber_delta;
    int st;
    if ((st = up(h, o_cam, &P.o_cam)) || (st = up(h, o_pt, &P.o_pt)) || (st = up(h, uv_d, &P.uv_d)) ||
        (st = up(h, pt_list, &P.pt_list)) || (st = up(h, pt_ptr, &P.pt_ptr)) ||
        (st = up(h, tile_obs, &P.tile_obs)) || (st = up(h, tile_pt, &P.tile_pt)) ||
        (st = up(h, tile_prob, &P.tile_prob)) || (st = up(h, prob_tile0, &P.prob_tile0)) ||
        (st = up(h, cam_ptr, &P.cam_ptr)) || (st = up(h, cam_obs, &P.cam_obs)))
        return st;
    void* part = nullptr;
    if (hipMalloc(&part, sizeof(double) * 9 * std::max<size_t>(tile_prob.size(), 1)) != hipSuccess) {
        h->err = "hipMalloc part";
        return NAV_BA_HIP_ERROR;
    }
    h->bufs.push_back(part);
    h->part = (double*)part;
    h->gpo_bytes = sizeof(double) * 6 * std::max<size_t>(no, 1);
    if (hipMalloc(&h->gpo, h->gpo_bytes) != hipSuccess) { h->err = "hipMalloc gpo"; return NAV_BA_HIP_ERROR; }
    h->prepared = true;
    return NAV_BA_OK;
}

int32_t nav_ba_cost_sizes(const nav_ba_cost_handle* h, int32_t* n_cam, int32_t* n_pt, int32_t* n_obs,
                          int32_t* n_problems) {
    if (!h || !h->prepared) return NAV_BA_NOT_PREPARED;
    if (n_cam) *n_cam = h->P.n_cam;
    if (n_pt) *n_pt = h->P.n_pt;
    if (n_obs) *n_obs = h->P.n_obs;
    if (n_problems) *n_problems = h->P.n_prob;
    return NAV_BA_OK;
}

int32_t nav_ba_cost_eval(nav_ba_cost_handle* h, int32_t dtype, const void* T_cw, const void* X, const void* intr,
                         void* cost, void* g_pose, void* g_X, void* g_intr, void* stream) {
    if (!h) return NAV_BA_BAD_ARG;
    if (!h->prepared) { h->err = "evaluation before nav_ba_cost_prepare"; return NAV_BA_NOT_PREPARED; }
    if (!T_cw || !X || !intr || !cost) { h->err = "null state or cost pointer"; return NAV_BA_BAD_ARG; }
    if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice"; return NAV_BA_HIP_ERROR; }
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NAV_BA_COST_F64)
        return eval_t<double>(h, (const double*)T_cw, (const double*)X, (const double*)intr, (double*)cost,
                              (double*)g_pose, (double*)g_X, (double*)g_intr, s);
    if (dtype == NAV_BA_COST_F32)
        return eval_t<float>(h, (const float*)T_cw, (const float*)X, (const float*)intr, (float*)cost,
                             (float*)g_pose, (float*)g_X, (float*)g_intr, s);
    h->err = "unknown dtype";
    return NAV_BA_BAD_ARG;
}

}  // extern "C"
