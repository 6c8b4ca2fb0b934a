// undistort_kernels.hip — key-point undistortion on the device
// (include/nav_kp_undistort.h):
//   radial-tangential pinhole: PinholeRadTan::UndistortKeyPoints
//     (core/sensor/camera/models/PinholeRadTan.cpp:11-24), i.e. OpenCV's
//     cv::undistortPoints with K, D = (k1 k2 p1 p2), the camera's R and P
//     and its default termination (5 fixed-point iterations);
//   Kannala-Brandt fisheye: KannalaBrandt8::UndistortKeyPoints
//     (KannalaBrandt8.cpp:231-242), i.e. cv::fisheye::undistortPoints with
//     D = (k1 k2 k3 k4), R and P (up to 10 Newton steps on theta).
// R is the camera's rectification (identity unless the calibration sets one)
// and P its projection (K unless set): GeometricCamera.h:47-48,65-66,108-109,
// set from the configuration at Calibration.cpp:63-93.  Both OpenCV calls form
// RR = P(:, 0:3) R in double; here on the host, once per call.
//
// One lane per key point, the double operations of that algorithm in its
// order, without fused multiply-adds (contraction is off in this file), so
// the result equals a plain double evaluation bit for bit for finite input.
// 8 bytes read and 8 written per point; about 150 FP64 instructions per
// point (5 iterations, each with one division), so the kernel is bound by
// FP64 issue rather than HBM at 3 M points (~13 µs against ~6 µs of bytes).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>

#include "host_scratch.h"
#include "../../include/nav_kp_undistort.h"

#pragma clang fp contract(off)

namespace navba {

struct KpParams {
    double cx, cy, ifx, ify;  // ifx = 1 / fx in double, as the reference's call forms it
    double k1, k2, p1, p2;
    double RR[9];             // P R, row-major
};

constexpr int kKpIters = 5;  // cv::undistortPoints default: TermCriteria(COUNT, 5, 0.01)

__global__ __launch_bounds__(256) void k_kp_undistort(const float2* __restrict__ uv, int n, KpParams q,
                                                      float2* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 p = uv[i];
    const double u = p.x, v = p.y;
    double x = (u - q.cx) * q.ifx;
    double y = (v - q.cy) * q.ify;
    const double x0 = x, y0 = y;
    for (int j = 0; j < kKpIters; ++j) {
        const double r2 = x * x + y * y;
        const double icdist = 1.0 / (1.0 + (q.k2 * r2 + q.k1) * r2);
        if (icdist < 0.0) {  // no undistortion possible: the normalised input point
            x = (u - q.cx) * q.ifx;
            y = (v - q.cy) * q.ify;
            break;
        }
        const double dx = 2.0 * q.p1 * x * y + q.p2 * (r2 + 2.0 * x * x);
        const double dy = q.p1 * (r2 + 2.0 * y * y) + 2.0 * q.p2 * x * y;
        x = (x0 - dx) * icdist;
        y = (y0 - dy) * icdist;
    }
    const double xx = q.RR[0] * x + q.RR[1] * y + q.RR[2];
    const double yy = q.RR[3] * x + q.RR[4] * y + q.RR[5];
    const double ww = 1.0 / (q.RR[6] * x + q.RR[7] * y + q.RR[8]);
    out[i] = make_float2((float)(xx * ww), (float)(yy * ww));
}

// Fisheye (Kannala-Brandt): OpenCV 4.2's fisheye::undistortPoints steps.
struct KbParams {
    double fx, fy, cx, cy;
    double k[4];
    double RR[9];  // P R, row-major
};

constexpr int kKbIters = 10;         // fisheye::undistortPoints: at most 10 Newton steps
constexpr double kKbEps = 1e-8;      // ... stopping once a step is below this
constexpr double kHalfPi = 1.5707963267948966;  // CV_PI / 2

__global__ __launch_bounds__(256) void k_kp_undistort_fisheye(const float2* __restrict__ uv, int n, KbParams q,
                                                              float2* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 p = uv[i];
    const double pw0 = ((double)p.x - q.cx) / q.fx, pw1 = ((double)p.y - q.cy) / q.fy;
    double scale = 1.0;
    double theta_d = sqrt(pw0 * pw0 + pw1 * pw1);
    theta_d = fmin(fmax(-kHalfPi, theta_d), kHalfPi);
    if (theta_d > kKbEps) {
        double theta = theta_d;
        for (int j = 0; j < kKbIters; ++j) {
            const double theta2 = theta * theta, theta4 = theta2 * theta2, theta6 = theta4 * theta2,
                         theta8 = theta6 * theta2;
            const double k0t2 = q.k[0] * theta2, k1t4 = q.k[1] * theta4, k2t6 = q.k[2] * theta6,
                         k3t8 = q.k[3] * theta8;
            const double fix = (theta * (1 + k0t2 + k1t4 + k2t6 + k3t8) - theta_d) /
                               (1 + 3 * k0t2 + 5 * k1t4 + 7 * k2t6 + 9 * k3t8);
            theta = theta - fix;
            if (fabs(fix) < kKbEps) break;
        }
        scale = tan(theta) / theta_d;
    }
    const double pu0 = pw0 * scale, pu1 = pw1 * scale;
    // RR (pu, 1) as a 3x3 times 3-vector product: ((0 + a0 b0) + a1 b1) + a2 b2
    const double pr0 = ((0.0 + q.RR[0] * pu0) + q.RR[1] * pu1) + q.RR[2] * 1.0;
    const double pr1 = ((0.0 + q.RR[3] * pu0) + q.RR[4] * pu1) + q.RR[5] * 1.0;
    const double pr2 = ((0.0 + q.RR[6] * pu0) + q.RR[7] * pu1) + q.RR[8] * 1.0;
    out[i] = make_float2((float)(pr0 / pr2), (float)(pr1 / pr2));
}

// RR = P3 R in double, P3 the left 3x3 of P (3 or 4 columns, row-major), K
// when P is null, R the identity when null; each entry a sum of three products
// in column order, without contraction.
static void rect_rr(double fx, double fy, double cx, double cy, const double* R, const double* P, int p_cols,
                    double RR[9]) {
    double P3[9] = {fx, 0.0, cx, 0.0, fy, cy, 0.0, 0.0, 1.0};
    if (P)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P3[3 * i + j] = P[p_cols * i + j];
    double R3[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    if (R) std::memcpy(R3, R, sizeof(R3));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            RR[3 * i + j] = (P3[3 * i] * R3[j] + P3[3 * i + 1] * R3[3 + j]) + P3[3 * i + 2] * R3[6 + j];
}

static KbParams kb_params(const float* prm, const double* R, const double* P, int p_cols) {
    KbParams q;
    q.fx = prm[0];
    q.fy = prm[1];
    q.cx = prm[2];
    q.cy = prm[3];
    for (int c = 0; c < 4; ++c) q.k[c] = prm[4 + c];
    rect_rr(q.fx, q.fy, q.cx, q.cy, R, P, p_cols, q.RR);
    return q;
}

static KpParams kp_params(const float* prm, const double* R, const double* P, int p_cols) {
    KpParams q;
    const double fx = prm[0], fy = prm[1], cx = prm[2], cy = prm[3];
    q.cx = cx;
    q.cy = cy;
    q.ifx = 1.0 / fx;
    q.ify = 1.0 / fy;
    q.k1 = prm[4];
    q.k2 = prm[5];
    q.p1 = prm[6];
    q.p2 = prm[7];
    rect_rr(fx, fy, cx, cy, R, P, p_cols, q.RR);
    return q;
}

static bool gfx950_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Staging of the allocation-free host entry: a file-scope device array of at
// most kNoAllocMaxPoints key points; every device that runs the library holds
// its own copy (at its own address) until the process ends.  Host code never
// passes the array's name to the device: each device's address is taken once,
// with that device current, by hipGetSymbolAddress and kept in a fixed slot
// table (no allocation on any call), guarded by the device's lock.
constexpr int kNoAllocMaxPoints = NAV_KP_NOALLOC_MAX_POINTS;  // 16 MiB of float2
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) float2 scratch[kNoAllocMaxPoints];

struct NoAllocSlot {
    std::mutex mu;
    float2* ptr = nullptr;  // this device's address of `scratch`
};
static NoAllocSlot g_noalloc[kNoAllocMaxDevices];

static bool rect_args(int32_t model, int32_t n, const float* uv, const float* params, const double* P,
                      int32_t p_cols, float* uv_out) {
    if (model != NAV_KP_RADTAN && model != NAV_KP_KB8) return false;
    if (n < 0 || !params || (n > 0 && (!uv || !uv_out))) return false;
    return !P || p_cols == 3 || p_cols == 4;
}

static int32_t launch_undistort(int32_t model, const float* uv, int32_t n, const float* params, const double* R,
                                const double* P, int32_t p_cols, float* uv_out, hipStream_t stream) {
    if (model == NAV_KP_RADTAN)
        hipLaunchKernelGGL(k_kp_undistort, dim3((n + 255) / 256), dim3(256), 0, stream,
                           reinterpret_cast<const float2*>(uv), n, kp_params(params, R, P, p_cols),
                           reinterpret_cast<float2*>(uv_out));
    else
        hipLaunchKernelGGL(k_kp_undistort_fisheye, dim3((n + 255) / 256), dim3(256), 0, stream,
                           reinterpret_cast<const float2*>(uv), n, kb_params(params, R, P, p_cols),
                           reinterpret_cast<float2*>(uv_out));
    return hipGetLastError() == hipSuccess ? NAV_BA_OK : NAV_BA_HIP_ERROR;
}

}  // namespace navba

using namespace navba;

extern "C" int32_t nav_ba_kp_undistort_rect(int32_t model, const float* uv, int32_t n, const float* params,
                                            const double* R, const double* P, int32_t p_cols, float* uv_out,
                                            void* stream) {
    if (!rect_args(model, n, uv, params, P, p_cols, uv_out)) return NAV_BA_BAD_ARG;
    if (n == 0) return NAV_BA_OK;
    return launch_undistort(model, uv, n, params, R, P, p_cols, uv_out, (hipStream_t)stream);
}

extern "C" int32_t nav_ba_kp_undistort_rect_host(int32_t model, const float* uv, int32_t n, const float* params,
                                                 const double* R, const double* P, int32_t p_cols, float* uv_out,
                                                 int32_t device) {
    if (!rect_args(model, n, uv, params, P, p_cols, uv_out)) return NAV_BA_BAD_ARG;
    if (!gfx950_device(device)) return NAV_BA_NO_DEVICE;
    if (n == 0) return NAV_BA_OK;
    if (hipSetDevice(device) != hipSuccess) return NAV_BA_HIP_ERROR;
    const size_t bytes = sizeof(float) * 2 * (size_t)n;
    HostScratch& hs = host_scratch(device);
    std::lock_guard<std::mutex> lock(hs.mu);
    float* d = (float*)host_scratch_get(hs, bytes);
    if (!d) return NAV_BA_HIP_ERROR;
    int32_t st = NAV_BA_OK;
    if (hipMemcpy(d, uv, bytes, hipMemcpyHostToDevice) != hipSuccess) st = NAV_BA_HIP_ERROR;
    if (st == NAV_BA_OK) st = launch_undistort(model, d, n, params, R, P, p_cols, d, nullptr);
    if (st == NAV_BA_OK && hipMemcpy(uv_out, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) st = NAV_BA_HIP_ERROR;
    return st;
}

extern "C" int32_t nav_ba_kp_undistort(const float* uv, int32_t n, const float* params, float* uv_out, void* stream) {
    return nav_ba_kp_undistort_rect(NAV_KP_RADTAN, uv, n, params, nullptr, nullptr, 3, uv_out, stream);
}

extern "C" int32_t nav_ba_kp_undistort_host(const float* uv, int32_t n, const float* params, float* uv_out,
                                            int32_t device) {
    return nav_ba_kp_undistort_rect_host(NAV_KP_RADTAN, uv, n, params, nullptr, nullptr, 3, uv_out, device);
}

extern "C" int32_t nav_ba_kp_undistort_fisheye(const float* uv, int32_t n, const float* params, float* uv_out,
                                               void* stream) {
    return nav_ba_kp_undistort_rect(NAV_KP_KB8, uv, n, params, nullptr, nullptr, 3, uv_out, stream);
}

extern "C" int32_t nav_ba_kp_undistort_fisheye_host(const float* uv, int32_t n, const float* params, float* uv_out,
                                                    int32_t device) {
    return nav_ba_kp_undistort_rect_host(NAV_KP_KB8, uv, n, params, nullptr, nullptr, 3, uv_out, device);
}

extern "C" int32_t nav_ba_kp_undistort_rect_host_noalloc(int32_t model, const float* uv, int32_t n,
                                                         const float* params, const double* R, const double* P,
                                                         int32_t p_cols, float* uv_out, int32_t device) {
    if (!rect_args(model, n, uv, params, P, p_cols, uv_out)) return NAV_BA_BAD_ARG;
    if (n > kNoAllocMaxPoints) return NAV_BA_TOO_LARGE;
    if (device < 0 || device >= kNoAllocMaxDevices || !gfx950_device(device)) return NAV_BA_NO_DEVICE;
    if (n == 0) return NAV_BA_OK;
    if (hipSetDevice(device) != hipSuccess) return NAV_BA_HIP_ERROR;
    NoAllocSlot& slot = g_noalloc[device];
    std::lock_guard<std::mutex> lock(slot.mu);
    if (!slot.ptr) {  // first call on this device: its address of the array
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NAV_BA_HIP_ERROR;
        slot.ptr = static_cast<float2*>(p);
    }
    float* d = reinterpret_cast<float*>(slot.ptr);
    const size_t bytes = sizeof(float) * 2 * (size_t)n;
    int32_t st = NAV_BA_OK;
    if (hipMemcpy(d, uv, bytes, hipMemcpyHostToDevice) != hipSuccess) st = NAV_BA_HIP_ERROR;
    if (st == NAV_BA_OK) st = launch_undistort(model, d, n, params, R, P, p_cols, d, nullptr);
    if (st == NAV_BA_OK && hipMemcpy(uv_out, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) st = NAV_BA_HIP_ERROR;
    return st;
}
