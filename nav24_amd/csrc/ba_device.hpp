// ba_device.hpp — device-side math of the BA kernels (gfx950).
//
// The Jacobian entries are templated on the compute type T (double for the
// plain FP64 path, float for the mixed path); transforms, residuals, weights
// and all sums are FP64 on both paths.
//
// Reference semantics restated (file:line):
//   edge error     e = z - pi(T X)          OptimizableTypes.hpp:102-107, Pinhole.cpp:30-36
//   Jacobians      P = -Jpi(Xc); J_X = P R; J_xi = P [-[Xc]x | I]
//                                           OptimizableTypes.cpp:140-161, Pinhole.cpp:70-80
//   Huber          rho(s) = s | 2 d sqrt(s) - d^2,  rho'(s) = 1 | d / sqrt(s)   (g2o kernel the
//                                           reference enables at BE_GraphOptim.cpp:191-195)
//   SE3 update     T <- exp(xi) T, quaternion w >= 0 and unit norm (g2o VertexSE3Expmap,
//                                           the vertex of OptimizableTypes.hpp:221-227)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace navba {

constexpr int kWave = 64;

// Camera as used by the kernels: R row-major and t, 12 values (FP64 in HBM).
struct CamRT { double R[9]; double t[3]; };

// Pose state: unit quaternion (x, y, z, w) and translation; 8 doubles.
struct PoseQT { double q[4]; double t[3]; double pad; };

struct Intrinsics { double fx, fy, cx, cy; };

// Linearisation of one edge: residual and depth in FP64, Jacobians in T.
template <typename T>
struct EdgeLin {
    double e0, e1;  // residual
    double z;       // camera-frame depth
    T JX[6];        // 2x3, d e / d X
    T Jxi[12];      // 2x6, d e / d xi (omega, upsilon)
};

__device__ __forceinline__ void transform(const double R[9], const double t[3], double x, double y, double z,
                                          double& xc, double& yc, double& zc) {
    xc = R[0] * x + R[1] * y + R[2] * z + t[0];
    yc = R[3] * x + R[4] * y + R[5] * z + t[1];
    zc = R[6] * x + R[7] * y + R[8] * z + t[2];
}

// Residual only, FP64.  Returns chi2 = e'e; depth in zc.
__device__ __forceinline__ double edge_residual(const double R[9], const double t[3], double x, double y, double z,
                                                double fx, double fy, double cx, double cy, double u, double v,
                                                double& e0, double& e1, double& zc) {
    double xc, yc;
    transform(R, t, x, y, z, xc, yc, zc);
    e0 = u - (fx * xc / zc + cx);
    e1 = v - (fy * yc / zc + cy);
    return e0 * e0 + e1 * e1;
}

// 1/z for the Jacobians.  FP64: v_rcp_f64 refined by two Newton-Raphson
// steps r <- r + r (1 - z r), without the scale / fix-up sequence of an IEEE
// division (on gfx950 it equalled 1.0 / z bit for bit on 2^20 random z over
// 64 binades); z = 0 or a non-finite z keeps the instruction's own result
// (inf, 0 or NaN, as the division gives).
template <typename T>
__device__ __forceinline__ T recip(T z) { return T(1) / z; }

template <>
__device__ __forceinline__ double recip<double>(double z) {
    const double r0 = __builtin_amdgcn_rcp(z);
    double r = r0;
    double e = fma(-z, r, 1.0);
    r = fma(r, e, r);
    e = fma(-z, r, 1.0);
    r = fma(r, e, r);
    return (z != 0.0 && isfinite(z)) ? r : r0;
}

// Full linearisation: residual (FP64) and both Jacobians (T).  The transform
// and the residual stay FP64 on both paths: scene coordinates tens of metres
// from the origin cancel in R X + t, and a pixel residual is the difference of
// two ~300 px values; in FP32 either costs 1e-5 relative on the final cost of
// the 300-keyframe loop.  The Jacobians take the FP64 camera-frame point
// rounded to T.  The constant zeros of P = -Jpi and of [-[Xc]x | I] are
// folded out.
template <typename T>
__device__ __forceinline__ double edge_linearize(const double R[9], const double t[3], double x, double y,
                                                 double z, double fx, double fy, double cx, double cy, double u,
                                                 double v, EdgeLin<T>& L) {
    double xcd, ycd, zcd;
    transform(R, t, x, y, z, xcd, ycd, zcd);
    L.e0 = u - (fx * xcd / zcd + cx);
    L.e1 = v - (fy * ycd / zcd + cy);
    L.z = zcd;
    const T xc = (T)xcd, yc = (T)ycd, zc = (T)zcd, fxt = (T)fx, fyt = (T)fy;
    const T iz = recip<T>(zc);
    const T P00 = -(fxt * iz);
    const T P02 = fxt * xc * iz * iz;
    const T P11 = -(fyt * iz);
    const T P12 = fyt * yc * iz * iz;
    const T R0 = (T)R[0], R1 = (T)R[1], R2 = (T)R[2], R3 = (T)R[3], R4 = (T)R[4], R5 = (T)R[5];
    const T R6 = (T)R[6], R7 = (T)R[7], R8 = (T)R[8];
    // J_X = P R
    L.JX[0] = P00 * R0 + P02 * R6;
    L.JX[1] = P00 * R1 + P02 * R7;
    L.JX[2] = P00 * R2 + P02 * R8;
    L.JX[3] = P11 * R3 + P12 * R6;
    L.JX[4] = P11 * R4 + P12 * R7;
    L.JX[5] = P11 * R5 + P12 * R8;
    // J_xi = P [[0, z, -y, 1, 0, 0], [-z, 0, x, 0, 1, 0], [y, -x, 0, 0, 0, 1]]
    L.Jxi[0] = P02 * yc;
    L.Jxi[1] = P00 * zc - P02 * xc;
    L.Jxi[2] = -(P00 * yc);
    L.Jxi[3] = P00;
    L.Jxi[4] = T(0);
    L.Jxi[5] = P02;
    L.Jxi[6] = P12 * yc - P11 * zc;
    L.Jxi[7] = -(P12 * xc);
    L.Jxi[8] = P11 * xc;
    L.Jxi[9] = T(0);
    L.Jxi[10] = P11;
    L.Jxi[11] = P12;
    return L.e0 * L.e0 + L.e1 * L.e1;
}

// ---- camera models of the edge (Calibration::project / projectJac,
// Calibration.cpp:235-247, dispatching to the model the calibration loaded,
// Calibration.cpp:78-87)
enum CameraModel { kCamPinhole = 0, kCamKB8 = 1 };

// J_xi has the pinhole's structural zeros at (0, 4) and (1, 3) (d e_u / d
// upsilon_y and d e_v / d upsilon_x); the fisheye Jacobian has none.
template <int CM>
struct CamTraits { static constexpr bool xi_zeros = (CM == kCamPinhole); };

// Kannala-Brandt projection of a camera-frame point as the fisheye model's
// Eigen overload computes it (KannalaBrandt8.cpp:80-98): theta and psi from
// the float atan2 of float arguments, the odd polynomial and the image
// coordinates in double.  The float atan2 is evaluated as the double atan2
// rounded to float, here and in the oracle.  No contraction into fused
// multiply-adds: the same double operations as the oracle's.
template <typename PP>
__device__ __forceinline__ void kb8_project(const PP& K, double xc, double yc, double zc, double& pu, double& pv) {
#pragma clang fp contract(off)
    const double x2y2 = xc * xc + yc * yc;
    const float rf = sqrtf((float)x2y2);
    const double th = (double)(float)atan2((double)rf, (double)(float)zc);
    const double psi = (double)(float)atan2((double)(float)yc, (double)(float)xc);
    const double t2 = th * th, t3 = th * t2, t5 = t3 * t2, t7 = t5 * t2, t9 = t7 * t2;
    const double r = th + K.k1 * t3 + K.k2 * t5 + K.k3 * t7 + K.k4 * t9;
    pu = K.fx * r * cos(psi) + K.cx;
    pv = K.fy * r * sin(psi) + K.cy;
}

// d pi / d Xc of the fisheye model (KannalaBrandt8.cpp:195-227), in double,
// with the same expressions (undefined on the optical axis, r = 0, as there).
template <typename PP>
__device__ __forceinline__ void kb8_jac(const PP& K, double x, double y, double z, double J[6]) {
#pragma clang fp contract(off)
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2;
    const double r = sqrt(r2);
    const double r3 = r2 * r;
    const double th = atan2(r, z);
    const double t2 = th * th, t3 = t2 * th, t4 = t2 * t2, t5 = t4 * th;
    const double t6 = t2 * t4, t7 = t6 * th, t8 = t4 * t4, t9 = t8 * th;
    const double f = th + t3 * K.k1 + t5 * K.k2 + t7 * K.k3 + t9 * K.k4;
    const double fd = 1 + 3 * K.k1 * t2 + 5 * K.k2 * t4 + 7 * K.k3 * t6 + 9 * K.k4 * t8;
    J[0] = K.fx * (fd * z * x2 / (r2 * (r2 + z2)) + f * y2 / r3);
    J[1] = K.fx * (fd * z * y * x / (r2 * (r2 + z2)) - f * y * x / r3);
    J[2] = -K.fx * fd * x / (r2 + z2);
    J[3] = K.fy * (fd * z * y * x / (r2 * (r2 + z2)) - f * y * x / r3);
    J[4] = K.fy * (fd * z * y2 / (r2 * (r2 + z2)) + f * x2 / r3);
    J[5] = -K.fy * fd * y / (r2 + z2);
}

// Residual of camera model CM; K carries fx fy cx cy and (fisheye) k1..k4.
template <int CM, typename PP>
__device__ __forceinline__ double edge_residual_m(const PP& K, const double R[9], const double t[3], double x,
                                                  double y, double z, double u, double v, double& e0, double& e1,
                                                  double& zc) {
    if constexpr (CM == kCamPinhole) {
        return edge_residual(R, t, x, y, z, K.fx, K.fy, K.cx, K.cy, u, v, e0, e1, zc);
    } else {
        double xc, yc;
        transform(R, t, x, y, z, xc, yc, zc);
        double pu, pv;
        kb8_project(K, xc, yc, zc, pu, pv);
        e0 = u - pu;
        e1 = v - pv;
        return e0 * e0 + e1 * e1;
    }
}

// Linearisation of camera model CM: P = -d pi / d Xc, J_X = P R, J_xi =
// P [[0, z, -y, 1, 0, 0], [-z, 0, x, 0, 1, 0], [y, -x, 0, 0, 0, 1]]
// (OptimizableTypes.cpp:140-161).  The fisheye Jacobian and its products are
// formed in double and rounded to T.
template <int CM, typename T, typename PP>
__device__ __forceinline__ double edge_linearize_m(const PP& K, const double R[9], const double t[3], double x,
                                                   double y, double z, double u, double v, EdgeLin<T>& L) {
    if constexpr (CM == kCamPinhole) {
        return edge_linearize(R, t, x, y, z, K.fx, K.fy, K.cx, K.cy, u, v, L);
    } else {
        double xcd, ycd, zcd;
        transform(R, t, x, y, z, xcd, ycd, zcd);
        double pu, pv;
        kb8_project(K, xcd, ycd, zcd, pu, pv);
        L.e0 = u - pu;
        L.e1 = v - pv;
        L.z = zcd;
        double J[6];
        kb8_jac(K, xcd, ycd, zcd, J);
        // the entries are formed in double and rounded to T: unlike the
        // pinhole's, the fisheye J_xi entries are sums of two products that can
        // cancel (in FP32 the mixed path's first step moved by 2e-8 relative)
        const double P00 = -J[0], P01 = -J[1], P02 = -J[2], P10 = -J[3], P11 = -J[4], P12 = -J[5];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            L.JX[c] = (T)(P00 * R[c] + P01 * R[3 + c] + P02 * R[6 + c]);
            L.JX[3 + c] = (T)(P10 * R[c] + P11 * R[3 + c] + P12 * R[6 + c]);
        }
        L.Jxi[0] = (T)(-(P01 * zcd) + P02 * ycd);
        L.Jxi[1] = (T)(P00 * zcd - P02 * xcd);
        L.Jxi[2] = (T)(-(P00 * ycd) + P01 * xcd);
        L.Jxi[3] = (T)P00;
        L.Jxi[4] = (T)P01;
        L.Jxi[5] = (T)P02;
        L.Jxi[6] = (T)(-(P11 * zcd) + P12 * ycd);
        L.Jxi[7] = (T)(P10 * zcd - P12 * xcd);
        L.Jxi[8] = (T)(-(P10 * ycd) + P11 * xcd);
        L.Jxi[9] = (T)P10;
        L.Jxi[10] = (T)P11;
        L.Jxi[11] = (T)P12;
        return L.e0 * L.e0 + L.e1 * L.e1;
    }
}

// Huber: returns rho(s); w = rho'(s).  Without the kernel rho = s, w = 1.
template <typename T>
__device__ __forceinline__ T robustify(T s, int robust, T delta, T& w) {
    if (!robust) { w = T(1); return s; }
    const T d2 = delta * delta;
    if (s <= d2) { w = T(1); return s; }
    const T r = sqrt(s);
    w = delta / r;
    return T(2) * r * delta - d2;
}

// 3x3 symmetric inverse by the adjugate; A given as (a00 a01 a02 a11 a12 a22),
// result in the same packing.
__device__ __forceinline__ void inv3_sym(const double a[6], double o[6]) {
    const double a00 = a[0], a01 = a[1], a02 = a[2], a11 = a[3], a12 = a[4], a22 = a[5];
    const double c00 = a11 * a22 - a12 * a12;
    const double c01 = a12 * a02 - a01 * a22;
    const double c02 = a01 * a12 - a11 * a02;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = c01 * id;
    o[2] = c02 * id;
    o[3] = (a00 * a22 - a02 * a02) * id;
    o[4] = (a02 * a01 - a00 * a12) * id;
    o[5] = (a00 * a11 - a01 * a01) * id;
}

// D = (Hll + lambda I)^-1 of one point (packed 00 01 02 11 12 22) from its
// block h = (Hll packed (6), bl (3)): the damped point block the Schur
// complement eliminates and the back-substitution applies.
__device__ __forceinline__ void point_dinv(const double* __restrict__ h, double lambda, double D[6]) {
    double a[6] = {h[0] + lambda, h[1], h[2], h[3] + lambda, h[4], h[5] + lambda};
    inv3_sym(a, D);
}

// ---- quaternions (FP64) ----
__device__ __forceinline__ void quat_to_R(const double q[4], double R[9]) {
    const double tx = 2 * q[0], ty = 2 * q[1], tz = 2 * q[2];
    const double twx = tx * q[3], twy = ty * q[3], twz = tz * q[3];
    const double txx = tx * q[0], txy = ty * q[0], txz = tz * q[0];
    const double tyy = ty * q[1], tyz = tz * q[1], tzz = tz * q[2];
    R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
    R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

// Rotation matrix -> quaternion (Shoemake 1987, trace / largest-diagonal branches).
__device__ __forceinline__ void R_to_quat(const double R[9], double q[4]) {
    const double tr = R[0] + R[4] + R[8];
    if (tr > 0.0) {
        double s = sqrt(tr + 1.0);
        q[3] = 0.5 * s;
        s = 0.5 / s;
        q[0] = (R[7] - R[5]) * s;
        q[1] = (R[2] - R[6]) * s;
        q[2] = (R[3] - R[1]) * s;
    } else {
        int i = 0;
        if (R[4] > R[0]) i = 1;
        if (R[8] > R[3 * i + i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double s = sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k + k] + 1.0);
        double c[3];
        c[i] = 0.5 * s;
        s = 0.5 / s;
        q[3] = (R[3 * k + j] - R[3 * j + k]) * s;
        c[j] = (R[3 * j + i] + R[3 * i + j]) * s;
        c[k] = (R[3 * k + i] + R[3 * i + k]) * s;
        q[0] = c[0]; q[1] = c[1]; q[2] = c[2];
    }
}

__device__ __forceinline__ void quat_normalize_pos(double q[4]) {
    if (q[3] < 0) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
    const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}

__device__ __forceinline__ void quat_rotate(const double q[4], const double v[3], double o[3]) {
    double ux = q[1] * v[2] - q[2] * v[1];
    double uy = q[2] * v[0] - q[0] * v[2];
    double uz = q[0] * v[1] - q[1] * v[0];
    ux += ux; uy += uy; uz += uz;
    o[0] = v[0] + q[3] * ux + (q[1] * uz - q[2] * uy);
    o[1] = v[1] + q[3] * uy + (q[2] * ux - q[0] * uz);
    o[2] = v[2] + q[3] * uz + (q[0] * uy - q[1] * ux);
}

// T_out = exp(xi) * T_in.  exp: theta < 1e-5 -> R = I + W + W^2, V = R;
// otherwise Rodrigues R and left Jacobian V.
__device__ inline void se3_oplus(const double xi[6], const PoseQT& in, PoseQT& out) {
    const double wx = xi[0], wy = xi[1], wz = xi[2];
    const double theta = sqrt(wx * wx + wy * wy + wz * wz);
    const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            W2[3 * i + j] = W[3 * i] * W[j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
    double R[9], V[9];
    if (theta < 0.00001) {
#pragma unroll
        for (int k = 0; k < 9; ++k) { R[k] = ((k % 4 == 0) ? 1.0 : 0.0) + W[k] + W2[k]; V[k] = R[k]; }
    } else {
        double s, c;
        sincos(theta, &s, &c);
        const double a = s / theta;
        const double b = (1 - c) / (theta * theta);
        const double cc = (theta - s) / (theta * theta * theta);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double I = (k % 4 == 0) ? 1.0 : 0.0;
            R[k] = I + a * W[k] + b * W2[k];
            V[k] = I + b * W[k] + cc * W2[k];
        }
    }
    double qi[4];
    R_to_quat(R, qi);
    quat_normalize_pos(qi);
    double ti[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ti[i] = V[3 * i] * xi[3] + V[3 * i + 1] * xi[4] + V[3 * i + 2] * xi[5];
    double rt[3];
    quat_rotate(qi, in.t, rt);
    out.t[0] = ti[0] + rt[0]; out.t[1] = ti[1] + rt[1]; out.t[2] = ti[2] + rt[2];
    const double* b = in.q;
    out.q[3] = qi[3] * b[3] - qi[0] * b[0] - qi[1] * b[1] - qi[2] * b[2];
    out.q[0] = qi[3] * b[0] + qi[0] * b[3] + qi[1] * b[2] - qi[2] * b[1];
    out.q[1] = qi[3] * b[1] + qi[1] * b[3] + qi[2] * b[0] - qi[0] * b[2];
    out.q[2] = qi[3] * b[2] + qi[2] * b[3] + qi[0] * b[1] - qi[1] * b[0];
    quat_normalize_pos(out.q);
    out.pad = 0.0;
}

// ---- reductions (fixed order: deterministic) ----
// DPP move of a 32-bit value (row_mask / bank_mask all rows, no bound control).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) { return __int_as_float(dpp_i<CTRL>(__float_as_int(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffll));
    const int hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes, returned in every lane (wave-uniform): DPP within
// each row of 16 lanes (quad swaps, then row rotations by 4 and 8), then the
// four row sums read from lanes 15, 31, 47, 63.  All lanes must be active.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x124>(v);  // row_ror:4
    v += dpp_move<0x128>(v);  // row_ror:8
    return (readlane(v, 15) + readlane(v, 31)) + (readlane(v, 47) + readlane(v, 63));
}

// The wave's sums of a[0..27) by a reduce-scatter: at distances 32, 16, 8,
// 4 and 2 each lane pair splits its current values in halves (28 -> 14 -> 7
// -> 4 -> 2 -> 1, zero padded), keeps one half and adds the partner's copy
// of it; then lanes 2i and 2i+1 add their values.  Lane l ends with the sum
// of index 14 b5 + 7 b4 + (4 b3 + 2 b2 + b1) (b_k = bit k of l; none when the
// last term is 7 or the index is 27).  A fixed order: the same bits every
// run.  About a third of the instructions of 27 separate wave sums.
__device__ __forceinline__ double wave_reduce_scatter27(const double (&a)[27], int lane, int& idx) {
    double v14[14], v7[7], v4[4], v2[2];
    const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0, h3 = (lane & 8) != 0, h2 = (lane & 4) != 0,
               h1 = (lane & 2) != 0;
#pragma unroll
    for (int j = 0; j < 14; ++j) {
        const double lo = a[j], hi = 14 + j < 27 ? a[14 + j] : 0.0;
        v14[j] = (h5 ? hi : lo) + __shfl_xor(h5 ? lo : hi, 32);
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const double lo = v14[j], hi = v14[7 + j];
        v7[j] = (h4 ? hi : lo) + __shfl_xor(h4 ? lo : hi, 16);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double lo = v7[j], hi = 4 + j < 7 ? v7[4 + j] : 0.0;
        v4[j] = (h3 ? hi : lo) + __shfl_xor(h3 ? lo : hi, 8);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double lo = v4[j], hi = v4[2 + j];
        v2[j] = (h2 ? hi : lo) + __shfl_xor(h2 ? lo : hi, 4);
    }
    double v = (h1 ? v2[1] : v2[0]) + __shfl_xor(h1 ? v2[0] : v2[1], 2);
    v += __shfl_xor(v, 1);
    const int local = 4 * (int)h3 + 2 * (int)h2 + (int)h1;
    const int i = 14 * (int)h5 + 7 * (int)h4 + local;
    idx = (local < 7 && i < 27) ? i : -1;
    return v;
}

// Sum over a block of NT threads; result valid in thread 0.  `sh` needs NT/64 slots.
template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v, T* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) r += sh[i];
    }
    return r;
}

}  // namespace navba
