// kp_undistort_oracle.cpp — CPU restatement of the key-point undistortion of
// the radial-tangential pinhole camera and of the Kannala-Brandt fisheye camera.
//
// TEST INFRASTRUCTURE ONLY: the checker of nav_ba_kp_undistort
// (include/nav_kp_undistort.h); the product library never links or calls it.
//
// Restates PinholeRadTan::UndistortKeyPoints
// (core/sensor/camera/models/PinholeRadTan.cpp:11-24), which calls
// cv::undistortPoints(pts, pts, K, D, R, P) of OpenCV >= 4.2.0
// (core/CMakeLists.txt:10; OpenCV is absent from this image) with K from
// (fx, fy, cx, cy), D = (k1, k2, p1, p2) and the camera's R and P, identity
// and K unless the calibration sets them (GeometricCamera.h:47-48,65-66,
// 108-109, Calibration.cpp:63-93, PinholeRadTan.hpp:13-18).  The restated steps are
// those of OpenCV's published undistortPoints: normalise by K (multiplying by
// 1/fx, 1/fy formed in double), a fixed number of fixed-point iterations
// x = (x0 - delta_tan(x)) / (1 + k1 r^2 + k2 r^4) (5 under the default
// termination; the iteration stops with the normalised input point when the
// radial factor's reciprocal is negative), then P R and the float result.
// Double arithmetic in the order written, no fused multiply-adds (compiled
// with -ffp-contract=off).
//
// Parity: unpinned against OpenCV itself (it cannot be run here); pinned by
// what the mathematics fixes (tests/test_kp_undistort_oracle.py):
//   - zero distortion returns the input point (u = fx ((u - cx) / fx) + cx
//     to within double rounding, then rounded to float);
//   - with p1 = p2 = 0 a point on the principal row stays on it (y = 0 is a
//     fixed point), and the radial model is symmetric in (x, y) -> (-x, -y);
//   - with many iterations the result inverts the forward Brown-Conrady
//     model (oracle_kp_distort) to 1e-9 px for mild distortion;
//   - the first iteration equals its closed form on hand-picked points;
//   - a radial factor that turns negative returns the input point;
//   - with R and P set, the result is the R = I, P = K result mapped by the
//     homography P R K^-1 (within the float rounding of the latter), for
//     both camera models; a 4th column of P is ignored.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

// RR = P3 R in double: P3 the left 3x3 of P (p_cols columns) or K when P is
// null, R or the identity; each entry ((p0 r0 + p1 r1) + p2 r2) in column order.
void form_rr(double fx, double fy, double cx, double cy, const double* R, const double* P, int p_cols,
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

void undistort_radtan(const float* uv, int32_t n, const float* params, int32_t iters, const double RR[9],
                      float* out) {
    const double fx = params[0], fy = params[1], cx = params[2], cy = params[3];
    const double k1 = params[4], k2 = params[5], p1 = params[6], p2 = params[7];
    const double ifx = 1.0 / fx, ify = 1.0 / fy;
    for (int32_t i = 0; i < n; ++i) {
        const double u = uv[2 * i], v = uv[2 * i + 1];
        double x = (u - cx) * ifx;
        double y = (v - cy) * ify;
        const double x0 = x, y0 = y;
        for (int32_t j = 0; j < iters; ++j) {
            const double r2 = x * x + y * y;
            const double icdist = 1.0 / (1.0 + (k2 * r2 + k1) * r2);
            if (icdist < 0.0) {
                x = (u - cx) * ifx;
                y = (v - cy) * ify;
                break;
            }
            const double dx = 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
            const double dy = p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
            x = (x0 - dx) * icdist;
            y = (y0 - dy) * icdist;
        }
        const double xx = RR[0] * x + RR[1] * y + RR[2];
        const double yy = RR[3] * x + RR[4] * y + RR[5];
        const double ww = 1.0 / (RR[6] * x + RR[7] * y + RR[8]);
        out[2 * i] = (float)(xx * ww);
        out[2 * i + 1] = (float)(yy * ww);
    }
}

}  // namespace

extern "C" {

// uv, out: n points (u, v); params: fx fy cx cy k1 k2 p1 p2; iters: 5 = the reference's call.
void oracle_kp_undistort(const float* uv, int32_t n, const float* params, int32_t iters, float* out) {
    double RR[9];
    form_rr(params[0], params[1], params[2], params[3], nullptr, nullptr, 3, RR);  // P R = K
    undistort_radtan(uv, n, params, iters, RR, out);
}

// Forward model (the one undistortion inverts), in double: normalised
// (x, y) -> distorted pixel, x_d = x (1 + k1 r^2 + k2 r^4) + 2 p1 x y + p2 (r^2 + 2 x^2),
// y_d = y (1 + k1 r^2 + k2 r^4) + p1 (r^2 + 2 y^2) + 2 p2 x y.
void oracle_kp_distort(const double* xy, int32_t n, const float* params, double* uv) {
    const double fx = params[0], fy = params[1], cx = params[2], cy = params[3];
    const double k1 = params[4], k2 = params[5], p1 = params[6], p2 = params[7];
    for (int32_t i = 0; i < n; ++i) {
        const double x = xy[2 * i], y = xy[2 * i + 1];
        const double r2 = x * x + y * y;
        const double rad = 1.0 + k1 * r2 + k2 * r2 * r2;
        const double xd = x * rad + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
        const double yd = y * rad + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
        uv[2 * i] = fx * xd + cx;
        uv[2 * i + 1] = fy * yd + cy;
    }
}

// Kannala-Brandt fisheye: KannalaBrandt8::UndistortKeyPoints
// (core/sensor/camera/models/KannalaBrandt8.cpp:231-242) calls
// cv::fisheye::undistortPoints(pts, pts, K, D = (k1..k4), R = I, P = K).
// Restated from OpenCV 4.2's published fisheye::undistortPoints (the
// reference's minimum version; later versions add a termination argument
// and mark non-converged points): pw = (p - c) / f; theta_d = |pw| clipped
// to [-pi/2, pi/2]; above 1e-8, up to 10 Newton steps on
// theta (1 + k1 t^2 + k2 t^4 + k3 t^6 + k4 t^8) = theta_d, stopping once a
// step is below 1e-8, and scale = tan(theta) / theta_d; then P R (pw scale, 1)
// (a 3x3 product summed from 0) and the float result.  Parity: unpinned
// against OpenCV (absent); pinned by tests/test_kp_undistort_oracle.py
// (equidistant closed form at k = 0, collinearity with the principal point,
// the principal point fixed, and round trips through oracle_kp_distort_fisheye).
static void undistort_kb8(const float* uv, int32_t n, const float* params, const double RR[9], float* out) {
    const double fx = params[0], fy = params[1], cx = params[2], cy = params[3];
    const double k[4] = {params[4], params[5], params[6], params[7]};
    const double half_pi = 1.5707963267948966;
    for (int32_t i = 0; i < n; ++i) {
        const double pw0 = ((double)uv[2 * i] - cx) / fx, pw1 = ((double)uv[2 * i + 1] - cy) / fy;
        double scale = 1.0;
        double theta_d = std::sqrt(pw0 * pw0 + pw1 * pw1);
        theta_d = std::min(std::max(-half_pi, theta_d), half_pi);
        if (theta_d > 1e-8) {
            double theta = theta_d;
            for (int j = 0; j < 10; ++j) {
                const double theta2 = theta * theta, theta4 = theta2 * theta2, theta6 = theta4 * theta2,
                             theta8 = theta6 * theta2;
                const double k0t2 = k[0] * theta2, k1t4 = k[1] * theta4, k2t6 = k[2] * theta6, k3t8 = k[3] * theta8;
                const double fix = (theta * (1 + k0t2 + k1t4 + k2t6 + k3t8) - theta_d) /
                                   (1 + 3 * k0t2 + 5 * k1t4 + 7 * k2t6 + 9 * k3t8);
                theta = theta - fix;
                if (std::fabs(fix) < 1e-8) break;
            }
            scale = std::tan(theta) / theta_d;
        }
        const double pu0 = pw0 * scale, pu1 = pw1 * scale;
        const double pr0 = ((0.0 + RR[0] * pu0) + RR[1] * pu1) + RR[2] * 1.0;
        const double pr1 = ((0.0 + RR[3] * pu0) + RR[4] * pu1) + RR[5] * 1.0;
        const double pr2 = ((0.0 + RR[6] * pu0) + RR[7] * pu1) + RR[8] * 1.0;
        out[2 * i] = (float)(pr0 / pr2);
        out[2 * i + 1] = (float)(pr1 / pr2);
    }
}

void oracle_kp_undistort_fisheye(const float* uv, int32_t n, const float* params, float* out) {
    double RR[9];
    form_rr(params[0], params[1], params[2], params[3], nullptr, nullptr, 3, RR);
    undistort_kb8(uv, n, params, RR, out);
}

// Either model (0 radial-tangential, 1 Kannala-Brandt) with the camera's R
// (3x3, null = identity) and P (p_cols = 3 or 4 columns, null = K), as the
// reference passes mR and mP to OpenCV (PinholeRadTan.cpp:20, KannalaBrandt8.cpp:238).
void oracle_kp_undistort_rect(int32_t model, const float* uv, int32_t n, const float* params, const double* R,
                              const double* P, int32_t p_cols, float* out) {
    double RR[9];
    form_rr(params[0], params[1], params[2], params[3], R, P, p_cols, RR);
    if (model == 0) undistort_radtan(uv, n, params, 5, RR, out);
    else undistort_kb8(uv, n, params, RR, out);
}

// Forward Kannala-Brandt model (the one fisheye undistortion inverts), in
// double: normalised pinhole (a, b) = (X / Z, Y / Z) -> distorted pixel,
// theta = atan(|(a, b)|), theta_d = theta (1 + k1 theta^2 + ... + k4 theta^8),
// u = fx theta_d / |(a, b)| a + cx (likewise v).
void oracle_kp_distort_fisheye(const double* ab, int32_t n, const float* params, double* uv) {
    const double fx = params[0], fy = params[1], cx = params[2], cy = params[3];
    const double k1 = params[4], k2 = params[5], k3 = params[6], k4 = params[7];
    for (int32_t i = 0; i < n; ++i) {
        const double a = ab[2 * i], b = ab[2 * i + 1];
        const double r = std::sqrt(a * a + b * b);
        const double th = std::atan(r), t2 = th * th;
        const double thd = th * (1 + t2 * (k1 + t2 * (k2 + t2 * (k3 + t2 * k4))));
        const double s = r > 0 ? thd / r : 1.0;
        uv[2 * i] = fx * s * a + cx;
        uv[2 * i + 1] = fy * s * b + cy;
    }
}

}  // extern "C"
