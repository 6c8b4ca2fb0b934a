/*
 * nav_kp_undistort.h — key-point undistortion of the radial-tangential
 * pinhole camera on the device: the step that produces the BA measurements.
 *
 * Replaces PinholeRadTan::UndistortKeyPoints
 * (core/sensor/camera/models/PinholeRadTan.cpp:11-24), reached per
 * observation through Calibration::undistort (core/sensor/camera/Calibration.cpp:135-149,
 * called at core/sensor/camera/Camera.cpp:175), whose result becomes the
 * measurement the BA edge reads (KeyPoint2D::getPointUd).  The reference calls
 * OpenCV's cv::undistortPoints(pts, pts, K, D, R, P) (OpenCV >= 4.2.0,
 * core/CMakeLists.txt:10; absent from this image) with
 *     K = [fx 0 cx; 0 fy cy; 0 0 1], D = (k1, k2, p1, p2)  (PinholeRadTan.hpp:13-18)
 *     R = the camera's rectification, I unless the calibration sets one, and
 *     P = its projection, K unless set  (GeometricCamera.h:47-48,65-66,108-109,
 *     set from the configuration at Calibration.cpp:63-93)
 * all float, key points cv::Point2f.  The published algorithm of that call,
 * in double precision:
 *     x0 = x = (u - cx) / fx,  y0 = y = (v - cy) / fy   (as (u - cx) * (1 / fx))
 *     5 times:  r2 = x^2 + y^2
 *               icdist = 1 / (1 + (k2 r2 + k1) r2)      (icdist < 0: x, y = x0, y0; stop)
 *               dx = 2 p1 x y + p2 (r2 + 2 x^2),  dy = p1 (r2 + 2 y^2) + 2 p2 x y
 *               x = (x0 - dx) icdist,  y = (y0 - dy) icdist
 *     (u', v') = (RR x) / (RR x)_z with RR = P(:, 0:3) R formed in double,
 *     rounded to float.
 * The device evaluates exactly these double operations (no fused
 * multiply-adds), so its output equals a plain double evaluation bit for bit.
 *
 * Stateless: no handle.  In place (uv_out == uv) is allowed, as the
 * reference's call is.  The same entry points exist for the Kannala-Brandt
 * fisheye camera below.
 */
#ifndef NAV_KP_UNDISTORT_H
#define NAV_KP_UNDISTORT_H

#include <stdint.h>
#include "nav_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* params: fx fy cx cy k1 k2 p1 p2 (the reference's vParams order,
 * PinholeRadTan(vParams), float).  uv, uv_out: n key points (u, v), float.
 * The entries without R and P are the camera's defaults, R = I and P = K. */

/* Device arrays, on `stream` (a hipStream_t, NULL = the null stream); returns
 * after the launch is queued.  NAV_BA_BAD_ARG for n < 0 or a null pointer
 * with n > 0, NAV_BA_HIP_ERROR if the launch fails. */
int32_t nav_ba_kp_undistort(const float* uv, int32_t n, const float* params, float* uv_out, void* stream);

/* Host arrays, through device `device` (copies in and out; blocks until done).
 * NAV_BA_NO_DEVICE without a usable gfx950 device. */
int32_t nav_ba_kp_undistort_host(const float* uv, int32_t n, const float* params, float* uv_out, int32_t device);

/* The Kannala-Brandt fisheye camera: KannalaBrandt8::UndistortKeyPoints
 * (core/sensor/camera/models/KannalaBrandt8.cpp:231-242), which calls
 * cv::fisheye::undistortPoints(pts, pts, K, D, R, P) with D = (k1, k2, k3, k4)
 * (KannalaBrandt8.cpp:27-29), R = I and P = K.  The published algorithm of that
 * call in OpenCV 4.2 (the reference's minimum version), in double precision:
 *     pw = ((u - cx) / fx, (v - cy) / fy),  theta_d = min(|pw|, pi / 2)
 *     theta_d > 1e-8:  theta = theta_d; up to 10 Newton steps on
 *         theta (1 + k1 theta^2 + k2 theta^4 + k3 theta^6 + k4 theta^8) = theta_d,
 *         stopping once a step is below 1e-8;  scale = tan(theta) / theta_d
 *     otherwise scale = 1;  (u', v') = (RR pw scale) / (RR pw scale)_z, RR = P R.
 * tan() comes from the device math library, so a result can differ from a
 * host evaluation by a float ulp (not observed in the tests).  params: fx fy
 * cx cy k1 k2 k3 k4 (the reference's vParameters order). */
int32_t nav_ba_kp_undistort_fisheye(const float* uv, int32_t n, const float* params, float* uv_out, void* stream);
int32_t nav_ba_kp_undistort_fisheye_host(const float* uv, int32_t n, const float* params, float* uv_out,
                                         int32_t device);

/* Either camera model with the camera's rectification R (3x3 row-major
 * double, NULL = identity) and projection P (p_cols = 3 or 4 columns,
 * row-major double, its left 3x3 used, NULL = K): the matrices the reference
 * passes to both OpenCV calls (mR, mP).  RR = P(:, 0:3) R is formed on the
 * host in double, each entry a sum of three products in column order.
 * model: NAV_KP_RADTAN (params fx fy cx cy k1 k2 p1 p2) or NAV_KP_KB8
 * (fx fy cx cy k1 k2 k3 k4).  Device arrays on `stream`, or host arrays
 * through `device` (_host).  NAV_BA_BAD_ARG for an unknown model, p_cols
 * other than 3 or 4 with P given, n < 0 or a null array with n > 0. */
enum nav_kp_model { NAV_KP_RADTAN = 0, NAV_KP_KB8 = 1 };

int32_t nav_ba_kp_undistort_rect(int32_t model, const float* uv, int32_t n, const float* params, const double* R,
                                 const double* P, int32_t p_cols, float* uv_out, void* stream);
int32_t nav_ba_kp_undistort_rect_host(int32_t model, const float* uv, int32_t n, const float* params,
                                      const double* R, const double* P, int32_t p_cols, float* uv_out,
                                      int32_t device);

/* nav_ba_kp_undistort_rect_host without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * key points are staged in a fixed device array of the library, at most
 * NAV_KP_NOALLOC_MAX_POINTS (NAV_BA_TOO_LARGE above), which every device the
 * library runs on holds until the process ends (16 MiB each).  Calls on one
 * device share that array and are serialised by a per-device lock; the array
 * keeps the last call's points.  Same arguments, results and errors as
 * nav_ba_kp_undistort_rect_host otherwise. */
#define NAV_KP_NOALLOC_MAX_POINTS (1 << 21)

int32_t nav_ba_kp_undistort_rect_host_noalloc(int32_t model, const float* uv, int32_t n, const float* params,
                                              const double* R, const double* P, int32_t p_cols, float* uv_out,
                                              int32_t device);

#ifdef __cplusplus
}
#endif
#endif /* NAV_KP_UNDISTORT_H */
