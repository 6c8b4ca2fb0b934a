// ba_oracle.cpp — FP64 CPU restatement of nav24's BA LM path (test infrastructure only).
//
// The checker for the GPU library; see ba_oracle.h for the reference files it
// follows.  Single-threaded on purpose: the reference runs one g2o
// SparseOptimizer on the caller's thread (BE_GraphOptim.cpp:40).
//
// Pins (no reference test, fixture or golden vector exists for this path):
//   project / edge Jacobians  : closed form + FP64 central finite differences (tests)
//   quat_from_R / R_from_quat : round trip, orthonormality, known rotations (tests)
//   se3_exp                   : Rodrigues closed form, exp(0)=I, exp(w)exp(-w)=I (tests)
//   huber                     : piecewise closed form, continuity at delta^2 (tests)
//   inv3                      : A*inv(A)=I (tests)
//   linear blocks / Schur     : brute-force dense system in long double (oracle_dense_step)
//   LM schedule               : hand-computed 1-camera/1-point trace, invariants (tests)
//   g2o-internal constants (tau=1e-5, +1e-3 scale guard, [1/3,2/3] crop, 10 trials,
//   small-angle exp branch) are restated from the published g2o sources and are
//   parity unpinned: nothing in this container can run g2o.

#include "ba_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

struct Quat { double x, y, z, w; };

// Rotate v by unit quaternion q, in the two-cross-product form used by
// Eigen's Quaternion*Vector3 (the form g2o's SE3Quat::map relies on,
// OptimizableTypes.hpp:106 -> SE3Quat::map = r*xyz + t).
inline void quat_rotate(const Quat& q, const double v[3], double out[3]) {
    double ux = q.y * v[2] - q.z * v[1];
    double uy = q.z * v[0] - q.x * v[2];
    double uz = q.x * v[1] - q.y * v[0];
    ux += ux; uy += uy; uz += uz;
    out[0] = v[0] + q.w * ux + (q.y * uz - q.z * uy);
    out[1] = v[1] + q.w * uy + (q.z * ux - q.x * uz);
    out[2] = v[2] + q.w * uz + (q.x * uy - q.y * ux);
}

// Hamilton product a*b.
inline Quat quat_mul(const Quat& a, const Quat& b) {
    Quat r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return r;
}

// g2o SE3Quat::normalizeRotation: flip to w >= 0, then unit norm.
inline void quat_normalize_pos(Quat& q) {
    if (q.w < 0) { q.x = -q.x; q.y = -q.y; q.z = -q.z; q.w = -q.w; }
    double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    q.x /= n; q.y /= n; q.z /= n; q.w /= n;
}

// Rotation matrix -> quaternion, Shoemake's method ("Quaternion Calculus and
// Fast Animation", SIGGRAPH 1987 course notes): trace branch when the trace is
// positive, otherwise the largest diagonal element picks the pivot axis.
inline Quat quat_from_R(const double R[9]) {
    auto m = [&](int i, int j) { return R[3 * i + j]; };
    Quat q;
    double tr = m(0, 0) + m(1, 1) + m(2, 2);
    if (tr > 0.0) {
        double s = std::sqrt(tr + 1.0);
        q.w = 0.5 * s;
        s = 0.5 / s;
        q.x = (m(2, 1) - m(1, 2)) * s;
        q.y = (m(0, 2) - m(2, 0)) * s;
        q.z = (m(1, 0) - m(0, 1)) * s;
    } else {
        int i = 0;
        if (m(1, 1) > m(0, 0)) i = 1;
        if (m(2, 2) > m(i, i)) i = 2;
        int j = (i + 1) % 3, k = (j + 1) % 3;
        double s = std::sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
        double c[3];
        c[i] = 0.5 * s;
        s = 0.5 / s;
        q.w = (m(k, j) - m(j, k)) * s;
        c[j] = (m(j, i) + m(i, j)) * s;
        c[k] = (m(k, i) + m(i, k)) * s;
        q.x = c[0]; q.y = c[1]; q.z = c[2];
    }
    return q;
}

inline void R_from_quat(const Quat& q, double R[9]) {
    const double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
    R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

struct Pose { Quat q; double t[3]; };

// g2o SE3Quat::exp: xi = (omega, upsilon).  Small-angle branch below 1e-5 uses
// R = I + W + W^2 and V = R (the published g2o code, kept as is).
inline Pose se3_exp(const double xi[6]) {
    const double wx = xi[0], wy = xi[1], wz = xi[2];
    const double theta = std::sqrt(wx * wx + wy * wy + wz * wz);
    const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            W2[3 * i + j] = W[3 * i + 0] * W[0 + j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
    double R[9], V[9];
    if (theta < 0.00001) {
        for (int k = 0; k < 9; ++k) R[k] = ((k % 4 == 0) ? 1.0 : 0.0) + W[k] + W2[k];
        for (int k = 0; k < 9; ++k) V[k] = R[k];
    } else {
        const double a = std::sin(theta) / theta;
        const double b = (1 - std::cos(theta)) / (theta * theta);
        const double c = (theta - std::sin(theta)) / std::pow(theta, 3);
        for (int k = 0; k < 9; ++k) {
            const double I = (k % 4 == 0) ? 1.0 : 0.0;
            R[k] = I + a * W[k] + b * W2[k];
            V[k] = I + b * W[k] + c * W2[k];
        }
    }
    Pose p;
    p.q = quat_from_R(R);
    const double* u = xi + 3;
    for (int i = 0; i < 3; ++i) p.t[i] = V[3 * i] * u[0] + V[3 * i + 1] * u[1] + V[3 * i + 2] * u[2];
    quat_normalize_pos(p.q);  // SE3Quat(Quaterniond, t) constructor normalises
    return p;
}

// g2o VertexSE3Expmap::oplusImpl: T <- exp(xi) * T (SE3Quat::operator*).
inline Pose se3_oplus(const double xi[6], const Pose& T) {
    Pose inc = se3_exp(xi);
    Pose r;
    double rt[3];
    quat_rotate(inc.q, T.t, rt);
    for (int i = 0; i < 3; ++i) r.t[i] = inc.t[i] + rt[i];
    r.q = quat_mul(inc.q, T.q);
    quat_normalize_pos(r.q);
    return r;
}

// g2o RobustKernelHuber::robustify.
inline void huber(double e, double delta, double rho[3]) {
    const double dsqr = delta * delta;
    if (e <= dsqr) {
        rho[0] = e; rho[1] = 1.0; rho[2] = 0.0;
    } else {
        const double sqrte = std::sqrt(e);
        rho[0] = 2 * sqrte * delta - dsqr;
        rho[1] = delta / sqrte;
        rho[2] = -0.5 * rho[1] / e;
    }
}

// 3x3 inverse by the adjugate (cofactor transpose over the determinant).
inline void inv3(const double A[9], double B[9]) {
    const double c00 = A[4] * A[8] - A[5] * A[7];
    const double c01 = A[5] * A[6] - A[3] * A[8];
    const double c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    const double id = 1.0 / det;
    B[0] = c00 * id;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    B[3] = c01 * id;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    B[6] = c02 * id;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// Camera of the edge: the pinhole (model 0) or the Kannala-Brandt fisheye
// (model 1, k1..k4), the model Calibration::loadParams picks
// (Calibration.cpp:78-87) and the edge reaches through Calibration::project /
// projectJac (Calibration.cpp:235-247).
struct Intr {
    double fx, fy, cx, cy;
    int model;
    double k1, k2, k3, k4;
};

// KannalaBrandt8::project, the Eigen overload the edge calls
// (KannalaBrandt8.cpp:80-98): theta = atan2f(sqrtf(float(x^2 + y^2)), float(z))
// and psi = atan2f(float(y), float(x)) in float, the odd polynomial and the
// image coordinates in double.  The float atan2 is restated as the double
// atan2 of the same float arguments rounded to float (the math library's
// atan2f may differ from that by a float ulp: parity unpinned to that ulp).
inline void kb8_project(const Intr& K, const double Xc[3], double& pu, double& pv) {
    const double x2y2 = Xc[0] * Xc[0] + Xc[1] * Xc[1];
    const float rf = std::sqrt((float)x2y2);
    const double th = (double)(float)std::atan2((double)rf, (double)(float)Xc[2]);
    const double psi = (double)(float)std::atan2((double)(float)Xc[1], (double)(float)Xc[0]);
    const double t2 = th * th, t3 = th * t2, t5 = t3 * t2, t7 = t5 * t2, t9 = t7 * t2;
    const double r = th + K.k1 * t3 + K.k2 * t5 + K.k3 * t7 + K.k4 * t9;
    pu = K.fx * r * std::cos(psi) + K.cx;
    pv = K.fy * r * std::sin(psi) + K.cy;
}

// KannalaBrandt8::projectJac, the Eigen overload (KannalaBrandt8.cpp:195-227),
// all in double.
inline void kb8_jac(const Intr& K, const double Xc[3], double J[6]) {
    const double x = Xc[0], y = Xc[1], z = Xc[2];
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2;
    const double r = std::sqrt(r2);
    const double r3 = r2 * r;
    const double th = std::atan2(r, z);
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

// EdgeSE3ProjectXYZ::computeError (OptimizableTypes.hpp:102-107): e = z -
// pi(T X); pi = Pinhole::project (Pinhole.cpp:30-36, (fx x/z + cx, fy y/z +
// cy)) or kb8_project.
inline void edge_error(const Pose& T, const double X[3], const Intr& K, const double uv[2],
                       double e[2], double Xc[3]) {
    double r[3];
    quat_rotate(T.q, X, r);
    Xc[0] = r[0] + T.t[0]; Xc[1] = r[1] + T.t[1]; Xc[2] = r[2] + T.t[2];
    if (K.model == 1) {
        double pu, pv;
        kb8_project(K, Xc, pu, pv);
        e[0] = uv[0] - pu;
        e[1] = uv[1] - pv;
        return;
    }
    e[0] = uv[0] - (K.fx * Xc[0] / Xc[2] + K.cx);
    e[1] = uv[1] - (K.fy * Xc[1] / Xc[2] + K.cy);
}

// EdgeSE3ProjectXYZ::linearizeOplus (OptimizableTypes.cpp:140-161) with
// Pinhole::projectJac (Pinhole.cpp:70-80) or kb8_jac:
//   P = -Jpi(Xc);  J_X = P * R;  J_xi = P * [-[Xc]x | I].
inline void edge_jac(const Pose& T, const double Xc[3], const Intr& K, double JX[6], double Jxi[12]) {
    const double x = Xc[0], y = Xc[1], z = Xc[2];
    double P[6] = {-(K.fx / z), -0.0, -(-K.fx * x / (z * z)),
                   -0.0, -(K.fy / z), -(-K.fy * y / (z * z))};
    if (K.model == 1) {
        double J[6];
        kb8_jac(K, Xc, J);
        for (int i = 0; i < 6; ++i) P[i] = -J[i];
    }
    double R[9];
    R_from_quat(T.q, R);
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c)
            JX[3 * r + c] = P[3 * r] * R[c] + P[3 * r + 1] * R[3 + c] + P[3 * r + 2] * R[6 + c];
    const double D[18] = {0, z, -y, 1, 0, 0,
                          -z, 0, x, 0, 1, 0,
                          y, -x, 0, 0, 0, 1};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 6; ++c)
            Jxi[6 * r + c] = P[3 * r] * D[c] + P[3 * r + 1] * D[6 + c] + P[3 * r + 2] * D[12 + c];
}

inline double round_f32(double v) { return (double)(float)v; }

// The prepared problem: the graph of BE_GraphOptim.cpp:79-211 in flat form.
struct Graph {
    int n_cam = 0, n_pt = 0;
    Intr K{};
    bool robust = false;
    double delta = 0;
    std::vector<Pose> pose;          // per camera
    std::vector<uint8_t> cam_fixed;  // per camera
    std::vector<double> X;           // per point, 3
    std::vector<uint8_t> pt_fixed;
    std::vector<uint8_t> pt_active;  // has >= 1 edge (BE_GraphOptim.cpp:204-209)
    // edges in g2o edge order: points ascending, then cameras ascending, then input order
    std::vector<int> e_cam, e_pt, e_orig;
    std::vector<double> e_uv;        // 2 per edge
    std::vector<int> pt_begin;       // edge range per point (n_pt + 1)
    // Hessian index mapping (SparseOptimizer::buildIndexMapping): free active
    // cameras first in camera order, then free active points in point order.
    std::vector<int> cam_idx, pt_idx;
    int n_fc = 0, n_fp = 0;
};

int build_graph(const nav_ba_problem* p, const nav_ba_options* o, Graph& g) {
    if (!p || !o || p->n_cam < 0 || p->n_pt < 0 || p->n_obs < 0) return NAV_BA_BAD_ARG;
    const bool compat = o->compat_fp32_io != 0;
    g.n_cam = p->n_cam; g.n_pt = p->n_pt;
    const bool kb = p->camera_model == 1;
    if (p->camera_model != 0 && !kb) return NAV_BA_BAD_ARG;
    g.K = {p->fx, p->fy, p->cx, p->cy, kb ? 1 : 0, kb ? p->k1 : 0.0, kb ? p->k2 : 0.0, kb ? p->k3 : 0.0,
           kb ? p->k4 : 0.0};
    if (compat)  // every calibration parameter is a float (GeometricCamera.h:117)
        g.K = {round_f32(p->fx), round_f32(p->fy), round_f32(p->cx), round_f32(p->cy), g.K.model,
               round_f32(g.K.k1), round_f32(g.K.k2), round_f32(g.K.k3), round_f32(g.K.k4)};
    g.robust = o->robust != 0;
    g.delta = o->huber_delta;
    g.pose.resize(g.n_cam);
    g.cam_fixed.assign(g.n_cam, 0);
    for (int c = 0; c < g.n_cam; ++c) {
        const double* T = p->T_cw + 12 * c;
        g.pose[c].q = quat_from_R(T);
        quat_normalize_pos(g.pose[c].q);  // SE3Quat(R, t) constructor (BE_GraphOptim.cpp:92-93)
        g.pose[c].t[0] = T[9]; g.pose[c].t[1] = T[10]; g.pose[c].t[2] = T[11];
        if (p->cam_fixed) g.cam_fixed[c] = p->cam_fixed[c] ? 1 : 0;
    }
    g.X.resize(3 * (size_t)g.n_pt);
    g.pt_fixed.assign(g.n_pt, 0);
    for (int i = 0; i < g.n_pt; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double v = p->X[3 * (size_t)i + k];
            g.X[3 * (size_t)i + k] = compat ? round_f32(v) : v;  // BE_GraphOptim.cpp:147
        }
        if (p->pt_fixed) g.pt_fixed[i] = p->pt_fixed[i] ? 1 : 0;
    }
    // edges, stable-sorted by (point, camera)
    std::vector<int> order(p->n_obs);
    for (int i = 0; i < p->n_obs; ++i) {
        if (p->obs_cam[i] < 0 || p->obs_cam[i] >= g.n_cam || p->obs_pt[i] < 0 || p->obs_pt[i] >= g.n_pt)
            return NAV_BA_BAD_ARG;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (p->obs_pt[a] != p->obs_pt[b]) return p->obs_pt[a] < p->obs_pt[b];
        return p->obs_cam[a] < p->obs_cam[b];
    });
    const size_t ne = order.size();
    g.e_cam.resize(ne); g.e_pt.resize(ne); g.e_orig.resize(ne); g.e_uv.resize(2 * ne);
    g.pt_begin.assign(g.n_pt + 1, 0);
    std::vector<int> cam_cnt(g.n_cam, 0);
    for (size_t k = 0; k < ne; ++k) {
        const int i = order[k];
        g.e_cam[k] = p->obs_cam[i]; g.e_pt[k] = p->obs_pt[i]; g.e_orig[k] = i;
        for (int d = 0; d < 2; ++d) {
            const double v = p->obs_uv[2 * (size_t)i + d];
            g.e_uv[2 * k + d] = compat ? round_f32(v) : v;  // cv::Point2f (Point2D.hpp:33-34)
        }
        g.pt_begin[g.e_pt[k] + 1]++;
        cam_cnt[g.e_cam[k]]++;
    }
    for (int i = 0; i < g.n_pt; ++i) g.pt_begin[i + 1] += g.pt_begin[i];
    g.pt_active.assign(g.n_pt, 0);
    for (int i = 0; i < g.n_pt; ++i) g.pt_active[i] = g.pt_begin[i + 1] > g.pt_begin[i];
    // index mapping: a vertex takes part if it is free and has an active edge
    // (an edge is active when not all of its vertices are fixed)
    std::vector<uint8_t> cam_in(g.n_cam, 0);
    for (size_t k = 0; k < ne; ++k)
        if (!(g.cam_fixed[g.e_cam[k]] && g.pt_fixed[g.e_pt[k]])) cam_in[g.e_cam[k]] = 1;
    g.cam_idx.assign(g.n_cam, -1);
    g.n_fc = 0;
    for (int c = 0; c < g.n_cam; ++c)
        if (!g.cam_fixed[c] && cam_in[c]) g.cam_idx[c] = g.n_fc++;
    g.pt_idx.assign(g.n_pt, -1);
    g.n_fp = 0;
    for (int i = 0; i < g.n_pt; ++i)
        if (!g.pt_fixed[i] && g.pt_active[i]) g.pt_idx[i] = g.n_fp++;
    int n_act_pts = 0;
    for (int i = 0; i < g.n_pt; ++i) n_act_pts += g.pt_active[i];
    (void)cam_cnt;
    if (g.n_cam == 0 || n_act_pts == 0) return NAV_BA_EMPTY;
    return NAV_BA_OK;
}

// activeRobustChi2 over all edges (edges with all vertices fixed included: they
// only add a constant).
// Order of the oracle's long sums: 0 as written (edges and points ascending);
// 1 reversed (the cost over edges and the Schur complement over points in
// descending order).  The reversed order is a test twin: the same mathematics
// with other roundings, which measures how far a problem's result moves with
// the order of its sums alone.
static int g_sum_order = 0;

double robust_chi2(const Graph& g, const std::vector<Pose>& pose, const std::vector<double>& X,
                   double* chi2_out = nullptr) {
    double chi = 0.0;
    double e[2], Xc[3], rho[3];
    const size_t ne = g.e_cam.size();
    for (size_t kk = 0; kk < ne; ++kk) {
        const size_t k = g_sum_order ? ne - 1 - kk : kk;
        edge_error(pose[g.e_cam[k]], &X[3 * (size_t)g.e_pt[k]], g.K, &g.e_uv[2 * k], e, Xc);
        const double c2 = e[0] * e[0] + e[1] * e[1];
        if (chi2_out) chi2_out[k] = c2;
        if (g.robust) { huber(c2, g.delta, rho); chi += rho[0]; }
        else chi += c2;
    }
    return chi;
}

// The linear system of BlockSolver<6,3>::buildSystem with
// BaseBinaryEdge::constructQuadraticForm (vertex 0 = point, vertex 1 = pose;
// robust weight rho' on both information and error, no rho'' term).
struct System {
    std::vector<double> Hpp, bp;  // per free camera: 36, 6
    std::vector<double> Hll, bl;  // per free point: 9, 3
    // Hpl blocks per (point, free camera), duplicates merged: point p's blocks
    // are hpl_cam/hpl_blk[hpl_begin[p] .. hpl_begin[p+1])
    std::vector<int> hpl_begin, hpl_cam;
    std::vector<double> hpl_blk;  // 18 per block, 6x3 row-major
};

void build_system(const Graph& g, const std::vector<Pose>& pose, const std::vector<double>& X, System& s) {
    s.Hpp.assign(36 * (size_t)g.n_fc, 0.0); s.bp.assign(6 * (size_t)g.n_fc, 0.0);
    s.Hll.assign(9 * (size_t)g.n_fp, 0.0); s.bl.assign(3 * (size_t)g.n_fp, 0.0);
    s.hpl_begin.assign(g.n_fp + 1, 0);
    s.hpl_cam.clear(); s.hpl_blk.clear();
    double e[2], Xc[3], JX[6], Jxi[12], rho[3];
    for (int p = 0; p < g.n_pt; ++p) {
        const int li = g.pt_idx[p];
        if (li >= 0) s.hpl_begin[li] = (int)s.hpl_cam.size();
        for (int k = g.pt_begin[p]; k < g.pt_begin[p + 1]; ++k) {
            const int c = g.e_cam[k];
            const int ci = g.cam_idx[c];
            const bool pt_free = li >= 0, cam_free = ci >= 0;
            if (!pt_free && !cam_free) continue;
            edge_error(pose[c], &X[3 * (size_t)p], g.K, &g.e_uv[2 * k], e, Xc);
            edge_jac(pose[c], Xc, g.K, JX, Jxi);
            double w = 1.0;
            if (g.robust) { huber(e[0] * e[0] + e[1] * e[1], g.delta, rho); w = rho[1]; }
            const double r[2] = {-e[0] * w, -e[1] * w};  // omega_r = -Omega e * rho'
            if (pt_free) {
                double* H = &s.Hll[9 * (size_t)li];
                double* b = &s.bl[3 * (size_t)li];
                for (int i = 0; i < 3; ++i) {
                    b[i] += JX[i] * r[0] + JX[3 + i] * r[1];
                    for (int j = 0; j < 3; ++j)
                        H[3 * i + j] += (JX[i] * w) * JX[j] + (JX[3 + i] * w) * JX[3 + j];
                }
                if (cam_free) {
                    double blk[18];
                    for (int i = 0; i < 6; ++i)
                        for (int j = 0; j < 3; ++j)
                            blk[3 * i + j] = (Jxi[i] * w) * JX[j] + (Jxi[6 + i] * w) * JX[3 + j];
                    const int nb = (int)s.hpl_cam.size();
                    if (nb > s.hpl_begin[li] && s.hpl_cam[nb - 1] == ci) {
                        for (int t = 0; t < 18; ++t) s.hpl_blk[18 * (size_t)(nb - 1) + t] += blk[t];
                    } else {
                        s.hpl_cam.push_back(ci);
                        s.hpl_blk.insert(s.hpl_blk.end(), blk, blk + 18);
                    }
                }
            }
            if (cam_free) {
                double* H = &s.Hpp[36 * (size_t)ci];
                double* b = &s.bp[6 * (size_t)ci];
                for (int i = 0; i < 6; ++i) {
                    b[i] += Jxi[i] * r[0] + Jxi[6 + i] * r[1];
                    for (int j = 0; j < 6; ++j)
                        H[6 * i + j] += (Jxi[i] * w) * Jxi[j] + (Jxi[6 + i] * w) * Jxi[6 + j];
                }
            }
        }
        if (li >= 0) s.hpl_begin[li + 1] = (int)s.hpl_cam.size();
    }
}

// Dense lower Cholesky in place (row-major n x n).  Returns false on a
// non-positive or non-finite pivot.
bool cholesky(std::vector<double>& A, int n) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double ljj = std::sqrt(d);
        A[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / ljj;
        }
    }
    return true;
}

void chol_solve(const std::vector<double>& L, int n, std::vector<double>& x) {
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
}

// BlockSolver<6,3>::solve: Schur complement onto the poses, dense Cholesky of
// the reduced camera system, back-substitution for the points.
// Fills S (n x n) and gs when asked; returns false if the factorisation fails.
bool schur_solve(const Graph& g, const System& s, double lambda, std::vector<double>& x,
                 std::vector<double>* S_out = nullptr, std::vector<double>* g_out = nullptr) {
    const int n = 6 * g.n_fc;
    std::vector<double> S((size_t)n * n, 0.0), coeff(n, 0.0);
    for (int c = 0; c < g.n_fc; ++c)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j)
                S[(size_t)(6 * c + i) * n + 6 * c + j] = s.Hpp[36 * (size_t)c + 6 * i + j] + (i == j ? lambda : 0.0);
    std::vector<double> Dinv(9 * (size_t)g.n_fp);
    for (int ll = 0; ll < g.n_fp; ++ll) {
        const int l = g_sum_order ? g.n_fp - 1 - ll : ll;
        double D[9];
        for (int t = 0; t < 9; ++t) D[t] = s.Hll[9 * (size_t)l + t] + ((t % 4 == 0) ? lambda : 0.0);
        double* Di = &Dinv[9 * (size_t)l];
        inv3(D, Di);
        const double* bl = &s.bl[3 * (size_t)l];
        double db[3];
        for (int i = 0; i < 3; ++i) db[i] = Di[3 * i] * bl[0] + Di[3 * i + 1] * bl[1] + Di[3 * i + 2] * bl[2];
        for (int a = s.hpl_begin[l]; a < s.hpl_begin[l + 1]; ++a) {
            const int i1 = s.hpl_cam[a];
            const double* Bi = &s.hpl_blk[18 * (size_t)a];
            double BD[18];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 3; ++c)
                    BD[3 * r + c] = Bi[3 * r] * Di[c] + Bi[3 * r + 1] * Di[3 + c] + Bi[3 * r + 2] * Di[6 + c];
            for (int r = 0; r < 6; ++r)
                coeff[6 * i1 + r] += Bi[3 * r] * db[0] + Bi[3 * r + 1] * db[1] + Bi[3 * r + 2] * db[2];
            for (int b = a; b < s.hpl_begin[l + 1]; ++b) {
                const int i2 = s.hpl_cam[b];
                const double* Bj = &s.hpl_blk[18 * (size_t)b];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c)
                        S[(size_t)(6 * i1 + r) * n + 6 * i2 + c] -=
                            BD[3 * r] * Bj[3 * c] + BD[3 * r + 1] * Bj[3 * c + 1] + BD[3 * r + 2] * Bj[3 * c + 2];
            }
        }
    }
    // mirror the upper block triangle (the part g2o fills) to the lower one
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) S[(size_t)i * n + j] = S[(size_t)j * n + i];
    std::vector<double> gs(n);
    for (int i = 0; i < n; ++i) gs[i] = s.bp[i] - coeff[i];
    if (S_out) *S_out = S;
    if (g_out) *g_out = gs;
    x.assign(n + 3 * (size_t)g.n_fp, 0.0);
    if (!cholesky(S, n)) return false;
    std::vector<double> xp = gs;
    chol_solve(S, n, xp);
    for (int i = 0; i < n; ++i) x[i] = xp[i];
    for (int l = 0; l < g.n_fp; ++l) {
        double cl[3] = {s.bl[3 * (size_t)l], s.bl[3 * (size_t)l + 1], s.bl[3 * (size_t)l + 2]};
        for (int a = s.hpl_begin[l]; a < s.hpl_begin[l + 1]; ++a) {
            const double* B = &s.hpl_blk[18 * (size_t)a];
            const double* d = &xp[6 * (size_t)s.hpl_cam[a]];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 6; ++r) cl[c] += B[3 * r + c] * -d[r];
        }
        const double* Di = &Dinv[9 * (size_t)l];
        for (int i = 0; i < 3; ++i)
            x[n + 3 * (size_t)l + i] = Di[3 * i] * cl[0] + Di[3 * i + 1] * cl[1] + Di[3 * i + 2] * cl[2];
    }
    return true;
}

void apply_update(const Graph& g, const std::vector<double>& x, std::vector<Pose>& pose, std::vector<double>& X) {
    const int n = 6 * g.n_fc;
    for (int c = 0; c < g.n_cam; ++c)
        if (g.cam_idx[c] >= 0) pose[c] = se3_oplus(&x[6 * (size_t)g.cam_idx[c]], pose[c]);
    for (int p = 0; p < g.n_pt; ++p)
        if (g.pt_idx[p] >= 0)
            for (int k = 0; k < 3; ++k) X[3 * (size_t)p + k] += x[n + 3 * (size_t)g.pt_idx[p] + k];
}

void write_back(const Graph& g, const nav_ba_problem* p, const nav_ba_options* o,
                const std::vector<Pose>& pose, const std::vector<double>& X, nav_ba_result* r) {
    if (r->T_cw) {
        for (int c = 0; c < g.n_cam; ++c) {
            double* T = r->T_cw + 12 * (size_t)c;
            R_from_quat(pose[c].q, T);  // BE_GraphOptim.cpp:220-223
            T[9] = pose[c].t[0]; T[10] = pose[c].t[1]; T[11] = pose[c].t[2];
        }
    }
    if (r->X) {
        for (int i = 0; i < g.n_pt; ++i)
            for (int k = 0; k < 3; ++k) {
                double v = g.pt_active[i] ? X[3 * (size_t)i + k] : p->X[3 * (size_t)i + k];
                // estimate().cast<float>() (BE_GraphOptim.cpp:239) for points in the graph
                if (g.pt_active[i] && o->compat_fp32_io) v = round_f32(v);
                r->X[3 * (size_t)i + k] = v;
            }
    }
}

double max_abs_diag(const Graph& g, const System& s) {
    double m = 0.0;
    for (int c = 0; c < g.n_fc; ++c)
        for (int j = 0; j < 6; ++j) m = std::max(std::fabs(s.Hpp[36 * (size_t)c + 7 * j]), m);
    for (int l = 0; l < g.n_fp; ++l)
        for (int j = 0; j < 3; ++j) m = std::max(std::fabs(s.Hll[9 * (size_t)l + 4 * j]), m);
    return m;
}

}  // namespace

extern "C" {

void oracle_set_sum_order(int32_t order) { g_sum_order = order ? 1 : 0; }

void oracle_quat_from_R(const double R[9], double q[4]) {
    Quat r = quat_from_R(R);
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}

void oracle_R_from_quat(const double q[4], double R[9]) { R_from_quat(Quat{q[0], q[1], q[2], q[3]}, R); }

void oracle_se3_exp(const double xi[6], double q[4], double t[3]) {
    Pose p = se3_exp(xi);
    q[0] = p.q.x; q[1] = p.q.y; q[2] = p.q.z; q[3] = p.q.w;
    t[0] = p.t[0]; t[1] = p.t[1]; t[2] = p.t[2];
}

void oracle_se3_oplus(const double xi[6], const double q_in[4], const double t_in[3],
                      double q_out[4], double t_out[3]) {
    Pose T{Quat{q_in[0], q_in[1], q_in[2], q_in[3]}, {t_in[0], t_in[1], t_in[2]}};
    Pose r = se3_oplus(xi, T);
    q_out[0] = r.q.x; q_out[1] = r.q.y; q_out[2] = r.q.z; q_out[3] = r.q.w;
    t_out[0] = r.t[0]; t_out[1] = r.t[1]; t_out[2] = r.t[2];
}

void oracle_huber(double s, double delta, double rho[3]) { huber(s, delta, rho); }

void oracle_inv3(const double A[9], double Ainv[9]) { inv3(A, Ainv); }

void oracle_edge(const double q[4], const double t[3], const double X[3],
                 double fx, double fy, double cx, double cy, const double uv[2],
                 double e[2], double J_X[6], double J_xi[12], double Xc[3]) {
    Pose T{Quat{q[0], q[1], q[2], q[3]}, {t[0], t[1], t[2]}};
    Intr K{fx, fy, cx, cy, 0, 0, 0, 0, 0};
    edge_error(T, X, K, uv, e, Xc);
    edge_jac(T, Xc, K, J_X, J_xi);
}

void oracle_edge_kb8(const double q[4], const double t[3], const double X[3], const double K8[8],
                     const double uv[2], double e[2], double J_X[6], double J_xi[12], double Xc[3]) {
    Pose T{Quat{q[0], q[1], q[2], q[3]}, {t[0], t[1], t[2]}};
    Intr K{K8[0], K8[1], K8[2], K8[3], 1, K8[4], K8[5], K8[6], K8[7]};
    edge_error(T, X, K, uv, e, Xc);
    edge_jac(T, Xc, K, J_X, J_xi);
}

double oracle_cost(const nav_ba_problem* p, const nav_ba_options* o, const double* T_cw, const double* X,
                   double* chi2_per_obs, uint8_t* outlier) {
    Graph g;
    nav_ba_problem q = *p;
    q.T_cw = T_cw; q.X = X;
    nav_ba_options oo = *o;
    oo.compat_fp32_io = 0;  // evaluate exactly at the given state
    if (build_graph(&q, &oo, g) == NAV_BA_BAD_ARG) return std::numeric_limits<double>::quiet_NaN();
    if (o->compat_fp32_io) {
        g.K = {round_f32(g.K.fx), round_f32(g.K.fy), round_f32(g.K.cx), round_f32(g.K.cy), g.K.model,
               round_f32(g.K.k1), round_f32(g.K.k2), round_f32(g.K.k3), round_f32(g.K.k4)};
        for (size_t k = 0; k < g.e_uv.size(); ++k) g.e_uv[k] = round_f32(g.e_uv[k]);
    }
    // Evaluate with the given rotation matrices' quaternions.
    std::vector<double> c2(g.e_cam.size());
    const double chi = robust_chi2(g, g.pose, g.X, c2.data());
    for (size_t k = 0; k < g.e_cam.size(); ++k) {
        const int i = g.e_orig[k];
        if (chi2_per_obs) chi2_per_obs[i] = c2[k];
        if (outlier) {
            double e[2], Xc[3];
            edge_error(g.pose[g.e_cam[k]], &g.X[3 * (size_t)g.e_pt[k]], g.K, &g.e_uv[2 * k], e, Xc);
            outlier[i] = (c2[k] > o->outlier_chi2 || !(Xc[2] > 0.0)) ? 1 : 0;
        }
    }
    return chi;
}

double oracle_linear_blocks(const nav_ba_problem* p, const nav_ba_options* o,
                            double* Hpp, double* bp, double* Hll, double* bl) {
    Graph g;
    if (build_graph(p, o, g) != NAV_BA_OK) return std::numeric_limits<double>::quiet_NaN();
    System s;
    build_system(g, g.pose, g.X, s);
    if (Hpp) std::memset(Hpp, 0, sizeof(double) * 36 * (size_t)g.n_cam);
    if (bp) std::memset(bp, 0, sizeof(double) * 6 * (size_t)g.n_cam);
    if (Hll) std::memset(Hll, 0, sizeof(double) * 9 * (size_t)g.n_pt);
    if (bl) std::memset(bl, 0, sizeof(double) * 3 * (size_t)g.n_pt);
    for (int c = 0; c < g.n_cam; ++c) {
        const int ci = g.cam_idx[c];
        if (ci < 0) continue;
        if (Hpp) std::memcpy(Hpp + 36 * (size_t)c, &s.Hpp[36 * (size_t)ci], 36 * sizeof(double));
        if (bp) std::memcpy(bp + 6 * (size_t)c, &s.bp[6 * (size_t)ci], 6 * sizeof(double));
    }
    for (int i = 0; i < g.n_pt; ++i) {
        const int li = g.pt_idx[i];
        if (li < 0) continue;
        if (Hll) std::memcpy(Hll + 9 * (size_t)i, &s.Hll[9 * (size_t)li], 9 * sizeof(double));
        if (bl) std::memcpy(bl + 3 * (size_t)i, &s.bl[3 * (size_t)li], 3 * sizeof(double));
    }
    return robust_chi2(g, g.pose, g.X);
}

int32_t oracle_reduced_system(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                              double* S, double* gv, int32_t max_n) {
    Graph g;
    if (build_graph(p, o, g) != NAV_BA_OK) return -1;
    System s;
    build_system(g, g.pose, g.X, s);
    std::vector<double> x, So, go;
    schur_solve(g, s, lambda, x, &So, &go);
    const int n = 6 * g.n_fc;
    if (n > max_n) return n;
    if (S) std::memcpy(S, So.data(), sizeof(double) * (size_t)n * n);
    if (gv) std::memcpy(gv, go.data(), sizeof(double) * (size_t)n);
    return n;
}

int32_t oracle_schur_step(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                          double* dx, int32_t max_n) {
    Graph g;
    if (build_graph(p, o, g) != NAV_BA_OK) return -1;
    System s;
    build_system(g, g.pose, g.X, s);
    std::vector<double> x;
    const bool ok = schur_solve(g, s, lambda, x);
    const int n = 6 * g.n_fc + 3 * g.n_fp;
    if (!ok) return -1;
    if (n <= max_n) std::memcpy(dx, x.data(), sizeof(double) * (size_t)n);
    return n;
}

int32_t oracle_dense_step(const nav_ba_problem* p, const nav_ba_options* o, double lambda,
                          double* dx, int32_t max_n) {
    Graph g;
    if (build_graph(p, o, g) != NAV_BA_OK) return -1;
    System s;
    build_system(g, g.pose, g.X, s);
    const int np = 6 * g.n_fc, n = np + 3 * g.n_fp;
    if (n > max_n) return n;
    std::vector<long double> H((size_t)n * n, 0.0L), b(n, 0.0L);
    for (int c = 0; c < g.n_fc; ++c)
        for (int i = 0; i < 6; ++i) {
            b[6 * c + i] = s.bp[6 * (size_t)c + i];
            for (int j = 0; j < 6; ++j) H[(size_t)(6 * c + i) * n + 6 * c + j] = s.Hpp[36 * (size_t)c + 6 * i + j];
        }
    for (int l = 0; l < g.n_fp; ++l)
        for (int i = 0; i < 3; ++i) {
            b[np + 3 * l + i] = s.bl[3 * (size_t)l + i];
            for (int j = 0; j < 3; ++j) H[(size_t)(np + 3 * l + i) * n + np + 3 * l + j] = s.Hll[9 * (size_t)l + 3 * i + j];
            for (int a = s.hpl_begin[l]; a < s.hpl_begin[l + 1]; ++a) {
                const int c = s.hpl_cam[a];
                for (int r = 0; r < 6; ++r) {
                    const long double v = s.hpl_blk[18 * (size_t)a + 3 * r + i];
                    H[(size_t)(6 * c + r) * n + np + 3 * l + i] = v;
                    H[(size_t)(np + 3 * l + i) * n + 6 * c + r] = v;
                }
            }
        }
    for (int i = 0; i < n; ++i) H[(size_t)i * n + i] += lambda;
    // Gaussian elimination with partial pivoting in long double
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int m = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs((double)H[(size_t)i * n + k]) > std::fabs((double)H[(size_t)m * n + k])) m = i;
        if (H[(size_t)m * n + k] == 0.0L) return -1;
        if (m != k) {
            for (int j = 0; j < n; ++j) std::swap(H[(size_t)k * n + j], H[(size_t)m * n + j]);
            std::swap(b[k], b[m]);
        }
        for (int i = k + 1; i < n; ++i) {
            const long double f = H[(size_t)i * n + k] / H[(size_t)k * n + k];
            if (f == 0.0L) continue;
            for (int j = k; j < n; ++j) H[(size_t)i * n + j] -= f * H[(size_t)k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        long double acc = b[i];
        for (int j = i + 1; j < n; ++j) acc -= H[(size_t)i * n + j] * (long double)dx[j];
        dx[i] = (double)(acc / H[(size_t)i * n + i]);
    }
    return n;
}

// SparseOptimizer::optimize(n) with OptimizationAlgorithmLevenberg::solve per
// iteration (g2o): lambda init (user or tau*max diag, tau = 1e-5), up to
// max_trials trials, rho = (chi0 - chi1) / (x'(lambda x + b) + 1e-3), good step
// lambda *= max(1/3, min(2/3, 1 - (2 rho - 1)^3)), bad step lambda *= nu, nu *= 2,
// Terminate when the trial limit is reached or rho == 0.
static int32_t solve_impl(const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r,
                          double* lambda_history, int32_t* trials_history, double* trace, int32_t max_trace,
                          int32_t* n_trace) {
    if (n_trace) *n_trace = 0;
    if (!p || !o || !r) return NAV_BA_BAD_ARG;
    Graph g;
    const int st = build_graph(p, o, g);
    if (st == NAV_BA_BAD_ARG) return st;
    r->iterations = 0; r->trials = 0; r->stop_reason = NAV_BA_STOP_MAX_ITER;
    r->n_outliers = 0; r->final_lambda = 0; r->solve_ms = 0;
    if (st == NAV_BA_EMPTY) {
        // early return before optimisation (BE_GraphOptim.cpp:59-63): nothing written back
        if (r->T_cw) std::memcpy(r->T_cw, p->T_cw, sizeof(double) * 12 * (size_t)p->n_cam);
        if (r->X) std::memcpy(r->X, p->X, sizeof(double) * 3 * (size_t)p->n_pt);
        if (r->outlier) std::memset(r->outlier, 0, (size_t)p->n_obs);
        r->stop_reason = NAV_BA_STOP_EMPTY;
        r->initial_cost = r->final_cost = 0.0;
        return NAV_BA_EMPTY;
    }
    const int max_trials = o->max_trials > 0 ? o->max_trials : 10;
    std::vector<Pose> pose = g.pose;
    std::vector<double> X = g.X;
    double lambda = 0.0, nu = 2.0;
    double chi0 = robust_chi2(g, pose, X);
    r->initial_cost = chi0;
    if (r->cost_history) r->cost_history[0] = chi0;
    System s;
    std::vector<double> x;
    for (int it = 0; it < o->n_iter; ++it) {
        double currentChi = robust_chi2(g, pose, X);
        build_system(g, pose, X, s);
        if (it == 0) {
            lambda = o->lambda_init > 0 ? o->lambda_init : 1e-5 * max_abs_diag(g, s);
            nu = 2.0;
        }
        if (lambda_history) lambda_history[it] = lambda;
        double rho = 0.0;
        int q = 0;
        do {
            std::vector<Pose> pose_bak = pose;
            std::vector<double> X_bak = X;
            const bool ok = schur_solve(g, s, lambda, x);
            if (!ok) std::fill(x.begin(), x.end(), 0.0);
            apply_update(g, x, pose, X);
            double tempChi = robust_chi2(g, pose, X);
            if (!ok) tempChi = std::numeric_limits<double>::max();
            // computeScale over the full x = [poses; points], b = [bp; bl]
            double scale = 0.0;
            const int npd = 6 * g.n_fc;
            for (int j = 0; j < npd; ++j) scale += x[j] * (lambda * x[j] + s.bp[j]);
            for (int j = 0; j < 3 * g.n_fp; ++j) scale += x[npd + j] * (lambda * x[npd + j] + s.bl[j]);
            scale += 1e-3;
            rho = (currentChi - tempChi) / scale;
            if (trace && n_trace && *n_trace < max_trace) {
                double* tr = trace + 7 * (size_t)(*n_trace)++;
                tr[0] = it; tr[1] = q; tr[2] = lambda; tr[3] = nu; tr[4] = rho; tr[5] = currentChi;
                tr[6] = tempChi;
            }
            if (rho > 0 && std::isfinite(tempChi)) {
                double alpha = 1. - std::pow((2 * rho - 1), 3);
                alpha = std::min(alpha, 2. / 3.);
                const double scaleFactor = std::max(1. / 3., alpha);
                lambda *= scaleFactor;
                nu = 2;
                currentChi = tempChi;
            } else {
                lambda *= nu;
                nu *= 2;
                pose = pose_bak;
                X = X_bak;
            }
            ++q;
        } while (rho < 0 && q < max_trials);
        r->trials += q;
        if (trials_history) trials_history[it] = q;
        r->iterations = it + 1;
        if (r->cost_history) r->cost_history[it + 1] = currentChi;
        if (q == max_trials || rho == 0) {
            r->stop_reason = NAV_BA_STOP_TERMINATE;
            break;
        }
    }
    r->final_lambda = lambda;
    r->final_cost = robust_chi2(g, pose, X);
    write_back(g, p, o, pose, X, r);
    // outlier flags at the final state (FP64): chi2 > threshold or depth <= 0
    int nout = 0;
    for (size_t k = 0; k < g.e_cam.size(); ++k) {
        double e[2], Xc[3];
        edge_error(pose[g.e_cam[k]], &X[3 * (size_t)g.e_pt[k]], g.K, &g.e_uv[2 * k], e, Xc);
        const double c2 = e[0] * e[0] + e[1] * e[1];
        const uint8_t f = (c2 > o->outlier_chi2 || !(Xc[2] > 0.0)) ? 1 : 0;
        if (r->outlier) r->outlier[g.e_orig[k]] = f;
        nout += f;
    }
    r->n_outliers = nout;
    return NAV_BA_OK;
}

int32_t oracle_ba_solve(const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r,
                        double* lambda_history, int32_t* trials_history) {
    return solve_impl(p, o, r, lambda_history, trials_history, nullptr, 0, nullptr);
}

int32_t oracle_ba_solve_trace(const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r,
                              double* trace, int32_t max_trace, int32_t* n_trace) {
    return solve_impl(p, o, r, nullptr, nullptr, trace, max_trace, n_trace);
}

}  // extern "C"
