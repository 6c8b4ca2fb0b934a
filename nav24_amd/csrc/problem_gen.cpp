// problem_gen.cpp — seeded synthetic bundle-adjustment problems (see problem_gen.h).
#include "problem_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Rng {
    std::mt19937_64 g;
    bool have = false;
    double spare = 0.0;
    explicit Rng(uint64_t s) : g(s) {}
    double u01() { return (double)(g() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * u01(); }
    int64_t randint(int64_t a, int64_t b) { return a + (int64_t)(g() % (uint64_t)(b - a + 1)); }
    // Box-Muller, both outputs used in turn.
    double normal() {
        if (have) { have = false; return spare; }
        const double u1 = 1.0 - u01();  // (0, 1]
        const double u2 = u01();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = 2.0 * kPi * u2;
        spare = r * std::sin(a);
        have = true;
        return r * std::cos(a);
    }
};

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a) {
    const double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / n, a.y / n, a.z / n};
}

// T_cw for a camera at centre C looking along f with world up +z; camera axes
// x right, y down, z forward (x = f cross up, y = f cross x).
void look_pose(V3 C, V3 f, double T[12]) {
    f = normalize(f);
    const V3 up{0, 0, 1};
    const V3 xc = normalize(cross(f, up));
    const V3 yc = cross(f, xc);
    const V3 rows[3] = {xc, yc, f};
    for (int i = 0; i < 3; ++i) {
        T[3 * i + 0] = rows[i].x; T[3 * i + 1] = rows[i].y; T[3 * i + 2] = rows[i].z;
        T[9 + i] = -(rows[i].x * C.x + rows[i].y * C.y + rows[i].z * C.z);
    }
}

// Rodrigues exponential of a rotation vector.
void rodrigues(const double w[3], double R[9]) {
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double K2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
    double a = 1.0, b = 0.5;
    if (th > 1e-12) { a = std::sin(th) / th; b = (1 - std::cos(th)) / (th * th); }
    for (int k = 0; k < 9; ++k) R[k] = ((k % 4 == 0) ? 1.0 : 0.0) + a * K[k] + b * K2[k];
}

struct Cam { double fx, fy, cx, cy, k1, k2, p1, p2; int w, h; bool dist; };

constexpr double kLoopMaxDepth = 25.0;  // metres: about 3 M observations for 150 000 points
constexpr double kLoopMinDepth = 1.0;   // metres: no feature tracked closer than this

// Pinhole (optionally radial-tangential / Brown-Conrady) projection.
bool project(const Cam& K, const double T[12], const double X[3], double uv[2], double min_depth,
             double max_depth = 1e300) {
    double Xc[3];
    for (int i = 0; i < 3; ++i) Xc[i] = T[3 * i] * X[0] + T[3 * i + 1] * X[1] + T[3 * i + 2] * X[2] + T[9 + i];
    if (!(Xc[2] > min_depth) || Xc[2] > max_depth) return false;
    double x = Xc[0] / Xc[2], y = Xc[1] / Xc[2];
    if (K.dist) {
        const double r2 = x * x + y * y;
        const double rad = 1 + K.k1 * r2 + K.k2 * r2 * r2;
        const double xd = x * rad + 2 * K.p1 * x * y + K.p2 * (r2 + 2 * x * x);
        const double yd = y * rad + K.p1 * (r2 + 2 * y * y) + 2 * K.p2 * x * y;
        x = xd; y = yd;
    }
    uv[0] = K.fx * x + K.cx;
    uv[1] = K.fy * y + K.cy;
    return uv[0] >= 0 && uv[0] < K.w && uv[1] >= 0 && uv[1] < K.h;
}

template <class T> T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

void sample_point(Rng& r, int geometry, double X[3]) {
    if (geometry == NAV_GEN_CIRCLE) {
        for (int k = 0; k < 3; ++k) X[k] = r.uniform(-2.0, 2.0);
    } else if (geometry == NAV_GEN_RANDOM_WALK) {
        X[0] = r.uniform(-20.0, 20.0); X[1] = r.uniform(-20.0, 20.0); X[2] = r.uniform(-5.0, 5.0);
    } else {  // solid tube of radius 15 m around the loop circle of radius 30 m
        for (;;) {
            X[0] = r.uniform(-45.0, 45.0); X[1] = r.uniform(-45.0, 45.0); X[2] = r.uniform(-15.0, 15.0);
            const double d = std::sqrt(X[0] * X[0] + X[1] * X[1]) - 30.0;
            if (d * d + X[2] * X[2] <= 225.0) break;
        }
    }
}

}  // namespace

extern "C" {

int32_t nav_gen_params_for(int32_t workload, uint64_t seed, nav_gen_params* o) {
    if (!o) return 2;
    nav_gen_params p{};
    p.noise_px = 1.0; p.outlier_rate = 0.05;
    p.rot_sigma = 0.05; p.trans_sigma = 0.1; p.point_sigma = 0.05;
    p.n_fixed = 2; p.seed = seed;
    switch (workload) {
        case 0: p.geometry = NAV_GEN_CIRCLE; p.n_cam = 10; p.n_pt = 1500; break;
        case 1: p.geometry = NAV_GEN_RANDOM_WALK; p.n_cam = 60; p.n_pt = 20000; break;
        case 2: p.geometry = NAV_GEN_LOOP; p.n_cam = 300; p.n_pt = 150000; p.outlier_rate = 0.03; break;
        case 3: {
            Rng r(seed ^ 0x9E3779B97F4A7C15ULL);
            p.geometry = NAV_GEN_CIRCLE;
            p.n_cam = (int32_t)r.randint(5, 30);
            p.n_pt = (int32_t)r.randint(300, 5000);
            p.obs_min = 2; p.obs_max = 8;
            break;
        }
        case 4: p.geometry = NAV_GEN_CIRCLE; p.n_cam = 20; p.n_pt = 5000; p.distortion = 1; break;
        default: return 2;
    }
    *o = p;
    return 0;
}

int32_t nav_gen_make(const nav_gen_params* p, nav_gen_problem** out) {
    if (!p || !out || p->n_cam <= 0 || p->n_pt < 0) return 2;
    Rng r(p->seed);
    const int nc = p->n_cam, np = p->n_pt;
    Cam K{450.0, 450.0, 320.0, 240.0, 0, 0, 0, 0, 640, 480, p->distortion != 0};
    if (K.dist) { K.k1 = r.uniform(-0.1, 0.1); K.k2 = r.uniform(-0.1, 0.1); K.p1 = r.uniform(-0.1, 0.1); K.p2 = r.uniform(-0.1, 0.1); }
    // ground-truth poses
    std::vector<double> Tgt(12 * (size_t)nc);
    if (p->geometry == NAV_GEN_CIRCLE) {
        for (int k = 0; k < nc; ++k) {
            const double a = 2 * kPi * k / nc;
            const V3 C{2.0 * std::cos(a), 2.0 * std::sin(a), 0.0};
            look_pose(C, sub(V3{0, 0, 0}, C), &Tgt[12 * (size_t)k]);
        }
    } else if (p->geometry == NAV_GEN_RANDOM_WALK) {
        V3 C{0, 0, 0};
        double yaw = 0.0;
        for (int k = 0; k < nc; ++k) {
            if (k > 0) {
                C.x += 0.5 * r.normal(); C.y += 0.5 * r.normal();
                yaw += 0.1 * r.normal();
            }
            look_pose(C, V3{std::cos(yaw), std::sin(yaw), 0.0}, &Tgt[12 * (size_t)k]);
        }
    } else {
        for (int k = 0; k < nc; ++k) {
            const double a = 2 * kPi * k / nc;
            const double rad = 30.0 + 0.2 * r.normal();
            const V3 C{rad * std::cos(a), rad * std::sin(a), 0.0};
            look_pose(C, V3{-std::sin(a), std::cos(a), 0.0}, &Tgt[12 * (size_t)k]);
        }
    }
    // points and their observing cameras
    std::vector<double> Xgt(3 * (size_t)np);
    std::vector<std::vector<int>> seen(nc);
    std::vector<int> vis;
    vis.reserve(nc);
    double uv[2];
    // loop: features are tracked up to a range, so a point is seen from the
    // part of the loop near it (about 20 views per point)
    const double max_depth = p->geometry == NAV_GEN_LOOP ? kLoopMaxDepth : 1e300;
    const double min_depth = p->geometry == NAV_GEN_LOOP ? kLoopMinDepth : 0.1;
    for (int i = 0; i < np; ++i) {
        double* X = &Xgt[3 * (size_t)i];
        int m = 0;
        for (int attempt = 0;; ++attempt) {
            sample_point(r, p->geometry, X);
            vis.clear();
            for (int c = 0; c < nc; ++c)
                if (project(K, &Tgt[12 * (size_t)c], X, uv, min_depth, max_depth)) vis.push_back(c);
            if ((int)vis.size() >= 2 || attempt >= 1000) break;
        }
        if (p->obs_min > 0) {
            const int hi = std::min(p->obs_max, nc);
            const int want = (int)r.randint(p->obs_min, std::max(p->obs_min, hi));
            m = std::min(want, (int)vis.size());
            for (int k = 0; k < m; ++k) {  // partial Fisher-Yates
                const int j = k + (int)r.randint(0, (int)vis.size() - 1 - k);
                std::swap(vis[k], vis[j]);
            }
            std::sort(vis.begin(), vis.begin() + m);
        } else {
            m = (int)vis.size();
        }
        for (int k = 0; k < m; ++k) seen[vis[k]].push_back(i);
    }
    std::vector<int32_t> ocam, opt;
    for (int c = 0; c < nc; ++c)
        for (int i : seen[c]) { ocam.push_back(c); opt.push_back(i); }
    const size_t no = ocam.size();
    std::vector<double> ouv(2 * no);
    std::vector<uint8_t> oout(no, 0);
    for (size_t k = 0; k < no; ++k) {
        project(K, &Tgt[12 * (size_t)ocam[k]], &Xgt[3 * (size_t)opt[k]], uv, -1e300);
        if (r.u01() < p->outlier_rate) {
            uv[0] = r.uniform(0.0, K.w); uv[1] = r.uniform(0.0, K.h);
            oout[k] = 1;
        } else {
            uv[0] += p->noise_px * r.normal(); uv[1] += p->noise_px * r.normal();
        }
        ouv[2 * k] = (double)(float)uv[0];
        ouv[2 * k + 1] = (double)(float)uv[1];
    }
    // initial state: perturb free poses (R <- Exp(dw) R, t <- t + dt) and all points
    std::vector<double> T0 = Tgt;
    std::vector<uint8_t> cfix(nc, 0);
    for (int c = 0; c < nc; ++c) {
        if (c < p->n_fixed) { cfix[c] = 1; continue; }
        double w[3] = {p->rot_sigma * r.normal(), p->rot_sigma * r.normal(), p->rot_sigma * r.normal()};
        double dR[9];
        rodrigues(w, dR);
        double* T = &T0[12 * (size_t)c];
        double R[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                R[3 * i + j] = dR[3 * i] * T[j] + dR[3 * i + 1] * T[3 + j] + dR[3 * i + 2] * T[6 + j];
        std::memcpy(T, R, sizeof R);
        for (int i = 0; i < 3; ++i) T[9 + i] += p->trans_sigma * r.normal();
    }
    std::vector<double> X0(3 * (size_t)np);
    for (size_t k = 0; k < X0.size(); ++k) X0[k] = (double)(float)(Xgt[k] + p->point_sigma * r.normal());

    nav_gen_problem* g = (nav_gen_problem*)std::calloc(1, sizeof(nav_gen_problem));
    g->n_cam = nc; g->n_pt = np; g->n_obs = (int32_t)no;
    g->T_cw = dup(T0); g->cam_fixed = dup(cfix);
    g->X = dup(X0); g->pt_fixed = dup(std::vector<uint8_t>(np, 0));
    g->obs_cam = dup(ocam); g->obs_pt = dup(opt); g->obs_uv = dup(ouv); g->obs_outlier = dup(oout);
    g->T_gt = dup(Tgt); g->X_gt = dup(Xgt);
    g->fx = K.fx; g->fy = K.fy; g->cx = K.cx; g->cy = K.cy;
    g->dist[0] = K.k1; g->dist[1] = K.k2; g->dist[2] = K.p1; g->dist[3] = K.p2;
    g->width = K.w; g->height = K.h;
    *out = g;
    return 0;
}

void nav_gen_free(nav_gen_problem* g) {
    if (!g) return;
    std::free(g->T_cw); std::free(g->cam_fixed); std::free(g->X); std::free(g->pt_fixed);
    std::free(g->obs_cam); std::free(g->obs_pt); std::free(g->obs_uv); std::free(g->obs_outlier);
    std::free(g->T_gt); std::free(g->X_gt);
    std::free(g);
}

uint64_t nav_gen_mt_draw(uint64_t seed, int32_t k) {
    std::mt19937_64 g(seed);
    uint64_t v = 0;
    for (int i = 0; i <= k; ++i) v = g();
    return v;
}

void nav_gen_normals(uint64_t seed, int32_t n, double* out) {
    Rng r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.normal();
}

}  // extern "C"
