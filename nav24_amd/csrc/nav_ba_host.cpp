// nav_ba_host.cpp — host runtime and C ABI of the MI355X BA back-end.
//
// Replaces NAV24::BE::GraphOptim::solve (core/backEnd/BE_GraphOptim.cpp:31-77):
//   prepare  = graph build (addFrameVertices :79-100, addMapPointVertices :117-211)
//              as flat arrays: points with observations compacted in ascending id,
//              observations ordered by (point, camera, input order), camera-major
//              index lists, linearisation tiles, Schur work items, co-visibility.
//   run      = optimizer.optimize(nIterations) (:66-68) on the device; the LM
//              iteration and each trial are replayed hipGraphs, the host only
//              reads a 64-byte status per trial to decide on another trial.
//   fetch    = recoverPose / recoverMapPoints (:213-244) plus the outlier flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nav_ba.h"
#include "../../include/nav_ba_debug.h"
#include "ba_kernels.h"

using namespace navba;

namespace {

constexpr const char* kVersion = "nav24_amd-ba 0.1 (gfx950)";

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct Graphs {
    hipGraphExec_t iter0 = nullptr, iter = nullptr, trial = nullptr;
};

// HIP event pairs around stage launches; read back after a stream sync.
struct StageTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> cat;
    size_t used = 0;
    double ms[16] = {};
    int launches[16] = {};
    void begin(int c, hipStream_t s) {
        if (ev.size() < 2 * (used + 1)) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            ev.push_back(a);
            ev.push_back(b);
            cat.push_back(0);
        }
        cat[used] = c;
        (void)hipEventRecord(ev[2 * used], s);
    }
    void end(hipStream_t s) { (void)hipEventRecord(ev[2 * used + 1], s); ++used; }
    void collect() {
        for (size_t i = 0; i < used; ++i) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) == hipSuccess) {
                ms[cat[i]] += t;
                launches[cat[i]] += 1;
            }
        }
        used = 0;
    }
    void clear() {
        used = 0;
        for (int i = 0; i < 16; ++i) { ms[i] = 0; launches[i] = 0; }
    }
    ~StageTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

}  // namespace

struct nav_ba_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bool prepared = false;
    nav_ba_options opt{};
    // host copy of the input shape and what fetch needs
    int n_cam = 0, n_pt_all = 0, n_obs = 0;
    bool empty = false;
    std::vector<int> pt_orig;          // local -> caller point index
    std::vector<double> X_in;          // caller points (for inactive points)
    std::vector<double> T_in;
    DevProblem P{};
    DevState cur{}, cand{}, init{};
    DevWork W{};
    std::vector<DevBuf> bufs;
    int nbuf = 0;
    Graphs graphs;
    bool timing = false;               // stage-by-stage launches with events
    bool timing_active = false;
    StageTimer timer;
    LMDev* hstatus = nullptr;          // pinned
    double* hfinal = nullptr;          // pinned [2]
    double* dfinal = nullptr;
    // results of the last run
    bool ran = false;
    int iterations = 0, trials = 0, stop_reason = 0;
    double initial_cost = 0, final_cost = 0, final_lambda = 0, solve_ms = 0;
    int n_outliers = 0;
    std::vector<double> cost_history;
};

namespace {

#define HIPTRY(h, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return NAV_BA_HIP_ERROR;                                                       \
        }                                                                                  \
    } while (0)

int fail(nav_ba_handle* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

// Device allocation slot `i` of at least `bytes` (grow on demand, reused across prepares).
int alloc(nav_ba_handle* h, void** out, size_t bytes) {
    if ((int)h->bufs.size() <= h->nbuf) h->bufs.resize(h->nbuf + 1);
    DevBuf& b = h->bufs[h->nbuf++];
    bytes = std::max<size_t>(bytes, 256);
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            h->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
            return NAV_BA_HIP_ERROR;
        }
        b.cap = bytes;
    }
    *out = b.p;
    return NAV_BA_OK;
}

template <class T>
int upload(nav_ba_handle* h, const T** out, const std::vector<T>& v) {
    void* p = nullptr;
    int st = alloc(h, &p, sizeof(T) * v.size());
    if (st) return st;
    if (!v.empty()) HIPTRY(h, hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, h->stream));
    *out = (const T*)p;
    return NAV_BA_OK;
}

// Shoemake rotation -> quaternion, then g2o's w >= 0 unit normalisation
// (SE3Quat(R, t) constructor used at BE_GraphOptim.cpp:92-93).
void host_quat_from_R(const double* R, double q[4]) {
    auto m = [&](int i, int j) { return R[3 * i + j]; };
    const double tr = m(0, 0) + m(1, 1) + m(2, 2);
    if (tr > 0.0) {
        double s = std::sqrt(tr + 1.0);
        q[3] = 0.5 * s;
        s = 0.5 / s;
        q[0] = (m(2, 1) - m(1, 2)) * s;
        q[1] = (m(0, 2) - m(2, 0)) * s;
        q[2] = (m(1, 0) - m(0, 1)) * s;
    } else {
        int i = 0;
        if (m(1, 1) > m(0, 0)) i = 1;
        if (m(2, 2) > m(i, i)) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double s = std::sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
        double c[3];
        c[i] = 0.5 * s;
        s = 0.5 / s;
        q[3] = (m(k, j) - m(j, k)) * s;
        c[j] = (m(j, i) + m(i, j)) * s;
        c[k] = (m(k, i) + m(i, k)) * s;
        q[0] = c[0]; q[1] = c[1]; q[2] = c[2];
    }
    if (q[3] < 0) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}

void host_R_from_quat(const double q[4], double R[9]) {
    const double tx = 2 * q[0], ty = 2 * q[1], tz = 2 * q[2];
    const double twx = tx * q[3], twy = ty * q[3], twz = tz * q[3];
    const double txx = tx * q[0], txy = ty * q[0], txz = tz * q[0];
    const double tyy = ty * q[1], tyz = tz * q[1], tzz = tz * q[2];
    R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
    R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

inline double rf32(double v) { return (double)(float)v; }

void destroy_graphs(nav_ba_handle* h) {
    if (h->graphs.iter0) (void)hipGraphExecDestroy(h->graphs.iter0);
    if (h->graphs.iter) (void)hipGraphExecDestroy(h->graphs.iter);
    if (h->graphs.trial) (void)hipGraphExecDestroy(h->graphs.trial);
    h->graphs = Graphs{};
}

bool mixed(const nav_ba_handle* h) { return h->opt.precision == NAV_BA_MIXED; }

// Enqueue one stage; with timing active it is bracketed by HIP events.
template <class F>
void stage(nav_ba_handle* h, int cat, F&& f) {
    if (h->timing_active) h->timer.begin(cat, h->stream);
    f();
    if (h->timing_active) h->timer.end(h->stream);
}

template <typename T>
void enqueue_lin(nav_ba_handle* h, bool lambda_init) {
    stage(h, NAV_BA_ST_LIN_PT, [&] { launch_lin_pt<T>(h->P, h->cur, h->W, h->stream); });
    stage(h, NAV_BA_ST_LIN_CAM, [&] { launch_lin_cam<T>(h->P, h->cur, h->W, h->stream); });
    if (lambda_init) stage(h, NAV_BA_ST_LAMBDA, [&] { launch_lambda_init<T>(h->P, h->W, h->stream); });
}

template <typename T>
void enqueue_trial(nav_ba_handle* h) {
    stage(h, NAV_BA_ST_HINV, [&] { launch_hinv<T>(h->P, h->W, h->stream); });
    if (h->P.n_fc > 0) {
        stage(h, NAV_BA_ST_SCHUR_ROW, [&] { launch_schur_row<T>(h->P, h->cur, h->W, h->stream); });
        stage(h, NAV_BA_ST_SCHUR_REDUCE, [&] { launch_schur_reduce<T>(h->P, h->W, h->stream); });
        stage(h, NAV_BA_ST_CHOLESKY, [&] { launch_cholesky_solve(h->P, h->W, h->stream); });
    }
    stage(h, NAV_BA_ST_BACKSUB, [&] { launch_backsub<T>(h->P, h->cur, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_UPDATE_CAM, [&] { launch_update_cam(h->P, h->cur, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_COST, [&] { launch_cost<T>(h->P, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_ACCEPT, [&] {
        launch_lm_accept(h->W, h->stream);
        launch_commit(h->P, h->cur, h->cand, h->W, h->stream);
    });
    (void)hipMemcpyAsync(h->hstatus, h->W.lm, sizeof(LMDev), hipMemcpyDeviceToHost, h->stream);
}

// One outer iteration (which = 0: with lambda initialisation, 1: without) or
// one more trial (which = 2): graph replay, or stage-by-stage when timing.
int launch_step(nav_ba_handle* h, int which) {
    if (!h->timing) {
        hipGraphExec_t g = which == 0 ? h->graphs.iter0 : which == 1 ? h->graphs.iter : h->graphs.trial;
        HIPTRY(h, hipGraphLaunch(g, h->stream));
        return NAV_BA_OK;
    }
    h->timing_active = true;
    const bool m = mixed(h);
    if (which == 0 || which == 1) {
        if (m) enqueue_lin<float>(h, which == 0); else enqueue_lin<double>(h, which == 0);
    }
    if (m) enqueue_trial<float>(h); else enqueue_trial<double>(h);
    h->timing_active = false;
    return NAV_BA_OK;
}

int capture(nav_ba_handle* h, hipGraphExec_t* out, int which) {
    hipGraph_t g = nullptr;
    HIPTRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const bool m = mixed(h);
    if (which == 0 || which == 1) {
        if (m) enqueue_lin<float>(h, which == 0); else enqueue_lin<double>(h, which == 0);
    }
    if (m) enqueue_trial<float>(h); else enqueue_trial<double>(h);
    HIPTRY(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(h, NAV_BA_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return NAV_BA_OK;
}

int validate(const nav_ba_problem* p) {
    if (!p || p->n_cam < 0 || p->n_pt < 0 || p->n_obs < 0) return NAV_BA_BAD_ARG;
    if ((p->n_cam > 0 && !p->T_cw) || (p->n_pt > 0 && !p->X)) return NAV_BA_BAD_ARG;
    if (p->n_obs > 0 && (!p->obs_cam || !p->obs_pt || !p->obs_uv)) return NAV_BA_BAD_ARG;
    for (int i = 0; i < p->n_obs; ++i)
        if (p->obs_cam[i] < 0 || p->obs_cam[i] >= p->n_cam || p->obs_pt[i] < 0 || p->obs_pt[i] >= p->n_pt)
            return NAV_BA_BAD_ARG;
    return NAV_BA_OK;
}

}  // namespace

extern "C" {

const char* nav_ba_version(void) { return kVersion; }

void nav_ba_default_options(nav_ba_options* o) {
    if (!o) return;
    o->n_iter = 5;
    o->robust = 0;
    o->huber_delta = (double)(float)std::sqrt(5.99);
    o->lambda_init = 0.0;
    o->precision = NAV_BA_FP64;
    o->compat_fp32_io = 1;
    o->outlier_chi2 = 5.991;
    o->max_trials = 10;
}

int32_t nav_ba_create(nav_ba_handle** out, int32_t device) {
    if (!out) return NAV_BA_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return NAV_BA_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return NAV_BA_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return NAV_BA_NO_DEVICE;
    nav_ba_handle* h = new nav_ba_handle();
    h->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipHostMalloc((void**)&h->hstatus, sizeof(LMDev), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->hfinal, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        nav_ba_destroy(h);
        return NAV_BA_HIP_ERROR;
    }
    *out = h;
    return NAV_BA_OK;
}

void nav_ba_destroy(nav_ba_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    destroy_graphs(h);
    for (auto& b : h->bufs)
        if (b.p) (void)hipFree(b.p);
    if (h->hstatus) (void)hipHostFree(h->hstatus);
    if (h->hfinal) (void)hipHostFree(h->hfinal);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* nav_ba_last_error(const nav_ba_handle* h) { return h ? h->err.c_str() : "null handle"; }

int32_t nav_ba_prepare(nav_ba_handle* h, const nav_ba_problem* p, const nav_ba_options* o) {
    if (!h) return NAV_BA_BAD_ARG;
    h->err.clear();
    h->prepared = false;
    h->ran = false;
    if (!o) return fail(h, NAV_BA_BAD_ARG, "null options");
    if (validate(p) != NAV_BA_OK) return fail(h, NAV_BA_BAD_ARG, "invalid problem (null array or index out of range)");
    if (o->precision != NAV_BA_FP64 && o->precision != NAV_BA_MIXED) return fail(h, NAV_BA_BAD_ARG, "unknown precision");
    if (o->n_iter < 0) return fail(h, NAV_BA_BAD_ARG, "negative iteration count");
    HIPTRY(h, hipSetDevice(h->device));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    destroy_graphs(h);
    h->opt = *o;
    if (h->opt.max_trials <= 0) h->opt.max_trials = 10;
    const bool compat = o->compat_fp32_io != 0;
    const int nc = p->n_cam, npa = p->n_pt, no = p->n_obs;
    h->n_cam = nc; h->n_pt_all = npa; h->n_obs = no;
    h->T_in.assign(p->T_cw, p->T_cw + 12 * (size_t)nc);
    h->X_in.assign(p->X, p->X + 3 * (size_t)npa);

    // ---- points with observations, compacted in ascending caller index
    std::vector<int> cnt_pt(npa, 0);
    for (int i = 0; i < no; ++i) cnt_pt[p->obs_pt[i]]++;
    std::vector<int> pt_local(npa, -1);
    h->pt_orig.clear();
    for (int i = 0; i < npa; ++i)
        if (cnt_pt[i] > 0) { pt_local[i] = (int)h->pt_orig.size(); h->pt_orig.push_back(i); }
    const int np = (int)h->pt_orig.size();
    h->empty = (nc == 0 || np == 0);
    if (h->empty) {
        h->prepared = true;
        return NAV_BA_OK;
    }

    // ---- observations ordered by (point, camera, input order): two stable counting sorts
    std::vector<int> by_cam(no), ord(no);
    {
        std::vector<int> off(nc + 1, 0);
        for (int i = 0; i < no; ++i) off[p->obs_cam[i] + 1]++;
        for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
        for (int i = 0; i < no; ++i) by_cam[off[p->obs_cam[i]]++] = i;
        std::vector<int> poff(np + 1, 0);
        for (int i = 0; i < no; ++i) poff[pt_local[p->obs_pt[i]] + 1]++;
        for (int q = 0; q < np; ++q) poff[q + 1] += poff[q];
        for (int k = 0; k < no; ++k) {
            const int i = by_cam[k];
            ord[poff[pt_local[p->obs_pt[i]]]++] = i;
        }
    }
    std::vector<int> o_cam(no), o_pt(no), o_orig(no), pt_ptr(np + 1, 0);
    std::vector<double> o_uv(2 * (size_t)no);
    for (int k = 0; k < no; ++k) {
        const int i = ord[k];
        o_cam[k] = p->obs_cam[i];
        o_pt[k] = pt_local[p->obs_pt[i]];
        o_orig[k] = i;
        o_uv[2 * (size_t)k] = compat ? rf32(p->obs_uv[2 * (size_t)i]) : p->obs_uv[2 * (size_t)i];
        o_uv[2 * (size_t)k + 1] = compat ? rf32(p->obs_uv[2 * (size_t)i + 1]) : p->obs_uv[2 * (size_t)i + 1];
        pt_ptr[o_pt[k] + 1]++;
    }
    for (int q = 0; q < np; ++q) pt_ptr[q + 1] += pt_ptr[q];
    std::vector<uint8_t> pt_free(np);
    for (int q = 0; q < np; ++q) pt_free[q] = (p->pt_fixed && p->pt_fixed[h->pt_orig[q]]) ? 0 : 1;

    // ---- free cameras: not fixed and with at least one (therefore active) edge
    std::vector<int> cnt_cam(nc, 0);
    for (int i = 0; i < no; ++i) cnt_cam[p->obs_cam[i]]++;
    std::vector<int> cam_free(nc, -1), free_cam;
    for (int c = 0; c < nc; ++c)
        if (!(p->cam_fixed && p->cam_fixed[c]) && cnt_cam[c] > 0) { cam_free[c] = (int)free_cam.size(); free_cam.push_back(c); }
    const int nfc = (int)free_cam.size();
    if (nfc > kMaxFreeCams) return fail(h, NAV_BA_TOO_LARGE, "more than 8192 free cameras");

    // ---- camera-major observation lists of free cameras (ascending observation index)
    std::vector<int> cam_ptr(nfc + 1, 0), cam_obs, cam_obs_b0;
    for (int k = 0; k < no; ++k) if (cam_free[o_cam[k]] >= 0) cam_ptr[cam_free[o_cam[k]] + 1]++;
    for (int f = 0; f < nfc; ++f) cam_ptr[f + 1] += cam_ptr[f];
    cam_obs.resize(cam_ptr[nfc]);
    {
        std::vector<int> pos(cam_ptr.begin(), cam_ptr.end() - 1);
        for (int k = 0; k < no; ++k) {
            const int f = cam_free[o_cam[k]];
            if (f >= 0) cam_obs[pos[f]++] = k;
        }
    }
    // first observation of the same point on the same camera (start of the
    // point's observations on cameras >= this one; duplicates are adjacent)
    cam_obs_b0.resize(cam_obs.size());
    for (size_t k = 0; k < cam_obs.size(); ++k) {
        int b = cam_obs[k];
        const int q = o_pt[b];
        while (b > pt_ptr[q] && o_cam[b - 1] == o_cam[cam_obs[k]]) --b;
        cam_obs_b0[k] = b;
    }

    // ---- point-aligned linearisation tiles
    std::vector<int> tile_obs{0}, tile_pt{0};
    for (int q = 0; q < np;) {
        int q1 = q;
        while (q1 < np && pt_ptr[q1 + 1] - pt_ptr[q] <= kTile) ++q1;
        if (q1 == q) q1 = q + 1;  // one point with more than kTile observations
        tile_obs.push_back(pt_ptr[q1]);
        tile_pt.push_back(q1);
        q = q1;
    }
    const int n_tiles = (int)tile_obs.size() - 1;

    // ---- Hpp work items: (free camera, chunk)
    std::vector<int> hf, ha0, ha1, cam_hitem_ptr{0};
    for (int f = 0; f < nfc; ++f) {
        for (int a = cam_ptr[f]; a < cam_ptr[f + 1]; a += kChunk) {
            hf.push_back(f); ha0.push_back(a); ha1.push_back(std::min(a + kChunk, cam_ptr[f + 1]));
        }
        cam_hitem_ptr.push_back((int)hf.size());
    }

    // ---- co-visibility through free points: column blocks g >= f of each S row
    std::vector<uint64_t> bits(((size_t)nfc * nfc + 63) / 64, 0);
    std::vector<int> fl;
    for (int q = 0; q < np; ++q) {
        if (!pt_free[q]) continue;
        fl.clear();
        for (int k = pt_ptr[q]; k < pt_ptr[q + 1]; ++k) {
            const int f = cam_free[o_cam[k]];
            if (f >= 0 && (fl.empty() || fl.back() != f)) fl.push_back(f);
        }
        for (size_t a = 0; a < fl.size(); ++a)
            for (size_t b = a + 1; b < fl.size(); ++b) {
                const size_t bit = (size_t)fl[a] * nfc + fl[b];
                bits[bit >> 6] |= 1ull << (bit & 63);
            }
    }
    std::vector<int> colptr{0}, colcam;
    for (int f = 0; f < nfc; ++f) {
        colcam.push_back(f);
        for (int g2 = f + 1; g2 < nfc; ++g2) {
            const size_t bit = (size_t)f * nfc + g2;
            if (bits[bit >> 6] >> (bit & 63) & 1ull) colcam.push_back(g2);
        }
        colptr.push_back((int)colcam.size());
    }

    // ---- Schur work items: per row, column groups of kMaxSlots, chunks of kChunk
    std::vector<int> sf, sa0, sa1, sg, sns, row_sitem_ptr{0}, row_nch;
    std::vector<long long> spofs;
    long long pofs = 0;
    int max_ns = 1;
    for (int f = 0; f < nfc; ++f) {
        const int ncols = colptr[f + 1] - colptr[f];
        const int ngroups = (ncols + kMaxSlots - 1) / kMaxSlots;
        const int nobs = cam_ptr[f + 1] - cam_ptr[f];
        const int nch = std::max(1, (nobs + kChunk - 1) / kChunk);
        row_nch.push_back(nch);
        for (int gi = 0; gi < ngroups; ++gi) {
            const int ns = std::min(kMaxSlots, ncols - gi * kMaxSlots);
            max_ns = std::max(max_ns, ns);
            for (int ch = 0; ch < nch; ++ch) {
                sf.push_back(f);
                sa0.push_back(cam_ptr[f] + ch * kChunk);
                sa1.push_back(std::min(cam_ptr[f] + (ch + 1) * kChunk, cam_ptr[f + 1]));
                sg.push_back(gi);
                sns.push_back(ns);
                spofs.push_back(pofs);
                pofs += 36LL * ns + 6;
            }
        }
        row_sitem_ptr.push_back((int)sf.size());
    }

    // ---- initial state
    std::vector<PoseQT> pose0(nc);
    for (int c = 0; c < nc; ++c) {
        const double* T = p->T_cw + 12 * (size_t)c;
        host_quat_from_R(T, pose0[c].q);
        pose0[c].t[0] = T[9]; pose0[c].t[1] = T[10]; pose0[c].t[2] = T[11];
        pose0[c].pad = 0;
    }
    std::vector<D4> X0(np);
    for (int q = 0; q < np; ++q) {
        const double* x = p->X + 3 * (size_t)h->pt_orig[q];
        X0[q] = compat ? D4{rf32(x[0]), rf32(x[1]), rf32(x[2]), 0.0} : D4{x[0], x[1], x[2], 0.0};
    }

    // ---- device upload
    h->nbuf = 0;
    DevProblem& P = h->P;
    P = DevProblem{};
    P.n_cam = nc; P.n_pt = np; P.n_obs = no; P.n_fc = nfc;
    P.robust = o->robust ? 1 : 0;
    P.delta = o->huber_delta;
    P.fx = compat ? rf32(p->fx) : p->fx; P.fy = compat ? rf32(p->fy) : p->fy;
    P.cx = compat ? rf32(p->cx) : p->cx; P.cy = compat ? rf32(p->cy) : p->cy;
    int st;
#define UP(dst, vec) if ((st = upload(h, &(dst), (vec))) != NAV_BA_OK) return st
    UP(P.o_cam, o_cam); UP(P.o_pt, o_pt); UP(P.o_uv, o_uv); UP(P.o_orig, o_orig);
    UP(P.pt_ptr, pt_ptr); UP(P.pt_free, pt_free); UP(P.cam_free, cam_free); UP(P.free_cam, free_cam);
    UP(P.cam_ptr, cam_ptr); UP(P.cam_obs, cam_obs); UP(P.cam_obs_b0, cam_obs_b0);
    UP(P.tile_obs, tile_obs); UP(P.tile_pt, tile_pt);
    UP(P.hitem_f, hf); UP(P.hitem_a0, ha0); UP(P.hitem_a1, ha1); UP(P.cam_hitem_ptr, cam_hitem_ptr);
    UP(P.sitem_f, sf); UP(P.sitem_a0, sa0); UP(P.sitem_a1, sa1); UP(P.sitem_g, sg); UP(P.sitem_nslots, sns);
    UP(P.sitem_pofs, spofs); UP(P.row_sitem_ptr, row_sitem_ptr); UP(P.row_nchunks, row_nch);
    UP(P.colptr, colptr); UP(P.colcam, colcam);
#undef UP
    P.n_tiles = n_tiles;
    P.n_hitems = (int)hf.size();
    P.n_sitems = (int)sf.size();
    P.max_ns = max_ns;
    P.n = 6 * nfc;
    P.n_pad = ((P.n + kNB - 1) / kNB) * kNB;
    P.nt = P.n_pad / kNB;

    const size_t tsz = mixed(h) ? sizeof(float) : sizeof(double);
    auto A = [&](void** dst, size_t bytes) { return alloc(h, dst, bytes); };
#define AL(dst, bytes) if ((st = A((void**)&(dst), (bytes))) != NAV_BA_OK) return st
    AL(h->init.pose, sizeof(PoseQT) * nc); AL(h->init.X, sizeof(D4) * np); h->init.cam = nullptr;
    AL(h->cur.pose, sizeof(PoseQT) * nc); AL(h->cur.cam, sizeof(CamRT) * nc); AL(h->cur.X, sizeof(D4) * np);
    AL(h->cand.pose, sizeof(PoseQT) * nc); AL(h->cand.cam, sizeof(CamRT) * nc); AL(h->cand.X, sizeof(D4) * np);
    DevWork& W = h->W;
    W.n_cost_blocks = (no + 256 * 4 - 1) / (256 * 4);
    if (W.n_cost_blocks == 0) W.n_cost_blocks = 1;
    W.n_pt_blocks = (np + 255) / 256;
    AL(W.w, tsz * no);
    AL(W.Hl, tsz * 9 * (size_t)np);
    AL(W.hpart, tsz * 27 * std::max<size_t>(hf.size(), 1));
    AL(W.Hc, sizeof(double) * 28 * std::max(nfc, 1));
    AL(W.ptsys, sizeof(double) * 12 * (size_t)np);
    AL(W.spart, sizeof(double) * std::max<long long>(pofs, 1));
    AL(W.S, sizeof(double) * std::max<size_t>((size_t)P.n_pad * P.n_pad, 1));
    AL(W.g, sizeof(double) * std::max(P.n_pad, 1));
    AL(W.y, sizeof(double) * std::max(P.n_pad, 1));
    AL(W.dp, sizeof(double) * std::max(P.n_pad, 1));
    AL(W.Lm, sizeof(double) * std::max<size_t>((size_t)P.n_pad * P.n_pad, 1));
    AL(W.red_cost, sizeof(double) * W.n_cost_blocks);
    AL(W.red_scale, sizeof(double) * (1 + W.n_pt_blocks));
    AL(W.red_aux, sizeof(double) * W.n_cost_blocks);
    AL(W.outlier, (size_t)no);
    AL(W.lm, sizeof(LMDev));
    AL(h->dfinal, 2 * sizeof(double));
#undef AL
    HIPTRY(h, hipMemcpyAsync(h->init.pose, pose0.data(), sizeof(PoseQT) * nc, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipMemcpyAsync(h->init.X, X0.data(), sizeof(D4) * np, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    // Schur rows may use more than the default 64 KiB of LDS
    if ((mixed(h) ? prepare_schur_row<float>(P) : prepare_schur_row<double>(P)) != 0)
        return fail(h, NAV_BA_TOO_LARGE, "Schur row workspace exceeds the LDS of one workgroup");
    const bool auto_lambda = !(o->lambda_init > 0);
    if ((st = capture(h, &h->graphs.iter, 1)) != NAV_BA_OK) return st;
    if (auto_lambda && (st = capture(h, &h->graphs.iter0, 0)) != NAV_BA_OK) return st;
    if ((st = capture(h, &h->graphs.trial, 2)) != NAV_BA_OK) return st;
    h->prepared = true;
    return NAV_BA_OK;
}

int32_t nav_ba_run(nav_ba_handle* h) {
    if (!h) return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "nav_ba_run before nav_ba_prepare");
    h->err.clear();
    h->ran = false;
    h->iterations = 0; h->trials = 0; h->stop_reason = NAV_BA_STOP_MAX_ITER;
    h->cost_history.assign(1, 0.0);
    if (h->empty) {
        h->stop_reason = NAV_BA_STOP_EMPTY;
        h->initial_cost = h->final_cost = 0.0;
        h->n_outliers = 0;
        h->ran = true;
        return NAV_BA_EMPTY;
    }
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nc = h->P.n_cam, np = h->P.n_pt;
    const bool m = mixed(h);
    HIPTRY(h, hipEventRecord(h->ev0, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * nc, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * np, hipMemcpyDeviceToDevice, s));
    launch_lm_reset(h->W.lm, h->opt.lambda_init > 0 ? h->opt.lambda_init : 0.0, s);
    launch_cam_rt(h->P, h->cur, s);
    if (m) launch_cost<float>(h->P, h->cur, h->W, s); else launch_cost<double>(h->P, h->cur, h->W, s);
    launch_lm_set_chi(h->W, s);
    HIPTRY(h, hipMemcpyAsync(h->hstatus, h->W.lm, sizeof(LMDev), hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipStreamSynchronize(s));
    h->initial_cost = h->hstatus->chi;
    h->cost_history[0] = h->initial_cost;
    const bool auto_lambda = !(h->opt.lambda_init > 0);
    int st;
    for (int it = 0; it < h->opt.n_iter; ++it) {
        if ((st = launch_step(h, (it == 0 && auto_lambda) ? 0 : 1)) != NAV_BA_OK) return st;
        HIPTRY(h, hipStreamSynchronize(s));
        int q = 1;
        double rho = h->hstatus->rho;
        while (rho < 0 && q < h->opt.max_trials) {
            if ((st = launch_step(h, 2)) != NAV_BA_OK) return st;
            HIPTRY(h, hipStreamSynchronize(s));
            rho = h->hstatus->rho;
            ++q;
        }
        h->trials += q;
        h->iterations = it + 1;
        h->cost_history.push_back(h->hstatus->chi);
        if (q == h->opt.max_trials || rho == 0) {
            h->stop_reason = NAV_BA_STOP_TERMINATE;
            break;
        }
    }
    h->final_lambda = h->hstatus->lambda;
    // outlier flags and the FP64 cost at the final state
    h->timing_active = h->timing;
    stage(h, NAV_BA_ST_FLAGS, [&] {
        launch_flags(h->P, h->cur, h->W, h->opt.outlier_chi2, s);
        launch_final(h->W, h->dfinal, s);
    });
    h->timing_active = false;
    HIPTRY(h, hipMemcpyAsync(h->hfinal, h->dfinal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipEventRecord(h->ev1, s));
    HIPTRY(h, hipStreamSynchronize(s));
    float ms = 0.f;
    HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->solve_ms = ms;
    if (h->timing) h->timer.collect();
    h->final_cost = h->hfinal[0];
    h->n_outliers = (int)h->hfinal[1];
    h->ran = true;
    return NAV_BA_OK;
}

int32_t nav_ba_fetch(nav_ba_handle* h, nav_ba_result* r) {
    if (!h || !r) return NAV_BA_BAD_ARG;
    if (!h->prepared || !h->ran) return fail(h, NAV_BA_NOT_PREPARED, "nav_ba_fetch before nav_ba_run");
    const int nc = h->n_cam, npa = h->n_pt_all;
    r->iterations = h->iterations; r->trials = h->trials; r->stop_reason = h->stop_reason;
    r->initial_cost = h->initial_cost; r->final_cost = h->final_cost; r->final_lambda = h->final_lambda;
    r->n_outliers = h->n_outliers; r->solve_ms = h->solve_ms;
    if (r->cost_history)
        for (size_t i = 0; i < h->cost_history.size(); ++i) r->cost_history[i] = h->cost_history[i];
    if (h->empty) {
        if (r->T_cw) std::memcpy(r->T_cw, h->T_in.data(), sizeof(double) * 12 * (size_t)nc);
        if (r->X) std::memcpy(r->X, h->X_in.data(), sizeof(double) * 3 * (size_t)npa);
        if (r->outlier) std::memset(r->outlier, 0, (size_t)h->n_obs);
        return NAV_BA_EMPTY;
    }
    HIPTRY(h, hipSetDevice(h->device));
    std::vector<PoseQT> pose(nc);
    std::vector<D4> X(h->P.n_pt);
    HIPTRY(h, hipMemcpyAsync(pose.data(), h->cur.pose, sizeof(PoseQT) * nc, hipMemcpyDeviceToHost, h->stream));
    HIPTRY(h, hipMemcpyAsync(X.data(), h->cur.X, sizeof(D4) * X.size(), hipMemcpyDeviceToHost, h->stream));
    if (r->outlier && h->n_obs > 0)
        HIPTRY(h, hipMemcpyAsync(r->outlier, h->W.outlier, (size_t)h->n_obs, hipMemcpyDeviceToHost, h->stream));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    if (r->T_cw)
        for (int c = 0; c < nc; ++c) {
            double* T = r->T_cw + 12 * (size_t)c;
            host_R_from_quat(pose[c].q, T);  // BE_GraphOptim.cpp:220-223
            T[9] = pose[c].t[0]; T[10] = pose[c].t[1]; T[11] = pose[c].t[2];
        }
    if (r->X) {
        std::memcpy(r->X, h->X_in.data(), sizeof(double) * 3 * (size_t)npa);
        const bool compat = h->opt.compat_fp32_io != 0;
        for (size_t q = 0; q < X.size(); ++q) {
            double* o = r->X + 3 * (size_t)h->pt_orig[q];
            o[0] = compat ? rf32(X[q].x) : X[q].x;  // estimate().cast<float>() (BE_GraphOptim.cpp:239)
            o[1] = compat ? rf32(X[q].y) : X[q].y;
            o[2] = compat ? rf32(X[q].z) : X[q].z;
        }
    }
    return NAV_BA_OK;
}

int32_t nav_ba_set_timing(nav_ba_handle* h, int32_t enable) {
    if (!h) return NAV_BA_BAD_ARG;
    h->timing = enable != 0;
    h->timer.clear();
    return NAV_BA_OK;
}

int32_t nav_ba_get_timing(const nav_ba_handle* h, nav_ba_timing* out) {
    if (!h || !out) return NAV_BA_BAD_ARG;
    for (int i = 0; i < 16; ++i) { out->ms[i] = h->timer.ms[i]; out->launches[i] = h->timer.launches[i]; }
    return NAV_BA_OK;
}

int32_t nav_ba_debug_reduced_system(nav_ba_handle* h, double lambda, double* S, double* g,
                                    int32_t max_n, int32_t* n_out) {
    if (!h || !n_out) return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "debug call before nav_ba_prepare");
    if (h->empty) { *n_out = 0; return NAV_BA_EMPTY; }
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    *n_out = P.n;
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    launch_lm_reset(h->W.lm, lambda, s);
    launch_cam_rt(P, h->cur, s);
    if (mixed(h)) {
        enqueue_lin<float>(h, false);
        launch_hinv<float>(P, h->W, s);
        if (P.n_fc > 0) { launch_schur_row<float>(P, h->cur, h->W, s); launch_schur_reduce<float>(P, h->W, s); }
    } else {
        enqueue_lin<double>(h, false);
        launch_hinv<double>(P, h->W, s);
        if (P.n_fc > 0) { launch_schur_row<double>(P, h->cur, h->W, s); launch_schur_reduce<double>(P, h->W, s); }
    }
    HIPTRY(h, hipStreamSynchronize(s));
    if (P.n > max_n) return NAV_BA_OK;
    std::vector<double> Sd((size_t)P.n_pad * P.n_pad), gd(P.n_pad);
    HIPTRY(h, hipMemcpy(Sd.data(), h->W.S, sizeof(double) * Sd.size(), hipMemcpyDeviceToHost));
    HIPTRY(h, hipMemcpy(gd.data(), h->W.g, sizeof(double) * gd.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < P.n; ++i) {
        if (g) g[i] = gd[i];
        for (int j = 0; j < P.n; ++j) {
            const double v = (j <= i) ? Sd[(size_t)i * P.n_pad + j] : Sd[(size_t)j * P.n_pad + i];
            if (S) S[(size_t)i * P.n + j] = v;
        }
    }
    return NAV_BA_OK;
}

int32_t nav_ba_solve(nav_ba_handle* h, const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r) {
    int st = nav_ba_prepare(h, p, o);
    if (st != NAV_BA_OK) return st;
    st = nav_ba_run(h);
    if (st != NAV_BA_OK && st != NAV_BA_EMPTY) return st;
    const int st2 = nav_ba_fetch(h, r);
    return st2 != NAV_BA_OK ? st2 : st;
}

}  // extern "C"
