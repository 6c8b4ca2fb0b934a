// nav_ba_host.cpp — host runtime and C ABI of the MI355X BA back-end.
//
// Replaces NAV24::BE::GraphOptim::solve (core/backEnd/BE_GraphOptim.cpp:31-77):
//   prepare  = graph build (addFrameVertices :79-100, addMapPointVertices :117-211)
//              as flat arrays: points with observations compacted in ascending id,
//              observations ordered by (point, camera, input order), camera-major
//              index lists, linearisation tiles, Schur work items, co-visibility.
//   run      = optimizer.optimize(nIterations) (:66-68) on the device.  One
//              replayed hipGraph advances every window by one LM trial (gated
//              linearisation, trial, accept/reject, next phase); the host only
//              watches a small status.
//   fetch    = recoverPose / recoverMapPoints (:213-244) plus the outlier flags.
// A batch of independent windows (the per-window calls nav24's front end
// makes, FE_SlamMonoV.cpp:161-165) is prepared as windows laid end to end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/nav_ba.h"
#include "../../include/nav_ba_debug.h"
#include "ba_kernels.h"

using namespace navba;

namespace {

constexpr const char* kVersion = "nav24_amd-ba 0.2 (gfx950)";
constexpr int kMaxDenseN = 16384;  // tiled dense reduced system of one window (2 GiB FP64)
// Conjugate gradients on the reduced system (above kMaxDenseN, or on request):
// dense S up to kPcgMaxN unknowns (19 GiB FP64), stop at ||r|| <= kPcgTol ||g||
// or kPcgMaxIter iterations; the host reads the stop flag every kPcgChunk.
constexpr int kPcgMaxN = 49152;
constexpr double kPcgTol = 1e-10;
constexpr int kPcgMaxIter = 2000;
constexpr int kPcgChunk = 8;
// Schur chunk tuner: problems with at least kTuneMinObs observations time each
// candidate chunk size once at prepare (best of kTuneReps launches) and take
// the largest size within kTuneNear of the fastest; the other row kernel is
// taken only if it beats that by kTuneMargin.  Sizes whose times differ by
// less than the timing noise give different roundings, so the rule prefers a
// stable pick (the largest near-best size) to the fastest of a noisy few.
constexpr int kTuneMinObs = 1 << 17;
constexpr int kTuneChunks[] = {128, 256, 512, 1024, 2048, 4096};
constexpr int kTuneReps = 3;
constexpr float kTuneNear = 1.05f;
constexpr float kTuneMargin = 0.95f;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
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

// Host copy of what fetch needs for one window.
struct WinHost {
    int n_cam = 0, n_pt = 0, n_obs = 0;
    int cam0 = 0, pt_local0 = 0, pt_local1 = 0, obs0 = 0;  // ranges in the concatenated problem
    std::vector<int> pt_orig;      // local point -> caller point index of this window
    std::vector<double> X_in, T_in;
    bool empty = false;
};

// ---- Schur work items -------------------------------------------------------------------
// Host structure the items are built from (camera-major lists of free cameras,
// per-row co-visible column blocks, point-major observations).
struct SchurGraph {
    int nfc, nc;
    const int* colptr;
    const int* colcam;
    const int* cam_ptr;
    const int* cam_obs;
    const int* cam_obs_b0;
    const int* o_pt;
    const int* o_cam;
    const int* pt_ptr;
    const uint8_t* pt_free;
    const int* free_cam;
    const int* cam_win;
    const int* cam_off;
};

// Host-side structure of n_win windows laid end to end (see ba_kernels.h):
// everything prepare uploads, built without touching the device.
struct HostLayout {
    int n_win = 0, nc = 0, np = 0, no = 0, nfc = 0;
    bool batched = false, all_empty = true;
    std::vector<WinHost> win;
    std::vector<int> cam_off, obs_off, pt_first, cam_win, pt_win;
    std::vector<int> o_cam, o_pt, o_orig, pt_ptr;
    std::vector<double> o_uv;
    std::vector<uint8_t> pt_free;
    std::vector<int> cam_free, free_cam, win_fc0, win_cam0;
    std::vector<int> cam_ptr, cam_obs, cam_obs_b0, c_pt;
    std::vector<double> c_uv;
    std::vector<int> tile_obs, tile_pt, tile_win, win_tile0;
    std::vector<int> cblk_o0, cblk_win, win_cblk0, win_pt0;
    std::vector<int> win_chol_order;
    std::vector<int> hf, ha0, ha1, cam_hitem_ptr;
    std::vector<int> colptr, colcam, zcolptr, zcolcam;
    std::vector<long long> win_S;
    std::vector<int> win_lda;
    long long S_size = 1;
    int n = 0, n_pad = 0;
    std::vector<PoseQT> pose0;
    std::vector<D4> X0;
    double fx = 0, fy = 0, cx = 0, cy = 0;
    int cam_model = NAV_BA_CAM_PINHOLE;
    double kb[4] = {0, 0, 0, 0};

    SchurGraph graph() const {
        return SchurGraph{nfc, nc, colptr.data(), colcam.data(), cam_ptr.data(), cam_obs.data(), cam_obs_b0.data(),
                          o_pt.data(), o_cam.data(), pt_ptr.data(), pt_free.data(), free_cam.data(), cam_win.data(),
                          cam_off.data()};
    }
};

}  // namespace

struct nav_ba_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bool prepared = false;
    bool batched = false;
    nav_ba_options opt{};
    HostLayout layout;  // the last prepare's host structure (fetch reads its windows)
    bool all_empty = false;
    DevProblem P{};
    DevState cur{}, cand{}, init{};
    DevWork W{};
    std::vector<DevBuf> bufs;
    int nbuf = 0;
    hipGraphExec_t step = nullptr;   // one LM step
    hipGraphExec_t step0 = nullptr;  // the first step of a solve
    bool timing = false;
    bool timing_active = false;
    int solver_mode = 0;  // 0 auto, 1 tiles, 2 blocks, 3 in-LDS per window, 4 conjugate gradients
    PcgState* hpcg = nullptr;  // pinned: the conjugate-gradient state, read between chunks
    PcgState pcg_last{};       // ... and after the last trial's solve
    int pcg_err = NAV_BA_OK;
    int schur_chunk = 0;  // observations per Schur work item: 0 = default or tuned
    int schur_chunk_used = 0;
    int schur_kernel = 0;  // Schur row kernel: 0 = default or tuned, 1 = rounds of kBlock, 2 = producer/consumer
    bool schur_pc_used = false;
    StageTimer timer;
    int* hdone = nullptr;      // pinned: windows finished
    LMDev* hlm = nullptr;      // pinned: the LM state of up to kBurstWindows windows, between bursts
    char* stage = nullptr;     // pinned staging of a prepare's uploads (grown on demand)
    size_t stage_cap = 0;
    char* stage_s = nullptr;   // pinned staging of the Schur work items (the tuner re-uploads them)
    size_t stage_s_cap = 0;
    // the cached index: the structure and settings of the last successful
    // prepare; a prepare with the same key only uploads the new state
    std::vector<char> prep_key;
    bool prep_cached = false;  // the last prepare took the cached index
    // results of the last run
    bool ran = false;
    std::vector<LMDev> lm_out;
    std::vector<double> hist_out, final_out;
    double solve_ms = 0;
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

// Device allocation slot of at least `bytes` (grow on demand, reused across prepares).
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

// Host arrays packed into a pinned staging buffer of the handle and sent with
// one copy into one device slot (each array 256-byte aligned, at least 256
// bytes, so an empty one still gets a distinct pointer): one pinned copy in
// place of one pageable copy per array.  A staging buffer is reused only after
// the stream has been synchronised since its last copy.
struct Packer {
    char** buf;
    size_t* cap;
    Packer(char** b, size_t* c) : buf(b), cap(c) {}
    struct Item {
        const void* src;
        size_t bytes, off;
        const void** dst;
    };
    std::vector<Item> items;
    size_t total = 0;
    template <class T>
    void add(const T** out, const std::vector<T>& v) {
        const size_t b = sizeof(T) * v.size();
        items.push_back({v.data(), b, total, (const void**)out});
        total += (std::max<size_t>(b, 1) + 255) & ~size_t(255);
    }
    int flush(nav_ba_handle* h) {
        void* d = nullptr;
        int st = alloc(h, &d, total);
        if (st) return st;
        if (*cap < total) {
            if (*buf) (void)hipHostFree(*buf);
            *buf = nullptr;
            *cap = 0;
            HIPTRY(h, hipHostMalloc((void**)buf, total, hipHostMallocDefault));
            *cap = total;
        }
        for (const Item& it : items) {
            if (it.bytes) std::memcpy(*buf + it.off, it.src, it.bytes);
            *it.dst = (const char*)d + it.off;
        }
        HIPTRY(h, hipMemcpyAsync(d, *buf, total, hipMemcpyHostToDevice, h->stream));
        return NAV_BA_OK;
    }
};

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

bool mixed(const nav_ba_handle* h) { return h->opt.precision == NAV_BA_MIXED; }

// Enqueue one stage; with timing active it is bracketed by HIP events.
template <class F>
void stage(nav_ba_handle* h, int cat, F&& f) {
    if (h->timing_active) h->timer.begin(cat, h->stream);
    f();
    if (h->timing_active) h->timer.end(h->stream);
}

// The first step of a solve (every window at iteration 0) also sets lambda.
template <typename T>
void enqueue_lin(nav_ba_handle* h, bool first) {
    stage(h, NAV_BA_ST_LIN_PT, [&] { launch_lin_pt<T>(h->P, h->cur, h->W, h->stream); });
    stage(h, NAV_BA_ST_LIN_CAM, [&] { launch_lin_cam<T>(h->P, h->cur, h->W, h->stream); });
    if (first && h->P.auto_lambda) stage(h, NAV_BA_ST_LAMBDA, [&] { launch_lambda_init<T>(h->P, h->W, h->stream); });
}

template <typename T>
void enqueue_reduced_system(nav_ba_handle* h) {
    if (h->P.n_fc > 0) {
        stage(h, NAV_BA_ST_SCHUR_ROW, [&] { launch_schur_row<T>(h->P, h->cur, h->W, h->stream); });
        if (!h->P.use_win)  // the per-window factorisation sums the rows itself
            stage(h, NAV_BA_ST_SCHUR_REDUCE, [&] { launch_schur_reduce<T>(h->P, h->W, h->stream); });
    }
}

// Conjugate gradients on the reduced system of one problem, driven from the
// host: the start, then chunks of kPcgChunk iterations with the stop flag read
// after each, then the step.  Not capturable (it reads device state), so a
// problem that uses it runs its steps eagerly.
void run_pcg(nav_ba_handle* h) {
    hipStream_t s = h->stream;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && h->pcg_err == NAV_BA_OK) {
            h->pcg_err = NAV_BA_HIP_ERROR;
            h->err = std::string("conjugate gradients: ") + hipGetErrorString(e);
        }
        return e == hipSuccess;
    };
    launch_pcg_start(h->P, h->W, s);
    for (int it = 0; it < kPcgMaxIter; it += kPcgChunk) {
        launch_pcg_iterations(h->P, h->W, kPcgChunk, kPcgTol, kPcgMaxIter, s);
        if (!chk(hipMemcpyAsync(h->hpcg, h->W.pcg_st, sizeof(PcgState), hipMemcpyDeviceToHost, s))) return;
        if (!chk(hipStreamSynchronize(s))) return;
        if (h->hpcg->done) break;
    }
    if (h->hpcg->iter >= 0) h->pcg_last = *h->hpcg;  // a solve ran (the window had not finished)
    launch_pcg_finish(h->P, h->W, s);
}

// One LM step of every window: gated linearisation, one trial, decision.
template <typename T>
void enqueue_step(nav_ba_handle* h, bool first) {
    enqueue_lin<T>(h, first);
    enqueue_reduced_system<T>(h);
    if (h->P.n_fc > 0)
        stage(h, NAV_BA_ST_CHOLESKY, [&] {
            if (h->P.use_win) launch_cholesky_small(h->P, h->cur, h->cand, h->W, h->stream);
            else if (h->P.use_pcg) run_pcg(h);
            else launch_cholesky_solve(h->P, h->cur, h->cand, h->W, h->stream);
        });
    // the candidate cameras first: the back-substitution also forms the trial cost
    if (!cholesky_updates_cameras(h->P))
        stage(h, NAV_BA_ST_UPDATE_CAM, [&] { launch_update_cam(h->P, h->cur, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_BACKSUB, [&] { launch_backsub<T>(h->P, h->cur, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_ACCEPT, [&] {
        launch_lm_accept(h->P, h->cur, h->cand, h->W, h->stream);
        launch_commit(h->P, h->cur, h->cand, h->W, h->stream);
    });
}

int launch_step(nav_ba_handle* h, bool first) {
    if (!h->timing && !h->P.use_pcg) {
        HIPTRY(h, hipGraphLaunch(first ? h->step0 : h->step, h->stream));
        return NAV_BA_OK;
    }
    // stage timing, or conjugate gradients: the step's launches one by one
    h->timing_active = h->timing;
    h->pcg_err = NAV_BA_OK;
    if (mixed(h)) enqueue_step<float>(h, first); else enqueue_step<double>(h, first);
    h->timing_active = false;
    if (h->pcg_err != NAV_BA_OK) return fail(h, h->pcg_err, h->err);
    return NAV_BA_OK;
}

// Two graphs: the first step of a solve (with the lambda initialisation) and every later one.
int capture_step(nav_ba_handle* h) {
    for (int first = 0; first < 2; ++first) {
        hipGraph_t g = nullptr;
        HIPTRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        if (mixed(h)) enqueue_step<float>(h, first); else enqueue_step<double>(h, first);
        HIPTRY(h, hipStreamEndCapture(h->stream, &g));
        hipError_t e = hipGraphInstantiate(first ? &h->step0 : &h->step, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess)
            return fail(h, NAV_BA_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
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

// Work items of one chunk size (camera-major observations per item) and their pair lists.
struct SchurItems {
    int chunk = 0, max_ns = 1;
    bool pc = false;                 // producer/consumer kernel (lanes bound per item)
    long long pofs = 0;
    std::vector<int> sf, sa0, sa1, sg, sns, srnd0, row_sitem_ptr, row_nch, sorder, rnd_pair0;
    std::vector<long long> spofs, row_pofs0;
    std::vector<uint8_t> lane_slot;  // [rounds][kBlock], or [items][kPcCons] for the producer/consumer kernel
    std::vector<uint8_t> rnd_group;  // [rounds] lanes per group (rounds of kBlock only)
    std::vector<int4> rnd_rows;      // [rounds]
    std::vector<int2> pairs;
};

// Lanes of one round over its slots, in aligned groups of grp: one group per
// slot with pairs, then one group at a time to the slot whose lanes carry the
// most pairs each.  Returns the largest share (pairs of one lane).
int assign_lanes(const std::vector<std::vector<int2>>& bucket, int ns, int lanes, int grp, std::vector<int>& nlane) {
    nlane.assign(ns, 0);
    int used = 0, worst = 0;
    // max-heap of (pairs per lane, -slot), kept in storage reused across calls
    // (the same heap operations, hence the same ties, as a priority queue)
    thread_local std::vector<std::pair<int, int>> pq;
    pq.clear();
    auto push = [&](std::pair<int, int> e) { pq.push_back(e); std::push_heap(pq.begin(), pq.end()); };
    for (int sl = 0; sl < ns; ++sl)
        if (!bucket[sl].empty()) {
            nlane[sl] = grp;
            used += grp;
            const int c = (int)bucket[sl].size();
            push({(c + grp - 1) / grp, -sl});
        }
    if (used > lanes) return 1 << 30;
    while (used + grp <= lanes && !pq.empty() && pq.front().first > 1) {
        const int sl = -pq.front().second;
        std::pop_heap(pq.begin(), pq.end());
        pq.pop_back();
        nlane[sl] += grp;
        used += grp;
        const int c = (int)bucket[sl].size();
        push({(c + nlane[sl] - 1) / nlane[sl], -sl});
    }
    if (!pq.empty()) worst = pq.front().first;
    return worst;
}

// Dispatch order of work items over the XCDs: workgroup i of a launch runs
// on XCD i % kXcds, each with its own L2.  The items of one window (a batch)
// or of one band of consecutive rows (one problem) gather the same points,
// weights and cameras, so they go to one XCD: windows are dealt to the XCDs
// largest first onto the least loaded, rows in kXcds bands of equal cost.
// Each XCD takes its items heaviest first; where its list has run out, a slot
// takes the next item of the XCD with the most left.
std::vector<int> xcd_order(const SchurGraph& G, const std::vector<int>& item_row, const std::vector<long long>& cost) {
    const int n = (int)item_row.size();
    std::vector<int> order;
    order.reserve(n);
    if (n == 0) return order;
    std::vector<int> xcd(n, 0);
    int n_win = 0;
    for (int f = 0; f < G.nfc; ++f) n_win = std::max(n_win, G.cam_win[G.free_cam[f]] + 1);
    if (n_win > 1) {
        std::vector<long long> wcost(n_win, 0);
        for (int i = 0; i < n; ++i) wcost[G.cam_win[G.free_cam[item_row[i]]]] += cost[i];
        std::vector<int> wo(n_win);
        for (int w = 0; w < n_win; ++w) wo[w] = w;
        std::stable_sort(wo.begin(), wo.end(), [&](int a, int b) { return wcost[a] > wcost[b]; });
        std::vector<long long> load(kXcds, 0);
        std::vector<int> wx(n_win, 0);
        for (int w : wo) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            wx[w] = x;
            load[x] += wcost[w];
        }
        for (int i = 0; i < n; ++i) xcd[i] = wx[G.cam_win[G.free_cam[item_row[i]]]];
    } else {
        std::vector<long long> rcost(G.nfc + 1, 0);
        long long total = 0;
        for (int i = 0; i < n; ++i) { rcost[item_row[i]] += cost[i]; total += cost[i]; }
        std::vector<int> rx(G.nfc, 0);
        long long run = 0;
        for (int f = 0; f < G.nfc; ++f) {
            rx[f] = total > 0 ? (int)std::min<long long>(kXcds - 1, (run + rcost[f] / 2) * kXcds / total) : 0;
            run += rcost[f];
        }
        for (int i = 0; i < n; ++i) xcd[i] = rx[item_row[i]];
    }
    std::vector<std::vector<int>> lists(kXcds);
    for (int i = 0; i < n; ++i) lists[xcd[i]].push_back(i);
    for (auto& l : lists) std::stable_sort(l.begin(), l.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<size_t> head(kXcds, 0);
    for (int pos = 0; pos < n; ++pos) {
        int x = pos % kXcds;
        if (head[x] == lists[x].size()) {
            size_t best = 0;
            for (int y = 0; y < kXcds; ++y)
                if (lists[y].size() - head[y] > best) { best = lists[y].size() - head[y]; x = y; }
        }
        order.push_back(lists[x][head[x]++]);
    }
    return order;
}

// Items: per S row, column groups of kMaxSlots slots, chunks of `chunk`
// observations taken in rounds of kBlock, dispatched heaviest first.
// Pairs of a round: the pairs (a, b) of each of its observations a with the
// later observations b of a's point on free cameras of the item's column
// group (a itself excluded: the diagonal pair is summed in phase A), grouped
// by the slot of b's camera, in order of a then b.  The slots with pairs share
// the round's lanes (assign_lanes; groups of 4 unless groups of 2 give a
// smaller largest share), and each slot's pairs are split into contiguous,
// nearly equal shares, one per lane; a slot's lanes are contiguous.  Entry:
// {b, a's row in the round}.  Per wave, row j holds every lane's pair j
// ({-1, 0} past a lane's share).  False when the pair entries exceed int32.
bool build_schur_items(const SchurGraph& G, int chunk, bool pc, SchurItems& S) {
    S = SchurItems{};
    S.chunk = chunk;
    S.pc = pc;
    const int lanes = kBlock;
    const int sg_n = kMaxSlots;
    S.row_sitem_ptr.push_back(0);
    for (int f = 0; f < G.nfc; ++f) {
        const int ncols = G.colptr[f + 1] - G.colptr[f];
        const int ngroups = (ncols + sg_n - 1) / sg_n;
        const int nobs = G.cam_ptr[f + 1] - G.cam_ptr[f];
        const int nch = std::max(1, (nobs + chunk - 1) / chunk);
        S.row_nch.push_back(nch);
        S.row_pofs0.push_back(S.pofs);
        for (int gi = 0; gi < ngroups; ++gi) {
            const int ns = std::min(sg_n, ncols - gi * sg_n);
            S.max_ns = std::max(S.max_ns, ns);
            for (int ch = 0; ch < nch; ++ch) {
                S.sf.push_back(f);
                S.sa0.push_back(G.cam_ptr[f] + ch * chunk);
                S.sa1.push_back(std::min(G.cam_ptr[f] + (ch + 1) * chunk, G.cam_ptr[f + 1]));
                S.sg.push_back(gi);
                S.sns.push_back(ns);
                S.spofs.push_back(S.pofs);
                S.pofs += 36LL * ns + 6;
            }
        }
        S.row_sitem_ptr.push_back((int)S.sf.size());
    }
    std::vector<long long> scost(S.sf.size(), 0);
    std::vector<int> slotmap(G.nc, -1);
    std::vector<std::vector<int2>> bucket;
    std::vector<int> nlane, nlane2, lane_s, lane_beg, lane_len;
    // the pairs of the round of camera-major observations [kb, ke), by slot
    auto round_pairs = [&](int kb, int ke, int ns) {
        for (int sl = 0; sl < ns; ++sl) bucket[sl].clear();
        for (int k = kb; k < ke; ++k) {
            const int a = G.cam_obs[k];
            const int q = G.o_pt[a];
            if (!G.pt_free[q]) continue;
            for (int b = G.cam_obs_b0[k]; b < G.pt_ptr[q + 1]; ++b) {
                if (b == a) continue;
                const int sl = slotmap[G.o_cam[b]];
                if (sl < 0) continue;
                bucket[sl].push_back(int2{b, k - kb});
            }
        }
    };
    // pair rows of a round whose lanes (lane_s / lane_beg / lane_len) are laid
    // out, for waves w0 .. w0 + nw - 1 of kWave lanes from lane l0
    auto emit_rows = [&](int l0, int nw, int rows[4]) -> bool {
        size_t tot = 0;
        for (int wv = 0; wv < 4; ++wv) rows[wv] = 0;
        for (int wv = 0; wv < nw; ++wv) {
            for (int l = 0; l < kWave; ++l) rows[wv] = std::max(rows[wv], lane_len[l0 + wv * kWave + l]);
            tot += (size_t)kWave * rows[wv];
        }
        if (S.pairs.size() + tot > (size_t)0x7fffffff) return false;
        S.rnd_pair0.push_back((int)S.pairs.size());
        S.rnd_rows.push_back(int4{rows[0], rows[1], rows[2], rows[3]});
        const size_t base = S.pairs.size();
        S.pairs.resize(base + tot);
        int2* out = S.pairs.data() + base;
        const int2* src[kWave];
        int len[kWave];
        for (int wv = 0; wv < nw; ++wv) {
            for (int l = 0; l < kWave; ++l) {
                const int t = l0 + wv * kWave + l;
                len[l] = lane_len[t];
                src[l] = len[l] > 0 ? bucket[lane_s[t]].data() + lane_beg[t] : nullptr;
            }
            for (int j = 0; j < rows[wv]; ++j)
                for (int l = 0; l < kWave; ++l) *out++ = j < len[l] ? src[l][j] : int2{-1, 0};
        }
        return true;
    };
    // a slot's pairs of one round split into contiguous, nearly equal shares
    // over its m lanes from lane t
    auto split = [&](int sl, int m, int t) {
        const int c = (int)bucket[sl].size();
        for (int u = 0, st = 0; u < m; ++u, ++t) {
            const int len = c / m + (u < c % m ? 1 : 0);
            lane_s[t] = sl;
            lane_beg[t] = st;
            lane_len[t] = len;
            st += len;
        }
    };
    std::vector<std::vector<int>> rcount;  // producer/consumer: [round][slot] pair counts of the item
    for (size_t it = 0; it < S.sf.size(); ++it) {
        const int f = S.sf[it], gi = S.sg[it], ns = S.sns[it];
        const int col0 = G.colptr[f] + gi * sg_n;
        for (int sl = 0; sl < ns; ++sl) slotmap[G.free_cam[G.colcam[col0 + sl]]] = sl;
        if (bucket.size() < (size_t)ns) bucket.resize(ns);
        S.srnd0.push_back((int)S.rnd_rows.size());
        long long cost = 0;
        if (pc) {
            // consumer lanes bound per item: one lane per slot with pairs, then one
            // at a time to the slot whose lanes carry the most pair rows summed over
            // the item's rounds
            const int nr = (S.sa1[it] - S.sa0[it] + kPcRound - 1) / kPcRound;
            rcount.assign(nr, std::vector<int>(ns, 0));
            std::vector<long long> tot(ns, 0);
            for (int r = 0; r < nr; ++r) {
                const int kb = S.sa0[it] + r * kPcRound;
                round_pairs(kb, std::min(kb + kPcRound, S.sa1[it]), ns);
                for (int sl = 0; sl < ns; ++sl) { rcount[r][sl] = (int)bucket[sl].size(); tot[sl] += rcount[r][sl]; }
            }
            auto load = [&](int sl, int m) {
                long long v = 0;
                for (int r = 0; r < nr; ++r) v += (rcount[r][sl] + m - 1) / m;
                return v;
            };
            nlane.assign(ns, 0);
            int used = 0;
            std::priority_queue<std::pair<long long, int>> pq;
            for (int sl = 0; sl < ns; ++sl)
                if (tot[sl] > 0) { nlane[sl] = 1; ++used; pq.push({load(sl, 1), -sl}); }
            while (used < kPcCons && !pq.empty() && pq.top().first > nr) {
                const int sl = -pq.top().second;
                pq.pop();
                ++nlane[sl];
                ++used;
                pq.push({load(sl, nlane[sl]), -sl});
            }
            lane_s.assign(kPcCons, 255);
            for (int sl = 0, t = 0; sl < ns; ++sl)
                for (int u = 0; u < nlane[sl]; ++u) lane_s[t++] = sl;
            for (int t = 0; t < kPcCons; ++t) S.lane_slot.push_back((uint8_t)lane_s[t]);
            for (int r = 0; r < nr; ++r) {
                const int kb = S.sa0[it] + r * kPcRound;
                round_pairs(kb, std::min(kb + kPcRound, S.sa1[it]), ns);
                lane_beg.assign(kPcCons, 0);
                lane_len.assign(kPcCons, 0);
                for (int sl = 0, t = 0; sl < ns; ++sl) {
                    if (nlane[sl] > 0) split(sl, nlane[sl], t);
                    t += nlane[sl];
                }
                int rows[4];
                if (!emit_rows(0, kPcCons / kWave, rows)) return false;
                S.rnd_group.push_back(1);
                cost += (long long)kWave * std::max(std::max(rows[0], rows[1]), 2) + 64;
            }
            for (int sl = 0; sl < ns; ++sl) slotmap[G.free_cam[G.colcam[col0 + sl]]] = -1;
            scost[it] = cost;
            continue;
        }
        for (int kb = S.sa0[it]; kb < S.sa1[it]; kb += lanes) {
            round_pairs(kb, std::min(kb + lanes, S.sa1[it]), ns);
            int grp = 4;
            const int worst4 = assign_lanes(bucket, ns, lanes, 4, nlane);
            if (assign_lanes(bucket, ns, lanes, 2, nlane2) < worst4) {
                grp = 2;
                nlane.swap(nlane2);
            }
            S.rnd_group.push_back((uint8_t)grp);
            // lanes laid out slot by slot (each slot's lanes whole, aligned groups);
            // a slot's pairs split into contiguous shares, one per lane
            lane_s.assign(lanes, 255);
            lane_beg.assign(lanes, 0);
            lane_len.assign(lanes, 0);
            for (int sl = 0, t = 0; sl < ns; ++sl) {
                split(sl, nlane[sl], t);
                t += nlane[sl];
            }
            int rows[4];
            if (!emit_rows(0, lanes / kWave, rows)) return false;
            for (int t = 0; t < lanes; ++t) S.lane_slot.push_back((uint8_t)lane_s[t]);
            const int rmax = std::max(std::max(rows[0], rows[1]), std::max(rows[2], rows[3]));
            cost += (long long)kWave * rmax + 64;
        }
        for (int sl = 0; sl < ns; ++sl) slotmap[G.free_cam[G.colcam[col0 + sl]]] = -1;
        scost[it] = cost;
    }
    S.sorder = xcd_order(G, S.sf, scost);
    return true;
}

// Upload the items and size the partial-sum buffer; these are the last device
// allocations of a prepare, so candidates of the tuner reuse the same slots.
int upload_schur_items(nav_ba_handle* h, const SchurItems& S) {
    DevProblem& P = h->P;
    int st;
    // staging: prepare synchronises the stream before the tuner, and the tuner after every timing
    Packer pk(&h->stage_s, &h->stage_s_cap);
#define UP(dst, vec) pk.add(&(dst), (vec))
    UP(P.sitem_f, S.sf); UP(P.sitem_a0, S.sa0); UP(P.sitem_a1, S.sa1); UP(P.sitem_g, S.sg); UP(P.sitem_nslots, S.sns);
    UP(P.sitem_pofs, S.spofs); UP(P.row_sitem_ptr, S.row_sitem_ptr); UP(P.row_nchunks, S.row_nch);
    UP(P.row_pofs0, S.row_pofs0);
    UP(P.sitem_order, S.sorder); UP(P.sitem_rnd0, S.srnd0);
    UP(P.rnd_lane_slot, S.lane_slot); UP(P.rnd_pair0, S.rnd_pair0); UP(P.rnd_rows, S.rnd_rows);
    UP(P.rnd_group, S.rnd_group); UP(P.pairs, S.pairs);
#undef UP
    if ((st = pk.flush(h)) != NAV_BA_OK) return st;
    if ((st = alloc(h, (void**)&h->W.spart, sizeof(double) * std::max<long long>(S.pofs, 1))) != NAV_BA_OK) return st;
    P.n_sitems = (int)S.sf.size();
    P.max_ns = S.max_ns;
    P.schur_pc = S.pc ? 1 : 0;
    h->schur_chunk_used = S.chunk;
    h->schur_pc_used = S.pc;
    return NAV_BA_OK;
}

// Process-wide memory of tuned chunk sizes, per device, precision and problem shape.
struct TuneKey {
    int device, mixed, n_win, nc, np, no, nfc, kernel, cam;
    bool operator<(const TuneKey& o) const {
        return std::tie(device, mixed, n_win, nc, np, no, nfc, kernel, cam) <
               std::tie(o.device, o.mixed, o.n_win, o.nc, o.np, o.no, o.nfc, o.kernel, o.cam);
    }
};
struct TuneChoice {
    int chunk;
    bool pc;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tuned;

bool cached_choice(const TuneKey& k, TuneChoice* c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(k);
    if (it == g_tuned.end()) return false;
    *c = it->second;
    return true;
}

void remember_choice(const TuneKey& k, TuneChoice c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[k] = c;
}

// Device time of the Schur stage (rows + reduction) of the uploaded items, best of kTuneReps.
int time_schur(nav_ba_handle* h, float* best_ms) {
    *best_ms = 1e30f;
    for (int r = 0; r <= kTuneReps; ++r) {
        HIPTRY(h, hipEventRecord(h->ev0, h->stream));
        const bool red = !h->P.use_win;  // the per-window factorisation sums the rows itself
        if (mixed(h)) {
            launch_schur_row<float>(h->P, h->cur, h->W, h->stream);
            if (red) launch_schur_reduce<float>(h->P, h->W, h->stream);
        } else {
            launch_schur_row<double>(h->P, h->cur, h->W, h->stream);
            if (red) launch_schur_reduce<double>(h->P, h->W, h->stream);
        }
        HIPTRY(h, hipEventRecord(h->ev1, h->stream));
        HIPTRY(h, hipEventSynchronize(h->ev1));
        float ms = 0.f;
        HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
        if (r > 0) *best_ms = std::min(*best_ms, ms);  // the first launch is a warm-up
    }
    return NAV_BA_OK;
}

// Pick the kernel and chunk size of the Schur items by timing candidates on
// the problem itself, linearised at its initial state: every chunk size with
// the default kernel (the largest within kTuneNear of the fastest is kept),
// then the other kernel at that size.  `S` holds the default items, uploaded
// from allocation slot `mark` on; it returns holding (and with the device
// holding) the chosen candidate's items.  Chunk sizes
// above the longest camera row give the same items as the longest row's size
// and are skipped.  `kernel` (1 or 2, also the default kernel then) and
// `chunk_only` (> 0) limit the candidates to one kernel or one size; 0
// allows all.
int tune_schur_items(nav_ba_handle* h, const SchurGraph& G, int mark, int kernel, int chunk_only, SchurItems& S) {
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    std::vector<LMDev> L(P.n_win);
    for (auto& l : L) { l = LMDev{}; l.lambda = 1e-4; l.nu = 2.0; l.phase = kPhaseLin; l.iter = 1; }
    HIPTRY(h, hipMemcpyAsync(h->W.lm, L.data(), sizeof(LMDev) * L.size(), hipMemcpyHostToDevice, s));
    launch_cam_rt(P, h->cur, s);
    if (mixed(h)) enqueue_lin<float>(h, false); else enqueue_lin<double>(h, false);
    int st;
    float best_ms;
    if ((st = time_schur(h, &best_ms)) != NAV_BA_OK) return st;
    const int def = S.chunk;
    const bool def_pc = S.pc;
    int longest = 0;
    for (int f = 0; f < G.nfc; ++f) longest = std::max(longest, G.cam_ptr[f + 1] - G.cam_ptr[f]);
    SchurItems C;
    // time one candidate; ms < 0 when it cannot run (pair entries beyond int32, LDS)
    auto trial = [&](int chunk, bool pc, float* ms) -> int {
        *ms = -1.f;
        if (!build_schur_items(G, chunk, pc, C)) return NAV_BA_OK;
        h->nbuf = mark;
        int st2;
        if ((st2 = upload_schur_items(h, C)) != NAV_BA_OK) return st2;
        if ((mixed(h) ? prepare_schur_row<float>(h->P) : prepare_schur_row<double>(h->P)) != 0) return NAV_BA_OK;
        return time_schur(h, ms);
    };
    // the sizes with the default kernel: the largest within kTuneNear of the fastest
    if (kernel == 0 || (kernel == 2) == def_pc) {
        std::vector<std::pair<int, float>> timed{{def, best_ms}};
        int prev = 0;
        for (int chunk : kTuneChunks) {
            if (prev >= longest) break;  // the same items as the previous size
            prev = chunk;
            if (chunk == def || (chunk_only > 0 && chunk != chunk_only)) continue;
            float ms;
            if ((st = trial(chunk, def_pc, &ms)) != NAV_BA_OK) return st;
            if (ms > 0.f) timed.push_back({chunk, ms});
        }
        float tmin = best_ms;
        for (const auto& c : timed) tmin = std::min(tmin, c.second);
        int pick = 0;
        for (const auto& c : timed)
            if (c.second <= kTuneNear * tmin && c.first > pick) { pick = c.first; best_ms = c.second; }
        if (pick != def && build_schur_items(G, pick, def_pc, C)) std::swap(S, C);
    }
    // the other kernel at that size
    if (kernel == 0) {
        float ms;
        if ((st = trial(S.chunk, !def_pc, &ms)) != NAV_BA_OK) return st;
        if (ms > 0.f && ms < kTuneMargin * best_ms) std::swap(S, C);
    }
    h->nbuf = mark;
    if ((st = upload_schur_items(h, S)) != NAV_BA_OK) return st;
    if ((mixed(h) ? prepare_schur_row<float>(h->P) : prepare_schur_row<double>(h->P)) != 0)
        return fail(h, NAV_BA_TOO_LARGE, "Schur row workspace exceeds the LDS of one workgroup");
    HIPTRY(h, hipStreamSynchronize(s));
    return NAV_BA_OK;
}

// Check the arguments of a prepare; NAV_BA_OK or an error with its message.
int check_args(const nav_ba_problem* probs, int n_win, const nav_ba_options* o, bool batched, std::string& err) {
    if (!o || !probs || n_win <= 0) { err = "null options or problems"; return NAV_BA_BAD_ARG; }
    if (!batched && n_win != 1) { err = "single-problem prepare takes one problem"; return NAV_BA_BAD_ARG; }
    for (int w = 0; w < n_win; ++w)
        if (validate(&probs[w]) != NAV_BA_OK) {
            err = "invalid problem " + std::to_string(w) + " (null array or index out of range)";
            return NAV_BA_BAD_ARG;
        }
    if (o->precision != NAV_BA_FP64 && o->precision != NAV_BA_MIXED) { err = "unknown precision"; return NAV_BA_BAD_ARG; }
    if (o->n_iter < 0) { err = "negative iteration count"; return NAV_BA_BAD_ARG; }
    for (int w = 0; w < n_win; ++w)
        if (probs[w].camera_model != NAV_BA_CAM_PINHOLE && probs[w].camera_model != NAV_BA_CAM_KB8) {
            err = "unknown camera model in problem " + std::to_string(w);
            return NAV_BA_BAD_ARG;
        }
    auto kb_differs = [&](const nav_ba_problem& a, const nav_ba_problem& b) {
        return a.camera_model == NAV_BA_CAM_KB8 && (a.k1 != b.k1 || a.k2 != b.k2 || a.k3 != b.k3 || a.k4 != b.k4);
    };
    for (int w = 1; w < n_win; ++w)
        if (probs[w].fx != probs[0].fx || probs[w].fy != probs[0].fy || probs[w].cx != probs[0].cx ||
            probs[w].cy != probs[0].cy || probs[w].camera_model != probs[0].camera_model ||
            kb_differs(probs[w], probs[0])) {
            err = "batched windows must share the camera model and its parameters";
            return NAV_BA_BAD_ARG;
        }
    return NAV_BA_OK;
}

// The graph build of BE_GraphOptim.cpp:79-211 as flat arrays (arguments
// already checked).  NAV_BA_OK or NAV_BA_TOO_LARGE with its message.
int build_layout(const nav_ba_problem* probs, int n_win, const nav_ba_options* o, bool batched, HostLayout& L,
                 std::string& err) {
    const bool compat = o->compat_fp32_io != 0;
    L = HostLayout{};
    L.n_win = n_win;
    L.batched = batched;

    // ---- concatenate the windows; keep only points with observations
    L.win.assign(n_win, WinHost());
    L.cam_off.assign(n_win + 1, 0);
    L.obs_off.assign(n_win + 1, 0);
    for (int w = 0; w < n_win; ++w) {
        L.cam_off[w + 1] = L.cam_off[w] + probs[w].n_cam;
        L.obs_off[w + 1] = L.obs_off[w] + probs[w].n_obs;
    }
    const int nc = L.nc = L.cam_off[n_win];
    const int no = L.no = L.obs_off[n_win];
    std::vector<int> g_cam(no), g_pt(no);
    L.pt_first.assign(n_win + 1, 0);
    L.cam_win.resize(nc);
    int np = 0;
    bool any = false;
    for (int w = 0; w < n_win; ++w) {
        const nav_ba_problem& p = probs[w];
        WinHost& wh = L.win[w];
        wh.n_cam = p.n_cam; wh.n_pt = p.n_pt; wh.n_obs = p.n_obs;
        wh.cam0 = L.cam_off[w]; wh.obs0 = L.obs_off[w];
        wh.T_in.assign(p.T_cw, p.T_cw + 12 * (size_t)p.n_cam);
        wh.X_in.assign(p.X, p.X + 3 * (size_t)p.n_pt);
        for (int c = 0; c < p.n_cam; ++c) L.cam_win[L.cam_off[w] + c] = w;
        std::vector<int> cnt(p.n_pt, 0), loc(p.n_pt, -1);
        for (int i = 0; i < p.n_obs; ++i) cnt[p.obs_pt[i]]++;
        wh.pt_local0 = np;
        for (int i = 0; i < p.n_pt; ++i)
            if (cnt[i] > 0) { loc[i] = np++; wh.pt_orig.push_back(i); L.pt_win.push_back(w); }
        wh.pt_local1 = np;
        wh.empty = (p.n_cam == 0 || wh.pt_orig.empty());
        any = any || !wh.empty;
        for (int i = 0; i < p.n_obs; ++i) {
            g_cam[L.obs_off[w] + i] = L.cam_off[w] + p.obs_cam[i];
            g_pt[L.obs_off[w] + i] = loc[p.obs_pt[i]];
        }
        L.pt_first[w + 1] = np;
    }
    L.np = np;
    L.all_empty = !any;
    if (L.all_empty) return NAV_BA_OK;

    // ---- observations ordered by (point, camera, input order): a stable counting
    // sort by point, then a stable insertion sort by camera inside each point's
    // observations (a few per point)
    std::vector<int> ord(no);
    L.pt_ptr.assign(np + 1, 0);
    {
        for (int i = 0; i < no; ++i) L.pt_ptr[g_pt[i] + 1]++;
        for (int q = 0; q < np; ++q) L.pt_ptr[q + 1] += L.pt_ptr[q];
        std::vector<int> pos(L.pt_ptr.begin(), L.pt_ptr.end() - 1);
        for (int i = 0; i < no; ++i) ord[pos[g_pt[i]]++] = i;
        for (int q = 0; q < np; ++q)
            for (int k = L.pt_ptr[q] + 1; k < L.pt_ptr[q + 1]; ++k) {
                const int i = ord[k], c = g_cam[i];
                int j = k;
                for (; j > L.pt_ptr[q] && g_cam[ord[j - 1]] > c; --j) ord[j] = ord[j - 1];
                ord[j] = i;
            }
    }
    L.o_cam.resize(no); L.o_pt.resize(no); L.o_orig.resize(no);
    L.o_uv.resize(2 * (size_t)no);
    // windows hold contiguous point ranges, so the window of ord[k] never decreases in k
    for (int k = 0, w = 0; k < no; ++k) {
        const int i = ord[k];
        while (i >= L.obs_off[w + 1]) ++w;
        const double* uv = probs[w].obs_uv + 2 * (size_t)(i - L.obs_off[w]);
        L.o_cam[k] = g_cam[i];
        L.o_pt[k] = g_pt[i];
        L.o_orig[k] = i;
        L.o_uv[2 * (size_t)k] = compat ? rf32(uv[0]) : uv[0];  // cv::Point2f measurement (Point2D.hpp:33)
        L.o_uv[2 * (size_t)k + 1] = compat ? rf32(uv[1]) : uv[1];
    }
    L.pt_free.resize(np);
    for (int q = 0; q < np; ++q) {
        const int w = L.pt_win[q];
        const int i = L.win[w].pt_orig[q - L.win[w].pt_local0];
        L.pt_free[q] = (probs[w].pt_fixed && probs[w].pt_fixed[i]) ? 0 : 1;
    }

    // ---- free cameras: not fixed and with at least one (therefore active) edge
    std::vector<int> cnt_cam(nc, 0);
    for (int i = 0; i < no; ++i) cnt_cam[g_cam[i]]++;
    L.cam_free.assign(nc, -1);
    L.win_fc0.assign(n_win + 1, 0);
    L.win_cam0.assign(n_win + 1, 0);
    for (int w = 0; w < n_win; ++w) {
        L.win_cam0[w] = L.cam_off[w];
        L.win_fc0[w] = (int)L.free_cam.size();
        for (int c = L.cam_off[w]; c < L.cam_off[w + 1]; ++c) {
            const int lc = c - L.cam_off[w];
            if (!(probs[w].cam_fixed && probs[w].cam_fixed[lc]) && cnt_cam[c] > 0) {
                L.cam_free[c] = (int)L.free_cam.size();
                L.free_cam.push_back(c);
            }
        }
    }
    L.win_cam0[n_win] = nc;
    L.win_fc0[n_win] = (int)L.free_cam.size();
    const int nfc = L.nfc = (int)L.free_cam.size();

    // ---- camera-major observation lists of free cameras (ascending observation index)
    L.cam_ptr.assign(nfc + 1, 0);
    for (int k = 0; k < no; ++k) if (L.cam_free[L.o_cam[k]] >= 0) L.cam_ptr[L.cam_free[L.o_cam[k]] + 1]++;
    for (int f = 0; f < nfc; ++f) L.cam_ptr[f + 1] += L.cam_ptr[f];
    L.cam_obs.resize(L.cam_ptr[nfc]);
    {
        std::vector<int> pos(L.cam_ptr.begin(), L.cam_ptr.end() - 1);
        for (int k = 0; k < no; ++k) {
            const int f = L.cam_free[L.o_cam[k]];
            if (f >= 0) L.cam_obs[pos[f]++] = k;
        }
    }
    L.c_pt.resize(L.cam_obs.size());
    L.c_uv.resize(2 * L.cam_obs.size());
    for (size_t k = 0; k < L.cam_obs.size(); ++k) {
        L.c_pt[k] = L.o_pt[L.cam_obs[k]];
        L.c_uv[2 * k] = L.o_uv[2 * (size_t)L.cam_obs[k]];
        L.c_uv[2 * k + 1] = L.o_uv[2 * (size_t)L.cam_obs[k] + 1];
    }
    // first observation of the same point on the same camera (start of the
    // point's observations on cameras >= this one; duplicates are adjacent)
    L.cam_obs_b0.resize(L.cam_obs.size());
    for (size_t k = 0; k < L.cam_obs.size(); ++k) {
        int b = L.cam_obs[k];
        const int q = L.o_pt[b];
        while (b > L.pt_ptr[q] && L.o_cam[b - 1] == L.o_cam[L.cam_obs[k]]) --b;
        L.cam_obs_b0[k] = b;
    }

    // ---- point-aligned linearisation tiles, never spanning two windows
    L.tile_obs = {0}; L.tile_pt = {0}; L.win_tile0 = {0};
    for (int w = 0; w < n_win; ++w) {
        for (int q = L.pt_first[w]; q < L.pt_first[w + 1];) {
            int q1 = q;
            while (q1 < L.pt_first[w + 1] && L.pt_ptr[q1 + 1] - L.pt_ptr[q] <= kTile) ++q1;
            if (q1 == q) q1 = q + 1;  // one point with more than kTile observations
            L.tile_obs.push_back(L.pt_ptr[q1]);
            L.tile_pt.push_back(q1);
            L.tile_win.push_back(w);
            q = q1;
        }
        L.win_tile0.push_back((int)L.tile_win.size());
    }
    // ---- window-aligned cost blocks (observations) and point blocks; windows
    // are contiguous, so each block starts where the previous one ends
    L.cblk_o0 = {0}; L.win_cblk0 = {0};
    L.win_pt0.assign(n_win + 1, 0);
    for (int w = 0; w < n_win; ++w) {
        const int o0 = L.pt_ptr[L.pt_first[w]], o1 = L.pt_ptr[L.pt_first[w + 1]];
        for (int a = o0; a < o1; a += kCostObs) { L.cblk_o0.push_back(std::min(a + kCostObs, o1)); L.cblk_win.push_back(w); }
        L.win_cblk0.push_back((int)L.cblk_win.size());
        L.win_pt0[w] = L.pt_first[w];
    }
    L.win_pt0[n_win] = np;

    // ---- Hpp work items: (free camera, chunk)
    L.cam_hitem_ptr = {0};
    for (int f = 0; f < nfc; ++f) {
        for (int a = L.cam_ptr[f]; a < L.cam_ptr[f + 1]; a += kChunk) {
            L.hf.push_back(f); L.ha0.push_back(a); L.ha1.push_back(std::min(a + kChunk, L.cam_ptr[f + 1]));
        }
        L.cam_hitem_ptr.push_back((int)L.hf.size());
    }

    // ---- co-visibility through free points, per window: column blocks g >= f of each S row
    // ... and the other blocks g > f of each row (zero in S, rewritten every trial
    // because the factorisation works in place)
    L.colptr = {0}; L.zcolptr = {0};
    {
        std::vector<uint64_t> bits;
        std::vector<int> fl;
        for (int w = 0; w < n_win; ++w) {
            const int f0 = L.win_fc0[w], nw = L.win_fc0[w + 1] - f0;
            bits.assign(((size_t)nw * nw + 63) / 64, 0);
            for (int q = L.pt_first[w]; q < L.pt_first[w + 1]; ++q) {
                if (!L.pt_free[q]) continue;
                fl.clear();
                for (int k = L.pt_ptr[q]; k < L.pt_ptr[q + 1]; ++k) {
                    const int f = L.cam_free[L.o_cam[k]];
                    if (f >= 0 && (fl.empty() || fl.back() != f - f0)) fl.push_back(f - f0);
                }
                for (size_t a = 0; a < fl.size(); ++a)
                    for (size_t b = a + 1; b < fl.size(); ++b) {
                        const size_t bit = (size_t)fl[a] * nw + fl[b];
                        bits[bit >> 6] |= 1ull << (bit & 63);
                    }
            }
            for (int lf = 0; lf < nw; ++lf) {
                L.colcam.push_back(f0 + lf);
                for (int lg = lf + 1; lg < nw; ++lg) {
                    const size_t bit = (size_t)lf * nw + lg;
                    if (bits[bit >> 6] >> (bit & 63) & 1ull) L.colcam.push_back(f0 + lg);
                    else L.zcolcam.push_back(f0 + lg);
                }
                L.colptr.push_back((int)L.colcam.size());
                L.zcolptr.push_back((int)L.zcolcam.size());
            }
        }
    }

    // ---- windows by free-camera count, largest first (the batched factorisation's order)
    L.win_chol_order.resize(n_win);
    for (int w = 0; w < n_win; ++w) L.win_chol_order[w] = w;
    std::stable_sort(L.win_chol_order.begin(), L.win_chol_order.end(), [&](int a, int b) {
        return L.win_fc0[a + 1] - L.win_fc0[a] > L.win_fc0[b + 1] - L.win_fc0[b];
    });

    // ---- dense reduced systems: one padded tiled system, or one small system per window
    L.win_S.assign(n_win, 0);
    L.win_lda.assign(n_win, 1);
    L.n = 6 * nfc;
    L.n_pad = ((L.n + kNB - 1) / kNB) * kNB;
    if (!batched) {
        if (L.n > kPcgMaxN) { err = "reduced camera system larger than 49152"; return NAV_BA_TOO_LARGE; }
        L.win_lda[0] = std::max(L.n_pad, 1);
        L.S_size = std::max<long long>((long long)L.n_pad * L.n_pad, 1);
    } else {
        L.S_size = 0;
        for (int w = 0; w < n_win; ++w) {
            const int nw = 6 * (L.win_fc0[w + 1] - L.win_fc0[w]);
            if (nw > kSmallMaxN) { err = "a batched window has more than 30 free cameras"; return NAV_BA_TOO_LARGE; }
            L.win_S[w] = L.S_size;
            L.win_lda[w] = std::max(nw, 1);
            L.S_size += (long long)L.win_lda[w] * L.win_lda[w];
        }
        L.n_pad = 0;
    }

    // ---- initial state: SE3Quat(R, t) per frame (BE_GraphOptim.cpp:92-93), points
    // rounded to float in compat mode (:147)
    L.pose0.resize(nc);
    for (int w = 0; w < n_win; ++w)
        for (int c = 0; c < probs[w].n_cam; ++c) {
            const double* T = probs[w].T_cw + 12 * (size_t)c;
            PoseQT& P0 = L.pose0[L.cam_off[w] + c];
            host_quat_from_R(T, P0.q);
            P0.t[0] = T[9]; P0.t[1] = T[10]; P0.t[2] = T[11];
            P0.pad = 0;
        }
    L.X0.resize(np);
    for (int q = 0; q < np; ++q) {
        const int w = L.pt_win[q];
        const double* x = probs[w].X + 3 * (size_t)L.win[w].pt_orig[q - L.win[w].pt_local0];
        L.X0[q] = compat ? D4{rf32(x[0]), rf32(x[1]), rf32(x[2]), 0.0} : D4{x[0], x[1], x[2], 0.0};
    }
    // intrinsics are float in the reference's calibration (GeometricCamera.h:117)
    L.fx = compat ? rf32(probs[0].fx) : probs[0].fx; L.fy = compat ? rf32(probs[0].fy) : probs[0].fy;
    L.cx = compat ? rf32(probs[0].cx) : probs[0].cx; L.cy = compat ? rf32(probs[0].cy) : probs[0].cy;
    // the fisheye coefficients are float parameters too (mvParameters[4..7])
    L.cam_model = probs[0].camera_model;
    const bool kb = L.cam_model == NAV_BA_CAM_KB8;
    const double kin[4] = {probs[0].k1, probs[0].k2, probs[0].k3, probs[0].k4};
    for (int i = 0; i < 4; ++i) L.kb[i] = kb ? (compat ? rf32(kin[i]) : kin[i]) : 0.0;
    return NAV_BA_OK;
}

// Every index of the layout and of the Schur items against the buffers they
// address, and the items' pairs against a direct enumeration of each row's
// pairs; "" if consistent, else the first violation.
std::string check_layout(const HostLayout& L, const SchurItems& S) {
    auto bad = [](const std::string& m) { return m; };
    const int no = L.no, np = L.np, nc = L.nc, nfc = L.nfc;
    {
        std::vector<char> hit(L.n_win, 0);
        if ((int)L.win_chol_order.size() != L.n_win) return bad("window order size");
        for (int i = 0; i < L.n_win; ++i) {
            const int w = L.win_chol_order[i];
            if (w < 0 || w >= L.n_win || hit[w]) return bad("window order is not a permutation");
            hit[w] = 1;
            if (i > 0) {
                const int p = L.win_chol_order[i - 1];
                if (L.win_fc0[p + 1] - L.win_fc0[p] < L.win_fc0[w + 1] - L.win_fc0[w]) return bad("window order");
            }
        }
    }
    for (int k = 0; k < no; ++k) {
        if (L.o_cam[k] < 0 || L.o_cam[k] >= nc || L.o_pt[k] < 0 || L.o_pt[k] >= np) return bad("observation index");
        if (k < L.pt_ptr[L.o_pt[k]] || k >= L.pt_ptr[L.o_pt[k] + 1]) return bad("pt_ptr range");
        if (k > 0 && L.o_pt[k] == L.o_pt[k - 1] && L.o_cam[k] < L.o_cam[k - 1]) return bad("camera order in a point");
        if (k > 0 && L.o_pt[k] == L.o_pt[k - 1] && L.o_cam[k] == L.o_cam[k - 1] && L.o_orig[k] < L.o_orig[k - 1])
            return bad("input order among a point's observations on one camera");
    }
    {
        std::vector<char> hit(no, 0);
        for (int k = 0; k < no; ++k) {
            const int i = L.o_orig[k];
            if (i < 0 || i >= no || hit[i]) return bad("observation order is not a permutation");
            hit[i] = 1;
        }
    }
    for (size_t k = 0; k < L.cam_obs.size(); ++k) {
        const int a = L.cam_obs[k];
        if (a < 0 || a >= no || L.c_pt[k] != L.o_pt[a]) return bad("camera-major list");
        const int b0 = L.cam_obs_b0[k];
        if (b0 > a || b0 < L.pt_ptr[L.o_pt[a]] || L.o_cam[b0] != L.o_cam[a]) return bad("cam_obs_b0");
    }
    // rounds of kBlock observations over kBlock lanes bound per round, or of
    // kPcRound observations with kPcCons consumer lanes bound per item
    const int lanes = S.pc ? kPcCons : kBlock;
    const int rlen = S.pc ? kPcRound : kBlock;
    if ((int)S.sf.size() != (int)S.srnd0.size()) return bad("item round table");
    {
        std::vector<char> hit(S.sf.size(), 0);
        if (S.sorder.size() != S.sf.size()) return bad("dispatch order size");
        for (int i : S.sorder) {
            if (i < 0 || i >= (int)S.sf.size() || hit[i]) return bad("dispatch order is not a permutation");
            hit[i] = 1;
        }
    }
    const size_t nrnd = S.rnd_rows.size();
    if (S.lane_slot.size() != (S.pc ? S.sf.size() : nrnd) * lanes || S.rnd_group.size() != nrnd ||
        S.rnd_pair0.size() != nrnd)
        return bad("round tables");
    std::vector<int> slotmap(nc, -1);
    std::vector<long long> seen;
    for (size_t it = 0; it < S.sf.size(); ++it) {
        const int f = S.sf[it], ns = S.sns[it], a0 = S.sa0[it], a1 = S.sa1[it];
        if (f < 0 || f >= nfc || ns < 1 || ns > kMaxSlots || a0 < L.cam_ptr[f] || a1 > L.cam_ptr[f + 1] || a0 > a1)
            return bad("item range");
        const int col0 = L.colptr[f] + S.sg[it] * kMaxSlots;
        if (col0 + ns > L.colptr[f + 1]) return bad("item slots");
        for (int sl = 0; sl < ns; ++sl) slotmap[L.free_cam[L.colcam[col0 + sl]]] = sl;
        const int nr = (a1 - a0 + rlen - 1) / rlen;
        if (S.srnd0[it] + nr > (int)nrnd) return bad("item rounds");
        if (it + 1 < S.sf.size() && S.srnd0[it + 1] != S.srnd0[it] + nr) return bad("item rounds");
        for (int r = 0; r < nr; ++r) {
            const int rnd = S.srnd0[it] + r, kb = a0 + r * rlen;
            const int grp = S.pc ? 1 : S.rnd_group[rnd];
            if (!S.pc && grp != 2 && grp != 4) return bad("lane group");
            // expected pairs of the round, as (b, row) keys
            std::vector<long long> want;
            for (int k = kb; k < std::min(kb + rlen, a1); ++k) {
                const int a = L.cam_obs[k], q = L.o_pt[a];
                if (!L.pt_free[q]) continue;
                for (int b = L.cam_obs_b0[k]; b < L.pt_ptr[q + 1]; ++b)
                    if (b != a && slotmap[L.o_cam[b]] >= 0) want.push_back((long long)b * 512 + (k - kb));
            }
            // the lanes' slots: contiguous per slot, aligned whole groups
            const uint8_t* ls = &S.lane_slot[(S.pc ? it : (size_t)rnd) * lanes];
            for (int t = 0; t < lanes; ++t) {
                if (ls[t] != 255 && ls[t] >= ns) return bad("lane slot");
                if (t > 0 && ls[t] != ls[t - 1]) {
                    if (t % grp != 0) return bad("lane group alignment");
                    if (ls[t - 1] != 255 && ls[t] != 255 && ls[t] < ls[t - 1]) return bad("lane slot order");
                    for (int u = 0; u < t - 1; ++u)
                        if (ls[u] == ls[t] && ls[t] != 255) return bad("slot lanes not contiguous");
                }
            }
            const int4 rw = S.rnd_rows[rnd];
            const int rowsv[4] = {rw.x, rw.y, rw.z, rw.w};
            if (rw.x < 0 || rw.y < 0 || rw.z < 0 || rw.w < 0 || (S.pc && (rw.z != 0 || rw.w != 0)))
                return bad("round rows");
            seen.clear();
            size_t e = (size_t)S.rnd_pair0[rnd];
            for (int wv = 0; wv < lanes / kWave; ++wv)
                for (int j = 0; j < rowsv[wv]; ++j)
                    for (int l = 0; l < kWave; ++l, ++e) {
                        if (e >= S.pairs.size()) return bad("pair list end");
                        const int2 pe = S.pairs[e];
                        if (pe.x < 0) continue;
                        const int t = wv * kWave + l;
                        if (ls[t] == 255) return bad("pair on an idle lane");
                        if (pe.x >= no || pe.y < 0 || pe.y >= rlen || kb + pe.y >= a1) return bad("pair index");
                        if (slotmap[L.o_cam[pe.x]] != ls[t]) return bad("pair camera is not the lane's slot");
                        if (L.o_pt[pe.x] != L.o_pt[L.cam_obs[kb + pe.y]]) return bad("pair point");
                        seen.push_back((long long)pe.x * 512 + pe.y);
                    }
            std::sort(want.begin(), want.end());
            std::sort(seen.begin(), seen.end());
            if (want != seen) return bad("round pairs differ from the direct enumeration");
        }
        for (int sl = 0; sl < ns; ++sl) slotmap[L.free_cam[L.colcam[col0 + sl]]] = -1;
    }
    return "";
}

// Build the device structure of n_win windows laid end to end (see ba_kernels.h).
// Everything a prepare's structure depends on: sizes, observation indices,
// fixed flags, the camera, the options (they are baked into the captured step
// graphs) and the handle's forced settings.  Poses, points and measurements
// are not part of it.
void make_prep_key(const nav_ba_handle* h, const nav_ba_problem* probs, int n_win, const nav_ba_options* o,
                   bool batched, std::vector<char>& key) {
    key.clear();
    auto put = [&](const void* p, size_t n) {
        const char* c = (const char*)p;
        key.insert(key.end(), c, c + n);
    };
    auto put_i = [&](long long v) { put(&v, sizeof v); };
    auto put_d = [&](double v) { put(&v, sizeof v); };
    put_i(n_win); put_i(batched); put_i(h->solver_mode); put_i(h->schur_chunk); put_i(h->schur_kernel);
    put_i(o->n_iter); put_i(o->robust); put_d(o->huber_delta); put_d(o->lambda_init); put_i(o->precision);
    put_i(o->compat_fp32_io); put_d(o->outlier_chi2); put_i(o->max_trials);
    for (int w = 0; w < n_win; ++w) {
        const nav_ba_problem& p = probs[w];
        put_i(p.n_cam); put_i(p.n_pt); put_i(p.n_obs); put_i(p.camera_model);
        put_d(p.fx); put_d(p.fy); put_d(p.cx); put_d(p.cy); put_d(p.k1); put_d(p.k2); put_d(p.k3); put_d(p.k4);
        put_i(p.cam_fixed != nullptr);
        if (p.cam_fixed) put(p.cam_fixed, (size_t)p.n_cam);
        put_i(p.pt_fixed != nullptr);
        if (p.pt_fixed) put(p.pt_fixed, (size_t)p.n_pt);
        put(p.obs_cam, sizeof(int32_t) * (size_t)p.n_obs);
        put(p.obs_pt, sizeof(int32_t) * (size_t)p.n_obs);
    }
}

// A prepare whose key equals the last one's: the host layout, work items,
// tuned choices and step graphs stay; the new initial poses, points and
// measurements (in the orders prepare chose) go to the device.
int refresh_windows(nav_ba_handle* h, const nav_ba_problem* probs, int n_win, const nav_ba_options* o) {
    HostLayout& L = h->layout;
    const bool compat = o->compat_fp32_io != 0;
    for (int w = 0; w < n_win; ++w) {
        const nav_ba_problem& p = probs[w];
        L.win[w].T_in.assign(p.T_cw, p.T_cw + 12 * (size_t)p.n_cam);
        L.win[w].X_in.assign(p.X, p.X + 3 * (size_t)p.n_pt);
        // a window's observations are contiguous in point-major order
        for (int k = L.pt_ptr[L.pt_first[w]]; k < L.pt_ptr[L.pt_first[w + 1]]; ++k) {
            const double* uv = p.obs_uv + 2 * (size_t)(L.o_orig[k] - L.obs_off[w]);
            L.o_uv[2 * (size_t)k] = compat ? rf32(uv[0]) : uv[0];
            L.o_uv[2 * (size_t)k + 1] = compat ? rf32(uv[1]) : uv[1];
        }
        for (int c = 0; c < p.n_cam; ++c) {
            const double* T = p.T_cw + 12 * (size_t)c;
            PoseQT& P0 = L.pose0[L.cam_off[w] + c];
            host_quat_from_R(T, P0.q);
            P0.t[0] = T[9]; P0.t[1] = T[10]; P0.t[2] = T[11];
            P0.pad = 0;
        }
        for (int q = L.pt_first[w]; q < L.pt_first[w + 1]; ++q) {
            const double* x = p.X + 3 * (size_t)L.win[w].pt_orig[q - L.win[w].pt_local0];
            L.X0[q] = compat ? D4{rf32(x[0]), rf32(x[1]), rf32(x[2]), 0.0} : D4{x[0], x[1], x[2], 0.0};
        }
    }
    for (size_t k = 0; k < L.cam_obs.size(); ++k) {
        L.c_uv[2 * k] = L.o_uv[2 * (size_t)L.cam_obs[k]];
        L.c_uv[2 * k + 1] = L.o_uv[2 * (size_t)L.cam_obs[k] + 1];
    }
    // through the pinned staging of the first prepare (it held these arrays and more)
    const size_t b_ouv = sizeof(double) * L.o_uv.size(), b_cuv = sizeof(double) * L.c_uv.size();
    const size_t b_pose = sizeof(PoseQT) * L.pose0.size(), b_x = sizeof(D4) * L.X0.size();
    const size_t need = b_ouv + b_cuv + b_pose + b_x;
    if (h->stage_cap < need) {
        if (h->stage) (void)hipHostFree(h->stage);
        h->stage = nullptr;
        h->stage_cap = 0;
        HIPTRY(h, hipHostMalloc((void**)&h->stage, need, hipHostMallocDefault));
        h->stage_cap = need;
    }
    char* sp = h->stage;
    std::memcpy(sp, L.o_uv.data(), b_ouv);
    std::memcpy(sp + b_ouv, L.c_uv.data(), b_cuv);
    std::memcpy(sp + b_ouv + b_cuv, L.pose0.data(), b_pose);
    std::memcpy(sp + b_ouv + b_cuv + b_pose, L.X0.data(), b_x);
    hipStream_t s = h->stream;
    HIPTRY(h, hipMemcpyAsync((void*)h->P.o_uv, sp, b_ouv, hipMemcpyHostToDevice, s));
    if (b_cuv) HIPTRY(h, hipMemcpyAsync((void*)h->P.c_uv, sp + b_ouv, b_cuv, hipMemcpyHostToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->init.pose, sp + b_ouv + b_cuv, b_pose, hipMemcpyHostToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->init.X, sp + b_ouv + b_cuv + b_pose, b_x, hipMemcpyHostToDevice, s));
    HIPTRY(h, hipStreamSynchronize(s));
    h->ran = false;
    h->prepared = true;
    return NAV_BA_OK;
}

int prepare_windows(nav_ba_handle* h, const nav_ba_problem* probs, int n_win, const nav_ba_options* o, bool batched) {
    if (!h) return NAV_BA_BAD_ARG;
    h->err.clear();
    std::string msg;
    int st = check_args(probs, n_win, o, batched, msg);
    if (st != NAV_BA_OK) {
        h->prepared = false;
        h->prep_key.clear();
        return fail(h, st, msg);
    }
    std::vector<char> pkey;
    make_prep_key(h, probs, n_win, o, batched, pkey);
    h->prep_cached = false;
    if (h->prepared && !h->all_empty && pkey == h->prep_key) {
        h->prep_cached = true;
        HIPTRY(h, hipSetDevice(h->device));
        HIPTRY(h, hipStreamSynchronize(h->stream));
        return refresh_windows(h, probs, n_win, o);
    }
    h->prep_key.clear();
    h->prepared = false;
    h->ran = false;
    h->schur_chunk_used = 0;
    HIPTRY(h, hipSetDevice(h->device));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    if (h->step) { (void)hipGraphExecDestroy(h->step); h->step = nullptr; }
    if (h->step0) { (void)hipGraphExecDestroy(h->step0); h->step0 = nullptr; }
    h->opt = *o;
    if (h->opt.max_trials <= 0) h->opt.max_trials = 10;
    h->batched = batched;
    HostLayout& L = h->layout;
    if ((st = build_layout(probs, n_win, o, batched, L, msg)) != NAV_BA_OK) return fail(h, st, msg);
    h->all_empty = L.all_empty;
    if (L.all_empty) {
        h->prepared = true;
        return NAV_BA_OK;
    }
    const int nc = L.nc, np = L.np, no = L.no, nfc = L.nfc;

    // ---- Schur work items and pair lists at the default chunk size (tuned below)
    const SchurGraph G = L.graph();
    const TuneKey key{h->device, mixed(h) ? 1 : 0, n_win, nc, np, no, nfc, h->schur_kernel, L.cam_model};
    // producer/consumer by default for batches and for the fisheye camera (whose
    // Jacobians leave the rounds kernel short of registers)
    TuneChoice choice{h->schur_chunk > 0 ? h->schur_chunk : kSchurChunk,
                      h->schur_kernel == 2 ||
                          (h->schur_kernel == 0 && (batched || L.cam_model == NAV_BA_CAM_KB8))};
    bool tune = (h->schur_chunk == 0 || h->schur_kernel == 0) && nfc > 0 && no >= kTuneMinObs;
    if (tune && cached_choice(key, &choice)) tune = false;  // this shape was tuned before on this device
    SchurItems SI;
    if (!build_schur_items(G, choice.chunk, choice.pc, SI))
        return fail(h, NAV_BA_TOO_LARGE, "more than 2^31 Schur pair entries");

    // ---- device upload
    h->nbuf = 0;
    DevProblem& P = h->P;
    P = DevProblem{};
    P.n_cam = nc; P.n_pt = np; P.n_obs = no; P.n_fc = nfc;
    P.robust = o->robust ? 1 : 0;
    P.delta = o->huber_delta;
    P.fx = L.fx; P.fy = L.fy; P.cx = L.cx; P.cy = L.cy;
    P.cam_model = L.cam_model == NAV_BA_CAM_KB8 ? kCamKB8 : kCamPinhole;
    P.k1 = L.kb[0]; P.k2 = L.kb[1]; P.k3 = L.kb[2]; P.k4 = L.kb[3];
    P.n_iter = o->n_iter;
    P.max_trials = h->opt.max_trials;
    P.auto_lambda = !(o->lambda_init > 0);
    P.lambda0 = o->lambda_init;
    Packer pk(&h->stage, &h->stage_cap);  // free again: every prepare ends with a stream synchronisation
#define UP(dst, vec) pk.add(&(dst), (vec))
    UP(P.o_cam, L.o_cam); UP(P.o_pt, L.o_pt); UP(P.o_uv, L.o_uv); UP(P.o_orig, L.o_orig);
    UP(P.pt_ptr, L.pt_ptr); UP(P.pt_free, L.pt_free); UP(P.cam_free, L.cam_free); UP(P.free_cam, L.free_cam);
    UP(P.cam_ptr, L.cam_ptr); UP(P.cam_obs, L.cam_obs); UP(P.cam_obs_b0, L.cam_obs_b0);
    UP(P.c_pt, L.c_pt); UP(P.c_uv, L.c_uv);
    UP(P.tile_obs, L.tile_obs); UP(P.tile_pt, L.tile_pt); UP(P.tile_win, L.tile_win);
    UP(P.hitem_f, L.hf); UP(P.hitem_a0, L.ha0); UP(P.hitem_a1, L.ha1); UP(P.cam_hitem_ptr, L.cam_hitem_ptr);
    UP(P.colptr, L.colptr); UP(P.colcam, L.colcam); UP(P.zcolptr, L.zcolptr); UP(P.zcolcam, L.zcolcam);
    UP(P.cam_win, L.cam_win); UP(P.pt_win, L.pt_win); UP(P.win_cam0, L.win_cam0); UP(P.win_fc0, L.win_fc0);
    UP(P.win_pt0, L.win_pt0); UP(P.win_S, L.win_S); UP(P.win_lda, L.win_lda);
    UP(P.cblk_o0, L.cblk_o0); UP(P.cblk_win, L.cblk_win); UP(P.win_cblk0, L.win_cblk0);
    UP(P.win_tile0, L.win_tile0); UP(P.win_chol_order, L.win_chol_order);
#undef UP
    if ((st = pk.flush(h)) != NAV_BA_OK) return st;
    P.n_tiles = (int)L.tile_win.size();
    P.n_hitems = (int)L.hf.size();
    P.n_win = n_win;
    P.max_win_fc = 0;
    for (int w = 0; w < n_win; ++w) P.max_win_fc = std::max(P.max_win_fc, L.win_fc0[w + 1] - L.win_fc0[w]);
    P.batched = batched ? 1 : 0;
    P.n_cblk = (int)L.cblk_win.size();
    P.n = L.n;
    P.n_pad = L.n_pad;
    P.nt = L.n_pad / kNB;
    // factorisation: in LDS per window (every batch; one problem of at most 30
    // free cameras unless another is asked for), 64x64 tiles, or on request
    // 6x6 blocks in global memory
    P.use_win = (batched || (nfc > 0 && 6 * nfc <= kSmallMaxN && (h->solver_mode == 0 || h->solver_mode == 3))) ? 1 : 0;
    P.use_pcg = (!batched && nfc > 0 && (h->solver_mode == 4 || L.n > kMaxDenseN)) ? 1 : 0;
    if (P.use_pcg) P.use_win = 0;
    P.use_blocks = (!P.use_win && !P.use_pcg && nfc > 0 && nfc <= kBlocksMaxFc && h->solver_mode == 2) ? 1 : 0;
    {
        int maxcam = 0;
        for (int w = 0; w < n_win; ++w) maxcam = std::max(maxcam, L.cam_off[w + 1] - L.cam_off[w]);
        P.commit_in_accept = maxcam <= kCommitInAcceptCams ? 1 : 0;
    }

    const size_t tsz = mixed(h) ? sizeof(float) : sizeof(double);
    const size_t nvec = batched ? std::max<size_t>(6 * (size_t)nfc, 1) : std::max<size_t>(L.n_pad, 1);
    auto A = [&](void** dst, size_t bytes) { return alloc(h, dst, bytes); };
#define AL(dst, bytes) if ((st = A((void**)&(dst), (bytes))) != NAV_BA_OK) return st
    AL(h->init.pose, sizeof(PoseQT) * nc); AL(h->init.X, sizeof(D4) * np); h->init.cam = nullptr;
    AL(h->cur.pose, sizeof(PoseQT) * nc); AL(h->cur.cam, sizeof(CamRT) * nc); AL(h->cur.X, sizeof(D4) * np);
    AL(h->cand.pose, sizeof(PoseQT) * nc); AL(h->cand.cam, sizeof(CamRT) * nc); AL(h->cand.X, sizeof(D4) * np);
    DevWork& W = h->W;
    AL(W.w, tsz * no);
    AL(W.Hl, sizeof(double) * 9 * (size_t)np);
    AL(W.hpart, sizeof(double) * 27 * std::max<size_t>(L.hf.size(), 1));
    AL(W.Hc, sizeof(double) * 28 * std::max(nfc, 1));
    AL(W.S, sizeof(double) * (batched ? (size_t)1 : (size_t)std::max<long long>(L.S_size, 1)));
    AL(W.g, sizeof(double) * nvec);
    AL(W.y, sizeof(double) * nvec);
    AL(W.dp, sizeof(double) * nvec);
    AL(W.Lm, sizeof(double) * ((batched || P.use_pcg) ? (size_t)1 : (size_t)std::max<long long>(L.S_size, 1)));
    AL(W.Li, sizeof(double) * 4096 * (batched ? (size_t)1 : (size_t)std::max(P.nt, 1)));
    AL(W.fpart, sizeof(double) * 64 * (batched ? (size_t)1 : (size_t)std::max(P.nt, 1) * std::max(P.nt, 1)));
    AL(W.red_cost, sizeof(double) * std::max(P.n_cblk, 1));
    AL(W.red_tcost, sizeof(double) * std::max(P.n_tiles, 1));
    AL(W.red_scale, sizeof(double) * (n_win + std::max(P.n_tiles, 1)));
    AL(W.red_aux, sizeof(double) * std::max(P.n_cblk, 1));
    AL(W.outlier, std::max<size_t>(no, 1));
    AL(W.hist, sizeof(double) * (size_t)n_win * (o->n_iter + 1));
    AL(W.final, sizeof(double) * 2 * (size_t)n_win);
    AL(W.done_count, sizeof(int));
    AL(W.lm, sizeof(LMDev) * n_win);
    AL(W.pcg, sizeof(double) * (P.use_pcg ? pcg_work_doubles(P.n) : (size_t)1));
    AL(W.pcg_st, sizeof(PcgState));
#undef AL
    P.xb0 = h->cur.X;
    P.xb1 = h->cand.X;
    const int schur_mark = h->nbuf;
    if ((st = upload_schur_items(h, SI)) != NAV_BA_OK) return st;
    HIPTRY(h, hipMemcpyAsync(h->init.pose, L.pose0.data(), sizeof(PoseQT) * nc, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipMemcpyAsync(h->init.X, L.X0.data(), sizeof(D4) * np, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    // Schur rows and the per-window factorisation may use more than 64 KiB of LDS
    if ((mixed(h) ? prepare_schur_row<float>(P) : prepare_schur_row<double>(P)) != 0)
        return fail(h, NAV_BA_TOO_LARGE, "Schur row workspace exceeds the LDS of one workgroup");
    if (!P.use_pcg && prepare_cholesky(P) != 0) return fail(h, NAV_BA_HIP_ERROR, "LDS attribute of the factorisation kernels");
    if (P.use_win && prepare_cholesky_small(P) != 0) return fail(h, NAV_BA_HIP_ERROR, "LDS attribute of the window factorisation");
    launch_fill_system(P, h->W, h->stream);
    if (tune) {
        if ((st = tune_schur_items(h, G, schur_mark, h->schur_kernel, h->schur_chunk, SI)) != NAV_BA_OK) return st;
        remember_choice(key, TuneChoice{SI.chunk, SI.pc});
    }
    if (!P.use_pcg && (st = capture_step(h)) != NAV_BA_OK) return st;
    h->prepared = true;
    h->prep_key.swap(pkey);
    return NAV_BA_OK;
}

int fetch_window(nav_ba_handle* h, int w, const std::vector<PoseQT>& pose, const std::vector<D4>& X,
                 const std::vector<uint8_t>& flags, nav_ba_result* r) {
    const WinHost& wh = h->layout.win[w];
    r->solve_ms = h->solve_ms;
    if (wh.empty || h->all_empty) {
        r->iterations = 0; r->trials = 0; r->stop_reason = NAV_BA_STOP_EMPTY; r->n_outliers = 0;
        r->initial_cost = r->final_cost = r->final_lambda = 0.0;
        if (r->T_cw) std::memcpy(r->T_cw, wh.T_in.data(), sizeof(double) * 12 * (size_t)wh.n_cam);
        if (r->X) std::memcpy(r->X, wh.X_in.data(), sizeof(double) * 3 * (size_t)wh.n_pt);
        if (r->outlier) std::memset(r->outlier, 0, (size_t)wh.n_obs);
        return NAV_BA_EMPTY;
    }
    const LMDev& L = h->lm_out[w];
    const double* hist = h->hist_out.data() + (size_t)w * (h->P.n_iter + 1);
    r->iterations = L.iter;
    r->trials = L.trials;
    r->stop_reason = L.stop == 1 ? NAV_BA_STOP_TERMINATE : NAV_BA_STOP_MAX_ITER;
    r->initial_cost = hist[0];
    r->final_cost = h->final_out[2 * (size_t)w];
    r->n_outliers = (int)h->final_out[2 * (size_t)w + 1];
    r->final_lambda = L.lambda;
    if (r->cost_history)
        for (int i = 0; i <= L.iter; ++i) r->cost_history[i] = hist[i];
    if (r->T_cw)
        for (int c = 0; c < wh.n_cam; ++c) {
            const PoseQT& q = pose[wh.cam0 + c];
            double* T = r->T_cw + 12 * (size_t)c;
            host_R_from_quat(q.q, T);  // BE_GraphOptim.cpp:220-223
            T[9] = q.t[0]; T[10] = q.t[1]; T[11] = q.t[2];
        }
    if (r->X) {
        std::memcpy(r->X, wh.X_in.data(), sizeof(double) * 3 * (size_t)wh.n_pt);
        const bool compat = h->opt.compat_fp32_io != 0;
        for (int q = wh.pt_local0; q < wh.pt_local1; ++q) {
            double* o = r->X + 3 * (size_t)wh.pt_orig[q - wh.pt_local0];
            o[0] = compat ? rf32(X[q].x) : X[q].x;  // estimate().cast<float>() (BE_GraphOptim.cpp:239)
            o[1] = compat ? rf32(X[q].y) : X[q].y;
            o[2] = compat ? rf32(X[q].z) : X[q].z;
        }
    }
    if (r->outlier && wh.n_obs > 0) std::memcpy(r->outlier, flags.data() + wh.obs0, (size_t)wh.n_obs);
    return NAV_BA_OK;
}

int fetch_all(nav_ba_handle* h, nav_ba_result* results, int n) {
    if (!h || !results) return NAV_BA_BAD_ARG;
    if (!h->prepared || !h->ran) return fail(h, NAV_BA_NOT_PREPARED, "fetch before nav_ba_run");
    if (n != (int)h->layout.win.size()) return fail(h, NAV_BA_BAD_ARG, "result count differs from the prepared window count");
    std::vector<PoseQT> pose;
    std::vector<D4> X, X1;
    std::vector<uint8_t> flags;
    if (!h->all_empty) {
        HIPTRY(h, hipSetDevice(h->device));
        pose.resize(h->P.n_cam);
        X.resize(h->P.n_pt);
        X1.resize(h->P.n_pt);
        flags.resize(h->P.n_obs);
        HIPTRY(h, hipMemcpyAsync(pose.data(), h->cur.pose, sizeof(PoseQT) * pose.size(), hipMemcpyDeviceToHost, h->stream));
        HIPTRY(h, hipMemcpyAsync(X.data(), h->P.xb0, sizeof(D4) * X.size(), hipMemcpyDeviceToHost, h->stream));
        HIPTRY(h, hipMemcpyAsync(X1.data(), h->P.xb1, sizeof(D4) * X1.size(), hipMemcpyDeviceToHost, h->stream));
        if (!flags.empty())
            HIPTRY(h, hipMemcpyAsync(flags.data(), h->W.outlier, flags.size(), hipMemcpyDeviceToHost, h->stream));
        HIPTRY(h, hipStreamSynchronize(h->stream));
        // each window's current points from the state buffer its LM state names
        for (int q = 0; q < h->P.n_pt; ++q)
            if (h->lm_out[h->layout.pt_win[q]].xsel) X[q] = X1[q];
    }
    int worst = NAV_BA_OK;
    for (int w = 0; w < n; ++w) {
        const int st = fetch_window(h, w, pose, X, flags, &results[w]);
        if (st != NAV_BA_OK) worst = st;
    }
    return worst;
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
        hipHostMalloc((void**)&h->hdone, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->hlm, sizeof(LMDev) * kBurstWindows, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->hpcg, sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
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
    if (h->step) (void)hipGraphExecDestroy(h->step);
    if (h->step0) (void)hipGraphExecDestroy(h->step0);
    for (auto& b : h->bufs)
        if (b.p) (void)hipFree(b.p);
    if (h->hdone) (void)hipHostFree(h->hdone);
    if (h->hlm) (void)hipHostFree(h->hlm);
    if (h->hpcg) (void)hipHostFree(h->hpcg);
    if (h->stage) (void)hipHostFree(h->stage);
    if (h->stage_s) (void)hipHostFree(h->stage_s);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* nav_ba_last_error(const nav_ba_handle* h) { return h ? h->err.c_str() : "null handle"; }

int32_t nav_ba_prepare(nav_ba_handle* h, const nav_ba_problem* p, const nav_ba_options* o) {
    return prepare_windows(h, p, 1, o, false);
}

int32_t nav_ba_prepare_batched(nav_ba_handle* h, const nav_ba_problem* windows, int32_t n_windows,
                               const nav_ba_options* o) {
    return prepare_windows(h, windows, n_windows, o, true);
}

int32_t nav_ba_run(nav_ba_handle* h) {
    if (!h) return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "nav_ba_run before nav_ba_prepare");
    h->err.clear();
    h->ran = false;
    if (h->all_empty) {
        h->ran = true;
        return NAV_BA_EMPTY;
    }
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    const bool m = mixed(h);
    HIPTRY(h, hipEventRecord(h->ev0, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    launch_lm_reset(P, h->W, s);
    launch_cam_rt(P, h->cur, s);
    if (m) launch_cost<float>(P, h->cur, h->W, 0, s); else launch_cost<double>(P, h->cur, h->W, 0, s);
    launch_lm_set_chi(P, h->W, s);
    // every window needs at most n_iter * max_trials steps.  Steps are
    // launched in bursts: n_iter first (one per iteration when every trial is
    // accepted); then, for up to kBurstWindows windows, half as many again as
    // the most iterations a running window still has (a step past a window's
    // end returns at once from every kernel, and costs about as much GPU time
    // as the host's check costs idle time), else 2.  The host checks the
    // count of finished windows after each burst.
    const long long max_steps = (long long)P.n_iter * P.max_trials;
    long long steps = 0;
    int st;
    long long burst = std::max(1, P.n_iter);
    const bool few = P.n_win <= kBurstWindows;
    while (steps < max_steps) {
        for (long long i = 0; i < burst && steps < max_steps; ++i, ++steps)
            if ((st = launch_step(h, steps == 0)) != NAV_BA_OK) return st;
        HIPTRY(h, hipMemcpyAsync(h->hdone, h->W.done_count, sizeof(int), hipMemcpyDeviceToHost, s));
        if (few) HIPTRY(h, hipMemcpyAsync(h->hlm, h->W.lm, sizeof(LMDev) * P.n_win, hipMemcpyDeviceToHost, s));
        HIPTRY(h, hipStreamSynchronize(s));
        if (*h->hdone >= P.n_win) break;
        burst = 2;
        if (few) {
            int rem = 0;
            for (int w = 0; w < P.n_win; ++w)
                if (h->hlm[w].phase != kPhaseDone) rem = std::max(rem, P.n_iter - h->hlm[w].iter);
            burst = std::max<long long>(2, rem + (rem + 1) / 2);
        }
    }
    // outlier flags and the FP64 cost at the final state
    h->timing_active = h->timing;
    stage(h, NAV_BA_ST_FLAGS, [&] {
        launch_flags(P, h->cur, h->W, h->opt.outlier_chi2, s);
        launch_final(P, h->W, s);
    });
    h->timing_active = false;
    HIPTRY(h, hipEventRecord(h->ev1, s));
    h->lm_out.resize(P.n_win);
    h->hist_out.resize((size_t)P.n_win * (P.n_iter + 1));
    h->final_out.resize(2 * (size_t)P.n_win);
    HIPTRY(h, hipMemcpyAsync(h->lm_out.data(), h->W.lm, sizeof(LMDev) * P.n_win, hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipMemcpyAsync(h->hist_out.data(), h->W.hist, sizeof(double) * h->hist_out.size(), hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipMemcpyAsync(h->final_out.data(), h->W.final, sizeof(double) * h->final_out.size(), hipMemcpyDeviceToHost, s));
    HIPTRY(h, hipStreamSynchronize(s));
    float ms = 0.f;
    HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->solve_ms = ms;
    if (h->timing) h->timer.collect();
    h->ran = true;
    return NAV_BA_OK;
}

int32_t nav_ba_fetch(nav_ba_handle* h, nav_ba_result* r) { return fetch_all(h, r, 1); }

int32_t nav_ba_fetch_batched(nav_ba_handle* h, nav_ba_result* results, int32_t n_windows) {
    return fetch_all(h, results, n_windows);
}

int32_t nav_ba_solve(nav_ba_handle* h, const nav_ba_problem* p, const nav_ba_options* o, nav_ba_result* r) {
    int st = nav_ba_prepare(h, p, o);
    if (st != NAV_BA_OK) return st;
    st = nav_ba_run(h);
    if (st != NAV_BA_OK && st != NAV_BA_EMPTY) return st;
    const int st2 = nav_ba_fetch(h, r);
    return st2 != NAV_BA_OK ? st2 : st;
}

int32_t nav_ba_solve_batched(nav_ba_handle* h, const nav_ba_problem* windows, int32_t n_windows,
                             const nav_ba_options* o, nav_ba_result* results) {
    int st = nav_ba_prepare_batched(h, windows, n_windows, o);
    if (st != NAV_BA_OK) return st;
    st = nav_ba_run(h);
    if (st != NAV_BA_OK && st != NAV_BA_EMPTY) return st;
    return nav_ba_fetch_batched(h, results, n_windows);
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

int32_t nav_ba_debug_set_solver(nav_ba_handle* h, int32_t mode) {
    if (!h || mode < 0 || mode > 4) return NAV_BA_BAD_ARG;
    h->solver_mode = mode;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_set_schur_chunk(nav_ba_handle* h, int32_t chunk) {
    if (!h || chunk < 0 || chunk > 4096 || chunk % kWave != 0) return NAV_BA_BAD_ARG;
    h->schur_chunk = chunk;
    return NAV_BA_OK;
}


int32_t nav_ba_debug_get_schur_chunk(const nav_ba_handle* h, int32_t* chunk) {
    if (!h || !chunk) return NAV_BA_BAD_ARG;
    *chunk = h->prepared ? h->schur_chunk_used : 0;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_prepare_cached(const nav_ba_handle* h, int32_t* cached) {
    if (!h || !cached) return NAV_BA_BAD_ARG;
    *cached = h->prep_cached ? 1 : 0;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_set_schur_kernel(nav_ba_handle* h, int32_t kernel) {
    if (!h || kernel < 0 || kernel > 2) return NAV_BA_BAD_ARG;
    h->schur_kernel = kernel;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_get_schur_kernel(const nav_ba_handle* h, int32_t* kernel) {
    if (!h || !kernel) return NAV_BA_BAD_ARG;
    *kernel = h->prepared && !h->all_empty ? (h->schur_pc_used ? 2 : 1) : 0;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_time_schur_rows(nav_ba_handle* h, double lambda, int32_t reps, double* avg_us) {
    if (!h || !avg_us || reps <= 0) return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "debug call before nav_ba_prepare");
    *avg_us = 0.0;
    if (h->all_empty || h->P.n_fc == 0) return NAV_BA_EMPTY;
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    std::vector<LMDev> L(P.n_win);
    for (auto& l : L) { l = LMDev{}; l.lambda = lambda; l.nu = 2.0; l.phase = kPhaseLin; l.iter = 1; }
    HIPTRY(h, hipMemcpyAsync(h->W.lm, L.data(), sizeof(LMDev) * L.size(), hipMemcpyHostToDevice, s));
    launch_cam_rt(P, h->cur, s);
    if (mixed(h)) enqueue_lin<float>(h, false); else enqueue_lin<double>(h, false);
    // one warm-up launch, then reps launches between one pair of events
    if (mixed(h)) launch_schur_row<float>(P, h->cur, h->W, s); else launch_schur_row<double>(P, h->cur, h->W, s);
    HIPTRY(h, hipEventRecord(h->ev0, s));
    for (int r = 0; r < reps; ++r) {
        if (mixed(h)) launch_schur_row<float>(P, h->cur, h->W, s);
        else launch_schur_row<double>(P, h->cur, h->W, s);
    }
    HIPTRY(h, hipEventRecord(h->ev1, s));
    HIPTRY(h, hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *avg_us = 1e3 * (double)ms / reps;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_pcg_stats(const nav_ba_handle* h, int32_t* iterations, double* rel_residual) {
    if (!h || !iterations || !rel_residual) return NAV_BA_BAD_ARG;
    if (!h->prepared) return NAV_BA_NOT_PREPARED;
    const PcgState& s = h->pcg_last;
    *iterations = h->P.use_pcg ? s.iter : 0;
    *rel_residual = (h->P.use_pcg && s.rr0 > 0.0) ? std::sqrt(s.rr / s.rr0) : 0.0;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_time_stage(nav_ba_handle* h, int32_t stage, double lambda, int32_t reps, double* avg_us) {
    if (!h || !avg_us || reps <= 0) return NAV_BA_BAD_ARG;
    if (stage != NAV_BA_ST_LIN_PT && stage != NAV_BA_ST_LIN_CAM && stage != NAV_BA_ST_COST &&
        stage != NAV_BA_ST_SCHUR_ROW)
        return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "debug call before nav_ba_prepare");
    *avg_us = 0.0;
    if (h->all_empty || (stage == NAV_BA_ST_SCHUR_ROW && h->P.n_fc == 0)) return NAV_BA_EMPTY;
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    std::vector<LMDev> L(P.n_win);
    for (auto& l : L) { l = LMDev{}; l.lambda = lambda; l.nu = 2.0; l.phase = kPhaseLin; l.iter = 1; }
    HIPTRY(h, hipMemcpyAsync(h->W.lm, L.data(), sizeof(LMDev) * L.size(), hipMemcpyHostToDevice, s));
    launch_cam_rt(P, h->cur, s);
    const bool m = mixed(h);
    if (m) enqueue_lin<float>(h, false); else enqueue_lin<double>(h, false);
    auto one = [&] {
        switch (stage) {
        case NAV_BA_ST_LIN_PT:
            if (m) launch_lin_pt<float>(P, h->cur, h->W, s); else launch_lin_pt<double>(P, h->cur, h->W, s);
            break;
        case NAV_BA_ST_LIN_CAM:
            if (m) launch_lin_cam<float>(P, h->cur, h->W, s); else launch_lin_cam<double>(P, h->cur, h->W, s);
            break;
        case NAV_BA_ST_COST:
            if (m) launch_cost<float>(P, h->cur, h->W, 1, s); else launch_cost<double>(P, h->cur, h->W, 1, s);
            break;
        default:
            if (m) launch_schur_row<float>(P, h->cur, h->W, s); else launch_schur_row<double>(P, h->cur, h->W, s);
        }
    };
    one();  // warm-up
    HIPTRY(h, hipEventRecord(h->ev0, s));
    for (int r = 0; r < reps; ++r) one();
    HIPTRY(h, hipEventRecord(h->ev1, s));
    HIPTRY(h, hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *avg_us = 1e3 * (double)ms / reps;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_check_layout(const nav_ba_problem* problems, int32_t n_windows, const nav_ba_options* o,
                                  int32_t batched, int32_t chunk, int32_t kernel, int64_t* n_pairs,
                                  int32_t* n_rounds, char* msg, int32_t msg_len) {
    auto say = [&](const std::string& m) {
        if (msg && msg_len > 0) {
            std::strncpy(msg, m.c_str(), (size_t)msg_len - 1);
            msg[msg_len - 1] = 0;
        }
    };
    say("");
    if (chunk < 0 || chunk > 4096 || chunk % kWave != 0) { say("chunk"); return NAV_BA_BAD_ARG; }
    if (kernel < 1 || kernel > 2) { say("kernel"); return NAV_BA_BAD_ARG; }
    std::string err;
    int st = check_args(problems, n_windows, o, batched != 0, err);
    if (st != NAV_BA_OK) { say(err); return st; }
    HostLayout L;
    if ((st = build_layout(problems, n_windows, o, batched != 0, L, err)) != NAV_BA_OK) { say(err); return st; }
    if (L.all_empty) { say("empty"); return NAV_BA_EMPTY; }
    SchurItems S;
    if (!build_schur_items(L.graph(), chunk > 0 ? chunk : kSchurChunk, kernel == 2, S)) {
        say("more than 2^31 Schur pair entries");
        return NAV_BA_TOO_LARGE;
    }
    long long np = 0;
    for (const int2& e : S.pairs) np += e.x >= 0 ? 1 : 0;
    if (n_pairs) *n_pairs = np;
    if (n_rounds) *n_rounds = (int32_t)S.rnd_rows.size();
    const std::string bad = check_layout(L, S);
    if (!bad.empty()) { say(bad); return -1; }
    return NAV_BA_OK;
}

int32_t nav_ba_debug_reduced_system(nav_ba_handle* h, double lambda, double* S, double* g,
                                    int32_t max_n, int32_t* n_out) {
    if (!h || !n_out) return NAV_BA_BAD_ARG;
    if (!h->prepared) return fail(h, NAV_BA_NOT_PREPARED, "debug call before nav_ba_prepare");
    if (h->batched) return fail(h, NAV_BA_BAD_ARG, "debug reduced system is for single problems");
    if (h->all_empty) { *n_out = 0; return NAV_BA_EMPTY; }
    HIPTRY(h, hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const DevProblem& P = h->P;
    *n_out = P.n;
    HIPTRY(h, hipMemcpyAsync(h->cur.pose, h->init.pose, sizeof(PoseQT) * P.n_cam, hipMemcpyDeviceToDevice, s));
    HIPTRY(h, hipMemcpyAsync(h->cur.X, h->init.X, sizeof(D4) * P.n_pt, hipMemcpyDeviceToDevice, s));
    LMDev L{};
    L.lambda = lambda;
    L.nu = 2.0;
    L.phase = kPhaseLin;
    L.iter = 1;
    HIPTRY(h, hipMemcpyAsync(h->W.lm, &L, sizeof(LMDev), hipMemcpyHostToDevice, s));
    launch_cam_rt(P, h->cur, s);
    if (mixed(h)) { enqueue_lin<float>(h, false); enqueue_reduced_system<float>(h); }
    else { enqueue_lin<double>(h, false); enqueue_reduced_system<double>(h); }
    if (P.use_win && P.n_fc > 0) {  // the step sums the system inside its factorisation: form S here
        if (mixed(h)) launch_schur_reduce<float>(P, h->W, s); else launch_schur_reduce<double>(P, h->W, s);
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

}  // extern "C"
