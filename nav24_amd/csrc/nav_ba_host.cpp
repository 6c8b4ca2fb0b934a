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
// Schur chunk tuner: problems with at least kTuneMinObs observations time each
// candidate chunk size once at prepare (best of kTuneReps launches) and keep a
// candidate only if it beats the current best by the margin.
constexpr int kTuneMinObs = 1 << 17;
constexpr int kTuneChunks[] = {128, 256, 512, 1024, 2048, 4096};
constexpr int kTuneReps = 3;
constexpr float kTuneMargin = 0.97f;

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

}  // namespace

struct nav_ba_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bool prepared = false;
    bool batched = false;
    nav_ba_options opt{};
    std::vector<WinHost> win;
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
    int solver_mode = 0;  // 0 auto, 1 tiles, 2 blocks
    int schur_chunk = 0;  // observations per Schur work item: 0 = default or tuned
    int schur_chunk_used = 0;
    int schur_run = kRunLen;  // Schur pairs per register run
    StageTimer timer;
    int* hdone = nullptr;      // pinned: windows finished
    char* stage = nullptr;     // pinned staging of a prepare's uploads (grown on demand)
    size_t stage_cap = 0;
    char* stage_s = nullptr;   // pinned staging of the Schur work items (the tuner re-uploads them)
    size_t stage_s_cap = 0;
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
        stage(h, NAV_BA_ST_SCHUR_REDUCE, [&] { launch_schur_reduce<T>(h->P, h->W, h->stream); });
    }
}

// One LM step of every window: gated linearisation, one trial, decision.
template <typename T>
void enqueue_step(nav_ba_handle* h, bool first) {
    enqueue_lin<T>(h, first);
    enqueue_reduced_system<T>(h);
    if (h->P.n_fc > 0)
        stage(h, NAV_BA_ST_CHOLESKY, [&] {
            if (h->P.batched) launch_cholesky_small(h->P, h->W, h->stream);
            else launch_cholesky_solve(h->P, h->cur, h->cand, h->W, h->stream);
        });
    stage(h, NAV_BA_ST_BACKSUB, [&] { launch_backsub<T>(h->P, h->cur, h->cand, h->W, h->stream); });
    if (!cholesky_updates_cameras(h->P))
        stage(h, NAV_BA_ST_UPDATE_CAM, [&] { launch_update_cam(h->P, h->cur, h->cand, h->W, h->stream); });
    stage(h, NAV_BA_ST_COST, [&] { launch_cost<T>(h->P, h->cand, h->W, 1, h->stream); });
    stage(h, NAV_BA_ST_ACCEPT, [&] {
        launch_lm_accept(h->P, h->W, h->stream);
        launch_commit(h->P, h->cur, h->cand, h->W, h->stream);
    });
}

int launch_step(nav_ba_handle* h, bool first) {
    if (!h->timing) {
        HIPTRY(h, hipGraphLaunch(first ? h->step0 : h->step, h->stream));
        return NAV_BA_OK;
    }
    h->timing_active = true;
    if (mixed(h)) enqueue_step<float>(h, first); else enqueue_step<double>(h, first);
    h->timing_active = false;
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

// Work items of one chunk size (camera-major observations per item) and their pair lists.
struct SchurItems {
    int chunk = 0, segs = 0, max_ns = 1;
    long long pofs = 0;
    std::vector<int> sf, sa0, sa1, sg, sns, row_sitem_ptr, row_nch, seg_batch, batch_ptr, sorder;
    std::vector<long long> spofs;
    std::vector<int2> pairs;
};

// Items: per S row, column groups of kMaxSlots, chunks of `chunk` observations,
// dispatched heaviest first (pair iterations + observations).  Pairs: per
// item, per (round, wave) segment, the pairs (a, b) of each lane's
// observation a with the later observations b of its point on free cameras of
// the item's column group (a itself excluded: the diagonal pair is summed in
// registers).  Entry: {b, lane | slot << 6 | local camera << 14}; {-1, 0} is
// padding.  A segment's pairs are grouped by slot into runs of at most
// `run` pairs (a slot's pairs split evenly), the runs sorted longest first
// (the runs of one slot spread apart) and packed 64 to a batch: pair j of the batch's run l at batch offset
// 64 j + l, R = the batch's longest run rows.  A lane then works through one
// run (one camera, one slot) and adds its 6x6 block to LDS once per batch.
// False when the pair count exceeds int32.
bool build_schur_items(const SchurGraph& G, int chunk, int run, SchurItems& S) {
    S = SchurItems{};
    S.chunk = chunk;
    S.row_sitem_ptr.push_back(0);
    for (int f = 0; f < G.nfc; ++f) {
        const int ncols = G.colptr[f + 1] - G.colptr[f];
        const int ngroups = (ncols + kMaxSlots - 1) / kMaxSlots;
        const int nobs = G.cam_ptr[f + 1] - G.cam_ptr[f];
        const int nch = std::max(1, (nobs + chunk - 1) / chunk);
        S.row_nch.push_back(nch);
        for (int gi = 0; gi < ngroups; ++gi) {
            const int ns = std::min(kMaxSlots, ncols - gi * kMaxSlots);
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
    const int rounds = (chunk + kBlock - 1) / kBlock;
    S.segs = rounds * (kBlock / kWave);
    S.seg_batch.assign(1, 0);
    S.batch_ptr.assign(1, 0);
    std::vector<long long> scost(S.sf.size(), 0);
    std::vector<int> slotmap(G.nc, -1);
    std::vector<std::vector<int2>> bucket;
    struct Run { int slot, start, len, u; };
    std::vector<Run> runs;
    for (size_t it = 0; it < S.sf.size(); ++it) {
        const int f = S.sf[it], gi = S.sg[it], ns = S.sns[it];
        const int col0 = G.colptr[f] + gi * kMaxSlots;
        const int w = G.cam_win[G.free_cam[f]];
        for (int sl = 0; sl < ns; ++sl) slotmap[G.free_cam[G.colcam[col0 + sl]]] = sl;
        if (bucket.size() < (size_t)ns) bucket.resize(ns);
        long long iters = 0;
        for (int r = 0; r < rounds; ++r)
            for (int wv = 0; wv < kBlock / kWave; ++wv) {
                for (int sl = 0; sl < ns; ++sl) bucket[sl].clear();
                const int kb = S.sa0[it] + r * kBlock + wv * kWave;
                for (int k = kb; k < std::min(kb + kWave, S.sa1[it]); ++k) {
                    const int a = G.cam_obs[k];
                    const int q = G.o_pt[a];
                    if (!G.pt_free[q]) continue;
                    for (int b = G.cam_obs_b0[k]; b < G.pt_ptr[q + 1]; ++b) {
                        if (b == a) continue;
                        const int sl = slotmap[G.o_cam[b]];
                        if (sl < 0) continue;
                        bucket[sl].push_back(int2{b, (k - kb) | (sl << 6) | ((G.o_cam[b] - G.cam_off[w]) << 14)});
                    }
                }
                runs.clear();
                for (int sl = 0; sl < ns; ++sl) {
                    const int c = (int)bucket[sl].size();
                    if (c == 0) continue;
                    const int nr = (c + run - 1) / run, base = c / nr, extra = c % nr;
                    for (int u = 0, st = 0; u < nr; ++u) {
                        const int len = base + (u < extra ? 1 : 0);
                        runs.push_back(Run{sl, st, len, u});
                        st += len;
                    }
                }
                // longest first; among equal lengths the slots' first runs, then
                // their second runs, ...: lanes next to each other hold
                // different slots, so their LDS adds hit different addresses
                std::stable_sort(runs.begin(), runs.end(), [](const Run& x, const Run& y) {
                    return x.len != y.len ? x.len > y.len : x.u < y.u;
                });
                for (size_t r0 = 0; r0 < runs.size(); r0 += kWave) {
                    const int R = runs[r0].len;
                    const size_t base = S.pairs.size();
                    if (base + (size_t)kWave * R > (size_t)0x7fffffff) return false;
                    S.pairs.resize(base + (size_t)kWave * R, int2{-1, 0});
                    for (size_t l = 0; l < (size_t)kWave && r0 + l < runs.size(); ++l) {
                        const Run& rn = runs[r0 + l];
                        for (int j = 0; j < rn.len; ++j) S.pairs[base + (size_t)kWave * j + l] = bucket[rn.slot][rn.start + j];
                    }
                    S.batch_ptr.push_back((int)S.pairs.size());
                    iters += R;
                }
                S.seg_batch.push_back((int)S.batch_ptr.size() - 1);
            }
        for (int sl = 0; sl < ns; ++sl) slotmap[G.free_cam[G.colcam[col0 + sl]]] = -1;
        scost[it] = (long long)kWave * iters + 4LL * (S.sa1[it] - S.sa0[it]);
    }
    S.sorder.resize(S.sf.size());
    for (size_t i = 0; i < S.sorder.size(); ++i) S.sorder[i] = (int)i;
    std::stable_sort(S.sorder.begin(), S.sorder.end(), [&](int x, int y) { return scost[x] > scost[y]; });
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
    UP(P.sitem_order, S.sorder);
    UP(P.seg_batch, S.seg_batch); UP(P.batch_ptr, S.batch_ptr); UP(P.pairs, S.pairs);
#undef UP
    if ((st = pk.flush(h)) != NAV_BA_OK) return st;
    if ((st = alloc(h, (void**)&h->W.spart, sizeof(double) * std::max<long long>(S.pofs, 1))) != NAV_BA_OK) return st;
    P.schur_segs = S.segs;
    P.n_sitems = (int)S.sf.size();
    P.max_ns = S.max_ns;
    h->schur_chunk_used = S.chunk;
    return NAV_BA_OK;
}

// Process-wide memory of tuned chunk sizes, per device, precision and problem shape.
struct TuneKey {
    int device, mixed, run, n_win, nc, np, no, nfc;
    bool operator<(const TuneKey& o) const {
        return std::tie(device, mixed, run, n_win, nc, np, no, nfc) <
               std::tie(o.device, o.mixed, o.run, o.n_win, o.nc, o.np, o.no, o.nfc);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;

bool cached_chunk(const TuneKey& k, int* chunk) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(k);
    if (it == g_tuned.end()) return false;
    *chunk = it->second;
    return true;
}

void remember_chunk(const TuneKey& k, int chunk) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[k] = chunk;
}

// Device time of the Schur stage (rows + reduction) of the uploaded items, best of kTuneReps.
int time_schur(nav_ba_handle* h, float* best_ms) {
    *best_ms = 1e30f;
    for (int r = 0; r <= kTuneReps; ++r) {
        HIPTRY(h, hipEventRecord(h->ev0, h->stream));
        if (mixed(h)) {
            launch_schur_row<float>(h->P, h->cur, h->W, h->stream);
            launch_schur_reduce<float>(h->P, h->W, h->stream);
        } else {
            launch_schur_row<double>(h->P, h->cur, h->W, h->stream);
            launch_schur_reduce<double>(h->P, h->W, h->stream);
        }
        HIPTRY(h, hipEventRecord(h->ev1, h->stream));
        HIPTRY(h, hipEventSynchronize(h->ev1));
        float ms = 0.f;
        HIPTRY(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
        if (r > 0) *best_ms = std::min(*best_ms, ms);  // the first launch is a warm-up
    }
    return NAV_BA_OK;
}

// Pick the chunk size of the Schur items by timing each candidate on the
// problem itself, linearised at its initial state.  `S` holds the items of
// the default size, uploaded from allocation slot `mark` on; it returns
// holding (and with the device holding) the fastest candidate's items.
int tune_schur_chunk(nav_ba_handle* h, const SchurGraph& G, int mark, SchurItems& S) {
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
    SchurItems C;
    for (int chunk : kTuneChunks) {
        if (chunk == def) continue;
        if (!build_schur_items(G, chunk, h->schur_run, C)) continue;
        h->nbuf = mark;
        if ((st = upload_schur_items(h, C)) != NAV_BA_OK) return st;
        float ms;
        if ((st = time_schur(h, &ms)) != NAV_BA_OK) return st;
        if (ms < kTuneMargin * best_ms) { best_ms = ms; std::swap(S, C); }
    }
    h->nbuf = mark;
    if ((st = upload_schur_items(h, S)) != NAV_BA_OK) return st;
    HIPTRY(h, hipStreamSynchronize(s));
    return NAV_BA_OK;
}

// Build the device structure of n_win windows laid end to end (see ba_kernels.h).
int prepare_windows(nav_ba_handle* h, const nav_ba_problem* probs, int n_win, const nav_ba_options* o, bool batched) {
    if (!h) return NAV_BA_BAD_ARG;
    h->err.clear();
    h->prepared = false;
    h->ran = false;
    h->schur_chunk_used = 0;
    if (!o || !probs || n_win <= 0) return fail(h, NAV_BA_BAD_ARG, "null options or problems");
    for (int w = 0; w < n_win; ++w)
        if (validate(&probs[w]) != NAV_BA_OK)
            return fail(h, NAV_BA_BAD_ARG, "invalid problem " + std::to_string(w) + " (null array or index out of range)");
    if (o->precision != NAV_BA_FP64 && o->precision != NAV_BA_MIXED) return fail(h, NAV_BA_BAD_ARG, "unknown precision");
    if (o->n_iter < 0) return fail(h, NAV_BA_BAD_ARG, "negative iteration count");
    for (int w = 1; w < n_win; ++w)
        if (probs[w].fx != probs[0].fx || probs[w].fy != probs[0].fy || probs[w].cx != probs[0].cx || probs[w].cy != probs[0].cy)
            return fail(h, NAV_BA_BAD_ARG, "batched windows must share the camera intrinsics");
    HIPTRY(h, hipSetDevice(h->device));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    if (h->step) { (void)hipGraphExecDestroy(h->step); h->step = nullptr; }
    if (h->step0) { (void)hipGraphExecDestroy(h->step0); h->step0 = nullptr; }
    h->opt = *o;
    if (h->opt.max_trials <= 0) h->opt.max_trials = 10;
    h->batched = batched;
    const bool compat = o->compat_fp32_io != 0;

    // ---- concatenate the windows; keep only points with observations
    h->win.assign(n_win, WinHost());
    std::vector<int> cam_off(n_win + 1, 0), obs_off(n_win + 1, 0);
    for (int w = 0; w < n_win; ++w) {
        cam_off[w + 1] = cam_off[w] + probs[w].n_cam;
        obs_off[w + 1] = obs_off[w] + probs[w].n_obs;
    }
    const int nc = cam_off[n_win];
    const int no = obs_off[n_win];
    std::vector<int> g_cam(no), g_pt(no), pt_first(n_win + 1, 0);
    std::vector<int> pt_win_all;  // local point -> window
    std::vector<int> cam_win(nc);
    int np = 0;
    bool any = false;
    for (int w = 0; w < n_win; ++w) {
        const nav_ba_problem& p = probs[w];
        WinHost& wh = h->win[w];
        wh.n_cam = p.n_cam; wh.n_pt = p.n_pt; wh.n_obs = p.n_obs;
        wh.cam0 = cam_off[w]; wh.obs0 = obs_off[w];
        wh.T_in.assign(p.T_cw, p.T_cw + 12 * (size_t)p.n_cam);
        wh.X_in.assign(p.X, p.X + 3 * (size_t)p.n_pt);
        for (int c = 0; c < p.n_cam; ++c) cam_win[cam_off[w] + c] = w;
        std::vector<int> cnt(p.n_pt, 0), loc(p.n_pt, -1);
        for (int i = 0; i < p.n_obs; ++i) cnt[p.obs_pt[i]]++;
        wh.pt_local0 = np;
        for (int i = 0; i < p.n_pt; ++i)
            if (cnt[i] > 0) { loc[i] = np++; wh.pt_orig.push_back(i); pt_win_all.push_back(w); }
        wh.pt_local1 = np;
        wh.empty = (p.n_cam == 0 || wh.pt_orig.empty());
        any = any || !wh.empty;
        for (int i = 0; i < p.n_obs; ++i) {
            g_cam[obs_off[w] + i] = cam_off[w] + p.obs_cam[i];
            g_pt[obs_off[w] + i] = loc[p.obs_pt[i]];
        }
        pt_first[w + 1] = np;
    }
    h->all_empty = !any;
    if (!batched && n_win != 1) return fail(h, NAV_BA_BAD_ARG, "single-problem prepare takes one problem");
    if (h->all_empty) {
        h->prepared = true;
        return NAV_BA_OK;
    }

    // ---- observations ordered by (point, camera, input order): two stable counting sorts
    std::vector<int> by_cam(no), ord(no);
    {
        std::vector<int> off(nc + 1, 0);
        for (int i = 0; i < no; ++i) off[g_cam[i] + 1]++;
        for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
        for (int i = 0; i < no; ++i) by_cam[off[g_cam[i]]++] = i;
        std::vector<int> poff(np + 1, 0);
        for (int i = 0; i < no; ++i) poff[g_pt[i] + 1]++;
        for (int q = 0; q < np; ++q) poff[q + 1] += poff[q];
        for (int k = 0; k < no; ++k) {
            const int i = by_cam[k];
            ord[poff[g_pt[i]]++] = i;
        }
    }
    std::vector<int> o_cam(no), o_pt(no), o_orig(no), pt_ptr(np + 1, 0);
    std::vector<double> o_uv(2 * (size_t)no);
    for (int k = 0; k < no; ++k) {
        const int i = ord[k];
        const int w = (int)(std::upper_bound(obs_off.begin(), obs_off.end(), i) - obs_off.begin()) - 1;
        const double* uv = probs[w].obs_uv + 2 * (size_t)(i - obs_off[w]);
        o_cam[k] = g_cam[i];
        o_pt[k] = g_pt[i];
        o_orig[k] = i;
        o_uv[2 * (size_t)k] = compat ? rf32(uv[0]) : uv[0];
        o_uv[2 * (size_t)k + 1] = compat ? rf32(uv[1]) : uv[1];
        pt_ptr[o_pt[k] + 1]++;
    }
    for (int q = 0; q < np; ++q) pt_ptr[q + 1] += pt_ptr[q];
    std::vector<uint8_t> pt_free(np);
    for (int q = 0; q < np; ++q) {
        const int w = pt_win_all[q];
        const int i = h->win[w].pt_orig[q - h->win[w].pt_local0];
        pt_free[q] = (probs[w].pt_fixed && probs[w].pt_fixed[i]) ? 0 : 1;
    }

    // ---- free cameras: not fixed and with at least one (therefore active) edge
    std::vector<int> cnt_cam(nc, 0);
    for (int i = 0; i < no; ++i) cnt_cam[g_cam[i]]++;
    std::vector<int> cam_free(nc, -1), free_cam, win_fc0(n_win + 1, 0), win_cam0(n_win + 1, 0);
    for (int w = 0; w < n_win; ++w) {
        win_cam0[w] = cam_off[w];
        win_fc0[w] = (int)free_cam.size();
        for (int c = cam_off[w]; c < cam_off[w + 1]; ++c) {
            const int lc = c - cam_off[w];
            if (!(probs[w].cam_fixed && probs[w].cam_fixed[lc]) && cnt_cam[c] > 0) {
                cam_free[c] = (int)free_cam.size();
                free_cam.push_back(c);
            }
        }
    }
    win_cam0[n_win] = nc;
    win_fc0[n_win] = (int)free_cam.size();
    const int nfc = (int)free_cam.size();
    int max_win_cams = 1;
    for (int w = 0; w < n_win; ++w) max_win_cams = std::max(max_win_cams, probs[w].n_cam);
    if (max_win_cams > kMaxWinCams) return fail(h, NAV_BA_TOO_LARGE, "more than 8192 cameras in a window");

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
    std::vector<int> c_pt(cam_obs.size());
    std::vector<double> c_uv(2 * cam_obs.size());
    for (size_t k = 0; k < cam_obs.size(); ++k) {
        c_pt[k] = o_pt[cam_obs[k]];
        c_uv[2 * k] = o_uv[2 * (size_t)cam_obs[k]];
        c_uv[2 * k + 1] = o_uv[2 * (size_t)cam_obs[k] + 1];
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

    // ---- point-aligned linearisation tiles, never spanning two windows
    std::vector<int> tile_obs{0}, tile_pt{0}, tile_win, win_tile0{0};
    for (int w = 0; w < n_win; ++w) {
        for (int q = pt_first[w]; q < pt_first[w + 1];) {
            int q1 = q;
            while (q1 < pt_first[w + 1] && pt_ptr[q1 + 1] - pt_ptr[q] <= kTile) ++q1;
            if (q1 == q) q1 = q + 1;  // one point with more than kTile observations
            tile_obs.push_back(pt_ptr[q1]);
            tile_pt.push_back(q1);
            tile_win.push_back(w);
            q = q1;
        }
        win_tile0.push_back((int)tile_win.size());
    }
    // ---- window-aligned cost blocks (observations) and point blocks; windows
    // are contiguous, so each block starts where the previous one ends
    std::vector<int> cblk_o0{0}, cblk_win, win_cblk0{0};
    std::vector<int> win_pt0(n_win + 1);
    for (int w = 0; w < n_win; ++w) {
        const int o0 = pt_ptr[pt_first[w]], o1 = pt_ptr[pt_first[w + 1]];
        for (int a = o0; a < o1; a += kCostObs) { cblk_o0.push_back(std::min(a + kCostObs, o1)); cblk_win.push_back(w); }
        win_cblk0.push_back((int)cblk_win.size());
        win_pt0[w] = pt_first[w];
    }
    win_pt0[n_win] = np;

    // ---- Hpp work items: (free camera, chunk)
    std::vector<int> hf, ha0, ha1, cam_hitem_ptr{0};
    for (int f = 0; f < nfc; ++f) {
        for (int a = cam_ptr[f]; a < cam_ptr[f + 1]; a += kChunk) {
            hf.push_back(f); ha0.push_back(a); ha1.push_back(std::min(a + kChunk, cam_ptr[f + 1]));
        }
        cam_hitem_ptr.push_back((int)hf.size());
    }

    // ---- co-visibility through free points, per window: column blocks g >= f of each S row
    // ... and the other blocks g > f of each row (zero in S, rewritten every trial
    // because the factorisation works in place)
    std::vector<int> colptr{0}, colcam, zcolptr{0}, zcolcam;
    {
        std::vector<uint64_t> bits;
        std::vector<int> fl;
        for (int w = 0; w < n_win; ++w) {
            const int f0 = win_fc0[w], nw = win_fc0[w + 1] - f0;
            bits.assign(((size_t)nw * nw + 63) / 64, 0);
            for (int q = pt_first[w]; q < pt_first[w + 1]; ++q) {
                if (!pt_free[q]) continue;
                fl.clear();
                for (int k = pt_ptr[q]; k < pt_ptr[q + 1]; ++k) {
                    const int f = cam_free[o_cam[k]];
                    if (f >= 0 && (fl.empty() || fl.back() != f - f0)) fl.push_back(f - f0);
                }
                for (size_t a = 0; a < fl.size(); ++a)
                    for (size_t b = a + 1; b < fl.size(); ++b) {
                        const size_t bit = (size_t)fl[a] * nw + fl[b];
                        bits[bit >> 6] |= 1ull << (bit & 63);
                    }
            }
            for (int lf = 0; lf < nw; ++lf) {
                colcam.push_back(f0 + lf);
                for (int lg = lf + 1; lg < nw; ++lg) {
                    const size_t bit = (size_t)lf * nw + lg;
                    if (bits[bit >> 6] >> (bit & 63) & 1ull) colcam.push_back(f0 + lg);
                    else zcolcam.push_back(f0 + lg);
                }
                colptr.push_back((int)colcam.size());
                zcolptr.push_back((int)zcolcam.size());
            }
        }
    }

    // ---- Schur work items and pair lists at the default chunk size (tuned below)
    const SchurGraph G{nfc, nc, colptr.data(), colcam.data(), cam_ptr.data(), cam_obs.data(), cam_obs_b0.data(),
                       o_pt.data(), o_cam.data(), pt_ptr.data(), pt_free.data(), free_cam.data(), cam_win.data(),
                       cam_off.data()};
    const TuneKey key{h->device, mixed(h) ? 1 : 0, h->schur_run, n_win, nc, np, no, nfc};
    int chunk0 = h->schur_chunk > 0 ? h->schur_chunk : kSchurChunk;
    bool tune = h->schur_chunk == 0 && nfc > 0 && no >= kTuneMinObs;
    if (tune && cached_chunk(key, &chunk0)) tune = false;  // this shape was tuned before on this device
    SchurItems SI;
    if (!build_schur_items(G, chunk0, h->schur_run, SI)) return fail(h, NAV_BA_TOO_LARGE, "more than 2^31 Schur pairs");

    // ---- dense reduced systems: one padded tiled system, or one small system per window
    std::vector<long long> win_S(n_win, 0);
    std::vector<int> win_lda(n_win, 1);
    long long S_size = 1;
    const int n = 6 * nfc;
    int n_pad = ((n + kNB - 1) / kNB) * kNB;
    if (!batched) {
        if (n > kMaxDenseN) return fail(h, NAV_BA_TOO_LARGE, "reduced camera system larger than 16384");
        win_lda[0] = std::max(n_pad, 1);
        S_size = std::max<long long>((long long)n_pad * n_pad, 1);
    } else {
        S_size = 0;
        for (int w = 0; w < n_win; ++w) {
            const int nw = 6 * (win_fc0[w + 1] - win_fc0[w]);
            if (nw > kSmallMaxN) return fail(h, NAV_BA_TOO_LARGE, "a batched window has more than 30 free cameras");
            win_S[w] = S_size;
            win_lda[w] = std::max(nw, 1);
            S_size += (long long)win_lda[w] * win_lda[w];
        }
        n_pad = 0;
    }

    // ---- initial state
    std::vector<PoseQT> pose0(nc);
    for (int w = 0; w < n_win; ++w)
        for (int c = 0; c < probs[w].n_cam; ++c) {
            const double* T = probs[w].T_cw + 12 * (size_t)c;
            PoseQT& P0 = pose0[cam_off[w] + c];
            host_quat_from_R(T, P0.q);
            P0.t[0] = T[9]; P0.t[1] = T[10]; P0.t[2] = T[11];
            P0.pad = 0;
        }
    std::vector<D4> X0(np);
    for (int q = 0; q < np; ++q) {
        const int w = pt_win_all[q];
        const double* x = probs[w].X + 3 * (size_t)h->win[w].pt_orig[q - h->win[w].pt_local0];
        X0[q] = compat ? D4{rf32(x[0]), rf32(x[1]), rf32(x[2]), 0.0} : D4{x[0], x[1], x[2], 0.0};
    }

    // ---- device upload
    h->nbuf = 0;
    DevProblem& P = h->P;
    P = DevProblem{};
    P.n_cam = nc; P.n_pt = np; P.n_obs = no; P.n_fc = nfc;
    P.robust = o->robust ? 1 : 0;
    P.delta = o->huber_delta;
    P.fx = compat ? rf32(probs[0].fx) : probs[0].fx; P.fy = compat ? rf32(probs[0].fy) : probs[0].fy;
    P.cx = compat ? rf32(probs[0].cx) : probs[0].cx; P.cy = compat ? rf32(probs[0].cy) : probs[0].cy;
    P.n_iter = o->n_iter;
    P.max_trials = h->opt.max_trials;
    P.auto_lambda = !(o->lambda_init > 0);
    P.lambda0 = o->lambda_init;
    int st;
    Packer pk(&h->stage, &h->stage_cap);  // free again: every prepare ends with a stream synchronisation
#define UP(dst, vec) pk.add(&(dst), (vec))
    UP(P.o_cam, o_cam); UP(P.o_pt, o_pt); UP(P.o_uv, o_uv); UP(P.o_orig, o_orig);
    UP(P.pt_ptr, pt_ptr); UP(P.pt_free, pt_free); UP(P.cam_free, cam_free); UP(P.free_cam, free_cam);
    UP(P.cam_ptr, cam_ptr); UP(P.cam_obs, cam_obs); UP(P.cam_obs_b0, cam_obs_b0);
    UP(P.c_pt, c_pt); UP(P.c_uv, c_uv);
    UP(P.tile_obs, tile_obs); UP(P.tile_pt, tile_pt); UP(P.tile_win, tile_win);
    UP(P.hitem_f, hf); UP(P.hitem_a0, ha0); UP(P.hitem_a1, ha1); UP(P.cam_hitem_ptr, cam_hitem_ptr);
    UP(P.colptr, colptr); UP(P.colcam, colcam); UP(P.zcolptr, zcolptr); UP(P.zcolcam, zcolcam);
    UP(P.cam_win, cam_win); UP(P.pt_win, pt_win_all); UP(P.win_cam0, win_cam0); UP(P.win_fc0, win_fc0);
    UP(P.win_pt0, win_pt0); UP(P.win_S, win_S); UP(P.win_lda, win_lda);
    UP(P.cblk_o0, cblk_o0); UP(P.cblk_win, cblk_win); UP(P.win_cblk0, win_cblk0);
    UP(P.win_tile0, win_tile0);
#undef UP
    if ((st = pk.flush(h)) != NAV_BA_OK) return st;
    P.n_tiles = (int)tile_win.size();
    P.n_hitems = (int)hf.size();
    P.n_win = n_win;
    P.max_win_cams = max_win_cams;
    P.max_win_fc = 0;
    for (int w = 0; w < n_win; ++w) P.max_win_fc = std::max(P.max_win_fc, win_fc0[w + 1] - win_fc0[w]);
    P.batched = batched ? 1 : 0;
    P.n_cblk = (int)cblk_win.size();
    P.n = n;
    P.n_pad = n_pad;
    P.nt = n_pad / kNB;
    P.use_blocks = (!batched && nfc > 0 && nfc <= kBlocksMaxFc &&
                    (h->solver_mode == 2 || (h->solver_mode == 0 && nfc <= kBlocksAutoFc))) ? 1 : 0;

    const size_t tsz = mixed(h) ? sizeof(float) : sizeof(double);
    const size_t nvec = batched ? std::max<size_t>(6 * (size_t)nfc, 1) : std::max<size_t>(n_pad, 1);
    auto A = [&](void** dst, size_t bytes) { return alloc(h, dst, bytes); };
#define AL(dst, bytes) if ((st = A((void**)&(dst), (bytes))) != NAV_BA_OK) return st
    AL(h->init.pose, sizeof(PoseQT) * nc); AL(h->init.X, sizeof(D4) * np); h->init.cam = nullptr;
    AL(h->cur.pose, sizeof(PoseQT) * nc); AL(h->cur.cam, sizeof(CamRT) * nc); AL(h->cur.X, sizeof(D4) * np);
    AL(h->cand.pose, sizeof(PoseQT) * nc); AL(h->cand.cam, sizeof(CamRT) * nc); AL(h->cand.X, sizeof(D4) * np);
    DevWork& W = h->W;
    AL(W.w, tsz * no);
    AL(W.Hl, sizeof(double) * 9 * (size_t)np);
    AL(W.hpart, sizeof(double) * 27 * std::max<size_t>(hf.size(), 1));
    AL(W.Hc, sizeof(double) * 28 * std::max(nfc, 1));
    AL(W.S, sizeof(double) * (size_t)std::max<long long>(S_size, 1));
    AL(W.g, sizeof(double) * nvec);
    AL(W.y, sizeof(double) * nvec);
    AL(W.dp, sizeof(double) * nvec);
    AL(W.Lm, sizeof(double) * (batched ? (size_t)1 : (size_t)std::max<long long>(S_size, 1)));
    AL(W.Li, sizeof(double) * 4096 * (batched ? (size_t)1 : (size_t)std::max(P.nt, 1)));
    AL(W.fpart, sizeof(double) * 64 * (batched ? (size_t)1 : (size_t)std::max(P.nt, 1) * std::max(P.nt, 1)));
    AL(W.red_cost, sizeof(double) * std::max(P.n_cblk, 1));
    AL(W.red_scale, sizeof(double) * (n_win + std::max(P.n_tiles, 1)));
    AL(W.red_aux, sizeof(double) * std::max(P.n_cblk, 1));
    AL(W.outlier, std::max<size_t>(no, 1));
    AL(W.hist, sizeof(double) * (size_t)n_win * (o->n_iter + 1));
    AL(W.final, sizeof(double) * 2 * (size_t)n_win);
    AL(W.done_count, sizeof(int));
    AL(W.lm, sizeof(LMDev) * n_win);
#undef AL
    const int schur_mark = h->nbuf;
    if ((st = upload_schur_items(h, SI)) != NAV_BA_OK) return st;
    HIPTRY(h, hipMemcpyAsync(h->init.pose, pose0.data(), sizeof(PoseQT) * nc, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipMemcpyAsync(h->init.X, X0.data(), sizeof(D4) * np, hipMemcpyHostToDevice, h->stream));
    HIPTRY(h, hipStreamSynchronize(h->stream));
    // Schur rows and the per-window factorisation may use more than 64 KiB of LDS
    if ((mixed(h) ? prepare_schur_row<float>(P) : prepare_schur_row<double>(P)) != 0)
        return fail(h, NAV_BA_TOO_LARGE, "Schur row workspace exceeds the LDS of one workgroup");
    if (prepare_cholesky(P) != 0) return fail(h, NAV_BA_HIP_ERROR, "LDS attribute of the factorisation kernels");
    if (batched && prepare_cholesky_small(P) != 0) return fail(h, NAV_BA_HIP_ERROR, "LDS attribute of the window factorisation");
    launch_fill_system(P, h->W, h->stream);
    if (tune) {
        if ((st = tune_schur_chunk(h, G, schur_mark, SI)) != NAV_BA_OK) return st;
        remember_chunk(key, SI.chunk);
    }
    if ((st = capture_step(h)) != NAV_BA_OK) return st;
    h->prepared = true;
    return NAV_BA_OK;
}

int fetch_window(nav_ba_handle* h, int w, const std::vector<PoseQT>& pose, const std::vector<D4>& X,
                 const std::vector<uint8_t>& flags, nav_ba_result* r) {
    const WinHost& wh = h->win[w];
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
    if (n != (int)h->win.size()) return fail(h, NAV_BA_BAD_ARG, "result count differs from the prepared window count");
    std::vector<PoseQT> pose;
    std::vector<D4> X;
    std::vector<uint8_t> flags;
    if (!h->all_empty) {
        HIPTRY(h, hipSetDevice(h->device));
        pose.resize(h->P.n_cam);
        X.resize(h->P.n_pt);
        flags.resize(h->P.n_obs);
        HIPTRY(h, hipMemcpyAsync(pose.data(), h->cur.pose, sizeof(PoseQT) * pose.size(), hipMemcpyDeviceToHost, h->stream));
        HIPTRY(h, hipMemcpyAsync(X.data(), h->cur.X, sizeof(D4) * X.size(), hipMemcpyDeviceToHost, h->stream));
        if (!flags.empty())
            HIPTRY(h, hipMemcpyAsync(flags.data(), h->W.outlier, flags.size(), hipMemcpyDeviceToHost, h->stream));
        HIPTRY(h, hipStreamSynchronize(h->stream));
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
        hipHostMalloc((void**)&h->hdone, sizeof(int), hipHostMallocDefault) != hipSuccess) {
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
    // launched in bursts (n_iter first: one per iteration when every trial is
    // accepted, then 2), the host checks the count of finished windows after
    // each burst; steps past a window's end return at once from every kernel.
    const long long max_steps = (long long)P.n_iter * P.max_trials;
    long long steps = 0;
    int st;
    long long burst = std::max(1, P.n_iter);
    while (steps < max_steps) {
        for (long long i = 0; i < burst && steps < max_steps; ++i, ++steps)
            if ((st = launch_step(h, steps == 0)) != NAV_BA_OK) return st;
        HIPTRY(h, hipMemcpyAsync(h->hdone, h->W.done_count, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPTRY(h, hipStreamSynchronize(s));
        if (*h->hdone >= P.n_win) break;
        burst = 2;
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
    if (!h || mode < 0 || mode > 2) return NAV_BA_BAD_ARG;
    h->solver_mode = mode;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_set_schur_chunk(nav_ba_handle* h, int32_t chunk) {
    if (!h || chunk < 0 || chunk > 4096 || chunk % kWave != 0) return NAV_BA_BAD_ARG;
    h->schur_chunk = chunk;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_set_schur_run(nav_ba_handle* h, int32_t run) {
    if (!h || run < 0 || run > 64) return NAV_BA_BAD_ARG;
    h->schur_run = run == 0 ? kRunLen : run;
    return NAV_BA_OK;
}

int32_t nav_ba_debug_get_schur_chunk(const nav_ba_handle* h, int32_t* chunk) {
    if (!h || !chunk) return NAV_BA_BAD_ARG;
    *chunk = h->prepared ? h->schur_chunk_used : 0;
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
