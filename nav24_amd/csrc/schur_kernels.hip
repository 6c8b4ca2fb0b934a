// schur_kernels.hip — reduced camera system rows (Schur complement onto the poses).
//
// g2o BlockSolver<6,3>::solve (the solver nav24 builds at
// core/backEnd/BE_GraphOptim.cpp:43-47) forms, per landmark p with blocks
// Hpl_ip (6x3) and D_p = Hll_p + lambda I:
//     S_ij -= Hpl_ip D_p^-1 Hpl_jp^T  (j >= i),   coef_i += Hpl_ip D_p^-1 bl_p.
// One workgroup owns one work item: block row f of S, restricted to one group
// of column cameras (the slots) and one chunk of camera f's observations,
// taken in rounds:
//   phase A  one lane per observation a of camera f on a free point p:
//            W_a = Hpl_a Dinv_p (FP64) and the point go to the lane's row of
//            an LDS buffer; the diagonal pair (a, a) adds W_a Hpl_a^T
//            (symmetric, 21 values) and Hpl_a Dinv_p bl_p (6) to the lane's
//            registers;
//   phase B  every lane is bound to one slot (column camera) for the round:
//            the host split each slot's pairs (a, b) — b a later observation
//            of a's point on that camera — evenly over the lanes it gave the
//            slot.  A lane walks its pairs (8-byte entries {b, a's row in the
//            buffer}, read coalesced, one row per step; its camera, the first
//            two entries and their weights are loaded before phase A, so their
//            latency hides behind it), forms W_a Hpl_b^T = (W_a (w_b J_X,b)^T)
//            J_xi,b from recomputed Jacobians and sums the 6x6 blocks in
//            registers;
//   flush    a slot's lanes come in aligned groups of 2 or 4 (per round);
//            each group adds its sums by DPP swaps, one lane per group stages
//            them in LDS, and each slot's groups are added in a fixed order to
//            the slot's accumulator (one writer per entry).
// A round is 256 observations for the whole workgroup: its 256 lanes share
// the slots (balance over many lanes), with workgroup barriers between the
// phases.  (Rounds of 64 observations per wave, each wave with its own
// accumulators and no barrier, measured 6 % slower on a batch of small
// windows, whose rows have under 30 slots.)
// Every sum has a fixed order: no atomics, so the rows are the same bits on
// every run.  Work items are dispatched heaviest first (a host-side permutation).
// T is the type of the Jacobian entries only.  Cameras, points, W, the pair
// products and the S-row slots are FP64 for both compute types: the point
// elimination cancels Hpp against Hpl Dinv Hpl^T, and with a few-view point's
// Dinv rounded to FP32 the final cost of an uneven window moved by 1e-3
// relative in a CPU emulation, against 1e-6 for FP32 Jacobians alone.
#include "ba_kernels.h"

namespace navba {

constexpr int kWbuf = 22;        // per observation row: W (18), point (3), padding (16-byte rows)
constexpr int kStageStride = 37; // odd stride of the staged rows (one per lane group, 36 entries)

// LDS of one work item.
struct SchurLayout {
    size_t acc, wbuf, lslot, lrange, red, total;
    __host__ __device__ SchurLayout(int ns) {
        size_t o = 0;
        acc = o;    o += sizeof(double) * 36 * (size_t)ns;            // [slot][36]
        wbuf = o;   o += sizeof(double) * kWbuf * kBlock;             // also the flush staging
        lslot = o;  o += sizeof(int) * kBlock;                        // slot of each lane this round
        lrange = o; o += sizeof(int) * 2 * (size_t)ns;                // first lane, lane count per slot
        red = o;    o += sizeof(double) * (kBlock / kWave) * 27;      // diagonal block and coefficient per wave
        total = (o + 15) & ~(size_t)15;
    }
};
static_assert(kStageStride * (kBlock / 2) <= kWbuf * kBlock, "flush staging fits in the observation buffer");

// Inputs of phase A for one observation: its point's state and Hll / bl and
// its measurement, loaded ahead of use (every load needs only the point index).
struct PaIn {
    double x, y, z, u, v, h[9];
    int live;  // observation in range and on a free point
    __device__ __forceinline__ const double* hl() const { return h; }
};
// The same with Hll / bl read where they are used (fewer live registers).
struct PaLazy {
    double x, y, z, u, v;
    const double* h;
    int live;
    __device__ __forceinline__ const double* hl() const { return h; }
};

__device__ __forceinline__ void pa_load_lazy(const DevProblem& P, const D4* __restrict__ X,
                                             const double* __restrict__ Hl, int k, int p, PaLazy& in) {
    in.live = P.pt_free[p];
    const D4 xp = X[p];
    in.x = xp.x; in.y = xp.y; in.z = xp.z;
    in.u = P.c_uv[2 * (size_t)k];
    in.v = P.c_uv[2 * (size_t)k + 1];
    in.h = Hl + 9 * (size_t)p;
}

__device__ __forceinline__ void pa_load(const DevProblem& P, const D4* __restrict__ X,
                                        const double* __restrict__ Hl, int k, int p, PaIn& in) {
    in.live = 0;
    if (p < 0) return;
    in.live = P.pt_free[p];
    const D4 xp = X[p];
    in.x = xp.x; in.y = xp.y; in.z = xp.z;
    in.u = P.c_uv[2 * (size_t)k];
    in.v = P.c_uv[2 * (size_t)k + 1];
#pragma unroll
    for (int i = 0; i < 9; ++i) in.h[i] = Hl[9 * (size_t)p + i];
}

// Phase A for one observation of camera (Rf, tf): W and the point into row
// wb, diagonal pair and coefficient into the registers.
template <typename T, int CM, typename In>
__device__ __forceinline__ void schur_phase_a(const DevProblem& P, const double Rf[9], const double tf[3],
                                              const In& in, double lambda, bool g0, double* wb, double coef[6],
                                              double dsum[21]) {
    if (!in.live) return;
    EdgeLin<T> La;
    const double sa = edge_linearize_m<CM>(P, Rf, tf, in.x, in.y, in.z, in.u, in.v, La);
    const double wa = (double)huber_weight<T>(sa, P);  // as k_lin_pt, without a gather of w
    double JXa[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) JXa[j] = (double)La.JX[j];
    double Ha[18];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double A0 = (double)La.Jxi[i] * wa, A1 = (double)La.Jxi[6 + i] * wa;
#pragma unroll
        for (int j = 0; j < 3; ++j)  // a pinhole J_xi is zero at (0, 4) and (1, 3)
            Ha[3 * i + j] = (CamTraits<CM>::xi_zeros && i == 4) ? A1 * JXa[3 + j]
                          : (CamTraits<CM>::xi_zeros && i == 3) ? A0 * JXa[j] : A0 * JXa[j] + A1 * JXa[3 + j];
    }
    double Dp[6];
    const double* hp = in.hl();
    point_dinv(hp, lambda, Dp);
    const double d00 = Dp[0], d01 = Dp[1], d02 = Dp[2], d11 = Dp[3], d12 = Dp[4], d22 = Dp[5];
    double W[18];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double h0 = Ha[3 * i], h1 = Ha[3 * i + 1], h2 = Ha[3 * i + 2];
        W[3 * i + 0] = h0 * d00 + h1 * d01 + h2 * d02;
        W[3 * i + 1] = h0 * d01 + h1 * d11 + h2 * d12;
        W[3 * i + 2] = h0 * d02 + h1 * d12 + h2 * d22;
    }
#pragma unroll
    for (int e = 0; e < 18; ++e) wb[e] = W[e];
    wb[18] = in.x; wb[19] = in.y; wb[20] = in.z;
    if (g0) {
        // Dinv bl of the point (the right-hand side's point term)
        const double b0 = hp[6], b1 = hp[7], b2 = hp[8];
        const double q0 = d00 * b0 + d01 * b1 + d02 * b2;
        const double q1 = d01 * b0 + d11 * b1 + d12 * b2;
        const double q2 = d02 * b0 + d12 * b1 + d22 * b2;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            coef[i] = fma(Ha[3 * i + 2], q2, fma(Ha[3 * i + 1], q1, fma(Ha[3 * i], q0, coef[i])));
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int c = i; c < 6; ++c, ++q)
                dsum[q] = fma(W[3 * i + 2], Ha[3 * c + 2], fma(W[3 * i + 1], Ha[3 * c + 1], fma(W[3 * i], Ha[3 * c], dsum[q])));
    }
}

// sacc += [M0 M1] [B0; B1], skipping the pinhole zeros of J_xi at (0, 4) and (1, 3).
template <int CM>
__device__ __forceinline__ void pair_add(double sacc[36], const double M0[6], const double M1[6], const double B0[6],
                                         const double B1[6]) {
    constexpr bool zx = CamTraits<CM>::xi_zeros;
#pragma unroll
    for (int rr = 0; rr < 6; ++rr)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double& a = sacc[6 * rr + c];
            if (zx && c == 4) a = fma(M1[rr], B1[c], a);
            else if (zx && c == 3) a = fma(M0[rr], B0[c], a);
            else a = fma(M1[rr], B1[c], fma(M0[rr], B0[c], a));
        }
}

// Phase B: the lane's pairs (entry rows of stride kWave from pl, nrow rows; e0,
// e1 and the weight w0 of e0 already loaded) into sacc, with the camera (R, t).
// Each load has a pair's work to land: at pair j the weight of entry j + 1
// and the entry j + 2 are issued, and used one pair later.
template <typename T, int CM>
__device__ __forceinline__ void schur_pairs(const DevProblem& P, const T* __restrict__ w, const double* wbuf,
                                            const int2* pl, int nrow, int2 e0, int2 e1, double w0,
                                            const double Rb[9], const double tb[3], double sacc[36]) {
    for (int j = 0; j < nrow && e0.x >= 0; ++j) {
        const double w1 = P.robust ? (double)w[e1.x >= 0 ? e1.x : 0] : 1.0;
        const int2 e2 = j + 2 < nrow ? pl[(size_t)(j + 2) * kWave] : int2{-1, 0};
        const double* wb = wbuf + e0.y * kWbuf;
        EdgeLin<T> Lb;
        edge_linearize_m<CM>(P, Rb, tb, wb[18], wb[19], wb[20], 0.0, 0.0, Lb);
        // the Huber weight of b folded into J_X,b: W_a (w_b J_X,b)^T J_xi,b
        double JXb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) JXb[q] = w0 * (double)Lb.JX[q];
        double M0[6], M1[6];
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) {
            const double v0 = wb[3 * rr], v1 = wb[3 * rr + 1], v2 = wb[3 * rr + 2];
            M0[rr] = v0 * JXb[0] + v1 * JXb[1] + v2 * JXb[2];
            M1[rr] = v0 * JXb[3] + v1 * JXb[4] + v2 * JXb[5];
        }
        double B0[6], B1[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) { B0[c] = (double)Lb.Jxi[c]; B1[c] = (double)Lb.Jxi[6 + c]; }
        pair_add<CM>(sacc, M0, M1, B0, B1);
        e0 = e1; w0 = w1;
        e1 = e2;
    }
}

// One pair's factors: W_a (w_b J_X,b)^T = [M0 M1] (6x2) and J_xi,b = [B0; B1]
// (2x6), with b's Jacobians at the point stored in a's W row wb.
template <typename T, int CM>
__device__ __forceinline__ void pair_factors(const DevProblem& P, const double* wb, double wgt, const double Rb[9],
                                             const double tb[3], double M0[6], double M1[6], double B0[6],
                                             double B1[6]) {
    EdgeLin<T> Lb;
    edge_linearize_m<CM>(P, Rb, tb, wb[18], wb[19], wb[20], 0.0, 0.0, Lb);
    double JXb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) JXb[q] = wgt * (double)Lb.JX[q];
#pragma unroll
    for (int rr = 0; rr < 6; ++rr) {
        const double v0 = wb[3 * rr], v1 = wb[3 * rr + 1], v2 = wb[3 * rr + 2];
        M0[rr] = v0 * JXb[0] + v1 * JXb[1] + v2 * JXb[2];
        M1[rr] = v0 * JXb[3] + v1 * JXb[4] + v2 * JXb[5];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) { B0[c] = (double)Lb.Jxi[c]; B1[c] = (double)Lb.Jxi[6 + c]; }
}

// Phase B two pairs at a time (the same sums, in the same order, as
// schur_pairs): the two pairs' Jacobians and products are independent, so
// their latencies overlap.  e[0..3] are the lane's first four entries (or
// {-1, 0} past its share) and wa, wb the weights of the first two.  Each
// step issues the weights of the next two entries and the entries two steps
// ahead.
template <typename T, int CM>
__device__ __forceinline__ void schur_pairs2(const DevProblem& P, const T* __restrict__ w, const double* wbuf,
                                             const int2* pl, int nrow, int2 eA, int2 eB, int2 eC, int2 eD,
                                             double wA, double wB, const double Rb[9], const double tb[3],
                                             double sacc[36]) {
    for (int j = 0; j < nrow && eA.x >= 0; j += 2) {
        const double wC = P.robust ? (double)w[eC.x >= 0 ? eC.x : 0] : 1.0;
        const double wD = P.robust ? (double)w[eD.x >= 0 ? eD.x : 0] : 1.0;
        const int2 eE = j + 4 < nrow ? pl[(size_t)(j + 4) * kWave] : int2{-1, 0};
        const int2 eF = j + 5 < nrow ? pl[(size_t)(j + 5) * kWave] : int2{-1, 0};
        double MA0[6], MA1[6], BA0[6], BA1[6], MB0[6], MB1[6], BB0[6], BB1[6];
        pair_factors<T, CM>(P, wbuf + eA.y * kWbuf, wA, Rb, tb, MA0, MA1, BA0, BA1);
        pair_factors<T, CM>(P, wbuf + eB.y * kWbuf, wB, Rb, tb, MB0, MB1, BB0, BB1);
        pair_add<CM>(sacc, MA0, MA1, BA0, BA1);
        if (eB.x >= 0) pair_add<CM>(sacc, MB0, MB1, BB0, BB1);
        eA = eC; eB = eD; wA = wC; wB = wD;
        eC = eE; eD = eF;
    }
}

// The sum of an aligned group of G lanes (2 or 4, the same in the whole
// workgroup), in every lane of the group.
__device__ __forceinline__ void group_sum(double sacc[36], int G) {
#pragma unroll
    for (int e = 0; e < 36; ++e) sacc[e] += dpp_move<0xB1>(sacc[e]);  // quad_perm [1,0,3,2]: lane pairs
    if (G == 4) {
#pragma unroll
        for (int e = 0; e < 36; ++e) sacc[e] += dpp_move<0x4E>(sacc[e]);  // quad_perm [2,3,0,1]: pairs of pairs
    }
}

// acc[s][e] += the staged group rows of slot s (lanes lfirst[s] .. + lcount[s]),
// four interleaved partial sums in a fixed order; tasks strided by `step` from `t0`.
__device__ __forceinline__ void flush_slots(double* acc, const double* stage, const int* lfirst, const int* lcount,
                                            int ns, int G, int t0, int step) {
    for (int task = t0; task < ns * 36; task += step) {
        const int s = task / 36, e = task % 36;
        const int n = lcount[s] / G;
        if (n == 0) continue;
        const double* sp = stage + (lfirst[s] / G) * kStageStride + e;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        int u = 0;
        for (; u + 4 <= n; u += 4) {
            v0 += sp[u * kStageStride];
            v1 += sp[(u + 1) * kStageStride];
            v2 += sp[(u + 2) * kStageStride];
            v3 += sp[(u + 3) * kStageStride];
        }
        for (; u < n; ++u) v0 += sp[u * kStageStride];
        acc[task] += (v0 + v1) + (v2 + v3);
    }
}

template <typename T, int CM>
__global__ __launch_bounds__(kBlock, 2) void k_schur_row(DevProblem P, const CamRT* __restrict__ cam,
                                                      const T* __restrict__ w,
                                                      const double* __restrict__ Hl, const LMDev* lm,
                                                      double* __restrict__ spart) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int it = P.sitem_order[blockIdx.x];
    const int f = P.sitem_f[it], a0 = P.sitem_a0[it], a1 = P.sitem_a1[it];
    const int g = P.sitem_g[it], ns = P.sitem_nslots[it];
    const int cf = P.free_cam[f];
    const int win = P.cam_win[cf];
    const D4* __restrict__ X = x_cur(P, lm, win);  // the window's current points
    if (!running(lm, win)) return;
    const double lambda = lm[win].lambda;
    const bool g0 = (g == 0);
    const SchurLayout L(ns);
    const int tid = threadIdx.x, lane = tid % kWave;
    const int wid = __builtin_amdgcn_readfirstlane(tid / kWave);
    double* acc = reinterpret_cast<double*>(smem + L.acc);
    double* wbuf = reinterpret_cast<double*>(smem + L.wbuf);
    int* lslot = reinterpret_cast<int*>(smem + L.lslot);
    int* lfirst = reinterpret_cast<int*>(smem + L.lrange);
    int* lcount = lfirst + ns;
    double* red = reinterpret_cast<double*>(smem + L.red);

    const int col0 = P.colptr[f] + g * kMaxSlots;
    for (int i = tid; i < ns * 36; i += kBlock) acc[i] = 0.0;

    double coef[6], dsum[21];
#pragma unroll
    for (int k = 0; k < 6; ++k) coef[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) dsum[k] = 0.0;

    const int rnd0 = P.sitem_rnd0[it];
    const int nrounds = (a1 - a0 + kBlock - 1) / kBlock;
    for (int r = 0; r < nrounds; ++r) {
        const int rnd = rnd0 + r;
        // ---- this lane's slot and first pairs: loads issued before phase A
        const int sl = P.rnd_lane_slot[(size_t)rnd * kBlock + tid];
        const int4 rows = P.rnd_rows[rnd];
        const int nrow = wid == 0 ? rows.x : wid == 1 ? rows.y : wid == 2 ? rows.z : rows.w;
        const int rofs = wid == 0 ? 0 : wid == 1 ? rows.x : wid == 2 ? rows.x + rows.y : rows.x + rows.y + rows.z;
        const int G = __builtin_amdgcn_readfirstlane(P.rnd_group[rnd]);
        const bool busy = sl != 255 && nrow > 0;
        const int2* pl = P.pairs + (size_t)P.rnd_pair0[rnd] + (size_t)rofs * kWave + lane;
        int2 e0 = int2{-1, 0}, e1 = int2{-1, 0};
        double w0 = 1.0;
        double Rb[9], tb[3];
        if (busy) {
            load_cam(cam, P.free_cam[P.colcam[col0 + sl]], Rb, tb);
            e0 = pl[0];
            if (nrow > 1) e1 = pl[kWave];
            if (P.robust && e0.x >= 0) w0 = (double)w[e0.x];
        }
        // ---- phase A: one lane per observation of camera f
        const int k = a0 + r * kBlock + tid;
        if (k < a1) {
            double Rf[9], tf[3];
            load_cam(cam, cf, Rf, tf);
            PaLazy in;
            pa_load_lazy(P, X, Hl, k, P.c_pt[k], in);
            schur_phase_a<T, CM>(P, Rf, tf, in, lambda, g0, wbuf + tid * kWbuf, coef, dsum);
        }
        lslot[tid] = sl;
        for (int s = tid; s < ns; s += kBlock) lcount[s] = 0;
        __syncthreads();
        // lanes of one slot are contiguous: record each slot's first lane and end
        if (sl != 255) {
            if (tid == 0 || lslot[tid - 1] != sl) lfirst[sl] = tid;
            if (tid == kBlock - 1 || lslot[tid + 1] != sl) lcount[sl] = tid + 1;
        }
        // ---- phase B: this lane's pairs of its slot
        double sacc[36];
#pragma unroll
        for (int e = 0; e < 36; ++e) sacc[e] = 0.0;
        if (busy)
            schur_pairs<T, CM>(P, w, wbuf, pl, nrow, e0, e1, w0, Rb, tb, sacc);
        group_sum(sacc, G);
        __syncthreads();  // the buffer's reads done; lane ranges set
        if (tid < ns && lcount[tid] > 0) lcount[tid] -= lfirst[tid];
        // ---- flush: one staged row per lane group, then per slot its groups in a fixed order
        if (tid % G == 0) {
#pragma unroll
            for (int e = 0; e < 36; ++e) wbuf[(tid / G) * kStageStride + e] = sacc[e];
        }
        __syncthreads();
        flush_slots(acc, wbuf, lfirst, lcount, ns, G, tid, kBlock);
        __syncthreads();
    }
    // ---- diagonal block (slot 0 of group 0) and coefficient: the 27 sums of
    // each wave by a reduce-scatter, then the waves in order; S-row slots out
    // (slot-major)
    if (g0) {
        double a27[27];
#pragma unroll
        for (int q = 0; q < 21; ++q) a27[q] = dsum[q];
#pragma unroll
        for (int k = 0; k < 6; ++k) a27[21 + k] = coef[k];
        int k27;
        const double v = wave_reduce_scatter27(a27, lane, k27);
        if ((lane & 1) == 0 && k27 >= 0) red[wid * 27 + k27] = v;
    }
    __syncthreads();
    double* out = spart + P.sitem_pofs[it];
    if (g0 && tid < 27) {
        double s = red[tid];
#pragma unroll
        for (int v = 1; v < kBlock / kWave; ++v) s += red[v * 27 + tid];
        if (tid < 21) {
            int i = 0, c = tid;  // packed upper-triangle index -> (i, c)
            while (c >= 6 - i) { c -= 6 - i; ++i; }
            c += i;
            acc[6 * i + c] += s;
            if (c != i) acc[6 * c + i] += s;
        } else {
            out[36 * ns + tid - 21] = s;
        }
    }
    __syncthreads();
    for (int i = tid; i < ns * 36; i += kBlock) out[i] = acc[i];
}

// ---- Producer/consumer variant: the item's lanes split by role, slots bound per item.
// Waves 0-1 (producers) run phase A on rounds of kPcRound observations, one
// per lane, into one of two W buffers; waves 2-3 (consumers) are bound to one
// slot each for the whole item and sum the pairs of the previous round from
// the other buffer, so the two phases overlap and one barrier per round
// separates them.  Each consumer lane keeps its slot's sum in registers
// across the rounds; the item ends with one flush: every consumer lane's
// row is staged in LDS and each (slot, entry) adds its slot's lanes in lane
// order.  The diagonal block and the coefficient are reduced from the
// producers' registers the same way (staged rows, then fixed partial sums).
// Host tables: rnd_lane_slot is [items][kPcCons], rnd_rows.x / .y are the
// pair rows of the two consumer waves.
struct PcLayout {
    size_t wbuf, lslot, lrange, red, total;
    __host__ __device__ PcLayout(int ns) {
        size_t o = 0;
        wbuf = o;   o += sizeof(double) * kWbuf * 2 * kPcRound;       // two W buffers; staging at the end
        lslot = o;  o += sizeof(int) * kPcCons;
        lrange = o; o += sizeof(int) * 2 * (size_t)ns;                // first consumer lane, end lane per slot
        red = o;    o += sizeof(double) * 4 * 27;                     // diagonal + coefficient, 4 partial sums
        total = (o + 15) & ~(size_t)15;
    }
};
static_assert(kPcCons * kStageStride <= kWbuf * 2 * kPcRound, "consumer rows fit in the W buffers");
static_assert(kPcRound * 27 <= kWbuf * 2 * kPcRound, "producer rows fit in the W buffers");
static_assert(kPcRound + kPcCons == kBlock && kPcRound == 2 * kWave, "two producer and two consumer waves");

template <typename T, int CM>
__global__ __launch_bounds__(kBlock) void k_schur_row_pc(DevProblem P, const CamRT* __restrict__ cam,
                                                         const T* __restrict__ w,
                                                         const double* __restrict__ Hl, const LMDev* lm,
                                                         double* __restrict__ spart) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int it = P.sitem_order[blockIdx.x];
    const int f = P.sitem_f[it], a0 = P.sitem_a0[it], a1 = P.sitem_a1[it];
    const int g = P.sitem_g[it], ns = P.sitem_nslots[it];
    const int cf = P.free_cam[f];
    const int win = P.cam_win[cf];
    const D4* __restrict__ X = x_cur(P, lm, win);  // the window's current points
    if (!running(lm, win)) return;
    const double lambda = lm[win].lambda;
    const bool g0 = (g == 0);
    const PcLayout L(ns);
    const int tid = threadIdx.x, lane = tid % kWave;
    double* wbuf = reinterpret_cast<double*>(smem + L.wbuf);
    int* lslot = reinterpret_cast<int*>(smem + L.lslot);
    int* lfirst = reinterpret_cast<int*>(smem + L.lrange);
    int* lend = lfirst + ns;
    double* red = reinterpret_cast<double*>(smem + L.red);
    const int col0 = P.colptr[f] + g * kMaxSlots;
    const int rnd0 = P.sitem_rnd0[it];
    const int nrounds = (a1 - a0 + kPcRound - 1) / kPcRound;
    for (int s = tid; s < ns; s += kBlock) { lfirst[s] = 0; lend[s] = 0; }

    // The roles split at wave boundaries and the branch is wave-uniform, so each
    // wave runs only its own loop; both loops pass the same barriers in the same
    // order (1 + nrounds, then 2 if g0, then 1).
    if (__builtin_amdgcn_readfirstlane(tid) < kPcRound) {
        // ---- producers: phase A of round r + 1 while the consumers sum round r
        double coef[6], dsum[21], Rf[9], tf[3];
#pragma unroll
        for (int k = 0; k < 6; ++k) coef[k] = 0.0;
#pragma unroll
        for (int k = 0; k < 21; ++k) dsum[k] = 0.0;
        load_cam(cam, cf, Rf, tf);
        // loads run ahead: the point index two rounds ahead, the gathers of the
        // next round, each issued before the barrier that precedes its use
        auto point_of = [&](int r) { const int k = a0 + r * kPcRound + tid; return k < a1 ? P.c_pt[k] : -1; };
        PaIn in;
        pa_load(P, X, Hl, a0 + tid, point_of(0), in);
        int p_next = point_of(1);
        schur_phase_a<T, CM>(P, Rf, tf, in, lambda, g0, wbuf + tid * kWbuf, coef, dsum);
        pa_load(P, X, Hl, a0 + kPcRound + tid, p_next, in);
        p_next = point_of(2);
        __syncthreads();
        for (int r = 0; r < nrounds; ++r) {
            // in: round r + 1; p_next: round r + 2
            schur_phase_a<T, CM>(P, Rf, tf, in, lambda, g0, wbuf + (((r + 1) & 1) * kPcRound + tid) * kWbuf, coef,
                                 dsum);
            pa_load(P, X, Hl, a0 + (r + 2) * kPcRound + tid, p_next, in);
            p_next = point_of(r + 3);
            __syncthreads();  // round r + 1's buffer written, round r's buffer read
        }
        if (g0) {
            // diagonal block and coefficient: rows staged, 4 partial sums of 32 lanes each
            double* row = wbuf + tid * 27;
#pragma unroll
            for (int q = 0; q < 21; ++q) row[q] = dsum[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) row[21 + q] = coef[q];
            __syncthreads();
            if (tid < 4 * 27) {
                const int q = tid % 27, part = tid / 27;
                const double* sp = wbuf + (part * (kPcRound / 4)) * 27 + q;
                double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
#pragma unroll
                for (int u = 0; u < kPcRound / 4; u += 4) {
                    v0 += sp[u * 27];
                    v1 += sp[(u + 1) * 27];
                    v2 += sp[(u + 2) * 27];
                    v3 += sp[(u + 3) * 27];
                }
                red[part * 27 + q] = (v0 + v1) + (v2 + v3);
            }
            __syncthreads();  // partial sums written, producer rows read
        }
        __syncthreads();      // consumer rows staged
    } else {
        // ---- consumers: bound to one slot for the item; pairs of round r
        const int c = tid - kPcRound, cw = c / kWave;
        const int sl = P.rnd_lane_slot[(size_t)it * kPcCons + c];
        lslot[c] = sl;
        double sacc[36], Rb[9], tb[3];
#pragma unroll
        for (int e = 0; e < 36; ++e) sacc[e] = 0.0;
        if (sl != 255) load_cam(cam, P.free_cam[P.colcam[col0 + sl]], Rb, tb);
        // round r's pair rows, first four entries and first two weights; the
        // next round's entries are loaded at the start of a round and their
        // weights at its end, so both land while the round runs
        int nrow = 0, nrow_n = 0;
        const int2* pl = nullptr;
        const int2* pl_n = nullptr;
        int2 e[4], n[4];
        double wa = 1.0, wb = 1.0, wa_n = 1.0, wb_n = 1.0;
        auto fetch_entries = [&](int rnd) {
            const int4 rows = P.rnd_rows[rnd];
            nrow_n = cw == 0 ? rows.x : rows.y;
            pl_n = P.pairs + (size_t)P.rnd_pair0[rnd] + (size_t)(cw == 0 ? 0 : rows.x) * kWave + lane;
#pragma unroll
            for (int q = 0; q < 4; ++q) n[q] = (sl != 255 && q < nrow_n) ? pl_n[q * kWave] : int2{-1, 0};
        };
        auto fetch_weights = [&]() {
            wa_n = 1.0; wb_n = 1.0;
            if (P.robust && sl != 255 && nrow_n > 0) {
                wa_n = (double)w[n[0].x >= 0 ? n[0].x : 0];
                wb_n = (double)w[n[1].x >= 0 ? n[1].x : 0];
            }
        };
        auto advance = [&]() {
            nrow = nrow_n; pl = pl_n;
#pragma unroll
            for (int q = 0; q < 4; ++q) e[q] = n[q];
            wa = wa_n; wb = wb_n;
        };
        fetch_entries(rnd0);
        fetch_weights();
        __syncthreads();
        // slots' consumer lanes are contiguous: each slot's first and end lane
        if (sl != 255) {
            if (c == 0 || lslot[c - 1] != sl) lfirst[sl] = c;
            if (c == kPcCons - 1 || lslot[c + 1] != sl) lend[sl] = c + 1;
        }
        for (int r = 0; r < nrounds; ++r) {
            advance();
            if (r + 1 < nrounds) fetch_entries(rnd0 + r + 1);
            // two pairs per step for the pinhole; the fisheye Jacobians leave no
            // registers for a second pair in flight
            if (sl != 255) {
                if constexpr (CamTraits<CM>::xi_zeros)
                    schur_pairs2<T, CM>(P, w, wbuf + (size_t)(r & 1) * kPcRound * kWbuf, pl, nrow, e[0], e[1], e[2],
                                        e[3], wa, wb, Rb, tb, sacc);
                else
                    schur_pairs<T, CM>(P, w, wbuf + (size_t)(r & 1) * kPcRound * kWbuf, pl, nrow, e[0], e[1], wa, Rb,
                                       tb, sacc);
            }
            if (r + 1 < nrounds) fetch_weights();
            __syncthreads();
        }
        if (g0) {
            __syncthreads();
            __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < 36; ++e) wbuf[c * kStageStride + e] = sacc[e];
        __syncthreads();
    }
    // ---- each (slot, entry) adds its slot's staged consumer rows in lane order
    double* out = spart + P.sitem_pofs[it];
    for (int task = tid; task < ns * 36; task += kBlock) {
        const int s = task / 36, e = task % 36;
        const int l0 = lfirst[s], n = lend[s] - l0;
        const double* sp = wbuf + l0 * kStageStride + e;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        int u = 0;
        for (; u + 4 <= n; u += 4) {
            v0 += sp[u * kStageStride];
            v1 += sp[(u + 1) * kStageStride];
            v2 += sp[(u + 2) * kStageStride];
            v3 += sp[(u + 3) * kStageStride];
        }
        for (; u < n; ++u) v0 += sp[u * kStageStride];
        double v = (v0 + v1) + (v2 + v3);
        if (g0 && s == 0) {
            // packed upper-triangle index of (min, max) of the entry's row and column
            const int i = e / 6, j = e % 6, lo = i < j ? i : j, hi = i < j ? j : i;
            const int q = lo * (11 - lo) / 2 + hi;
            v += (red[q] + red[27 + q]) + (red[54 + q] + red[81 + q]);
        }
        out[task] = v;
    }
    if (g0 && tid < 6) out[36 * ns + tid] = (red[21 + tid] + red[48 + tid]) + (red[75 + tid] + red[102 + tid]);
}

size_t schur_row_smem(const DevProblem& P) {
    return P.schur_pc ? PcLayout(P.max_ns).total : SchurLayout(P.max_ns).total;
}

template <typename T>
int prepare_schur_row(const DevProblem& P) {
    const size_t bytes = schur_row_smem(P);
    if (bytes > 160 * 1024) return -1;
    const bool kb = P.cam_model == kCamKB8;
    const void* fn = P.schur_pc ? (kb ? (const void*)k_schur_row_pc<T, kCamKB8> : (const void*)k_schur_row_pc<T, kCamPinhole>)
                                : (kb ? (const void*)k_schur_row<T, kCamKB8> : (const void*)k_schur_row<T, kCamPinhole>);
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e == hipSuccess ? 0 : -2;
}

template <typename T>
void launch_schur_row(const DevProblem& P, const DevState& s, const DevWork& W, hipStream_t st) {
    if (P.n_sitems == 0) return;
    const bool kb = P.cam_model == kCamKB8;
    auto fn = P.schur_pc ? (kb ? k_schur_row_pc<T, kCamKB8> : k_schur_row_pc<T, kCamPinhole>)
                         : (kb ? k_schur_row<T, kCamKB8> : k_schur_row<T, kCamPinhole>);
    hipLaunchKernelGGL(fn, dim3(P.n_sitems), dim3(kBlock), schur_row_smem(P), st, P, s.cam, (const T*)W.w,
                       (const double*)W.Hl, (const LMDev*)W.lm, W.spart);
}

template void launch_schur_row<double>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);
template void launch_schur_row<float>(const DevProblem&, const DevState&, const DevWork&, hipStream_t);
template int prepare_schur_row<double>(const DevProblem&);
template int prepare_schur_row<float>(const DevProblem&);

}  // namespace navba
