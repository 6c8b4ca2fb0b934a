// chol_kernels.hip — FP64 dense Cholesky of the reduced camera system and the
// two triangular solves (the factorisation g2o's LinearSolverEigen performs at
// BE_GraphOptim.cpp:43; here a plain LL^T, which for S + lambda I > 0 gives the
// same step as LDL^T).
//
// Right-looking, 64x64 tiles, the matrix padded with the identity to a tile
// multiple.  One launch per block column k:
//   every workgroup (i, j), k < j <= i, solves L_ik and L_jk from the updated
//   A_ik, A_jk and the finished L_kk by substitution (rows in parallel), then
//   updates A_ij -= L_ik L_jk^T;  workgroups (i, k+1) also store L_ik, and the
//   workgroup (k+1, k+1) factors its freshly updated tile in LDS.
// The factor goes to a separate buffer, so no launch overwrites a tile that
// another workgroup of the same launch still reads.
#include "ba_kernels.h"

namespace navba {

typedef double TileLds[64][65];

__device__ __forceinline__ void tile_load(TileLds& T, const double* __restrict__ src, int lda) {
    for (int idx = threadIdx.x; idx < 4096; idx += 256) T[idx >> 6][idx & 63] = src[(size_t)(idx >> 6) * lda + (idx & 63)];
}

__device__ __forceinline__ void tile_store(const TileLds& T, double* __restrict__ dst, int lda) {
    for (int idx = threadIdx.x; idx < 4096; idx += 256) dst[(size_t)(idx >> 6) * lda + (idx & 63)] = T[idx >> 6][idx & 63];
}

// In-LDS Cholesky of the lower triangle of A (64x64, 256 threads): A <- L with
// a zero upper triangle.  Four 16-column panels; each panel is factored by
// wave 0 alone, lane r holding row r of the panel in registers (pivots and
// column entries are broadcast with readlane, fully unrolled, no branches),
// then all four waves apply the panel to the trailing lower triangle with
// 4x4 register blocks.  Returns true (in every thread) if a pivot was not
// positive and finite (it is then replaced by 1 so the values stay finite).
__device__ __forceinline__ bool tile_potrf(TileLds& A, int* bad_flag) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) *bad_flag = 0;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int c0 = 16 * p;
        if (wid == 0) {
            double a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = A[lane][c0 + j];
            bool bad = false;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                double d = readlane(a[j], c0 + j);
                const bool ok = (d > 0.0) && isfinite(d);
                bad |= !ok;
                d = ok ? d : 1.0;
                double r = __builtin_amdgcn_rsq(d);  // 1/sqrt(d): estimate and two Newton steps
                r = r * (1.5 - 0.5 * d * r * r);
                r = r * (1.5 - 0.5 * d * r * r);
                const double l = a[j] * r;
                a[j] = (lane > c0 + j) ? l : ((lane == c0 + j) ? d * r : a[j]);
#pragma unroll
                for (int c = j + 1; c < 16; ++c) {
                    const double lc = readlane(a[j], c0 + c);
                    a[c] = (lane >= c0 + c) ? a[c] - a[j] * lc : a[c];
                }
            }
            if (lane >= c0) {
#pragma unroll
                for (int j = 0; j < 16; ++j) A[lane][c0 + j] = a[j];
            }
            if (bad && lane == 0) *bad_flag = 1;
        }
        __syncthreads();
        if (p < 3) {
            // A[r][c] -= sum_j L[r][c0+j] L[c][c0+j] for c0+16 <= c <= r
            const int r0 = 4 * (tid >> 4), cc0 = 4 * (tid & 15);
            if (cc0 + 3 >= c0 + 16 && r0 + 3 >= cc0) {
                double acc[4][4] = {};
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    double x[4], y[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) { x[u] = A[r0 + u][c0 + j]; y[u] = A[cc0 + u][c0 + j]; }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * y[v];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (cc0 + v >= c0 + 16 && r0 + u >= cc0 + v) A[r0 + u][cc0 + v] -= acc[u][v];
            }
            __syncthreads();
        }
    }
    for (int idx = tid; idx < 4096; idx += 256)
        if ((idx & 63) > (idx >> 6)) A[idx >> 6][idx & 63] = 0.0;
    __syncthreads();
    return *bad_flag != 0;
}

// y_k of the forward substitution L y = g, for the diagonal tile k whose
// factor D is in LDS: r = g_k - sum_{j<k} L_kj y_j (rows of the factor in
// Lm, y_j already in global memory), then D y_k = r column by column in wave 0
// with readlane broadcasts.
__device__ __forceinline__ void tile_forward(const TileLds& D, const double* __restrict__ Lm, int lda, int k,
                             const double* __restrict__ g, double* __restrict__ y, double* rv) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    {
        const int r = tid >> 2, q = tid & 3;
        const double* row = Lm + (size_t)(64 * k + r) * lda;
        double s = 0.0;
        for (int m = q; m < 64 * k; m += 4) s += row[m] * y[m];
        s += dpp_move<0xB1>(s);
        s += dpp_move<0x4E>(s);
        if (q == 0) rv[r] = g[64 * k + r] - s;
    }
    __syncthreads();
    if (wid == 0) {
        double rm = rv[lane];
        const double dm = 1.0 / D[lane][lane];
        double ym = 0.0;
#pragma unroll 8
        for (int c = 0; c < 64; ++c) {
            const double lmc = D[lane][c];
            const double yc = readlane(rm * dm, c);
            if (lane == c) ym = yc;
            if (lane > c) rm -= lmc * yc;
        }
        y[64 * k + lane] = ym;
    }
    __syncthreads();
}

// B <- B L^-T for a 64x64 tile B and lower-triangular L (both in LDS),
// inv[c] = 1 / L[c][c].  Four threads per row share the dot products.
__device__ __forceinline__ void tile_trsm(const TileLds& L, const double* inv, TileLds& B) {
    const int tid = threadIdx.x;
    const int r = tid >> 2, q = tid & 3;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = 0.0;
#pragma unroll
    for (int c = 0; c < 64; ++c) {
        double part = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (4 * i + 3 < c || (4 * i < c && q + 4 * i < c)) part += x[i] * L[c][q + 4 * i];
        part += dpp_move<0xB1>(part);  // sum over the 4 lanes of the row (quad)
        part += dpp_move<0x4E>(part);
        const double xc = (B[r][c] - part) * inv[c];
        if (q == (c & 3)) x[c >> 2] = xc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) B[r][q + 4 * i] = x[i];
    __syncthreads();
}

// acc (4x4 per thread) = A B^T, rows 4*(tid>>4)+u, columns (tid&15)+16v.
__device__ __forceinline__ void tile_gemm_nt(const TileLds& A, const TileLds& B, double acc[4][4]) {
    const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int m = 0; m < 64; ++m) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = A[4 * tr + u][m]; b[u] = B[tc + 16 * u][m]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
}

__global__ __launch_bounds__(256) void k_chol_first(const double* __restrict__ S, double* __restrict__ Lm, int lda,
                                                    const double* __restrict__ g, double* __restrict__ y, LMDev* lm) {
    __shared__ TileLds A;
    __shared__ double rv[64];
    __shared__ int bad;
    tile_load(A, S, lda);
    __syncthreads();
    if (tile_potrf(A, &bad) && threadIdx.x == 0) atomicOr(&lm->fail, 1);
    tile_store(A, Lm, lda);
    tile_forward(A, Lm, lda, 0, g, y, rv);
}

__global__ __launch_bounds__(256) void k_chol_step(double* __restrict__ S, double* __restrict__ Lm, int lda, int k,
                                                   const double* __restrict__ g, double* __restrict__ y, LMDev* lm) {
    __shared__ TileLds Lk;
    __shared__ TileLds Bi;
    __shared__ TileLds Bj;
    __shared__ double inv[64];
    __shared__ double rv[64];
    __shared__ int bad;
    const int idx = blockIdx.x;
    int ii = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((ii + 1) * (ii + 2) / 2 <= idx) ++ii;
    while (ii * (ii + 1) / 2 > idx) --ii;
    const int jj = idx - ii * (ii + 1) / 2;
    const int i = k + 1 + ii, j = k + 1 + jj;
    const bool same = (i == j);
    tile_load(Lk, Lm + (size_t)(64 * k) * lda + 64 * k, lda);
    tile_load(Bi, S + (size_t)(64 * i) * lda + 64 * k, lda);
    if (!same) tile_load(Bj, S + (size_t)(64 * j) * lda + 64 * k, lda);
    __syncthreads();
    if (threadIdx.x < 64) inv[threadIdx.x] = 1.0 / Lk[threadIdx.x][threadIdx.x];
    __syncthreads();
    tile_trsm(Lk, inv, Bi);
    if (!same) tile_trsm(Lk, inv, Bj);
    if (j == k + 1) tile_store(Bi, Lm + (size_t)(64 * i) * lda + 64 * k, lda);
    double acc[4][4];
    tile_gemm_nt(Bi, same ? Bi : Bj, acc);
    const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
    double* Aij = S + (size_t)(64 * i) * lda + 64 * j;
    if (same && i == k + 1) {
        __syncthreads();  // Lk is free again: reuse it for the tile to factor
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                Lk[4 * tr + u][tc + 16 * v] = Aij[(size_t)(4 * tr + u) * lda + tc + 16 * v] - acc[u][v];
        __syncthreads();
        if (tile_potrf(Lk, &bad) && threadIdx.x == 0) atomicOr(&lm->fail, 1);
        tile_store(Lk, Lm + (size_t)(64 * i) * lda + 64 * i, lda);
        __syncthreads();  // L_i,k (stored above by this workgroup) is read back by the forward step
        tile_forward(Lk, Lm, lda, i, g, y, rv);
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) Aij[(size_t)(4 * tr + u) * lda + tc + 16 * v] -= acc[u][v];
    }
}

// Backward substitution L^T x = y with the tiled factor (the forward half ran
// inside the factorisation launches); one workgroup.  Within a diagonal tile
// (staged in LDS) one wave solves column by column: lane m keeps the running
// right-hand side r_m, the solved value of step c is broadcast with readlane
// and lanes m < c update r_m -= L[c][m] x_c (no reductions on the dependency
// chain).  dp = 0 when the factorisation failed.
__global__ __launch_bounds__(256) void k_chol_solve(const double* __restrict__ Lm, int lda, int nt,
                                                    const double* __restrict__ y, double* __restrict__ dp,
                                                    const LMDev* lm) {
    __shared__ TileLds D;
    __shared__ double part[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (lm->fail) {
        for (int i = tid; i < 64 * nt; i += 256) dp[i] = 0.0;
        return;
    }
    for (int k = nt - 1; k >= 0; --k) {
        tile_load(D, Lm + (size_t)(64 * k) * lda + 64 * k, lda);
        double s = 0.0;
        for (int m = 64 * (k + 1) + wid; m < 64 * nt; m += 4) s += Lm[(size_t)m * lda + 64 * k + lane] * dp[m];
        part[wid][lane] = s;
        __syncthreads();
        if (wid == 0) {
            double rm = y[64 * k + lane] - (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
            const double dm = 1.0 / D[lane][lane];
            double xm = 0.0;
#pragma unroll 8
            for (int c = 63; c >= 0; --c) {
                const double lcm = D[c][lane];
                const double xc = readlane(rm * dm, c);
                if (lane == c) xm = xc;
                if (lane < c) rm -= lcm * xc;
            }
            dp[64 * k + lane] = xm;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- per-window systems
// One workgroup (512 threads) per running window.  The window's reduced
// system (n = 6 nb unknowns, nb <= 30 free cameras) is held in LDS as a packed
// lower triangle of 6x6 camera blocks (block (I, J), J <= I, at
// (I(I+1)/2 + J) * kBS doubles, row-major inside; the odd stride kBS = 37
// spreads neighbouring blocks over LDS banks).  Right-looking block Cholesky,
// per block column K:
//   diag   L_KK = chol(A_KK), M_K = L_KK^-1 and y_K = M_K y_K, in registers of
//          one thread (1/sqrt by rsq + two Newton steps) - done by the thread
//          that applied the last update to A_KK, right after it;
//   panel  L_IK = A_IK M_K^T and y_I -= L_IK y_K, one thread per row of a block;
//   update A_IJ -= L_IK L_JK^T for K < J <= I, one thread per 6x6 block (36
//          register accumulators).
// Two barriers per block column.  Backward substitution x_K = M_K^T (y_K -
// sum_{I>K} L_IK^T x_I) runs block by block.  A pivot that is not positive and
// finite marks the window failed (dp = 0, lm.fail = 1), as the tiled path does.
constexpr int kBS = 37;
constexpr int kSmallThreads = 512;

__device__ __forceinline__ int blk_off(int I, int J) { return (I * (I + 1) / 2 + J) * kBS; }

// Lower-triangular 6x6 Cholesky and inverse of the block at a (row-major 6x6,
// lower part read), result L written back to a (lower part) and M = L^-1 to m
// (full 6x6, upper part zero); y <- M y.  Returns false on a bad pivot.
__device__ __forceinline__ bool diag_factor(double* a, double* m, double* y) {
    double l[6][6], inv[6];
    bool ok_all = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = a[6 * j + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
        const bool ok = (d > 0.0) && isfinite(d);
        ok_all = ok_all && ok;
        d = ok ? d : 1.0;
        double r = __builtin_amdgcn_rsq(d);
        r = r * (1.5 - 0.5 * d * r * r);
        r = r * (1.5 - 0.5 * d * r * r);
        l[j][j] = d * r;
        inv[j] = r;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double v = a[6 * i + j];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= l[i][k] * l[j][k];
            l[i][j] = v * r;
        }
    }
    double M[6][6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
#pragma unroll
        for (int i = 0; i < 6; ++i) M[i][j] = 0.0;
        M[j][j] = inv[j];
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double v = 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) v += l[i][k] * M[k][j];
            M[i][j] = -v * inv[i];
        }
    }
    double yn[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k <= i; ++k) v += M[i][k] * y[k];
        yn[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        y[i] = yn[i];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (j <= i) a[6 * i + j] = l[i][j];
            m[6 * i + j] = M[i][j];
        }
    }
    return ok_all;
}

__global__ __launch_bounds__(kSmallThreads) void k_chol_small(DevProblem P, const double* __restrict__ S,
                                                              const double* __restrict__ g,
                                                              double* __restrict__ dp, LMDev* lm) {
    extern __shared__ double sm[];
    __shared__ double xk[6];
    __shared__ int bad;
    const int win = blockIdx.x;
    if (!running(lm, win)) return;
    const int nb = P.win_fc0[win + 1] - P.win_fc0[win];
    if (nb == 0) return;
    const int n = 6 * nb;
    const int nblk = nb * (nb + 1) / 2;
    double* Lb = sm;                      // nblk * kBS
    double* Mb = sm + (size_t)nblk * kBS;  // nb * kBS: M_K = L_KK^-1
    double* yv = Mb + (size_t)nb * kBS;    // n
    const double* Sw = S + P.win_S[win];
    const int lda = P.win_lda[win];
    const int g0 = 6 * P.win_fc0[win];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nblk * 36; idx += kSmallThreads) {
        const int b = idx / 36, e = idx - 36 * (idx / 36);
        int I = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
        while ((I + 1) * (I + 2) / 2 <= b) ++I;
        while (I * (I + 1) / 2 > b) --I;
        const int J = b - I * (I + 1) / 2;
        const int r = e / 6, c = e - 6 * (e / 6);
        Lb[b * kBS + e] = Sw[(size_t)(6 * I + r) * lda + 6 * J + c];
    }
    for (int i = tid; i < n; i += kSmallThreads) yv[i] = g[g0 + i];
    if (tid == 0) bad = 0;
    __syncthreads();
    if (tid == 0) {
        if (!diag_factor(Lb + blk_off(0, 0), Mb, yv)) bad = 1;
    }
    __syncthreads();
    for (int K = 0; K < nb; ++K) {
        // panel: L_IK = A_IK M_K^T, y_I -= L_IK y_K
        const int rows = 6 * (nb - K - 1);
        const double* M = Mb + K * kBS;
        for (int t = tid; t < rows; t += kSmallThreads) {
            const int I = K + 1 + t / 6, r = t - 6 * (t / 6);
            double* row = Lb + blk_off(I, K) + 6 * r;
            double a[6], o[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) a[c] = row[c];
            double yr = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double v = 0.0;
#pragma unroll
                for (int m = 0; m <= c; ++m) v += a[m] * M[6 * c + m];
                o[c] = v;
                yr += v * yv[6 * K + c];
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) row[c] = o[c];
            yv[6 * I + r] -= yr;
        }
        __syncthreads();
        // trailing update A_IJ -= L_IK L_JK^T; block (K+1, K+1) is item 0 and its
        // thread factors it right away
        const int rem = nb - K - 1;
        const int nupd = rem * (rem + 1) / 2;
        for (int q = tid; q < nupd; q += kSmallThreads) {
            // q -> (I, J) with I, J in K+1.., J <= I, column-major from the diagonal
            int jj = 0, qq = q, len = rem;
            while (qq >= len) { qq -= len; ++jj; --len; }
            const int J = K + 1 + jj, I = J + qq;
            const double* A = Lb + blk_off(I, K);
            const double* B = Lb + blk_off(J, K);
            double* C = Lb + blk_off(I, J);
            double acc[36];
#pragma unroll
            for (int e = 0; e < 36; ++e) acc[e] = C[e];
            double bm[36];
#pragma unroll
            for (int e = 0; e < 36; ++e) bm[e] = B[e];
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double ar[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) ar[m] = A[6 * r + m];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    double v = acc[6 * r + c];
#pragma unroll
                    for (int m = 0; m < 6; ++m) v -= ar[m] * bm[6 * c + m];
                    acc[6 * r + c] = v;
                }
            }
#pragma unroll
            for (int e = 0; e < 36; ++e) C[e] = acc[e];
            if (q == 0) {  // block (K+1, K+1): this thread's own writes are visible to it
                double yk[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) yk[i] = yv[6 * J + i];
                if (!diag_factor(C, Mb + J * kBS, yk)) bad = 1;
#pragma unroll
                for (int i = 0; i < 6; ++i) yv[6 * J + i] = yk[i];
            }
        }
        __syncthreads();
    }
    if (bad) {
        for (int i = tid; i < n; i += kSmallThreads) dp[g0 + i] = 0.0;
        if (tid == 0) lm[win].fail = 1;
        return;
    }
    // backward: x_K = M_K^T (y_K - sum_{I>K} L_IK^T x_I); y holds the updated right-hand sides
    for (int K = nb - 1; K >= 0; --K) {
        if (tid < 6) {
            const double* M = Mb + K * kBS;
            double v = 0.0;
            for (int m = tid; m < 6; ++m) v += M[6 * m + tid] * yv[6 * K + m];
            xk[tid] = v;
        }
        __syncthreads();
        if (tid < 6) {
            yv[6 * K + tid] = xk[tid];
            dp[g0 + 6 * K + tid] = xk[tid];
        }
        for (int t = tid; t < 6 * K; t += kSmallThreads) {
            const int J = t / 6, c = t - 6 * (t / 6);
            const double* L = Lb + blk_off(K, J);
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r) v += L[6 * r + c] * xk[r];
            yv[t] -= v;
        }
        __syncthreads();
    }
}

size_t cholesky_small_smem(const DevProblem& P) {
    const size_t nb = (size_t)P.max_win_fc;
    return sizeof(double) * (nb * (nb + 1) / 2 * kBS + nb * kBS + 6 * nb);
}

void launch_cholesky_small(const DevProblem& P, const DevWork& W, hipStream_t st) {
    hipLaunchKernelGGL(k_chol_small, dim3(P.n_win), dim3(kSmallThreads), cholesky_small_smem(P), st, P,
                       (const double*)W.S, (const double*)W.g, W.dp, W.lm);
}

int prepare_cholesky_small(const DevProblem& P) {
    const size_t smem = cholesky_small_smem(P);
    if (smem > 160 * 1024) return -1;
    return hipFuncSetAttribute((const void*)k_chol_small, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) ==
                   hipSuccess ? 0 : -1;
}

void launch_cholesky_solve(const DevProblem& P, const DevWork& W, hipStream_t st) {
    if (P.n_fc == 0) return;
    const int nt = P.nt, lda = P.n_pad;
    hipLaunchKernelGGL(k_chol_first, dim3(1), dim3(256), 0, st, (const double*)W.S, W.Lm, lda, (const double*)W.g,
                       W.y, W.lm);
    for (int k = 0; k + 1 < nt; ++k) {
        const int rem = nt - k - 1;
        hipLaunchKernelGGL(k_chol_step, dim3(rem * (rem + 1) / 2), dim3(256), 0, st, W.S, W.Lm, lda, k,
                           (const double*)W.g, W.y, W.lm);
    }
    hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(256), 0, st, (const double*)W.Lm, lda, nt, (const double*)W.y,
                       W.dp, (const LMDev*)W.lm);
}

}  // namespace navba
