// chol_kernels.hip — FP64 dense Cholesky of the reduced camera system and the
// two triangular solves (the factorisation g2o's LinearSolverEigen performs at
// BE_GraphOptim.cpp:43; here a plain LL^T, which for S + lambda I > 0 gives the
// same step as LDL^T).
//
// One problem: right-looking, 64x64 tiles, the matrix padded with the identity
// to a tile multiple, one launch per block column k.  Each diagonal tile is
// factored together with the inverse of its factor, M_k = L_kk^-1, so that
//   panel    L_ik = A_ik M_k^T                      (a 64x64x64 product),
//   forward  y_k = M_k (g_k - sum_{j<k} L_kj y_j)   (matrix-vector),
//   backward x_k = M_k^T (y_k - sum_{m>k} L_mk^T x_m)
// need no substitution chains.  On the reduced systems of the 60-keyframe
// problem (n = 348, condition 4e6, lambda 1e-4 .. 1e-10) this changes the step
// by at most 4e-13 relative against a substitution-based factorisation.
// Launch k: workgroup (i, j), k < j <= i, forms L_ik (and L_jk) from the
// updated A_ik, A_jk and M_k, then updates A_ij -= L_ik L_jk^T; workgroups
// (i, k+1) store L_ik, and workgroup (k+1, k+1) factors its freshly updated
// tile, inverts it and runs that tile's forward step.  The factor goes to a
// separate buffer, so no workgroup overwrites a tile another one still reads;
// the diagonal factors L_kk themselves are never stored (only M_k is used).
// The tile products run on FP64 MFMA (v_mfma_f64_16x16x4); the chain through
// the pivots runs in registers.  The backward step is a single launch.
#include "ba_kernels.h"

namespace navba {

typedef double TileLds[64][65];

// A 64x64 tile through registers, 16 elements per thread: fetch issues the
// global loads, put writes LDS (several fetches can be in flight before the
// first put).
__device__ __forceinline__ void tile_fetch(double v[16], const double* __restrict__ src, int lda) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int idx = threadIdx.x + 256 * s;
        v[s] = src[(size_t)(idx >> 6) * lda + (idx & 63)];
    }
}

__device__ __forceinline__ void tile_put(TileLds& T, const double v[16]) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int idx = threadIdx.x + 256 * s;
        T[idx >> 6][idx & 63] = v[s];
    }
}

__device__ __forceinline__ void tile_load(TileLds& T, const double* __restrict__ src, int lda) {
    double v[16];
    tile_fetch(v, src, lda);
    tile_put(T, v);
}

__device__ __forceinline__ void tile_store(const TileLds& T, double* __restrict__ dst, int lda) {
    double v[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int idx = threadIdx.x + 256 * s;
        v[s] = T[idx >> 6][idx & 63];
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int idx = threadIdx.x + 256 * s;
        dst[(size_t)(idx >> 6) * lda + (idx & 63)] = v[s];
    }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One FP64 16x16x4 matrix FMA of a wave, D = A B + C, with A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k and D[i][j] in element i / 4 of lane
// j + 16 (i % 4) (checked on gfx950 against a host product).
__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Index of the calling wave in its workgroup, as a scalar (wave-uniform
// branches and block indices stay on the scalar unit).
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Orders the LDS accesses of one wave around a hand-off between its lanes
// (the wave's LDS instructions execute in program order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// M[C..C+15][C..C+15] = L^-1 of the 16x16 diagonal block of the factor in A
// (lower triangle, rpiv = 1 / diag), by one wave: its two 8x8 diagonal blocks
// by substitution (one lane per column), then M21 = -M22 (L21 M11) on FP64
// MFMA (8x8 products, zero padded to 16x16).  The block's upper right 8x8
// quarter is not written (it must already hold zeros).
__device__ __forceinline__ void wave_inv16(const TileLds& A, const double* rpiv, TileLds& M, int C) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, kq = lane >> 4, r8 = lane & 7;
    const int R = C + 8;
    if (lane < 16) {
        const int b = C + 8 * (lane >> 3), j = lane & 7;
        double d[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) d[i] = (i == j) ? rpiv[b + j] : 0.0;
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) s += A[b + i][b + k] * d[k];  // d[k] = 0 for k < j
            if (i > j) d[i] = -s * rpiv[b + i];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) M[b + i][b + j] = d[i];
    }
    wave_sync();
    // T = L21 M11 (8x8): rows i = 4v + kq < 8, columns r16 < 8 of the 16x16
    // product (operand rows and columns 8..15 zero; row 8 + r8 read instead)
    const bool in = r16 < 8;
    dbl4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m0 = 0; m0 < 8; m0 += 4) {
        const double va = A[R + r8][C + m0 + kq], vb = M[C + m0 + kq][C + r16];
        t = mfma16x16x4(in ? va : 0.0, in ? vb : 0.0, t);
    }
    if (in) {
        M[R + kq][C + r16] = t[0];
        M[R + 4 + kq][C + r16] = t[1];
    }
    wave_sync();
    dbl4 u = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m0 = 0; m0 < 8; m0 += 4) {
        const double va = M[R + r8][R + m0 + kq], vb = M[R + m0 + kq][C + r16];
        u = mfma16x16x4(in ? va : 0.0, in ? vb : 0.0, u);
    }
    wave_sync();
    if (in) {
        M[R + kq][C + r16] = -u[0];
        M[R + 4 + kq][C + r16] = -u[1];
    }
    wave_sync();
}

// 16x16 block product of one wave, acc += X[xr.., xc..] Y[yr.., yc..]; the
// result's element v in lane (r16, kq) is row 4v + kq, column r16.
__device__ __forceinline__ dbl4 blk_mul(const TileLds& X, int xr, int xc, const TileLds& Y, int yr, int yc, dbl4 acc) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
    double a[4], b[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        a[kk] = X[xr + r16][xc + 4 * kk + kq];
        b[kk] = Y[yr + 4 * kk + kq][yc + r16];
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) acc = mfma16x16x4(a[kk], b[kk], acc);
    return acc;
}

// Block (r, c) of M (16x16) <- sign * acc, in the layout of blk_mul.
__device__ __forceinline__ void blk_put(TileLds& M, int r, int c, dbl4 acc, double sign) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) M[16 * r + 4 * v + kq][16 * c + r16] = sign * acc[v];
}

// Block (i, j) of the tile inverse, M_ij = -M_ii T_ij, where block (i, j) of M
// holds T_ij = sum_{j<=k<i} L_ik M_kj (one wave; in place).
__device__ __forceinline__ void blk_finish(TileLds& M, int i, int j) {
    const dbl4 z = {0.0, 0.0, 0.0, 0.0};
    const dbl4 t = blk_mul(M, 16 * i, 16 * i, M, 16 * i, 16 * j, z);
    wave_sync();
    blk_put(M, i, j, t, -1.0);
}

// Block (i, j) of M += L_ik M_kj (or = when first), one wave.
__device__ __forceinline__ void blk_term(const TileLds& A, TileLds& M, int i, int k, int j, bool first) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    if (!first) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = M[16 * i + 4 * v + kq][16 * j + r16];
    }
    acc = blk_mul(A, 16 * i, 16 * k, M, 16 * k, 16 * j, acc);
    wave_sync();
    blk_put(M, i, j, acc, 1.0);
}

// The value of v in lane l (l wave-uniform), as a uniform double.
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// In-LDS Cholesky of the SPD tile A (64x64, 256 threads) and the inverse of
// its factor: the lower triangle of A becomes L, rpiv[j] = 1 / L_jj and
// M = L^-1 (upper triangle zero).  The strict upper triangle of A's four
// 16x16 diagonal blocks is left with scratch values, the rest of A's upper
// triangle is unchanged.  Wave w owns columns 16w..16w+15, lane l row l.
// Four stages, one barrier between stages; in stage s:
//   waves w >= s apply panel s-1 (final since the last barrier) to their
//     columns, rows 16w.., on FP64 MFMA (16x16 blocks, K = 16, in place);
//   wave s then factors its 16 columns in registers.  Pivot j: 1/sqrt(d_j)
//     by rsq and two Goldschmidt steps; the next pivot is formed from
//     wave-uniform values read before this pivot's column update,
//     d_{j+1} = a_{j+1,j+1} - (a_{j+1,j} / sqrt(d_j))^2, so that the chain
//     through the pivots holds no cross-lane step; columns j+1 and j+2 are
//     updated with uniform multipliers, the others with the scaled column
//     broadcast through `scratch` (16 x 64 doubles of LDS, not M);
//   wave s-1, whose panel is final, forms the inverse's block row s-1 in 16x16
//     blocks, M_ii = L_ii^-1 and M_ij = -M_ii T_ij, and adds its terms to
//     the sums T_mj = sum_{j<=k<m} L_mk M_kj of the rows below (kept in M's
//     blocks (m, j)); stage 0's idle waves clear M's upper triangle.
// After the last stage wave 3 forms block row 3.  A pivot that is not
// positive and finite is replaced by 1 (the values stay finite); the return
// value, the same in every thread, says whether one was.
__device__ __forceinline__ bool tile_potrf_inv(TileLds& A, double* rpiv, double* scratch, TileLds& M, int nst) {
    const int lane = threadIdx.x & 63, w = wave_id();
    const int r16 = lane & 15, kq = lane >> 4;
    const int c0 = 16 * w;
    double (*cb)[64] = reinterpret_cast<double (*)[64]>(scratch);
    bool bad = false;
    for (int s = 0; s < nst; ++s) {
        if (s == 0 && w > 0) {
            for (int idx = (int)threadIdx.x - 64; idx < 4096; idx += 192) {
                const int i = idx >> 6, j = idx & 63;
                if ((j >> 3) > (i >> 3)) M[i][j] = 0.0;
            }
        }
        if (s == 0 && w >= nst) {
            // identity padding rows (the last tile of the system): their part of
            // the factor is the identity already in A; M's rows are the identity
            for (int idx = lane; idx < 16 * 64; idx += kWave) {
                const int i = c0 + (idx >> 6), j = idx & 63;
                if ((j >> 3) <= (i >> 3)) M[i][j] = (i == j) ? 1.0 : 0.0;
            }
            if (lane < 16) rpiv[c0 + lane] = 1.0;
        }
        if (s > 0 && w == s - 1) {
            // block row s-1 of the inverse and its terms of the rows below
            wave_inv16(A, rpiv, M, 16 * w);
            for (int j = 0; j < w; ++j) blk_finish(M, w, j);
            for (int m = w + 1; m < nst; ++m)
                for (int j = 0; j <= w; ++j) blk_term(A, M, m, w, j, j == w);
        }
        if (s > 0 && w >= s && w < nst) {
            const int p = 16 * (s - 1);
            double bv[4], av[4][4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) bv[kk] = A[c0 + r16][p + 4 * kk + kq];
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
                if (bi >= w)
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) av[bi][kk] = A[16 * bi + r16][p + 4 * kk + kq];
#pragma unroll
            for (int bi = 0; bi < 4; ++bi) {
                if (bi >= w) {
                    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) acc = mfma16x16x4(av[bi][kk], bv[kk], acc);
#pragma unroll
                    for (int v = 0; v < 4; ++v) A[16 * bi + 4 * v + kq][c0 + r16] -= acc[v];
                }
            }
        }
        if (w == s) {
            wave_sync();
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = A[lane][c0 + c];
            double rp[16];
            double d = lane_value(x[0], c0);
            double a1 = lane_value(x[0], c0 + 1), a2 = lane_value(x[0], c0 + 2), ad = lane_value(x[1], c0 + 1);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const bool ok = (d > 0.0) && isfinite(d);
                bad |= !ok;
                d = ok ? d : 1.0;
                double H = __builtin_amdgcn_rsq(d);
                double g = d * H;
#pragma unroll
                for (int it = 0; it < 2; ++it) {
                    const double e = fma(-g, H, 1.0);
                    const double gh = 0.5 * g, Hh = 0.5 * H;
                    g = fma(gh, e, g);
                    H = fma(Hh, e, H);
                }
                rp[j] = H;
                const double l1 = a1 * H, l2 = a2 * H;
                const double dn = fma(-l1, l1, ad);
                x[j] = (lane == c0 + j) ? g : x[j] * H;
                if (j + 3 < 16) cb[j][lane] = x[j];
                if (j + 1 < 16) x[j + 1] = fma(-x[j], l1, x[j + 1]);
                if (j + 2 < 16) x[j + 2] = fma(-x[j], l2, x[j + 2]);
                if (j + 1 < 16) {
                    d = dn;
                    a1 = (j + 2 < 16) ? lane_value(x[j + 1], c0 + j + 2) : 0.0;
                    a2 = (j + 3 < 16) ? lane_value(x[j + 1], c0 + j + 3) : 0.0;
                    ad = (j + 2 < 16) ? lane_value(x[j + 2], c0 + j + 2) : 0.0;
                }
                if (j + 3 < 16) {
                    wave_sync();
#pragma unroll
                    for (int c = j + 3; c < 16; ++c) x[c] = fma(-x[j], cb[j][c0 + c], x[c]);
                }
            }
            if (lane >= c0) {
#pragma unroll
                for (int c = 0; c < 16; ++c) A[lane][c0 + c] = x[c];
            }
            if (lane == 0) {
#pragma unroll
                for (int j = 0; j < 16; ++j) rpiv[c0 + j] = rp[j];
            }
        }
        if (s < nst - 1) __syncthreads();
    }
    bad = __syncthreads_or(bad);
    if (w == nst - 1) {
        wave_inv16(A, rpiv, M, 16 * w);
        for (int j = 0; j < w; ++j) blk_finish(M, w, j);
    }
    __syncthreads();
    return bad;
}

// acc[b] = block (wave, b) of X Y^T in 16x16 blocks; with kTriY, Y is lower
// triangular and the zero part of the sum is skipped.
template <bool kTriY>
__device__ __forceinline__ void tile_mfma_nt(const TileLds& X, const TileLds& Y, dbl4 acc[4]) {
    const int lane = threadIdx.x & 63, w = wave_id(), r16 = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m0 = 0; m0 < 64; m0 += 4) {
        const double a = X[16 * w + r16][m0 + kq];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (!kTriY || m0 < 16 * b + 16) acc[b] = mfma16x16x4(a, Y[16 * b + r16][m0 + kq], acc[b]);
    }
}

// B <- B M^T (the panel solve B L^-T with M = L^-1 lower triangular, in LDS).
__device__ __forceinline__ void tile_right_mul_t(TileLds& B, const TileLds& M) {
    dbl4 acc[4];
    tile_mfma_nt<true>(B, M, acc);
    const int lane = threadIdx.x & 63, w = wave_id(), r16 = lane & 15, kq = lane >> 4;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) B[16 * w + 4 * v + kq][16 * b + r16] = acc[b][v];
    __syncthreads();
}

// out = B v for a 64x64 tile B in LDS and a 64-vector v in global memory
// (four threads per row, summed by DPP in a fixed order).
__device__ __forceinline__ void tile_gemv(const TileLds& B, const double* __restrict__ v, double* __restrict__ out) {
    const int tid = threadIdx.x;
    const int r = tid >> 2, q = tid & 3;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int m = q; m < 64; m += 16) {
        s0 += B[r][m] * v[m];
        s1 += B[r][m + 4] * v[m + 4];
        s2 += B[r][m + 8] * v[m + 8];
        s3 += B[r][m + 12] * v[m + 12];
    }
    double s = (s0 + s1) + (s2 + s3);
    s += dpp_move<0xB1>(s);
    s += dpp_move<0x4E>(s);
    if (q == 0) out[r] = s;
}

// y_k = M_k (g_k - sum_{j<=k-1} L_kj y_j).  The products L_kj y_j were formed
// by the workgroups that computed L_kj (tile_gemv into fpart[k][j], 64 values
// each); the nterms of them are summed here in a fixed order, so the row block
// of the factor is not read again.  M_k is in LDS.
__device__ __forceinline__ void tile_forward(const TileLds& M, const double* __restrict__ fpart, int nterms,
                                             const double* __restrict__ g, double* __restrict__ y, double* rv) {
    const int tid = threadIdx.x;
    const int r = tid >> 2, q = tid & 3;
    {
        double s0 = 0.0, s1 = 0.0;
        int j = q;
        for (; j + 4 < nterms; j += 8) {
            s0 += fpart[64 * j + r];
            s1 += fpart[64 * (j + 4) + r];
        }
        if (j < nterms) s0 += fpart[64 * j + r];
        double s = s0 + s1;
        s += dpp_move<0xB1>(s);
        s += dpp_move<0x4E>(s);
        if (q == 0) rv[r] = g[r] - s;
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
    for (int m = q; m < 64; m += 8) {  // M[r][m] = 0 for m > r
        s0 += M[r][m] * rv[m];
        s1 += M[r][m + 4] * rv[m + 4];
    }
    double s = s0 + s1;
    s += dpp_move<0xB1>(s);
    s += dpp_move<0x4E>(s);
    if (q == 0) y[r] = s;
    __syncthreads();
}

// 16-column stages of diagonal tile i that hold rows of the n-row system (the
// last tile's padding rows beyond are the identity and are not factored).
__device__ __forceinline__ int tile_stages(int n, int i) { return min(4, (n - 64 * i + 15) >> 4); }

__global__ __launch_bounds__(256) void k_chol_first(const double* __restrict__ S, double* __restrict__ Lm,
                                                    double* __restrict__ Li, int lda, int n, const double* __restrict__ g,
                                                    double* __restrict__ y, LMDev* lm) {
    __shared__ TileLds A;
    __shared__ TileLds M;
    __shared__ double scratch[16 * 64];
    __shared__ double rv[64];
    __shared__ double rpiv[64];
    tile_load(A, S, lda);
    if (lm->phase == kPhaseDone) return;  // a step launched after the solve has finished
    __syncthreads();
    if (tile_potrf_inv(A, rpiv, scratch, M, tile_stages(n, 0)) && threadIdx.x == 0) atomicOr(&lm->fail, 1);
    tile_store(M, Li, 64);
    tile_forward(M, nullptr, 0, g, y, rv);
}

__global__ __launch_bounds__(256) void k_chol_step(double* __restrict__ S, double* __restrict__ Lm,
                                                   double* __restrict__ Li, int lda, int n, int k,
                                                   const double* __restrict__ g, double* __restrict__ y,
                                                   double* __restrict__ fpart, int nt, LMDev* lm) {
    __shared__ TileLds Mk;
    __shared__ TileLds Bi;
    __shared__ TileLds Bj;
    __shared__ double rv[64];
    __shared__ double rpiv[64];
    const int idx = blockIdx.x;
    int ii = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((ii + 1) * (ii + 2) / 2 <= idx) ++ii;
    while (ii * (ii + 1) / 2 > idx) --ii;
    const int jj = idx - ii * (ii + 1) / 2;
    const int i = k + 1 + ii, j = k + 1 + jj;
    const bool same = (i == j);
    // every global load of the launch in flight at once: M_k, A_ik, A_jk and
    // the tile A_ij to update (no other workgroup writes any of them)
    const int lane = threadIdx.x & 63, w = wave_id(), r16 = lane & 15, kq = lane >> 4;
    double* Aij = S + (size_t)(64 * i) * lda + 64 * j;
    double vm[16], vi[16], vj[16], cur[4][4];
    tile_fetch(vm, Li + (size_t)4096 * k, 64);
    tile_fetch(vi, S + (size_t)(64 * i) * lda + 64 * k, lda);
    if (!same) tile_fetch(vj, S + (size_t)(64 * j) * lda + 64 * k, lda);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) cur[b][v] = Aij[(size_t)(16 * w + 4 * v + kq) * lda + 16 * b + r16];
    if (lm->phase == kPhaseDone) return;  // a step launched after the solve has finished (loads in flight meanwhile)
    tile_put(Mk, vm);
    tile_put(Bi, vi);
    if (!same) tile_put(Bj, vj);
    __syncthreads();
    tile_right_mul_t(Bi, Mk);  // L_ik
    if (!same) tile_right_mul_t(Bj, Mk);  // L_jk
    if (j == k + 1) {
        tile_store(Bi, Lm + (size_t)(64 * i) * lda + 64 * k, lda);
        tile_gemv(Bi, y + 64 * k, fpart + ((size_t)i * nt + k) * 64);  // L_ik y_k for the forward step of tile i
    }
    dbl4 acc[4];
    tile_mfma_nt<false>(Bi, same ? Bi : Bj, acc);
    if (same && i == k + 1) {
        // Mk is free (the last product that read it ended with a barrier): it
        // takes the updated tile to factor, Bj its inverse
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) Mk[16 * w + 4 * v + kq][16 * b + r16] = cur[b][v] - acc[b][v];
        __syncthreads();
        if (tile_potrf_inv(Mk, rpiv, &Bi[0][0], Bj, tile_stages(n, i)) && threadIdx.x == 0) atomicOr(&lm->fail, 1);
        tile_store(Bj, Li + (size_t)4096 * i, 64);
        __syncthreads();  // fpart[i][k] (written above by this workgroup) is read back by the forward step
        tile_forward(Bj, fpart + (size_t)i * nt * 64, k + 1, g + 64 * i, y + 64 * i, rv);
    } else {
        // a diagonal tile keeps its upper triangle (never read; left unchanged
        // so that it does not drift over the trials of a solve)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (!same || 16 * w + 4 * v + kq >= 16 * b + r16)
                    Aij[(size_t)(16 * w + 4 * v + kq) * lda + 16 * b + r16] = cur[b][v] - acc[b][v];
    }
}

// Backward step x = L^-T y for up to kBackSingleMax tiles in one workgroup of
// 512 threads, tile by tile
// from the last.  The right-hand sides r (the forward result) are held in LDS
// and updated in place:
//   x_k  = M_k^T r_k      (64x64 product, 8 partial sums per entry),
//   r_i -= L_ki^T x_k     (all columns left of tile k: 128 per pass, 4 partial
//                          sums per column);
// r_k is final when step k starts (every later step has subtracted its
// term), and x_k takes its place in LDS.  The L_ki loads of a step's first
// pass are issued before x_k is formed.  One workgroup keeps the tile chain
// inside one launch.  dp = 0 when the factorisation failed.  The same
// workgroup then forms the candidate poses from x (the camera update of the
// problem's one window) and the camera part of x'(lambda x + b).
constexpr int kBackThreads = 512;

// NP: 128-column passes of the widest update (64 (nt - 1) columns), so that
// the L loads of a step fit in registers.
constexpr int kBackPassMax = (64 * (kBackSingleMax - 1) + 127) / 128;

template <int NP>
__global__ __launch_bounds__(kBackThreads) void k_chol_back(DevProblem P, const double* __restrict__ Lm,
                                                            const double* __restrict__ Li, int lda, int nt,
                                                            const double* __restrict__ y, double* __restrict__ dp,
                                                            const LMDev* lm, const PoseQT* __restrict__ pose,
                                                            const double* __restrict__ Hc, PoseQT* __restrict__ pn,
                                                            CamRT* __restrict__ cn, double* __restrict__ red_scale) {
    extern __shared__ double rs[];  // [64 nt]
    __shared__ double part[8][64];
    __shared__ double red[4][128 * NP];
    __shared__ double xk[64];
    __shared__ double shs[kBackThreads / kWave];
    const int tid = threadIdx.x, n = 64 * nt;
    // the camera inputs (pose, b_p) of the first kBackThreads cameras load
    // while the solve runs
    const CamIn cin = camera_inputs(P, pose, Hc, max(0, min(tid, P.n_cam - 1)));
    if (lm->phase == kPhaseDone) return;
    const int fail = lm->fail;
    if (fail) {
        for (int i = tid; i < n; i += kBackThreads) dp[i] = 0.0;
    } else {
    const int c = tid & 63, grp = tid >> 6;   // x_k: entry c, rows grp + 8u
    const int pc = tid & 127, mq = tid >> 7;  // updates: column pc of each pass, rows 16 mq + u
    // M_k of the next step is loaded during the current one
    double mv[8];
    {
        const double* Mk = Li + (size_t)4096 * (nt - 1);
#pragma unroll
        for (int u = 0; u < 8; ++u) mv[u] = Mk[(grp + 8 * u) * 64 + c];
    }
    for (int i = tid; i < n; i += kBackThreads) rs[i] = y[i];
    __syncthreads();
    for (int k = nt - 1; k >= 0; --k) {
        const int ncols = 64 * k;
        const double* Lk = Lm + (size_t)(64 * k) * lda;
        // every L_ki load of the step, all passes, before x_k is formed
        double lv[NP][16];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int col = 128 * q + pc;
            if (col < ncols) {
#pragma unroll
                for (int u = 0; u < 16; ++u) lv[q][u] = Lk[(size_t)(16 * mq + u) * lda + col];
            }
        }
        {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int u = 0; u < 8; u += 4) {
                s0 += mv[u] * rs[64 * k + grp + 8 * u];
                s1 += mv[u + 1] * rs[64 * k + grp + 8 * u + 8];
                s2 += mv[u + 2] * rs[64 * k + grp + 8 * u + 16];
                s3 += mv[u + 3] * rs[64 * k + grp + 8 * u + 24];
            }
            part[grp][c] = (s0 + s1) + (s2 + s3);
        }
        if (k > 0) {
            const double* Mk = Li + (size_t)4096 * (k - 1);
#pragma unroll
            for (int u = 0; u < 8; ++u) mv[u] = Mk[(grp + 8 * u) * 64 + c];
        }
        __syncthreads();
        if (tid < 64) {
            const double x = ((part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid])) +
                             ((part[4][tid] + part[5][tid]) + (part[6][tid] + part[7][tid]));
            xk[tid] = x;
            rs[64 * k + tid] = x;  // r_k is not read again
            dp[64 * k + tid] = x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int col = 128 * q + pc;
            if (col < ncols) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                for (int u = 0; u < 16; u += 4) {
                    s0 += lv[q][u] * xk[16 * mq + u];
                    s1 += lv[q][u + 1] * xk[16 * mq + u + 1];
                    s2 += lv[q][u + 2] * xk[16 * mq + u + 2];
                    s3 += lv[q][u + 3] * xk[16 * mq + u + 3];
                }
                red[mq][col] = (s0 + s1) + (s2 + s3);
            }
        }
        __syncthreads();
        for (int col = tid; col < ncols; col += kBackThreads)
            rs[col] -= (red[0][col] + red[1][col]) + (red[2][col] + red[3][col]);
        __syncthreads();
    }
    }
    const double lambda = lm->lambda;
    double sc = 0.0;
    for (int c = tid; c < P.n_cam; c += kBackThreads)
        sc += camera_candidate(c == tid ? cin : camera_inputs(P, pose, Hc, c), lambda, fail, c,
                               [&](int f, int i) { return rs[6 * f + i]; }, pn, cn);
    const double s = block_sum<double, kBackThreads>(sc, shs);
    if (tid == 0) red_scale[0] = s;
}

// Backward step for many tiles (more than kBackSingleMax): one launch per tile
// k from the last, so the reads of the factor spread over k + 1 CUs.  The
// right-hand sides r of the tiles still to solve live in y (the forward
// result) and are updated in place: every workgroup of launch k forms
// x_k = M_k^T r_k (a 64x64 product, recomputed by each workgroup instead of
// waiting for one); workgroup k writes it to dp, workgroup i < k subtracts
// L_ki^T x_k from r_i.  r_k is only read during launch k, and final when it
// starts (launches j > k subtracted their terms).  dp = 0 when the
// factorisation failed.
__global__ __launch_bounds__(256) void k_chol_back_tile(const double* __restrict__ Lm, const double* __restrict__ Li,
                                                   int lda, int nt, int k, double* __restrict__ y,
                                                   double* __restrict__ dp, const LMDev* lm) {
    __shared__ double part[4][64];
    __shared__ double xk[64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int i = blockIdx.x;  // 0..k: i == k writes x_k, i < k updates r_i
    // every global load of the launch is issued before the first use: the
    // L_ki column slice (independent of x_k), M_k, r_k and the failure flag
    double lv[16];
    if (i < k) {
        const double* L = Lm + (size_t)(64 * k) * lda + 64 * i + lane;
#pragma unroll
        for (int t = 0; t < 16; ++t) lv[t] = L[(size_t)(wid + 4 * t) * lda];
    }
    double mv[16], rv[16];
    {
        const double* Mk = Li + (size_t)4096 * k;
        const double* rk = y + 64 * k;
#pragma unroll
        for (int t = 0; t < 16; ++t) {  // rows r = wid + 4t; M_k[r][c] = 0 for r < c
            mv[t] = Mk[(wid + 4 * t) * 64 + lane];
            rv[t] = rk[wid + 4 * t];
        }
    }
    if (lm->phase == kPhaseDone) return;  // a step launched after the solve has finished
    const int fail = lm->fail;
    if (fail) {
        if (k == nt - 1 && tid < 64) dp[64 * i + tid] = 0.0;  // launch nt-1 has one workgroup per tile
        return;
    }
    // x_k[c] = sum_r M_k[r][c] r_k[r]: lanes over c, waves over r
    {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int t = 0; t < 16; t += 4) {
            t0 += mv[t] * rv[t];
            t1 += mv[t + 1] * rv[t + 1];
            t2 += mv[t + 2] * rv[t + 2];
            t3 += mv[t + 3] * rv[t + 3];
        }
        part[wid][lane] = (t0 + t1) + (t2 + t3);
    }
    __syncthreads();
    if (wid == 0) xk[lane] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    __syncthreads();
    if (i == k) {
        if (tid < 64) dp[64 * k + tid] = xk[tid];
        return;
    }
    // r_i[c] -= sum_m L[64k + m][64i + c] x_k[m]: lanes over c, waves over m = wid + 4t
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int t = 0; t < 16; t += 4) {
        s0 += lv[t] * xk[wid + 4 * t];
        s1 += lv[t + 1] * xk[wid + 4 * t + 4];
        s2 += lv[t + 2] * xk[wid + 4 * t + 8];
        s3 += lv[t + 3] * xk[wid + 4 * t + 12];
    }
    __syncthreads();
    part[wid][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wid == 0) y[64 * i + lane] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// ---------------------------------------------------------------- per-window systems
// One workgroup (512 threads) per running window, and for a single problem of
// at most 30 free cameras.  The window's reduced system (n = 6 nb unknowns)
// is summed from the Schur rows' partials straight into LDS, as a packed lower
// triangle of 6x6 camera blocks (block (I, J), J <= I, at (I(I+1)/2 + J) * kBS
// doubles, row-major inside; the odd stride kBS = 37 spreads neighbouring
// blocks over LDS banks).  Right-looking block Cholesky, per block column K:
//   panel  L_IK = A_IK M_K^T and y_I -= L_IK y_K, one thread per row of a
//          block (block row K+1 in the first lanes of wave 0);
//   next   wave 0: A_{K+1,K+1} -= L_{K+1,K} L_{K+1,K}^T, one entry per lane,
//          then L = chol(A_{K+1,K+1}), M_{K+1} = L^-1 and y_{K+1} = M_{K+1}
//          y_{K+1} in registers of lane 0 (1/sqrt by rsq + two Newton steps;
//          six lanes with a row each measured no faster);
//   update meanwhile, waves 1..7: A_IJ -= L_IK L_JK^T for the other blocks,
//          K < J <= I, one thread per 6x6 block (36 register accumulators).
// Two barriers per block column; the chain through the columns is wave 0's
// panel rows, its diagonal update and the 6x6 factor.  Backward substitution
// x_K = M_K^T (y_K - sum_{I>K} L_IK^T x_I) runs block by block in wave 0, with
// no workgroup barrier; then the window's candidate poses are formed from the
// step in LDS (the camera update, fused).  A pivot that is not positive and
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

// Block column C of a window's system in LDS, by one wave: first the update
// by panel K (A_IC -= L_IK L_CK^T for I >= C; none when K < 0), then its
// factorisation in registers, one row per lane and slot (rows 6C + lane +
// 64 q): for each of its six columns j, the pivot l_jj = sqrt(d_j) (1/sqrt by
// rsq and two Newton steps, kept in dinv), the rows below scaled by 1/l_jj,
// the forward step y_j = y_j / l_jj and y_r -= l_rj y_j below, and the rank-1
// update of the column block's later columns; the pivots and multipliers come
// from lanes 0..5 as wave-uniform values.  The factor L_IC overwrites the
// column block and y its rows.  Returns false on a pivot that is not
// positive and finite (replaced by 1), the same in every lane.
template <int kQ>
__device__ __forceinline__ bool win_panel_q(double* Lb, double* yv, double* dinv, int n, int C, int K, int lane) {
    const int r0 = 6 * C;
    double v[kQ][6], yq[kQ];
    double u[6][6];  // L_CK, wave-uniform
    if (K >= 0) {
#pragma unroll
        for (int e = 0; e < 36; ++e) u[e / 6][e % 6] = Lb[blk_off(C, K) + e];
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int r = r0 + lane + 64 * q;
        yq[q] = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) v[q][c] = 0.0;
        if (r < n) {
            const int I = r / 6, i = r - 6 * (r / 6);
            const double* a = Lb + blk_off(I, C) + 6 * i;
#pragma unroll
            for (int c = 0; c < 6; ++c) v[q][c] = a[c];
            yq[q] = yv[r];
            if (K >= 0) {
                const double* l = Lb + blk_off(I, K) + 6 * i;
                double lk[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) lk[m] = l[m];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    double w = v[q][c];
#pragma unroll
                    for (int m = 0; m < 6; ++m) w -= lk[m] * u[c][m];
                    v[q][c] = w;
                }
            }
        }
    }
    bool ok_all = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = lane_value(v[0][j], j);
        const bool ok = (d > 0.0) && isfinite(d);
        ok_all = ok_all && ok;
        d = ok ? d : 1.0;
        double rs = __builtin_amdgcn_rsq(d);
        rs = rs * (1.5 - 0.5 * d * rs * rs);
        rs = rs * (1.5 - 0.5 * d * rs * rs);
        const double yj = lane_value(yq[0], j) * rs;
        // unmasked: a row past n (zero) is never written back, and the diagonal
        // block's rows (lanes 0..5 of slot 0) take extra operations only above
        // the diagonal, which nothing reads; only their right-hand side, final
        // once their own pivot is passed, is masked
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            v[q][j] *= rs;
            if (q > 0 || lane > j) yq[q] -= v[q][j] * yj;
        }
        if (lane == j) {
            v[0][j] = d * rs;
            yq[0] = yj;
        }
#pragma unroll
        for (int c = j + 1; c < 6; ++c) {
            const double lcj = lane_value(v[0][j], c);
#pragma unroll
            for (int q = 0; q < kQ; ++q) v[q][c] -= v[q][j] * lcj;
        }
        if (lane == 0) dinv[r0 + j] = rs;
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int r = r0 + lane + 64 * q;
        if (r < n) {
            const int I = r / 6, i = r - 6 * (r / 6);
            double* a = Lb + blk_off(I, C) + 6 * i;
#pragma unroll
            for (int c = 0; c < 6; ++c) a[c] = v[q][c];
            yv[r] = yq[q];
        }
    }
    return ok_all;
}

// win_panel_q with as many row slots as block column C has rows (n - 6C).
__device__ __forceinline__ bool win_panel(double* Lb, double* yv, double* dinv, int n, int C, int K, int lane) {
    static_assert(kSmallMaxN <= 192, "three row slots of 64 lanes");
    const int rows = n - 6 * C;
    if (rows > 128) return win_panel_q<3>(Lb, yv, dinv, n, C, K, lane);
    if (rows > 64) return win_panel_q<2>(Lb, yv, dinv, n, C, K, lane);
    return win_panel_q<1>(Lb, yv, dinv, n, C, K, lane);
}

// Part of one (row, slot) block of a window's reduced system from the Schur
// rows' partials: NE = 36 entries (the whole block, 16-byte loads chunk by
// chunk) or NE = 6 (one row of the block, the loads of up to seven chunks in
// flight at once).  Task index q: block q / (36 / NE), part q % (36 / NE).
// The chunks of an entry are added in the one-problem reduction's order
// (k_schur_reduce): in groups of eight into eight partial sums, the rest into
// the first, combined pairwise.
template <int NE>
__device__ __forceinline__ void sum_block_part(const DevProblem& P, const double* __restrict__ spart,
                                               const double* __restrict__ Hc, double* Lb, const int* srow,
                                               const int* rcol, const int* rns, const int* rnch,
                                               const long long* rp0, int nb, int fc0, double lambda, int q) {
    constexpr int kParts = 36 / NE;
    const int b = q / kParts, part = q - kParts * (q / kParts);
    int lo = 0, hi = nb - 1;  // the row: largest lf with srow[lf] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (srow[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int lf = lo, f = fc0 + lf, s = b - srow[lf];
    const int nch = rnch[lf];
    const long long stride = 36LL * rns[lf] + 6;  // doubles; even, so every item is 16-byte aligned
    const int e0 = NE * part;
    const double* sp = spart + rp0[lf] + 36 * s + e0;
    const int lg = s == 0 ? lf : P.colcam[rcol[lf] + s] - fc0;
    double a[NE];
    double h[NE];  // the diagonal block's Hpp + lambda I entries, loaded ahead
    if (nch < 8) {
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int i = (e0 + k) / 6, c = (e0 + k) % 6;
            h[k] = s == 0 ? Hc[28 * (size_t)f + sym6(i, c)] + (i == c ? lambda : 0.0) : 0.0;
        }
        // fewer than eight chunks: one partial sum in chunk order, then the
        // seven empty ones added pairwise
        const double z = 0.0;
#pragma unroll
        for (int e = 0; e < NE; ++e) a[e] = 0.0;
        if constexpr (NE == 36) {
            for (int ch = 0; ch < nch; ++ch) {
                const double2* p2 = reinterpret_cast<const double2*>(sp + ch * stride);
                double2 v[18];
#pragma unroll
                for (int k = 0; k < 18; ++k) v[k] = p2[k];
#pragma unroll
                for (int k = 0; k < 18; ++k) {
                    a[2 * k] += v[k].x;
                    a[2 * k + 1] += v[k].y;
                }
            }
        } else {
            double2 v[7][3];
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) {
                const double2* p2 = reinterpret_cast<const double2*>(sp + min(ch, nch - 1) * stride);
#pragma unroll
                for (int k = 0; k < 3; ++k) v[ch][k] = p2[k];
            }
#pragma unroll
            for (int ch = 0; ch < 7; ++ch)
                if (ch < nch)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        a[2 * k] += v[ch][k].x;
                        a[2 * k + 1] += v[ch][k].y;
                    }
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) a[e] = ((a[e] + z) + (z + z)) + ((z + z) + (z + z));
    } else {
        // eight chunks or more (small chunks of long camera rows): entry by
        // entry, written as it is summed
        for (int k = 0; k < NE; ++k) {
            double w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = 0.0;
            int ch = 0;
            for (; ch + 8 <= nch; ch += 8) {
#pragma unroll
                for (int u = 0; u < 8; ++u) w[u] += sp[(ch + u) * stride + k];
            }
            for (; ch < nch; ++ch) w[0] += sp[ch * stride + k];
            const double sum = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
            const int i = (e0 + k) / 6, c = (e0 + k) % 6;
            if (s == 0) Lb[blk_off(lf, lf) + 6 * i + c] = Hc[28 * (size_t)f + sym6(i, c)] + (i == c ? lambda : 0.0) - sum;
            else Lb[blk_off(lg, lf) + 6 * c + i] = -sum;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int i = (e0 + k) / 6, c = (e0 + k) % 6;
        if (s == 0) Lb[blk_off(lf, lf) + 6 * i + c] = h[k] - a[k];
        else Lb[blk_off(lg, lf) + 6 * c + i] = -a[k];
    }
}

__global__ __launch_bounds__(kSmallThreads) void k_chol_small(DevProblem P, const double* __restrict__ spart,
                                                              const double* __restrict__ Hc,
                                                              double* __restrict__ dp, LMDev* lm,
                                                              const PoseQT* __restrict__ pose, PoseQT* __restrict__ pn,
                                                              CamRT* __restrict__ cn, double* __restrict__ red_scale) {
    constexpr int kRows = kSmallMaxN / 6;
    extern __shared__ double sm[];
    __shared__ double xk[6];
    __shared__ double ssum[kSmallThreads / 64];
    __shared__ int bad;
    __shared__ int srow[kRows + 1];
    __shared__ int rcol[kRows], rns[kRows], rnch[kRows];
    __shared__ long long rp0[kRows];
    const int win = P.win_chol_order[blockIdx.x];
    if (!running(lm, win)) return;
    const int fc0 = P.win_fc0[win];
    const int nb = P.win_fc0[win + 1] - fc0;
    const int n = 6 * nb;
    const int nblk = nb * (nb + 1) / 2;
    double* Lb = sm;                      // nblk * kBS
    double* yv = sm + (size_t)nblk * kBS;  // n: right-hand side, then the step
    double* dinv = yv + n;                 // n: 1 / l_jj
    const int g0 = 6 * fc0;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = wave_id();
    const double lambda = lm[win].lambda;
    const int fail0 = lm[win].fail;
    if (nb > 0) {
    // ---- the window's reduced system, summed from the Schur rows' partials
    // straight into LDS: per S row f, S_ff = Hpp + lambda I - sum and S_gf =
    // -sum^T over the row's chunk items, and g_f = bp_f - coefficient, the
    // same sums in the same order as the one-problem reduction
    // (k_schur_reduce), so both give the same bits; blocks without
    // co-visibility stay zero.  A row of at most 30 slots is one column group,
    // whose chunk items lie back to back (36 doubles per slot, then 6 of the
    // coefficient).  One thread per block, or per row of a block, sums its
    // entries (sum_block_part), and one more per row the row's coefficient.
    for (int i = tid; i < nblk * kBS; i += kSmallThreads) Lb[i] = 0.0;
    if (wid == 0) {
        int v = 0;
        if (lane < nb) {
            const int f = fc0 + lane;
            const int c0 = P.colptr[f];
            v = P.colptr[f + 1] - c0;
            rcol[lane] = c0;
            rns[lane] = v;
            rnch[lane] = P.row_nchunks[f];
            rp0[lane] = P.row_pofs0[f];
        }
        int incl = v;  // slots before each row: a wave prefix sum (nb <= 30 lanes)
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const int o = __shfl_up(incl, d);
            if (lane >= d) incl += o;
        }
        if (lane < nb) srow[lane] = incl - v;
        if (lane == nb - 1) srow[nb] = incl;
        if (lane == 0) bad = 0;
    }
    __syncthreads();
    const int nslots = srow[nb];
    // blocks in six row tasks (all chunk loads of a task in flight at once)
    // where they fit one pass, whole blocks otherwise
    const bool rows6 = 6 * nslots + nb <= kSmallThreads;
    const int ntask = rows6 ? 6 * nslots : nslots;
    if (tid < ntask) {
        if (rows6) sum_block_part<6>(P, spart, Hc, Lb, srow, rcol, rns, rnch, rp0, nb, fc0, lambda, tid);
        else sum_block_part<36>(P, spart, Hc, Lb, srow, rcol, rns, rnch, rp0, nb, fc0, lambda, tid);
    } else if (tid < ntask + nb) {
        // right-hand side of row lf: g = bp - coefficient, the coefficient of the
        // row's chunks in four partial sums (chunks congruent mod 4), as the reduction
        const int lf = tid - ntask, f = fc0 + lf;
        const int nch = rnch[lf];
        const long long stride = 36LL * rns[lf] + 6;
        const double* cq = spart + rp0[lf] + 36 * rns[lf];
        double bp[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) bp[e] = Hc[28 * (size_t)f + 21 + e];
        double cp[4][6];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 6; ++e) cp[q][e] = 0.0;
        for (int c0 = 0; c0 < nch; c0 += 4) {
            double w[4][6];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double* p = cq + min(c0 + q, nch - 1) * stride;
#pragma unroll
                for (int e = 0; e < 6; ++e) w[q][e] = p[e];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c0 + q < nch)
#pragma unroll
                    for (int e = 0; e < 6; ++e) cp[q][e] += w[q][e];
        }
#pragma unroll
        for (int e = 0; e < 6; ++e) yv[6 * lf + e] = bp[e] - ((cp[0][e] + cp[1][e]) + (cp[2][e] + cp[3][e]));
    }
    __syncthreads();
    // ---- factorisation, right-looking by block columns of six.  Wave 0 keeps
    // the chain: it updates block column C = K+1 with panel K and factors it
    // in registers (win_panel), while waves 1..7 apply panel K to the block
    // columns beyond C; one barrier per block column.
    if (wid == 0 && !win_panel(Lb, yv, dinv, n, 0, -1, lane) && lane == 0) bad = 1;
    __syncthreads();
    for (int K = 0; K + 1 < nb; ++K) {
        if (wid == 0) {
            if (!win_panel(Lb, yv, dinv, n, K + 1, K, lane) && lane == 0) bad = 1;
        } else {
            // A_IJ -= L_IK L_JK^T for K+2 <= J <= I by waves 1..7: one 6x6 block
            // per thread (36 register accumulators) while the blocks outnumber
            // the threads' sixths, else one block row per thread, so that the
            // last columns' few blocks do not hold the barrier for a whole
            // block's 216 products.  Either way each entry subtracts its six
            // products in the same order.
            const int rem = nb - K - 2;
            const int nupd = rem * (rem + 1) / 2;
            if (6 * nupd > kSmallThreads - 64) {
                for (int q = tid - 64; q < nupd; q += kSmallThreads - 64) {
                    // q -> (I, J), column-major from the diagonal
                    int jj = 0, qq = q, len = rem;
                    while (qq >= len) { qq -= len; ++jj; --len; }
                    const int J = K + 2 + jj, I = J + qq;
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
                }
            } else if (tid - 64 < 6 * nupd) {
                const int q = (tid - 64) / 6, r = (tid - 64) - 6 * ((tid - 64) / 6);
                int jj = 0, qq = q, len = rem;
                while (qq >= len) { qq -= len; ++jj; --len; }
                const int J = K + 2 + jj, I = J + qq;
                const double* A = Lb + blk_off(I, K) + 6 * r;
                const double* B = Lb + blk_off(J, K);
                double* C = Lb + blk_off(I, J) + 6 * r;
                double ar[6], acc[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) ar[m] = A[m];
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[c] = C[c];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    double v = acc[c];
#pragma unroll
                    for (int m = 0; m < 6; ++m) v -= ar[m] * B[6 * c + m];
                    acc[c] = v;
                }
#pragma unroll
                for (int c = 0; c < 6; ++c) C[c] = acc[c];
            }
        }
        __syncthreads();
    }
    if (bad) {
        for (int i = tid; i < n; i += kSmallThreads) dp[g0 + i] = 0.0;
        if (tid == 0) lm[win].fail = 1;
    } else if (wid == 0) {
        // backward, by wave 0 alone: x_K = L_KK^-T (y_K - sum_{I>K} L_IK^T x_I),
        // the 6x6 solve in lane 0; y holds the updated right-hand sides, then the step
        for (int K = nb - 1; K >= 0; --K) {
            if (lane == 0) {
                const double* L = Lb + blk_off(K, K);
                double x[6];
#pragma unroll
                for (int i = 5; i >= 0; --i) {
                    double v = yv[6 * K + i];
#pragma unroll
                    for (int k = i + 1; k < 6; ++k) v -= L[6 * k + i] * x[k];
                    x[i] = v * dinv[6 * K + i];
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    xk[i] = x[i];
                    yv[6 * K + i] = x[i];
                    dp[g0 + 6 * K + i] = x[i];
                }
            }
            wave_sync();
            for (int t = lane; t < 6 * K; t += 64) {
                const int J = t / 6, c = t - 6 * (t / 6);
                const double* L = Lb + blk_off(K, J);
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < 6; ++r) v += L[6 * r + c] * xk[r];
                yv[t] -= v;
            }
            wave_sync();
        }
    }
    }  // nb > 0
    __syncthreads();
    // candidate poses of the window's cameras from the step in LDS, and their
    // part of x'(lambda x + b): the camera update (k_update_cam), with the same
    // threads and the same sum
    const int fail = (nb > 0 && bad) ? 1 : fail0;
    double sc = 0.0;
    if (tid < 256)
        for (int c = P.win_cam0[win] + tid; c < P.win_cam0[win + 1]; c += 256)
            sc += camera_candidate(camera_inputs(P, pose, Hc, c), lambda, fail, c,
                                   [&](int f, int i) { return yv[6 * (f - fc0) + i]; }, pn, cn);
    const double s = block_sum<double, 256>(sc, ssum);
    if (tid == 0) red_scale[win] = s;
}

size_t cholesky_small_smem(const DevProblem& P) {
    const size_t nb = (size_t)P.max_win_fc;
    return sizeof(double) * (nb * (nb + 1) / 2 * kBS + 12 * nb);
}

// One workgroup per window, largest windows first: with the largest window's
// LDS a CU may hold one workgroup, and the small windows then fill in behind
// the large ones (P.win_chol_order).
void launch_cholesky_small(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_chol_small, dim3(P.n_win), dim3(kSmallThreads), cholesky_small_smem(P), st, P,
                       (const double*)W.spart, (const double*)W.Hc, W.dp, W.lm, (const PoseQT*)cur.pose, cand.pose,
                       cand.cam, W.red_scale);
}

int prepare_cholesky_small(const DevProblem& P) {
    const size_t smem = cholesky_small_smem(P);
    if (smem > 160 * 1024) return -1;
    return hipFuncSetAttribute((const void*)k_chol_small, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) ==
                   hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- one mid-size problem, on request
// One workgroup (512 threads) factors the whole system in 6x6 camera blocks,
// the per-window algorithm with the matrix in global memory (it stays
// in the L2 of the CU's XCD: 490 KB at 58 free cameras).  The lower block
// triangle is packed column by column (block (I, J), J <= I, at
// J nb - J(J-1)/2 + (I - J)) and entry-major (entry e of block b at
// e * nblk + b), so the threads of a trailing update, which own consecutive
// blocks, read and write consecutive addresses.  The panel column of the
// current step is staged in LDS.  One CU runs the whole chain of nb block
// steps through L2 (3.5 us per step at 11 free cameras, against 530 us in all
// at 58 and 349 us for the tiles there), so the per-window kernel in LDS
// replaced it for small systems; it stays as a switch for tests and sizes
// between 31 and 128 free cameras.
__device__ __forceinline__ int colstart(int J, int nb) { return J * nb - J * (J - 1) / 2; }

// (I, J) of packed block b; J found from the column starts
__device__ __forceinline__ void block_ij(int b, int nb, int& I, int& J) {
    const double B = 2.0 * nb + 1.0;
    int j = (int)((B - sqrt(B * B - 8.0 * b)) * 0.5);
    if (j < 0) j = 0;
    while (j > 0 && colstart(j, nb) > b) --j;
    while (j + 1 < nb && colstart(j + 1, nb) <= b) ++j;
    J = j;
    I = j + (b - colstart(j, nb));
}

__global__ __launch_bounds__(kSmallThreads) void k_chol_blocks(const double* __restrict__ S, int lda, int nb,
                                                               const double* __restrict__ g,
                                                               double* __restrict__ Lg, double* __restrict__ dp,
                                                               LMDev* lm) {
    extern __shared__ double sm[];
    __shared__ double xk[6];
    __shared__ int bad;
    if (lm->phase == kPhaseDone) return;
    const int n = 6 * nb;
    const int nblk = nb * (nb + 1) / 2;
    double* Pn = sm;                                      // (nb - 1) * kBS: panel column L_IK, I > K
    double* Mb = sm + (size_t)(nb > 1 ? nb - 1 : 1) * kBS;  // nb * kBS: M_K = L_KK^-1
    double* yv = Mb + (size_t)nb * kBS;                   // n
    double* dg = yv + n;                                  // kBS: diagonal block being factored
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nblk * 36; idx += kSmallThreads) {
        const int b = idx / 36, e = idx - 36 * (idx / 36);
        int I, J;
        block_ij(b, nb, I, J);
        const int r = e / 6, c = e - 6 * (e / 6);
        Lg[(size_t)e * nblk + b] = S[(size_t)(6 * I + r) * lda + 6 * J + c];
    }
    for (int i = tid; i < n; i += kSmallThreads) yv[i] = g[i];
    if (tid == 0) bad = 0;
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int e = 0; e < 36; ++e) dg[e] = Lg[(size_t)e * nblk];
        if (!diag_factor(dg, Mb, yv)) bad = 1;
#pragma unroll
        for (int e = 0; e < 36; ++e) Lg[(size_t)e * nblk] = dg[e];
    }
    __syncthreads();
    for (int K = 0; K < nb; ++K) {
        // panel: L_IK = A_IK M_K^T into Lg and LDS, y_I -= L_IK y_K
        const int rows = 6 * (nb - K - 1);
        const double* M = Mb + K * kBS;
        const int cK = colstart(K, nb);
        for (int t = tid; t < rows; t += kSmallThreads) {
            const int I = K + 1 + t / 6, r = t - 6 * (t / 6);
            const int bIK = cK + (I - K);
            double a[6], o[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) a[c] = Lg[(size_t)(6 * r + c) * nblk + bIK];
            double yr = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double v = 0.0;
#pragma unroll
                for (int m = 0; m <= c; ++m) v += a[m] * M[6 * c + m];
                o[c] = v;
                yr += v * yv[6 * K + c];
            }
            double* prow = Pn + (I - K - 1) * kBS + 6 * r;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                Lg[(size_t)(6 * r + c) * nblk + bIK] = o[c];
                prow[c] = o[c];
            }
            yv[6 * I + r] -= yr;
        }
        __syncthreads();
        // trailing update: the trailing blocks are the packed tail from column K+1
        const int rem = nb - K - 1;
        const int nupd = rem * (rem + 1) / 2;
        const int base = K + 1 < nb ? colstart(K + 1, nb) : 0;
        for (int q = tid; q < nupd; q += kSmallThreads) {
            int I, J;
            block_ij(base + q, nb, I, J);
            const double* A = Pn + (I - K - 1) * kBS;
            const double* B = Pn + (J - K - 1) * kBS;
            double acc[36];
#pragma unroll
            for (int e = 0; e < 36; ++e) acc[e] = Lg[(size_t)e * nblk + base + q];
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
            if (q == 0) {  // block (K+1, K+1): factor it now, through LDS
#pragma unroll
                for (int e = 0; e < 36; ++e) dg[e] = acc[e];
            } else {
#pragma unroll
                for (int e = 0; e < 36; ++e) Lg[(size_t)e * nblk + base + q] = acc[e];
            }
        }
        if (tid == 0 && nupd > 0) {
            const int J = K + 1;
            double yk[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) yk[i] = yv[6 * J + i];
            if (!diag_factor(dg, Mb + J * kBS, yk)) bad = 1;
#pragma unroll
            for (int i = 0; i < 6; ++i) yv[6 * J + i] = yk[i];
#pragma unroll
            for (int e = 0; e < 36; ++e) Lg[(size_t)e * nblk + base] = dg[e];
        }
        __syncthreads();
    }
    if (bad) {
        for (int i = tid; i < n; i += kSmallThreads) dp[i] = 0.0;
        if (tid == 0) lm->fail = 1;
        return;
    }
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
            dp[6 * K + tid] = xk[tid];
        }
        for (int t = tid; t < 6 * K; t += kSmallThreads) {
            const int J = t / 6, c = t - 6 * (t / 6);
            const int bKJ = colstart(J, nb) + (K - J);
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r) v += Lg[(size_t)(6 * r + c) * nblk + bKJ] * xk[r];
            yv[t] -= v;
        }
        __syncthreads();
    }
}

size_t cholesky_blocks_smem(int nb) {
    return sizeof(double) * ((size_t)(nb > 1 ? nb - 1 : 1) * kBS + (size_t)nb * kBS + 6 * (size_t)nb + kBS);
}

int prepare_cholesky(const DevProblem& P) {
    if (!P.use_blocks) {
        // the tiled path's backward step holds the right-hand sides in LDS
        const size_t smem = sizeof(double) * (size_t)std::max(P.n_pad, 64);
        if (smem + 9 * 1024 > 160 * 1024) return -1;
        const void* kerns[] = {(const void*)k_chol_back<1>, (const void*)k_chol_back<2>, (const void*)k_chol_back<3>,
                               (const void*)k_chol_back<4>, (const void*)k_chol_back<kBackPassMax>};
        for (const void* kf : kerns)
            if (hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) != hipSuccess) return -1;
        return 0;
    }
    const size_t smem = cholesky_blocks_smem(P.n_fc);
    if (smem > 160 * 1024) return -1;
    return hipFuncSetAttribute((const void*)k_chol_blocks, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) ==
                   hipSuccess ? 0 : -1;
}

void launch_cholesky_solve(const DevProblem& P, const DevState& cur, const DevState& cand, const DevWork& W,
                           hipStream_t st) {
    if (P.n_fc == 0) return;
    const int nt = P.nt, lda = P.n_pad;
    if (P.use_blocks) {
        hipLaunchKernelGGL(k_chol_blocks, dim3(1), dim3(kSmallThreads), cholesky_blocks_smem(P.n_fc), st,
                           (const double*)W.S, lda, P.n_fc, (const double*)W.g, W.Lm, W.dp, W.lm);
        return;
    }
    hipLaunchKernelGGL(k_chol_first, dim3(1), dim3(256), 0, st, (const double*)W.S, W.Lm, W.Li, lda, P.n,
                       (const double*)W.g, W.y, W.lm);
    for (int k = 0; k + 1 < nt; ++k) {
        const int rem = nt - k - 1;
        hipLaunchKernelGGL(k_chol_step, dim3(rem * (rem + 1) / 2), dim3(256), 0, st, W.S, W.Lm, W.Li, lda, P.n, k,
                           (const double*)W.g, W.y, W.fpart, nt, W.lm);
    }
    if (nt <= kBackSingleMax) {
        const int np = std::max(1, (64 * (nt - 1) + 127) / 128);
        auto kern = np == 1 ? k_chol_back<1> : np == 2 ? k_chol_back<2> : np == 3 ? k_chol_back<3>
                  : np == 4 ? k_chol_back<4> : k_chol_back<kBackPassMax>;
        hipLaunchKernelGGL(kern, dim3(1), dim3(kBackThreads), sizeof(double) * (size_t)lda, st, P,
                           (const double*)W.Lm, (const double*)W.Li, lda, nt, (const double*)W.y, W.dp,
                           (const LMDev*)W.lm, (const PoseQT*)cur.pose, (const double*)W.Hc, cand.pose, cand.cam,
                           W.red_scale);
    } else {
        for (int k = nt - 1; k >= 0; --k)
            hipLaunchKernelGGL(k_chol_back_tile, dim3(k + 1), dim3(256), 0, st, (const double*)W.Lm,
                               (const double*)W.Li, lda, nt, k, W.y, W.dp, (const LMDev*)W.lm);
    }
}

}  // namespace navba
