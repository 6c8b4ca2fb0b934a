// pcg_kernels.hip — block-Jacobi preconditioned conjugate gradients on the
// reduced camera system of one problem: S dp = g, S symmetric and positive
// definite (Hpp + lambda I minus the Schur terms; both triangles written by
// the Schur reduction in this mode), the preconditioner the inverse of each
// 6x6 camera block of S.  Stops at ||r|| <= tol ||g|| or after max_iter
// iterations, and uses the last iterate as the step.  This is the solver for
// systems beyond the dense factorisation (more than kMaxDenseN unknowns), and
// on request for any size.
//
// Vectors are in free-camera order, 6 per camera.  Every sum has a fixed
// order (per-thread partials, DPP wave sums, per-workgroup partials added in
// index order by one workgroup), so a solve gives the same bits every run.
// One iteration is five launches: matvec (q = S p, and p.q per workgroup),
// alpha, update (x, r, z = Dinv r, and r.z, r.r per workgroup), beta (stop
// test), direction (p = z + beta p).  The host launches them in chunks and
// reads the stop flag between chunks.
#include <algorithm>

#include "ba_kernels.h"

namespace navba {

constexpr int kPcgThreads = 256;
constexpr int kPcgRows = kPcgThreads / kWave;  // matvec rows per workgroup (one wave each)

// Dinv = S_ff^-1 for the 6x6 block at S (row stride lda), from S_ff = L L^T:
// Dinv = L^-T L^-1.  false when a pivot is not positive and finite.
__device__ bool block_inverse6(const double* S, size_t lda, double Dinv[36]) {
    double l[6][6];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = S[(size_t)j * lda + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
        ok = ok && d > 0.0 && isfinite(d);
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        l[j][j] = ljj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double v = S[(size_t)i * lda + j];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= l[i][k] * l[j][k];
            l[i][j] = v / ljj;
        }
    }
    // M = L^-1 (lower), column by column
    double m[6][6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
#pragma unroll
        for (int i = 0; i < 6; ++i) m[i][j] = 0.0;
        m[j][j] = 1.0 / l[j][j];
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double v = 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) v += l[i][k] * m[k][j];
            m[i][j] = -v / l[i][i];
        }
    }
    // Dinv = M^T M
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) v += m[k][i] * m[k][j];
            Dinv[6 * i + j] = v;
        }
    return ok;
}

// z_f = Dinv_f r_f, and this block's r.z and r.r
__device__ __forceinline__ void precondition(const double* Dinv, const double r[6], double z[6], double& rz,
                                             double& rr) {
    rz = 0.0;
    rr = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) v += Dinv[6 * i + j] * r[j];
        z[i] = v;
        rz += r[i] * v;
        rr += r[i] * r[i];
    }
}

// Sum of part[0..n) by one workgroup in a fixed order; valid in thread 0.
__device__ __forceinline__ double sum_parts(const double* part, int n, double* sh) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) s += part[i];
    return block_sum<double, kPcgThreads>(s, sh);
}

__global__ void k_pcg_reset(PcgState* st) {
    if (threadIdx.x == 0) {
        st->fail = 0;
        st->done = 0;
        st->iter = -1;  // stays -1 when the window has finished (no solve)
    }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_start(const double* __restrict__ S, int lda, int nb,
                                                           const double* __restrict__ g, double* __restrict__ Dinv,
                                                           double* __restrict__ x, double* __restrict__ r,
                                                           double* __restrict__ z, double* __restrict__ p,
                                                           double* __restrict__ part_rz,
                                                           double* __restrict__ part_rr, PcgState* st,
                                                           const LMDev* lm) {
    __shared__ double sh[kPcgThreads / kWave];
    if (!running(lm, 0)) return;
    const int f = blockIdx.x * kPcgThreads + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (f < nb) {
        double D[36];
        if (!block_inverse6(S + (size_t)(6 * f) * lda + 6 * f, (size_t)lda, D)) st->fail = 1;
        double rf[6], zf[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) rf[i] = g[6 * f + i];
        precondition(D, rf, zf, rz, rr);
#pragma unroll
        for (int k = 0; k < 36; ++k) Dinv[36 * (size_t)f + k] = D[k];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            x[6 * f + i] = 0.0;
            r[6 * f + i] = rf[i];
            z[6 * f + i] = zf[i];
            p[6 * f + i] = zf[i];
        }
    }
    const double srz = block_sum<double, kPcgThreads>(rz, sh);
    __syncthreads();
    const double srr = block_sum<double, kPcgThreads>(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = srz;
        part_rr[blockIdx.x] = srr;
    }
}

// After the start: r.z, ||g||^2, and the stop flag (a finished window, a
// zero right-hand side or a block that is not positive definite).
__global__ __launch_bounds__(kPcgThreads) void k_pcg_begin(const double* __restrict__ part_rz,
                                                           const double* __restrict__ part_rr, int nparts,
                                                           PcgState* st, const LMDev* lm) {
    __shared__ double sh[kPcgThreads / kWave];
    if (!running(lm, 0)) {
        if (threadIdx.x == 0) st->done = 1;
        return;
    }
    const double rz = sum_parts(part_rz, nparts, sh);
    __syncthreads();
    const double rr = sum_parts(part_rr, nparts, sh);
    if (threadIdx.x == 0) {
        st->rz = rz;
        st->rr0 = rr;
        st->rr = rr;
        st->iter = 0;
        st->done = (st->fail || !(rr > 0.0)) ? 1 : 0;
    }
}

// q = S p over full rows (one wave per row, four partial sums per lane), and
// this workgroup's part of p.q
__global__ __launch_bounds__(kPcgThreads) void k_pcg_matvec(const double* __restrict__ S, int lda, int n,
                                                            const double* __restrict__ p, double* __restrict__ q,
                                                            double* __restrict__ part_pq, const PcgState* st) {
    __shared__ double rows[kPcgRows];
    if (st->done) return;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int i = blockIdx.x * kPcgRows + wid;
    double v = 0.0;
    if (i < n) {
        const double* row = S + (size_t)i * lda;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int j = lane;
        for (; j + 3 * kWave < n; j += 4 * kWave) {
            s0 += row[j] * p[j];
            s1 += row[j + kWave] * p[j + kWave];
            s2 += row[j + 2 * kWave] * p[j + 2 * kWave];
            s3 += row[j + 3 * kWave] * p[j + 3 * kWave];
        }
        for (; j < n; j += kWave) s0 += row[j] * p[j];
        v = wave_sum((s0 + s1) + (s2 + s3));
        if (lane == 0) q[i] = v;
    }
    if (lane == 0) rows[wid] = i < n ? p[i] * v : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kPcgRows; ++k) s += rows[k];
        part_pq[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_alpha(const double* __restrict__ part_pq, int nparts,
                                                           PcgState* st) {
    __shared__ double sh[kPcgThreads / kWave];
    if (st->done) return;
    const double pq = sum_parts(part_pq, nparts, sh);
    if (threadIdx.x == 0) {
        if (!(pq > 0.0) || !isfinite(pq)) {
            st->fail = 1;
            st->done = 1;
        } else {
            st->alpha = st->rz / pq;
        }
    }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_update(int nb, const double* __restrict__ Dinv,
                                                            double* __restrict__ x, double* __restrict__ r,
                                                            double* __restrict__ z, const double* __restrict__ p,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ part_rz,
                                                            double* __restrict__ part_rr, const PcgState* st) {
    __shared__ double sh[kPcgThreads / kWave];
    if (st->done) return;
    const double alpha = st->alpha;
    const int f = blockIdx.x * kPcgThreads + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (f < nb) {
        double rf[6], zf[6], D[36];
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = Dinv[36 * (size_t)f + k];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            x[6 * f + i] += alpha * p[6 * f + i];
            rf[i] = r[6 * f + i] - alpha * q[6 * f + i];
            r[6 * f + i] = rf[i];
        }
        precondition(D, rf, zf, rz, rr);
#pragma unroll
        for (int i = 0; i < 6; ++i) z[6 * f + i] = zf[i];
    }
    const double srz = block_sum<double, kPcgThreads>(rz, sh);
    __syncthreads();
    const double srr = block_sum<double, kPcgThreads>(rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = srz;
        part_rr[blockIdx.x] = srr;
    }
}

// The stop test ||r||^2 <= tol^2 ||g||^2 or max_iter, else beta = r'z' / r.z
__global__ __launch_bounds__(kPcgThreads) void k_pcg_beta(const double* __restrict__ part_rz,
                                                          const double* __restrict__ part_rr, int nparts,
                                                          double tol2, int max_iter, PcgState* st) {
    __shared__ double sh[kPcgThreads / kWave];
    if (st->done) return;
    const double rz = sum_parts(part_rz, nparts, sh);
    __syncthreads();
    const double rr = sum_parts(part_rr, nparts, sh);
    if (threadIdx.x == 0) {
        st->iter += 1;
        st->rr = rr;
        if (rr <= tol2 * st->rr0 || st->iter >= max_iter) {
            st->done = 1;
        } else {
            st->beta = rz / st->rz;
            st->rz = rz;
        }
    }
}

__global__ __launch_bounds__(kPcgThreads) void k_pcg_direction(int n, const double* __restrict__ z,
                                                               double* __restrict__ p, const PcgState* st) {
    if (st->done) return;
    const double beta = st->beta;
    const int i = blockIdx.x * kPcgThreads + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// The step: the last iterate, or zero and the window's failure flag
__global__ __launch_bounds__(kPcgThreads) void k_pcg_finish(int n, const double* __restrict__ x,
                                                            double* __restrict__ dp, const PcgState* st, LMDev* lm) {
    if (!running(lm, 0)) return;
    const bool fail = st->fail != 0;
    const int i = blockIdx.x * kPcgThreads + threadIdx.x;
    if (i < n) dp[i] = fail ? 0.0 : x[i];
    if (fail && i == 0) lm[0].fail = 1;
}

namespace {
int parts_of(int count) { return (count + kPcgThreads - 1) / kPcgThreads; }
}  // namespace

size_t pcg_work_doubles(int n) {
    const int nb = n / 6;
    const size_t np = (size_t)std::max(parts_of(nb), (n + kPcgRows - 1) / kPcgRows);
    return 36 * (size_t)nb + 5 * (size_t)n + 3 * np;
}

void launch_pcg_start(const DevProblem& P, const DevWork& W, hipStream_t st) {
    const int n = P.n, nb = n / 6;
    double* x = W.pcg + 36 * (size_t)nb;
    double* r = x + n;
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    double* part = q + n;
    const size_t np = (size_t)std::max(parts_of(nb), (n + kPcgRows - 1) / kPcgRows);
    double* part_rz = part;
    double* part_rr = part + np;
    hipLaunchKernelGGL(k_pcg_reset, dim3(1), dim3(64), 0, st, W.pcg_st);
    hipLaunchKernelGGL(k_pcg_start, dim3(parts_of(nb)), dim3(kPcgThreads), 0, st, (const double*)W.S, P.n_pad, nb,
                       (const double*)W.g, W.pcg, x, r, z, p, part_rz, part_rr, W.pcg_st, (const LMDev*)W.lm);
    hipLaunchKernelGGL(k_pcg_begin, dim3(1), dim3(kPcgThreads), 0, st, (const double*)part_rz,
                       (const double*)part_rr, parts_of(nb), W.pcg_st, (const LMDev*)W.lm);
}

void launch_pcg_iterations(const DevProblem& P, const DevWork& W, int count, double tol, int max_iter,
                           hipStream_t st) {
    const int n = P.n, nb = n / 6;
    double* x = W.pcg + 36 * (size_t)nb;
    double* r = x + n;
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    double* part = q + n;
    const size_t np = (size_t)std::max(parts_of(nb), (n + kPcgRows - 1) / kPcgRows);
    double* part_rz = part;
    double* part_rr = part + np;
    double* part_pq = part + 2 * np;
    const int mv_blocks = (n + kPcgRows - 1) / kPcgRows;
    for (int k = 0; k < count; ++k) {
        hipLaunchKernelGGL(k_pcg_matvec, dim3(mv_blocks), dim3(kPcgThreads), 0, st, (const double*)W.S, P.n_pad, n,
                           (const double*)p, q, part_pq, (const PcgState*)W.pcg_st);
        hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(kPcgThreads), 0, st, (const double*)part_pq, mv_blocks,
                           W.pcg_st);
        hipLaunchKernelGGL(k_pcg_update, dim3(parts_of(nb)), dim3(kPcgThreads), 0, st, nb, (const double*)W.pcg, x,
                           r, z, (const double*)p, (const double*)q, part_rz, part_rr, (const PcgState*)W.pcg_st);
        hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(kPcgThreads), 0, st, (const double*)part_rz,
                           (const double*)part_rr, parts_of(nb), tol * tol, max_iter, W.pcg_st);
        hipLaunchKernelGGL(k_pcg_direction, dim3((n + kPcgThreads - 1) / kPcgThreads), dim3(kPcgThreads), 0, st, n,
                           (const double*)z, p, (const PcgState*)W.pcg_st);
    }
}

void launch_pcg_finish(const DevProblem& P, const DevWork& W, hipStream_t st) {
    const int n = P.n, nb = n / 6;
    const double* x = W.pcg + 36 * (size_t)nb;
    hipLaunchKernelGGL(k_pcg_finish, dim3((n + kPcgThreads - 1) / kPcgThreads), dim3(kPcgThreads), 0, st, n, x, W.dp,
                       (const PcgState*)W.pcg_st, W.lm);
}

}  // namespace navba
