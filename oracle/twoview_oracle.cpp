// twoview_oracle.cpp — CPU restatement of the scoring of two-view RANSAC
// hypotheses.
//
// TEST INFRASTRUCTURE ONLY: the checker of nav_tv_score_homography /
// nav_tv_score_fundamental (include/nav_twoview.h); the product library never
// links or calls it.
//
// Restates TwoViewReconstruction::CheckHomography
// (core/operators/mapInit/OP_2ViewReconstruction.cpp:447-528),
// ::CheckFundamental (:530-611) and the keep-the-best loop of FindHomography /
// FindFundamental (:290-314, :341-365): one sequential pass over the matches
// per hypothesis, float arithmetic in the order written, no fused
// multiply-adds (compiled with -ffp-contract=off; x86-64 float has no excess
// precision).
//
// The reference has no tests of its own for this path.  Pinned by what the
// geometry fixes (tests/test_twoview_oracle.py):
//   - H = I with x2 = x1: every chi is 0, every match an inlier, the score the
//     float sum of 2 N copies of th_score;
//   - a translation H with integer offsets: chi = |d|^2 / sigma^2 exactly, so
//     the inlier flags and the score follow from |d|^2 against th_score;
//   - a general projective H: chi equals the float64 transfer error computed
//     with matrix products (x2in1 = H12 x2 dehomogenised), which a transposed
//     matrix or swapped image fails;
//   - F from K, R, t of a synthetic two-view scene: exact correspondences give
//     near-zero point-to-epipolar-line distances, and perturbed ones the float64
//     distances |l . x| / |l_(0,1)| with l = F x1 and l = F^T x2;
//   - the keep-the-best rule: equal maxima keep the first, NaN scores and
//     scores of 0 never win, no winner leaves every flag false.
//
// Also restates TwoViewReconstruction::CheckRT (:1244-1352) with Triangulate
// (:1179-1194) for a list of motion candidates (R, t).  The reference takes
// the 3D point from OpenCV's float SVD of the 4x4 system (cv::SVD::compute)
// and forms P2 = K [R|t], O2 = -R^T t and R X + t with cv::gemm, and the norms
// and dot products with cv::norm / Mat::dot.  OpenCV is absent here, so the
// restatement fixes its own reading of those calls: matrix products and the
// norm and dot sums accumulate float products in double and round to float
// (the published behaviour of OpenCV's 32F gemm / norm / dot kernels), and the
// null vector comes from a one-sided (Hestenes) Jacobi SVD in double, rounded
// to float.  Parity against OpenCV's bits is unpinned; pinned instead
// (tests/test_twoview_oracle.py) by: the null vector against LAPACK's SVD
// (numpy) on random and rank-3 systems; exact correspondences triangulating
// to the generating points; the cheirality, reprojection and parallax
// decisions against a float64 evaluation of the same geometry away from the
// thresholds; the parallax against the float64 angle at the reference's index
// into the sorted cosines.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" {

// kp1, kp2: (u, v) float; matches: n pairs (i1, i2); H21, H12 row-major.
float oracle_tv_check_homography(const float* kp1, const float* kp2, const int32_t* matches, int32_t n,
                                 const float* H21, const float* H12, float sigma, float th, uint8_t* inliers) {
    const float h11 = H21[0], h12 = H21[1], h13 = H21[2], h21 = H21[3], h22 = H21[4], h23 = H21[5], h31 = H21[6],
                h32 = H21[7], h33 = H21[8];
    const float h11inv = H12[0], h12inv = H12[1], h13inv = H12[2], h21inv = H12[3], h22inv = H12[4],
                h23inv = H12[5], h31inv = H12[6], h32inv = H12[7], h33inv = H12[8];
    float score = 0;
    const float invSigmaSquare = 1.0f / (sigma * sigma);
    for (int32_t i = 0; i < n; ++i) {
        bool bIn = true;
        const float u1 = kp1[2 * matches[2 * i]], v1 = kp1[2 * matches[2 * i] + 1];
        const float u2 = kp2[2 * matches[2 * i + 1]], v2 = kp2[2 * matches[2 * i + 1] + 1];

        const float w2in1inv = 1.0f / (h31inv * u2 + h32inv * v2 + h33inv);
        const float u2in1 = (h11inv * u2 + h12inv * v2 + h13inv) * w2in1inv;
        const float v2in1 = (h21inv * u2 + h22inv * v2 + h23inv) * w2in1inv;
        const float squareDist1 = (u1 - u2in1) * (u1 - u2in1) + (v1 - v2in1) * (v1 - v2in1);
        const float chiSquare1 = squareDist1 * invSigmaSquare;
        if (chiSquare1 > th)
            bIn = false;
        else
            score += th - chiSquare1;

        const float w1in2inv = 1.0f / (h31 * u1 + h32 * v1 + h33);
        const float u1in2 = (h11 * u1 + h12 * v1 + h13) * w1in2inv;
        const float v1in2 = (h21 * u1 + h22 * v1 + h23) * w1in2inv;
        const float squareDist2 = (u2 - u1in2) * (u2 - u1in2) + (v2 - v1in2) * (v2 - v1in2);
        const float chiSquare2 = squareDist2 * invSigmaSquare;
        if (chiSquare2 > th)
            bIn = false;
        else
            score += th - chiSquare2;

        inliers[i] = bIn ? 1 : 0;
    }
    return score;
}

// F21 row-major; th = th_f (a term's threshold), th_score what a term adds below it.
float oracle_tv_check_fundamental(const float* kp1, const float* kp2, const int32_t* matches, int32_t n,
                                  const float* F21, float sigma, float th, float th_score, uint8_t* inliers) {
    const float f11 = F21[0], f12 = F21[1], f13 = F21[2], f21 = F21[3], f22 = F21[4], f23 = F21[5], f31 = F21[6],
                f32 = F21[7], f33 = F21[8];
    float score = 0;
    const float invSigmaSquare = 1.0f / (sigma * sigma);
    for (int32_t i = 0; i < n; ++i) {
        bool bIn = true;
        const float u1 = kp1[2 * matches[2 * i]], v1 = kp1[2 * matches[2 * i] + 1];
        const float u2 = kp2[2 * matches[2 * i + 1]], v2 = kp2[2 * matches[2 * i + 1] + 1];

        const float a2 = f11 * u1 + f12 * v1 + f13;  // l2 = F21 x1
        const float b2 = f21 * u1 + f22 * v1 + f23;
        const float c2 = f31 * u1 + f32 * v1 + f33;
        const float num2 = a2 * u2 + b2 * v2 + c2;
        const float squareDist1 = num2 * num2 / (a2 * a2 + b2 * b2);
        const float chiSquare1 = squareDist1 * invSigmaSquare;
        if (chiSquare1 > th)
            bIn = false;
        else
            score += th_score - chiSquare1;

        const float a1 = f11 * u2 + f21 * v2 + f31;  // l1 = x2^T F21
        const float b1 = f12 * u2 + f22 * v2 + f32;
        const float c1 = f13 * u2 + f23 * v2 + f33;
        const float num1 = a1 * u1 + b1 * v1 + c1;
        const float squareDist2 = num1 * num1 / (a1 * a1 + b1 * b1);
        const float chiSquare2 = squareDist2 * invSigmaSquare;
        if (chiSquare2 > th)
            bIn = false;
        else
            score += th_score - chiSquare2;

        inliers[i] = bIn ? 1 : 0;
    }
    return score;
}

// All n_hyp hypotheses in turn and the reference's keep-the-best rule.
// homography != 0: M0 = H21s, M1 = H12s (9 floats each per hypothesis), th = th_score;
// otherwise M0 = F21s, th = th_f.  scores: n_hyp; inliers: n (the kept one's, all 0
// if none); *best = -1 and *best_score = 0 if no score is above 0.
void oracle_tv_find(int32_t homography, const float* kp1, const float* kp2, const int32_t* matches, int32_t n,
                    const float* M0, const float* M1, int32_t n_hyp, float sigma, float th, float th_score,
                    float* scores, int32_t* best, float* best_score, uint8_t* inliers, uint8_t* work) {
    float score = 0.0f;
    *best = -1;
    for (int32_t i = 0; i < n; ++i) inliers[i] = 0;
    for (int32_t it = 0; it < n_hyp; ++it) {
        const float cur = homography
                              ? oracle_tv_check_homography(kp1, kp2, matches, n, M0 + 9 * it, M1 + 9 * it, sigma,
                                                           th_score, work)
                              : oracle_tv_check_fundamental(kp1, kp2, matches, n, M0 + 9 * it, sigma, th, th_score,
                                                            work);
        scores[it] = cur;
        if (cur > score) {
            *best = it;
            for (int32_t i = 0; i < n; ++i) inliers[i] = work[i];
            score = cur;
        }
    }
    *best_score = score;
}

// Right singular vector of the smallest singular value of the 4x4 row-major A:
// one-sided Jacobi rotations on A's columns (pairs (0,1) (0,2) (0,3) (1,2)
// (1,3) (2,3) per sweep, up to 30 sweeps, a pair skipped once its columns are
// orthogonal to 1e-15 of their norms), V accumulating the rotations; the
// column of V whose rotated A column has the smallest norm (the first of
// equal ones).  Double, left to right, no fused multiply-adds.
void oracle_tv_null_vector4(const double* A, double* x) {
    double a[16], v[16];
    for (int i = 0; i < 16; ++i) {
        a[i] = A[i];
        v[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int r = 0; r < 4; ++r) {
                    al = al + a[4 * r + p] * a[4 * r + p];
                    be = be + a[4 * r + q] * a[4 * r + q];
                    ga = ga + a[4 * r + p] * a[4 * r + q];
                }
                if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < 4; ++r) {
                    const double ap = a[4 * r + p], aq = a[4 * r + q];
                    a[4 * r + p] = c * ap - s * aq;
                    a[4 * r + q] = s * ap + c * aq;
                    const double vp = v[4 * r + p], vq = v[4 * r + q];
                    v[4 * r + p] = c * vp - s * vq;
                    v[4 * r + q] = s * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    int k = 0;
    double best = 0.0;
    for (int c = 0; c < 4; ++c) {
        double nrm = 0.0;
        for (int r = 0; r < 4; ++r) nrm = nrm + a[4 * r + c] * a[4 * r + c];
        if (c == 0 || nrm < best) {
            best = nrm;
            k = c;
        }
    }
    for (int r = 0; r < 4; ++r) x[r] = v[4 * r + k];
}

// 3x3 (or 3x4) float product with double accumulation, rounded to float.
static float dot3d(const float* a, const float* b, int sb) {
    double r = 0.0;
    for (int k = 0; k < 3; ++k) r = r + (double)a[k] * (double)b[sb * k];
    return (float)r;
}

// Triangulate (:1179-1194): A rows u1 P1_2 - P1_0, v1 P1_2 - P1_1, u2 P2_2 -
// P2_0, v2 P2_2 - P2_1 in float; x = null vector (float); X = x_0..2 / x_3.
void oracle_tv_triangulate(float u1, float v1, float u2, float v2, const float* P1, const float* P2, float* X) {
    double A[16];
    for (int c = 0; c < 4; ++c) {
        A[c] = (double)(u1 * P1[8 + c] - P1[c]);
        A[4 + c] = (double)(v1 * P1[8 + c] - P1[4 + c]);
        A[8 + c] = (double)(u2 * P2[8 + c] - P2[c]);
        A[12 + c] = (double)(v2 * P2[8 + c] - P2[4 + c]);
    }
    double x[4];
    oracle_tv_null_vector4(A, x);
    const float w = (float)x[3];
    for (int r = 0; r < 3; ++r) X[r] = (float)x[r] / w;
}

// CheckRT for n_cand candidates (R: 9 floats, t: 3 floats each, row-major).
// K: 3x3 row-major float.  inliers: per match.  Outputs per candidate c:
// n_good[c], parallax[c] (degrees; 0 if none), cos_sel[c] (the cosine at the
// reference's index, NaN if none), p3d[c][n_kp1][3] (0 where not set),
// good[c][n_kp1] (vbGood), and per match cosp[c][n] / passed[c][n].
void oracle_tv_check_rt(const float* kp1, const float* kp2, const int32_t* matches, int32_t n,
                        const uint8_t* inliers, const float* K, const float* R, const float* t, int32_t n_cand,
                        float th2, float min_cos, int32_t min_tri, int32_t n_kp1, int32_t* n_good, float* parallax,
                        float* cos_sel, float* p3d, uint8_t* good, float* cosp, uint8_t* passed) {
    const float fx = K[0], fy = K[4], cx = K[2], cy = K[5];
    for (int32_t c = 0; c < n_cand; ++c) {
        const float* Rc = R + 9 * c;
        const float* tc = t + 3 * c;
        float P1[12] = {0}, P2[12], Rt[12];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) P1[4 * r + q] = K[3 * r + q];
        for (int r = 0; r < 3; ++r) {
            for (int q = 0; q < 3; ++q) Rt[4 * r + q] = Rc[3 * r + q];
            Rt[4 * r + 3] = tc[r];
        }
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 4; ++q) P2[4 * r + q] = dot3d(K + 3 * r, Rt + q, 4);  // P2 = K [R|t]
        float O2[3];  // -(R^T t): row r of R^T is column r of R
        for (int r = 0; r < 3; ++r) {
            const float col[3] = {Rc[r], Rc[3 + r], Rc[6 + r]};
            O2[r] = -dot3d(col, tc, 1);
        }
        float* P3 = p3d + (size_t)3 * n_kp1 * c;
        uint8_t* G = good + (size_t)n_kp1 * c;
        for (int32_t i = 0; i < 3 * n_kp1; ++i) P3[i] = 0.0f;
        for (int32_t i = 0; i < n_kp1; ++i) G[i] = 0;
        std::vector<float> vcos;
        int32_t ng = 0;
        for (int32_t i = 0; i < n; ++i) {
            cosp[(size_t)n * c + i] = 0.0f;
            passed[(size_t)n * c + i] = 0;
            if (!inliers[i]) continue;
            const int32_t i1 = matches[2 * i], i2 = matches[2 * i + 1];
            const float u1 = kp1[2 * i1], v1 = kp1[2 * i1 + 1], u2 = kp2[2 * i2], v2 = kp2[2 * i2 + 1];
            float X[3];
            oracle_tv_triangulate(u1, v1, u2, v2, P1, P2, X);
            if (!std::isfinite(X[0]) || !std::isfinite(X[1]) || !std::isfinite(X[2])) continue;
            const float n2[3] = {X[0] - O2[0], X[1] - O2[1], X[2] - O2[2]};
            const float d1 = (float)std::sqrt(((0.0 + (double)X[0] * X[0]) + (double)X[1] * X[1]) + (double)X[2] * X[2]);
            const float d2 =
                (float)std::sqrt(((0.0 + (double)n2[0] * n2[0]) + (double)n2[1] * n2[1]) + (double)n2[2] * n2[2]);
            const double dt = ((0.0 + (double)X[0] * n2[0]) + (double)X[1] * n2[1]) + (double)X[2] * n2[2];
            const float cosP = (float)(dt / (double)(d1 * d2));
            if (X[2] <= 0 && cosP < min_cos) continue;
            float X2[3];
            for (int r = 0; r < 3; ++r) X2[r] = dot3d(Rc + 3 * r, X, 1) + tc[r];
            if (X2[2] <= 0 && cosP < min_cos) continue;
            const float iz1 = 1.0f / X[2];
            const float im1x = fx * X[0] * iz1 + cx, im1y = fy * X[1] * iz1 + cy;
            const float e1 = (im1x - u1) * (im1x - u1) + (im1y - v1) * (im1y - v1);
            if (e1 > th2) continue;
            const float iz2 = 1.0f / X2[2];
            const float im2x = fx * X2[0] * iz2 + cx, im2y = fy * X2[1] * iz2 + cy;
            const float e2 = (im2x - u2) * (im2x - u2) + (im2y - v2) * (im2y - v2);
            if (e2 > th2) continue;
            vcos.push_back(cosP);
            P3[3 * i1] = X[0];
            P3[3 * i1 + 1] = X[1];
            P3[3 * i1 + 2] = X[2];
            ++ng;
            if (cosP < min_cos) G[i1] = 1;
            cosp[(size_t)n * c + i] = cosP;
            passed[(size_t)n * c + i] = 1;
        }
        n_good[c] = ng;
        if (ng > 0) {
            std::sort(vcos.begin(), vcos.end());
            const size_t idx = (size_t)std::min(min_tri, (int32_t)vcos.size() - 1);
            cos_sel[c] = vcos[idx];
            parallax[c] = static_cast<float>(std::acos(vcos[idx]) * 180 / 3.14159265358979323846);
        } else {
            cos_sel[c] = NAN;
            parallax[c] = 0.0f;
        }
    }
}

// ---- hypotheses of FindHomography / FindFundamental (:266-366)
//
// Normalize (:1195-1240) as the reference writes it: float sums in input
// order.  ComputeH21 (:368-407) / ComputeF21 (:410-446) take the right
// singular vector of the smallest singular value of A (2N x 9 or N x 9) from
// OpenCV's float SVD; here it is the eigenvector of the smallest eigenvalue
// of A^T A, by inverse iteration in double (A^T A + mu I, mu = 1e-13 trace,
// factored once by Cholesky, the diagonal's reciprocals formed once; 40 iterations from (1, ..., 1)), unit length,
// sign with the largest-magnitude entry positive, rounded to float.  F's rank-2
// step (SVD of Fpre, smallest singular value set to 0) is Fpre (I - v v^T)
// with v that eigenvector of Fpre^T Fpre.  H21 = T2^-1 Hn T1, H12 = H21^-1,
// F21 = T2^T Fn T1: products summed in double and rounded to float (OpenCV's
// 32F gemm), 3x3 inverses by cofactors over the determinant in double
// (OpenCV's small-matrix invert).  Pinned (tests/test_twoview_oracle.py)
// against LAPACK's SVD and noise-free scenes; unpinned against OpenCV's bits.

void oracle_tv_normalize(const float* kp, int32_t n, float* out, float* T) {
    float meanX = 0, meanY = 0;
    for (int32_t i = 0; i < n; ++i) {
        meanX += kp[2 * i];
        meanY += kp[2 * i + 1];
    }
    meanX = meanX / n;
    meanY = meanY / n;
    float meanDevX = 0, meanDevY = 0;
    for (int32_t i = 0; i < n; ++i) {
        out[2 * i] = kp[2 * i] - meanX;
        out[2 * i + 1] = kp[2 * i + 1] - meanY;
        meanDevX += std::fabs(out[2 * i]);
        meanDevY += std::fabs(out[2 * i + 1]);
    }
    meanDevX = meanDevX / n;
    meanDevY = meanDevY / n;
    const float sX = 1.0f / meanDevX, sY = 1.0f / meanDevY;
    for (int32_t i = 0; i < n; ++i) {
        out[2 * i] = out[2 * i] * sX;
        out[2 * i + 1] = out[2 * i + 1] * sY;
    }
    const float Tm[9] = {sX, 0.0f, -meanX * sX, 0.0f, sY, -meanY * sY, 0.0f, 0.0f, 1.0f};
    for (int k = 0; k < 9; ++k) T[k] = Tm[k];
}

// Smallest eigenvector of the symmetric positive semi-definite n x n M (row-major, n <= 9).
void oracle_tv_small_eigvec(const double* M, int n, double* v) {
    double L[81];
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr = tr + M[n * i + i];
    const double mu = 1e-13 * tr;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = M[n * i + j] + (i == j ? mu : 0.0);
            for (int k = 0; k < j; ++k) a = a - L[9 * i + k] * L[9 * j + k];
            if (i == j) {
                L[9 * i + i] = std::sqrt(a > 0.0 ? a : mu);
            } else {
                L[9 * i + j] = a / L[9 * j + j];
            }
        }
    double x[9], rd[9];
    for (int i = 0; i < n; ++i) {
        x[i] = 1.0;
        rd[i] = 1.0 / L[9 * i + i];  // the diagonal's reciprocals, formed once
    }
    for (int it = 0; it < 40; ++it) {
        double y[9];
        for (int i = 0; i < n; ++i) {  // L y = x
            double a = x[i];
            for (int k = 0; k < i; ++k) a = a - L[9 * i + k] * y[k];
            y[i] = a * rd[i];
        }
        for (int i = n - 1; i >= 0; --i) {  // L^T x = y
            double a = y[i];
            for (int k = i + 1; k < n; ++k) a = a - L[9 * k + i] * x[k];
            x[i] = a * rd[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm = nrm + x[i] * x[i];
        const double r = 1.0 / std::sqrt(nrm);
        for (int i = 0; i < n; ++i) x[i] = x[i] * r;
    }
    int imax = 0;
    for (int i = 1; i < n; ++i)
        if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    const double sg = x[imax] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) v[i] = x[i] * sg;
}

// C = A B for 3x3 float matrices, products summed in double, rounded to float.
static void mul3(const float* A, const float* B, float* C) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a = a + (double)A[3 * r + k] * (double)B[3 * k + c];
            C[3 * r + c] = (float)a;
        }
}

// 3x3 float inverse by cofactors in double; false if the determinant is 0.
void oracle_tv_inv3f(const float* S, float* D) {
    const double d0 = (double)S[4] * S[8] - (double)S[5] * S[7];
    const double d1 = (double)S[3] * S[8] - (double)S[5] * S[6];
    const double d2 = (double)S[3] * S[7] - (double)S[4] * S[6];
    double det = (double)S[0] * d0 - (double)S[1] * d1 + (double)S[2] * d2;
    if (det == 0.0) {
        for (int k = 0; k < 9; ++k) D[k] = 0.0f;
        return;
    }
    det = 1.0 / det;
    D[0] = (float)(d0 * det);
    D[1] = (float)(((double)S[2] * S[7] - (double)S[1] * S[8]) * det);
    D[2] = (float)(((double)S[1] * S[5] - (double)S[2] * S[4]) * det);
    D[3] = (float)(-d1 * det);
    D[4] = (float)(((double)S[0] * S[8] - (double)S[2] * S[6]) * det);
    D[5] = (float)(((double)S[2] * S[3] - (double)S[0] * S[5]) * det);
    D[6] = (float)(d2 * det);
    D[7] = (float)(((double)S[1] * S[6] - (double)S[0] * S[7]) * det);
    D[8] = (float)(((double)S[0] * S[4] - (double)S[1] * S[3]) * det);
}

// A^T A of the rows of ComputeH21 (homography) or ComputeF21 for 8 normalised pairs.
static void hyp_normal(bool hom, const float* p1, const float* p2, double* M) {
    for (int k = 0; k < 81; ++k) M[k] = 0.0;
    for (int i = 0; i < 8; ++i) {
        const float u1 = p1[2 * i], v1 = p1[2 * i + 1], u2 = p2[2 * i], v2 = p2[2 * i + 1];
        float rows[2][9];
        int nr;
        if (hom) {
            const float r0[9] = {0.0f, 0.0f, 0.0f, -u1, -v1, -1.0f, v2 * u1, v2 * v1, v2};
            const float r1[9] = {u1, v1, 1.0f, 0.0f, 0.0f, 0.0f, -u2 * u1, -u2 * v1, -u2};
            for (int k = 0; k < 9; ++k) { rows[0][k] = r0[k]; rows[1][k] = r1[k]; }
            nr = 2;
        } else {
            const float r0[9] = {u2 * u1, u2 * v1, u2, v2 * u1, v2 * v1, v2, u1, v1, 1.0f};
            for (int k = 0; k < 9; ++k) rows[0][k] = r0[k];
            nr = 1;
        }
        for (int r = 0; r < nr; ++r)
            for (int a = 0; a < 9; ++a)
                for (int b = 0; b < 9; ++b) M[9 * a + b] = M[9 * a + b] + (double)rows[r][a] * (double)rows[r][b];
    }
}

// Every hypothesis of FindHomography and FindFundamental from the RANSAC sets
// (n_sets x 8 indices into matches).  Outputs n_sets x 9 floats each.
void oracle_tv_hypotheses(const float* kp1, int32_t n1, const float* kp2, int32_t n2, const int32_t* matches,
                          const int32_t* sets, int32_t n_sets, float* H21, float* H12, float* F21) {
    std::vector<float> q1(2 * (size_t)n1), q2(2 * (size_t)n2);
    float T1[9], T2[9], T2inv[9], T2t[9];
    oracle_tv_normalize(kp1, n1, q1.data(), T1);
    oracle_tv_normalize(kp2, n2, q2.data(), T2);
    oracle_tv_inv3f(T2, T2inv);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T2t[3 * r + c] = T2[3 * c + r];
    for (int32_t it = 0; it < n_sets; ++it) {
        float p1[16], p2[16];
        for (int j = 0; j < 8; ++j) {
            const int32_t m = sets[8 * it + j];
            p1[2 * j] = q1[2 * matches[2 * m]];
            p1[2 * j + 1] = q1[2 * matches[2 * m] + 1];
            p2[2 * j] = q2[2 * matches[2 * m + 1]];
            p2[2 * j + 1] = q2[2 * matches[2 * m + 1] + 1];
        }
        double M[81], v[9];
        float Hn[9], tmp[9];
        hyp_normal(true, p1, p2, M);
        oracle_tv_small_eigvec(M, 9, v);
        for (int k = 0; k < 9; ++k) Hn[k] = (float)v[k];
        mul3(T2inv, Hn, tmp);
        mul3(tmp, T1, H21 + 9 * it);
        oracle_tv_inv3f(H21 + 9 * it, H12 + 9 * it);
        hyp_normal(false, p1, p2, M);
        oracle_tv_small_eigvec(M, 9, v);
        float Fp[9];
        for (int k = 0; k < 9; ++k) Fp[k] = (float)v[k];
        double G[9], w[3];  // Fpre^T Fpre
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc = acc + (double)Fp[3 * k + a] * (double)Fp[3 * k + b];
                G[3 * a + b] = acc;
            }
        oracle_tv_small_eigvec(G, 3, w);
        float Fn[9];  // Fpre (I - w w^T)
        for (int r = 0; r < 3; ++r) {
            double fw = 0.0;
            for (int k = 0; k < 3; ++k) fw = fw + (double)Fp[3 * r + k] * w[k];
            for (int c = 0; c < 3; ++c) Fn[3 * r + c] = (float)((double)Fp[3 * r + c] - fw * w[c]);
        }
        mul3(T2t, Fn, tmp);
        mul3(tmp, T1, F21 + 9 * it);
    }
}

}  // extern "C"
