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
#include <cmath>
#include <cstdint>

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

}  // extern "C"
