/*
 * nav_twoview.h — the per-match work of two-view map initialisation on the
 * device, which produces BA's first state: the RANSAC hypotheses and their
 * scoring, and triangulation with the cheirality test of the motion candidates.
 *
 * Replaces the scoring loops of TwoViewReconstruction::FindHomography and
 * ::FindFundamental (core/operators/mapInit/OP_2ViewReconstruction.cpp:266-315
 * and :317-366), i.e. CheckHomography (:447-528) and CheckFundamental
 * (:530-611) called once per RANSAC iteration, followed by the keep-the-best
 * rule `if (currentScore > score) { best = it; score = currentScore; }`
 * starting from score = 0.  The hypotheses themselves (8-point sets, the SVD
 * in ComputeH21 / ComputeF21, the normalisation T2inv Hn T1, H21.inv()) stay
 * with the caller; every hypothesis is scored in one launch and the best is
 * chosen on the device.
 *
 * Per match i (u1, v1) = kp1[matches[i].first], (u2, v2) = kp2[matches[i].second],
 * all float as in the reference, evaluated left to right without fused
 * multiply-adds:
 *   homography:  w = 1 / (h31' u2 + h32' v2 + h33'),  x2in1 = ((h11' u2 + h12' v2 + h13') w, ...)
 *                chi1 = |x1 - x2in1|^2 * (1 / sigma^2), and chi2 likewise with H21 and x1;
 *                each chi <= th_score adds (th_score - chi) to the score, and
 *                a match is an inlier when neither chi exceeds th_score.
 *   fundamental: (a2, b2, c2) = F21 x1, chi1 = (a2 u2 + b2 v2 + c2)^2 / (a2^2 + b2^2) * (1 / sigma^2),
 *                (a1, b1, c1) = F21^T x2, chi2 likewise; each chi <= th_f adds
 *                (th_score - chi) to the score; inlier when neither exceeds th_f.
 * The score of one hypothesis is summed in float in the reference's order
 * (match 0 term 1, match 0 term 2, match 1 term 1, ...), so it equals the
 * reference's float sum bit for bit.  A NaN chi counts as the reference's
 * does (not above the threshold, so the NaN enters the score, and such a score
 * never wins).  The reference's defaults: sigma 1, th_score 5.991, th_f 3.841
 * (OP_2ViewReconstruction.hpp:32-44).
 */
#ifndef NAV_TWOVIEW_H
#define NAV_TWOVIEW_H

#include <stdint.h>
#include "nav_ba.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs (device arrays on `stream`, a hipStream_t, NULL = the null stream):
 *   kp1, kp2       key points (u, v) float of image 1 and image 2 (cv::KeyPoint::pt)
 *   matches        n_matches pairs (index into kp1, index into kp2) int32, the
 *                  reference's mvMatches12; indices are not range-checked
 *   H21, H12       n_hyp row-major 3x3 float matrices each (H12 = H21^-1 as the caller formed it)
 *   F21            n_hyp row-major 3x3 float matrices
 * Outputs (device):
 *   scores         n_hyp floats, the score of every hypothesis
 *   best           one int32: the index the reference keeps (the first one
 *                  whose score is above every earlier score and above 0), -1 if none
 *   best_score     one float: that score, 0 if none
 *   inliers        n_matches bytes, 1 where the best hypothesis holds the
 *                  match as an inlier, all 0 if none (vbMatchesInliers)
 * n_matches, n_hyp >= 0; n_matches <= NAV_TV_MAX_MATCHES.  Returns after the
 * launches are queued; NAV_BA_BAD_ARG for sizes out of range, a null pointer
 * or sigma <= 0, NAV_BA_HIP_ERROR if a launch fails. */
#define NAV_TV_MAX_MATCHES (1 << 24)

int32_t nav_tv_score_homography(const float* kp1, const float* kp2, const int32_t* matches, int32_t n_matches,
                                const float* H21, const float* H12, int32_t n_hyp, float sigma, float th_score,
                                float* scores, int32_t* best, float* best_score, uint8_t* inliers, void* stream);

int32_t nav_tv_score_fundamental(const float* kp1, const float* kp2, const int32_t* matches, int32_t n_matches,
                                 const float* F21, int32_t n_hyp, float sigma, float th_f, float th_score,
                                 float* scores, int32_t* best, float* best_score, uint8_t* inliers, void* stream);

/* Host arrays, through device `device` (copies in and out; blocks until done);
 * n1, n2 are the key-point counts of the two images, and matches are checked
 * against them.  NAV_BA_NO_DEVICE without a usable gfx950 device. */
int32_t nav_tv_score_homography_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2,
                                     const int32_t* matches, int32_t n_matches, const float* H21, const float* H12,
                                     int32_t n_hyp, float sigma, float th_score, float* scores, int32_t* best,
                                     float* best_score, uint8_t* inliers, int32_t device);

int32_t nav_tv_score_fundamental_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2,
                                      const int32_t* matches, int32_t n_matches, const float* F21, int32_t n_hyp,
                                      float sigma, float th_f, float th_score, float* scores, int32_t* best,
                                      float* best_score, uint8_t* inliers, int32_t device);

/* Motion candidates: TwoViewReconstruction::CheckRT (OP_2ViewReconstruction.cpp:1244-1352)
 * with Triangulate (:1179-1194) for n_cand candidates (R, t) at once, as
 * ReconstructF (4 candidates, :612-700) and ReconstructH (8) call it.  Per
 * inlier match: the 4x4 system of Triangulate, its null vector, then the
 * reference's float tests in order (finite point; parallax cosine between the
 * rays from both camera centres; depth > 0 in both cameras unless
 * cos >= min_parallax_cos; squared reprojection error <= th2 in both images).
 * The reference takes the null vector from OpenCV's float SVD; here it comes
 * from a one-sided Jacobi SVD in double, rounded to float, and matrix
 * products, norms and dot products sum float products in double, rounded to
 * float.  So results equal the reference's up to OpenCV's own rounding, and
 * a match close to a threshold can fall on the other side.
 * Inputs:  kp1 (n_kp1 points), kp2, matches (n_matches pairs), inliers
 *          (n_matches bytes, vbMatchesInliers), K (3x3 row-major float,
 *          host), R (n_cand row-major 3x3, host), t (n_cand x 3, host),
 *          th2 (the reference passes 4 sigma^2), min_parallax_cos
 *          (mMinParallaxCRT, 0.99998), min_triangulated (mMinTriangulated, 50).
 * Outputs per candidate c (device for nav_tv_check_rt, host for _host):
 *   n_good[c]         CheckRT's return value
 *   parallax[c]       its parallax (degrees): acos of cos_sel[c], 0 if n_good is 0
 *   cos_sel[c]        the cosine at index min(min_triangulated, n_good - 1) of the
 *                     ascending passing cosines (NaN if none)
 *   p3d[c][n_kp1][3]  vP3D: the point at the image-1 key point of each passing
 *                     match, 0 elsewhere
 *   good[c][n_kp1]    vbGood: 1 where a passing match has cos < min_parallax_cos
 *   cos_parallax[c][n_matches], passed[c][n_matches] (device entry only):
 *                     each match's cosine and pass flag (also the selection's input)
 * 1 <= n_cand <= NAV_TV_MAX_CANDIDATES.  The device entry does not range-check
 * the match indices; the _host entry does. */
#define NAV_TV_MAX_CANDIDATES 8

int32_t nav_tv_check_rt(const float* kp1, int32_t n_kp1, const float* kp2, const int32_t* matches,
                        const uint8_t* inliers, int32_t n_matches, const float* K, const float* R, const float* t,
                        int32_t n_cand, float th2, float min_parallax_cos, int32_t min_triangulated,
                        int32_t* n_good, float* parallax, float* cos_sel, float* p3d, uint8_t* good,
                        float* cos_parallax, uint8_t* passed, void* stream);

int32_t nav_tv_check_rt_host(const float* kp1, int32_t n_kp1, const float* kp2, int32_t n_kp2,
                             const int32_t* matches, const uint8_t* inliers, int32_t n_matches, const float* K,
                             const float* R, const float* t, int32_t n_cand, float th2, float min_parallax_cos,
                             int32_t min_triangulated, int32_t* n_good, float* parallax, float* cos_sel, float* p3d,
                             uint8_t* good, int32_t device);

/* RANSAC hypotheses: what FindHomography / FindFundamental form per iteration
 * (OP_2ViewReconstruction.cpp:266-366) before scoring, for n_sets sets of 8
 * match indices (the reference's mvSets, drawn by the caller):
 *   Normalize both images' key points (:1195-1240; float sums in input order),
 *   ComputeH21 / ComputeF21 (:368-446) on the 8 normalised pairs,
 *   H21 = T2^-1 Hn T1, H12 = H21^-1, F21 = T2^T Fn T1.
 * The reference takes each solution from OpenCV's float SVD.  Here it is the
 * eigenvector of the smallest eigenvalue of A^T A by inverse iteration in
 * double (40 iterations on A^T A + 1e-13 trace I), unit length with its
 * largest-magnitude entry positive, rounded to float; F's rank-2 step
 * removes the smallest-eigenvalue direction of Fpre^T Fpre.  A hypothesis is
 * the reference's up to its scale and to float rounding, except for a
 * degenerate 8-set whose two smallest singular values are close.
 * Outputs H21, H12, F21: n_sets row-major 3x3 floats each.
 *
 * nav_tv_normalize: Normalize on the host (n >= 1 points -> normalised points
 * and T, row-major), for callers of the device entry.
 * nav_tv_hypotheses: device arrays (q1, q2 the normalised key points, matches,
 * sets, outputs) on `stream`; T1, T2 on the host; set entries are not
 * range-checked.
 * nav_tv_hypotheses_host: host arrays; normalises on the host, checks every
 * match and set index; NAV_BA_NO_DEVICE without a gfx950 device. */
int32_t nav_tv_normalize(const float* kp, int32_t n, float* out, float* T);

int32_t nav_tv_hypotheses(const float* q1, const float* q2, const int32_t* matches, const int32_t* sets,
                          int32_t n_sets, const float* T1, const float* T2, float* H21, float* H12, float* F21,
                          void* stream);

int32_t nav_tv_hypotheses_host(const float* kp1, int32_t n1, const float* kp2, int32_t n2, const int32_t* matches,
                               int32_t n_matches, const int32_t* sets, int32_t n_sets, float* H21, float* H12,
                               float* F21, int32_t device);

#ifdef __cplusplus
}
#endif
#endif /* NAV_TWOVIEW_H */
