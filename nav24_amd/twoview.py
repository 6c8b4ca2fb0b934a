"""Scoring of two-view RANSAC hypotheses on the GPU.

Host-side mirror of the scoring half of TwoViewReconstruction
(core/operators/mapInit/OP_2ViewReconstruction.hpp:94, .cpp:266-366,
:447-611): built from the same parameters (sigma, the chi-square thresholds
mThChiSqScore and mThChiSqF of Params2VR) and exposing FindHomography /
FindFundamental, which score all RANSAC hypotheses at once (CheckHomography /
CheckFundamental of each) and keep the best by the reference's rule, with
the reference's outputs (inlier flags, score, kept matrix); FindHomographyScores /
FindFundamentalScores also return every score.  CheckRT (:1244-1352) runs
for all motion candidates of ReconstructF / ReconstructH at once.  The
hypotheses come from the caller (the 8-point solves stay on the host).  The
arithmetic runs in the C ABI (include/nav_twoview.h); without the built library
or a gfx950 device the calls raise.

    tv = TwoViewReconstructionHip(sigma=1.0)
    scores, best, best_score, inliers = tv.FindHomographyScores(kp1, kp2, matches, H21s, H12s)
    scores, best, best_score, inliers = tv.FindFundamentalScores(kp1, kp2, matches, F21s)
    n_good, p3d, good, parallax = tv.CheckRT(kp1, kp2, matches, inliers, K, Rs, ts)   # 4 or 8 candidates
    H21s, H12s, F21s = tv.Hypotheses(kp1, kp2, matches, sets)                         # from the RANSAC sets
    tv.score_device("F", kp1_t, kp2_t, matches_t, F21s_t, out=(scores_t, best_t, best_score_t, inliers_t))
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from ._loader import load_library
from .abi import NAV_BA_OK, STATUS_NAMES

# Params2VR defaults (OP_2ViewReconstruction.hpp:32-44)
DEF_TH_CHISQ_SCORE = 5.991
DEF_TH_CHISQ_F = 3.841
DEF_2VR_SIGMA = 1.0
DEF_MIN_PARALLAX_CRT = 0.99998
DEF_MIN_TRIANGULATED = 50
MAX_CANDIDATES = 8  # NAV_TV_MAX_CANDIDATES

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = load_library("libnav24_ba.so")
        fp, ip, u8 = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        v, i32, f32 = C.c_void_p, C.c_int32, C.c_float
        L.nav_tv_score_homography.argtypes = [v, v, v, i32, v, v, i32, f32, f32, v, v, v, v, v]
        L.nav_tv_score_homography.restype = i32
        L.nav_tv_score_fundamental.argtypes = [v, v, v, i32, v, i32, f32, f32, f32, v, v, v, v, v]
        L.nav_tv_score_fundamental.restype = i32
        L.nav_tv_score_homography_host.argtypes = [fp, i32, fp, i32, ip, i32, fp, fp, i32, f32, f32, fp, ip, fp,
                                                   u8, i32]
        L.nav_tv_score_homography_host.restype = i32
        L.nav_tv_score_fundamental_host.argtypes = [fp, i32, fp, i32, ip, i32, fp, i32, f32, f32, f32, fp, ip, fp,
                                                    u8, i32]
        L.nav_tv_score_fundamental_host.restype = i32
        L.nav_tv_check_rt.argtypes = [v, i32, v, v, v, i32, fp, fp, fp, i32, f32, f32, i32] + [v] * 8
        L.nav_tv_check_rt.restype = i32
        L.nav_tv_check_rt_host.argtypes = [fp, i32, fp, i32, ip, u8, i32, fp, fp, fp, i32, f32, f32, i32, ip, fp, fp,
                                           fp, u8, i32]
        L.nav_tv_check_rt_host.restype = i32
        L.nav_tv_hypotheses_host.argtypes = [fp, i32, fp, i32, ip, i32, ip, i32, fp, fp, fp, i32]
        L.nav_tv_hypotheses_host.restype = i32
        L.nav_tv_hypotheses.argtypes = [v, v, v, v, i32, fp, fp, v, v, v, v]
        L.nav_tv_hypotheses.restype = i32
        L.nav_tv_normalize.argtypes = [fp, i32, fp, fp]
        L.nav_tv_normalize.restype = i32
        _lib = L
    return _lib


def _check(st: int, what: str) -> None:
    if st != NAV_BA_OK:
        raise RuntimeError(f"{what}: {STATUS_NAMES.get(st, st)}")


def _pts(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).reshape(-1, 2)


def _mats(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).reshape(-1, 9)


class TwoViewReconstructionHip:
    """Hypothesis scoring of the reference's TwoViewReconstruction on device
    `device`.  Key points are (n, 2) float (cv::KeyPoint::pt), matches (m, 2)
    int pairs (index in image 1, index in image 2), the reference's
    mvMatches12; matrices 3x3 row-major float."""

    def __init__(self, sigma: float = DEF_2VR_SIGMA, th_chisq_score: float = DEF_TH_CHISQ_SCORE,
                 th_chisq_f: float = DEF_TH_CHISQ_F, min_parallax_crt: float = DEF_MIN_PARALLAX_CRT,
                 min_triangulated: int = DEF_MIN_TRIANGULATED, device: int = 0):
        if not sigma > 0:
            raise ValueError("sigma must be > 0")
        self.sigma = float(sigma)
        self.th_score = float(th_chisq_score)
        self.th_f = float(th_chisq_f)
        self.min_parallax_crt = float(min_parallax_crt)
        self.min_triangulated = int(min_triangulated)
        self.device = device

    @property
    def th2(self) -> float:
        """CheckRT's reprojection threshold as ReconstructF / ReconstructH pass it: 4 sigma^2 (float)."""
        s = np.float32(self.sigma)
        return float(np.float32(4.0) * (s * s))

    # ---- RANSAC hypotheses from the 8-match sets ----------------------------
    def Hypotheses(self, kp1, kp2, matches, sets):
        """The hypotheses FindHomography / FindFundamental form from the RANSAC
        sets (k, 8) of match indices (mvSets): (H21s, H12s, F21s) float32
        (k, 3, 3), normalisation and 8-point solves as ComputeH21 / ComputeF21
        (OP_2ViewReconstruction.cpp:266-366, :368-446) on the device."""
        k1, k2 = _pts(kp1), _pts(kp2)
        m = np.ascontiguousarray(matches, np.int32).reshape(-1, 2)
        st = np.ascontiguousarray(sets, np.int32).reshape(-1, 8)
        ns = st.shape[0]
        H21 = np.zeros((ns, 9), np.float32)
        H12 = np.zeros((ns, 9), np.float32)
        F21 = np.zeros((ns, 9), np.float32)
        fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
        r = lib().nav_tv_hypotheses_host(k1.ctypes.data_as(fp), k1.shape[0], k2.ctypes.data_as(fp), k2.shape[0],
                                         m.ctypes.data_as(ip), m.shape[0], st.ctypes.data_as(ip), ns,
                                         H21.ctypes.data_as(fp), H12.ctypes.data_as(fp), F21.ctypes.data_as(fp),
                                         self.device)
        _check(r, "nav_tv_hypotheses_host")
        return H21.reshape(ns, 3, 3), H12.reshape(ns, 3, 3), F21.reshape(ns, 3, 3)

    # ---- CheckRT for a list of motion candidates ---------------------------
    def CheckRT(self, kp1, kp2, matches, inliers, K, Rs, ts):
        """CheckRT (OP_2ViewReconstruction.cpp:1244-1352) of every candidate
        (Rs (k, 3, 3), ts (k, 3), k <= 8) with th2 = 4 sigma^2.  Returns
        (nGood int32[k], vP3D float32[k, n_kp1, 3], vbGood bool[k, n_kp1],
        parallax float32[k])."""
        k1, k2 = _pts(kp1), _pts(kp2)
        m = np.ascontiguousarray(matches, np.int32).reshape(-1, 2)
        inl = np.ascontiguousarray(inliers, np.uint8).reshape(-1)
        if inl.size != m.shape[0]:
            raise ValueError("one inlier flag per match")
        Kf = np.ascontiguousarray(K, np.float32).reshape(9)
        R = _mats(Rs)
        t = np.ascontiguousarray(ts, np.float32).reshape(-1, 3)
        nc, nk = R.shape[0], k1.shape[0]
        if t.shape[0] != nc or not 1 <= nc <= MAX_CANDIDATES:
            raise ValueError(f"1 to {MAX_CANDIDATES} candidates, as many R as t")
        ng = np.zeros(nc, np.int32)
        par = np.zeros(nc, np.float32)
        cs = np.zeros(nc, np.float32)
        p3d = np.zeros((nc, nk, 3), np.float32)
        good = np.zeros((nc, nk), np.uint8)
        fp, ip, u8 = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        st = lib().nav_tv_check_rt_host(k1.ctypes.data_as(fp), nk, k2.ctypes.data_as(fp), k2.shape[0],
                                        m.ctypes.data_as(ip), inl.ctypes.data_as(u8), m.shape[0],
                                        Kf.ctypes.data_as(fp), R.ctypes.data_as(fp), t.ctypes.data_as(fp), nc,
                                        self.th2, self.min_parallax_crt, self.min_triangulated,
                                        ng.ctypes.data_as(ip), par.ctypes.data_as(fp), cs.ctypes.data_as(fp),
                                        p3d.ctypes.data_as(fp), good.ctypes.data_as(u8), self.device)
        _check(st, "nav_tv_check_rt_host")
        self.last_cos_sel = cs
        return ng, p3d, good.astype(bool), par

    def check_rt_device(self, kp1, kp2, matches, inliers, K, Rs, ts, out: Optional[dict] = None):
        """CheckRT on device tensors (key points float32 (n, 2), matches int32
        (m, 2), inliers uint8 (m,), cuda, contiguous); K, Rs, ts on the host.
        Returns a dict of tensors n_good, parallax, cos_sel, p3d, good,
        cos_parallax, passed on torch's current stream."""
        import torch

        for t_, dt, nm in ((kp1, torch.float32, "kp1"), (kp2, torch.float32, "kp2"), (matches, torch.int32, "matches"),
                           (inliers, torch.uint8, "inliers")):
            if t_.dtype != dt or not t_.is_cuda or not t_.is_contiguous():
                raise ValueError(f"{nm}: contiguous {dt} cuda tensor")
        n, nk = matches.numel() // 2, kp1.numel() // 2
        if inliers.numel() != n:
            raise ValueError("one inlier flag per match")
        Kf = np.ascontiguousarray(K, np.float32).reshape(9)
        R = _mats(Rs)
        t = np.ascontiguousarray(ts, np.float32).reshape(-1, 3)
        nc = R.shape[0]
        if t.shape[0] != nc or not 1 <= nc <= MAX_CANDIDATES:
            raise ValueError(f"1 to {MAX_CANDIDATES} candidates, as many R as t")
        dev = kp1.device
        if out is None:
            out = dict(n_good=torch.empty(nc, dtype=torch.int32, device=dev),
                       parallax=torch.empty(nc, dtype=torch.float32, device=dev),
                       cos_sel=torch.empty(nc, dtype=torch.float32, device=dev),
                       p3d=torch.empty((nc, nk, 3), dtype=torch.float32, device=dev),
                       good=torch.empty((nc, nk), dtype=torch.uint8, device=dev),
                       cos_parallax=torch.empty((nc, n), dtype=torch.float32, device=dev),
                       passed=torch.empty((nc, n), dtype=torch.uint8, device=dev))
        fp = C.POINTER(C.c_float)
        p = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
        st = lib().nav_tv_check_rt(p(kp1), nk, p(kp2), p(matches), p(inliers), n, Kf.ctypes.data_as(fp),
                                   R.ctypes.data_as(fp), t.ctypes.data_as(fp), nc, self.th2, self.min_parallax_crt,
                                   self.min_triangulated, p(out["n_good"]), p(out["parallax"]), p(out["cos_sel"]),
                                   p(out["p3d"]), p(out["good"]), p(out["cos_parallax"]), p(out["passed"]),
                                   C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        _check(st, "nav_tv_check_rt")
        return out

    # ---- all hypotheses of one RANSAC run ---------------------------------
    def FindHomographyScores(self, kp1, kp2, matches, H21s, H12s) -> Tuple[np.ndarray, int, np.float32, np.ndarray]:
        """Scores of every (H21, H12) pair, the index FindHomography keeps (-1
        if none scores above 0), its score and its inlier flags (uint8)."""
        return self._find(True, kp1, kp2, matches, H21s, H12s)

    def FindFundamentalScores(self, kp1, kp2, matches, F21s) -> Tuple[np.ndarray, int, np.float32, np.ndarray]:
        """As FindHomographyScores for fundamental-matrix hypotheses."""
        return self._find(False, kp1, kp2, matches, F21s, None)

    # ---- the reference's outputs: FindHomography / FindFundamental ----------
    def FindHomography(self, kp1, kp2, matches, H21s, H12s):
        """(vbMatchesInliers bool[m], score, H21) as FindHomography leaves them:
        H21 None and every flag false when no hypothesis scores above 0."""
        _, best, score, inl = self._find(True, kp1, kp2, matches, H21s, H12s)
        return inl.astype(bool), float(score), (None if best < 0 else _mats(H21s)[best].reshape(3, 3).copy())

    def FindFundamental(self, kp1, kp2, matches, F21s):
        """(vbMatchesInliers bool[m], score, F21) as FindFundamental leaves them."""
        _, best, score, inl = self._find(False, kp1, kp2, matches, F21s, None)
        return inl.astype(bool), float(score), (None if best < 0 else _mats(F21s)[best].reshape(3, 3).copy())

    def _find(self, hom: bool, kp1, kp2, matches, M0, M1):
        k1, k2 = _pts(kp1), _pts(kp2)
        m = np.ascontiguousarray(matches, np.int32).reshape(-1, 2)
        A = _mats(M0)
        B = _mats(M1) if hom else None
        if hom and B.shape != A.shape:
            raise ValueError("H21s and H12s differ in count")
        n_hyp = A.shape[0]
        scores = np.zeros(n_hyp, np.float32)
        best = np.zeros(1, np.int32)
        bs = np.zeros(1, np.float32)
        inl = np.zeros(m.shape[0], np.uint8)
        fp, ip, u8 = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        common = (k1.ctypes.data_as(fp), k1.shape[0], k2.ctypes.data_as(fp), k2.shape[0], m.ctypes.data_as(ip),
                  m.shape[0])
        outs = (scores.ctypes.data_as(fp), best.ctypes.data_as(ip), bs.ctypes.data_as(fp), inl.ctypes.data_as(u8),
                self.device)
        if hom:
            st = lib().nav_tv_score_homography_host(*common, A.ctypes.data_as(fp), B.ctypes.data_as(fp), n_hyp,
                                                    self.sigma, self.th_score, *outs)
            _check(st, "nav_tv_score_homography_host")
        else:
            st = lib().nav_tv_score_fundamental_host(*common, A.ctypes.data_as(fp), n_hyp, self.sigma, self.th_f,
                                                     self.th_score, *outs)
            _check(st, "nav_tv_score_fundamental_host")
        return scores, int(best[0]), bs[0], inl

    # ---- device tensors -----------------------------------------------------
    def score_device(self, kind: str, kp1, kp2, matches, M0, M1=None, out: Optional[tuple] = None):
        """kind "H" (M0 = H21s, M1 = H12s) or "F" (M0 = F21s): torch tensors on
        cuda — key points (n, 2) float32, matches (m, 2) int32, matrices (k, 3, 3)
        or (k, 9) float32, all contiguous.  Returns (scores (k,) float32, best (1,)
        int32, best_score (1,) float32, inliers (m,) uint8) on torch's current
        stream; `out` may supply those four tensors."""
        import torch

        hom = kind == "H"
        if kind not in ("H", "F"):
            raise ValueError('kind: "H" or "F"')
        dev = kp1.device

        def need(t, dtype, name):
            if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name}: contiguous {dtype} cuda tensor")

        need(kp1, torch.float32, "kp1")
        need(kp2, torch.float32, "kp2")
        need(matches, torch.int32, "matches")
        need(M0, torch.float32, "M0")
        if hom:
            if M1 is None:
                raise ValueError("H12s needed for kind H")
            need(M1, torch.float32, "M1")
            if M1.numel() != M0.numel():
                raise ValueError("H21s and H12s differ in count")
        if M0.numel() % 9 or matches.numel() % 2 or kp1.numel() % 2 or kp2.numel() % 2:
            raise ValueError("shapes: key points (n, 2), matches (m, 2), matrices (k, 9)")
        n_hyp, n = M0.numel() // 9, matches.numel() // 2
        if out is None:
            out = (torch.empty(n_hyp, dtype=torch.float32, device=dev), torch.empty(1, dtype=torch.int32, device=dev),
                   torch.empty(1, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.uint8, device=dev))
        scores, best, bs, inl = out
        if scores.numel() < n_hyp or inl.numel() < n:
            raise ValueError("out tensors too small")
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        if hom:
            st = lib().nav_tv_score_homography(p(kp1), p(kp2), p(matches), n, p(M0), p(M1), n_hyp, self.sigma,
                                               self.th_score, p(scores), p(best), p(bs), p(inl), stream)
        else:
            st = lib().nav_tv_score_fundamental(p(kp1), p(kp2), p(matches), n, p(M0), n_hyp, self.sigma, self.th_f,
                                                self.th_score, p(scores), p(best), p(bs), p(inl), stream)
        _check(st, "nav_tv_score_" + ("homography" if hom else "fundamental"))
        return scores, best, bs, inl
