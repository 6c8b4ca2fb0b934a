"""Seeded synthetic two-view scenes and RANSAC hypotheses for the two-view
scoring path (tests and benchmark input; numpy only).

A scene is what TwoViewReconstruction::Reconstruct receives
(core/operators/mapInit/OP_2ViewReconstruction.cpp:68-97): key points of two
images and the match list mvMatches12, here with a pinhole camera K, a second
camera at (R, t), points on a plane (`planar`, the homography case) or spread
in depth (the fundamental case), 1 px Gaussian noise and a share of outlier
matches.  Hypotheses follow FindHomography / FindFundamental (:266-366): 8
distinct matches per RANSAC iteration, Hartley-normalised DLT (ComputeH21) or
8-point F with rank 2 enforced (ComputeF21), mapped back with the
normalisations and rounded to float; H12 = H21^-1.  numpy's SVD stands in for
OpenCV's, so the matrices are not the reference's bit for bit; the scoring
that consumes them is what is compared.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class TwoViewScene:
    kp1: np.ndarray      # (n1, 2) float32
    kp2: np.ndarray      # (n2, 2) float32
    matches: np.ndarray  # (m, 2) int32: (index in kp1, index in kp2)
    K: np.ndarray        # (3, 3) float64
    R: np.ndarray
    t: np.ndarray
    inlier: np.ndarray   # (m,) bool: generated as a true correspondence


def synthetic_two_view(n_matches: int, seed: int = 0, planar: bool = False, outlier_frac: float = 0.2,
                       noise_px: float = 1.0, extra_kp: int = 100) -> TwoViewScene:
    rng = np.random.default_rng(seed)
    K = np.array([[458.654, 0.0, 367.215], [0.0, 457.296, 248.375], [0.0, 0.0, 1.0]])
    ang = np.array([0.02, -0.05, 0.01])
    th = np.linalg.norm(ang)
    a = ang / th
    Ax = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + np.sin(th) * Ax + (1 - np.cos(th)) * Ax @ Ax
    t = np.array([0.25, 0.03, 0.02])
    n = n_matches
    x = rng.uniform(-1.5, 1.5, n)
    y = rng.uniform(-1.0, 1.0, n)
    if planar:  # on the plane Z = 4 + 0.2 X - 0.1 Y
        Xx, Xy = 3.0 * x, 3.0 * y
        X = np.stack([Xx, Xy, 4.0 + 0.2 * Xx - 0.1 * Xy], 1)
    else:
        z = rng.uniform(2.0, 10.0, n)
        X = np.stack([x * z / 2.0, y * z / 2.0, z], 1)
    p1 = (K @ X.T).T
    X2 = (R @ X.T).T + t
    p2 = (K @ X2.T).T
    uv1 = p1[:, :2] / p1[:, 2:] + rng.normal(0, noise_px, (n, 2))
    uv2 = p2[:, :2] / p2[:, 2:] + rng.normal(0, noise_px, (n, 2))
    inl = rng.random(n) >= outlier_frac
    nout = int((~inl).sum())
    uv2[~inl] = rng.uniform([0, 0], [752, 480], (nout, 2))
    # the key point lists hold the matched points in shuffled order plus unmatched ones
    n1, n2 = n + extra_kp, n + extra_kp
    kp1 = rng.uniform([0, 0], [752, 480], (n1, 2))
    kp2 = rng.uniform([0, 0], [752, 480], (n2, 2))
    i1 = rng.permutation(n1)[:n]
    i2 = rng.permutation(n2)[:n]
    kp1[i1] = uv1
    kp2[i2] = uv2
    order = np.argsort(i1)  # mvMatches12 is built in increasing index of image 1
    matches = np.stack([i1[order], i2[order]], 1).astype(np.int32)
    return TwoViewScene(kp1.astype(np.float32), kp2.astype(np.float32), matches, K, R, t, inl[order])


def _normalize(p: np.ndarray):
    """Hartley normalisation as TwoViewReconstruction::Normalize: mean 0, mean
    absolute deviation 1 per axis."""
    m = p.mean(0)
    d = np.abs(p - m).mean(0)
    s = 1.0 / d
    T = np.array([[s[0], 0, -m[0] * s[0]], [0, s[1], -m[1] * s[1]], [0, 0, 1.0]])
    return (p - m) * s, T


def _h_dlt(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    A = []
    for (u1, v1), (u2, v2) in zip(a, b):
        A.append([0, 0, 0, -u1, -v1, -1, v2 * u1, v2 * v1, v2])
        A.append([u1, v1, 1, 0, 0, 0, -u2 * u1, -u2 * v1, -u2])
    _, _, vt = np.linalg.svd(np.array(A))
    return vt[-1].reshape(3, 3)


def _f_8pt(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    A = [[u2 * u1, u2 * v1, u2, v2 * u1, v2 * v1, v2, u1, v1, 1] for (u1, v1), (u2, v2) in zip(a, b)]
    _, _, vt = np.linalg.svd(np.array(A))
    F = vt[-1].reshape(3, 3)
    u, w, vt = np.linalg.svd(F)
    w[2] = 0
    return u @ np.diag(w) @ vt


def ransac_hypotheses(scene: TwoViewScene, n_hyp: int, seed: int = 0):
    """(H21s (k, 3, 3), H12s, F21s) float32 from 8-match sets drawn as the
    reference draws them (8 distinct matches per iteration)."""
    rng = np.random.default_rng(seed + 1)
    m = scene.matches
    a1, T1 = _normalize(scene.kp1.astype(np.float64))  # over all key points of an image, as the reference
    a2, T2 = _normalize(scene.kp2.astype(np.float64))
    n1, n2 = a1[m[:, 0]], a2[m[:, 1]]
    H21s, H12s, F21s = [], [], []
    for _ in range(n_hyp):
        idx = rng.choice(len(m), 8, replace=False)
        Hn = _h_dlt(n1[idx], n2[idx])
        H21 = (np.linalg.inv(T2) @ Hn @ T1).astype(np.float32)
        H21s.append(H21)
        H12s.append(np.linalg.inv(H21.astype(np.float64)).astype(np.float32))
        Fn = _f_8pt(n1[idx], n2[idx])
        F21s.append((T2.T @ Fn @ T1).astype(np.float32))
    return np.stack(H21s), np.stack(H12s), np.stack(F21s)


def motion_candidates(scene: TwoViewScene, n_extra: int = 0, seed: int = 0):
    """(R (k, 3, 3), t (k, 3)) float32: the four motions DecomposeE yields for
    the scene's essential matrix (R and its 180-degree twist about t, with +t
    and -t; OP_2ViewReconstruction.cpp:1356-1376), then n_extra perturbed
    motions (ReconstructH checks 8 candidates)."""
    th = scene.t / np.linalg.norm(scene.t)
    Rtw = (2 * np.outer(th, th) - np.eye(3)) @ scene.R
    R = [scene.R, Rtw, scene.R, Rtw]
    t = [scene.t, scene.t, -scene.t, -scene.t]
    rng = np.random.default_rng(seed)
    for _ in range(n_extra):
        w = rng.normal(0, 0.02, 3)
        a = np.linalg.norm(w)
        Wx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
        R.append((np.eye(3) + np.sin(a) / a * Wx + (1 - np.cos(a)) / a ** 2 * Wx @ Wx) @ scene.R)
        t.append(scene.t + rng.normal(0, 0.02, 3))
    return np.stack(R).astype(np.float32), np.stack(t).astype(np.float32)
