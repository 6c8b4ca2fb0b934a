"""FP64 CPU restatement of the differentiable robust reprojection cost.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py as the checker of
the device cost and adjoint (include/nav_ba_cost.h); the product library
never loads it.

    C = sum_obs rho(|z - pi_d(T_cw X)|^2)

with pi_d the pinhole projection (Pinhole.cpp:30-36) of the Brown-Conrady
distorted normalised coordinates (the model PinholeRadTan.cpp:11-24 removes
from key points), rho the Huber kernel of the BA edge (BE_GraphOptim.cpp:191-195)
or the identity.  Pose gradients are taken in the tangent of the left
perturbation T <- exp(xi) T, xi = (omega, upsilon), the update and column order
of the reference edge (OptimizableTypes.cpp:140-161, OptimizableTypes.hpp:221-227).

Pins (tests/test_cost_oracle.py): every gradient against FP64 central finite
differences (eps 1e-6), and, at zero distortion, the pose and point gradients
against -2 rho' J^T e with the BA oracle's edge Jacobians (oracle_edge).
The distortion and intrinsics gradients have no reference counterpart: they
are pinned by the finite differences only.
"""
from __future__ import annotations

import numpy as np


def _project(T_cw, X, obs_cam, obs_pt, intr):
    R = T_cw[:, :9].reshape(-1, 3, 3)[obs_cam]
    t = T_cw[:, 9:][obs_cam]
    Xw = X[obs_pt]
    Xc = np.einsum("nij,nj->ni", R, Xw) + t
    iz = 1.0 / Xc[:, 2]
    x, y = Xc[:, 0] * iz, Xc[:, 1] * iz
    fx, fy, cx, cy, k1, k2, p1, p2 = intr
    r2 = x * x + y * y
    rad = 1.0 + k1 * r2 + k2 * r2 * r2
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return R, Xc, iz, x, y, r2, rad, xd, yd


def _huber(s, robust, delta):
    if not robust:
        return s.copy(), np.ones_like(s)
    d2 = delta * delta
    big = s > d2
    r = np.sqrt(np.where(big, s, 1.0))
    rho = np.where(big, 2 * delta * r - d2, s)
    w = np.where(big, delta / r, 1.0)
    return rho, w


def cost_terms(T_cw, X, obs_cam, obs_pt, obs_uv, intr, robust=True, delta=np.sqrt(5.991)):
    """Per-observation robust cost rho(|e|^2)."""
    _, _, _, _, _, _, _, xd, yd = _project(T_cw, X, obs_cam, obs_pt, intr)
    fx, fy, cx, cy = intr[:4]
    e0 = obs_uv[:, 0] - (fx * xd + cx)
    e1 = obs_uv[:, 1] - (fy * yd + cy)
    rho, _ = _huber(e0 * e0 + e1 * e1, robust, delta)
    return rho


def cost_and_grad(T_cw, X, obs_cam, obs_pt, obs_uv, intr, robust=True, delta=np.sqrt(5.991)):
    """Total cost and its gradient: (C, dC/dxi [n_cam, 6], dC/dX [n_pt, 3], dC/dintr [8])."""
    R, Xc, iz, x, y, r2, rad, xd, yd = _project(T_cw, X, obs_cam, obs_pt, intr)
    fx, fy, cx, cy, k1, k2, p1, p2 = intr
    e0 = obs_uv[:, 0] - (fx * xd + cx)
    e1 = obs_uv[:, 1] - (fy * yd + cy)
    rho, w = _huber(e0 * e0 + e1 * e1, robust, delta)
    gpu, gpv = -2 * w * e0, -2 * w * e1
    gi = np.zeros(8)
    gi[0] = np.sum(gpu * xd)
    gi[1] = np.sum(gpv * yd)
    gi[2] = np.sum(gpu)
    gi[3] = np.sum(gpv)
    gxd, gyd = gpu * fx, gpv * fy
    gi[4] = np.sum((gxd * x + gyd * y) * r2)
    gi[5] = np.sum((gxd * x + gyd * y) * r2 * r2)
    gi[6] = np.sum(gxd * 2 * x * y + gyd * (r2 + 2 * y * y))
    gi[7] = np.sum(gxd * (r2 + 2 * x * x) + gyd * 2 * x * y)
    kr = k1 + 2 * k2 * r2
    dxdx = rad + 2 * x * x * kr + 2 * p1 * y + 6 * p2 * x
    dxdy = 2 * x * y * kr + 2 * p1 * x + 2 * p2 * y
    dydy = rad + 2 * y * y * kr + 6 * p1 * y + 2 * p2 * x
    gx = gxd * dxdx + gyd * dxdy
    gy = gxd * dxdy + gyd * dydy
    gc = np.stack([gx * iz, gy * iz, -(gx * x + gy * y) * iz], axis=1)
    gXo = np.einsum("nji,nj->ni", R, gc)  # R^T gc
    gpo = np.concatenate([np.cross(Xc, gc), gc], axis=1)
    g_pose = np.zeros((T_cw.shape[0], 6))
    np.add.at(g_pose, obs_cam, gpo)
    g_X = np.zeros((X.shape[0], 3))
    np.add.at(g_X, obs_pt, gXo)
    return float(np.sum(rho)), g_pose, g_X, gi


def so3_exp(w):
    """Rodrigues' formula."""
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * (K @ K)


def se3_left(xi, T):
    """exp(xi) T for one pose row [R (9) | t (3)], xi = (omega, upsilon); the
    translation part of exp uses V = I + (1-cos)/th^2 K + (th - sin)/th^3 K^2."""
    w, u = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    Rw = so3_exp(w)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        V = np.eye(3) + 0.5 * K
    else:
        V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * K + (th - np.sin(th)) / th ** 3 * (K @ K)
    R = T[:9].reshape(3, 3)
    out = np.empty(12)
    out[:9] = (Rw @ R).reshape(9)
    out[9:] = Rw @ T[9:] + V @ u
    return out
