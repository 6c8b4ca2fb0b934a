"""Pins of the FP64 differentiable-cost oracle (CPU only).

Every gradient against FP64 central finite differences (eps 1e-6): poses in
the tangent of the left perturbation exp(xi) T, points, intrinsics and
distortion.  At zero distortion the pose and point gradients must equal
-2 rho' J^T e with the BA oracle's own edge Jacobians (the reference edge,
OptimizableTypes.cpp:140-161): a sign or column-order mistake in either fails.
"""
import numpy as np
import pytest

from nav24_amd import problems
from oracle import cost_oracle as co
from oracle import oracle


def small(seed=0, n_cam=4, n_pt=30, dist=True):
    p = problems.params_for(4, seed)
    p.n_cam, p.n_pt = n_cam, n_pt
    prob = problems.make(p)
    intr = np.array([prob.fx, prob.fy, prob.cx, prob.cy, *prob.dist])
    if not dist:
        intr[4:] = 0.0
    return prob, intr


def total(prob, T, X, intr, robust):
    return float(np.sum(co.cost_terms(T, X, prob.obs_cam, prob.obs_pt, prob.obs_uv, intr, robust)))


@pytest.mark.parametrize("robust", [False, True])
def test_gradient_matches_central_differences(robust):
    prob, intr = small(seed=1)
    T, X = prob.T_cw.copy(), prob.X.copy()
    c, gp, gx, gi = co.cost_and_grad(T, X, prob.obs_cam, prob.obs_pt, prob.obs_uv, intr, robust)
    assert c == pytest.approx(total(prob, T, X, intr, robust), rel=1e-14)
    eps = 1e-6
    scale = max(abs(c), 1.0)
    # poses: xi tangent, left perturbation
    for cam in range(prob.n_cam):
        for k in range(6):
            xi = np.zeros(6)
            xi[k] = eps
            Tp, Tm = T.copy(), T.copy()
            Tp[cam] = co.se3_left(xi, T[cam])
            Tm[cam] = co.se3_left(-xi, T[cam])
            fd = (total(prob, Tp, X, intr, robust) - total(prob, Tm, X, intr, robust)) / (2 * eps)
            assert abs(fd - gp[cam, k]) <= 1e-5 * scale, (cam, k, fd, gp[cam, k])
    for q in range(0, prob.n_pt, 3):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[q, k] += eps
            Xm[q, k] -= eps
            fd = (total(prob, T, Xp, intr, robust) - total(prob, T, Xm, intr, robust)) / (2 * eps)
            assert abs(fd - gx[q, k]) <= 1e-5 * scale, (q, k, fd, gx[q, k])
    for k in range(8):
        ip, im = intr.copy(), intr.copy()
        ip[k] += eps
        im[k] -= eps
        fd = (total(prob, T, X, ip, robust) - total(prob, T, X, im, robust)) / (2 * eps)
        assert abs(fd - gi[k]) <= 1e-5 * scale * max(1.0, abs(gi[k]) / scale), (k, fd, gi[k])


def test_zero_distortion_matches_reference_edge_jacobians():
    prob, intr = small(seed=2, dist=False)
    T, X = prob.T_cw, prob.X
    delta = np.sqrt(5.991)
    c, gp, gx, gi = co.cost_and_grad(T, X, prob.obs_cam, prob.obs_pt, prob.obs_uv, intr, True, delta)
    gp_ref = np.zeros_like(gp)
    gx_ref = np.zeros_like(gx)
    K = tuple(intr[:4])
    for a in range(prob.n_obs):
        cam, q = prob.obs_cam[a], prob.obs_pt[a]
        qt = oracle.quat_from_R(T[cam, :9].reshape(3, 3))
        e, JX, Jxi, Xc = oracle.edge(qt, T[cam, 9:], X[q], K, prob.obs_uv[a])
        s = float(e @ e)
        w = 1.0 if s <= delta * delta else delta / np.sqrt(s)
        gp_ref[cam] += 2 * w * (Jxi.T @ e)
        gx_ref[q] += 2 * w * (JX.T @ e)
    # e = z - pi, so dC/dparam = 2 w e^T de/dparam with de/dparam = J (the edge Jacobians)
    np.testing.assert_allclose(gp, gp_ref, rtol=1e-9, atol=1e-9 * np.abs(gp_ref).max())
    np.testing.assert_allclose(gx, gx_ref, rtol=1e-9, atol=1e-9 * np.abs(gx_ref).max())


def test_se3_left_is_a_group_action():
    rng = np.random.default_rng(0)
    T = small()[0].T_cw[2]
    xi = rng.normal(size=6) * 0.3
    back = co.se3_left(-xi, co.se3_left(xi, T))
    np.testing.assert_allclose(back, T, atol=1e-12)
    # the rotation part equals the BA oracle's exp map
    q, t = oracle.se3_exp(xi)
    R = oracle.R_from_quat(q)
    np.testing.assert_allclose(co.se3_left(xi, np.r_[np.eye(3).reshape(9), 0, 0, 0])[:9], R.reshape(9), atol=1e-12)
    np.testing.assert_allclose(co.se3_left(xi, np.r_[np.eye(3).reshape(9), 0, 0, 0])[9:], t, atol=1e-12)
