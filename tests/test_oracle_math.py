"""Pins of the oracle's closed-form pieces (CPU only).

Each check is chosen so that a plausible mistake (a dropped term, a sign, a
transposed operand) fails it: Jacobians against FP64 central differences of
the residual under the exp-map perturbation the update uses, SE3 exp against
an independent Rodrigues formula, quaternion conversions through round trips
and special rotations, Huber through its piecewise closed form and continuity.
"""
import numpy as np
import pytest

from oracle import oracle

K = (450.0, 450.0, 320.0, 240.0)


def rodrigues(w):
    th = np.linalg.norm(w)
    Kx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-12:
        return np.eye(3) + Kx
    return np.eye(3) + np.sin(th) / th * Kx + (1 - np.cos(th)) / th**2 * Kx @ Kx


def random_pose(rng):
    R = rodrigues(rng.normal(size=3))
    q = oracle.quat_from_R(R)
    q = q * np.sign(q[3]) / np.linalg.norm(q)
    return q, rng.normal(size=3)


def test_project_special_case_identity_pose():
    q = np.array([0, 0, 0, 1.0]); t = np.zeros(3)
    e, JX, Jxi, Xc = oracle.edge(q, t, np.array([0, 0, 3.0]), K, np.array([320.0, 240.0]))
    assert np.allclose(e, 0) and np.allclose(Xc, [0, 0, 3])
    # R = I: J_X = -Jpi
    Jpi = np.array([[450 / 3, 0, 0], [0, 450 / 3, 0]])
    assert np.allclose(JX, -Jpi)


def test_residual_closed_form():
    rng = np.random.default_rng(0)
    q, t = random_pose(rng)
    X = rng.normal(size=3) + np.array([0, 0, 8.0])
    uv = np.array([300.0, 200.0])
    e, _, _, Xc = oracle.edge(q, t, X, K, uv)
    R = oracle.R_from_quat(q)
    xc = R @ X + t
    assert np.allclose(Xc, xc, rtol=1e-14)
    assert np.allclose(e, uv - np.array([450 * xc[0] / xc[2] + 320, 450 * xc[1] / xc[2] + 240]), rtol=1e-13)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jacobians_match_central_differences(seed):
    rng = np.random.default_rng(seed)
    q, t = random_pose(rng)
    R = oracle.R_from_quat(q)
    X = R.T @ (np.array([0.3, -0.2, 6.0]) + rng.normal(size=3) * 0.5 - t)
    uv = np.array([310.0, 250.0])
    e0, JX, Jxi, _ = oracle.edge(q, t, X, K, uv)
    eps = 1e-6
    num_X = np.zeros((2, 3))
    for k in range(3):
        d = np.zeros(3); d[k] = eps
        ep = oracle.edge(q, t, X + d, K, uv)[0]
        em = oracle.edge(q, t, X - d, K, uv)[0]
        num_X[:, k] = (ep - em) / (2 * eps)
    num_xi = np.zeros((2, 6))
    for k in range(6):
        d = np.zeros(6); d[k] = eps
        qp, tp = oracle.se3_oplus(d, q, t)
        qm, tm = oracle.se3_oplus(-d, q, t)
        num_xi[:, k] = (oracle.edge(qp, tp, X, K, uv)[0] - oracle.edge(qm, tm, X, K, uv)[0]) / (2 * eps)
    scale = np.abs(JX).max()
    assert np.abs(JX - num_X).max() <= 1e-6 * scale
    assert np.abs(Jxi - num_xi).max() <= 1e-6 * np.abs(Jxi).max()


@pytest.mark.parametrize("w", [[0.3, -0.2, 0.5], [1e-3, 2e-3, -1e-3], [2.5, 0.1, -1.0]])
def test_se3_exp_matches_rodrigues(w):
    xi = np.array(list(w) + [0.1, -0.4, 0.25])
    q, t = oracle.se3_exp(xi)
    R = oracle.R_from_quat(q)
    w = np.array(w)
    assert np.allclose(R, rodrigues(w), atol=1e-14)
    th = np.linalg.norm(w)
    Kx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    V = np.eye(3) + (1 - np.cos(th)) / th**2 * Kx + (th - np.sin(th)) / th**3 * Kx @ Kx
    assert np.allclose(t, V @ xi[3:], atol=1e-14)
    assert q[3] >= 0 and abs(np.linalg.norm(q) - 1) < 1e-15


def test_se3_exp_zero_and_inverse():
    q, t = oracle.se3_exp(np.zeros(6))
    assert np.array_equal(q, [0, 0, 0, 1.0]) and np.array_equal(t, np.zeros(3))
    xi = np.array([0.4, -0.1, 0.2, 0, 0, 0])
    q1, t1 = oracle.se3_oplus(xi, np.array([0, 0, 0, 1.0]), np.zeros(3))
    q2, t2 = oracle.se3_oplus(-xi, q1, t1)
    assert np.allclose(q2, [0, 0, 0, 1], atol=1e-15) and np.allclose(t2, 0, atol=1e-15)


def test_se3_small_angle_branch_is_the_published_one():
    # below 1e-5 the update uses R = I + W + W^2 (no 1/2 on W^2)
    w = np.array([4e-6, -3e-6, 2e-6])
    q, _ = oracle.se3_exp(np.concatenate([w, np.zeros(3)]))
    Kx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    R_g2o = np.eye(3) + Kx + Kx @ Kx
    qn = oracle.quat_from_R(R_g2o)
    qn = qn / np.linalg.norm(qn)
    assert np.allclose(q, qn, atol=1e-17)


@pytest.mark.parametrize("R", [
    np.eye(3),
    np.diag([1.0, -1.0, -1.0]),          # 180 deg about x: trace <= 0 branch, i = 0
    np.diag([-1.0, 1.0, -1.0]),          # i = 1
    np.diag([-1.0, -1.0, 1.0]),          # i = 2
    rodrigues(np.array([0.0, 3.1, 0.05])),
])
def test_quaternion_conversions(R):
    q = oracle.quat_from_R(R)
    assert abs(np.linalg.norm(q) - 1) < 1e-15
    assert np.allclose(oracle.R_from_quat(q), R, atol=1e-15)


def test_huber_closed_form_and_continuity():
    d = np.sqrt(5.991)
    for s in [0.0, 1.0, 5.0, d * d, 7.0, 400.0]:
        rho = oracle.huber(s, d)
        if s <= d * d:
            assert np.array_equal(rho, [s, 1.0, 0.0])
        else:
            assert np.isclose(rho[0], 2 * np.sqrt(s) * d - d * d) and np.isclose(rho[1], d / np.sqrt(s))
    lo, hi = oracle.huber(d * d, d), oracle.huber(d * d * (1 + 1e-12), d)
    assert abs(lo[0] - hi[0]) < 1e-9 and abs(lo[1] - hi[1]) < 1e-9


def test_inv3():
    rng = np.random.default_rng(4)
    A = rng.normal(size=(3, 3)); A = A @ A.T + 3 * np.eye(3)
    assert np.allclose(oracle.inv3(A) @ A, np.eye(3), atol=1e-14)
