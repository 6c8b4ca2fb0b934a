"""Seeded problem generator: known-answer draws and workload shapes (CPU)."""
import json
import os

import numpy as np

from nav24_amd import problems

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_mt19937_64_standard_known_answer():
    # the C++ standard's check value: 10000th draw of a default-seeded mt19937_64
    assert problems.mt_draw(5489, 9999) == 9981545732273789042


def test_box_muller_matches_closed_form_of_raw_draws():
    d = [problems.mt_draw(0, k) for k in range(4)]
    u = [(x >> 11) * 2.0**-53 for x in d]
    z = []
    for i in (0, 2):
        r = np.sqrt(-2.0 * np.log(1.0 - u[i]))
        a = 2.0 * np.pi * u[i + 1]
        z += [r * np.cos(a), r * np.sin(a)]
    assert np.allclose(problems.normals(0, 4), z, rtol=1e-15, atol=1e-15)


def test_rng_fixture():
    kat = json.load(open(os.path.join(GOLD, "rng_kat.json")))
    assert [str(problems.mt_draw(0, k)) for k in range(3)] == kat["mt19937_64_seed0_draws_0_1_2"]
    assert np.array_equal(problems.normals(0, 8), np.array(kat["normals_seed0_first8"]))


def test_workload0_is_reproducible_and_matches_fixture():
    a, b = problems.workload(0), problems.workload(0)
    for k in ("T_cw", "X", "obs_cam", "obs_pt", "obs_uv"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    g = json.load(open(os.path.join(GOLD, "workload0_oracle.json")))
    assert (a.n_cam, a.n_pt, a.n_obs) == (g["n_cam"], g["n_pt"], g["n_obs"])
    assert int(a.obs_outlier.sum()) == g["n_generated_outliers"]
    assert a.obs_uv.sum() == g["uv_sum"] and a.X.sum() == g["X_sum"] and a.T_cw.sum() == g["T_sum"]


def test_generated_geometry_is_consistent():
    p = problems.workload(0, seed=4, n_pt=200)
    # ground-truth projections of inlier observations are within noise
    R = p.T_gt[:, :9].reshape(-1, 3, 3)
    t = p.T_gt[:, 9:]
    Xc = np.einsum("nij,nj->ni", R[p.obs_cam], p.X_gt[p.obs_pt]) + t[p.obs_cam]
    uv = np.stack([450 * Xc[:, 0] / Xc[:, 2] + 320, 450 * Xc[:, 1] / Xc[:, 2] + 240], 1)
    inl = p.obs_outlier == 0
    err = np.linalg.norm(uv[inl] - p.obs_uv[inl], axis=1)
    assert np.all(Xc[:, 2] > 0.1) and np.median(err) < 2.0 and err.max() < 8.0
    # rotations are orthonormal; fixed cameras are unperturbed
    assert np.allclose(np.einsum("nij,nkj->nik", R, R), np.eye(3), atol=1e-12)
    assert np.array_equal(p.T_cw[:2], p.T_gt[:2]) and np.all(p.cam_fixed[:2] == 1)
    # values handed to the solver are FP32-representable
    assert np.array_equal(p.obs_uv, p.obs_uv.astype(np.float32).astype(np.float64))
    assert np.array_equal(p.X, p.X.astype(np.float32).astype(np.float64))


def test_batched_window_sizes_in_range():
    for w in range(6):
        pr = problems.params_for(3, 100 + w)
        assert 5 <= pr.n_cam <= 30 and 300 <= pr.n_pt <= 5000
    p = problems.workload(3, seed=100, n_pt=300)
    k = np.bincount(p.obs_pt, minlength=p.n_pt)
    assert k.min() >= 2 and k.max() <= min(8, p.n_cam)
