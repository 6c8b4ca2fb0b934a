"""Problems in the reference's two call patterns of the back-end, solved with
the library's own defaults (nav_ba_default_options: compat FP32 roundings on,
robust off, delta (float)sqrt(5.99), lambda from 1e-5 max|diag H|):

  map initialisation  2 frames, the first fixed, every point seen by both,
                      10 iterations   (OP_MapInitialization.cpp:198-204)
  local window        11 frames, frame 0 and every point it sees fixed,
                      5 iterations    (FE_SlamMonoV.cpp:154-165)

The generator's values are FP32-representable; the inputs are moved off
float values (points, measurements, intrinsics), so the compat roundings of
BE_GraphOptim.cpp:147, Point2D.hpp:33 and GeometricCamera.h:117 change them."""
import numpy as np

from nav24_amd import problems
from nav24_amd.abi import Options, Problem


def off_float(prob, seed):
    """Inputs moved by a few parts in 1e9: no longer float values."""
    rng = np.random.default_rng(seed)
    X = prob.X * (1.0 + 3e-9 * rng.standard_normal(prob.X.shape))
    uv = prob.obs_uv + 1e-6 * rng.standard_normal(prob.obs_uv.shape)
    return Problem(T_cw=prob.T_cw, X=X, obs_cam=prob.obs_cam, obs_pt=prob.obs_pt, obs_uv=uv,
                   fx=prob.fx + 1.0 / 3.0, fy=prob.fy - 1.0 / 7.0, cx=prob.cx + 0.1, cy=prob.cy - 0.2,
                   cam_fixed=prob.cam_fixed, pt_fixed=prob.pt_fixed)


def keep_points(prob, keep):
    """The problem restricted to the points in `keep` (bool per point)."""
    new = -np.ones(prob.n_pt, np.int64)
    new[keep] = np.arange(int(keep.sum()))
    ob = keep[prob.obs_pt]
    return Problem(T_cw=prob.T_cw, X=prob.X[keep], obs_cam=prob.obs_cam[ob], obs_pt=new[prob.obs_pt[ob]],
                   obs_uv=prob.obs_uv[ob], fx=prob.fx, fy=prob.fy, cx=prob.cx, cy=prob.cy,
                   cam_fixed=prob.cam_fixed, pt_fixed=None if prob.pt_fixed is None else prob.pt_fixed[keep])


def map_init_problem(seed=0, n_pt=6000):
    p = problems.workload(1, seed=seed, n_cam=2, n_pt=n_pt, n_fixed=1)
    seen = np.zeros((2, p.n_pt), bool)
    seen[p.obs_cam, p.obs_pt] = True
    q = keep_points(p, seen[0] & seen[1])
    q.cam_fixed = np.array([1, 0], np.uint8)
    q.pt_fixed = np.zeros(q.n_pt, np.uint8)
    return off_float(q, seed)


def window_problem(seed=0, n_pt=5000, n_cam=11):
    p = problems.workload(1, seed=seed, n_cam=n_cam, n_pt=n_pt, n_fixed=1)
    p = keep_points(p, np.bincount(p.obs_pt, minlength=p.n_pt) > 0)
    cam_fixed = np.zeros(p.n_cam, np.uint8)
    cam_fixed[0] = 1
    pt_fixed = np.zeros(p.n_pt, np.uint8)
    pt_fixed[p.obs_pt[p.obs_cam == 0]] = 1
    p.cam_fixed, p.pt_fixed = cam_fixed, pt_fixed
    return off_float(p, seed + 100)


def prerounded(prob, opts):
    """The same problem with its inputs rounded to float beforehand and compat
    off: solving it is the compat-mode solve without the final FP32 rounding of
    the points (tested on the CPU for the oracle)."""
    f = lambda v: float(np.float32(v))
    q = Problem(T_cw=prob.T_cw, X=prob.X.astype(np.float32).astype(np.float64), obs_cam=prob.obs_cam,
                obs_pt=prob.obs_pt, obs_uv=prob.obs_uv.astype(np.float32).astype(np.float64),
                fx=f(prob.fx), fy=f(prob.fy), cx=f(prob.cx), cy=f(prob.cy),
                cam_fixed=prob.cam_fixed, pt_fixed=prob.pt_fixed)
    o = Options(**{k: getattr(opts, k) for k in ("n_iter", "robust", "huber_delta", "lambda_init", "precision",
                                                 "outlier_chi2", "max_trials")}, compat_fp32_io=False)
    return q, o


def unrounded_oracle(oracle_mod, prob, opts):
    """The oracle's compat-mode solve without the final FP32 rounding of the points."""
    return oracle_mod.solve(*prerounded(prob, opts))


def float_points_explained(got_X, ref_X, got_unrounded_X, ref_unrounded_X):
    """Returned points are float values.  Every coordinate where the library's
    float differs from the oracle's must be one where the two FP64 results
    round to different floats: the library's float is the rounding of its own
    FP64 point, the oracle's of its own.  Returns (number of float differences,
    number explained, largest difference of the FP64 points relative to the
    point's norm)."""
    g = np.asarray(got_X, np.float64)
    r = np.asarray(ref_X, np.float64)
    gu = np.asarray(got_unrounded_X, np.float64)
    ru = np.asarray(ref_unrounded_X, np.float64)
    rnd = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    assert np.array_equal(rnd(gu), g), "the library's float points are not its FP64 points rounded"
    assert np.array_equal(rnd(ru), r), "the oracle's float points are not its FP64 points rounded"
    diff = g != r
    n3 = lambda a: np.linalg.norm(a.reshape(-1, 3), axis=1)  # noqa: E731
    rel = float(np.max(n3(gu - ru) / np.maximum(n3(ru), 1e-300)))
    explained = diff & (rnd(gu) != rnd(ru))
    return int(diff.sum()), int(explained.sum()), rel


def float_points_equal(got_X, ref_X, ref_unrounded_X, rel=1e-10):
    """Returned points are float values; they must equal the oracle's bit for
    bit, except where the oracle's unrounded value lies within `rel` of a float
    rounding midpoint (both sides of such a value are then correct roundings
    of FP64 results that agree to `rel`); those may differ by one float ulp.
    Returns (number of differences, number of differences allowed)."""
    g = np.asarray(got_X, np.float64)
    r = np.asarray(ref_X, np.float64)
    diff = g != r
    if not diff.any():
        return 0, 0
    u = np.asarray(ref_unrounded_X, np.float64)[diff]
    gf, rf = g[diff].astype(np.float32), r[diff].astype(np.float32)
    one_ulp = (np.nextafter(rf, np.float32(np.inf)) == gf) | (np.nextafter(rf, np.float32(-np.inf)) == gf)
    mid = (gf.astype(np.float64) + rf.astype(np.float64)) / 2.0
    near = np.abs(u - mid) <= rel * np.abs(u)
    return int(diff.sum()), int((one_ulp & near).sum())
