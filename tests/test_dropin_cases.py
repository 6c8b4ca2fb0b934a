"""CPU checks of the drop-in call-pattern problems and of the helpers the GPU
tests use to compare FP32-rounded points (tests/dropin_cases.py)."""
import numpy as np

from dropin_cases import float_points_equal, map_init_problem, unrounded_oracle, window_problem
from nav24_amd.abi import Options
from oracle import oracle


def _opts(n_iter):
    o = Options()  # the library's defaults: compat on, robust off, automatic lambda
    o.n_iter = n_iter
    return o


def test_cases_have_the_reference_shape():
    m = map_init_problem()
    assert m.n_cam == 2 and list(m.cam_fixed) == [1, 0] and not m.pt_fixed.any()
    assert np.all(np.bincount(m.obs_pt, minlength=m.n_pt) == 2)
    w = window_problem()
    assert w.n_cam == 11 and w.cam_fixed[0] == 1 and not w.cam_fixed[1:].any()
    seen0 = np.zeros(w.n_pt, bool)
    seen0[w.obs_pt[w.obs_cam == 0]] = True
    assert np.array_equal(w.pt_fixed.astype(bool), seen0) and 0 < seen0.sum() < w.n_pt


def test_inputs_are_off_float_so_compat_rounding_acts():
    for prob, n in ((map_init_problem(), 10), (window_problem(), 5)):
        assert not np.array_equal(prob.X.astype(np.float32).astype(np.float64), prob.X)
        assert not np.array_equal(prob.obs_uv.astype(np.float32).astype(np.float64), prob.obs_uv)
        assert float(np.float32(prob.fx)) != prob.fx
        o = _opts(n)
        a = oracle.solve(prob, o)
        o.compat_fp32_io = False
        b = oracle.solve(prob, o)
        assert a.final_cost != b.final_cost
        assert np.array_equal(a.X.astype(np.float32).astype(np.float64), a.X)


def test_unrounded_oracle_rounds_to_the_compat_points():
    for prob, n in ((map_init_problem(), 10), (window_problem(), 5)):
        o = _opts(n)
        r = oracle.solve(prob, o)
        u = unrounded_oracle(oracle, prob, o)
        assert np.array_equal(u.X.astype(np.float32).astype(np.float64), r.X)
        assert np.array_equal(u.T_cw, r.T_cw) and u.final_cost == r.final_cost
        assert np.array_equal(u.outlier, r.outlier)


def test_float_points_equal_allows_only_midpoint_ties():
    x = np.float32(1.5)
    up = np.nextafter(x, np.float32(2))
    mid = (float(x) + float(up)) / 2
    ref = np.array([float(x), 2.0])
    assert float_points_equal(ref.copy(), ref, ref) == (0, 0)
    # one ulp apart at a value that sits on the midpoint: allowed
    assert float_points_equal(np.array([float(up), 2.0]), ref, np.array([mid, 2.0])) == (1, 1)
    # one ulp apart far from the midpoint: not allowed
    assert float_points_equal(np.array([float(up), 2.0]), ref, np.array([float(x), 2.0])) == (1, 0)
    # two ulps apart: never allowed
    up2 = np.nextafter(up, np.float32(2))
    assert float_points_equal(np.array([float(up2), 2.0]), ref, np.array([mid, 2.0])) == (1, 0)
