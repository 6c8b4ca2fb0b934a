"""The cached index of nav_ba_prepare (and nav_ba_solve): a problem with the
same structure as the last one on the handle (sizes, observation indices,
fixed flags, camera, options) keeps the layout, work items, tuned choices and
step graphs and only uploads the new poses, points and measurements.  Its
results must be the bits a fresh handle gives for the new problem; any change
of structure or options must take the full prepare."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, Problem, workload_options
from nav24_amd.ba import GraphOptimHip
from parity_util import say

pytestmark = pytest.mark.gpu


def moved(p, seed, dx=0.02, duv=0.3):
    """The same structure with new initial poses' translations, points and measurements."""
    rng = np.random.default_rng(seed)
    T = p.T_cw.copy()
    T[:, 9:] += rng.normal(0, dx, (p.n_cam, 3))
    return Problem(T_cw=T, X=p.X + rng.normal(0, dx, p.X.shape), obs_cam=p.obs_cam.copy(), obs_pt=p.obs_pt.copy(),
                   obs_uv=p.obs_uv + rng.normal(0, duv, p.obs_uv.shape), fx=p.fx, fy=p.fy, cx=p.cx, cy=p.cy,
                   cam_fixed=p.cam_fixed, pt_fixed=p.pt_fixed, camera_model=p.camera_model, kb=p.kb)


def same(a, b):
    return (np.array_equal(a.T_cw, b.T_cw) and np.array_equal(a.X, b.X) and
            np.array_equal(a.cost_history, b.cost_history) and np.array_equal(a.outlier, b.outlier))


@pytest.fixture
def handle(solver):  # a handle of its own: the shared one may hold any earlier structure
    h = GraphOptimHip(0)
    yield h
    h.close()


@pytest.mark.parametrize("compat", [False, True])
def test_same_structure_reuses_the_index_and_gives_a_fresh_handles_bits(handle, compat):
    solver = handle
    a = problems.workload(3, seed=5)
    b = moved(a, 1)
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    o.compat_fp32_io = compat
    say(f"cached index: window 5 ({a.n_obs} observations), compat {compat}")
    solver.solve(a, o)
    assert not solver.prepare_cached
    got = solver.solve(b, o)
    assert solver.prepare_cached
    fresh = GraphOptimHip(0)
    try:
        ref = fresh.solve(b, o)
    finally:
        fresh.close()
    assert same(got, ref)


def test_structure_or_option_changes_take_the_full_prepare(handle):
    solver = handle
    a = problems.workload(3, seed=6)
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    solver.solve(a, o)
    solver.solve(moved(a, 2), o)
    assert solver.prepare_cached
    b = moved(a, 3)
    b.obs_pt = b.obs_pt.copy()
    b.obs_pt[7] = (b.obs_pt[7] + 1) % b.n_pt          # one observation on another point
    c = moved(a, 4)
    c.cam_fixed = np.zeros(c.n_cam, np.uint8)          # other fixed cameras
    for prob, opts, what in ((b, o, "an observation index"), (c, o, "the fixed flags")):
        got = solver.solve(prob, opts)
        assert not solver.prepare_cached, what
        fresh = GraphOptimHip(0)
        try:
            assert same(got, fresh.solve(prob, opts)), what
        finally:
            fresh.close()
    o2 = workload_options(3)
    o2.precision = NAV_BA_FP64
    o2.n_iter = 7
    solver.solve(c, o2)
    assert not solver.prepare_cached, "the options"


def test_batched_same_structure(handle):
    solver = handle
    wins = [problems.workload(3, seed=s) for s in range(4)]
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    solver.solve_batched(wins, o)
    new = [moved(w, 10 + i) for i, w in enumerate(wins)]
    got = solver.solve_batched(new, o)
    assert solver.prepare_cached
    fresh = GraphOptimHip(0)
    try:
        ref = fresh.solve_batched(new, o)
    finally:
        fresh.close()
    assert all(same(g, r) for g, r in zip(got, ref))
