"""Failure paths through the C ABI on the GPU: calls out of order, bad
arguments, empty problems, and inputs that make every LM trial fail.

A NaN measurement makes the initial cost NaN; g2o's rule then rejects every
trial (rho is NaN, so neither rho > 0 nor rho < 0 holds) and each iteration
ends after one rejected trial (OptimizationAlgorithmLevenberg, restated in the
oracle's solve_impl).  The GPU must take the same path: the same iterations,
trials and stop reason, the inputs returned unchanged, and the other windows
of a batch unaffected."""
import math

import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import (NAV_BA_BAD_ARG, NAV_BA_EMPTY, NAV_BA_FP64, NAV_BA_NOT_PREPARED, Problem,
                           workload_options)
from nav24_amd.ba import BAError, GraphOptimHip
from parity_util import check_fp64_against_oracle, say

pytestmark = pytest.mark.gpu


def test_run_and_fetch_before_prepare():
    s = GraphOptimHip(0)
    try:
        with pytest.raises(BAError) as e:
            s.run()
        assert e.value.status == NAV_BA_NOT_PREPARED
        with pytest.raises(BAError) as e:
            s.fetch()
        assert e.value.status == NAV_BA_NOT_PREPARED
    finally:
        s.close()


def test_bad_arguments_are_rejected(solver):
    prob = problems.workload(0, seed=0)
    o = workload_options(0)
    bad = Problem(T_cw=prob.T_cw, X=prob.X, obs_cam=prob.obs_cam.copy(), obs_pt=prob.obs_pt, obs_uv=prob.obs_uv)
    bad.obs_cam[3] = prob.n_cam  # one past the last camera
    with pytest.raises(BAError) as e:
        solver.solve(bad, o)
    assert e.value.status == NAV_BA_BAD_ARG
    o.precision = 7
    with pytest.raises(BAError) as e:
        solver.solve(prob, o)
    assert e.value.status == NAV_BA_BAD_ARG
    for call in (lambda: solver.set_solver(5), lambda: solver.set_schur_kernel(3), lambda: solver.set_schur_chunk(100)):
        with pytest.raises(BAError) as e:
            call()
        assert e.value.status == NAV_BA_BAD_ARG
    # the handle still works after the rejected calls
    o.precision = NAV_BA_FP64
    assert solver.solve(prob, o).status == 0


def test_empty_problem_returns_its_inputs(solver):
    prob = problems.workload(0, seed=1)
    empty = Problem(T_cw=prob.T_cw, X=prob.X, obs_cam=np.zeros(0, np.int32), obs_pt=np.zeros(0, np.int32),
                    obs_uv=np.zeros((0, 2)))
    o = workload_options(0)
    r = solver.solve(empty, o)
    assert r.status == NAV_BA_EMPTY and r.iterations == 0
    assert np.array_equal(r.T_cw, prob.T_cw) and np.array_equal(r.X, prob.X)


def _with_nan_measurement(prob, k=5):
    uv = prob.obs_uv.copy()
    uv[k, 0] = np.nan
    return Problem(T_cw=prob.T_cw, X=prob.X, obs_cam=prob.obs_cam, obs_pt=prob.obs_pt, obs_uv=uv,
                   cam_fixed=prob.cam_fixed, pt_fixed=prob.pt_fixed)


@pytest.mark.parametrize("solver_mode", [0, 1, 4])
def test_nan_measurement_rejects_every_trial(solver, oracle_mod, solver_mode):
    """Every trial fails (the factorisation meets a non-finite pivot, or the
    cost is NaN): the GPU keeps the oracle's schedule and returns the inputs,
    through the per-window kernel (mode 0 on this 10-keyframe problem), the tiles
    and conjugate gradients."""
    prob = _with_nan_measurement(problems.workload(0, seed=1))
    o = workload_options(0)
    o.precision = NAV_BA_FP64
    ref = oracle_mod.solve(prob, o)
    say(f"NaN measurement, solver mode {solver_mode}: oracle {ref.iterations} iterations, {ref.trials} trials")
    solver.set_solver(solver_mode)
    try:
        got = solver.solve(prob, o)
    finally:
        solver.set_solver(0)
    assert (got.iterations, got.trials, got.stop_reason) == (ref.iterations, ref.trials, ref.stop_reason)
    assert math.isnan(got.final_cost) and math.isnan(ref.final_cost)
    assert np.array_equal(got.T_cw, ref.T_cw) and np.array_equal(got.X, ref.X)


def test_nan_window_leaves_the_batch_alone(solver, oracle_mod):
    wins = [problems.workload(3, seed=s) for s in range(3)]
    wins[1] = _with_nan_measurement(wins[1])
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    got = solver.solve_batched(wins, o)
    for i in (0, 2):
        check_fp64_against_oracle(oracle_mod, wins[i], o, got[i], oracle_mod.solve(wins[i], o), label=f"window {i}")
    ref = oracle_mod.solve(wins[1], o)
    assert (got[1].iterations, got[1].trials) == (ref.iterations, ref.trials)
    assert math.isnan(got[1].final_cost)
