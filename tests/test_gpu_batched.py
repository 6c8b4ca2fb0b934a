"""Batched windows (nav_ba_solve_batched) vs the FP64 oracle, window by window.

A batch is the loop of independent per-window solves of FE_SlamMonoV.cpp:161-165
run as one device solve; every window must reproduce what the oracle gives
for that window alone: same iterations, trials and stop reason (its own LM
schedule), final cost and cost history at the FP64 bar, identical outlier
flags.  Windows are uneven (5-30 keyframes, 300-5000 points).
"""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import (NAV_BA_EMPTY, NAV_BA_FP64, NAV_BA_MIXED, NAV_BA_STOP_EMPTY, NAV_BA_TOO_LARGE, Problem,
                           workload_options)
from nav24_amd.ba import BAError
from parity_util import check_fp64_against_oracle, rel, say

pytestmark = pytest.mark.gpu


def windows(seeds):
    return [problems.workload(3, seed=s) for s in seeds]


@pytest.mark.parametrize("kernel", [1, 2])
def test_batched_fp64_windows_match_oracle(solver, oracle_mod, kernel):
    say(f"batched fp64 workload3 seeds 0..7, Schur-row kernel {kernel}")
    wins = windows(range(8))
    opts = workload_options(3)
    opts.precision = NAV_BA_FP64
    solver.set_schur_kernel(kernel)
    try:
        got = solver.solve_batched(wins, opts)
        assert solver.schur_kernel == kernel
    finally:
        solver.set_schur_kernel(0)
    for w, (prob, g) in enumerate(zip(wins, got)):
        ref = oracle_mod.solve(prob, opts)
        check_fp64_against_oracle(oracle_mod, prob, opts, g, ref, label=f"window {w} ({prob.n_cam} kf, {prob.n_pt} pts)")


def test_batched_auto_lambda_and_robust_off(solver, oracle_mod):
    say("batched fp64 auto lambda, robust off, workload3 seeds 20..23")
    wins = windows(range(20, 24))
    opts = workload_options(3)
    opts.precision = NAV_BA_FP64
    opts.lambda_init = 0.0
    opts.robust = False
    opts.n_iter = 6
    got = solver.solve_batched(wins, opts)
    for w, (prob, g) in enumerate(zip(wins, got)):
        ref = oracle_mod.solve(prob, opts)
        check_fp64_against_oracle(oracle_mod, prob, opts, g, ref, label=f"window {w}")


def test_batched_mixed_close_to_fp64(solver, oracle_mod):
    say("batched mixed workload3 seeds 0..7")
    wins = windows(range(8))
    opts = workload_options(3)
    assert opts.precision == NAV_BA_MIXED
    got = solver.solve_batched(wins, opts)
    ref_opts = workload_options(3)
    ref_opts.precision = NAV_BA_FP64
    for w, (prob, g) in enumerate(zip(wins, got)):
        ref = oracle_mod.solve(prob, ref_opts)
        say(f"  window {w} mixed {g.final_cost!r} oracle {ref.final_cost!r} rel {rel(g.final_cost, ref.final_cost):.2e}")
        assert rel(g.final_cost, ref.final_cost) <= 1e-5


def test_batched_with_empty_and_camera_only_windows(solver, oracle_mod):
    say("batched with an empty window and an all-points-fixed window")
    a, b = windows([30, 31])
    empty = Problem(T_cw=a.T_cw[:3].copy(), X=a.X[:10].copy(), obs_cam=np.zeros(0, np.int32),
                    obs_pt=np.zeros(0, np.int32), obs_uv=np.zeros((0, 2)))
    b.pt_fixed = np.ones(b.n_pt, np.uint8)
    opts = workload_options(3)
    opts.precision = NAV_BA_FP64
    got = solver.solve_batched([a, empty, b], opts)
    assert got[1].status == NAV_BA_EMPTY and got[1].stop_reason == NAV_BA_STOP_EMPTY
    assert np.array_equal(got[1].T_cw, empty.T_cw) and np.array_equal(got[1].X, empty.X)
    for w, prob in ((0, a), (2, b)):
        ref = oracle_mod.solve(prob, opts)
        check_fp64_against_oracle(oracle_mod, prob, opts, got[w], ref, label=f"window {w}")
    np.testing.assert_array_equal(got[2].X, b.X)


def test_batched_rejects_window_over_30_free_cameras(solver):
    say("batched window with 40 keyframes -> TOO_LARGE")
    p = problems.params_for(0, 3)
    p.n_cam, p.n_pt = 40, 600
    big = problems.make(p)
    opts = workload_options(3)
    with pytest.raises(BAError) as e:
        solver.solve_batched(windows([0]) + [big], opts)
    assert e.value.status == NAV_BA_TOO_LARGE


def test_batched_prepare_once_run_twice(solver, oracle_mod):
    say("batched prepare once, run twice")
    wins = windows(range(40, 44))
    opts = workload_options(3)
    opts.precision = NAV_BA_FP64
    solver.prepare_batched(wins, opts)
    solver.run()
    first = solver.fetch_batched()
    solver.run()
    second = solver.fetch_batched()
    # every run restarts from the uploaded initial state; the off-diagonal
    # Schur blocks are summed with LDS atomics, so runs may differ in the last
    # bits, and each run must meet the oracle bar on its own
    for w, prob in enumerate(wins):
        ref = oracle_mod.solve(prob, opts)
        for run, res in (("first", first[w]), ("second", second[w])):
            check_fp64_against_oracle(oracle_mod, prob, opts, res, ref, label=f"window {w} {run} run")
