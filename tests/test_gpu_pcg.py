"""Block-Jacobi preconditioned conjugate gradients on the reduced camera
system (solver mode 4; automatic above 16384 unknowns), on the GPU.

g2o's BlockSolver hands the reduced system to a linear solver; the reference
configures a Cholesky one (BE_GraphOptim.cpp:43-47).  Conjugate gradients stop
at a relative residual of 1e-10, so their step differs from the factored step
by up to the system's condition number times that, and the LM path moves with
it: the bars below are the factored path's widened to 1e-8 for the final cost
and the history (measured: at most 1.2e-11 and 2.5e-10 on these problems),
with the same schedule and flags."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, workload_options
from parity_util import rel, say

pytestmark = pytest.mark.gpu


def _hist_rel(a, b):
    n = min(len(a), len(b))
    return float(np.max(np.abs(np.asarray(a[:n]) - np.asarray(b[:n])) / np.abs(np.asarray(b[:n]))))


@pytest.mark.parametrize("workload,kw,iters", [(1, {"n_pt": 6000}, 8), (0, {}, 10), (1, {"n_cam": 24, "n_pt": 3000}, 6)])
def test_pcg_solve_matches_oracle(solver, oracle_mod, workload, kw, iters):
    prob = problems.workload(workload, seed=2, **kw)
    o = workload_options(workload)
    o.precision = NAV_BA_FP64
    o.n_iter = iters
    say(f"conjugate gradients: workload {workload} {kw}, {prob.n_obs} observations, {iters} iterations")
    ref = oracle_mod.solve(prob, o)
    solver.set_solver(4)
    try:
        got = solver.solve(prob, o)
        it, res = solver.pcg_stats()
    finally:
        solver.set_solver(0)
    say(f"  gpu {got.final_cost!r} oracle {ref.final_cost!r} rel {rel(got.final_cost, ref.final_cost):.2e}, "
        f"history {_hist_rel(got.cost_history, ref.cost_history):.2e}; last solve {it} iterations, residual {res:.1e}")
    assert 0 < it < 2000 and res <= 1e-10
    assert (got.iterations, got.trials, got.stop_reason) == (ref.iterations, ref.trials, ref.stop_reason)
    assert rel(got.final_cost, ref.final_cost) <= 1e-8
    assert _hist_rel(got.cost_history, ref.cost_history) <= 1e-8
    assert np.array_equal(got.outlier, ref.outlier)


def test_pcg_is_deterministic(solver):
    prob = problems.workload(1, seed=4, n_pt=4000)
    o = workload_options(1)
    o.precision = NAV_BA_FP64
    o.n_iter = 5
    solver.set_solver(4)
    try:
        a = solver.solve(prob, o)
        b = solver.solve(prob, o)
    finally:
        solver.set_solver(0)
    assert a.final_cost == b.final_cost and np.array_equal(a.T_cw, b.T_cw) and np.array_equal(a.X, b.X)


def test_pcg_is_chosen_above_the_dense_limit(solver):
    """2 800 keyframes (two fixed), 20 000 points of 2-8 views: 16 788 pose
    unknowns, beyond the dense factorisation's 16 384.  No oracle at this size
    (its dense Cholesky would take minutes): the solve must pick conjugate
    gradients, converge on the last trial, and lower the cost monotonically
    over its accepted iterations."""
    prob = problems.workload(3, seed=1, n_cam=2800, n_pt=20000)
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    o.n_iter = 4
    say(f"large system: {prob.n_cam} cameras, {prob.n_obs} observations")
    got = solver.solve(prob, o)
    it, res = solver.pcg_stats()
    say(f"  cost {got.initial_cost:.6e} -> {got.final_cost:.6e}, {got.iterations} iterations, {got.trials} trials; "
        f"last solve {it} iterations, residual {res:.1e}")
    assert it > 0 and res <= 1e-10
    h = np.asarray(got.cost_history)
    assert np.all(np.diff(h) <= 0) and h[-1] < h[0]
