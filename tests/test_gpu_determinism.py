"""Reproducibility of the FP64 path: every device sum is in a fixed order, so
the same prepared problem solved twice gives the same bits (poses, points,
cost history, flags), for one problem and for a batch of windows."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
from parity_util import say

pytestmark = pytest.mark.gpu


def _same(a, b):
    return (np.array_equal(a.T_cw, b.T_cw) and np.array_equal(a.X, b.X) and
            np.array_equal(a.cost_history, b.cost_history) and np.array_equal(a.outlier, b.outlier) and
            a.final_cost == b.final_cost and a.trials == b.trials)


@pytest.mark.parametrize("prec,kernel", [(NAV_BA_FP64, 1), (NAV_BA_MIXED, 1), (NAV_BA_FP64, 2)])
def test_single_problem_runs_are_bit_identical(solver, prec, kernel):
    prob = problems.workload(1, seed=0)
    o = workload_options(1)
    o.precision = prec
    o.n_iter = 6
    say(f"determinism: workload 1, precision {prec}, Schur-row kernel {kernel}, two runs of one prepare "
        "and a second prepare")
    solver.set_schur_kernel(kernel)
    try:
        solver.prepare(prob, o)
        solver.run()
        a = solver.fetch()
        solver.run()
        b = solver.fetch()
        c = solver.solve(prob, o)
    finally:
        solver.set_schur_kernel(0)
    say(f"  final costs {a.final_cost!r} {b.final_cost!r} {c.final_cost!r}")
    assert _same(a, b) and _same(a, c)


@pytest.mark.parametrize("kernel", [1, 2])
def test_batched_runs_are_bit_identical(solver, kernel):
    wins = [problems.workload(3, seed=s) for s in range(12)]
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    say(f"determinism: 12 batched windows, FP64, Schur-row kernel {kernel}, two runs")
    solver.set_schur_kernel(kernel)
    try:
        solver.prepare_batched(wins, o)
        solver.run()
        a = solver.fetch_batched()
        solver.run()
        b = solver.fetch_batched()
    finally:
        solver.set_schur_kernel(0)
    bad = [i for i in range(len(wins)) if not _same(a[i], b[i])]
    say(f"  windows that differ: {bad}")
    assert not bad
