"""GPU path vs the FP64 oracle on the same seeded problems (C ABI calls)."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
from parity_util import check_fp64_against_oracle, rel as _rel, say as _say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lam", [1e-4, 1.0])
def test_reduced_system_matches_oracle(solver, oracle_mod, lam):
    _say(f"reduced system workload0 n_pt=300 lambda={lam}")
    prob = problems.workload(0, seed=3, n_pt=300)
    opts = workload_options(0)
    solver.prepare(prob, opts)
    S, g = solver.debug_reduced_system(lam)
    So, go = oracle_mod.reduced_system(prob, opts, lam)
    assert S.shape == So.shape
    scale = np.abs(So).max()
    assert np.abs(S - So).max() <= 1e-11 * scale
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()


@pytest.mark.parametrize("workload,seed,over", [(0, 0, {}), (0, 7, {"n_pt": 400}), (1, 0, {})])
def test_fp64_solve_matches_oracle(solver, oracle_mod, workload, seed, over):
    _say(f"fp64 solve workload={workload} seed={seed} {over}")
    prob = problems.workload(workload, seed=seed, **over)
    opts = workload_options(workload)
    opts.precision = NAV_BA_FP64
    ref = oracle_mod.solve(prob, opts)
    got = solver.solve(prob, opts)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)


@pytest.mark.parametrize("mode", [1, 2])
def test_fp64_both_factorisations_match_oracle(solver, oracle_mod, mode):
    _say(f"fp64 solve workload=1 seed=2 factorisation mode {mode} (1 tiles, 2 blocks)")
    prob = problems.workload(1, seed=2, n_pt=6000)
    opts = workload_options(1)
    opts.precision = NAV_BA_FP64
    opts.n_iter = 8
    ref = oracle_mod.solve(prob, opts)
    solver.set_solver(mode)
    try:
        got = solver.solve(prob, opts)
    finally:
        solver.set_solver(0)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_fp64_every_factorisation_of_a_small_problem(solver, oracle_mod, mode):
    """22 free cameras: automatic (the per-window kernel in LDS), 64x64 tiles,
    6x6 blocks in global memory and the per-window kernel by request each
    reach the oracle's result."""
    _say(f"fp64 solve workload=1 n_cam=24 n_pt=3000 factorisation mode {mode}")
    prob = problems.workload(1, seed=5, n_cam=24, n_pt=3000)
    opts = workload_options(1)
    opts.precision = NAV_BA_FP64
    opts.n_iter = 8
    ref = oracle_mod.solve(prob, opts)
    solver.set_solver(mode)
    try:
        got = solver.solve(prob, opts)
    finally:
        solver.set_solver(0)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)


@pytest.mark.parametrize("kernel", [1, 2])
def test_window_factorisation_equals_a_batch_of_one(solver, kernel):
    """A single window through the per-window factorisation and the same
    window as a batch of one run the same kernels on the same sums (given the
    same Schur-row kernel and chunk): the results agree to the last bit."""
    prob = problems.workload(3, seed=6)
    opts = workload_options(3)
    opts.precision = NAV_BA_FP64
    _say(f"window 6 single vs batch of one, Schur-row kernel {kernel}")
    solver.set_schur_kernel(kernel)
    solver.set_schur_chunk(256)
    try:
        a = solver.solve(prob, opts)
        (b,) = solver.solve_batched([prob], opts)
    finally:
        solver.set_schur_kernel(0)
        solver.set_schur_chunk(0)
    _say(f"  single {a.final_cost!r} batch {b.final_cost!r}")
    assert a.iterations == b.iterations and a.trials == b.trials
    assert a.final_cost == b.final_cost
    assert np.array_equal(a.T_cw, b.T_cw) and np.array_equal(a.X, b.X)
    assert np.array_equal(a.outlier, b.outlier)


@pytest.mark.parametrize("n_cam", [24, 19, 20])
def test_fp64_tiled_partial_last_tile(solver, oracle_mod, n_cam):
    """Tiled factorisation whose last 64-row tile holds 4, 38 and 44 rows of
    the system (22, 17 and 18 free cameras): one, three and three of its four
    16-column stages are factored, the rest is identity padding."""
    _say(f"fp64 solve workload=1 n_cam={n_cam} n_pt=3000 tiled factorisation")
    prob = problems.workload(1, seed=8, n_cam=n_cam, n_pt=3000)
    opts = workload_options(1)
    opts.precision = NAV_BA_FP64
    opts.n_iter = 6
    ref = oracle_mod.solve(prob, opts)
    solver.set_solver(1)
    try:
        got = solver.solve(prob, opts)
    finally:
        solver.set_solver(0)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)


def test_loop_tiled_factorisation_fp64_and_mixed(solver, oracle_mod):
    """300-keyframe loop (298 free cameras, 1788 pose unknowns: the tiled
    multi-workgroup factorisation), fewer points so the oracle stays quick."""
    _say("loop workload=2 n_pt=12000 fp64 and mixed, 4 iterations")
    prob = problems.workload(2, seed=0, n_pt=12000)
    opts = workload_options(2)
    opts.precision = NAV_BA_FP64
    opts.n_iter = 4
    ref = oracle_mod.solve(prob, opts)
    got = solver.solve(prob, opts)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)
    opts.precision = NAV_BA_MIXED
    mix = solver.solve(prob, opts)
    _say(f"  mixed {mix.final_cost!r} oracle {ref.final_cost!r} rel {_rel(mix.final_cost, ref.final_cost):.2e}")
    assert _rel(mix.final_cost, ref.final_cost) <= 1e-5


def test_mixed_solve_close_to_oracle(solver, oracle_mod):
    _say("mixed solve workload0")
    prob = problems.workload(0, seed=1)
    opts = workload_options(0)
    opts.precision = NAV_BA_MIXED
    ref_opts = workload_options(0)
    ref = oracle_mod.solve(prob, ref_opts)
    got = solver.solve(prob, opts)
    _say(f"  cost mixed {got.final_cost!r} oracle {ref.final_cost!r}")
    assert _rel(got.final_cost, ref.final_cost) <= 1e-5
