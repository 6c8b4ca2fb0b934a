"""GPU path vs the FP64 oracle on the same seeded problems (C ABI calls)."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import Problem
from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, Options, workload_options

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _sensitivity(oracle_mod, prob, opts):
    """Relative change of the oracle's own final cost and cost history when a
    few camera translations move by one ulp: the problem's rounding sensitivity."""
    ref = oracle_mod.solve(prob, opts)
    base = ref.final_cost
    worst = worst_hist = 0.0
    for direction in (np.inf, -np.inf):
        q = Problem(**{k: getattr(prob, k) for k in (
            "T_cw", "X", "obs_cam", "obs_pt", "obs_uv", "fx", "fy", "cx", "cy", "cam_fixed", "pt_fixed")})
        q.T_cw = q.T_cw.copy()
        q.T_cw[q.n_cam // 2:, 9:] = np.nextafter(q.T_cw[q.n_cam // 2:, 9:], direction)
        r = oracle_mod.solve(q, opts)
        worst = max(worst, _rel(r.final_cost, base))
        n = min(len(r.cost_history), len(ref.cost_history))
        worst_hist = max(worst_hist, float(np.max(np.abs(r.cost_history[:n] - ref.cost_history[:n])
                                                  / ref.cost_history[:n])))
    return worst, worst_hist


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
    _say(f"  cost gpu {got.final_cost!r} oracle {ref.final_cost!r} iters {got.iterations}/{ref.iterations} "
         f"trials {got.trials}/{ref.trials} solve_ms {got.solve_ms:.2f}")
    assert got.iterations == ref.iterations
    assert got.trials == ref.trials
    _say(f"  rel cost diff {_rel(got.final_cost, ref.final_cost):.3e} "
         f"max rel history diff {np.max(np.abs(got.cost_history - ref.cost_history) / ref.cost_history):.3e} "
         f"max pose diff {np.abs(got.T_cw - ref.T_cw).max():.3e} flags diff {(got.outlier != ref.outlier).sum()}")
    # 1e-9 relative, or four times the problem's own 1-ulp sensitivity when the
    # problem is that ill-conditioned (two-view points pulled away by outliers)
    tol, tol_hist = 1e-9, 1e-9
    hist_diff = np.max(np.abs(got.cost_history - ref.cost_history) / ref.cost_history)
    if _rel(got.final_cost, ref.final_cost) > tol or hist_diff > tol_hist:
        sens, sens_hist = _sensitivity(oracle_mod, prob, opts)
        _say(f"  oracle 1-ulp sensitivity: final {sens:.3e}, history {sens_hist:.3e}")
        tol, tol_hist = max(tol, 4 * sens), max(tol_hist, 4 * sens_hist)
    assert _rel(got.final_cost, ref.final_cost) <= tol
    assert hist_diff <= tol_hist
    assert np.array_equal(got.outlier, ref.outlier)
    np.testing.assert_allclose(got.T_cw, ref.T_cw, atol=1e-6)


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
