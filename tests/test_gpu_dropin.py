"""The drop-in's own default path on the GPU: nav_ba_solve with
nav_ba_default_options() (compat FP32 roundings on, robust off, delta
(float)sqrt(5.99), automatic lambda), on the reference's two call patterns and
on workloads 0 and 1, against the FP64 oracle in the same mode.  Returned
points are float values and must equal the oracle's bit for bit (one float ulp
only where the FP64 value sits on a rounding midpoint)."""
import numpy as np
import pytest

from dropin_cases import (float_points_equal, float_points_explained, map_init_problem, off_float, prerounded,
                          unrounded_oracle, window_problem)
from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, workload_options
from parity_util import check_fp64_against_oracle, say

pytestmark = pytest.mark.gpu


def _check_points(solver, oracle_mod, prob, opts, got, ref):
    """Float points bit-equal to the oracle's; where they are not, the library's
    and the oracle's FP64 points (the same solves on inputs rounded beforehand,
    compat off) must straddle a float rounding boundary and agree to 1e-6 of
    the point's norm."""
    assert np.array_equal(got.X.astype(np.float32).astype(np.float64), got.X)
    u = unrounded_oracle(oracle_mod, prob, opts)
    nd, nok = float_points_equal(got.X, ref.X, u.X)
    if nd == nok:
        say(f"  points: {nd} differ from the oracle's float values, {nok} of them at a rounding midpoint")
        return
    gu = solver.solve(*prerounded(prob, opts))
    nd, nex, rel = float_points_explained(got.X, ref.X, gu.X, u.X)
    say(f"  points: {nd} of {got.X.size} float values differ from the oracle's; {nex} are FP64 points that agree to "
        f"{rel:.2e} and straddle a float rounding boundary")
    assert nd == nex and rel <= 1e-6


@pytest.mark.parametrize("case", ["map_init", "window"])
def test_default_options_on_reference_call_patterns(solver, oracle_mod, case):
    from nav24_amd.ba import default_options
    o = default_options()
    assert o.compat_fp32_io and not o.robust and o.lambda_init <= 0 and o.precision == NAV_BA_FP64
    assert o.huber_delta == float(np.float32(np.sqrt(5.99)))
    if case == "map_init":
        prob, o.n_iter = map_init_problem(), 10
    else:
        prob = window_problem()
        assert o.n_iter == 5
    say(f"default options, {case}: {prob.n_cam} cameras, {prob.n_pt} points, {prob.n_obs} observations, "
        f"{o.n_iter} iterations")
    ref = oracle_mod.solve(prob, o)
    got = solver.solve(prob, o)
    check_fp64_against_oracle(oracle_mod, prob, o, got, ref, label=case)
    _check_points(solver, oracle_mod, prob, o, got, ref)


@pytest.mark.parametrize("workload", [0, 1])
def test_compat_roundings_on_workloads(solver, oracle_mod, workload):
    prob = off_float(problems.workload(workload, seed=0), 7 + workload)
    o = workload_options(workload)
    o.precision = NAV_BA_FP64
    o.compat_fp32_io = True
    say(f"compat roundings on workload {workload}: {prob.n_obs} observations")
    ref = oracle_mod.solve(prob, o)
    got = solver.solve(prob, o)
    check_fp64_against_oracle(oracle_mod, prob, o, got, ref, label=f"workload {workload}")
    _check_points(solver, oracle_mod, prob, o, got, ref)
