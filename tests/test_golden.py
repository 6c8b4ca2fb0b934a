"""The oracle reproduces its committed outputs (regression of the checker)."""
import json
import os

import numpy as np

from nav24_amd import problems
from nav24_amd.abi import Problem, workload_options
from oracle import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_tiny_problem_fixture():
    d = np.load(os.path.join(GOLD, "tiny_problem.npz"))
    prob = Problem(T_cw=d["T_cw"], X=d["X"], obs_cam=d["obs_cam"], obs_pt=d["obs_pt"], obs_uv=d["obs_uv"],
                   cam_fixed=d["cam_fixed"], pt_fixed=d["pt_fixed"])
    r, lam, trials = oracle.solve(prob, workload_options(0), with_history=True)
    assert np.allclose(r.T_cw, d["out_T_cw"], rtol=0, atol=1e-12)
    assert np.allclose(r.X, d["out_X"], rtol=0, atol=1e-12)
    assert np.array_equal(r.outlier, d["out_outlier"])
    assert np.allclose(r.cost_history, d["cost_history"], rtol=1e-12)
    assert np.array_equal(trials, d["trials"])


def test_workload0_oracle_fixture():
    g = json.load(open(os.path.join(GOLD, "workload0_oracle.json")))
    r, lam, _ = oracle.solve(problems.workload(0), workload_options(0), with_history=True)
    assert (r.iterations, r.trials, r.n_outliers) == (g["iterations"], g["trials"], g["n_outliers"])
    assert np.allclose(r.cost_history, g["cost_history"], rtol=1e-12)
    assert np.allclose(lam, g["lambda_history"], rtol=1e-12)
