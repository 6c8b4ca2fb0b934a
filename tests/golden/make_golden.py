"""Writes the committed fixtures under tests/golden/ (run from the repo root).

  rng_kat.json          raw mt19937_64 draws and the generator's first normals
  tiny_problem.npz      a 4-camera / 25-point problem and the oracle's outputs
  workload0_oracle.json the 10-keyframe circle workload: sizes and oracle trace

These pin the generator and the oracle against silent changes; the oracle's
correctness itself is pinned by tests/test_oracle_*.py (closed forms, brute
force, schedule invariants) — no reference test vectors exist for this path.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nav24_amd import problems  # noqa: E402
from nav24_amd.abi import workload_options  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    kat = {
        "mt19937_64_seed5489_draw9999": str(problems.mt_draw(5489, 9999)),
        "mt19937_64_seed0_draws_0_1_2": [str(problems.mt_draw(0, k)) for k in range(3)],
        "normals_seed0_first8": problems.normals(0, 8).tolist(),
    }
    with open(os.path.join(OUT, "rng_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    prob = problems.workload(0, seed=21, n_cam=4, n_pt=25)
    opts = workload_options(0)
    r, lam, trials = oracle.solve(prob, opts, with_history=True)
    np.savez(os.path.join(OUT, "tiny_problem.npz"), T_cw=prob.T_cw, X=prob.X, obs_cam=prob.obs_cam,
             obs_pt=prob.obs_pt, obs_uv=prob.obs_uv, cam_fixed=prob.cam_fixed, pt_fixed=prob.pt_fixed,
             out_T_cw=r.T_cw, out_X=r.X, out_outlier=r.outlier, cost_history=r.cost_history,
             lambda_history=lam, trials=trials)

    prob = problems.workload(0, seed=0)
    r, lam, trials = oracle.solve(prob, opts, with_history=True)
    summ = {
        "n_cam": prob.n_cam, "n_pt": prob.n_pt, "n_obs": prob.n_obs,
        "n_generated_outliers": int(prob.obs_outlier.sum()),
        "uv_sum": float(prob.obs_uv.sum()), "X_sum": float(prob.X.sum()), "T_sum": float(prob.T_cw.sum()),
        "iterations": r.iterations, "trials": r.trials, "n_outliers": r.n_outliers,
        "initial_cost": r.initial_cost, "final_cost": r.final_cost,
        "cost_history": r.cost_history.tolist(), "lambda_history": lam.tolist(),
    }
    with open(os.path.join(OUT, "workload0_oracle.json"), "w") as f:
        json.dump(summ, f, indent=1)


if __name__ == "__main__":
    main()
