"""Oracle solves of generated windows in a pool of worker processes (CPU only;
each worker regenerates its windows from their seeds and loads the oracle
itself).  Test infrastructure: the GPU parity tests of the full 512-window
batch use it as the checker."""
from __future__ import annotations

import hashlib
import multiprocessing as mp
import os
import sys
from concurrent.futures import ProcessPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _solve_window(args):
    seed, opt_fields, perturb = args
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import numpy as np

    from nav24_amd import problems
    from nav24_amd.abi import Options
    from oracle import oracle
    p = problems.workload(3, seed=seed)
    up, down = np.inf, -np.inf
    if perturb in ("X+", "X+t-"):
        p.X = np.nextafter(p.X, up)
    if perturb == "X-":
        p.X = np.nextafter(p.X, down)
    if perturb == "t+":
        p.T_cw[:, 9:] = np.nextafter(p.T_cw[:, 9:], up)
    if perturb in ("t-", "X+t-"):
        p.T_cw[:, 9:] = np.nextafter(p.T_cw[:, 9:], down)
    r = oracle.solve(p, Options(**opt_fields), reversed_sums=perturb == "order")
    return {"seed": seed, "final_cost": r.final_cost, "cost_history": r.cost_history.tolist(),
            "iterations": r.iterations, "trials": r.trials, "stop_reason": r.stop_reason,
            "n_outliers": r.n_outliers, "outlier": r.outlier.tobytes(),
            "outlier_sha1": hashlib.sha1(r.outlier.tobytes()).hexdigest()}


def workers() -> int:
    """Pool size: the host's cores, at most 16 (the CPU share of a one-GPU host)."""
    return max(1, min(16, os.cpu_count() or 1))


def solve_windows(seeds, opts, perturb=None):
    """The oracle's solve of window `seed` of the uneven-window generator, for
    each seed, in worker processes started fresh (spawn).  perturb: None, or
    one ulp on every point coordinate ("X+", "X-"), on every camera
    translation ("t+", "t-"), or both ("X+t-"); or the unperturbed window with
    the oracle's long sums in reversed order ("order")."""
    fields = {k: getattr(opts, k) for k in ("n_iter", "robust", "huber_delta", "lambda_init", "precision",
                                            "compat_fp32_io", "outlier_chi2", "max_trials")}
    with ProcessPoolExecutor(max_workers=workers(), mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_solve_window, [(s, fields, perturb) for s in seeds], chunksize=4))
