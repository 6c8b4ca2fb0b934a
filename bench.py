#!/usr/bin/env python3
"""bench.py — bundle-adjustment LM throughput on one MI355X.

Headline (the metric's own comparison, at equal final cost): 512 uneven local
windows (5-30 keyframes, 300-5000 points, 2-8 views per point, 10 LM
iterations each) solved together by the batched FP64 path, against the same
windows solved one at a time by the plain FP64 path.  One step = one batched
solve of all windows (nav_ba_run) with the batch resident in HBM (prepare is
timed apart).  `value` is window LM iterations/s; beside it observations/s,
the per-window final-cost agreement with the plain path, the mixed-precision
batch, the roofline of the batch's dominant kernel and the FP64 oracle on the
host's cores.  Further legs: the single random-walk problem (60 keyframes,
~565 k observations), the 300-keyframe loop (~3 M observations) against the
FP64 oracle's solve, the differentiable cost, key-point undistortion, one
window per call, and two-view initialisation.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BA LM iters/sec + obs/sec at equal final cost; mixed/batched vs plain FP64"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 matrix

WORKLOAD_NAMES = {
    0: "circle-10kf-1500pt", 1: "randomwalk-60kf-20000pt", 2: "loop-300kf-150000pt",
    3: "batched-512-windows",
}


def problem_stats(prob):
    """Sizes that the algorithmic byte / flop counts use."""
    kp = np.bincount(prob.obs_pt, minlength=prob.n_pt)
    active = kp > 0
    n_pt = int(active.sum())
    free_cam = np.ones(prob.n_cam, bool) if prob.cam_fixed is None else prob.cam_fixed == 0
    # observations on free cameras per point (Schur pair structure)
    kf = np.bincount(prob.obs_pt[free_cam[prob.obs_cam]], minlength=prob.n_pt).astype(np.float64)
    return dict(n_obs=prob.n_obs, n_pt=n_pt, n_cam=prob.n_cam, n_fc=int(free_cam.sum()),
                sum_k=float(kf.sum()), sum_pairs=float((kf * (kf + 1) / 2).sum()))


def stage_work(st, fp64=True):
    """Algorithmic bytes (HBM-bound stages) or flops (Schur rows) per launch."""
    no, npt, nc = st["n_obs"], st["n_pt"], st["n_cam"]
    if fp64:
        lin_pt = 32 * no + 113 * npt + 96 * nc     # o_cam, o_pt, uv, w out | X, Hll+bl out, ptr, free | cams
        lin_cam = 20 * st["sum_k"] + 32 * npt + 96 * nc  # camera-major point index + uv | X | cams
        cost = 24 * no + 32 * npt + 96 * nc        # o_cam, o_pt, uv | X | cams
    else:
        lin_pt = 28 * no + 77 * npt + 96 * nc
        lin_cam = 28 * no + 32 * npt + 96 * nc
        cost = 24 * no + 32 * npt + 96 * nc
    schur_flops = 108 * st["sum_k"] + 36 * st["sum_k"] + 216 * st["sum_pairs"]
    return {
        "lin_pt": ("hbm", lin_pt), "lin_cam": ("hbm", lin_cam), "cost": ("hbm", cost),
        "schur_row": ("fp64-valu", schur_flops),
    }


# kernel names as the profiler reports them: <scalar type, camera model 0 = pinhole>
STAGE_KERNEL = {"schur_row": "k_schur_row<double, 0>", "lin_pt": "k_lin_pt<double, 0>",
                "lin_cam": "k_lin_cam<double, 0>", "cost": "k_cost<double, 0>"}
SCHUR_KERNEL = {1: "k_schur_row<double, 0>", 2: "k_schur_row_pc<double, 0>"}


def measured_traffic(workload_name, stage):
    """HBM bytes per launch of a kernel (a stage name of STAGE_KERNEL or a
    kernel name) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of this workload (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload_name}_fp64.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"].get(STAGE_KERNEL.get(stage, stage))
    except (OSError, ValueError, KeyError):
        return None


def roofline_entry(timing, st, fp64):
    work = stage_work(st, fp64)
    total = {k: v[0] for k, v in timing.items() if v[1] > 0}
    dominant = max(total, key=total.get)
    cand = dominant if dominant in work else max((k for k in work if k in total), key=lambda k: total[k])
    ms, n = timing[cand]
    avg_s = ms / n / 1e3
    bound, amount = work[cand]
    if bound == "hbm":
        achieved = amount / avg_s / 1e9
        peak, unit = HBM_PEAK_GBS, "GB/s"
    else:  # the Schur rows run on the FP64 vector units (FP64 for both compute types)
        achieved = amount / avg_s / 1e12
        peak, unit = FP64_PEAK_TFLOPS, "TFLOP/s"
    return {
        "kernel": cand, "dominant_stage": dominant, "bound": bound, "achieved": achieved, "peak": peak,
        "unit": unit, "frac": achieved / peak, "traffic": None, "avg_launch_us": avg_s * 1e6,
        "algorithmic_per_launch": amount,
        "stages_ms_per_step": {k: round(v[0], 4) for k, v in timing.items() if v[1] > 0},
    }


def hbm_stage_rates(solver, st, fp64=True, reps=10):
    """Achieved HBM rate of the residual / Jacobian / accumulation kernels on
    their algorithmic bytes: `reps` launches of each between HIP events on the
    handle's stream, every window running and doing the work (a solve's stage
    timing would also count the launches that a window in its trial phase
    skips); the lin_cam stage includes its per-camera chunk reduction."""
    work = stage_work(st, fp64)
    out = {}
    for k in ("lin_pt", "lin_cam", "cost"):
        us = solver.time_stage(k, 1e-4, reps)
        if us <= 0:
            continue
        gbs = work[k][1] / (us * 1e-6) / 1e9
        out[k] = {"avg_launch_us": round(us, 2), "algorithmic_bytes": int(work[k][1]), "achieved_GBs": round(gbs, 1),
                  "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def _median_time(fn, reps):
    """Median wall time of `reps` calls of fn (and the times)."""
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def _sum_stats(stats):
    out = {k: 0.0 for k in ("n_obs", "n_pt", "n_cam", "n_fc", "sum_k", "sum_pairs")}
    for st in stats:
        for k in out:
            out[k] += st[k]
    return out


def _oracle_window(args):
    """Oracle solve of uneven window `seed` (for the all-cores CPU baseline)."""
    seed, n_iter = args
    sys.path.insert(0, ROOT)
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, workload_options
    from oracle import oracle
    o = workload_options(3)
    o.precision = NAV_BA_FP64
    r = oracle.solve(problems.workload(3, seed=seed), o)
    return r.iterations


def headline_leg(n_windows, steps, warmup, with_cpu):
    """The metric's comparison: n_windows uneven windows in one batched FP64
    solve against the plain FP64 path one window at a time, at equal final
    cost; the mixed batch beside it."""
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
    from nav24_amd.ba import GraphOptimHip

    wins = [problems.workload(3, seed=w) for w in range(n_windows)]
    n_obs = np.array([w.n_obs for w in wins], np.float64)
    st = _sum_stats([problem_stats(w) for w in wins])
    o64 = workload_options(3)
    o64.precision = NAV_BA_FP64
    s = GraphOptimHip(0)
    t0 = time.perf_counter()
    s.prepare_batched(wins, o64)
    prep = time.perf_counter() - t0
    for _ in range(warmup):
        s.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    r64 = s.fetch_batched()
    it64 = np.array([x.iterations for x in r64], np.float64)
    value = float(it64.sum()) / dt
    # roofline of the dominant kernel, the Schur rows, with every window running
    # (its own HIP events on the handle's stream, 10 launches at the initial
    # linearisation, lambda 1e-4)
    schur_us = s.time_schur_rows(1e-4, 10)
    flops = stage_work(st)["schur_row"][1]
    ach = flops / (schur_us * 1e-6) / 1e12
    kname = SCHUR_KERNEL.get(s.schur_kernel, STAGE_KERNEL["schur_row"])
    roof = {"kernel": kname, "bound": "fp64-valu", "achieved": round(ach, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 5),
            "traffic": measured_traffic(WORKLOAD_NAMES[3], kname), "avg_launch_us": round(schur_us, 2),
            "algorithmic_flops_per_launch": flops,
            "note": "FP64 vector ALU (no MFMA): 108 k + 36 k + 216 k(k+1)/2 flops per point of k free views"}
    s.set_timing(True)
    s.run()
    stages = {k: round(v[0], 4) for k, v in s.timing().items() if v[1] > 0}
    s.set_timing(False)
    s.close()
    # plain FP64 path, one window at a time (each prepared before its timed run)
    s = GraphOptimHip(0)
    tot = 0.0
    iters = 0
    plain = []
    for w in wins:
        s.prepare(w, o64)
        s.run()  # warm-up of this window's graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run()
        torch.cuda.synchronize()
        tot += time.perf_counter() - t0
        r = s.fetch()
        iters += r.iterations
        plain.append(r)
    s.close()
    plain_rate = iters / tot
    rel_bp = np.array([abs(a.final_cost - b.final_cost) / b.final_cost for a, b in zip(r64, plain)])
    # the mixed batch
    om = workload_options(3)
    om.precision = NAV_BA_MIXED
    s = GraphOptimHip(0)
    s.prepare_batched(wins, om)
    for _ in range(warmup):
        s.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run()
    torch.cuda.synchronize()
    dtm = (time.perf_counter() - t0) / steps
    rm = s.fetch_batched()
    s.close()
    itm = np.array([x.iterations for x in rm], np.float64)
    rel_m = np.array([abs(a.final_cost - b.final_cost) / b.final_cost for a, b in zip(rm, plain)])
    out = {
        "value": round(value, 1), "unit": "LM iterations/s", "ms_per_step": round(dt * 1e3, 3),
        "config": {"workload": f"{n_windows} uneven windows (5-30 keyframes, 300-5000 points, 2-8 views per point),"
                               " 10 LM iterations each, one batched solve", "n_windows": n_windows,
                   "n_obs": int(n_obs.sum()), "n_cam": int(st["n_cam"]), "precision": "fp64",
                   "huber_delta": "sqrt(5.991)", "lambda_init": 1e-4, "fixed_cameras_per_window": 2},
        "obs_per_s": round(float((it64 * n_obs).sum()) / dt, 1),
        "prepare_ms": round(prep * 1e3, 1),
        "speedup_vs_plain_fp64": round(value / plain_rate, 2),
        "plain_fp64_one_window_at_a_time": {"value": round(plain_rate, 1), "unit": "LM iterations/s",
                                            "ms_total": round(tot * 1e3, 2)},
        "equal_final_cost": {"windows": n_windows, "within_1e-9": int((rel_bp <= 1e-9).sum()),
                             "within_1e-5": int((rel_bp <= 1e-5).sum()), "max_rel": float(rel_bp.max()),
                             "median_rel": float(np.median(rel_bp)),
                             "note": "batched FP64 against the plain FP64 path per window; the two factor each "
                                     "window's system in different kernels (another order of FP64 additions)"},
        "mixed_batched": {"value": round(float(itm.sum()) / dtm, 1), "unit": "LM iterations/s",
                          "ms_per_step": round(dtm * 1e3, 3), "speedup_vs_plain_fp64":
                              round(float(itm.sum()) / dtm / plain_rate, 2),
                          "final_cost_rel_vs_plain_fp64": {"max": float(rel_m.max()), "median": float(np.median(rel_m)),
                                                           "p99": float(np.quantile(rel_m, 0.99)),
                                                           "windows_within_1e-5": int((rel_m <= 1e-5).sum())}},
        "roofline": roof, "stages_ms_per_step_stage_timed": stages,
    }
    if with_cpu:
        from oracle import oracle
        sample = wins[:16]
        done = {}

        def one_core():
            done["iters"] = sum(oracle.solve(w, o64).iterations for w in sample)
        med, ts = _median_time(one_core, 5)
        it_s = done["iters"]
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        ncpu = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        with ProcessPoolExecutor(max_workers=ncpu, mp_context=mp.get_context("spawn")) as ex:
            all_it = sum(ex.map(_oracle_window, [(w, 10) for w in range(n_windows)], chunksize=8))
        t_all = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(it_s / med, 1), "unit": "LM iterations/s", "cores": 1, "kind": "port",
            "sample": f"windows 0-15 of the batch ({int(n_obs[:16].sum())} observations, {it_s} window LM "
                      f"iterations) solved one after another by the FP64 oracle, median of 5 runs "
                      f"({', '.join(f'{t:.2f}' for t in ts)} s)",
            "all_cores": {"value": round(all_it / t_all, 1), "unit": "LM iterations/s", "cores": ncpu,
                          "sample": f"all {n_windows} windows, one FP64 oracle solve each, {ncpu} host processes "
                                    f"(oracle x {ncpu} cores), {t_all:.2f} s including process start"},
            "cpu": _cpu_model(), "host_cpus": os.cpu_count()}
    return out


def _pose_point_rmse(a_T, b_T, a_X=None, b_X=None):
    """RMS rotation angle (rad) and translation (m) differences of two pose
    sets, and the RMS point difference (m) when points are given."""
    Ra = a_T[:, :9].reshape(-1, 3, 3)
    Rb = b_T[:, :9].reshape(-1, 3, 3)
    tr = np.einsum("nij,nij->n", Ra, Rb)  # trace(Ra Rb^T)
    ang = np.arccos(np.clip((tr - 1.0) / 2.0, -1.0, 1.0))
    dt = np.linalg.norm(a_T[:, 9:] - b_T[:, 9:], axis=1)
    out = {"rotation_rad": float(np.sqrt(np.mean(ang ** 2))), "translation_m": float(np.sqrt(np.mean(dt ** 2)))}
    if a_X is not None:
        out["point_m"] = float(np.sqrt(np.mean(np.linalg.norm(a_X - b_X, axis=1) ** 2)))
    return out


def single_leg(workload, steps, warmup, with_cpu):
    """One problem (the 60-keyframe random walk): FP64 and mixed LM solves
    resident in HBM, prepare and end-to-end nav_ba_solve times, per-stage
    roofline, the factorisation stage, and the FP64 oracle on one core."""
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
    from nav24_amd.ba import GraphOptimHip

    prob = problems.workload(workload, seed=0)
    opts = workload_options(workload)
    opts.precision = NAV_BA_FP64
    st = problem_stats(prob)
    solver = GraphOptimHip(0)
    t0 = time.perf_counter()
    solver.prepare(prob, opts)
    prep = time.perf_counter() - t0
    for _ in range(warmup):
        solver.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = solver.fetch()
    iters = res.iterations
    value = iters * steps / dt
    solver.set_timing(True)
    for _ in range(steps):
        solver.run()
    timing = solver.timing()
    solver.set_timing(False)
    per_step = {k: (v[0] / steps, v[1] / steps) for k, v in timing.items()}
    roof = roofline_entry(per_step, st, True)
    if roof["kernel"] == "schur_row":
        roof["kernel"] = SCHUR_KERNEL.get(solver.schur_kernel, STAGE_KERNEL["schur_row"])
    roof["achieved"] = round(roof["achieved"], 3)
    roof["frac"] = round(roof["frac"], 5)
    roof["traffic"] = measured_traffic(WORKLOAD_NAMES.get(workload, ""), roof["kernel"])
    roof["hbm_kernels"] = hbm_stage_rates(solver, st, True)
    n = 6 * st["n_fc"]
    chol_ms = timing.get("cholesky", (0.0, 0))[0] / steps
    if chol_ms > 0 and res.trials > 0:
        flops = n ** 3 / 3.0 + 2.0 * n * n
        ach = flops / (chol_ms / res.trials / 1e3) / 1e12
        roof["cholesky_stage"] = {"n": n, "flops_per_trial": flops, "us_per_trial": round(chol_ms / res.trials * 1e3, 2),
                                  "achieved": round(ach, 4), "unit": "TFLOP/s",
                                  "frac": round(ach / FP64_PEAK_TFLOPS, 6)}
    e2e, _ = _median_time(lambda: solver.solve(prob, opts), 5)
    out = {"workload": WORKLOAD_NAMES.get(workload, str(workload)), "n_cam": prob.n_cam, "n_pt": prob.n_pt,
           "n_obs": prob.n_obs, "lm_iterations": iters, "trials": res.trials,
           "value": round(value, 2), "unit": "LM iterations/s", "ms_per_solve": round(dt / steps * 1e3, 3),
           "obs_per_s": round(prob.n_obs * value, 1), "prepare_ms": round(prep * 1e3, 2),
           "end_to_end_solve_ms": round(e2e * 1e3, 2), "final_cost": res.final_cost, "roofline": roof}
    solver.close()
    mopts = workload_options(workload)
    mopts.precision = NAV_BA_MIXED
    ms = GraphOptimHip(0)
    ms.prepare(prob, mopts)
    for _ in range(warmup):
        ms.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ms.run()
    torch.cuda.synchronize()
    mdt = time.perf_counter() - t0
    mres = ms.fetch()
    ms.close()
    out["mixed"] = {"value": round(mres.iterations * steps / mdt, 2), "unit": "LM iterations/s",
                    "ms_per_solve": round(mdt / steps * 1e3, 3),
                    "final_cost_rel_vs_fp64": abs(mres.final_cost - res.final_cost) / res.final_cost,
                    "speedup_vs_fp64": round((mres.iterations / mdt) / (iters / dt), 3)}
    if with_cpu:
        from oracle import oracle  # checker and CPU baseline only
        holder = {}

        def solve_once():
            holder["r"] = oracle.solve(prob, opts)
        med, ts = _median_time(solve_once, 3)
        ref = holder["r"]
        out["cpu_baseline"] = {"value": round(ref.iterations / med, 3), "unit": "LM iterations/s", "cores": 1,
                               "kind": "port",
                               "sample": f"the full solve ({ref.iterations} LM iterations, {prob.n_obs} observations),"
                                         f" FP64 oracle, 1 thread, median of 3 ({', '.join(f'{t:.2f}' for t in ts)} s)"}
        out["final_cost_rel_vs_oracle"] = abs(res.final_cost - ref.final_cost) / ref.final_cost
        out["outlier_flags_identical"] = bool(np.array_equal(res.outlier, ref.outlier))
    return out


def kb8_leg(steps, warmup, with_cpu):
    """The 60-keyframe random walk seen through the Kannala-Brandt fisheye
    camera (the same scene, observations re-projected with KannalaBrandt8's
    model, pixel noise and outliers): FP64 solve resident in HBM, the Schur
    rows' rate, and the FP64 oracle on one core with the final-cost agreement."""
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, workload_options
    from nav24_amd.ba import GraphOptimHip

    prob = problems.with_kb8(problems.workload(1, seed=0), seed=0)
    opts = workload_options(1)
    opts.precision = NAV_BA_FP64
    st = problem_stats(prob)
    solver = GraphOptimHip(0)
    t0 = time.perf_counter()
    solver.prepare(prob, opts)
    prep = time.perf_counter() - t0
    for _ in range(warmup):
        solver.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = solver.fetch()
    schur_us = solver.time_schur_rows(1e-4, 10)
    flops = stage_work(st)["schur_row"][1]
    out = {"workload": "randomwalk-60kf-20000pt, Kannala-Brandt camera k1..k4 = " + str(list(problems.KB8_COEFFS)),
           "n_obs": prob.n_obs, "lm_iterations": res.iterations, "trials": res.trials,
           "value": round(res.iterations * steps / dt, 2), "unit": "LM iterations/s",
           "ms_per_solve": round(dt / steps * 1e3, 3), "prepare_ms": round(prep * 1e3, 2),
           "final_cost": res.final_cost,
           "schur_rows": {"kernel": SCHUR_KERNEL.get(solver.schur_kernel), "avg_launch_us": round(schur_us, 2),
                          "achieved_TFLOPs": round(flops / (schur_us * 1e-6) / 1e12, 3)}}
    solver.close()
    if with_cpu:
        from oracle import oracle  # checker and CPU baseline only
        t0 = time.perf_counter()
        ref = oracle.solve(prob, opts)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(ref.iterations / tc, 3), "unit": "LM iterations/s", "cores": 1,
                               "kind": "port", "sample": f"the full solve ({ref.iterations} LM iterations), FP64 "
                                                         f"oracle, 1 thread, one run ({tc:.2f} s)"}
        out["final_cost_rel_vs_oracle"] = abs(res.final_cost - ref.final_cost) / ref.final_cost
        out["outlier_flags_identical"] = bool(np.array_equal(res.outlier, ref.outlier))
    return out


def loop_leg(steps, warmup, with_cpu):
    """300 keyframes on a 30 m loop, 150 000 points, 30 LM iterations: FP64
    and mixed on the GPU, each against the FP64 oracle's solve of the same
    problem (tests/golden/loop_oracle.json: final cost, poses, every 50th
    point), with that solve's own one-ulp sensitivity."""
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
    from nav24_amd.ba import GraphOptimHip

    prob = problems.workload(2, seed=0)
    st = problem_stats(prob)
    out = {"workload": WORKLOAD_NAMES[2], "n_cam": prob.n_cam, "n_pt": prob.n_pt, "n_obs": prob.n_obs}
    fx = None
    fx_path = os.path.join(ROOT, "tests", "golden", "loop_oracle.json")
    if os.path.exists(fx_path):
        with open(fx_path) as f:
            fx = json.load(f)
        if fx.get("n_obs") != prob.n_obs:
            fx = None
    res = {}
    for name, prec in (("fp64", NAV_BA_FP64), ("mixed", NAV_BA_MIXED)):
        o = workload_options(2)
        o.precision = prec
        s = GraphOptimHip(0)
        t0 = time.perf_counter()
        s.prepare(prob, o)
        prep = time.perf_counter() - t0
        for _ in range(warmup):
            s.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        r = s.fetch()
        if name == "fp64":
            out["hbm_kernels"] = hbm_stage_rates(s, st, True)
        s.close()
        res[name] = r
        entry = {"lm_iterations_per_s": round(r.iterations / dt, 2), "obs_per_s": round(prob.n_obs * r.iterations / dt, 1),
                 "ms_per_solve": round(dt * 1e3, 2), "lm_iterations": r.iterations, "trials": r.trials,
                 "final_cost": r.final_cost, "prepare_ms": round(prep * 1e3, 1)}
        if fx is not None:
            Xs = np.asarray(fx["X_sample"])
            entry["vs_fp64_oracle"] = {
                "final_cost_rel": abs(r.final_cost - fx["final_cost"]) / fx["final_cost"],
                "rmse": _pose_point_rmse(r.T_cw, np.asarray(fx["T_cw"]), r.X[::fx["point_stride"]], Xs),
                "point_rmse_sample": f"every {fx['point_stride']}th point ({len(Xs)})"}
        out[name] = entry
    if fx is not None:
        out["oracle"] = {"final_cost": fx["final_cost"], "trials": fx["trials"],
                         "one_ulp_sensitivity_final": fx["one_ulp_sensitivity"]["final"]}
    m, f = res["mixed"], res["fp64"]
    out["mixed"]["final_cost_rel_vs_fp64_gpu"] = abs(m.final_cost - f.final_cost) / f.final_cost
    out["mixed"]["speedup_vs_fp64"] = round(out["mixed"]["lm_iterations_per_s"] / out["fp64"]["lm_iterations_per_s"], 3)
    if with_cpu:
        from oracle import oracle
        o = workload_options(2)
        o.precision = NAV_BA_FP64
        o.n_iter = 1
        holder = {}

        def one_iter():
            holder["r"] = oracle.solve(prob, o)
        med, ts = _median_time(one_iter, 3)
        r1 = holder["r"]
        out["cpu_baseline"] = {"value": round(r1.iterations / med, 4), "unit": "LM iterations/s", "cores": 1,
                               "kind": "port",
                               "sample": f"the first LM iteration of the same problem ({r1.trials} trial, "
                                         f"{prob.n_obs} observations), FP64 oracle, 1 thread, median of 3 "
                                         f"({', '.join(f'{t:.2f}' for t in ts)} s)"}
    return out


def diffcost_leg(n_problems, steps, with_cpu=False):
    """Robust reprojection cost with radial-tangential distortion and its
    hand-written adjoint (poses, points, intrinsics, distortion) over a batch of
    20-keyframe, 5 000-point problems resident in HBM: forward (cost only) and
    forward + backward throughput, FP32 against FP64."""
    import torch
    from nav24_amd import problems
    from nav24_amd.diffcost import CostBatch

    probs = [problems.workload(4, seed=s) for s in range(n_problems)]
    T = np.concatenate([p.T_cw for p in probs])
    X = np.concatenate([p.X for p in probs])
    intr = np.array([probs[0].fx, probs[0].fy, probs[0].cx, probs[0].cy, *probs[0].dist])
    b = CostBatch(probs)
    out = {"workload": f"{n_problems} x (20 kf, 5000 pts), rad-tan distortion", "n_obs": b.n_obs,
           "n_cam": b.n_cam, "n_pt": b.n_pt}
    res = {}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        args = [torch.tensor(a, device="cuda", dtype=dt) for a in (T, X, intr)]
        for grad, key in ((False, "forward"), (True, "forward_backward")):
            b.eval(*args, grad=grad)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                r = b.eval(*args, grad=grad)
            torch.cuda.synchronize()
            dt_s = (time.perf_counter() - t0) / steps
            out.setdefault(name, {})[key] = {"ms": round(dt_s * 1e3, 4), "obs_per_s": round(b.n_obs / dt_s, 1)}
            if grad:
                res[name] = r
    if with_cpu:
        from oracle import cost_oracle as co

        def oracle_all():
            c0 = p0 = 0
            for p in probs:
                co.cost_and_grad(T[c0:c0 + p.n_cam], X[p0:p0 + p.n_pt], p.obs_cam, p.obs_pt, p.obs_uv, intr, True)
                c0 += p.n_cam
                p0 += p.n_pt
        med, ts = _median_time(oracle_all, 3)
        out["cpu_baseline"] = {"value": round(b.n_obs / med, 1), "unit": "observations/s (forward + backward)",
                               "cores": 1, "kind": "port",
                               "sample": f"all {len(probs)} problems, FP64 numpy oracle, median of 3 "
                                         f"({', '.join(f'{t:.2f}' for t in ts)} s)"}
    rel = lambda a, r: float((a.double() - r).abs().max() / r.abs().max())
    out["fp32"]["rel_err_vs_fp64"] = {"cost": rel(res["fp32"].cost, res["fp64"].cost),
                                      "pose": rel(res["fp32"].g_pose, res["fp64"].g_pose),
                                      "point": rel(res["fp32"].g_X, res["fp64"].g_X),
                                      "intrinsics": rel(res["fp32"].g_intr, res["fp64"].g_intr)}
    # the same against FP64 at the FP32-rounded inputs: the arithmetic's part of
    # the error, without the rounding of the poses, points and intrinsics
    r32in = b.eval(*[torch.tensor(a, device="cuda", dtype=torch.float32).double() for a in (T, X, intr)], grad=True)
    out["fp32"]["rel_err_vs_fp64_same_inputs"] = {"cost": rel(res["fp32"].cost, r32in.cost),
                                                  "pose": rel(res["fp32"].g_pose, r32in.g_pose),
                                                  "point": rel(res["fp32"].g_X, r32in.g_X),
                                                  "intrinsics": rel(res["fp32"].g_intr, r32in.g_intr)}
    b.close()
    return out


def undistort_leg(n_points, steps, with_cpu):
    """Key-point undistortion of the radial-tangential camera
    (PinholeRadTan::UndistortKeyPoints): n_points key points resident in HBM,
    one per observation of the 3 M-observation loop; each of K launches timed
    with HIP events on torch's stream.  8 bytes read and 8 written per point.  With the
    CPU baseline: the oracle on one host thread over the same points, and
    whether the device result equals it."""
    import torch
    from nav24_amd.undistort import PinholeRadTanHip

    prm = [450.0, 450.0, 320.0, 240.0, -0.28, 0.07, 5e-4, 3e-4]
    r = np.random.default_rng(0)
    uv = np.stack([r.uniform(0, 640, n_points), r.uniform(0, 480, n_points)], 1).astype(np.float32)
    cam = PinholeRadTanHip(prm)
    src = torch.from_numpy(uv).cuda()
    dst = torch.empty_like(src)
    cam.undistort_device(src, out=dst)
    torch.cuda.synchronize()
    # each launch between its own pair of events (back to back, the host's
    # per-call overhead would be timed instead of the kernel)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record()
        cam.undistort_device(src, out=dst)
        b.record()
    torch.cuda.synchronize()
    us = sum(a.elapsed_time(b) for a, b in ev) * 1e3 / steps
    nbytes = 16 * n_points
    out = {"workload": f"{n_points} key points, 640x480 pinhole, k1 k2 p1 p2 = {prm[4:]}", "n_points": n_points,
           "us_per_launch": round(us, 2), "points_per_s": round(n_points / (us * 1e-6), 1),
           "roofline": {"bound": "fp64 issue (5 iterations with a division each)", "algorithmic_bytes": nbytes,
                        "achieved_GBs": round(nbytes / (us * 1e-6) / 1e9, 1),
                        "frac_of_hbm_peak": round(nbytes / (us * 1e-6) / (HBM_PEAK_GBS * 1e9), 4)}}
    if with_cpu:
        from oracle import oracle
        t0 = time.perf_counter()
        ref = oracle.kp_undistort(uv, prm)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n_points / dt, 1), "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"the same {n_points} points, FP64 oracle, 1 thread, {dt:.3f} s"}
        out["equal_to_oracle"] = bool(np.array_equal(dst.cpu().numpy(), ref))
    # the Kannala-Brandt fisheye camera (KannalaBrandt8::UndistortKeyPoints) on the same points
    from nav24_amd.undistort import KannalaBrandt8Hip
    kb = [380.0, 380.0, 320.0, 240.0, 0.05, -0.01, 0.002, -0.0003]
    fish = KannalaBrandt8Hip(kb)
    fish.undistort_device(src, out=dst)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record()
        fish.undistort_device(src, out=dst)
        b.record()
    torch.cuda.synchronize()
    us_f = sum(a.elapsed_time(b) for a, b in ev) * 1e3 / steps
    out["fisheye"] = {"params": kb, "us_per_launch": round(us_f, 2), "points_per_s": round(n_points / (us_f * 1e-6), 1),
                      "achieved_GBs": round(nbytes / (us_f * 1e-6) / 1e9, 1)}
    if with_cpu:
        t0 = time.perf_counter()
        ref = oracle.kp_undistort_fisheye(uv, kb)
        dt = time.perf_counter() - t0
        got = dst.cpu().numpy()
        out["fisheye"]["cpu_baseline"] = {"value": round(n_points / dt, 1), "unit": "points/s", "cores": 1,
                                          "kind": "port", "sample": f"the same points, 1 thread, {dt:.3f} s"}
        out["fisheye"]["values_differing_from_oracle"] = int((got != ref).sum())
    return out


def window_leg(n_windows, reps, with_cpu):
    """One local window per call, as the reference's front end solves them
    (FE_SlamMonoV.cpp:154-165): nav_ba_solve end to end (upload and ordering,
    the LM solve, results back to the host) for windows 0..n_windows-1 of the
    uneven-window generator, `reps` calls each, with the prepare / run / fetch
    split of the same calls.  Two cases: a new structure every call (another
    window solved on the handle in between: host ordering, work items and
    graph capture each time), and the same structure with new poses, points
    and measurements (the cached index: only the state is uploaded).  The FP64
    oracle on window 0 is the CPU baseline."""
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import Problem, workload_options
    from nav24_amd.ba import GraphOptimHip

    def moved(p, seed):
        rng = np.random.default_rng(seed)
        T = p.T_cw.copy()
        T[:, 9:] += rng.normal(0, 0.01, (p.n_cam, 3))
        return Problem(T_cw=T, X=p.X + rng.normal(0, 0.01, p.X.shape), obs_cam=p.obs_cam, obs_pt=p.obs_pt,
                       obs_uv=p.obs_uv + rng.normal(0, 0.2, p.obs_uv.shape), cam_fixed=p.cam_fixed,
                       pt_fixed=p.pt_fixed)

    s = GraphOptimHip(0)
    o = workload_options(3)
    wins = [problems.workload(3, seed=wi) for wi in range(n_windows)]
    other = problems.workload(3, seed=n_windows + 1)

    def timed(prob, before):
        """Seconds of one solve, then of prepare / run / fetch on a second call."""
        if before is not None:
            s.solve(before, o)
        t0 = time.perf_counter()
        r = s.solve(prob, o)
        t_solve = time.perf_counter() - t0
        cached = s.prepare_cached
        if before is not None:
            s.solve(before, o)
        a = time.perf_counter()
        s.prepare(prob, o)
        b = time.perf_counter()
        s.run()
        torch.cuda.synchronize()
        c = time.perf_counter()
        s.fetch()
        d = time.perf_counter()
        return r, cached, t_solve, b - a, c - b, d - c

    per = []
    for wi, w in enumerate(wins):
        s.solve(w, o)  # warm-up
        row = {"n_cam": w.n_cam, "n_obs": w.n_obs}
        for case, runs in (("new_structure", [(w, other if k % 2 == 0 else wins[(wi + 1) % n_windows])
                                              for k in range(reps)]),
                           ("cached_structure", [(moved(w, 100 * wi + k), None) for k in range(reps)])):
            if case == "cached_structure":
                s.solve(w, o)
            res = [timed(p, before) for p, before in runs]
            row[case] = {"ms_per_solve": round(float(np.median([x[2] for x in res])) * 1e3, 3),
                         "prepare_ms": round(float(np.median([x[3] for x in res])) * 1e3, 3),
                         "run_ms": round(float(np.median([x[4] for x in res])) * 1e3, 3),
                         "fetch_ms": round(float(np.median([x[5] for x in res])) * 1e3, 3),
                         "cached_prepares": int(sum(x[1] for x in res)), "lm_iterations": res[-1][0].iterations}
        per.append(row)
    s.close()
    new_ms = np.array([x["new_structure"]["ms_per_solve"] for x in per])
    cached_ms = np.array([x["cached_structure"]["ms_per_solve"] for x in per])
    out = {"workload": f"windows 0..{n_windows - 1} of the uneven-window generator, one solve call each",
           "precision": "mixed (the window options)",
           "median_ms_per_solve": round(float(np.median(new_ms)), 3), "max_ms_per_solve": round(float(new_ms.max()), 3),
           "median_ms_per_solve_cached": round(float(np.median(cached_ms)), 3),
           "max_ms_per_solve_cached": round(float(cached_ms.max()), 3), "windows": per}
    if with_cpu:
        from oracle import oracle
        w = problems.workload(3, seed=0)
        o64 = workload_options(3)
        t0 = time.perf_counter()
        ref = oracle.solve(w, o64)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dt * 1e3, 3), "unit": "ms per solve of window 0", "cores": 1,
                               "kind": "port", "sample": f"window 0 ({w.n_obs} observations, {ref.iterations} "
                                                         f"LM iterations), FP64 oracle, 1 thread"}
    return out


def twoview_leg(n_matches, n_hyp, steps, with_cpu):
    """Two-view RANSAC scoring (TwoViewReconstruction::FindHomography /
    FindFundamental scoring loops): n_matches matches of a synthetic planar
    scene and n_hyp homography and n_hyp fundamental hypotheses resident in
    HBM; each scoring call (all hypotheses, the kept one, its inlier flags)
    timed between its own HIP events on torch's stream.  With the CPU
    baseline: the oracle on one host thread over the same input, and whether
    scores, kept index and flags equal it."""
    import torch
    from nav24_amd.twoview import TwoViewReconstructionHip
    from nav24_amd.twoview_synth import ransac_hypotheses, synthetic_two_view

    sc = synthetic_two_view(n_matches, seed=1, planar=True)
    H21, H12, F21 = ransac_hypotheses(sc, n_hyp, seed=1)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    k1, k2, mt, h21, h12, f21 = d(sc.kp1), d(sc.kp2), d(sc.matches), d(H21), d(H12), d(F21)
    tv = TwoViewReconstructionHip()
    outH = tv.score_device("H", k1, k2, mt, h21, h12)
    outF = tv.score_device("F", k1, k2, mt, f21)
    torch.cuda.synchronize()
    res = {"workload": f"{n_matches} matches (planar scene, 20 % outliers), {n_hyp} H and {n_hyp} F hypotheses",
           "n_matches": n_matches, "n_hyp": n_hyp}
    for kind, out, call in (("homography", outH, lambda: tv.score_device("H", k1, k2, mt, h21, h12, out=outH)),
                            ("fundamental", outF, lambda: tv.score_device("F", k1, k2, mt, f21, out=outF))):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record()
            call()
            b.record()
        torch.cuda.synchronize()
        us = sum(a.elapsed_time(b) for a, b in ev) * 1e3 / steps
        res[kind] = {"us_per_call": round(us, 2),
                     "match_hypotheses_per_s": round(n_matches * n_hyp / (us * 1e-6), 1),
                     "kept": int(out[1].item()), "inliers": int(out[3].sum().item())}
    if with_cpu:
        from oracle import oracle
        reps = 20
        for kind, args in (("homography", (H21, H12)), ("fundamental", (F21,))):
            t0 = time.perf_counter()
            for _ in range(reps):
                ref = oracle.tv_find(sc.kp1, sc.kp2, sc.matches, *args)
            dt = (time.perf_counter() - t0) / reps
            out = outH if kind == "homography" else outF
            res[kind]["cpu_baseline"] = {"value": round(n_matches * n_hyp / dt, 1), "unit": "match-hypotheses/s",
                                         "us_per_call": round(dt * 1e6, 1), "cores": 1, "kind": "port",
                                         "sample": f"the same input, oracle, 1 thread, {reps} calls"}
            res[kind]["equal_to_oracle"] = bool(
                np.array_equal(out[0].cpu().numpy().view(np.uint32), ref[0].view(np.uint32))
                and out[1].item() == ref[1] and np.array_equal(out[3].cpu().numpy(), ref[3]))
    # the hypotheses themselves from 200 RANSAC sets (normalisation on the host,
    # 8-point solves on the device), host arrays in and out, end to end
    rng = np.random.default_rng(1)
    sets = np.stack([rng.choice(n_matches, 8, replace=False) for _ in range(n_hyp)]).astype(np.int32)
    tv.Hypotheses(sc.kp1, sc.kp2, sc.matches, sets)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        hyp = tv.Hypotheses(sc.kp1, sc.kp2, sc.matches, sets)
    dt = (time.perf_counter() - t0) / reps
    res["hypotheses"] = {"sets": n_hyp, "us_per_call_host_arrays": round(dt * 1e6, 1)}
    if with_cpu:
        from oracle import oracle
        t0 = time.perf_counter()
        ref = oracle.tv_hypotheses(sc.kp1, sc.kp2, sc.matches, sets)
        dtc = time.perf_counter() - t0
        res["hypotheses"]["cpu_baseline"] = {"value": round(dtc * 1e6, 1), "unit": "us per call", "cores": 1,
                                             "kind": "port", "sample": "the same sets, oracle, 1 thread, 1 call"}
        res["hypotheses"]["equal_to_oracle"] = bool(all(
            np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(hyp, ref)))
    # CheckRT for the motion candidates (4 as ReconstructF, 8 as ReconstructH)
    # on the fundamental matrix's inliers, each call between its own events
    from nav24_amd.twoview_synth import motion_candidates
    inl = outF[3]
    for n_c in (4, 8):
        R, t = motion_candidates(sc, n_c - 4, seed=1)
        rt = tv.check_rt_device(k1, k2, mt, inl, sc.K, R, t)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record()
            tv.check_rt_device(k1, k2, mt, inl, sc.K, R, t, out=rt)
            b.record()
        torch.cuda.synchronize()
        us = sum(a.elapsed_time(b) for a, b in ev) * 1e3 / steps
        n_inl = int(inl.sum().item())
        entry = {"us_per_call": round(us, 2), "inlier_matches": n_inl,
                 "triangulations_per_s": round(n_inl * n_c / (us * 1e-6), 1),
                 "n_good": rt["n_good"].cpu().tolist()}
        if with_cpu:
            from oracle import oracle
            inl_h = inl.cpu().numpy()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                ref = oracle.tv_check_rt(sc.kp1, sc.kp2, sc.matches, inl_h, sc.K, R, t, th2=tv.th2)
            dt = (time.perf_counter() - t0) / reps
            entry["cpu_baseline"] = {"value": round(n_inl * n_c / dt, 1), "unit": "triangulations/s",
                                     "us_per_call": round(dt * 1e6, 1), "cores": 1, "kind": "port",
                                     "sample": f"the same input, oracle, 1 thread, {reps} calls"}
            entry["equal_to_oracle"] = bool(
                np.array_equal(rt["n_good"].cpu().numpy(), ref["n_good"])
                and np.array_equal(rt["p3d"].cpu().numpy().view(np.uint32), ref["p3d"].view(np.uint32))
                and np.array_equal(rt["good"].cpu().numpy(), ref["good"]))
        res[f"check_rt_{n_c}_candidates"] = entry
    return res


def copy_calibration(mb, steps):
    """A plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it
    scales the same counters on the solver kernels."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"copy_mib": mb, "bytes_read": n * 4, "bytes_written": n * 4,
                      "ms": round(dt * 1e3, 4), "GBps_read_plus_write": round(2 * n * 4 / dt / 1e9, 1)}), flush=True)


def batched_profile(n_windows, steps, warmup, precision="mixed"):
    import torch
    from nav24_amd import problems
    from nav24_amd.abi import NAV_BA_FP64, NAV_BA_MIXED, workload_options
    from nav24_amd.ba import GraphOptimHip
    wins = [problems.workload(3, seed=w) for w in range(n_windows)]
    o = workload_options(3)
    o.precision = NAV_BA_MIXED if precision == "mixed" else NAV_BA_FP64
    s = GraphOptimHip(0)
    s.prepare_batched(wins, o)
    for _ in range(warmup):
        s.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    s.set_timing(True)
    s.run()
    tm = s.timing()
    s.set_timing(False)
    # last: the headline's roofline timing of the Schur rows (1 warm-up and 10
    # timed launches back to back, every window running)
    schur_us = s.time_schur_rows(1e-4, 10)
    kernel, chunk = s.schur_kernel, s.schur_chunk
    s.close()
    print(json.dumps({"batched_windows": n_windows, "ms_per_batch": round(dt * 1e3, 3),
                      "stages_ms": {k: round(v[0], 4) for k, v in tm.items() if v[1] > 0},
                      "schur_kernel": SCHUR_KERNEL.get(kernel), "schur_chunk": chunk,
                      "schur_rows_avg_us": round(schur_us, 2)}), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=1, help="single-problem leg: generator workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--windows", type=int, default=512, help="headline: number of uneven windows in the batch")
    ap.add_argument("--loop", type=int, default=1, help="loop leg (300 keyframes, 150 000 points): 0 = skip")
    ap.add_argument("--kb8", type=int, default=1, help="fisheye-camera leg (the random walk through a "
                                                        "Kannala-Brandt camera): 0 = skip")
    ap.add_argument("--diffcost", type=int, default=64,
                    help="differentiable-cost leg: number of distorted-camera problems (0 = skip)")
    ap.add_argument("--undistort", type=int, default=3_000_000,
                    help="key-point undistortion leg: number of key points (0 = skip)")
    ap.add_argument("--window-solves", type=int, default=8,
                    help="per-call window leg: number of windows solved one call at a time (0 = skip)")
    ap.add_argument("--twoview", type=int, default=1000,
                    help="two-view RANSAC scoring leg: number of matches, 200 H and 200 F hypotheses (0 = skip)")
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling: run only the batched solve (K timed runs) and print its timing")
    ap.add_argument("--single-only", action="store_true",
                    help="profiling: run only the single-problem leg of --workload and print it")
    ap.add_argument("--window-only", action="store_true",
                    help="profiling: run only the per-call window leg (--window-solves windows) and print it")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="profiling (--batched-only): precision of the batched solve")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="profiling: K device-to-device copies of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs is designed but not run.", file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401  (device synchronisation inside the legs)
    if args.copy_mb > 0:
        copy_calibration(args.copy_mb, args.steps)
        return
    if args.batched_only:
        batched_profile(args.windows, args.steps, args.warmup, args.precision)
        return
    if args.single_only:
        print(json.dumps(single_leg(args.workload, args.steps, args.warmup, not args.no_cpu)), flush=True)
        return
    if args.window_only:
        print(json.dumps(window_leg(max(args.window_solves, 1), 10, not args.no_cpu)), flush=True)
        return
    with_cpu = not args.no_cpu
    head = headline_leg(args.windows, args.steps, args.warmup, with_cpu)
    line = {
        "metric": METRIC, "value": head.pop("value"), "unit": head.pop("unit"), "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head.pop("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": head.pop("config"),
    }
    roof = head.pop("roofline")
    cpu = head.pop("cpu_baseline", None)
    line.update(head)
    line["roofline"] = roof
    if cpu is not None:
        line["cpu_baseline"] = cpu
    line["single_problem"] = single_leg(args.workload, args.steps, args.warmup, with_cpu)
    if args.kb8:
        line["kb8"] = kb8_leg(args.steps, args.warmup, with_cpu)
    if args.loop:
        line["loop"] = loop_leg(max(1, min(args.steps, 3)), 1, with_cpu)
    if args.diffcost > 0:
        line["diffcost"] = diffcost_leg(args.diffcost, max(args.steps, 5), with_cpu)
    if args.undistort > 0:
        line["undistort"] = undistort_leg(args.undistort, max(args.steps, 20), with_cpu)
    if args.window_solves > 0:
        line["window_solve"] = window_leg(args.window_solves, 10, with_cpu)
    if args.twoview > 0:
        line["twoview"] = twoview_leg(args.twoview, 200, max(args.steps, 20), with_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
