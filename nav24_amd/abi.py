"""ctypes mirror of include/nav_ba.h (the C ABI of the BA back-end).

The structures here are byte-for-byte the C structs; `ProblemArrays` keeps the
numpy buffers a `nav_ba_problem` points into alive for as long as it lives.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

NAV_BA_OK = 0
NAV_BA_EMPTY = 1
NAV_BA_BAD_ARG = 2
NAV_BA_NOT_PD_TERMINATED = 3
NAV_BA_HIP_ERROR = 4
NAV_BA_NO_DEVICE = 5
NAV_BA_TOO_LARGE = 6
NAV_BA_NOT_PREPARED = 7

NAV_BA_FP64 = 0
NAV_BA_MIXED = 1

NAV_BA_CAM_PINHOLE = 0
NAV_BA_CAM_KB8 = 1

NAV_BA_STOP_MAX_ITER = 0
NAV_BA_STOP_TERMINATE = 1
NAV_BA_STOP_EMPTY = 2

STATUS_NAMES = {
    0: "OK", 1: "EMPTY", 2: "BAD_ARG", 3: "NOT_PD_TERMINATED", 4: "HIP_ERROR",
    5: "NO_DEVICE", 6: "TOO_LARGE", 7: "NOT_PREPARED",
}

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)


class NavBaProblem(C.Structure):
    _fields_ = [
        ("n_cam", C.c_int32), ("n_pt", C.c_int32), ("n_obs", C.c_int32),
        ("T_cw", _dp), ("cam_fixed", _u8p), ("X", _dp), ("pt_fixed", _u8p),
        ("obs_cam", _i32p), ("obs_pt", _i32p), ("obs_uv", _dp),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("camera_model", C.c_int32),
        ("k1", C.c_double), ("k2", C.c_double), ("k3", C.c_double), ("k4", C.c_double),
    ]


class NavBaOptions(C.Structure):
    _fields_ = [
        ("n_iter", C.c_int32), ("robust", C.c_int32), ("huber_delta", C.c_double),
        ("lambda_init", C.c_double), ("precision", C.c_int32), ("compat_fp32_io", C.c_int32),
        ("outlier_chi2", C.c_double), ("max_trials", C.c_int32),
    ]


class NavBaResult(C.Structure):
    _fields_ = [
        ("T_cw", _dp), ("X", _dp), ("outlier", _u8p), ("cost_history", _dp),
        ("iterations", C.c_int32), ("trials", C.c_int32), ("stop_reason", C.c_int32),
        ("n_outliers", C.c_int32), ("initial_cost", C.c_double), ("final_cost", C.c_double),
        ("final_lambda", C.c_double), ("solve_ms", C.c_double),
    ]


def ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Problem:
    """One BA problem as numpy arrays in the C ABI's layout."""
    T_cw: np.ndarray            # (n_cam, 12) float64
    X: np.ndarray               # (n_pt, 3) float64
    obs_cam: np.ndarray         # (n_obs,) int32
    obs_pt: np.ndarray          # (n_obs,) int32
    obs_uv: np.ndarray          # (n_obs, 2) float64
    fx: float = 450.0
    fy: float = 450.0
    cx: float = 320.0
    cy: float = 240.0
    cam_fixed: Optional[np.ndarray] = None   # (n_cam,) uint8
    pt_fixed: Optional[np.ndarray] = None    # (n_pt,) uint8
    camera_model: int = 0                    # NAV_BA_CAM_PINHOLE or NAV_BA_CAM_KB8
    kb: tuple = (0.0, 0.0, 0.0, 0.0)         # Kannala-Brandt k1..k4
    # generator extras (not part of the ABI)
    T_gt: Optional[np.ndarray] = None
    X_gt: Optional[np.ndarray] = None
    obs_outlier: Optional[np.ndarray] = None
    dist: Optional[np.ndarray] = None
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.T_cw = np.ascontiguousarray(self.T_cw, dtype=np.float64).reshape(-1, 12)
        self.X = np.ascontiguousarray(self.X, dtype=np.float64).reshape(-1, 3)
        self.obs_cam = np.ascontiguousarray(self.obs_cam, dtype=np.int32).reshape(-1)
        self.obs_pt = np.ascontiguousarray(self.obs_pt, dtype=np.int32).reshape(-1)
        self.obs_uv = np.ascontiguousarray(self.obs_uv, dtype=np.float64).reshape(-1, 2)
        if self.cam_fixed is not None:
            self.cam_fixed = np.ascontiguousarray(self.cam_fixed, dtype=np.uint8).reshape(-1)
        if self.pt_fixed is not None:
            self.pt_fixed = np.ascontiguousarray(self.pt_fixed, dtype=np.uint8).reshape(-1)

    @property
    def n_cam(self) -> int:
        return self.T_cw.shape[0]

    @property
    def n_pt(self) -> int:
        return self.X.shape[0]

    @property
    def n_obs(self) -> int:
        return self.obs_cam.shape[0]

    def c_struct(self) -> NavBaProblem:
        p = NavBaProblem()
        p.n_cam, p.n_pt, p.n_obs = self.n_cam, self.n_pt, self.n_obs
        p.T_cw = ptr(self.T_cw, C.c_double)
        p.cam_fixed = ptr(self.cam_fixed, C.c_uint8)
        p.X = ptr(self.X, C.c_double)
        p.pt_fixed = ptr(self.pt_fixed, C.c_uint8)
        p.obs_cam = ptr(self.obs_cam, C.c_int32)
        p.obs_pt = ptr(self.obs_pt, C.c_int32)
        p.obs_uv = ptr(self.obs_uv, C.c_double)
        p.fx, p.fy, p.cx, p.cy = self.fx, self.fy, self.cx, self.cy
        p.camera_model = int(self.camera_model)
        p.k1, p.k2, p.k3, p.k4 = (float(v) for v in self.kb)
        return p


@dataclass
class Options:
    """Solver settings (nav_ba_options).  Defaults are the reference's
    (nIterations core/backEnd/Problem.hpp:18,36; mbRobust Problem.hpp:46;
    Huber delta BE_GraphOptim.cpp:125)."""
    n_iter: int = 5
    robust: bool = False
    huber_delta: float = float(np.float32(np.sqrt(5.99)))
    lambda_init: float = 0.0
    precision: int = NAV_BA_FP64
    compat_fp32_io: bool = True
    outlier_chi2: float = 5.991
    max_trials: int = 10

    def c_struct(self) -> NavBaOptions:
        o = NavBaOptions()
        o.n_iter = int(self.n_iter)
        o.robust = int(bool(self.robust))
        o.huber_delta = float(self.huber_delta)
        o.lambda_init = float(self.lambda_init)
        o.precision = int(self.precision)
        o.compat_fp32_io = int(bool(self.compat_fp32_io))
        o.outlier_chi2 = float(self.outlier_chi2)
        o.max_trials = int(self.max_trials)
        return o


def workload_options(workload: int) -> Options:
    """Solver settings of the named workloads: Huber delta sqrt(5.991),
    iterations 10/20/30/10/10, user lambda 1e-4 except on the 300-keyframe
    loop, which names no initial lambda and gets g2o's automatic one
    (1e-5 max diag H): with 1e-4 its first iteration rejects ten trials and
    stops."""
    iters = {0: 10, 1: 20, 2: 30, 3: 10, 4: 10}[workload]
    prec = NAV_BA_MIXED if workload in (2, 3) else NAV_BA_FP64
    lam = 0.0 if workload == 2 else 1e-4
    return Options(n_iter=iters, robust=True, huber_delta=float(np.sqrt(5.991)),
                   lambda_init=lam, precision=prec, compat_fp32_io=False)


@dataclass
class Result:
    T_cw: np.ndarray
    X: np.ndarray
    outlier: np.ndarray
    cost_history: np.ndarray
    iterations: int
    trials: int
    stop_reason: int
    n_outliers: int
    initial_cost: float
    final_cost: float
    final_lambda: float
    solve_ms: float
    status: int = 0


class ResultBuffers:
    """Caller-owned output arrays for one nav_ba_result."""

    def __init__(self, prob: Problem, n_iter: int):
        self.T_cw = np.zeros((prob.n_cam, 12), np.float64)
        self.X = np.zeros((prob.n_pt, 3), np.float64)
        self.outlier = np.zeros(prob.n_obs, np.uint8)
        self.cost_history = np.full(max(n_iter, 0) + 1, np.nan, np.float64)
        self.c = NavBaResult()
        self.c.T_cw = ptr(self.T_cw, C.c_double)
        self.c.X = ptr(self.X, C.c_double)
        self.c.outlier = ptr(self.outlier, C.c_uint8)
        self.c.cost_history = ptr(self.cost_history, C.c_double)

    def result(self, status: int) -> Result:
        c = self.c
        return Result(self.T_cw, self.X, self.outlier, self.cost_history[: c.iterations + 1],
                      c.iterations, c.trials, c.stop_reason, c.n_outliers, c.initial_cost,
                      c.final_cost, c.final_lambda, c.solve_ms, status)
