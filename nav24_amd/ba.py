"""Python binding of the MI355X BA back-end (libnav24_ba.so, include/nav_ba.h).

`GraphOptimHip` mirrors the reference back-end `NAV24::BE::GraphOptim`
(core/backEnd/BE_GraphOptim.hpp:19-34): `solve(problem)` refines the poses
and points of a problem in one call.  The problem is the flat structure of
arrays of `nav24_amd.abi.Problem` instead of PR_VBA's object graph
(core/backEnd/Problem.hpp:41-47).  Errors raise `BAError` carrying the C
status; there is no CPU fallback — without the built library or a gfx950
device the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from ._loader import load_library
from .abi import (NAV_BA_EMPTY, NAV_BA_OK, NAV_BA_STOP_EMPTY, STATUS_NAMES, NavBaOptions, NavBaProblem, NavBaResult,
                  Options, Problem, Result, ResultBuffers)

_dp = C.POINTER(C.c_double)

# Every entry point declared in include/nav_ba.h and include/nav_ba_debug.h.
EXPORTED = [
    "nav_ba_version", "nav_ba_create", "nav_ba_destroy", "nav_ba_last_error",
    "nav_ba_default_options", "nav_ba_solve", "nav_ba_prepare", "nav_ba_run", "nav_ba_fetch",
    "nav_ba_prepare_batched", "nav_ba_fetch_batched", "nav_ba_solve_batched",
    "nav_ba_debug_reduced_system", "nav_ba_set_timing", "nav_ba_get_timing", "nav_ba_debug_set_solver",
    "nav_ba_debug_set_schur_chunk", "nav_ba_debug_get_schur_chunk", "nav_ba_debug_check_layout",
    "nav_ba_debug_time_schur_rows", "nav_ba_debug_time_stage", "nav_ba_debug_pcg_stats", "nav_ba_debug_set_schur_kernel", "nav_ba_debug_get_schur_kernel",
    "nav_ba_debug_prepare_cached",
]

STAGES = ["lin_pt", "lin_cam", "lambda_init", "hinv", "schur_row", "schur_reduce", "cholesky",
          "backsub", "update_cam", "cost", "accept", "flags"]


class NavBaTiming(C.Structure):
    _fields_ = [("ms", C.c_double * 16), ("launches", C.c_int32 * 16)]

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = load_library("libnav24_ba.so")
        H = C.c_void_p
        L.nav_ba_version.restype = C.c_char_p
        L.nav_ba_version.argtypes = []
        L.nav_ba_create.argtypes = [C.POINTER(H), C.c_int32]
        L.nav_ba_create.restype = C.c_int32
        L.nav_ba_destroy.argtypes = [H]
        L.nav_ba_destroy.restype = None
        L.nav_ba_last_error.argtypes = [H]
        L.nav_ba_last_error.restype = C.c_char_p
        L.nav_ba_default_options.argtypes = [C.POINTER(NavBaOptions)]
        L.nav_ba_default_options.restype = None
        L.nav_ba_solve.argtypes = [H, C.POINTER(NavBaProblem), C.POINTER(NavBaOptions), C.POINTER(NavBaResult)]
        L.nav_ba_solve.restype = C.c_int32
        L.nav_ba_prepare.argtypes = [H, C.POINTER(NavBaProblem), C.POINTER(NavBaOptions)]
        L.nav_ba_prepare.restype = C.c_int32
        L.nav_ba_run.argtypes = [H]
        L.nav_ba_run.restype = C.c_int32
        L.nav_ba_fetch.argtypes = [H, C.POINTER(NavBaResult)]
        L.nav_ba_fetch.restype = C.c_int32
        L.nav_ba_prepare_batched.argtypes = [H, C.POINTER(NavBaProblem), C.c_int32, C.POINTER(NavBaOptions)]
        L.nav_ba_prepare_batched.restype = C.c_int32
        L.nav_ba_fetch_batched.argtypes = [H, C.POINTER(NavBaResult), C.c_int32]
        L.nav_ba_fetch_batched.restype = C.c_int32
        L.nav_ba_solve_batched.argtypes = [H, C.POINTER(NavBaProblem), C.c_int32, C.POINTER(NavBaOptions),
                                           C.POINTER(NavBaResult)]
        L.nav_ba_solve_batched.restype = C.c_int32
        L.nav_ba_debug_reduced_system.argtypes = [H, C.c_double, _dp, _dp, C.c_int32, C.POINTER(C.c_int32)]
        L.nav_ba_debug_reduced_system.restype = C.c_int32
        L.nav_ba_set_timing.argtypes = [H, C.c_int32]
        L.nav_ba_set_timing.restype = C.c_int32
        L.nav_ba_get_timing.argtypes = [H, C.POINTER(NavBaTiming)]
        L.nav_ba_get_timing.restype = C.c_int32
        L.nav_ba_debug_set_solver.argtypes = [H, C.c_int32]
        L.nav_ba_debug_set_solver.restype = C.c_int32
        L.nav_ba_debug_set_schur_chunk.argtypes = [H, C.c_int32]
        L.nav_ba_debug_set_schur_chunk.restype = C.c_int32
        L.nav_ba_debug_get_schur_chunk.argtypes = [H, C.POINTER(C.c_int32)]
        L.nav_ba_debug_get_schur_chunk.restype = C.c_int32
        L.nav_ba_debug_prepare_cached.argtypes = [H, C.POINTER(C.c_int32)]
        L.nav_ba_debug_prepare_cached.restype = C.c_int32
        L.nav_ba_debug_set_schur_kernel.argtypes = [H, C.c_int32]
        L.nav_ba_debug_set_schur_kernel.restype = C.c_int32
        L.nav_ba_debug_get_schur_kernel.argtypes = [H, C.POINTER(C.c_int32)]
        L.nav_ba_debug_get_schur_kernel.restype = C.c_int32
        L.nav_ba_debug_check_layout.argtypes = [C.POINTER(NavBaProblem), C.c_int32, C.POINTER(NavBaOptions),
                                                C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                                                C.POINTER(C.c_int32), C.c_char_p, C.c_int32]
        L.nav_ba_debug_check_layout.restype = C.c_int32
        L.nav_ba_debug_time_schur_rows.argtypes = [H, C.c_double, C.c_int32, C.POINTER(C.c_double)]
        L.nav_ba_debug_time_schur_rows.restype = C.c_int32
        L.nav_ba_debug_time_stage.argtypes = [H, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_double)]
        L.nav_ba_debug_time_stage.restype = C.c_int32
        L.nav_ba_debug_pcg_stats.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.nav_ba_debug_pcg_stats.restype = C.c_int32
        _lib = L
    return _lib


class BAError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def default_options() -> Options:
    o = NavBaOptions()
    lib().nav_ba_default_options(C.byref(o))
    return Options(n_iter=o.n_iter, robust=bool(o.robust), huber_delta=o.huber_delta,
                   lambda_init=o.lambda_init, precision=o.precision, compat_fp32_io=bool(o.compat_fp32_io),
                   outlier_chi2=o.outlier_chi2, max_trials=o.max_trials)


class GraphOptimHip:
    """GPU bundle-adjustment back-end on one gfx950 device (one HIP stream)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        st = lib().nav_ba_create(C.byref(self._h), device)
        if st != NAV_BA_OK:
            raise BAError(st, f"nav_ba_create(device={device}) failed")
        self._keep = None
        self._prob: Optional[Problem] = None
        self._opts: Optional[Options] = None

    def close(self):
        if self._h:
            lib().nav_ba_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, allow_empty: bool = False) -> int:
        if st == NAV_BA_OK or (allow_empty and st == NAV_BA_EMPTY):
            return st
        raise BAError(st, lib().nav_ba_last_error(self._h).decode())

    def solve(self, prob: Problem, opts: Options) -> Result:
        """GraphOptim::solve (BE_GraphOptim.cpp:31-77) in one call."""
        ps, os_ = prob.c_struct(), opts.c_struct()
        rb = ResultBuffers(prob, opts.n_iter)
        st = lib().nav_ba_solve(self._h, C.byref(ps), C.byref(os_), C.byref(rb.c))
        self._check(st, allow_empty=True)
        return rb.result(st)

    def prepare(self, prob: Problem, opts: Options) -> None:
        ps, os_ = prob.c_struct(), opts.c_struct()
        self._keep = (ps, os_)
        self._prob, self._opts = prob, opts
        self._check(lib().nav_ba_prepare(self._h, C.byref(ps), C.byref(os_)))

    def run(self) -> int:
        return self._check(lib().nav_ba_run(self._h), allow_empty=True)

    def fetch(self) -> Result:
        if self._prob is None:  # no prepare on this handle: the library says so
            self._check(lib().nav_ba_fetch(self._h, C.byref(NavBaResult())))
        rb = ResultBuffers(self._prob, self._opts.n_iter)
        st = lib().nav_ba_fetch(self._h, C.byref(rb.c))
        self._check(st, allow_empty=True)
        return rb.result(st)

    # -- batches of independent windows (FE_SlamMonoV.cpp:161-165, one solve per window)

    def _batch(self, windows: Sequence[Problem], opts: Options):
        arr = (NavBaProblem * len(windows))(*[w.c_struct() for w in windows])
        return arr, opts.c_struct()

    def _results(self, windows: Sequence[Problem], n_iter: int):
        bufs = [ResultBuffers(w, n_iter) for w in windows]
        return bufs, (NavBaResult * len(windows))(*[b.c for b in bufs])

    def solve_batched(self, windows: Sequence[Problem], opts: Options) -> List[Result]:
        """One GraphOptim::solve per window, all windows in one device solve."""
        arr, os_ = self._batch(windows, opts)
        bufs, res = self._results(windows, opts.n_iter)
        st = lib().nav_ba_solve_batched(self._h, arr, len(windows), C.byref(os_), res)
        self._check(st, allow_empty=True)
        return [self._window_result(b, res[i]) for i, b in enumerate(bufs)]

    def prepare_batched(self, windows: Sequence[Problem], opts: Options) -> None:
        arr, os_ = self._batch(windows, opts)
        self._keep = (arr, os_)
        self._prob, self._opts = list(windows), opts
        self._check(lib().nav_ba_prepare_batched(self._h, arr, len(windows), C.byref(os_)))

    def fetch_batched(self) -> List[Result]:
        windows = self._prob
        bufs, res = self._results(windows, self._opts.n_iter)
        st = lib().nav_ba_fetch_batched(self._h, res, len(windows))
        self._check(st, allow_empty=True)
        return [self._window_result(b, res[i]) for i, b in enumerate(bufs)]

    @staticmethod
    def _window_result(b: ResultBuffers, c: NavBaResult) -> Result:
        b.c = c
        st = NAV_BA_EMPTY if c.stop_reason == NAV_BA_STOP_EMPTY else NAV_BA_OK
        return b.result(st)

    def set_timing(self, enable: bool) -> None:
        """Stage-by-stage launches with HIP events around each stage (no graphs)."""
        self._check(lib().nav_ba_set_timing(self._h, int(bool(enable))))

    def timing(self) -> dict:
        """{stage: (total_ms, launches)} accumulated since set_timing(True)."""
        t = NavBaTiming()
        self._check(lib().nav_ba_get_timing(self._h, C.byref(t)))
        return {name: (t.ms[i], t.launches[i]) for i, name in enumerate(STAGES)}

    def set_solver(self, mode: int) -> None:
        """Reduced-system solver for the next prepare: 0 auto, 1 64x64 tiles, 2 6x6 camera
        blocks in one workgroup, 3 the per-window kernel in LDS, 4 conjugate gradients."""
        self._check(lib().nav_ba_debug_set_solver(self._h, int(mode)))

    def set_schur_chunk(self, chunk: int) -> None:
        """Observations per Schur work item for the next prepare: 0 automatic
        (tuned at prepare on large problems), or a multiple of 64 up to 4096."""
        self._check(lib().nav_ba_debug_set_schur_chunk(self._h, int(chunk)))

    def set_schur_kernel(self, kernel: int) -> None:
        """Schur-row kernel for the next prepare: 0 automatic, 1 rounds of 256
        observations (lanes bound per round), 2 producer/consumer (lanes bound
        per work item)."""
        self._check(lib().nav_ba_debug_set_schur_kernel(self._h, int(kernel)))

    @property
    def schur_kernel(self) -> int:
        """Schur-row kernel chosen by the last prepare (1 or 2; 0 if none)."""
        v = C.c_int32(0)
        self._check(lib().nav_ba_debug_get_schur_kernel(self._h, C.byref(v)))
        return v.value

    @property
    def prepare_cached(self) -> bool:
        """Whether the last prepare reused the cached index (the same structure
        as the prepare before it; only the new state was uploaded)."""
        v = C.c_int32(0)
        self._check(lib().nav_ba_debug_prepare_cached(self._h, C.byref(v)))
        return bool(v.value)

    def time_schur_rows(self, lam: float = 1e-4, reps: int = 10) -> float:
        """Average device time (us) of the Schur-row kernel over all windows of
        the prepared problem at its initial linearisation (every window running)."""
        v = C.c_double(0.0)
        self._check(lib().nav_ba_debug_time_schur_rows(self._h, lam, int(reps), C.byref(v)), allow_empty=True)
        return v.value

    def pcg_stats(self):
        """(iterations, relative residual) of the last trial's conjugate-gradient
        solve; (0, 0.0) when the problem is factored instead."""
        it, res = C.c_int32(0), C.c_double(0.0)
        self._check(lib().nav_ba_debug_pcg_stats(self._h, C.byref(it), C.byref(res)))
        return it.value, res.value

    def time_stage(self, stage: str, lam: float = 1e-4, reps: int = 10) -> float:
        """Average device time (us) of one stage's launch ("lin_pt", "lin_cam",
        "cost" or "schur_row") over all windows at the initial linearisation,
        every window running and doing the stage's work."""
        v = C.c_double(0.0)
        self._check(lib().nav_ba_debug_time_stage(self._h, STAGES.index(stage), lam, int(reps), C.byref(v)),
                    allow_empty=True)
        return v.value

    @property
    def schur_chunk(self) -> int:
        """Observations per Schur work item chosen by the last prepare."""
        v = C.c_int32(0)
        self._check(lib().nav_ba_debug_get_schur_chunk(self._h, C.byref(v)))
        return v.value

    def debug_reduced_system(self, lam: float):
        n = C.c_int32(0)
        self._check(lib().nav_ba_debug_reduced_system(self._h, lam, None, None, 0, C.byref(n)), allow_empty=True)
        S = np.zeros((max(n.value, 1), max(n.value, 1)))
        g = np.zeros(max(n.value, 1))
        self._check(lib().nav_ba_debug_reduced_system(self._h, lam, S.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                                                      n.value, C.byref(n)), allow_empty=True)
        return S[: n.value, : n.value], g[: n.value]


def check_layout(windows: Sequence[Problem], opts: Options, batched: bool = False, chunk: int = 0,
                 kernel: int = 1) -> dict:
    """Host-only check of the structure prepare builds for these windows (no
    device needed) with Schur-row kernel `kernel` (1 or 2): raises BAError with
    the first violation, else returns the Schur pair-entry and round counts."""
    arr = (NavBaProblem * len(windows))(*[w.c_struct() for w in windows])
    os_ = opts.c_struct()
    n_pairs, n_rounds = C.c_int64(0), C.c_int32(0)
    msg = C.create_string_buffer(256)
    st = lib().nav_ba_debug_check_layout(arr, len(windows), C.byref(os_), int(bool(batched)), int(chunk),
                                         int(kernel), C.byref(n_pairs), C.byref(n_rounds), msg, 256)
    if st != NAV_BA_OK:
        raise BAError(st, msg.value.decode())
    return {"pairs": n_pairs.value, "rounds": n_rounds.value}


def version() -> str:
    return lib().nav_ba_version().decode()
