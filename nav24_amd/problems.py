"""Seeded synthetic BA problems (binding of csrc/problem_gen.cpp).

Workloads (seed 0 unless stated):
  0  10 keyframes on a 2 m circle looking at a 4 m cube of 1 500 points
  1  60 keyframes on a random walk, 20 000 points in a 40x40x10 m box
  2  300 keyframes on a 30 m-radius loop, 150 000 points within 15 m
  3  one batched window (window w: seed = base + w)
  4  20 keyframes, 5 000 points, radial-tangential distortion
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import Problem
from ._loader import load_library

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)


class GenParams(C.Structure):
    _fields_ = [
        ("geometry", C.c_int32), ("n_cam", C.c_int32), ("n_pt", C.c_int32),
        ("noise_px", C.c_double), ("outlier_rate", C.c_double), ("rot_sigma", C.c_double),
        ("trans_sigma", C.c_double), ("point_sigma", C.c_double), ("n_fixed", C.c_int32),
        ("obs_min", C.c_int32), ("obs_max", C.c_int32), ("distortion", C.c_int32),
        ("seed", C.c_uint64),
    ]


class GenProblem(C.Structure):
    _fields_ = [
        ("n_cam", C.c_int32), ("n_pt", C.c_int32), ("n_obs", C.c_int32),
        ("T_cw", _dp), ("cam_fixed", _u8p), ("X", _dp), ("pt_fixed", _u8p),
        ("obs_cam", _i32p), ("obs_pt", _i32p), ("obs_uv", _dp), ("obs_outlier", _u8p),
        ("T_gt", _dp), ("X_gt", _dp),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("dist", C.c_double * 4), ("width", C.c_int32), ("height", C.c_int32),
    ]


_lib = None


def _gen():
    global _lib
    if _lib is None:
        lib = load_library("libnav24_gen.so")
        lib.nav_gen_params_for.argtypes = [C.c_int32, C.c_uint64, C.POINTER(GenParams)]
        lib.nav_gen_params_for.restype = C.c_int32
        lib.nav_gen_make.argtypes = [C.POINTER(GenParams), C.POINTER(C.POINTER(GenProblem))]
        lib.nav_gen_make.restype = C.c_int32
        lib.nav_gen_free.argtypes = [C.POINTER(GenProblem)]
        lib.nav_gen_free.restype = None
        lib.nav_gen_mt_draw.argtypes = [C.c_uint64, C.c_int32]
        lib.nav_gen_mt_draw.restype = C.c_uint64
        lib.nav_gen_normals.argtypes = [C.c_uint64, C.c_int32, _dp]
        lib.nav_gen_normals.restype = None
        _lib = lib
    return _lib


def params_for(workload: int, seed: int = 0) -> GenParams:
    p = GenParams()
    if _gen().nav_gen_params_for(workload, seed, C.byref(p)) != 0:
        raise ValueError(f"unknown workload {workload}")
    return p


def make(params: GenParams) -> Problem:
    out = C.POINTER(GenProblem)()
    if _gen().nav_gen_make(C.byref(params), C.byref(out)) != 0:
        raise ValueError("problem generation failed")
    g = out.contents
    try:
        nc, npt, no = g.n_cam, g.n_pt, g.n_obs

        def arr(p, n, dt, shape):
            if n == 0:
                return np.zeros(shape, dt)
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).reshape(shape).copy()

        prob = Problem(
            T_cw=arr(g.T_cw, nc * 12, np.float64, (nc, 12)),
            X=arr(g.X, npt * 3, np.float64, (npt, 3)),
            obs_cam=arr(g.obs_cam, no, np.int32, (no,)),
            obs_pt=arr(g.obs_pt, no, np.int32, (no,)),
            obs_uv=arr(g.obs_uv, no * 2, np.float64, (no, 2)),
            fx=g.fx, fy=g.fy, cx=g.cx, cy=g.cy,
            cam_fixed=arr(g.cam_fixed, nc, np.uint8, (nc,)),
            pt_fixed=arr(g.pt_fixed, npt, np.uint8, (npt,)),
            T_gt=arr(g.T_gt, nc * 12, np.float64, (nc, 12)),
            X_gt=arr(g.X_gt, npt * 3, np.float64, (npt, 3)),
            obs_outlier=arr(g.obs_outlier, no, np.uint8, (no,)),
            dist=np.array(list(g.dist), np.float64),
        )
    finally:
        _gen().nav_gen_free(out)
    return prob


def workload(workload: int, seed: int = 0, **overrides) -> Problem:
    """Generate a named workload; keyword overrides change GenParams fields
    (e.g. n_pt=200 for a smaller problem of the same geometry)."""
    p = params_for(workload, seed)
    for k, v in overrides.items():
        setattr(p, k, v)
    return make(p)


def mt_draw(seed: int, k: int) -> int:
    return int(_gen().nav_gen_mt_draw(seed, k))


def normals(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    _gen().nav_gen_normals(seed, n, out.ctypes.data_as(_dp))
    return out
