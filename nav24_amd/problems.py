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


# Kannala-Brandt coefficients of the fisheye test camera (a mild wide-angle
# lens; the reference's Calibration reads them as mvParameters[4..7]).
KB8_COEFFS = (0.02, -0.005, 0.001, -0.0002)


def kb8_project(Xc: np.ndarray, K) -> np.ndarray:
    """Kannala-Brandt image coordinates of camera-frame points (n, 3), in
    double throughout (the generator's model; the edge's residual rounds its
    angles to float)."""
    fx, fy, cx, cy, k1, k2, k3, k4 = K
    x, y, z = Xc[:, 0], Xc[:, 1], Xc[:, 2]
    th = np.arctan2(np.hypot(x, y), z)
    psi = np.arctan2(y, x)
    r = th + k1 * th ** 3 + k2 * th ** 5 + k3 * th ** 7 + k4 * th ** 9
    return np.stack([fx * r * np.cos(psi) + cx, fy * r * np.sin(psi) + cy], 1)


def with_kb8(prob: Problem, coeffs=KB8_COEFFS, noise_px: float = 1.0, seed: int = 0) -> Problem:
    """The same scene seen through a Kannala-Brandt camera: every inlier
    observation re-projected from the generator's ground truth with the
    fisheye model plus N(0, noise_px) pixel noise, every outlier kept at a
    uniform random pixel; the initial state is unchanged."""
    rng = np.random.default_rng(seed)
    K = (prob.fx, prob.fy, prob.cx, prob.cy) + tuple(coeffs)
    T = prob.T_gt[prob.obs_cam]
    R = T[:, :9].reshape(-1, 3, 3)
    Xc = np.einsum("nij,nj->ni", R, prob.X_gt[prob.obs_pt]) + T[:, 9:]
    uv = kb8_project(Xc, K) + rng.normal(0.0, noise_px, (prob.n_obs, 2))
    out = prob.obs_outlier.astype(bool) if prob.obs_outlier is not None else np.zeros(prob.n_obs, bool)
    uv[out] = rng.uniform([0.0, 0.0], [2 * prob.cx, 2 * prob.cy], (int(out.sum()), 2))
    return Problem(T_cw=prob.T_cw, X=prob.X, obs_cam=prob.obs_cam, obs_pt=prob.obs_pt, obs_uv=uv,
                   fx=prob.fx, fy=prob.fy, cx=prob.cx, cy=prob.cy, cam_fixed=prob.cam_fixed, pt_fixed=prob.pt_fixed,
                   camera_model=1, kb=tuple(coeffs), T_gt=prob.T_gt, X_gt=prob.X_gt, obs_outlier=prob.obs_outlier)
