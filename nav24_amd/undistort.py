"""Key-point undistortion of the radial-tangential pinhole camera and of the
Kannala-Brandt fisheye camera on the GPU.

Host-side mirror of PinholeRadTan (core/sensor/camera/models/PinholeRadTan.hpp:11-21,
PinholeRadTan.cpp:11-24): built from the same parameter vector (fx, fy, cx,
cy, k1, k2, p1, p2, float) and exposing `UndistortKeyPoints`, which returns
the key points with their undistorted coordinates, as the reference's
cv::undistortPoints call does.  The arithmetic runs in the C ABI
(include/nav_kp_undistort.h); without the built library or a gfx950 device
the calls raise.

    cam = PinholeRadTanHip([fx, fy, cx, cy, k1, k2, p1, p2])
    cam.setRectificationMat(R); cam.setProjectionMat(P)   # optional, as Calibration.cpp:88-93
    uv_ud = cam.UndistortKeyPoints(uv)            # numpy (n, 2) float32, host
    cam.undistort_device(uv_t, out=uv_t)          # torch (n, 2) float32 on cuda, in place

KannalaBrandt8Hip (KannalaBrandt8.cpp:231-242, cv::fisheye::undistortPoints)
has the same calls, with params fx fy cx cy k1 k2 k3 k4.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from ._loader import load_library
from .abi import NAV_BA_OK, STATUS_NAMES

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = load_library("libnav24_ba.so")
        fp = C.POINTER(C.c_float)
        L.nav_ba_kp_undistort.argtypes = [C.c_void_p, C.c_int32, fp, C.c_void_p, C.c_void_p]
        L.nav_ba_kp_undistort.restype = C.c_int32
        L.nav_ba_kp_undistort_host.argtypes = [fp, C.c_int32, fp, fp, C.c_int32]
        L.nav_ba_kp_undistort_host.restype = C.c_int32
        L.nav_ba_kp_undistort_fisheye.argtypes = [C.c_void_p, C.c_int32, fp, C.c_void_p, C.c_void_p]
        L.nav_ba_kp_undistort_fisheye.restype = C.c_int32
        L.nav_ba_kp_undistort_fisheye_host.argtypes = [fp, C.c_int32, fp, fp, C.c_int32]
        L.nav_ba_kp_undistort_fisheye_host.restype = C.c_int32
        dp = C.POINTER(C.c_double)
        L.nav_ba_kp_undistort_rect.argtypes = [C.c_int32, C.c_void_p, C.c_int32, fp, dp, dp, C.c_int32, C.c_void_p,
                                               C.c_void_p]
        L.nav_ba_kp_undistort_rect.restype = C.c_int32
        L.nav_ba_kp_undistort_rect_host.argtypes = [C.c_int32, fp, C.c_int32, fp, dp, dp, C.c_int32, fp, C.c_int32]
        L.nav_ba_kp_undistort_rect_host.restype = C.c_int32
        _lib = L
    return _lib


def _check(st: int, what: str) -> None:
    if st != NAV_BA_OK:
        raise RuntimeError(f"{what}: {STATUS_NAMES.get(st, st)}")


class PinholeRadTanHip:
    """Radial-tangential pinhole camera (the reference's PinholeRadTan) whose
    key-point undistortion runs on device `device`."""

    model = 0  # NAV_KP_RADTAN

    def __init__(self, params: Sequence[float], device: int = 0):
        p = np.asarray(params, np.float32).reshape(-1)
        if p.size < 4:
            raise ValueError("params: fx fy cx cy [k1 k2 p1 p2]")
        full = np.zeros(8, np.float32)  # missing distortion coefficients are zero, as mD_cv's
        full[: min(8, p.size)] = p[:8]
        self.params = full
        self.device = device
        self.R: Optional[np.ndarray] = None  # identity (GeometricCamera.h:47,65)
        self.P: Optional[np.ndarray] = None  # K (GeometricCamera.h:48,66)

    def setRectificationMat(self, R) -> None:
        """The camera's rectification (3x3), GeometricCamera::setRectificationMat."""
        self.R = np.ascontiguousarray(R, np.float64).reshape(3, 3)

    def setProjectionMat(self, P) -> None:
        """The camera's projection (3x3 or 3x4), GeometricCamera::setProjectionMat."""
        P = np.ascontiguousarray(P, np.float64)
        if P.shape not in ((3, 3), (3, 4)):
            raise ValueError("P: 3x3 or 3x4")
        self.P = P

    def _rp(self):
        dp = C.POINTER(C.c_double)
        R = None if self.R is None else self.R.ctypes.data_as(dp)
        P = None if self.P is None else self.P.ctypes.data_as(dp)
        return R, P, 3 if self.P is None else self.P.shape[1]

    def UndistortKeyPoints(self, uv: np.ndarray) -> np.ndarray:
        """(n, 2) key points (u, v) -> their undistorted coordinates, float32."""
        a = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.empty_like(a)
        fp = C.POINTER(C.c_float)
        st = lib().nav_ba_kp_undistort_rect_host(self.model, a.ctypes.data_as(fp), a.shape[0],
                                                 self.params.ctypes.data_as(fp), *self._rp(),
                                                 out.ctypes.data_as(fp), self.device)
        _check(st, "nav_ba_kp_undistort_rect_host")
        return out

    def undistort_device(self, uv, out: Optional[object] = None):
        """Device tensors: uv (n, 2) float32 contiguous on cuda; the result goes
        to `out` (may be uv itself) or a new tensor, on torch's current stream."""
        import torch

        if uv.dtype != torch.float32 or not uv.is_cuda or not uv.is_contiguous() or uv.dim() != 2 or uv.shape[1] != 2:
            raise ValueError("uv: contiguous (n, 2) float32 cuda tensor")
        if out is None:
            out = torch.empty_like(uv)
        elif out.shape != uv.shape or out.dtype != torch.float32 or not out.is_contiguous() or not out.is_cuda:
            raise ValueError("out: contiguous (n, 2) float32 cuda tensor")
        stream = torch.cuda.current_stream(uv.device).cuda_stream
        st = lib().nav_ba_kp_undistort_rect(self.model, C.c_void_p(uv.data_ptr()), uv.shape[0],
                                            self.params.ctypes.data_as(C.POINTER(C.c_float)), *self._rp(),
                                            C.c_void_p(out.data_ptr()), C.c_void_p(stream))
        _check(st, "nav_ba_kp_undistort_rect")
        return out


class KannalaBrandt8Hip(PinholeRadTanHip):
    """Kannala-Brandt fisheye camera (the reference's KannalaBrandt8,
    core/sensor/camera/models/KannalaBrandt8.cpp:231-242): params fx fy cx cy
    k1 k2 k3 k4; the same calls."""

    model = 1  # NAV_KP_KB8
