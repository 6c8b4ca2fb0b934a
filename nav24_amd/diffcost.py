"""Differentiable robust reprojection cost (include/nav_ba_cost.h) for PyTorch.

`CostBatch` prepares the observation structure of a batch of problems once;
`CostBatch.eval` runs the device cost and hand-written adjoint on torch
tensors (HIP device memory, torch's current stream).  `BACost` is the
`torch.autograd.Function` over it:

    cost = BACost.apply(xi, X, intr, T_cw, batch)      # sum of the problems' costs

with xi [n_cam, 6] the pose tangents at T_cw (left perturbation exp(xi) T_cw,
xi = (omega, upsilon), the update of the reference edge's pose vertex,
OptimizableTypes.hpp:221-227), X [n_pt, 3] the points and intr [8] =
(fx, fy, cx, cy, k1, k2, p1, p2).  The device kernel returns the
left-tangent gradient g at T = exp(xi) T_cw (the gradient for a perturbation
exp(d) T); backward maps it to d cost / d xi = J_l(xi)^T g, J_l the SE3 left
Jacobian, by the chain rule through `se3_left`, so the two agree only at
xi = 0 and every xi gets its true gradient.  There is no CPU fallback:
without the built library or a gfx950 device the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._loader import load_library
from .abi import NAV_BA_OK, STATUS_NAMES, NavBaProblem, Problem

NAV_BA_COST_F64 = 0
NAV_BA_COST_F32 = 1

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = load_library("libnav24_ba.so")
        H, vp = C.c_void_p, C.c_void_p
        L.nav_ba_cost_create.argtypes = [C.POINTER(H), C.c_int32]
        L.nav_ba_cost_create.restype = C.c_int32
        L.nav_ba_cost_destroy.argtypes = [H]
        L.nav_ba_cost_destroy.restype = None
        L.nav_ba_cost_last_error.argtypes = [H]
        L.nav_ba_cost_last_error.restype = C.c_char_p
        L.nav_ba_cost_prepare.argtypes = [H, C.POINTER(NavBaProblem), C.c_int32, C.c_int32, C.c_double]
        L.nav_ba_cost_prepare.restype = C.c_int32
        L.nav_ba_cost_eval.argtypes = [H, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.nav_ba_cost_eval.restype = C.c_int32
        L.nav_ba_cost_sizes.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]
        L.nav_ba_cost_sizes.restype = C.c_int32
        _lib = L
    return _lib


class CostError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


@dataclass
class CostResult:
    cost: "object"             # [n_problems]
    g_pose: Optional[object]   # [n_cam, 6]
    g_X: Optional[object]      # [n_pt, 3]
    g_intr: Optional[object]   # [8]


class CostBatch:
    """Observation structure of a batch of problems, resident on one device."""

    def __init__(self, problems: Sequence[Problem], robust: bool = True, huber_delta: float = float(np.sqrt(5.991)),
                 device: int = 0):
        self._h = C.c_void_p()
        st = lib().nav_ba_cost_create(C.byref(self._h), device)
        if st != NAV_BA_OK:
            raise CostError(st, f"nav_ba_cost_create(device={device}) failed")
        self.device = device
        self.problems = list(problems)
        arr = (NavBaProblem * len(self.problems))(*[p.c_struct() for p in self.problems])
        self._check(lib().nav_ba_cost_prepare(self._h, arr, len(self.problems), int(bool(robust)), float(huber_delta)))
        n = [C.c_int32() for _ in range(4)]
        self._check(lib().nav_ba_cost_sizes(self._h, *[C.byref(v) for v in n]))
        self.n_cam, self.n_pt, self.n_obs, self.n_problems = (v.value for v in n)

    def _check(self, st):
        if st != NAV_BA_OK:
            raise CostError(st, lib().nav_ba_cost_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().nav_ba_cost_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eval(self, T_cw, X, intr, grad: bool = True) -> CostResult:
        """Cost per problem and (if grad) the gradients, as torch tensors on
        the inputs' device, computed on torch's current stream."""
        import torch
        dt = T_cw.dtype
        if dt not in (torch.float64, torch.float32) or X.dtype != dt or intr.dtype != dt:
            raise CostError(2, "T_cw, X and intr must share dtype float64 or float32")
        for name, t, shape in (("T_cw", T_cw, (self.n_cam, 12)), ("X", X, (self.n_pt, 3)), ("intr", intr, (8,))):
            if not t.is_cuda or tuple(t.shape) != shape or not t.is_contiguous():
                raise CostError(2, f"{name} must be a contiguous device tensor of shape {shape}")
        dev = T_cw.device
        cost = torch.empty(self.n_problems, dtype=dt, device=dev)
        gp = torch.empty((self.n_cam, 6), dtype=dt, device=dev) if grad else None
        gx = torch.empty((self.n_pt, 3), dtype=dt, device=dev) if grad else None
        gi = torch.empty(8, dtype=dt, device=dev) if grad else None
        ptr = (lambda t: C.c_void_p(t.data_ptr()) if t is not None else None)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self._check(lib().nav_ba_cost_eval(self._h, NAV_BA_COST_F64 if dt == torch.float64 else NAV_BA_COST_F32,
                                           ptr(T_cw), ptr(X), ptr(intr), ptr(cost), ptr(gp), ptr(gx), ptr(gi),
                                           C.c_void_p(stream)))
        return CostResult(cost, gp, gx, gi)


def se3_left(xi, T_cw):
    """exp(xi) T_cw for every pose row [R (9) | t (3)], xi = (omega, upsilon), in torch.
    Below theta^2 = 1e-6 the Rodrigues and V coefficients are their series to
    theta^2 (error below 1e-13), so the map and its autograd derivative are
    finite at xi = 0."""
    import torch
    w, u = xi[:, :3], xi[:, 3:]
    th2 = (w * w).sum(1, keepdim=True)
    small = th2 < 1e-6
    th2s = torch.where(small, torch.ones_like(th2), th2)
    ths = torch.sqrt(th2s)
    a = torch.where(small, 1.0 - th2 / 6.0, torch.sin(ths) / ths)
    b = torch.where(small, 0.5 - th2 / 24.0, (1 - torch.cos(ths)) / th2s)
    c = torch.where(small, 1.0 / 6.0 - th2 / 120.0, (ths - torch.sin(ths)) / (th2s * ths))
    Z = torch.zeros_like(w[:, 0])
    K = torch.stack([torch.stack([Z, -w[:, 2], w[:, 1]], 1), torch.stack([w[:, 2], Z, -w[:, 0]], 1),
                     torch.stack([-w[:, 1], w[:, 0], Z], 1)], 1)
    K2 = K @ K
    eye = torch.eye(3, dtype=xi.dtype, device=xi.device).expand_as(K)
    Rw = eye + a[..., None] * K + b[..., None] * K2
    V = eye + b[..., None] * K + c[..., None] * K2
    R = T_cw[:, :9].reshape(-1, 3, 3)
    t = T_cw[:, 9:]
    Rn = Rw @ R
    tn = (Rw @ t[..., None])[..., 0] + (V @ u[..., None])[..., 0]
    return torch.cat([Rn.reshape(-1, 9), tn], 1).contiguous()


def tangent_to_entry_grad(g, T):
    """A gradient with respect to the 12 entries [R | t] of each pose T that
    acts on every tangent perturbation exactly as the left-tangent gradient g
    = (g_omega, g_upsilon) does.  A left perturbation d moves R by [d_omega]x R
    and t by [d_omega]x t + d_upsilon, so g . d = a . d_omega + g_upsilon . dt
    with a = g_omega - t x g_upsilon, and a . d_omega = <[a]x R / 2, dR>
    (Frobenius) because [d_omega]x = dR R^T."""
    import torch
    R = T[:, :9].reshape(-1, 3, 3)
    t = T[:, 9:]
    a = g[:, :3] - torch.linalg.cross(t, g[:, 3:], dim=1)
    Z = torch.zeros_like(a[:, 0])
    A = torch.stack([torch.stack([Z, -a[:, 2], a[:, 1]], 1), torch.stack([a[:, 2], Z, -a[:, 0]], 1),
                     torch.stack([-a[:, 1], a[:, 0], Z], 1)], 1)
    GR = 0.5 * (A @ R)
    return torch.cat([GR.reshape(-1, 9), g[:, 3:]], 1)


try:
    import torch

    class BACost(torch.autograd.Function):
        """Sum over the batch of the robust reprojection costs, differentiable
        in the pose tangents xi, the points X and the intrinsics/distortion."""

        @staticmethod
        def forward(ctx, xi, X, intr, T_cw, batch):
            with torch.no_grad():
                T = se3_left(xi, T_cw)
                need = any(ctx.needs_input_grad[:3])
                r = batch.eval(T, X.contiguous(), intr.contiguous(), grad=need)
            if need:
                ctx.save_for_backward(r.g_pose, r.g_X, r.g_intr, xi.detach(), T_cw.detach(), T)
            ctx.need = need
            return r.cost.sum()

        @staticmethod
        def backward(ctx, g):
            gp, gx, gi, xi, T_cw, T = ctx.saved_tensors
            g_xi = None
            if ctx.needs_input_grad[0]:
                # d cost / d xi = (d T / d xi)^T G_T, G_T the entry gradient of the
                # left-tangent gradient gp at T: J_l(xi)^T gp
                with torch.enable_grad():
                    x = xi.clone().requires_grad_(True)
                    Tx = se3_left(x, T_cw)
                    (g_xi,) = torch.autograd.grad(Tx, x, tangent_to_entry_grad(gp, T))
                g_xi = g * g_xi
            return g_xi, g * gx, g * gi, None, None
except ImportError:  # torch is optional for the C ABI users
    BACost = None
