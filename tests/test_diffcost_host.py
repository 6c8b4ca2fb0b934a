"""Host-side pins of the autograd binding of the differentiable cost
(nav24_amd/diffcost.py): the pose-tangent chain rule BACost.backward applies
to the device kernel's left-tangent gradient, checked with torch-only costs
(no GPU).  For a cost C of the pose entries, the left-tangent gradient at
T = exp(xi) T_cw is g = d C(exp(d) T) / d d at d = 0; backward must turn it
into d C(exp(xi) T_cw) / d xi for any xi, not only xi = 0."""
import numpy as np
import pytest
import torch

from nav24_amd.diffcost import se3_left, tangent_to_entry_grad


def _poses(n, seed):
    g = torch.Generator().manual_seed(seed)
    xi0 = torch.randn(n, 6, generator=g, dtype=torch.float64)
    T0 = torch.cat([torch.eye(3, dtype=torch.float64).reshape(1, 9).repeat(n, 1),
                    torch.zeros(n, 3, dtype=torch.float64)], 1)
    return se3_left(xi0, T0)


def _cost(T, W):
    # a smooth, non-quadratic function of the 12 entries of every pose
    return (torch.sin(T @ W) ** 2).sum() + (T[:, 9:] ** 3).sum()


@pytest.mark.parametrize("scale", [0.0, 1e-4, 0.1, 0.5, 2.0])
def test_backward_chain_rule_matches_autograd(scale):
    n = 7
    T_cw = _poses(n, 1)
    W = torch.randn(12, 5, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    xi = scale * torch.randn(n, 6, generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    # reference: autograd of C(exp(xi) T_cw) with respect to xi
    x = xi.clone().requires_grad_(True)
    (g_ref,) = torch.autograd.grad(_cost(se3_left(x, T_cw), W), x)
    # what the device returns: the left-tangent gradient at T = exp(xi) T_cw
    T = se3_left(xi, T_cw).detach()
    d = torch.zeros(n, 6, dtype=torch.float64, requires_grad=True)
    (g_left,) = torch.autograd.grad(_cost(se3_left(d, T), W), d)
    # the backward mapping
    x2 = xi.clone().requires_grad_(True)
    (g_xi,) = torch.autograd.grad(se3_left(x2, T_cw), x2, tangent_to_entry_grad(g_left, T))
    np.testing.assert_allclose(g_xi.numpy(), g_ref.numpy(), rtol=1e-10, atol=1e-10)
    if scale >= 0.1:
        # away from xi = 0 the left-tangent gradient itself is not d/dxi
        assert np.abs(g_left.numpy() - g_ref.numpy()).max() > 1e-3


def test_se3_left_series_branch_is_continuous():
    T_cw = _poses(3, 4)
    u = torch.randn(3, 6, generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    u[:, :3] /= torch.linalg.norm(u[:, :3], dim=1, keepdim=True)
    for th in (1e-3 * (1 - 1e-9), 1e-3 * (1 + 1e-9)):  # theta^2 on both sides of 1e-6
        xi = u.clone()
        xi[:, :3] *= th
        Ta = se3_left(xi, T_cw)
        xi_b = u.clone()
        xi_b[:, :3] *= 1e-3
        np.testing.assert_allclose(Ta.numpy(), se3_left(xi_b, T_cw).numpy(), atol=1e-11)


def test_se3_left_gradient_finite_at_zero():
    T_cw = _poses(2, 6)
    x = torch.zeros(2, 6, dtype=torch.float64, requires_grad=True)
    (g,) = torch.autograd.grad(se3_left(x, T_cw).sum(), x)
    assert torch.isfinite(g).all()
