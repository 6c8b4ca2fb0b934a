"""Device cost and hand-written adjoint (include/nav_ba_cost.h) against the
FP64 CPU oracle (oracle/cost_oracle.py), FP32 against FP64, and the
torch.autograd.Function against FP64 central differences (gradcheck)."""
import numpy as np
import pytest

from nav24_amd import problems
from oracle import cost_oracle as co
from parity_util import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    from nav24_amd import diffcost
    from nav24_amd.abi import NAV_BA_NO_DEVICE
    import ctypes
    h = ctypes.c_void_p()
    if diffcost.lib().nav_ba_cost_create(ctypes.byref(h), 0) == NAV_BA_NO_DEVICE:
        pytest.skip("no gfx950 device")
    diffcost.lib().nav_ba_cost_destroy(h)
    return torch.device("cuda:0")


def batch_problems(n=3, n_pt=400, n_cam=6):
    out = []
    for s in range(n):
        p = problems.params_for(4, s)
        p.n_pt, p.n_cam = n_pt, n_cam
        out.append(problems.make(p))
    return out


def concat(probs):
    T = np.concatenate([p.T_cw for p in probs])
    X = np.concatenate([p.X for p in probs])
    intr = np.array([probs[0].fx, probs[0].fy, probs[0].cx, probs[0].cy, *probs[0].dist])
    return T, X, intr


def oracle_batch(probs, T, X, intr, robust=True):
    """Per-problem costs and the concatenated gradients from the CPU oracle."""
    costs, gps, gxs, gi = [], [], [], np.zeros(8)
    c0 = p0 = 0
    for p in probs:
        c, gp, gx, g8 = co.cost_and_grad(T[c0:c0 + p.n_cam], X[p0:p0 + p.n_pt], p.obs_cam, p.obs_pt, p.obs_uv,
                                         intr, robust)
        costs.append(c); gps.append(gp); gxs.append(gx); gi += g8
        c0 += p.n_cam; p0 += p.n_pt
    return np.array(costs), np.concatenate(gps), np.concatenate(gxs), gi


@pytest.mark.parametrize("robust", [True, False])
def test_fp64_cost_and_gradient_match_oracle(dev, robust):
    from nav24_amd.diffcost import CostBatch
    say(f"cost fp64 vs oracle, robust={robust}")
    probs = batch_problems()
    T, X, intr = concat(probs)
    b = CostBatch(probs, robust=robust)
    r = b.eval(torch.tensor(T, device=dev), torch.tensor(X, device=dev), torch.tensor(intr, device=dev))
    c, gp, gx, gi = oracle_batch(probs, T, X, intr, robust)
    got = [r.cost.cpu().numpy(), r.g_pose.cpu().numpy(), r.g_X.cpu().numpy(), r.g_intr.cpu().numpy()]
    for name, g, ref in zip(("cost", "pose", "point", "intr"), got, (c, gp, gx, gi)):
        err = np.abs(g - ref).max() / max(np.abs(ref).max(), 1e-300)
        say(f"  {name}: max rel err {err:.2e}")
        assert err <= 1e-11, name
    b.close()


def test_fp32_close_to_fp64_and_cost_only_path(dev):
    from nav24_amd.diffcost import CostBatch
    say("cost fp32 vs fp64; forward-only path; determinism")
    probs = batch_problems(n=4, n_pt=600)
    T, X, intr = concat(probs)
    b = CostBatch(probs)
    r64 = b.eval(*(torch.tensor(a, device=dev) for a in (T, X, intr)))
    r32 = b.eval(*(torch.tensor(a, device=dev, dtype=torch.float32) for a in (T, X, intr)))
    # against FP64 at the same, float-rounded inputs: the FP32 path forms the
    # projection and residual in FP64 and only the adjoint in FP32, so what is
    # left is FP32 rounding of the gradient terms (about 1e-7 relative)
    r64f = b.eval(*(torch.tensor(a, device=dev, dtype=torch.float32).double() for a in (T, X, intr)))
    for name, a, ref in (("cost", r32.cost, r64f.cost), ("pose", r32.g_pose, r64f.g_pose),
                         ("point", r32.g_X, r64f.g_X), ("intr", r32.g_intr, r64f.g_intr)):
        err = (a.double() - ref).abs().max().item() / ref.abs().max().item()
        say(f"  fp32 {name}: max rel err {err:.2e} (FP64 at the same inputs)")
        assert err <= 1e-5, name
    # against FP64 at the original inputs the rounding of the inputs to float
    # adds its own part (poses and points of ~10 m to 6e-8 relative)
    for name, a, ref in (("cost", r32.cost, r64.cost), ("pose", r32.g_pose, r64.g_pose),
                         ("point", r32.g_X, r64.g_X), ("intr", r32.g_intr, r64.g_intr)):
        err = (a.double() - ref).abs().max().item() / ref.abs().max().item()
        say(f"  fp32 {name}: max rel err {err:.2e} (FP64 at the original inputs)")
        assert err <= 2e-3, name
    fwd = b.eval(*(torch.tensor(a, device=dev) for a in (T, X, intr)), grad=False)
    assert fwd.g_pose is None and torch.equal(fwd.cost, r64.cost)
    again = b.eval(*(torch.tensor(a, device=dev) for a in (T, X, intr)))
    assert torch.equal(again.g_pose, r64.g_pose) and torch.equal(again.g_X, r64.g_X)
    b.close()


def test_points_without_observations_get_zero_gradient(dev):
    from nav24_amd.diffcost import CostBatch
    probs = batch_problems(n=1, n_pt=100)
    p = probs[0]
    keep = p.obs_pt != 0  # drop every observation of point 0
    from nav24_amd.abi import Problem
    q = Problem(T_cw=p.T_cw, X=p.X, obs_cam=p.obs_cam[keep], obs_pt=p.obs_pt[keep], obs_uv=p.obs_uv[keep])
    T, X, intr = concat([p])
    b = CostBatch([q])
    r = b.eval(*(torch.tensor(a, device=dev) for a in (T, X, intr)))
    assert torch.count_nonzero(r.g_X[0]) == 0 and torch.count_nonzero(r.g_X[1:]) > 0
    b.close()


@pytest.mark.parametrize("xi_scale", [0.0, 0.3])
def test_autograd_gradcheck_fp64(dev, xi_scale):
    """gradcheck at xi = 0 and at random xi of norm about 0.3-0.5 per pose,
    where d cost / d xi is J_l(xi)^T times the kernel's left-tangent gradient."""
    from nav24_amd.diffcost import BACost, CostBatch
    say(f"autograd gradcheck fp64 (eps 1e-6), xi scale {xi_scale}")
    probs = batch_problems(n=2, n_pt=12, n_cam=3)
    T, X, intr = concat(probs)
    b = CostBatch(probs)
    Tt = torch.tensor(T, device=dev)
    rng = np.random.default_rng(11)
    xi0 = xi_scale * rng.standard_normal((T.shape[0], 6)) / np.sqrt(2.0)
    xi0[:, 3:] *= 0.2  # translations of a few centimetres keep every point in front
    xi = torch.tensor(xi0, dtype=torch.float64, device=dev, requires_grad=True)
    Xt = torch.tensor(X, device=dev, requires_grad=True)
    it = torch.tensor(intr, device=dev, requires_grad=True)
    f = lambda a, x, k: BACost.apply(a, x, k, Tt, b)
    assert torch.autograd.gradcheck(f, (xi, Xt, it), eps=1e-6, atol=1e-4, rtol=1e-5)
    b.close()
