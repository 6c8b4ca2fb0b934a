"""GPU key-point undistortion (nav_ba_kp_undistort) against the oracle: equal
float results (the device evaluates the same double operations without fused
multiply-adds), through the host entry and on device tensors, in place."""
import numpy as np
import pytest

from oracle import oracle
from parity_util import say as _say

pytestmark = pytest.mark.gpu

CAM = [450.0, 450.0, 320.0, 240.0]
PARAMS = {
    "none": CAM + [0.0, 0.0, 0.0, 0.0],
    "mild": CAM + [-0.1, 0.01, 1e-4, -2e-4],
    "strong": CAM + [-0.28, 0.07, 5e-4, 3e-4],
    "barrel_fails": CAM + [-5.0, 0.0, 0.0, 0.0],  # radial factor negative at the corners
    "pincushion": [458.6, 457.3, 367.2, 248.4, 0.05, -0.01, -8e-4, 6e-4],
}


def _pts(n, seed):
    r = np.random.default_rng(seed)
    return np.stack([r.uniform(-80, 720, n), r.uniform(-80, 560, n)], 1).astype(np.float32)


@pytest.mark.parametrize("name", sorted(PARAMS))
def test_host_entry_equals_oracle(name):
    from nav24_amd.undistort import PinholeRadTanHip
    _say(f"kp undistort host entry, params {name}, 200000 points")
    uv = _pts(200000, 7)
    got = PinholeRadTanHip(PARAMS[name]).UndistortKeyPoints(uv)
    ref = oracle.kp_undistort(uv, PARAMS[name])
    assert np.array_equal(got, ref)


def test_device_entry_in_place_and_sizes():
    import torch
    from nav24_amd.undistort import PinholeRadTanHip
    _say("kp undistort device tensors: in place, n = 1, 255, 256, 257, 3 000 000")
    cam = PinholeRadTanHip(PARAMS["strong"])
    for n in (1, 255, 256, 257, 3_000_000):
        uv = _pts(n, n)
        t = torch.from_numpy(uv).cuda()
        out = cam.undistort_device(t)
        cam.undistort_device(t, out=t)  # in place
        torch.cuda.synchronize()
        ref = oracle.kp_undistort(uv, PARAMS["strong"])
        assert np.array_equal(out.cpu().numpy(), ref), n
        assert np.array_equal(t.cpu().numpy(), ref), n


def test_empty_input_and_argument_errors():
    import ctypes
    from nav24_amd.abi import NAV_BA_BAD_ARG, NAV_BA_OK
    from nav24_amd.undistort import PinholeRadTanHip, lib
    _say("kp undistort: n = 0 and argument checks")
    out = PinholeRadTanHip(PARAMS["mild"]).UndistortKeyPoints(np.zeros((0, 2), np.float32))
    assert out.shape == (0, 2)
    prm = (ctypes.c_float * 8)(*PARAMS["mild"])
    assert lib().nav_ba_kp_undistort(None, 0, prm, None, None) == NAV_BA_OK
    assert lib().nav_ba_kp_undistort(None, -3, prm, None, None) == NAV_BA_BAD_ARG


KB_PARAMS = {
    "zero": [380.0, 380.0, 320.0, 240.0, 0.0, 0.0, 0.0, 0.0],
    "typical": [380.0, 380.0, 320.0, 240.0, 0.05, -0.01, 0.002, -0.0003],
    "strong": [286.2, 286.4, 321.4, 244.9, -0.0124, 0.0574, -0.0573, 0.0120],
}


def _ulps(a, b):
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)


@pytest.mark.parametrize("name", sorted(KB_PARAMS))
def test_fisheye_host_entry_matches_oracle(name):
    from nav24_amd.undistort import KannalaBrandt8Hip
    uv = _pts(200000, 11)
    got = KannalaBrandt8Hip(KB_PARAMS[name]).UndistortKeyPoints(uv)
    ref = oracle.kp_undistort_fisheye(uv, KB_PARAMS[name])
    neq = int((got != ref).sum())
    _say(f"kp undistort fisheye host entry, params {name}, 200000 points: {neq} values differ from the oracle")
    assert np.all(_ulps(got, ref) <= 1.0)  # tan() of the device and host libraries may differ in the last bit


def test_fisheye_device_entry_in_place():
    import torch
    from nav24_amd.undistort import KannalaBrandt8Hip
    cam = KannalaBrandt8Hip(KB_PARAMS["typical"])
    for n in (1, 257, 1_000_000):
        uv = _pts(n, 100 + n)
        t = torch.from_numpy(uv).cuda()
        cam.undistort_device(t, out=t)
        torch.cuda.synchronize()
        ref = oracle.kp_undistort_fisheye(uv, KB_PARAMS["typical"])
        got = t.cpu().numpy()
        _say(f"kp undistort fisheye device in place n={n}: {int((got != ref).sum())} values differ")
        assert np.all(_ulps(got, ref) <= 1.0), n


@pytest.mark.parametrize("model", [0, 1])
def test_rectified_camera_equals_oracle(model):
    """A calibration with a rectification R and a 3x4 projection P
    (Calibration.cpp:88-93): host entry and device tensors equal the oracle
    bit for bit, and differ from the unrectified result."""
    import torch
    from nav24_amd.undistort import KannalaBrandt8Hip, PinholeRadTanHip
    from test_kp_undistort_oracle import RECT_P, RECT_R
    prm = PARAMS["strong"] if model == 0 else [380.0, 380.0, 320.0, 240.0, 0.05, -0.01, 0.002, -0.0003]
    _say(f"kp undistort with R and P, model {model}, 200000 points")
    cam = (PinholeRadTanHip if model == 0 else KannalaBrandt8Hip)(prm)
    uv = _pts(200000, 11)
    plain = cam.UndistortKeyPoints(uv)
    cam.setRectificationMat(RECT_R)
    cam.setProjectionMat(RECT_P)
    got = cam.UndistortKeyPoints(uv)
    ref = oracle.kp_undistort_rect(model, uv, prm, RECT_R, RECT_P)
    assert np.array_equal(got, ref)
    assert np.abs(got - plain).max() > 1.0
    t = torch.from_numpy(uv).cuda()
    cam.undistort_device(t, out=t)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), ref)
