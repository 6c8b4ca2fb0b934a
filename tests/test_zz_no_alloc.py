"""The allocation-free undistortion entry (nav_ba_kp_undistort_rect_host_noalloc)
against the entry it varies (nav_ba_kp_undistort_rect_host), through the C
ABI: equal float results on two calls in a row, the second smaller, for both
camera models with and without a rectification; sizes beyond the limit are
refused."""
import ctypes as C

import numpy as np
import pytest

from parity_util import say

pytestmark = pytest.mark.gpu

MAXN = 1 << 21


def _lib():
    from nav24_amd.abi import NAV_BA_NO_DEVICE
    from nav24_amd.undistort import lib
    L = lib()
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.nav_ba_kp_undistort_rect_host_noalloc.argtypes = [C.c_int32, fp, C.c_int32, fp, dp, dp, C.c_int32, fp,
                                                        C.c_int32]
    L.nav_ba_kp_undistort_rect_host_noalloc.restype = C.c_int32
    probe = np.zeros((1, 2), np.float32)
    prm = np.float32([450, 450, 320, 240, 0, 0, 0, 0])
    st = L.nav_ba_kp_undistort_rect_host_noalloc(0, probe.ctypes.data_as(fp), 1, prm.ctypes.data_as(fp), None, None,
                                                 3, probe.ctypes.data_as(fp), 0)
    if st == NAV_BA_NO_DEVICE:
        pytest.skip("no gfx950 device")
    return L


def _call(L, name, model, uv, prm, R, P):
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    out = np.empty_like(uv)
    Ra = None if R is None else np.ascontiguousarray(R, np.float64)
    Pa = None if P is None else np.ascontiguousarray(P, np.float64)
    st = getattr(L, name)(model, uv.ctypes.data_as(fp), uv.shape[0], prm.ctypes.data_as(fp),
                          None if Ra is None else Ra.ctypes.data_as(dp), None if Pa is None else Pa.ctypes.data_as(dp),
                          3 if Pa is None else Pa.shape[1], out.ctypes.data_as(fp), 0)
    return st, out


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("rect", [False, True])
def test_no_alloc_equals_the_allocating_entry(model, rect):
    from nav24_amd.abi import NAV_BA_OK
    from test_kp_undistort_oracle import RECT_P, RECT_R
    L = _lib()
    prm = np.float32([450, 450, 320, 240, -0.28, 0.07, 5e-4, 3e-4] if model == 0 else
                     [380, 380, 320, 240, 0.05, -0.01, 0.002, -0.0003])
    R, P = (RECT_R, RECT_P) if rect else (None, None)
    rng = np.random.default_rng(5 + model)
    for n in (MAXN, 40_000):  # two calls in a row, the second smaller
        say(f"no-alloc undistortion: model {model}, rectified {rect}, {n} points")
        uv = np.stack([rng.uniform(-80, 720, n), rng.uniform(-80, 560, n)], 1).astype(np.float32)
        st_a, a = _call(L, "nav_ba_kp_undistort_rect_host_noalloc", model, uv, prm, R, P)
        st_b, b = _call(L, "nav_ba_kp_undistort_rect_host", model, uv, prm, R, P)
        assert st_a == NAV_BA_OK and st_b == NAV_BA_OK
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_no_alloc_refuses_more_than_its_array():
    from nav24_amd.abi import NAV_BA_TOO_LARGE
    L = _lib()
    fp = C.POINTER(C.c_float)
    prm = np.float32([450, 450, 320, 240, 0, 0, 0, 0])
    uv = np.zeros((1, 2), np.float32)
    # the size is checked before any array is read
    st = L.nav_ba_kp_undistort_rect_host_noalloc(0, uv.ctypes.data_as(fp), MAXN + 1, prm.ctypes.data_as(fp), None,
                                                 None, 3, uv.ctypes.data_as(fp), 0)
    assert st == NAV_BA_TOO_LARGE
