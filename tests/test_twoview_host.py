"""Host-side parts of the two-view C ABI that need no GPU: Normalize
(OP_2ViewReconstruction.cpp:1195-1240, float sums in input order) equal to the
oracle's restatement bit for bit, and the argument checks the *_host entries
make before they look for a device."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle
from nav24_amd.abi import NAV_BA_BAD_ARG
from nav24_amd.twoview import lib

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)
u8 = C.POINTER(C.c_uint8)


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (1000, 2), (20000, 3)])
def test_normalize_equals_oracle(n, seed):
    rng = np.random.default_rng(seed)
    kp = rng.uniform([-5, 0], [760, 485], (n, 2)).astype(np.float32)
    out = np.zeros_like(kp)
    T = np.zeros(9, np.float32)
    assert lib().nav_tv_normalize(kp.ctypes.data_as(fp), n, out.ctypes.data_as(fp), T.ctypes.data_as(fp)) == 0
    ref, RT = oracle.tv_normalize(kp)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(T.view(np.uint32), RT.reshape(-1).view(np.uint32))


def test_host_entries_reject_bad_arguments_before_the_device():
    L = lib()
    kp = np.zeros((4, 2), np.float32)
    m = np.array([[0, 0], [1, 9]], np.int32)  # 9 is beyond the 4 key points of image 2
    sets = np.zeros((1, 8), np.int32)
    H = np.zeros((1, 9), np.float32)
    out = np.zeros(9, np.float32)
    best = np.zeros(1, np.int32)
    inl = np.zeros(2, np.uint8)
    # hypotheses: a match index out of range, then a set index out of range
    assert L.nav_tv_hypotheses_host(kp.ctypes.data_as(fp), 4, kp.ctypes.data_as(fp), 4, m.ctypes.data_as(ip), 2,
                                    sets.ctypes.data_as(ip), 1, out.ctypes.data_as(fp), out.ctypes.data_as(fp),
                                    out.ctypes.data_as(fp), 0) == NAV_BA_BAD_ARG
    m[1, 1] = 1
    sets[0, 5] = 2  # only matches 0 and 1 exist
    assert L.nav_tv_hypotheses_host(kp.ctypes.data_as(fp), 4, kp.ctypes.data_as(fp), 4, m.ctypes.data_as(ip), 2,
                                    sets.ctypes.data_as(ip), 1, out.ctypes.data_as(fp), out.ctypes.data_as(fp),
                                    out.ctypes.data_as(fp), 0) == NAV_BA_BAD_ARG
    # scoring: sigma must be positive
    assert L.nav_tv_score_fundamental_host(kp.ctypes.data_as(fp), 4, kp.ctypes.data_as(fp), 4, m.ctypes.data_as(ip),
                                           2, H.ctypes.data_as(fp), 1, 0.0, 3.841, 5.991, out.ctypes.data_as(fp),
                                           best.ctypes.data_as(ip), out.ctypes.data_as(fp), inl.ctypes.data_as(u8),
                                           0) == NAV_BA_BAD_ARG
    # CheckRT: 0 and 9 candidates are out of range
    K = np.eye(3, dtype=np.float32).reshape(-1)
    R = np.tile(np.eye(3, dtype=np.float32).reshape(-1), 9)
    t = np.zeros(27, np.float32)
    ng = np.zeros(9, np.int32)
    f9 = np.zeros(9, np.float32)
    p3 = np.zeros(9 * 4 * 3, np.float32)
    gd = np.zeros(9 * 4, np.uint8)
    for nc in (0, 9):
        assert L.nav_tv_check_rt_host(kp.ctypes.data_as(fp), 4, kp.ctypes.data_as(fp), 4, m.ctypes.data_as(ip),
                                      inl.ctypes.data_as(u8), 2, K.ctypes.data_as(fp), R.ctypes.data_as(fp),
                                      t.ctypes.data_as(fp), nc, 4.0, 0.99998, 50, ng.ctypes.data_as(ip),
                                      f9.ctypes.data_as(fp), f9.ctypes.data_as(fp), p3.ctypes.data_as(fp),
                                      gd.ctypes.data_as(u8), 0) == NAV_BA_BAD_ARG
    assert L.nav_tv_normalize(kp.ctypes.data_as(fp), 0, out.ctypes.data_as(fp), out.ctypes.data_as(fp)) == \
        NAV_BA_BAD_ARG
