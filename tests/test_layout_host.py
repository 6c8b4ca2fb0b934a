"""Host-side structure of a prepare, checked on the CPU without a device
(nav_ba_debug_check_layout): every index of the flat graph arrays and of the
Schur work items against the buffer it addresses, and every Schur round's
pairs against a direct enumeration of the row's pairs.  Covers wide rows
(column groups), narrow rows, batches of windows, fixed points and cameras,
duplicate observations, every chunk size and both Schur-row kernels (rounds
with lanes bound per round; producer/consumer with lanes bound per item)."""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_BAD_ARG, Problem, workload_options
from nav24_amd.ba import BAError, check_layout


def with_duplicates(p, frac=0.05, seed=0):
    """The problem with a share of its observations repeated (same camera and
    point, measurement moved by a pixel), at random places in the input order."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(p.n_obs, int(frac * p.n_obs), replace=False)
    cam = np.concatenate([p.obs_cam, p.obs_cam[idx]])
    pt = np.concatenate([p.obs_pt, p.obs_pt[idx]])
    uv = np.concatenate([p.obs_uv, p.obs_uv[idx] + 1.0])
    perm = rng.permutation(len(cam))
    return Problem(T_cw=p.T_cw, X=p.X, obs_cam=cam[perm], obs_pt=pt[perm], obs_uv=uv[perm],
                   cam_fixed=p.cam_fixed, pt_fixed=p.pt_fixed)


KERNELS = pytest.mark.parametrize("kernel", [1, 2])


@KERNELS
@pytest.mark.parametrize("chunk", [0, 64, 512, 2048])
def test_wide_rows(chunk, kernel):
    p = problems.workload(1, seed=2, n_pt=4000)
    r = check_layout([p], workload_options(1), chunk=chunk, kernel=kernel)
    assert r["pairs"] > 0


@KERNELS
def test_column_groups(kernel):
    p = problems.workload(1, seed=6, n_cam=160, n_pt=3000)
    check_layout([p], workload_options(1), chunk=256, kernel=kernel)


@KERNELS
def test_narrow_rows_with_fixed_points_and_duplicates(kernel):
    p = problems.workload(0, seed=1, n_pt=600)
    p.pt_fixed = (np.arange(p.n_pt) % 7 == 0).astype(np.uint8)
    q = with_duplicates(p)
    r = check_layout([q], workload_options(0), kernel=kernel)
    assert r["pairs"] > 0


@KERNELS
@pytest.mark.parametrize("chunk", [64, 256, 2048])
def test_batch_of_windows(chunk, kernel):
    wins = [problems.workload(3, seed=s) for s in range(12)]
    wins[3] = with_duplicates(wins[3], seed=3)
    r = check_layout(wins, workload_options(3), batched=True, chunk=chunk, kernel=kernel)
    assert r["rounds"] > 0


@KERNELS
def test_loop_rows(kernel):
    p = problems.workload(2, seed=0, n_pt=8000)
    check_layout([p], workload_options(2), chunk=1024, kernel=kernel)


def test_kernels_enumerate_the_same_pairs():
    wins = [problems.workload(3, seed=s) for s in range(4)]
    a = check_layout(wins, workload_options(3), batched=True, chunk=512, kernel=1)
    b = check_layout(wins, workload_options(3), batched=True, chunk=512, kernel=2)
    assert a["pairs"] == b["pairs"] and b["rounds"] >= a["rounds"]


def test_bad_chunk_rejected():
    p = problems.workload(0, seed=0, n_pt=100)
    with pytest.raises(BAError) as e:
        check_layout([p], workload_options(0), chunk=100)
    assert e.value.status == NAV_BA_BAD_ARG
    with pytest.raises(BAError) as e:
        check_layout([p], workload_options(0), kernel=3)
    assert e.value.status == NAV_BA_BAD_ARG


def test_camera_model_checked_on_the_host():
    a = problems.workload(3, seed=0)
    b = problems.with_kb8(problems.workload(3, seed=1))
    check_layout([b], workload_options(3), batched=True)
    with pytest.raises(BAError) as e:
        check_layout([a, b], workload_options(3), batched=True)
    assert e.value.status == NAV_BA_BAD_ARG
    c = problems.with_kb8(problems.workload(3, seed=2), coeffs=(0.0, 0.0, 0.0, 0.1))
    with pytest.raises(BAError):
        check_layout([b, c], workload_options(3), batched=True)
    bad = problems.workload(0, seed=0, n_pt=100)
    bad.camera_model = 7
    with pytest.raises(BAError) as e:
        check_layout([bad], workload_options(0))
    assert e.value.status == NAV_BA_BAD_ARG
