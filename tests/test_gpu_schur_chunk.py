"""Schur work items: forced sizes, both row kernels, column groups and the
prepare-time tuner against the oracle.

The chunk size and the kernel (rounds with lanes bound per round, or
producer/consumer waves with lanes bound per item) only change how the
observations of one camera row are split over workgroups and lanes (and so
the order of FP64 additions), never what is summed: the reduced system must
match the oracle's for every choice, and the tuned solve must match the
oracle like the default one.
"""
import numpy as np
import pytest

from nav24_amd import problems
from nav24_amd.abi import NAV_BA_FP64, workload_options
from parity_util import check_fp64_against_oracle, say as _say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("chunk", [64, 128, 512, 2048])
def test_reduced_system_any_chunk(solver, oracle_mod, chunk, kernel):
    _say(f"reduced system workload1 n_pt=3000 schur chunk {chunk} kernel {kernel}")
    prob = problems.workload(1, seed=4, n_pt=3000)
    opts = workload_options(1)
    solver.set_schur_chunk(chunk)
    solver.set_schur_kernel(kernel)
    try:
        solver.prepare(prob, opts)
        assert solver.schur_chunk == chunk and solver.schur_kernel == kernel
        S, g = solver.debug_reduced_system(1e-4)
    finally:
        solver.set_schur_chunk(0)
        solver.set_schur_kernel(0)
    So, go = oracle_mod.reduced_system(prob, opts, 1e-4)
    assert S.shape == So.shape
    assert np.abs(S - So).max() <= 1e-11 * np.abs(So).max()
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()


def max_covisible(prob):
    """Largest number of free cameras sharing a free point with one free camera."""
    free = np.ones(prob.n_cam, bool) if prob.cam_fixed is None else prob.cam_fixed == 0
    seen = np.zeros((prob.n_cam, prob.n_pt), bool)
    seen[prob.obs_cam, prob.obs_pt] = True
    seen = seen[free]
    co = (seen.astype(np.int32) @ seen.T.astype(np.int32)) > 0
    return int(co.sum(1).max())


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("chunk", [256, 1024])
def test_reduced_system_column_groups(solver, oracle_mod, chunk, kernel):
    """Rows with more co-visible cameras than one work item's 112 slots are
    split into column groups (each item then recomputes its W_a)."""
    prob = problems.workload(1, seed=6, n_cam=160, n_pt=6000)
    mc = max_covisible(prob)
    _say(f"reduced system workload1 160 cameras n_pt=6000 (up to {mc} co-visible) schur chunk {chunk} kernel {kernel}")
    assert mc > 112
    opts = workload_options(1)
    solver.set_schur_chunk(chunk)
    solver.set_schur_kernel(kernel)
    try:
        solver.prepare(prob, opts)
        S, g = solver.debug_reduced_system(1e-4)
    finally:
        solver.set_schur_chunk(0)
        solver.set_schur_kernel(0)
    So, go = oracle_mod.reduced_system(prob, opts, 1e-4)
    assert np.abs(S - So).max() <= 1e-11 * np.abs(So).max()
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()


def test_tuned_chunk_solve_matches_oracle(solver, oracle_mod):
    prob = problems.workload(1, seed=5)
    assert prob.n_obs >= 1 << 17, "the tuner runs from 131072 observations on"
    _say(f"tuned schur chunk workload1 seed=5 n_obs={prob.n_obs}")
    opts = workload_options(1)
    opts.precision = NAV_BA_FP64
    opts.n_iter = 6
    solver.set_schur_chunk(0)
    got = solver.solve(prob, opts)
    chunk, kernel = solver.schur_chunk, solver.schur_kernel
    _say(f"  tuned chunk {chunk} kernel {kernel}")
    assert chunk in (128, 256, 512, 1024, 2048, 4096) and kernel in (1, 2)
    ref = oracle_mod.solve(prob, opts)
    check_fp64_against_oracle(oracle_mod, prob, opts, got, ref)
    # a second prepare of the same shape reuses the tuned choice
    solver.prepare(prob, opts)
    assert solver.schur_chunk == chunk and solver.schur_kernel == kernel


def test_chunk_argument_checked(solver):
    from nav24_amd.ba import BAError
    for bad in (-64, 100, 8192):
        with pytest.raises(BAError):
            solver.set_schur_chunk(bad)
    for bad in (-1, 3):
        with pytest.raises(BAError):
            solver.set_schur_kernel(bad)
