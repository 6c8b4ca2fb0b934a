"""Shared fixtures.  GPU tests are marked `gpu` and call the HIP library
through its C ABI; without a gfx950 device they skip (here), with one they
must run the HIP path (there is no CPU fallback)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def solver():
    from nav24_amd.abi import NAV_BA_NO_DEVICE
    from nav24_amd.ba import BAError, GraphOptimHip
    try:
        s = GraphOptimHip(0)
    except BAError as e:
        if e.status == NAV_BA_NO_DEVICE:
            pytest.skip("no gfx950 device")
        raise
    yield s
    s.close()


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle
