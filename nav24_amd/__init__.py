"""nav24_amd — MI355X-native bundle-adjustment back-end for nav24.

Drop-in for nav24's BE::GraphOptim::solve (core/backEnd/BE_GraphOptim.cpp:31-77):
C-ABI library libnav24_ba.so (include/nav_ba.h) with hand-written gfx950 HIP
kernels, plus Python bindings (ba.py, problems.py, autograd.py).
"""
from .abi import Options, Problem, Result, workload_options  # noqa: F401

__all__ = ["Options", "Problem", "Result", "workload_options"]
