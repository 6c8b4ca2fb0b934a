"""Locate and load the package's in-tree shared libraries (nav24_amd/_lib/)."""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")

_cache: dict = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load_library(name: str) -> ctypes.CDLL:
    """Load nav24_amd/_lib/<name>; raise if it has not been built (run
    `python -c "import __graft_entry__ as g; g.build()"` or nav24_amd.build)."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run nav24_amd/build.py first")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
