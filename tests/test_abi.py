"""The C-ABI library loads and exports every declared entry point (CPU; no
compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from nav24_amd import ba
from nav24_amd._loader import lib_path
from nav24_amd.abi import NAV_BA_NO_DEVICE, NavBaOptions, NavBaProblem, NavBaResult

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(headers=None):
    names = []
    if headers is None:
        headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    for h in headers:
        src = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:const char\*|int32_t|void)\s+(nav_\w+)\s*\(", src, re.M)
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(lib_path("libnav24_ba.so"))
    names = declared()
    assert sorted(declared(["nav_ba.h", "nav_ba_debug.h"])) == sorted(ba.EXPORTED)
    assert "nav_ba_cost_eval" in names and "nav_tv_score_fundamental_host" in names
    for n in names:
        assert hasattr(lib, n), n


def test_cost_handle_reports_no_device_here():
    from nav24_amd import diffcost
    h = ctypes.c_void_p()
    st = diffcost.lib().nav_ba_cost_create(ctypes.byref(h), 0)
    if st == NAV_BA_NO_DEVICE:
        assert not h.value
    else:  # with a gfx950 device the handle is real
        diffcost.lib().nav_ba_cost_destroy(h)


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of every field as gcc lays out include/nav_ba.h."""
    structs = {"nav_ba_problem": NavBaProblem, "nav_ba_options": NavBaOptions, "nav_ba_result": NavBaResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nav_ba.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for cname, cls in structs.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_version_and_default_options():
    assert "gfx950" in ba.version()
    o = ba.default_options()
    assert (o.n_iter, o.robust, o.precision, o.max_trials) == (5, False, 0, 10)
    assert o.huber_delta == float(__import__("numpy").float32(5.99 ** 0.5))
    assert o.compat_fp32_io and o.outlier_chi2 == 5.991


def test_create_without_device_reports_no_device():
    h = ctypes.c_void_p()
    st = ba.lib().nav_ba_create(ctypes.byref(h), 0)
    if st == 0:
        ba.lib().nav_ba_destroy(h)
        pytest.skip("a gfx950 device is present")
    assert st == NAV_BA_NO_DEVICE
