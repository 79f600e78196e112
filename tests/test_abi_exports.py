"""The C-ABI library loads (no GPU needed) and exports every function that
include/*.h declares; the ctypes structs match the C struct sizes."""
import ctypes as C
import os
import re

import grace_amd as ga

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


IO_HEADER = "grmhd_io.h"   # exported by libgrmhd_io.so, the HDF5 state-dump library


def declared_functions(io=False):
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h") or (h == IO_HEADER) != io:
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(grmhd_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    L = ga.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_symbol_list_matches_headers():
    assert set(ga.ABI_SYMBOLS) == declared_functions()


def test_io_library_exports_declared_symbols():
    L = ga.io_lib()
    names = declared_functions(io=True)
    assert len(names) == 4
    assert not [n for n in sorted(names) if not hasattr(L, n)]


def test_struct_layouts():
    # grmhd_params ends with int32 flat_metric after a double; total size is a multiple of 8
    assert C.sizeof(ga.Params) % 8 == 0
    assert C.sizeof(ga.Grid) == 24
    assert C.sizeof(ga.Instance) == 72  # 3 int32 + pad, int64, 6 int32, 3 double
    p = ga.default_params()
    assert p.rho_fl == 1e-14 and p.c2p_tol == 1e-10 and p.max_sigma == 100 and p.fd_order == 6


def test_setup_functions_run_on_host():
    from grace_amd import workloads as W
    st = W.c2p_states(64)
    assert st["cons"].shape == (10, 64)
    wl = W.blast(nb=(1, 1, 1), n=8)
    assert wl.state.shape == (1, 23, 14, 14, 14)
