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


def _c_layouts(tmp_path):
    """sizeof and field offsets of the header's structs, from the C compiler."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    structs = {"grmhd_params": ga.Params, "grmhd_grid": ga.Grid, "grmhd_instance": ga.Instance,
               "grmhd_forest": ga.Forest, "grmhd_c2p_stats_t": ga.C2PStats, "grmhd_tuning": ga.Tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "grmhd_abi.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f in cls._fields_:
            lines.append(f'printf("{name} {f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, field, val = line.split()
        got[(name, field)] = int(val)
    return structs, got


def test_struct_layouts(tmp_path):
    structs, got = _c_layouts(tmp_path)
    for name, cls in structs.items():
        assert C.sizeof(cls) == got[(name, "size")], name
        for f in cls._fields_:
            assert getattr(cls, f[0]).offset == got[(name, f[0])], (name, f[0])
    p = ga.default_params()
    assert p.rho_fl == 1e-14 and p.c2p_tol == 1e-10 and p.max_sigma == 100 and p.fd_order == 6
    assert p.excise_by_radius == 0 and p.excision_cks == 0


def test_setup_functions_run_on_host():
    from grace_amd import workloads as W
    st = W.c2p_states(64)
    assert st["cons"].shape == (10, 64)
    wl = W.blast(nb=(1, 1, 1), n=8)
    assert wl.state.shape == (1, 23, 14, 14, 14)
