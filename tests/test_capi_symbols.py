"""The C-ABI library loads without a GPU and exports every function include/alamo_amd.h declares."""
import os
import re

from alamo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "alamo_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(alamo_[a-z0-9_]+)\s*\(", src)))


def test_all_declared_symbols_exported():
    names = declared()
    assert len(names) >= 20
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.SIGNATURES) >= set(names)


def test_defaults_and_errors_without_gpu():
    import ctypes
    d = _lib.MGDesc()
    assert _lib.lib().alamo_mg_desc_default(ctypes.byref(d)) == 0
    assert d.cycle == 1 and d.pre_smooth == 2 and d.sweeps_per_smooth == 2 and abs(d.omega - 2 / 3) < 1e-15
    assert d.max_iter == 200 and d.bottom_max_iter == 200 and d.bottom_tol_rel == 1e-4
    bad = _lib.MGDesc()
    _lib.lib().alamo_mg_desc_default(ctypes.byref(bad))
    bad.n_cell[0] = 1
    h = ctypes.c_void_p()
    rc = _lib.lib().alamo_mg_create(ctypes.byref(bad), ctypes.byref(h))
    assert rc == 1 and b"n_cell" in _lib.lib().alamo_last_error()


def test_bottom_solver_option_without_gpu():
    import ctypes
    import pytest
    from alamo_amd.operator import bottom_solver_code
    d = _lib.MGDesc()
    assert _lib.lib().alamo_mg_desc_default(ctypes.byref(d)) == 0
    assert d.bottom_solver == 0  # bicgstab, the choice the reference forces (Linear.H:258)
    assert [bottom_solver_code(n) for n in ("bicgstab", "smoother", "dense")] == [0, 1, 2]
    with pytest.raises(ValueError):
        bottom_solver_code("cg")
    for q in range(3):
        d.n_cell[q], d.dx[q] = 8, 0.125
    d.bottom_solver = 7
    h = ctypes.c_void_p()
    assert _lib.lib().alamo_mg_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"bottom_solver" in _lib.lib().alamo_last_error()


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of alamo_mg_desc and alamo_solve_stats have the C header's size and
    field offsets (checked with the host C compiler)."""
    import ctypes
    import subprocess
    fields = [f for f, _ in _lib.MGDesc._fields_]
    sfields = [f for f, _ in _lib.SolveStats._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "alamo_amd.h"', "int main(void) {",
           '  printf("%zu\\n", sizeof(alamo_mg_desc));', '  printf("%zu\\n", sizeof(alamo_solve_stats));']
    src += ['  printf("%%zu\\n", offsetof(alamo_mg_desc, %s));' % f for f in fields]
    src += ['  printf("%%zu\\n", offsetof(alamo_solve_stats, %s));' % f for f in sfields]
    src += ["  return 0;", "}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)])
    vals = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    assert vals[0] == ctypes.sizeof(_lib.MGDesc) and vals[1] == ctypes.sizeof(_lib.SolveStats)
    assert vals[2:2 + len(fields)] == [getattr(_lib.MGDesc, f).offset for f in fields]
    assert vals[2 + len(fields):] == [getattr(_lib.SolveStats, f).offset for f in sfields]
