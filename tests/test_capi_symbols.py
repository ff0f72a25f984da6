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
