"""alamo_mg_fsmooth_noalloc (no reference counterpart: the entry point for callers that cannot
allocate, e.g. an external multigrid driver captured in a HIP graph) against alamo_mg_fsmooth,
both called through the C ABI: bit-identical results (the same sweep kernels, ping-pong through the
library's static device scratch instead of a per-call allocation), on two calls in a row, the
second with a smaller level, including an odd sweep count (result copied back from the scratch)
and the l1 smoother; a level beyond the scratch's size and an unprepared handle are rejected."""
import ctypes

import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd._lib import lib
from alamo_amd.operator import Elastic, LAYOUT_AOS
from helpers import random_iso, random_bc, to_dev, to_host, rng

pytestmark = pytest.mark.gpu

MAX_DOUBLES = 8388608  # ALAMO_NOALLOC_MAX_DOUBLES


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def both(op, mglev, x0, b, cuda):
    """(alamo_mg_fsmooth result, alamo_mg_fsmooth_noalloc result) from the same x0"""
    xa, xb, bd = to_dev(x0, cuda), to_dev(x0, cuda), to_dev(b, cuda)
    assert lib().alamo_mg_fsmooth(op.h, 0, mglev, ptr(xa), ptr(bd), None) == 0
    rc = lib().alamo_mg_fsmooth_noalloc(op.h, 0, mglev, ptr(xb), ptr(bd), None)
    assert rc == 0, lib().alamo_last_error()
    return to_host(xa), to_host(xb)


@pytest.mark.parametrize("smoother,sweeps", [("jacobi", 2), ("l1", 3)])
def test_noalloc_matches_fsmooth(cuda, smoother, sweeps):
    ncell = (64, 48, 40)
    N = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    bct, bcv = random_bc(81, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    op = Elastic(ncell, dx, sym=P.ISO, bc=(bct, bcv), smoother=smoother, sweeps_per_smooth=sweeps)
    op.SetModel(random_iso(N, 82), LAYOUT_AOS)
    op.prepareForSolve()
    # call 1: the finest level; call 2: the smaller level 1 (the scratch keeps call 1's values)
    for mglev in (0, 1):
        n = op.LevelNodes(mglev)
        print("noalloc", smoother, sweeps, "level", mglev, "nodes", n, flush=True)
        x0 = rng(83 + mglev).uniform(-1, 1, 3 * n)
        b = rng(85 + mglev).uniform(-1, 1, 3 * n)
        a, c = both(op, mglev, x0, b, cuda)
        assert np.array_equal(a, c)
        assert not np.array_equal(a, x0)


def test_noalloc_rejects_large_level_and_unprepared(cuda):
    ncell = (144, 144, 144)  # 145^3 nodes: 3 x 145^3 doubles exceed the scratch
    N = 145 ** 3
    assert 3 * N > MAX_DOUBLES
    op = Elastic(ncell, (1 / 144,) * 3, sym=P.ISO)
    op.SetModel(P.iso_field(np.ones(N), np.ones(N)))
    x = to_dev(np.zeros(3 * N), cuda)
    rc = lib().alamo_mg_fsmooth_noalloc(op.h, 0, 0, ptr(x), ptr(x), None)
    assert rc == 3 and b"prepare" in lib().alamo_last_error()   # not prepared yet: ALAMO_ERR_STATE
    op.prepareForSolve()
    rc = lib().alamo_mg_fsmooth_noalloc(op.h, 0, 0, ptr(x), ptr(x), None)
    assert rc == 1 and b"scratch" in lib().alamo_last_error()   # too large: ALAMO_ERR_ARG
    # level 1 (73^3 nodes) fits
    n1 = op.LevelNodes(1)
    y = to_dev(np.ones(3 * n1), cuda)
    assert lib().alamo_mg_fsmooth_noalloc(op.h, 0, 1, ptr(y), ptr(y), None) == 0
