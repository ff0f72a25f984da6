"""GPU parity of the l1-Jacobi smoother (smoother="l1"; no reference counterpart: the smoother of
the cycles is a preconditioner choice that leaves the operator, and so the converged solution,
unchanged).  One Fsmooth call (2 sweeps x += (b - A x) / d, d the row's absolute sum with the sign
of its diagonal) against the oracle's sweeps with the oracle's brute-force-pinned d, on:

* isotropic levels served by the row kernel with random BC tables (face rows by k_boundary) and
  with all-Dirichlet boundaries (face rows inside the row kernel's launch), and on the coarse
  level 1 (the restricted coefficients);
* MajorMinor rows (k_mm_interior and k_boundary) and small levels (node-parallel rows).
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd.operator import Elastic, LAYOUT_AOS
from helpers import random_iso, random_mm, random_diag, random_bc, to_dev, to_host, rng

pytestmark = pytest.mark.gpu


def close(gpu, ref, rtol):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(gpu - ref).max()
    assert err <= rtol * scale, "max err %.3e vs scale %.3e" % (err, scale)


CASES = [
    (P.ISO, (128, 100, 80), "random"),
    (P.ISO, (130, 96, 84), "dirichlet"),
    (P.ISO, (6, 5, 4), "random"),
    (P.MAJORMINOR, (12, 10, 8), "random"),
    (P.DIAGONAL, (6, 6, 4), "random"),
]


@pytest.mark.parametrize("sym,ncell,bcs", CASES)
def test_l1_fsmooth(cuda, ora, sym, ncell, bcs):
    print("l1 fsmooth", sym, ncell, bcs, flush=True)
    nodes = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    C = {P.ISO: random_iso, P.MAJORMINOR: random_mm, P.DIAGONAL: random_diag}[sym](nodes, 51)
    if sym == P.ISO:
        # a stiffness jump of 1:50 across a plane, so some rows are not diagonally dominant
        lam = C[0::2].reshape(ncell[2] + 1, ncell[1] + 1, ncell[0] + 1)
        lam[:, :, ncell[0] // 2:] *= 50.0
        C[0::2] = lam.ravel()
        C[1::2] = C[1::2] * np.where(np.arange(nodes) % (ncell[0] + 1) >= ncell[0] // 2, 50.0, 1.0)
    bct, bcv = random_bc(52, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN)) if bcs == "random" else P.bc_constant()
    op = Elastic(ncell, dx, sym=sym, bc=(bct, bcv), smoother="l1")
    op.SetModel(C, LAYOUT_AOS)
    x0 = rng(53).uniform(-1, 1, 3 * nodes)
    b = rng(54).uniform(-1, 1, 3 * nodes)
    d = ora.l1_diag(ncell, dx, sym, C, bct)
    x = to_dev(x0, cuda)
    op.Fsmooth(0, 0, x, to_dev(b, cuda))
    close(to_host(x), ora.jacobi_l1(ncell, dx, sym, C, bct, d, 2, x0, b), 1e-11)
    if sym == P.ISO and op.NMGLevels() > 1:
        # level 1: the restricted coefficients of the hierarchy
        nc1 = op.LevelNCell(1)
        n1 = op.LevelNodes(1)
        dx1 = tuple(2 * h for h in dx)
        C1 = ora.coeff_level(ncell, dx, sym, C, 1)
        d1 = ora.l1_diag(nc1, dx1, sym, C1, bct)
        y0 = rng(55).uniform(-1, 1, 3 * n1)
        c = rng(56).uniform(-1, 1, 3 * n1)
        y = to_dev(y0, cuda)
        op.Fsmooth(0, 1, y, to_dev(c, cuda))
        close(to_host(y), ora.jacobi_l1(nc1, dx1, sym, C1, bct, d1, 2, y0, c), 1e-11)


def test_l1_switch_restores_reference_sweeps(cuda, ora):
    """UpdateOpts(smoother="jacobi") after l1 gives the reference's damped point Jacobi bit for bit"""
    ncell = (64, 40, 36)
    nodes = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    C = random_iso(nodes, 57)
    bct, bcv = P.bc_constant()
    x0 = rng(58).uniform(-1, 1, 3 * nodes)
    b = to_dev(rng(59).uniform(-1, 1, 3 * nodes), cuda)
    ref = Elastic(ncell, dx, sym=P.ISO, bc=(bct, bcv))
    ref.SetModel(C)
    xr = to_dev(x0, cuda)
    ref.Fsmooth(0, 0, xr, b)
    op = Elastic(ncell, dx, sym=P.ISO, bc=(bct, bcv), smoother="l1")
    op.SetModel(C)
    xl = to_dev(x0, cuda)
    op.Fsmooth(0, 0, xl, b)
    op.UpdateOpts(smoother="jacobi")
    xj = to_dev(x0, cuda)
    op.Fsmooth(0, 0, xj, b)
    assert np.array_equal(to_host(xj), to_host(xr))
    assert not np.array_equal(to_host(xl), to_host(xr))
