"""The transposed level-0 operator (alamo_mg_apply_transpose), which back-propagates the residual
training loss through A.

A^T is pinned against the transpose of the oracle's operator, assembled column by column on small
ragged grids (every boundary type, all three symmetries; FP64, 1e-12 of the scale), and at larger
sizes by the adjoint identity <A^T r, x> = <r, A x> with A x from the GPU's own Fapply.
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd.operator import Elastic, LAYOUT_AOS, apply_transpose
from alamo_amd._lib import AlamoError
from helpers import random_iso, random_mm, random_diag, random_bc, to_dev, to_host, rng

pytestmark = pytest.mark.gpu


def _coeffs(sym, nodes, seed):
    if sym == P.ISO:
        return random_iso(nodes, seed)
    if sym == P.MAJORMINOR:
        return random_mm(nodes, seed)
    return random_diag(nodes, seed)


def _build(sym, ncell, seed):
    nodes = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    C = _coeffs(sym, nodes, seed)
    bc = random_bc(seed)
    op = Elastic(ncell, dx, sym=sym, bc=bc)
    op.SetModel(C, LAYOUT_AOS)
    return op, C, bc, nodes, dx


@pytest.mark.parametrize("sym,ncell,seed", [(P.ISO, (4, 5, 6), 21), (P.ISO, (4, 4, 4), 22),
                                            (P.MAJORMINOR, (5, 4, 4), 23), (P.DIAGONAL, (4, 6, 5), 24)])
def test_transpose_matches_dense_oracle(cuda, ora, sym, ncell, seed):
    op, C, (t, v), nodes, dx = _build(sym, ncell, seed)
    print("transpose vs dense oracle", sym, ncell, flush=True)
    n3 = 3 * nodes
    A = np.empty((n3, n3))
    e = np.zeros(n3)
    for j in range(n3):
        e[j] = 1.0
        A[:, j] = ora.fapply(ncell, dx, sym, C, t, e)
        e[j] = 0.0
    for s in range(2):
        r = rng(100 + s).uniform(-1, 1, n3)
        out = to_dev(np.zeros(n3), cuda)
        apply_transpose(op, to_dev(r, cuda), out)
        ref = A.T @ r
        err = np.abs(to_host(out) - ref).max()
        assert err <= 1e-12 * np.abs(ref).max(), "max err %.3e" % err


@pytest.mark.parametrize("sym,ncell", [(P.ISO, (64, 48, 40)), (P.MAJORMINOR, (24, 20, 16)),
                                       (P.DIAGONAL, (16, 18, 20))])
def test_transpose_adjoint_identity(cuda, sym, ncell):
    op, C, bc, nodes, dx = _build(sym, ncell, 31)
    print("adjoint identity", sym, ncell, flush=True)
    r = rng(32).uniform(-1, 1, 3 * nodes)
    x = rng(33).uniform(-1, 1, 3 * nodes)
    Ax = to_dev(np.zeros(3 * nodes), cuda)
    op.Fapply(0, 0, Ax, to_dev(x, cuda))
    ATr = to_dev(np.zeros(3 * nodes), cuda)
    apply_transpose(op, to_dev(r, cuda), ATr)
    ax, atr = to_host(Ax), to_host(ATr)
    lhs, rhs = float(atr @ x), float(r @ ax)
    scale = float(np.abs(atr * x).sum())
    print("<A^T r, x> %.17e  <r, A x> %.17e" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_transpose_rejects_psi_and_aliasing(cuda):
    op, C, bc, nodes, dx = _build(P.ISO, (8, 8, 8), 41)
    r = to_dev(rng(42).uniform(-1, 1, 3 * nodes), cuda)
    with pytest.raises(AlamoError):
        apply_transpose(op, r, r)
    op.SetPsi(np.full(8 * 8 * 8, 0.5))
    with pytest.raises(AlamoError):
        apply_transpose(op, r, to_dev(np.zeros(3 * nodes), cuda))
