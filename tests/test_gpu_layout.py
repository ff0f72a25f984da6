"""Pitched node layout (row_align): a handle whose large levels have their x rows padded to a
multiple of `row_align` nodes gives the compact handle's results through every entry point.

Operator rows, boundary values, the Newton right-hand side and the grid transfers run the same
kernels with the same arithmetic on both layouts, so they must agree bit for bit (after
dropping the pads); the pads themselves must stay zero.  Solves differ only in the order of
the flat reductions (the pads move the block boundaries): FP64 solves agree to 1e-10 relative
with iteration counts within one; mixed-precision Krylov solves, whose FP32 inner vectors round
differently once a coefficient differs in its last bits, agree to 1e-8 (the bar of every
mixed-vs-FP64 comparison) with counts within three.  The learned stage rounds its input to bf16
after a scale that differs in the last bits, so its output agrees to 1e-3 relative.
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd.operator import (Elastic, Linear, Krylov, Net, precond_apply, LAYOUT_AOS, PREC_MIXED, CYCLE_V,
                                PRECOND_LEARNED)
from helpers import random_iso, random_mm, random_bc, to_dev, to_host, rng

pytestmark = pytest.mark.gpu


def _pair(sym, ncell, seed, bc, align, **kw):
    nodes = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    C = random_iso(nodes, seed) if sym == P.ISO else random_mm(nodes, seed)
    ops = []
    for a in (0, align):
        op = Elastic(ncell, dx, sym=sym, bc=bc, row_align=a, **kw)
        op.SetModel(C, LAYOUT_AOS)
        ops.append(op)
    return ops, C, nodes, dx


def _pads_zero(op, x, mglev=0, ncomp=3):
    N = [v + 1 for v in op.LevelNCell(mglev)]
    sy = op.LevelLayout(mglev)[0]
    a = to_host(x).reshape(ncomp, N[2], N[1], sy)
    assert not np.any(a[..., N[0]:]), "pad entries were written"


def _all_dirichlet():
    t, v = P.bc_constant()
    return t, rng(5).uniform(-0.1, 0.1, (26, 3))


def test_layout_pitch(cuda):
    ops, *_ = _pair(P.ISO, (128, 100, 80), 1, random_bc(1), 32)
    cmp_, pad = ops
    assert cmp_.LevelLayout(0) == (129, 129 * 101, 129 * 101 * 81)
    assert pad.LevelLayout(0) == (160, 160 * 101, 160 * 101 * 81)
    # level 1 (65 nodes per row) is below 4 x 32 and stays compact
    assert pad.LevelLayout(1) == (65, 65 * 51, 65 * 51 * 41)
    x = np.arange(3 * 129 * 101 * 81, dtype=np.float64)
    assert np.array_equal(pad.from_layout(pad.to_layout(x)), x)


CASES = [
    # sym, ncell, bc, row_align: the row kernel with general faces; all faces Dirichlet (the x-hi
    # face node written from the last tile's halo, n0 = 128); x faces Dirichlet with the face
    # node inside the last tile (n0 = 130); MajorMinor rows
    (P.ISO, (128, 100, 80), "random", 32),
    (P.ISO, (128, 100, 80), "dirichlet", 32),
    (P.ISO, (130, 97, 83), "xdirichlet", 32),
    (P.MAJORMINOR, (64, 40, 36), "random", 16),
]


def _bc(kind, seed):
    if kind == "random":
        return random_bc(seed, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    if kind == "dirichlet":
        return _all_dirichlet()
    t, v = random_bc(seed, (P.TRACTION, P.NEUMANN, P.DISPLACEMENT))
    t[0, :] = P.DISPLACEMENT
    t[3, :] = P.DISPLACEMENT
    return t, v


@pytest.mark.parametrize("sym,ncell,bck,align", CASES)
def test_layout_operator_rows(cuda, sym, ncell, bck, align):
    (cmp_, pad), C, nodes, dx = _pair(sym, ncell, 31, _bc(bck, 31), align)
    print("layout rows", sym, ncell, bck, align, pad.LevelLayout(0), flush=True)
    x = rng(32).uniform(-1, 1, 3 * nodes)
    b = rng(33).uniform(-1, 1, 3 * nodes)
    n3p = 3 * pad.LevelSize(0)
    xc, bc_ = to_dev(x, cuda), to_dev(b, cuda)
    xp, bp = to_dev(pad.to_layout(x), cuda), to_dev(pad.to_layout(b), cuda)
    for name in ("apply", "resid", "diag", "fsmooth"):
        oc = to_dev(np.zeros(3 * nodes), cuda)
        op_ = to_dev(np.zeros(n3p), cuda)
        if name == "apply":
            cmp_.Fapply(0, 0, oc, xc)
            pad.Fapply(0, 0, op_, xp)
        elif name == "resid":
            cmp_.correctionResidual(0, 0, oc, xc, bc_)
            pad.correctionResidual(0, 0, op_, xp, bp)
        elif name == "diag":
            cmp_.Diagonal(0, 0, oc)
            pad.Diagonal(0, 0, op_)
        else:
            oc.copy_(xc)
            op_.copy_(xp)
            cmp_.Fsmooth(0, 0, oc, bc_)
            pad.Fsmooth(0, 0, op_, bp)
        _pads_zero(pad, op_)
        got, ref = pad.from_layout(to_host(op_)), to_host(oc)
        assert np.array_equal(got, ref), "%s: max diff %.3e" % (name, np.abs(got - ref).max())


def test_layout_transfers_and_coefficients(cuda):
    ncell = (128, 100, 80)
    (cmp_, pad), C, nodes, dx = _pair(P.ISO, ncell, 41, random_bc(41), 32)
    for op in (cmp_, pad):
        op.prepareForSolve()
    f = rng(42).uniform(-1, 1, 3 * nodes)
    n1 = cmp_.LevelNodes(1)
    cc = to_dev(np.zeros(3 * n1), cuda)
    cp = to_dev(np.zeros(3 * pad.LevelSize(1)), cuda)
    cmp_.restriction(0, 1, cc, to_dev(f, cuda))
    pad.restriction(0, 1, cp, to_dev(pad.to_layout(f), cuda))
    assert np.array_equal(pad.from_layout(to_host(cp), 1), to_host(cc))
    g = rng(43).uniform(-1, 1, 3 * n1)
    fc = to_dev(np.zeros(3 * nodes), cuda)
    fp = to_dev(np.zeros(3 * pad.LevelSize(0)), cuda)
    cmp_.interpolation(0, 0, fc, to_dev(g, cuda))
    pad.interpolation(0, 0, fp, to_dev(pad.to_layout(g, 1), cuda))
    _pads_zero(pad, fp)
    assert np.array_equal(pad.from_layout(to_host(fp)), to_host(fc))
    for lev in (0, 1):
        kc = to_dev(np.zeros(2 * cmp_.LevelNodes(lev)), cuda)
        kp = to_dev(np.zeros(2 * pad.LevelSize(lev)), cuda)
        cmp_.GetCoeff(lev, kc)
        pad.GetCoeff(lev, kp)
        assert np.array_equal(pad.from_layout(to_host(kp), lev, 2), to_host(kc))


def test_layout_bc_init_and_newton_rhs(cuda):
    ncell = (128, 100, 80)
    t, v = random_bc(51, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    (cmp_, pad), C, nodes, dx = _pair(P.ISO, ncell, 51, (t, v), 32)
    u = rng(52).uniform(-0.01, 0.01, 3 * nodes)
    F0 = rng(53).uniform(-0.01, 0.01, 9 * nodes)
    res = []
    for op in (cmp_, pad):
        n3 = 3 * op.LevelSize(0)
        b = to_dev(np.zeros(n3), cuda)
        op.BCInit(b)
        rhs = to_dev(np.zeros(n3), cuda)
        op.NewtonRHS(to_dev(op.to_layout(u), cuda), b, to_dev(op.to_layout(F0, ncomp=9, aos=True), cuda), rhs)
        res.append((op.from_layout(to_host(b)), op.from_layout(to_host(rhs))))
        if op is pad:
            _pads_zero(pad, rhs)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("precision,krylov", [(0, False), (PREC_MIXED, True)])
def test_layout_solve(cuda, precision, krylov):
    ncell = (128, 96, 80)  # coarsens to an (8, 6, 5) bottom level
    t, v = _all_dirichlet()
    (cmp_, pad), C, nodes, dx = _pair(P.ISO, ncell, 61, (t, v), 32, precision=precision, cycle=CYCLE_V,
                                      pre_smooth=1, post_smooth=1, max_iter=200)
    out = []
    for op in (cmp_, pad):
        n3 = 3 * op.LevelSize(0)
        b = to_dev(np.zeros(n3), cuda)
        op.BCInit(b)
        sol = to_dev(np.zeros(n3), cuda)
        print("layout solve", precision, krylov, op.LevelLayout(0), flush=True)
        s = Krylov(op, restart=16, max_iter=200) if krylov else Linear(op)
        st = s.solve(sol, b, 1e-10, 0.0)
        assert st["converged"] == 1
        out.append((st["iterations"], op.from_layout(to_host(sol))))
    err = np.linalg.norm(out[0][1] - out[1][1]) / np.linalg.norm(out[0][1])
    print("iterations", out[0][0], out[1][0], "rel diff", err)
    assert abs(out[0][0] - out[1][0]) <= (3 if precision == PREC_MIXED else 1)
    assert err <= (1e-8 if precision == PREC_MIXED else 1e-10)


def test_layout_learned_stage(cuda):
    ncell = (128, 96, 80)
    t, v = _all_dirichlet()
    (cmp_, pad), C, nodes, dx = _pair(P.ISO, ncell, 71, (t, v), 32, precision=PREC_MIXED, cycle=CYCLE_V,
                                      pre_smooth=1, post_smooth=1)
    net = Net(seed=3)
    r = rng(72).uniform(-1, 1, 3 * nodes)
    z = []
    for op in (cmp_, pad):
        net.attach(op)
        zd = to_dev(np.zeros(3 * op.LevelSize(0)), cuda)
        print("learned stage", op.LevelLayout(0), flush=True)
        precond_apply(op, PRECOND_LEARNED, to_dev(op.to_layout(r), cuda), zd)
        if op is pad:
            _pads_zero(pad, zd)
        z.append(op.from_layout(to_host(zd)))
    err = np.linalg.norm(z[0] - z[1]) / np.linalg.norm(z[0])
    print("rel diff", err)
    assert err <= 1e-3
    # training-pair generation and a few Adam steps on the pitched handle
    st = Net(seed=4).train(pad, npairs=2, steps=4, lr=1e-3, seed=9, solve_max_iter=20, solve_tol=1e-4)
    assert st["pairs"] == 2 and np.isfinite(st["first_loss"]) and np.isfinite(st["last_loss"])
