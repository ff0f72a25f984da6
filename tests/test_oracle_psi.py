"""Pins the oracle's psi terms (Operator/Elastic.cpp:118-137,191-195,243-260,340-363,839-907;
Solver/Nonlocal/Newton.H:234-266) with checks that do not restate its formulas:

* psi = 1 everywhere reduces every psi path to the plain one (psi_small = 0: exactly)
* linear u, constant C, psi linear in space: the interior row is (C : grad u) . grad psi (1 - psi_small)
  exactly (the cell gradient of a linear field is exact, the second differences vanish)
* quadratic u, constant C, constant psi = c: interior rows are (c (1 - s) + s) C : grad grad u exactly
* boundary rows with constant psi = c inside and the ghost cells at 1: traction rows scale by
  (1 - s) (n_in c + 8 - n_in) / 8 + s with n_in = 4, 2, 1 interior cells on faces, edges, corners;
  displacement and Neumann rows do not scale
* the diagonal equals (A e_{n,p})_{n,p} with the psi operator (brute force over all unknowns)
* restriction: psi = f(x) g(y) h(z) restricts to the product of the 1-D rules (injection at the
  lo index, [1/4, 1/2, 1/4] elsewhere), which covers every face / edge / corner branch
* Newton right-hand side: psi = 1 is the plain path; linear u, constant C, linear psi gives
  rhs = b - DW . grad psi inside (no psi_small) and the unscaled boundary rows
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from helpers import random_bc

NC = (6, 5, 7)
DX = (0.3, 0.25, 0.2)


def grid(ncell=NC, dx=DX):
    N = [n + 1 for n in ncell]
    k, j, i = np.meshgrid(*[np.arange(n) for n in N[::-1]], indexing="ij")
    return i.ravel() * dx[0], j.ravel() * dx[1], k.ravel() * dx[2], i.ravel(), j.ravel(), k.ravel()


def cells(ncell=NC, dx=DX):
    k, j, i = np.meshgrid(*[np.arange(n) for n in ncell[::-1]], indexing="ij")
    return (i.ravel() + 0.5) * dx[0], (j.ravel() + 0.5) * dx[1], (k.ravel() + 0.5) * dx[2]


def interior(ncell=NC):
    x, y, z, i, j, k = grid(ncell)
    return (i > 0) & (j > 0) & (k > 0) & (i < ncell[0]) & (j < ncell[1]) & (k < ncell[2])


def coef(sym, seed):
    r = np.random.default_rng(seed)
    if sym == P.ISO:
        return r.uniform(0.5, 2, 2)
    if sym == P.MAJORMINOR:
        return P.cubic_mm(1.68, 1.21, 0.75, P.euler_R(*r.uniform(0, 6, 3)))
    return r.uniform(0.5, 2, 9)


def dense(sym, c):
    from oracle_ctypes import c4
    return np.array([[[[c4(sym, c, i, j, k, l) for l in range(3)] for k in range(3)] for j in range(3)]
                     for i in range(3)])


def ncells(ncell=NC):
    return int(np.prod(ncell))


@pytest.mark.parametrize("sym", [P.ISO, P.MAJORMINOR, P.DIAGONAL])
def test_psi_one_is_plain(ora, sym):
    r = np.random.default_rng(1)
    nodes = int(np.prod([n + 1 for n in NC]))
    C = np.concatenate([coef(sym, s) for s in range(nodes)])
    bct, _ = random_bc(2)
    u = r.uniform(-1, 1, 3 * nodes)
    one = np.ones(ncells())
    assert np.array_equal(ora.fapply_psi(NC, DX, sym, C, bct, one, 0.0, u), ora.fapply(NC, DX, sym, C, bct, u))
    assert np.array_equal(ora.diagonal_psi(NC, DX, sym, C, bct, one, 0.0), ora.diagonal(NC, DX, sym, C, bct))
    b = r.uniform(-1, 1, 3 * nodes)
    F0 = r.uniform(-0.01, 0.01, 9 * nodes)
    assert np.array_equal(ora.newton_rhs_psi(NC, DX, sym, C, F0, bct, one, u, b),
                          ora.newton_rhs(NC, DX, sym, C, F0, bct, u, b))


@pytest.mark.parametrize("sym", [P.ISO, P.MAJORMINOR])
def test_linear_u_linear_psi(ora, sym):
    x, y, z, *_ = grid()
    nodes = x.size
    c = coef(sym, 3)
    G = np.random.default_rng(4).uniform(-1, 1, (3, 3))
    u = np.concatenate([G[p, 0] * x + G[p, 1] * y + G[p, 2] * z + 0.1 * p for p in range(3)])
    gp = np.array([0.7, -0.4, 0.9])
    cx, cy, cz = cells()
    psi = 1.5 + gp[0] * cx + gp[1] * cy + gp[2] * cz
    s = 1e-3
    f = ora.fapply_psi(NC, DX, sym, np.tile(c, nodes), random_bc(5)[0], psi, s, u).reshape(3, nodes)
    S = np.einsum("ijkl,kl->ij", dense(sym, c), G)
    want = (1 - s) * S @ gp
    inn = interior()
    assert np.abs(f[:, inn] - want[:, None]).max() < 1e-12 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("sym", [P.ISO, P.MAJORMINOR, P.DIAGONAL])
def test_quadratic_u_constant_psi(ora, sym):
    x, y, z, *_ = grid()
    nodes = x.size
    c = coef(sym, 6)
    H = np.random.default_rng(7).uniform(-1, 1, (3, 3, 3))
    H = 0.5 * (H + H.transpose(0, 2, 1))
    X = np.stack([x, y, z])
    u = np.concatenate([0.5 * np.einsum("an,ab,bn->n", X, H[p], X) for p in range(3)])
    cval, s = 0.37, 1e-2
    f = ora.fapply_psi(NC, DX, sym, np.tile(c, nodes), random_bc(8)[0], np.full(ncells(), cval), s, u)
    want = (cval * (1 - s) + s) * np.einsum("ijkl,klj->i", dense(sym, c), H)
    inn = interior()
    assert np.abs(f.reshape(3, nodes)[:, inn] - want[:, None]).max() < 1e-10 * np.abs(want).max()


def test_boundary_rows_ghost_cells_are_one(ora):
    """Traction rows scale by the node's psi_avg, whose outside cells read 1."""
    x, y, z, i, j, k = grid()
    nodes = x.size
    r = np.random.default_rng(9)
    C = np.concatenate([coef(P.ISO, q) for q in range(nodes)])
    u = r.uniform(-1, 1, 3 * nodes)
    cval, s = 0.25, 1e-3
    for types in ((P.TRACTION,), (P.DISPLACEMENT,), (P.NEUMANN,)):
        bct, _ = random_bc(10, types)
        f0 = ora.fapply(NC, DX, P.ISO, C, bct, u).reshape(3, nodes)
        f1 = ora.fapply_psi(NC, DX, P.ISO, C, bct, np.full(ncells(), cval), s, u).reshape(3, nodes)
        on = [(i == 0) | (i == NC[0]), (j == 0) | (j == NC[1]), (k == 0) | (k == NC[2])]
        nb = on[0].astype(int) + on[1] + on[2]
        n_in = np.where(nb == 1, 4, np.where(nb == 2, 2, 1))
        fac = (1 - s) * (n_in * cval + 8 - n_in) / 8 + s
        bnd = nb > 0
        if types[0] == P.TRACTION:
            assert np.allclose(f1[:, bnd], f0[:, bnd] * fac[bnd], rtol=1e-13, atol=1e-13)
        else:
            assert np.array_equal(f1[:, bnd], f0[:, bnd])


@pytest.mark.parametrize("sym", [P.ISO, P.MAJORMINOR])
def test_diagonal_psi_brute_force(ora, sym):
    nc, dx = (3, 2, 3), (0.5, 0.4, 0.3)
    nodes = int(np.prod([n + 1 for n in nc]))
    r = np.random.default_rng(11)
    C = np.concatenate([coef(sym, q) for q in range(nodes)])
    bct, _ = random_bc(12, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    psi = r.uniform(0.1, 1.0, int(np.prod(nc)))
    D = ora.diagonal_psi(nc, dx, sym, C, bct, psi, 1e-3)
    for q in range(3 * nodes):
        e = np.zeros(3 * nodes)
        e[q] = 1.0
        assert abs(ora.fapply_psi(nc, dx, sym, C, bct, psi, 1e-3, e)[q] - D[q]) < 1e-11 * max(1.0, abs(D[q]))


def restrict_1d(f):
    n = f.size // 2
    out = np.empty(n)
    out[0] = f[0]
    for I in range(1, n):
        out[I] = 0.25 * f[2 * I - 1] + 0.5 * f[2 * I] + 0.25 * f[2 * I + 1]
    return out


def test_psi_restriction_separable(ora):
    nc, dx = (8, 8, 16), (0.1, 0.2, 0.1)  # two coarser levels: (4, 4, 8), (2, 2, 4)
    r = np.random.default_rng(13)
    fx, fy, fz = r.uniform(0.2, 1, nc[0]), r.uniform(0.2, 1, nc[1]), r.uniform(0.2, 1, nc[2])
    psi = (fz[:, None, None] * fy[None, :, None] * fx[None, None, :]).ravel()
    got = ora.psi_level(nc, dx, psi, 1)
    want = (restrict_1d(fz)[:, None, None] * restrict_1d(fy)[None, :, None] * restrict_1d(fx)[None, None, :]).ravel()
    assert np.allclose(got, want, rtol=1e-14, atol=0)
    got2 = ora.psi_level(nc, dx, psi, 2)
    want2 = (restrict_1d(restrict_1d(fz))[:, None, None] * restrict_1d(restrict_1d(fy))[None, :, None] *
             restrict_1d(restrict_1d(fx))[None, None, :]).ravel()
    assert np.allclose(got2, want2, rtol=1e-14, atol=0)
    assert np.allclose(ora.psi_level(nc, dx, np.full(psi.size, 0.3), 2), 0.3, rtol=1e-15)


def test_newton_rhs_linear_psi(ora):
    x, y, z, i, j, k = grid()
    nodes = x.size
    c = coef(P.ISO, 14)
    G = np.random.default_rng(15).uniform(-0.01, 0.01, (3, 3))
    u = np.concatenate([G[p, 0] * x + G[p, 1] * y + G[p, 2] * z for p in range(3)])
    b = np.random.default_rng(16).uniform(-1, 1, 3 * nodes)
    gp = np.array([0.3, 0.8, -0.5])
    cx, cy, cz = cells()
    psi = 2.0 + gp[0] * cx + gp[1] * cy + gp[2] * cz
    bct, _ = random_bc(17, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    C = np.tile(c, nodes)
    rhs = ora.newton_rhs_psi(NC, DX, P.ISO, C, None, bct, psi, u, b).reshape(3, nodes)
    plain = ora.newton_rhs(NC, DX, P.ISO, C, None, bct, u, b).reshape(3, nodes)
    DW = np.einsum("ijkl,kl->ij", dense(P.ISO, c), G)
    inn = interior()
    want = b.reshape(3, nodes)[:, inn] - (DW @ gp)[:, None]
    assert np.abs(rhs[:, inn] - want).max() < 1e-12
    assert np.array_equal(rhs[:, ~inn], plain[:, ~inn])


def test_mg_solve_psi_converges_to_the_operator(ora):
    """The multigrid solve with psi on every level reaches the psi operator's solution."""
    nc = (8, 8, 8)
    dx = (1 / 8,) * 3
    nodes = 9 ** 3
    r = np.random.default_rng(18)
    C = np.concatenate([coef(P.ISO, q) for q in range(nodes)])
    cx, cy, cz = cells(nc, dx)
    psi = np.where((cx - 0.5) ** 2 + (cy - 0.5) ** 2 + (cz - 0.5) ** 2 < 0.09, 0.05, 1.0)
    t, v = P.bc_constant(vals={"XHI": [0.01, 0, 0]})
    b = ora.bc_init(nc, t, v)
    x, it, fr = ora.mg_solve_psi(nc, dx, P.ISO, C, t, psi, 1e-8, b, tol_rel=1e-11)
    assert it > 0
    res = b - ora.fapply_psi(nc, dx, P.ISO, C, t, psi, 1e-8, x)
    assert np.abs(res).max() <= 1e-10 * np.abs(b).max()
    # the void changes the solution
    x0, *_ = ora.mg_solve(nc, dx, P.ISO, C, t, b, tol_rel=1e-11)
    assert np.linalg.norm(x - x0) > 1e-3 * np.linalg.norm(x0)
