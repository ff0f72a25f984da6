"""Pins the oracle's Elastic::SetUniform rows (Operator/Elastic.cpp:245) and the affine model's
Newton::W / DW (Solver/Nonlocal/Newton.H:528-611, Model/Solid/Affine/Isotropic.H:38-45) with
checks that do not restate its formulas:

* SetUniform with a constant C changes nothing, bit for bit (grad C is exactly 0);
* SetUniform with a varying C: quadratic u has exact central differences, so the interior rows
  are the textbook isotropic operator mu lap u_i + (lambda + mu) d_i div u at each node's own
  (lambda, mu); a linear u leaves 0 in the interior; the boundary rows never change;
* DW / W for a linear u (exact one-sided and central differences at every node) against Hooke's
  law sigma = lambda tr(E) I + mu (E + E^T), E = grad u - F0, for node-varying (lambda, mu) and a
  non-symmetric F0, and against the cubic law (C11, C12, C44 on the crystal axes) for MajorMinor;
  w = E : sigma / 2.
"""
import numpy as np

import oracle_ctypes as ora
from alamo_amd import problems as P
from helpers import rng, random_bc

NC = (6, 5, 7)
DX = (0.3, 0.25, 0.2)
N = int(np.prod([n + 1 for n in NC]))


def grid(ncell=NC, dx=DX):
    Nn = [n + 1 for n in ncell]
    k, j, i = np.meshgrid(*[np.arange(n) for n in Nn[::-1]], indexing="ij")
    return i.ravel() * dx[0], j.ravel() * dx[1], k.ravel() * dx[2], i.ravel(), j.ravel(), k.ravel()


def interior(ncell=NC):
    _, _, _, i, j, k = grid(ncell)
    return (i > 0) & (j > 0) & (k > 0) & (i < ncell[0]) & (j < ncell[1]) & (k < ncell[2])


def iso_random(seed):
    r = rng(seed)
    lam, mu = r.uniform(0.5, 2.0, N), r.uniform(0.5, 2.0, N)
    return lam, mu, P.iso_field(lam, mu)


def test_uniform_constant_model_is_bit_identical():
    C = P.iso_field(np.full(N, 1.3), np.full(N, 0.7))
    bct, _ = random_bc(4)
    u = rng(5).standard_normal(3 * N)
    f = ora.fapply(NC, DX, P.ISO, C, bct, u)
    fu = ora.fapply_uniform(NC, DX, P.ISO, C, bct, u)
    assert np.array_equal(f, fu)


def test_uniform_rows_closed_forms():
    lam, mu, C = iso_random(6)
    bct, _ = random_bc(7)
    x, y, z, _, _, _ = grid()
    X = np.stack([x, y, z])
    r = rng(8)
    A = [(lambda M: M + M.T)(r.standard_normal((3, 3))) for _ in range(3)]  # u_p = x^T A_p x / 2
    u = np.concatenate([0.5 * np.einsum("an,ab,bn->n", X, A[p], X) for p in range(3)])
    fu = ora.fapply_uniform(NC, DX, P.ISO, C, bct, u).reshape(3, N)
    lap = np.array([np.trace(A[p]) for p in range(3)])
    grad_div = np.array([sum(A[k][k, i] for k in range(3)) for i in range(3)])
    inn = interior()
    for i in range(3):
        expect = mu * lap[i] + (lam + mu) * grad_div[i]
        assert np.allclose(fu[i][inn], expect[inn], rtol=1e-9, atol=1e-9)
    # the boundary rows are those of the full operator, and grad(C) . grad(u) is what differs inside
    f = ora.fapply(NC, DX, P.ISO, C, bct, u).reshape(3, N)
    assert np.array_equal(f[:, ~inn], fu[:, ~inn])
    assert np.abs(f[:, inn] - fu[:, inn]).max() > 1e-3
    H = r.standard_normal((3, 3))
    ul = np.concatenate([H[p] @ X for p in range(3)])
    assert np.abs(ora.fapply_uniform(NC, DX, P.ISO, C, bct, ul).reshape(3, N)[:, inn]).max() < 1e-11
    assert np.abs(ora.fapply(NC, DX, P.ISO, C, bct, ul).reshape(3, N)[:, inn]).max() > 1e-3


def linear_u(seed):
    x, y, z, _, _, _ = grid()
    X = np.stack([x, y, z])
    H = rng(seed).standard_normal((3, 3))
    return H, np.concatenate([H[p] @ X for p in range(3)])


def test_dw_w_isotropic_hooke():
    lam, mu, C = iso_random(9)
    H, u = linear_u(10)
    F0 = rng(11).standard_normal((N, 3, 3)) * 0.1
    dw = ora.model_dw(NC, DX, P.ISO, C, F0.ravel(), u).reshape(3, 3, N)
    w = ora.model_w(NC, DX, P.ISO, C, F0.ravel(), u)
    E = H[None] - F0
    tr = np.trace(E, axis1=1, axis2=2)
    sig = lam[:, None, None] * tr[:, None, None] * np.eye(3)[None] + mu[:, None, None] * (E + E.transpose(0, 2, 1))
    assert np.allclose(dw.transpose(2, 0, 1), sig, rtol=1e-11, atol=1e-12)
    assert np.allclose(w, 0.5 * np.einsum("nab,nab->n", E, sig), rtol=1e-11, atol=1e-12)
    # F0 = None is F0 = 0
    assert np.array_equal(ora.model_dw(NC, DX, P.ISO, C, None, u),
                          ora.model_dw(NC, DX, P.ISO, C, np.zeros(9 * N), u))


def test_dw_w_cubic_hooke():
    C11, C12, C44 = 1.68, 1.21, 0.75
    c = np.asarray(ora.cubic(C11, C12, C44))  # crystal axes = lab axes
    C = np.tile(c, N)
    H, u = linear_u(12)
    dw = ora.model_dw(NC, DX, P.MAJORMINOR, C, None, u).reshape(3, 3, N)
    E = 0.5 * (H + H.T)
    sig = np.zeros((3, 3))
    for a in range(3):
        sig[a, a] = C11 * E[a, a] + C12 * (np.trace(E) - E[a, a])
        for b in range(3):
            if a != b:
                sig[a, b] = 2.0 * C44 * E[a, b]
    assert np.allclose(dw, sig[:, :, None], rtol=1e-11, atol=1e-12)
    w = ora.model_w(NC, DX, P.MAJORMINOR, C, None, u)
    assert np.allclose(w, 0.5 * np.sum(H * sig), rtol=1e-11)
