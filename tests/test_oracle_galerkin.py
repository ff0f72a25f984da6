"""Pins the oracle's Galerkin coarse operators (coarse_op GALERKIN; no reference counterpart) that
the GPU's R A P kernels are checked against:

* the stencil's product with random vectors equals the composition restrict(A(interpolate(x)))
  of the pinned transfers and operator, on levels 1 and 2 (level 2 from level 1's stencil), with
  random boundary tables;
* rows of interior coarse nodes of a constant-coefficient operator annihilate constant and linear
  fields (A kills them and trilinear prolongation reproduces them exactly).
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from helpers import random_iso, random_mm, random_bc, rng


def nodes(ncell):
    return int(np.prod([c + 1 for c in ncell]))


@pytest.mark.parametrize("sym", [P.ISO, P.MAJORMINOR])
def test_stencil_equals_composition(ora, sym):
    ncell, dx = (8, 8, 12), (0.125, 0.1, 0.09)
    N = nodes(ncell)
    C = (random_iso if sym == P.ISO else random_mm)(N, 61)
    bct, _ = random_bc(62, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    nc1 = tuple(c // 2 for c in ncell)
    dx1 = tuple(2 * h for h in dx)
    S1 = ora.galerkin(ncell, dx, sym, C, bct, 1)
    x = rng(63).standard_normal(3 * nodes(nc1))
    comp = ora.restrict(ncell, dx, ora.fapply(ncell, dx, sym, C, bct, ora.interp_add(ncell, dx, x, np.zeros(3 * N))))
    got = ora.stencil_apply(nc1, dx1, S1, x)
    assert np.abs(got - comp).max() <= 1e-12 * np.abs(comp).max()
    nc2 = tuple(c // 2 for c in nc1)
    dx2 = tuple(2 * h for h in dx1)
    S2 = ora.galerkin(ncell, dx, sym, C, bct, 2)
    y = rng(64).standard_normal(3 * nodes(nc2))
    comp2 = ora.restrict(nc1, dx1, ora.stencil_apply(nc1, dx1, S1, ora.interp_add(nc1, dx1, y, np.zeros(3 * nodes(nc1)))))
    got2 = ora.stencil_apply(nc2, dx2, S2, y)
    assert np.abs(got2 - comp2).max() <= 1e-12 * np.abs(comp2).max()


def test_interior_rows_annihilate_linear_fields(ora):
    ncell, dx = (12, 12, 16), (1 / 12, 1 / 12, 1 / 16)
    N = nodes(ncell)
    C = P.iso_field(np.full(N, 1.7), np.full(N, 0.6))
    bct, _ = random_bc(65, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    nc1 = tuple(c // 2 for c in ncell)
    dx1 = tuple(2 * h for h in dx)
    S1 = ora.galerkin(ncell, dx, P.ISO, C, bct, 1)
    Nn = [c + 1 for c in nc1]
    k, j, i = np.meshgrid(*[np.arange(n) for n in Nn[::-1]], indexing="ij")
    X = [i.ravel() * dx1[0], j.ravel() * dx1[1], k.ravel() * dx1[2]]
    G = rng(66).standard_normal((3, 3))
    u = np.concatenate([0.3 * (a + 1) + G[a, 0] * X[0] + G[a, 1] * X[1] + G[a, 2] * X[2] for a in range(3)])
    f = ora.stencil_apply(nc1, dx1, S1, u).reshape(3, -1)
    inner = ((i >= 2) & (i <= nc1[0] - 2) & (j >= 2) & (j <= nc1[1] - 2) & (k >= 2) & (k <= nc1[2] - 2)).ravel()
    assert inner.sum() > 10
    scale = np.abs(S1).max() * np.abs(u).max()
    assert np.abs(f[:, inner]).max() <= 1e-12 * scale
