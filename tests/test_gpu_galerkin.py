"""GPU parity of the Galerkin coarse operators (coarse_op="galerkin": A_l = R A_{l-1} P on the
levels >= 1 of the cycles; no reference counterpart, the preconditioner alone changes) and the
solve of the sharp two-phase field they exist for.

* Fapply / Diagonal / one l1 Fsmooth call on levels 1 and 2 of a Galerkin handle against the
  oracle's Galerkin stencils (probed compositions), isotropic and MajorMinor, random boundary
  tables;
* FGMRES with the Galerkin V-cycle and the l1 smoother on a 64^3-cell two-phase 1:1000 field with
  a ~4-cell interface (width 0.5), FP64 and mixed precision: the oracle's FP64 operator evaluates
  the true residual (<= 1e-9 |b|), mixed against FP64 to 1e-8.
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd._lib import AlamoError
from alamo_amd.operator import Elastic, Krylov, LAYOUT_AOS, PREC_MIXED, CYCLE_V
from helpers import random_iso, random_mm, random_bc, to_dev, to_host, rng

pytestmark = pytest.mark.gpu


def close(gpu, ref, rtol):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(gpu - ref).max()
    assert err <= rtol * scale, "max err %.3e vs scale %.3e" % (err, scale)


@pytest.mark.parametrize("sym,ncell", [(P.ISO, (32, 24, 16)), (P.MAJORMINOR, (16, 16, 24))])
def test_galerkin_levels(cuda, ora, sym, ncell):
    print("galerkin levels", sym, ncell, flush=True)
    N = int(np.prod([n + 1 for n in ncell]))
    dx = tuple(1.0 / n for n in ncell)
    C = (random_iso if sym == P.ISO else random_mm)(N, 71)
    bct, bcv = random_bc(72, (P.DISPLACEMENT, P.TRACTION, P.NEUMANN))
    op = Elastic(ncell, dx, sym=sym, bc=(bct, bcv), smoother="l1", coarse_op="galerkin", bottom_solver="dense")
    op.SetModel(C, LAYOUT_AOS)
    for lev in (1, 2):
        nc = op.LevelNCell(lev)
        dxl = tuple(h * 2 ** lev for h in dx)
        n = op.LevelNodes(lev)
        S = ora.galerkin(ncell, dx, sym, C, bct, lev)
        x = rng(73 + lev).uniform(-1, 1, 3 * n)
        out = to_dev(np.zeros(3 * n), cuda)
        op.Fapply(0, lev, out, to_dev(x, cuda))
        Sx = ora.stencil_apply(nc, dxl, S, x)
        close(to_host(out), Sx, 1e-12)
        Sr = S.reshape(27, 3, 3, n)
        D = np.stack([Sr[13, q, q] for q in range(3)]).ravel()
        dg = to_dev(np.zeros(3 * n), cuda)
        op.Diagonal(0, lev, dg)
        close(to_host(dg), D, 1e-12)
        d1 = np.abs(Sr).sum(axis=(0, 2)).ravel() * np.where(D < 0, -1.0, 1.0)
        b = rng(75 + lev).uniform(-1, 1, 3 * n)
        xs = x.copy()
        for _ in range(op.desc.sweeps_per_smooth):
            xs = xs + (b - ora.stencil_apply(nc, dxl, S, xs)) / d1
        xd = to_dev(x, cuda)
        op.Fsmooth(0, lev, xd, to_dev(b, cuda))
        close(to_host(xd), xs, 1e-11)


def test_galerkin_needs_direct_or_smoother_bottom(cuda):
    ncell = (16, 16, 16)
    N = 17 ** 3
    op = Elastic(ncell, (1 / 16,) * 3, sym=P.ISO, coarse_op="galerkin", bottom_solver="bicgstab")
    op.SetModel(random_iso(N, 77))
    with pytest.raises(AlamoError, match="bottom_solver"):
        op.prepareForSolve()


def test_sharp_two_phase_fgmres(cuda, ora):
    import torch
    n = 64
    pr = P.two_phase_problem(n=n, corr_cells=8, contrast=1000.0, seed=11, width=0.5)
    print("interface", pr["interface"], flush=True)
    N = (n + 1) ** 3
    b = ora.bc_init(pr["ncell"], pr["bct"], pr["bcv"])
    rhs = ora.newton_rhs(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["F0"], pr["bct"], np.zeros(3 * N), b)
    x = {}
    for prec in (0, PREC_MIXED):
        op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]), precision=prec, cycle=CYCLE_V,
                     pre_smooth=1, post_smooth=1, sweeps_per_smooth=3, sweeps_finest=1, smoother="l1",
                     coarse_op="galerkin", bottom_solver="dense")
        op.SetModel(pr["C"])
        sol = to_dev(np.zeros(3 * N), cuda)
        st = Krylov(op, restart=16, max_iter=400).solve(sol, to_dev(rhs, cuda), 1e-10, 0.0)
        print("precision", prec, "iterations", st["iterations"], "rel res", st["final_res_inf"] / st["rhs_inf"],
              flush=True)
        x[prec] = to_host(sol)
        r = ora.residual(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], x[prec], rhs)
        assert np.abs(r).max() <= 1e-9 * np.abs(rhs).max()
    err = np.linalg.norm(x[PREC_MIXED] - x[0]) / np.linalg.norm(x[0])
    print("mixed vs FP64", err)
    assert err <= 1e-8
