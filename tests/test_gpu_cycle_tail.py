"""The FP32 cycle's last smoothed level and dense bottom in one single-workgroup launch
(k_cycle_tail: pre-smoothing, residual, restriction, dense product, interpolation and
post-smoothing in LDS) against the same steps as separate launches (ALAMO_CYCLE_TAIL=0): through
the F- and V-cycle preconditioners and whole mixed-precision solves.  Both paths evaluate the same
rows in FP32, possibly with different fused multiply-adds, so they agree to FP32 rounding."""
import os

import numpy as np
import pytest
import torch

from alamo_amd import problems as P
from alamo_amd.operator import Elastic, Krylov, Linear, precond_apply, PRECOND_MG, PREC_MIXED
from helpers import rng, to_dev, to_host

pytestmark = pytest.mark.gpu


class _env:
    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update({k: str(v) for k, v in self.kv.items()})

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _op(n, seed, **kw):
    pr = P.two_phase_problem(n=n, corr_cells=4, contrast=100.0, seed=seed, width=1.0)
    opts = dict(cycle=1, precision=PREC_MIXED, pre_smooth=1, post_smooth=1, sweeps_per_smooth=3, sweeps_finest=2,
                smoother="l1", bottom_solver="dense", max_coarsening_level=-1)
    opts.update(kw)
    op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=P.bc_constant(), **opts)
    op.SetModel(pr["C"])
    op.prepareForSolve()
    return op, pr


def _rhs(op, pr, cuda):
    N = int(np.prod([c + 1 for c in pr["ncell"]]))
    b = torch.zeros(3 * N, dtype=torch.float64, device=cuda)
    op.BCInit(b)
    rhs = torch.zeros_like(b)
    op.NewtonRHS(torch.zeros_like(b), b, to_dev(pr["F0"], cuda), rhs)
    return rhs


@pytest.mark.parametrize("n,kw", [(64, dict(max_coarsening_level=4)), (32, {}), (64, dict(pre_smooth=1, post_smooth=2)),
                                  (64, dict(cycle=0, sweeps_finest=1)), (16, dict(sweeps_per_smooth=2)),
                                  (8, dict(max_coarsening_level=1, post_smooth=2, sweeps_finest=1)),
                                  (8, dict(max_coarsening_level=1, cycle=0))])
def test_tail_preconditioner_matches_separate_launches(cuda, n, kw):
    """One cycle (the FGMRES preconditioner: F-cycle with interpolated guesses, or the V-cycle from a
    first sweep) with the tail launch equals the separate launches to FP32 rounding.  The two-level
    cases (8 cells on 4, odd level-0 sweep counts) keep level 0 out of the tail: the preconditioner
    aliases level 0's buffers to the Krylov vectors by the parity of its sweeps."""
    op, pr = _op(n, seed=7 + n, **kw)
    N = (n + 1) ** 3
    r = to_dev(rng(3).standard_normal(3 * N), cuda)
    print("case n=%d %s levels %d" % (n, kw, op.NMGLevels()), flush=True)
    z = {}
    for name, v in (("separate", 0), ("tail", 1)):
        zz = torch.zeros_like(r)
        with _env(ALAMO_CYCLE_TAIL=v):
            precond_apply(op, PRECOND_MG, r, zz)
        torch.cuda.synchronize()
        z[name] = to_host(zz)
    scale = np.abs(z["separate"]).max()
    err = np.abs(z["tail"] - z["separate"]).max()
    print("tail vs separate launches: max diff %.3e of %.3e" % (err, scale), flush=True)
    assert np.isfinite(z["tail"]).all()
    assert err <= 2e-5 * scale


@pytest.mark.parametrize("krylov", [True, False])
def test_tail_mixed_solves(cuda, krylov):
    """Mixed FGMRES (F-cycle preconditioner) and the mixed multigrid solve itself (F-cycles whose
    V-cycle from the last smoothed level starts with the first sweep from zero) reach the 1e-10 FP64
    true residual with the tail as without it, in about as many iterations, to the same solution."""
    n = 64
    op, pr = _op(n, seed=5, max_coarsening_level=4, max_iter=400)
    rhs = _rhs(op, pr, cuda)
    res = {}
    for name, v in (("separate", 0), ("tail", 1)):
        x = torch.zeros_like(rhs)
        with _env(ALAMO_CYCLE_TAIL=v):
            if krylov:
                st = Krylov(op, restart=24, max_iter=300).solve(x, rhs, 1e-10, 0.0)
            else:
                st = Linear(op).solve(x, rhs, 1e-10, 0.0)
        print(name, "iterations", st["iterations"], "final", st["final_res_inf"] / st["rhs_inf"], flush=True)
        assert st["converged"] == 1
        assert st["final_res_inf"] <= 1e-10 * st["rhs_inf"]
        res[name] = (st["iterations"], to_host(x))
    assert abs(res["tail"][0] - res["separate"][0]) <= max(2, 0.1 * res["separate"][0])
    xs, xt = res["separate"][1], res["tail"][1]
    err = np.linalg.norm(xt - xs) / np.linalg.norm(xs)
    print("relative L2 distance of the two solutions %.3e" % err, flush=True)
    assert err <= 1e-7
