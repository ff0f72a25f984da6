"""The benchmark configurations as parity cases; each prints the relative L2 displacement error
of the GPU solve against an FP64 reference solve.

* Eshelby sphere, 64^3 cells (65^3 nodes), V-cycle with 2 pre / 2 post Jacobi sweeps, 1e-10:
  FP64 and mixed (FP32 cycles in FP64 refinement) against the oracle's FP64 solve.
* 64-grain Voronoi polycrystal, rotated cubic grains (C11 1.68, C12 1.21, C44 0.75), u_z on top,
  bottom fixed: with Neumann lateral faces a mixed V-cycle solve at 64^3 against the oracle; with
  the configuration's traction-free lateral faces FGMRES (point Jacobi V-cycle, dense bottom,
  which ends the hierarchy at 4 cells) at 128^3 and at the configuration's 256^3, the oracle
  evaluating the GPU solution's true residual.
* Two-phase 1:1000 field at 512^3 cells (513^3 nodes, correlation 16 cells, seed 5, erf interface
  w = 0.7): plain FP64 multigrid, FP64 / mixed / learned FGMRES; the learned and mixed solves
  against the FP64 FGMRES solve, the oracle evaluating the mixed solution's true residual (a full
  oracle solve would take hours).
Tolerance: relative L2 error 1e-8 (the target after refinement); true residuals 1e-9 of |b|
(1e-10 target plus the oracle's own rounding).
"""
import numpy as np
import pytest

from alamo_amd import problems as P
from alamo_amd.operator import Elastic, Linear, Krylov, LAYOUT_AOS, PREC_MIXED, CYCLE_V
from helpers import eshelby_rhs, to_dev, to_host

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_config_eshelby_65_vcycle(cuda, ora, precision):
    n = 64
    pr, bc, rhs = eshelby_rhs(ora, n)
    xref, it, *_ = ora.mg_solve(pr["ncell"], pr["dx"], P.ISO, pr["C"], bc[0], rhs, tol_rel=1e-12, cycle=0, nu1=1,
                                nu2=1, max_iter=400)
    op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=bc, precision=precision, cycle=CYCLE_V, pre_smooth=1,
                 post_smooth=1, max_iter=400)
    op.SetModel(pr["C"])
    sol = to_dev(np.zeros_like(rhs), cuda)
    print("eshelby 64^3 V-cycle precision", precision, flush=True)
    st = Linear(op).solve(sol, to_dev(rhs, cuda), 1e-10, 0.0)
    err = _rel(to_host(sol), xref)
    print("GPU iterations", st["iterations"], "oracle iterations", it, "rel L2 error vs oracle FP64", err)
    assert st["converged"] == 1 and err <= 1e-8


def _polycrystal(ora, n, lateral=P.NEUMANN):
    pr = P.polycrystal_problem(n=n, ngrains=64, seed=64, lateral=lateral)
    N = (n + 1) ** 3
    b = ora.bc_init(pr["ncell"], pr["bct"], pr["bcv"])
    return pr, N, b


def _poly_op(pr, precision, records=False, **kw):
    """the polycrystal handle: a per-grain model, or (records) the same stiffness as full 21-slot
    records per node, the form an integrator hands over"""
    op = Elastic(pr["ncell"], pr["dx"], sym=P.MAJORMINOR, bc=(pr["bct"], pr["bcv"]), precision=precision,
                 cycle=CYCLE_V, pre_smooth=1, post_smooth=1, max_iter=400, **kw)
    if records:
        op.SetModel(pr["table"][pr["gid"]].ravel(), LAYOUT_AOS)
    else:
        op.SetModelGrains(pr["gid"], pr["table"])
    return op


def test_config_polycrystal_64_vs_oracle(cuda, ora):
    n = 64
    pr, N, b = _polycrystal(ora, n)
    C = pr["table"][pr["gid"]].ravel()
    rhs = ora.newton_rhs(pr["ncell"], pr["dx"], P.MAJORMINOR, C, np.zeros(9 * N), pr["bct"], np.zeros(3 * N), b)
    xref, it, *_ = ora.mg_solve(pr["ncell"], pr["dx"], P.MAJORMINOR, C, pr["bct"], rhs, tol_rel=1e-12, cycle=0,
                                nu1=1, nu2=1, max_iter=600)
    op = _poly_op(pr, PREC_MIXED)
    sol = to_dev(np.zeros(3 * N), cuda)
    print("polycrystal 64^3 mixed", flush=True)
    st = Linear(op).solve(sol, to_dev(rhs, cuda), 1e-10, 0.0)
    err = _rel(to_host(sol), xref)
    print("GPU iterations", st["iterations"], "oracle iterations", it, "rel L2 error vs oracle FP64", err)
    assert st["converged"] == 1 and err <= 1e-8


@pytest.mark.parametrize("n,records", [(128, False), (256, False), (128, True)])
def test_config_polycrystal_true_residual(cuda, ora, n, records):
    pr, N, b = _polycrystal(ora, n, lateral=P.TRACTION)
    C = pr["table"][pr["gid"]].ravel()
    rhs = ora.newton_rhs(pr["ncell"], pr["dx"], P.MAJORMINOR, C, np.zeros(9 * N), pr["bct"], np.zeros(3 * N), b)
    x = {}
    for prec in (0, PREC_MIXED):
        op = _poly_op(pr, prec, records, sweeps_per_smooth=3, sweeps_finest=2, bottom_solver="dense")
        full = op.NMGLevels()
        op.prepareForSolve()
        # the 2-cell level is singular with traction-free sides: the hierarchy ends at 4 cells
        assert op.NMGLevels() == full - 1 and op.LevelNCell(op.NMGLevels() - 1)[0] == 4
        sol = to_dev(np.zeros(3 * N), cuda)
        print("traction-free polycrystal %d^3 FGMRES precision" % n, prec, "records" if records else "grains",
              flush=True)
        st = Krylov(op, restart=16, max_iter=400).solve(sol, to_dev(rhs, cuda), 1e-10, 0.0)
        print("iterations", st["iterations"], "seconds", st["seconds"])
        assert st["converged"] == 1
        x[prec] = to_host(sol)
        del op
    r = ora.residual(pr["ncell"], pr["dx"], P.MAJORMINOR, C, pr["bct"], x[PREC_MIXED], rhs)
    err = _rel(x[PREC_MIXED], x[0])
    print("oracle |r|/|b|", np.abs(r).max() / np.abs(rhs).max(), "rel L2 mixed vs FP64", err)
    assert np.abs(r).max() <= 1e-9 * np.abs(rhs).max()
    assert err <= 1e-8


def test_config_two_phase_513_learned_vs_plain(cuda, ora):
    """The 512^3-cell two-phase 1:1000 field (correlation 16 cells, seed 5, the benchmark's erf
    interface w = 0.7): plain FP64 multigrid (V-cycles; the reference's damped point Jacobi and the
    l1 smoother), FP64 and mixed FGMRES, and the learned-preconditioned mixed FGMRES (network trained
    on two 65^3 fields of the same kind), all to 1e-10.  Every method's outcome is printed; the
    FGMRES methods must converge, agree with the FP64 solve to 1e-8, and the oracle's FP64 operator
    evaluates the mixed solution's true residual on the host."""
    import torch
    from alamo_amd.operator import Net, PRECOND_LEARNED, PRECOND_MG
    n = 512
    print("generating the 513^3 two-phase field", flush=True)
    pr = P.two_phase_problem(n=n, corr_cells=16, contrast=1000.0, seed=5, width=0.7)
    print("interface", pr["interface"], flush=True)
    N = (n + 1) ** 3
    C = to_dev(pr["C"], cuda)
    F0 = to_dev(pr["F0"], cuda)
    del pr["F0"], pr["eta"]
    net = Net(seed=11)
    for f in range(2):
        prt = P.two_phase_problem(n=64, corr_cells=8, contrast=1000.0, seed=40 + f, width=0.7)
        opt = Elastic(prt["ncell"], prt["dx"], sym=P.ISO, bc=(prt["bct"], prt["bcv"]), precision=PREC_MIXED,
                      cycle=CYCLE_V, pre_smooth=1, post_smooth=1, sweeps_per_smooth=4, sweeps_finest=1,
                      smoother="l1", bottom_solver="dense")
        opt.SetModel(prt["C"])
        print("train", f, net.train(opt, npairs=8, steps=200, lr=1e-3, seed=60 + f), flush=True)
        del opt
    rhs, x, res = None, {}, {}
    methods = [("fp64-mg-jacobi", 0, None, "jacobi"), ("fp64-mg-l1", 0, None, "l1"),
               ("fp64-fgmres", 0, PRECOND_MG, "l1"), ("mixed-fgmres", PREC_MIXED, PRECOND_MG, "l1"),
               ("learned-mixed-fgmres", PREC_MIXED, PRECOND_LEARNED, "l1")]
    for name, prec, kind, sm in methods:
        krylov = kind is not None
        op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]), precision=prec, cycle=CYCLE_V,
                     pre_smooth=1, post_smooth=1, sweeps_per_smooth=4 if krylov else 2,
                     sweeps_finest=1 if krylov else 0, smoother=sm, bottom_solver="dense" if krylov else "bicgstab",
                     max_iter=300)
        op.SetModel(C, layout=0)
        if rhs is None:
            b = torch.zeros(3 * N, dtype=torch.float64, device=cuda)
            op.BCInit(b)
            rhs = torch.zeros_like(b)
            op.NewtonRHS(torch.zeros_like(b), b, F0, rhs)
            del b, F0
            torch.cuda.synchronize()
        if kind == PRECOND_LEARNED:
            net.attach(op)
        sol = torch.zeros(3 * N, dtype=torch.float64, device=cuda)
        print("513^3", name, flush=True)
        s = Krylov(op, restart=16, precond=kind, max_iter=1000) if krylov else Linear(op)
        st = s.solve(sol, rhs, 1e-10, 0.0, raise_on_fail=False)
        res[name] = st
        print("  converged", st["converged"], "iterations", st["iterations"], "seconds", st["seconds"],
              "rel res", st["final_res_inf"] / st["rhs_inf"], flush=True)
        if krylov:
            assert st["converged"] == 1, name
            x[name] = sol
        del op, s
        torch.cuda.empty_cache()
    for name in ("mixed-fgmres", "learned-mixed-fgmres"):
        err = float(torch.linalg.norm(x[name] - x["fp64-fgmres"]) / torch.linalg.norm(x["fp64-fgmres"]))
        print("rel L2", name, "vs FP64 FGMRES", err, flush=True)
        assert err <= 1e-8
    u = to_host(x["mixed-fgmres"])
    bh = to_host(rhs)
    del x
    torch.cuda.empty_cache()
    r = ora.residual(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], u, bh)
    print("oracle |b - A u|/|b| of the mixed solution", np.abs(r).max() / np.abs(bh).max(), flush=True)
    assert np.abs(r).max() <= 1e-9 * np.abs(bh).max()


def test_config_training_129_2000_pairs(cuda):
    """The training configuration at its stated size: 2000 (residual, cycle) pairs on 128^3-cell
    (129^3-node) two-phase fields (correlation 8 cells, contrast 1:100, lambda = mu per phase, the
    benchmark's erf interface w = 0.7), 10 fields x 200 pairs, one Adam step per pair, trained on the
    GPU by the library's BF16 MFMA forward / backward kernels on the residual loss (the stage's squared
    residual ratio 1 - <q, Ay>^2 / (|q|^2 |Ay|^2), 1 for a network that removes nothing).  The loss
    must fall from about 1 to below 0.2 of its first value (measured 0.999 -> 0.035) and stay finite."""
    import time
    from alamo_amd.operator import Net, CYCLE_F
    net = Net(seed=2024)
    t0 = time.time()
    first = last = None
    for f in range(10):
        pr = P.two_phase_problem(n=128, corr_cells=8, contrast=100.0, seed=1 + f, width=0.7)
        op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]), precision=PREC_MIXED,
                     cycle=CYCLE_F, pre_smooth=1, post_smooth=1, sweeps_per_smooth=3, sweeps_finest=2,
                     smoother="l1", bottom_solver="dense")
        op.SetModel(pr["C"])
        st = net.train(op, npairs=200, steps=200, lr=1e-3, seed=100 + f, solve_max_iter=0, loss="residual")
        print("field %d: pairs %d, loss %.4f -> %.4f, train %.1fs" % (f, st["pairs"], st["first_loss"],
                                                                      st["last_loss"], st["seconds_train"]),
              flush=True)
        assert st["pairs"] == 200
        assert np.isfinite(st["first_loss"]) and np.isfinite(st["last_loss"])
        first = st["first_loss"] if first is None else first
        last = st["last_loss"]
        del op
    print("2000 pairs: loss %.4f -> %.4f in %.1f s" % (first, last, time.time() - t0), flush=True)
    assert first > 0.9
    assert last < 0.2 * first


def test_bench_field_257_true_residual(cuda, ora):
    """The benchmark's workload: the 256^3-cell (257^3-node) two-phase 1:1000 field (correlation 8
    cells, seed 256, erf interface w = 0.7), solved by mixed and FP64 FGMRES(24) with the benchmark's
    l1 F-cycle (2 sweeps on the finest level, 3 below, coarsest level 4 cells, dense bottom) to 1e-10.
    The oracle's FP64 operator evaluates the mixed solution's true residual on the host, and the two
    solutions agree to 1e-8."""
    import torch
    from alamo_amd.operator import CYCLE_F
    n = 256
    print("generating the 257^3 two-phase field", flush=True)
    pr = P.two_phase_problem(n=n, corr_cells=8, contrast=1000.0, seed=256, width=0.7)
    print("interface", pr["interface"], flush=True)
    N = (n + 1) ** 3
    x, st = {}, {}
    rhs = None
    for prec in (PREC_MIXED, 0):
        op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]), precision=prec, cycle=CYCLE_F,
                     pre_smooth=1, post_smooth=1, sweeps_per_smooth=3, sweeps_finest=2, smoother="l1",
                     bottom_solver="dense", max_coarsening_level=6)
        op.SetModel(pr["C"])
        if rhs is None:
            b = torch.zeros(3 * N, dtype=torch.float64, device=cuda)
            op.BCInit(b)
            rhs = torch.zeros_like(b)
            op.NewtonRHS(torch.zeros_like(b), b, to_dev(pr["F0"], cuda), rhs)
            del b
        sol = torch.zeros_like(rhs)
        st[prec] = Krylov(op, restart=24, max_iter=400).solve(sol, rhs, 1e-10, 0.0)
        print("precision", prec, "iterations", st[prec]["iterations"], "seconds", st[prec]["seconds"], flush=True)
        assert st[prec]["converged"] == 1
        x[prec] = to_host(sol)
        del op, sol
        torch.cuda.empty_cache()
    rh = to_host(rhs)
    r = ora.residual(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], x[PREC_MIXED], rh)
    rel = np.abs(r).max() / np.abs(rh).max()
    err = _rel(x[PREC_MIXED], x[0])
    print("oracle |b - A u|inf / |b|inf %.3e, rel L2 mixed vs FP64 %.3e" % (rel, err), flush=True)
    assert rel <= 1e-9
    assert err <= 1e-8
