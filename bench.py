"""Benchmark: time to a 1e-10 relative FP64 residual on a 257^3-node two-phase field, and
Fapply HBM bandwidth.

A step is one complete solve A u = b (initial guess 0) of the 256^3-cell / 257^3-node
two-phase isotropic field (Gaussian random field g, correlation length 8 cells, phase fraction
eta = (1 + erf(g / width)) / 2 with width 0.7: a 10-90 % interface of about 6 cells; stiffness
lambda = mu = 1000^eta, dilatational eigenstrain 1e-3 eta, u = 0 on all 26 boundary regions),
through the C ABI, with b, C and u resident in HBM.

Prints one JSON line (rank 0).  `value` is the time-to-solution of the fastest converged solver
method in seconds (lower is better; null when none converged); every method, converged or not,
is listed in config.methods with its iterations, final residual and relative L2 distance to an
FP64 reference solve.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=256, help="cells per direction")
    p.add_argument("--corr", type=float, default=8.0)
    p.add_argument("--contrast", type=float, default=1000.0)
    p.add_argument("--seed", type=int, default=256)
    p.add_argument("--width", type=float, default=0.7,
                   help="interface parameter w of eta = (1 + erf(g / w)) / 2 (0: thresholded)")
    p.add_argument("--krylov-smoother", default="l1", choices=["jacobi", "l1"],
                   help="smoother of the FGMRES cycle (plain multigrid methods: 'jacobi' unless the method "
                        "name ends in -l1)")
    p.add_argument("--mg-max-iter", type=int, default=1500, help="cycle / iteration limit of the solves")
    p.add_argument("--fapply-reps", type=int, default=100)
    p.add_argument("--restart", type=int, default=24, help="FGMRES restart length")
    p.add_argument("--krylov-smooth", type=int, default=1,
                   help="Fsmooth calls before and after the coarse correction in the FGMRES cycle")
    p.add_argument("--krylov-sweeps", type=int, default=3, help="smoothing sweeps per Fsmooth call in the FGMRES cycle")
    p.add_argument("--krylov-sweeps-finest", type=int, default=2,
                   help="Jacobi sweeps per Fsmooth call on the finest level of the FGMRES cycle (0: --krylov-sweeps)")
    p.add_argument("--krylov-coarsest", type=int, default=4,
                   help="cells per direction of the coarsest level of the FGMRES cycle (0: coarsen to 2 cells)")
    p.add_argument("--krylov-bottom", default="dense", choices=["bicgstab", "dense", "smoother"],
                   help="bottom solver of the FGMRES cycle (the plain multigrid methods keep bicgstab)")
    p.add_argument("--skip-cpu", action="store_true")
    p.add_argument("--coarse-op", default="restrict", choices=["restrict", "galerkin"],
                   help="coarse-level operators of the FGMRES cycle (methods ending in -gal: galerkin)")
    p.add_argument("--methods",
                   default="fp64,fp64v,fp64-l1,fgmres-fp64,fgmres-mixed,fgmres-mixedv,fgmres-mixed-gal,"
                           "fgmres-learned-mixed",
                   help="fp64/mixed: F-cycles with 2+2 Fsmooth calls (the reference's MLMG settings); "
                        "fp64v/mixedv: V-cycles with 1+1 Fsmooth calls (2 pre / 2 post Jacobi sweeps), the "
                        "reference's damped point Jacobi, or the l1 smoother with the suffix -l1; "
                        "fgmres-*: flexible GMRES(--restart) preconditioned by one cycle (V with the suffix v, "
                        "else F) with --krylov-smooth Fsmooth calls of --krylov-sweeps sweeps each side "
                        "(--krylov-sweeps-finest on the finest level); "
                        "fgmres-learned-*: the same with the learned correction stage after each cycle; "
                        "the suffix -gal selects Galerkin coarse operators")
    p.add_argument("--train-n", type=int, default=128, help="cells per direction of the training fields")
    p.add_argument("--train-fields", type=int, default=2)
    p.add_argument("--train-pairs", type=int, default=8, help="(residual, error) pairs per training field")
    p.add_argument("--train-steps", type=int, default=300, help="Adam steps per training field")
    p.add_argument("--train-loss", default="residual", choices=["residual", "error"],
                   help="network training loss: the stage's squared residual ratio 1 - cos^2(q, A y), or the "
                        "MSE against the cycle's remaining error (one FGMRES solve per pair)")
    p.add_argument("--verbose", type=int, default=0)
    p.add_argument("--fapply-only", action="store_true",
                   help="only time Fapply on the finest level (plus a device copy of the same bytes), "
                        "for counter runs")
    return p.parse_args()


def fapply_only(a):
    import numpy as np
    import torch
    from alamo_amd import problems as P
    from alamo_amd.operator import Elastic
    dev = torch.device("cuda:0")
    n = a.n
    N = (n + 1) ** 3
    pr = P.two_phase_problem(n=n, corr_cells=a.corr, contrast=a.contrast, seed=a.seed, width=a.width)
    op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]))
    op.SetModel(pr["C"])
    op.prepareForSolve()
    u = torch.rand(3 * N, dtype=torch.float64, device=dev)
    f = torch.empty_like(u)
    c = torch.empty_like(u)
    for _ in range(20):
        op.Fapply(0, 0, f, u)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.fapply_reps):
        op.Fapply(0, 0, f, u)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.fapply_reps
    # calibration copy: reads and writes 3 N doubles
    for _ in range(3):
        c.copy_(u)
    torch.cuda.synchronize()
    del c
    att = measure_stream(a, N, dev)
    print(json.dumps({"fapply_ms": ms, "fapply_gbs": 64.0 * N / (ms * 1e-3) / 1e9, "nodes": N,
                      "copy_bytes_each_way": 24 * N, "attainable_gbs": att}))


def main():
    a = parse()
    if a.fapply_only:
        fapply_only(a)
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the solver has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from alamo_amd import problems as P
    from alamo_amd.operator import (Elastic, Linear, Krylov, Net, PREC_FP64, PREC_MIXED, CYCLE_F, CYCLE_V,
                                    PRECOND_MG, PRECOND_LEARNED)

    dev = torch.device("cuda:0")
    n = a.n
    t0 = time.time()
    pr = P.two_phase_problem(n=n, corr_cells=a.corr, contrast=a.contrast, seed=a.seed, width=a.width)
    print("problem generated in %.1fs: %s" % (time.time() - t0, json.dumps(pr["interface"])), file=sys.stderr,
          flush=True)
    N = (n + 1) ** 3
    bc = (pr["bct"], pr["bcv"])
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)

    # the FGMRES cycle stops coarsening at --krylov-coarsest cells (its dense bottom); the plain
    # multigrid methods keep the reference's full depth
    mcl = int(round(math.log2(n / a.krylov_coarsest))) if a.krylov_coarsest > 0 else -1

    def make_op(prec, vcyc, krylov, smoother, coarse_op="restrict"):
        nsm = a.krylov_smooth if krylov else (1 if vcyc else 2)
        return Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=bc, precision=prec, cycle=CYCLE_V if vcyc else CYCLE_F,
                       pre_smooth=nsm, post_smooth=nsm, sweeps_per_smooth=a.krylov_sweeps if krylov else 2,
                       sweeps_finest=a.krylov_sweeps_finest if krylov else 0, smoother=smoother,
                       max_iter=a.mg_max_iter, verbose=a.verbose, coarse_op=coarse_op,
                       bottom_solver=a.krylov_bottom if krylov or coarse_op == "galerkin" else "bicgstab",
                       max_coarsening_level=mcl if krylov else -1)

    results = {}
    sols = {}
    rhs = None
    fapply = None
    net, train_info = None, None
    for meth in a.methods.split(","):
        krylov = meth.startswith("fgmres-")
        base = meth[len("fgmres-"):] if krylov else meth
        learned = base.startswith("learned-")
        if learned:
            base = base[len("learned-"):]
        gal = base.endswith("-gal")
        if gal:
            base = base[:-4]
        l1 = base.endswith("-l1")
        if l1:
            base = base[:-3]
        smoother = a.krylov_smoother if krylov else ("l1" if l1 else "jacobi")
        coarse_op = "galerkin" if gal else (a.coarse_op if krylov else "restrict")
        prec = PREC_MIXED if base.startswith("mixed") else PREC_FP64
        vcyc = base.endswith("v")
        if learned and net is None:
            net, train_info = train_network(a, P, Elastic, Net, prec, CYCLE_V if vcyc else CYCLE_F, smoother,
                                            coarse_op)
        op = make_op(prec, vcyc, krylov, smoother, coarse_op)
        op.SetModel(pr["C"])
        op.prepareForSolve()
        if rhs is None:
            b = torch.zeros(3 * N, dtype=torch.float64, device=dev)
            op.BCInit(b)
            rhs = torch.zeros_like(b)
            op.NewtonRHS(torch.zeros_like(b), b, T(pr["F0"]), rhs)
            del b
        sol = torch.zeros(3 * N, dtype=torch.float64, device=dev)
        if learned:
            net.attach(op)
        lin = (Krylov(op, restart=a.restart, precond=PRECOND_LEARNED if learned else PRECOND_MG,
                      max_iter=a.mg_max_iter, verbose=a.verbose) if krylov else Linear(op))
        ok = True
        st = None
        for _ in range(max(1, a.warmup)):
            sol.zero_()
            op.prepareForSolve()
            torch.cuda.synchronize()
            tw = time.perf_counter()
            st = lin.solve(sol, rhs, 1e-10, 0.0, raise_on_fail=False)
            torch.cuda.synchronize()
            tw = time.perf_counter() - tw
            ok = st["converged"] == 1
            if not ok:
                break
        rec = dict(precision="f64" if prec == PREC_FP64 else "f32-in-f64", smoother=smoother, coarse_op=coarse_op)
        if ok:
            stats = []
            torch.cuda.synchronize()
            ts = time.perf_counter()
            for _ in range(a.steps):
                # a step: coefficient restriction to every level, diagonals, then the solve from u0 = 0
                sol.zero_()
                op.prepareForSolve()
                stats.append(lin.solve(sol, rhs, 1e-10, 0.0, raise_on_fail=False))
            torch.cuda.synchronize()
            te = time.perf_counter()
            st = stats[-1]
            rec.update(seconds=(te - ts) / len(stats), iterations=st["iterations"],
                       final_rel_res=st["final_res_inf"] / max(st["rhs_inf"], 1e-300),
                       converged=all(x["converged"] for x in stats))
        else:
            # not converged (or diverged): one attempt, not timed as a step
            rec.update(seconds=None, attempt_seconds=tw, iterations=st["iterations"],
                       final_rel_res=st["final_res_inf"] / max(st["rhs_inf"], 1e-300), converged=False,
                       rc=st["rc"])
        results[meth] = rec
        if rec["converged"]:
            sols[meth] = sol.detach().clone()
        if learned:
            rec["fallback_events"] = st["fallback_events"]
            md, mw = net.layer_times_bwd(op, reps=5)
            rec["layers"] = layer_report(net.layer_times(op, reps=5), n, md, mw)
        print("%s: %s s/solve, %d iterations, final %.2e%s" % (meth, "%.4f" % rec["seconds"] if rec["seconds"] else "-",
                                                               rec["iterations"], rec["final_rel_res"],
                                                               "" if rec["converged"] else " (not converged)"),
              file=sys.stderr, flush=True)
        if fapply is None:
            fapply = measure_fapply(a, op, N, dev)
        del op, lin, sol
        torch.cuda.empty_cache()

    # reference solution: FP64 FGMRES to 1e-13 (untimed); relative L2 distance of every method's solution
    ref = reference_solve(a, make_op, pr["C"], rhs, N, dev)
    for k, x in sols.items():
        results[k]["rel_l2_vs_fp64_ref"] = float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref))

    conv = [k for k in results if results[k]["converged"]]
    best = min(conv, key=lambda k: results[k]["seconds"]) if conv else None
    val = results[best]["seconds"] if best else None

    # the metric's comparison: the learned-preconditioned solve against the plain FP64 multigrid
    # (the fastest converged stationary cycle on the GPU), both to the same 1e-10 FP64 true residual
    plain = [k for k in conv if not k.startswith("fgmres-") and results[k]["precision"] == "f64"]
    learned = [k for k in conv if k.startswith("fgmres-learned-")]
    lvp = None
    if plain and learned:
        pk = min(plain, key=lambda k: results[k]["seconds"])
        lk = min(learned, key=lambda k: results[k]["seconds"])
        lvp = {"learned": lk, "learned_s": results[lk]["seconds"], "learned_iterations": results[lk]["iterations"],
               "plain_mg": pk, "plain_mg_s": results[pk]["seconds"], "plain_mg_iterations": results[pk]["iterations"],
               "speedup": results[pk]["seconds"] / results[lk]["seconds"],
               "unconverged_plain": [k for k in results if not k.startswith("fgmres-") and k not in conv]}

    # the value method's solution, checked by the oracle's FP64 operator on the host
    oracle_res = None
    if best is not None and not a.skip_cpu:
        oracle_res = oracle_residual(pr, sols[best].cpu().numpy(), rhs.cpu().numpy())

    cpu = None
    if not a.skip_cpu and best is not None:
        cpu = cpu_baseline(pr, rhs.cpu().numpy(), a, results)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "fapply_iso_f64_257_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    out = {
        "metric": "time-to-1e-10 residual at 256^3 nodes (learned vs plain MG); Fapply HBM GB/s",
        "value": val,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": val * 1e3 if val is not None else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": results[best]["precision"] if best else None,
        "data": "synthetic",
        "config": {
            "workload": "two-phase GRF (corr %g cells, contrast 1:%g, seed %d) isotropic field, eta = (1 + "
                        "erf(g/%g))/2: 10-90%% interface %.1f cells, %.0f%% of nodes within 5%% of a pure phase; %d^3 "
                        "cells / %d^3 nodes, Dirichlet box, eigenstrain load; solve to |b-Au|inf/|b|inf <= 1e-10"
                        % (a.corr, a.contrast, a.seed, a.width, pr["interface"]["interface_cells_10_90"],
                           100 * pr["interface"]["pure_phase_frac"], n, n + 1),
            "interface": pr["interface"],
            "method": best,
            "fgmres": {"restart": a.restart, "cycle_smooth_calls": a.krylov_smooth,
                       "sweeps_per_call": a.krylov_sweeps,
                       "sweeps_per_call_finest": a.krylov_sweeps_finest or a.krylov_sweeps,
                       "smoother": a.krylov_smoother, "bottom_solver": a.krylov_bottom},
            "oracle_true_rel_res": oracle_res,
            "learned_vs_plain_mg": lvp,
            "methods": results,
            "training": train_info,
            "fapply_gbs": fapply["gbs"] if fapply else None,
            "fsmooth_sweep_ms": fapply["fsmooth_sweep_ms"] if fapply else None,
            "fsmooth_sweep_gbs": fapply["fsmooth_sweep_gbs"] if fapply else None,
            "fapply_ms": fapply["ms"] if fapply else None,
            "fapply_timing": "HIP events around %d launches on the launch stream" % a.fapply_reps,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": fapply["gbs"] if fapply else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (fapply["gbs"] / HBM_PEAK_GBS) if fapply else None,
            "traffic": traffic,
            "attainable": fapply["attainable_gbs"] if fapply else None,
            "frac_of_attainable": (fapply["gbs"] / fapply["attainable_gbs"]) if fapply else None,
            "attainable_note": "a plain stream with Fapply's bytes per node (5 FP64 reads, 3 writes, "
                               "alamo_stream_probe) timed the same way on the same device",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def measure_fapply(a, op, N, dev):
    """Fapply (level 0, FP64) and one FP64 Jacobi sweep: mean durations from HIP events on the
    launch stream (the default stream, where the library launches)"""
    import torch
    u = torch.rand(3 * N, dtype=torch.float64, device=dev)
    f = torch.empty_like(u)
    # warm-up launches first: the first launches after other work run ~15% slower (clocks, caches)
    for _ in range(20):
        op.Fapply(0, 0, f, u)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.fapply_reps):
        op.Fapply(0, 0, f, u)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.fapply_reps
    bytes_per_launch = 64.0 * N  # (3 u + 3 f + 2 C) x 8 B per node, isotropic FP64
    out = dict(ms=ms, gbs=bytes_per_launch / (ms * 1e-3) / 1e9, bytes=bytes_per_launch)
    # Fsmooth (level 0, FP64): fused sweeps, (3 x + 3 b + 2 C + 3 x') x 8 B per node and sweep (the
    # point diagonal is recomputed from lambda, mu; the l1 diagonal from them and their differences)
    b = torch.rand_like(u)
    for _ in range(3):
        op.Fsmooth(0, 0, f, b)
    torch.cuda.synchronize()
    reps = max(1, a.fapply_reps // 5)
    e0.record()
    for _ in range(reps):
        op.Fsmooth(0, 0, f, b)
    e1.record()
    torch.cuda.synchronize()
    ms_sweep = e0.elapsed_time(e1) / reps / op.desc.sweeps_per_smooth
    out["fsmooth_sweep_ms"] = ms_sweep
    out["fsmooth_sweep_gbs"] = 88.0 * N / (ms_sweep * 1e-3) / 1e9
    del u, f, b
    out["attainable_gbs"] = measure_stream(a, N, dev)
    return out


def measure_stream(a, N, dev):
    """the attainable rate of Fapply's traffic: the library's streaming kernel with the same five
    FP64 reads and three writes per node over N nodes (alamo_stream_probe), HIP events"""
    import ctypes
    import torch
    from alamo_amd._lib import check, lib
    src = torch.rand(5 * N, dtype=torch.float64, device=dev)
    dst = torch.empty(3 * N, dtype=torch.float64, device=dev)
    call = lambda: check(lib().alamo_stream_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                                  N, None))
    for _ in range(20):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.fapply_reps):
        call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.fapply_reps
    del src, dst
    return 64.0 * N / (ms * 1e-3) / 1e9


def reference_solve(a, make_op, C, rhs, N, dev):
    """FP64 FGMRES with the l1 Galerkin V-cycle to |b - A u|_inf <= 1e-13 |b|_inf (not timed)"""
    import torch
    from alamo_amd.operator import Krylov, PREC_FP64, PRECOND_MG
    op = make_op(PREC_FP64, True, True, "l1", "galerkin")
    op.SetModel(C)
    sol = torch.zeros(3 * N, dtype=torch.float64, device=dev)
    st = Krylov(op, restart=a.restart, precond=PRECOND_MG, max_iter=2000).solve(sol, rhs, 1e-13, 0.0,
                                                                                 raise_on_fail=False)
    print("reference solve: %d iterations, final %.2e" % (st["iterations"], st["final_res_inf"] / st["rhs_inf"]),
          file=sys.stderr, flush=True)
    del op
    return sol


def oracle_residual(pr, x, b):
    """|b - A x|_inf / |b|_inf with the oracle's FP64 operator (checker only)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle_ctypes as ora
    from alamo_amd import problems as P
    r = ora.residual(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], x, b)
    return float(np.abs(r).max() / np.abs(b).max())


MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 matrix-core peak of MI355X


def layer_report(ms, n, ms_data=None, ms_wgrad=None):
    """per-layer throughput of the network's convolutions at n^3 cells: forward, backward-data and
    weight gradient each 2 27 Cin Cout FLOPs per voxel; bytes (Cin + Cout) 2 per voxel (weight
    gradient: the two inputs), bound min(MFMA peak, HBM x AI)"""
    spec = [(16, 16, 0), (16, 32, 1), (32, 64, 2), (64, 32, 2), (32, 16, 1), (16, 16, 0)]
    out = []
    for l, ((cin, cout, lev), t) in enumerate(zip(spec, ms)):
        V = (n // (2 ** lev) + 1) ** 3
        flops = 2.0 * 27 * cin * cout * V
        byts = (cin + cout) * 2.0 * V
        bound = min(MFMA_BF16_PEAK_TFLOPS, HBM_PEAK_GBS * flops / byts / 1e3)
        rec = {"cin": cin, "cout": cout, "level": lev, "bound_tflops": bound}
        for key, tt in (("fwd", t), ("bwd_data", None if ms_data is None else ms_data[l]),
                        ("wgrad", None if ms_wgrad is None else ms_wgrad[l])):
            if tt is None:
                continue
            tf = flops / (tt * 1e-3) / 1e12
            rec[key] = {"ms": tt, "tflops": tf, "frac": tf / bound}
        out.append(rec)
    return out


def train_network(a, P, Elastic, Net, precision, cycle, smoother, coarse_op):
    """Train the correction network on random two-phase fields of train_n^3 cells (seeds other
    than the benchmark field's, the same interface width), outside the timed region."""
    t0 = time.time()
    net = Net(seed=12345)
    stats = []
    for f in range(a.train_fields):
        pr = P.two_phase_problem(n=a.train_n, corr_cells=a.corr, contrast=a.contrast, seed=a.seed + 1 + f,
                                 width=a.width)
        nsm = a.krylov_smooth
        op = Elastic(pr["ncell"], pr["dx"], sym=P.ISO, bc=(pr["bct"], pr["bcv"]), precision=precision, cycle=cycle,
                     pre_smooth=nsm, post_smooth=nsm, sweeps_per_smooth=a.krylov_sweeps,
                     sweeps_finest=a.krylov_sweeps_finest, smoother=smoother, bottom_solver=a.krylov_bottom,
                     coarse_op=coarse_op)
        op.SetModel(pr["C"])
        stats.append(net.train(op, npairs=a.train_pairs, steps=a.train_steps, lr=1e-3, seed=77 + f,
                               solve_max_iter=0 if a.train_loss == "residual" else 100, solve_tol=1e-6,
                               verbose=a.verbose, loss=a.train_loss))
        del op
    info = {"loss": a.train_loss, "fields": a.train_fields, "cells": a.train_n, "pairs_per_field": a.train_pairs,
            "steps_per_field": a.train_steps, "first_loss": stats[0]["first_loss"],
            "last_loss": stats[-1]["last_loss"], "seconds": time.time() - t0}
    print("network trained: %s" % json.dumps(info), file=sys.stderr, flush=True)
    return net, info


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pr, rhs, a, results):
    """The value method's algorithm (FGMRES(restart) right-preconditioned by the same cycle, l1
    smoother, BiCGStab bottom to 1e-10) on the oracle, FP64, on the host cores:

    * at the full 257^3 size: setup (hierarchy, diagonals) and one iteration timed, scaled to the
      GPU method's iteration count, with every host thread and with one;
    * a full solve of the same field type at 129^3 nodes (same correlation length, width and
      seed), timed end to end with every host thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes
    import numpy as np
    import oracle_ctypes as ora
    from alamo_amd import problems as P
    gomp = ctypes.CDLL("libgomp.so.1")  # the OpenMP runtime the oracle is linked against
    cores = int(gomp.omp_get_max_threads())
    best = min((k for k in results if results[k]["converged"]), key=lambda k: results[k]["seconds"])
    gpu_its = results[best]["iterations"]
    fcyc = not best.endswith("v")
    kw = dict(restart=a.restart, nu1=a.krylov_smooth, nu2=a.krylov_smooth, sweeps=a.krylov_sweeps,
              sweeps_finest=a.krylov_sweeps_finest, smoother=1 if a.krylov_smoother == "l1" else 0,
              cycle=1 if fcyc else 0)
    cyc_name = "%s %s-cycle" % (a.krylov_smoother, "F" if fcyc else "V")

    def timed(p, b, max_iter):
        nc = p["ncell"][0]
        mcl = int(round(math.log2(nc / a.krylov_coarsest))) if a.krylov_coarsest > 0 else -1
        t0 = time.perf_counter()
        x, it, fr = ora.krylov_solve(p["ncell"], p["dx"], P.ISO, p["C"], p["bct"], b, tol_rel=1e-10,
                                     max_iter=max_iter, max_level=mcl, **kw)
        return time.perf_counter() - t0, it, fr

    def extrapolated(nthr):
        t_setup, _, _ = timed(pr, rhs, 0)
        print("cpu baseline: %d thread(s), setup %.1fs" % (nthr, t_setup), file=sys.stderr, flush=True)
        t_one, _, _ = timed(pr, rhs, 1)
        per_it = max(t_one - t_setup, 1e-9)
        print("cpu baseline: %d thread(s), one iteration %.1fs" % (nthr, per_it), file=sys.stderr, flush=True)
        return {"value": t_setup + gpu_its * per_it, "unit": "s", "cores": nthr,
                "sample": "oracle FP64 FGMRES(%d) + %s on the full %d^3-node field, %d thread(s): setup+"
                          "residual (%.2fs) and one iteration (%.2fs) timed, scaled to the GPU method's %d "
                          "iterations" % (a.restart, cyc_name, a.n + 1, nthr, t_setup, per_it, gpu_its)}

    out = extrapolated(cores)
    out.update(kind="port", cpu_model=cpu_model(), method=best,
               note="the oracle runs the value method's algorithm in FP64 (the GPU method: %s)" % results[best][
                   "precision"])
    # a full solve at 129^3 nodes
    nh = a.n // 2
    ph = P.two_phase_problem(n=nh, corr_cells=a.corr, contrast=a.contrast, seed=a.seed, width=a.width)
    Nh = (nh + 1) ** 3
    bh = ora.newton_rhs(ph["ncell"], ph["dx"], P.ISO, ph["C"], ph["F0"], ph["bct"], np.zeros(3 * Nh),
                        ora.bc_init(ph["ncell"], ph["bct"], ph["bcv"]))
    t, it, fr = timed(ph, bh, 1000)
    out["full_solve"] = {"nodes": "%d^3" % (nh + 1), "seconds": t, "iterations": abs(it), "converged": it > 0,
                         "final_rel_res": fr / float(np.abs(bh).max()), "cores": cores}
    print("cpu baseline: %d^3 full solve %.1fs, %d iterations" % (nh + 1, t, abs(it)), file=sys.stderr, flush=True)
    # the reference's own loops run one rank per core without threads: the 257^3 sample on one core
    gomp.omp_set_num_threads(1)
    try:
        out["single_core"] = extrapolated(1)
    finally:
        gomp.omp_set_num_threads(cores)
    return out


if __name__ == "__main__":
    main()
