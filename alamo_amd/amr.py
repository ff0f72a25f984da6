"""AMR over the C ABI: the two-level helpers of ``Operator<Grid::Node>`` (src/Operator/Operator.cpp)
and the composite multi-level solve (``Composite``).

For the helpers, fine boxes are inclusive cell boxes in fine index space (refinement ratio 2, lo even, hi odd);
fine fields live on the box's nodal box grown by one node, ``[lo - 1, hi + 2]^3``, 3 SoA planes.
Fields are torch tensors on the GPU (memory only); masks are int32 tensors of the coarse nodes.
"""
import ctypes

import numpy as np

from ._lib import check, dev_ptr, lib
from .problems import ISO

CRSE_NODE, CRSE_FINE_NODE, FINE_NODE = 0, 1, 2  # amrex::nodelap_detail


def _iv(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
    return a, a.ctypes.data_as(ctypes.c_void_p)


def _nodes(n):
    return int(np.prod([int(v) + 1 for v in np.asarray(n).reshape(3)]))


def grown_fine_nodes(lo, hi):
    """nodes per direction of a fine box's grown nodal box"""
    return tuple(int(hi[d]) - int(lo[d]) + 4 for d in range(3))


def fine_mask(ncell_crse, boxes_lo, boxes_hi, mask, stream=None):
    """m_nd_fine_mask (MLNodeLinOp::buildMasks, read at Operator.cpp:236,553)"""
    nc, pnc = _iv(ncell_crse)
    lo, plo = _iv(boxes_lo)
    hi, phi = _iv(boxes_hi)
    check(lib().alamo_amr_fine_mask(pnc, len(lo) // 3, plo, phi, dev_ptr(mask, _nodes(nc), "fine_mask mask", "int32"),
                                    stream))


def reflux(ncell_crse, lo, hi, mask, crse_res, fine_res, stream=None):
    """Operator<Grid::Node>::reflux (Operator.cpp:526-633) for one fine box"""
    nc, pnc = _iv(ncell_crse)
    l, pl = _iv(lo)
    h, ph = _iv(hi)
    nf = int(np.prod(grown_fine_nodes(l, h)))
    check(lib().alamo_amr_reflux(pnc, pl, ph, dev_ptr(mask, _nodes(nc), "reflux mask", "int32"),
                                 dev_ptr(crse_res, 3 * _nodes(nc), "reflux crse"), dev_ptr(fine_res, 3 * nf, "reflux fine"),
                                 stream))


def average_down(ncell_crse, lo, hi, crse_sol, fine_sol, stream=None):
    """Operator<Grid::Node>::averageDownSolutionRHS (Operator.cpp:435-447)"""
    nc, pnc = _iv(ncell_crse)
    l, pl = _iv(lo)
    h, ph = _iv(hi)
    nf = int(np.prod(grown_fine_nodes(l, h)))
    check(lib().alamo_amr_average_down(pnc, pl, ph, dev_ptr(crse_sol, 3 * _nodes(nc), "average_down crse"),
                                       dev_ptr(fine_sol, 3 * nf, "average_down fine"), stream))


def fixup_residual_mask(ncell_crse, mask, resmsk, stream=None):
    """Operator<Grid::Node>::fixUpResidualMask (Operator.cpp:230-251)"""
    nc, pnc = _iv(ncell_crse)
    check(lib().alamo_amr_fixup_residual_mask(pnc, dev_ptr(mask, _nodes(nc), "fixup mask", "int32"),
                                              dev_ptr(resmsk, _nodes(nc), "fixup resmsk", "int32"), stream))


class Composite:
    """Solver::Nonlocal::Linear::solve over several AMR levels (Linear.H:131-168 with amrex::MLMG's
    composite cycle) on an explicit mesh (Integrator.cpp:121-139: explicitmesh.lo<l> / hi<l>).

    ``boxes[l - 1]`` lists the inclusive cell boxes ``(lo, hi)`` of AMR level l >= 1 in that level's
    own index space (refinement ratio 2).  Every level is a full-domain field of 3 SoA planes of
    ``(n_cell 2^l + 1)^3`` nodes; ``Level(l)`` is level l's ``Elastic`` operator (SetModel, BCInit,
    Fapply, Stress ... on it), ``solve`` runs the composite V-cycles.  Options are ``Elastic``'s."""

    def __init__(self, ncell, dx, boxes, sym=ISO, bc=None, **opts):
        from .operator import Elastic, make_desc
        self.desc = make_desc(ncell, dx, sym, bc, **opts)
        self.nlevels = len(boxes) + 1
        nbox = np.array([len(bl) for bl in boxes], dtype=np.int32)
        lo = np.ascontiguousarray(np.array([b[0] for bl in boxes for b in bl], dtype=np.int32).reshape(-1))
        hi = np.ascontiguousarray(np.array([b[1] for bl in boxes for b in bl], dtype=np.int32).reshape(-1))
        if lo.size == 0:
            lo = hi = np.zeros(3, dtype=np.int32)
        self.boxes = [[(tuple(b[0]), tuple(b[1])) for b in bl] for bl in boxes]
        h = ctypes.c_void_p()
        check(lib().alamo_amr_create(ctypes.byref(self.desc), self.nlevels, nbox.ctypes.data_as(ctypes.c_void_p),
                                     lo.ctypes.data_as(ctypes.c_void_p), hi.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.byref(h)))
        self.h = h
        self.levels = []
        for l in range(self.nlevels):
            hl = ctypes.c_void_p()
            check(lib().alamo_amr_level(h, l, ctypes.byref(hl)))
            self.levels.append(Elastic._view(hl, self, [int(n) << l for n in ncell], [float(x) / (1 << l) for x in dx],
                                             sym, self.desc))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().alamo_amr_destroy(h)
            except Exception:
                pass
            self.h = None

    def Level(self, amrlev):
        return self.levels[amrlev]

    def Nodes(self, amrlev):
        return _nodes(self.levels[amrlev].ncell)

    def Masks(self, amrlev, region=None, fine_mask=None, stream=None):
        """the level's region mask (2 inside, 1 on the boundary, 0 outside) and its coverage by the
        next level (the nodal fine mask: 2 fine, 1 coarse/fine, 0 coarse), int32 device tensors"""
        n = self.Nodes(amrlev)
        check(lib().alamo_amr_masks(self.h, amrlev,
                                    None if region is None else dev_ptr(region, n, "Masks region", "int32"),
                                    None if fine_mask is None else dev_ptr(fine_mask, n, "Masks fine_mask", "int32"),
                                    stream))

    def solve(self, sols, rhss, tol_rel, tol_abs, raise_on_fail=True):
        """composite solve: sols[l] initial guess in, solution out; rhss[l] right-hand sides (BCInit
        values in level 0's boundary nodes).  Returns (final |r|_inf, SolveStats)."""
        from ._lib import SolveStats, AlamoError
        if len(sols) != self.nlevels or len(rhss) != self.nlevels:
            raise AlamoError(1, "need one solution and one right-hand side per AMR level")
        ps = (ctypes.c_void_p * self.nlevels)(*[dev_ptr(sols[l], 3 * self.Nodes(l), "solve sol[%d]" % l)
                                                for l in range(self.nlevels)])
        pb = (ctypes.c_void_p * self.nlevels)(*[dev_ptr(rhss[l], 3 * self.Nodes(l), "solve rhs[%d]" % l)
                                                for l in range(self.nlevels)])
        st = SolveStats()
        rc = lib().alamo_amr_solve(self.h, ps, pb, float(tol_rel), float(tol_abs), ctypes.byref(st))
        if raise_on_fail:
            check(rc)
        return st.final_res_inf, st


def write_plotfile(path, ncell0, prob_lo, prob_hi, boxes, names, fields, time=0.0, steps=None):
    """AMReX plotfile of nodal fields (Integrator::WritePlotFile's "<prefix>node" output,
    Integrator.cpp:808-1054): ``fields[l]`` holds ``len(names)`` SoA planes of level l's
    ``(ncell0 2^l + 1)^3`` nodes, either all numpy arrays (host) or all device tensors;
    ``boxes[l - 1]`` are level l's inclusive cell boxes ``(lo, hi)`` (level 0: the whole domain)."""
    nl = len(fields)
    if len(boxes) != nl - 1:
        raise ValueError("boxes must list the boxes of every level >= 1")
    nbox = np.array([0] + [len(bl) for bl in boxes], dtype=np.int32)
    lo = np.array([b[0] for bl in boxes for b in bl] or [(0, 0, 0)], dtype=np.int32).reshape(-1)
    hi = np.array([b[1] for bl in boxes for b in bl] or [(0, 0, 0)], dtype=np.int32).reshape(-1)
    nm = [s.encode() for s in names]
    cnames = (ctypes.c_char_p * len(nm))(*nm)
    host = all(isinstance(f, np.ndarray) for f in fields)
    keep = []
    if host:
        for f in fields:
            keep.append(np.ascontiguousarray(f, dtype=np.float64))
        ptrs = (ctypes.c_void_p * nl)(*[k.ctypes.data for k in keep])
    else:
        for f in fields:
            if not (f.is_cuda and f.is_contiguous() and str(f.dtype) == "torch.float64"):
                raise ValueError("device fields must be contiguous float64 CUDA tensors")
        ptrs = (ctypes.c_void_p * nl)(*[f.data_ptr() for f in fields])
    for l, f in enumerate(keep if host else fields):
        need = len(names) * int(np.prod([(int(n) << l) + 1 for n in ncell0]))
        if (f.size if host else f.numel()) < need:
            raise ValueError("field of level %d holds fewer than %d values" % (l, need))
    st = np.asarray(steps if steps is not None else [0] * nl, dtype=np.int32)
    fn = lib().alamo_write_plotfile_host if host else lib().alamo_write_plotfile
    c = lambda a, t: np.ascontiguousarray(a, dtype=t)
    nc, pl, ph = c(ncell0, np.int32), c(prob_lo, np.float64), c(prob_hi, np.float64)
    check(fn(str(path).encode(), nl, nc.ctypes.data_as(ctypes.c_void_p), pl.ctypes.data_as(ctypes.c_void_p),
             ph.ctypes.data_as(ctypes.c_void_p), nbox.ctypes.data_as(ctypes.c_void_p),
             lo.ctypes.data_as(ctypes.c_void_p), hi.ctypes.data_as(ctypes.c_void_p), len(names), cnames, ptrs,
             float(time), st.ctypes.data_as(ctypes.c_void_p)))
