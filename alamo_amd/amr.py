"""Two-level AMR helpers of ``Operator<Grid::Node>`` (src/Operator/Operator.cpp) over the C ABI.

Fine boxes are inclusive cell boxes in fine index space (refinement ratio 2, lo even, hi odd);
fine fields live on the box's nodal box grown by one node, ``[lo - 1, hi + 2]^3``, 3 SoA planes.
Fields are torch tensors on the GPU (memory only); masks are int32 tensors of the coarse nodes.
"""
import ctypes

import numpy as np

from ._lib import check, dev_ptr, lib

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
