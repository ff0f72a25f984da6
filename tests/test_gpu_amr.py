"""GPU parity of the two-level AMR helpers against the oracle: masks, average-down and the
residual-mask fix-up bit-exact; reflux bit-exact too (same stencil, same summation order),
checked at 1e-15 of the field scale."""
import numpy as np
import pytest

from alamo_amd import amr
from alamo_amd._lib import AlamoError
from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu

NC = np.array([16, 12, 20])
BOXES = [((4, 2, 8), (15, 13, 27)), ((16, 2, 8), (23, 9, 19)), ((0, 0, 0), (7, 7, 7)), ((24, 16, 32), (31, 23, 39))]


def _int_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)


def _mask(ora, cuda):
    import torch
    lo = [b[0] for b in BOXES]
    hi = [b[1] for b in BOXES]
    m = torch.zeros(int(np.prod(NC + 1)), dtype=torch.int32, device=cuda)
    amr.fine_mask(NC, lo, hi, m)
    torch.cuda.synchronize()
    return m, ora.amr_fine_mask(NC, lo, hi)


def test_fine_mask(cuda, ora):
    print("amr fine mask", flush=True)
    m, ref = _mask(ora, cuda)
    np.testing.assert_array_equal(m.cpu().numpy(), ref)


@pytest.mark.parametrize("b", range(len(BOXES)))
def test_reflux_and_average_down(cuda, ora, b):
    lo, hi = BOXES[b]
    m, mref = _mask(ora, cuda)
    g = ora.amr_grown_fine_nodes(lo, hi)
    r = np.random.default_rng(b)
    fine = r.uniform(-1, 1, 3 * int(np.prod(g)))
    crse = r.uniform(-1, 1, 3 * int(np.prod(NC + 1)))
    print("amr reflux box", b, flush=True)
    c = to_dev(crse, cuda)
    amr.reflux(NC, lo, hi, m, c, to_dev(fine, cuda))
    ref = ora.amr_reflux(NC, lo, hi, mref, crse, fine)
    got = to_host(c)
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()
    c = to_dev(crse, cuda)
    amr.average_down(NC, lo, hi, c, to_dev(fine, cuda))
    np.testing.assert_array_equal(to_host(c), ora.amr_average_down(NC, lo, hi, crse, fine))


def test_fixup_residual_mask(cuda, ora):
    m, mref = _mask(ora, cuda)
    res = np.random.default_rng(9).integers(0, 2, mref.size).astype(np.int32)
    rd = _int_dev(res, cuda)
    amr.fixup_residual_mask(NC, m, rd)
    np.testing.assert_array_equal(rd.cpu().numpy(), ora.amr_fixup_residual_mask(mref, res))


def test_rejects_misaligned_boxes(cuda):
    import torch
    m = torch.zeros(int(np.prod(NC + 1)), dtype=torch.int32, device=cuda)
    with pytest.raises(AlamoError):
        amr.fine_mask(NC, [(1, 0, 0)], [(8, 7, 7)], m)  # odd lo
    with pytest.raises(AlamoError):
        amr.fine_mask(NC, [(0, 0, 0)], [(33, 7, 7)], m)  # beyond the refined domain
