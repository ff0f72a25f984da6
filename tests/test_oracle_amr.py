"""CPU checks that pin the oracle's two-level AMR helpers (oracle/amr_oracle.cpp).

The node mask is AMReX's (absent here): checked against its definition from cell coverage,
evaluated independently with numpy, and against integer identities.  Reflux with one fine box
covering the whole domain must equal the single-level restriction (pinned in
test_oracle_mg.py), and reflux reproduces constants and leaves uncovered nodes untouched.
"""
import numpy as np
import pytest

import oracle_ctypes as ora

NC = np.array([16, 12, 20])
BOXES = [((4, 2, 8), (15, 13, 27)), ((16, 2, 8), (23, 9, 19)), ((0, 0, 0), (7, 7, 7)), ((24, 16, 32), (31, 23, 39))]


def _mask_numpy(nc, boxes):
    cov = np.zeros((nc[2], nc[1], nc[0]), dtype=bool)
    for lo, hi in boxes:
        cov[lo[2] // 2:(hi[2] + 1) // 2, lo[1] // 2:(hi[1] + 1) // 2, lo[0] // 2:(hi[0] + 1) // 2] = True
    c = np.pad(cov, 1, mode="edge")  # cells outside mirror the boundary cell
    n = np.zeros((nc[2] + 1, nc[1] + 1, nc[0] + 1), dtype=int)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                n += c[dz:dz + nc[2] + 1, dy:dy + nc[1] + 1, dx:dx + nc[0] + 1]
    return np.where(n == 8, 2, np.where(n > 0, 1, 0)).ravel()


@pytest.mark.parametrize("nb", [0, 1, 2, 4])
def test_fine_mask_definition(nb):
    boxes = BOXES[:nb]
    lo = [b[0] for b in boxes] or np.zeros((0, 3))
    hi = [b[1] for b in boxes] or np.zeros((0, 3))
    m = ora.amr_fine_mask(NC, lo, hi)
    np.testing.assert_array_equal(m, _mask_numpy(NC, boxes))
    if nb == 0:
        assert not m.any()


def test_fine_mask_single_box_counts():
    lo, hi = (4, 2, 8), (15, 13, 27)  # coarsened cells [2,7] x [1,6] x [4,13]
    m = ora.amr_fine_mask(NC, [lo], [hi]).reshape(NC[2] + 1, NC[1] + 1, NC[0] + 1)
    e = np.array([6, 6, 10])  # coarse cells of the box
    assert (m == 2).sum() == np.prod(e - 1)                   # strictly inside the nodal box
    assert (m >= 1).sum() == np.prod(e + 1)                   # the box's nodal box
    assert (m == 1).sum() == np.prod(e + 1) - np.prod(e - 1)  # its surface
    full = ora.amr_fine_mask(NC, [(0, 0, 0)], [tuple(2 * NC - 1)])
    assert (full == 2).all()  # mirrored outside cells: the domain boundary is fine too


def _fine_field(lo, hi, seed):
    g = ora.amr_grown_fine_nodes(lo, hi)
    return np.random.default_rng(seed).uniform(-1, 1, 3 * int(np.prod(g))), g


def test_reflux_whole_domain_equals_restriction():
    lo, hi = (0, 0, 0), tuple(2 * NC - 1)
    fine, g = _fine_field(lo, hi, 1)
    mask = ora.amr_fine_mask(NC, [lo], [hi])
    crse = np.random.default_rng(2).uniform(-1, 1, 3 * int(np.prod(NC + 1)))
    got = ora.amr_reflux(NC, lo, hi, mask, crse, fine)
    # valid fine nodes: drop the ghost layer of the grown box
    f = fine.reshape(3, g[2], g[1], g[0])[:, 1:-1, 1:-1, 1:-1].reshape(3, -1)
    ref = ora.restrict(2 * NC, np.ones(3), f.ravel())
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15)


def test_reflux_constants_and_untouched_nodes():
    lo, hi = BOXES[0]
    mask = ora.amr_fine_mask(NC, [b[0] for b in BOXES], [b[1] for b in BOXES])
    g = ora.amr_grown_fine_nodes(lo, hi)
    fine = np.concatenate([np.full(int(np.prod(g)), c) for c in (1.5, -2.0, 0.25)])
    crse = np.random.default_rng(3).uniform(-1, 1, 3 * int(np.prod(NC + 1)))
    got = ora.amr_reflux(NC, lo, hi, mask, crse, fine).reshape(3, NC[2] + 1, NC[1] + 1, NC[0] + 1)
    c0 = crse.reshape(got.shape)
    sl = (slice(lo[2] // 2, (hi[2] + 1) // 2 + 1), slice(lo[1] // 2, (hi[1] + 1) // 2 + 1),
          slice(lo[0] // 2, (hi[0] + 1) // 2 + 1))
    for n, c in enumerate((1.5, -2.0, 0.25)):
        assert np.all(got[n][sl] == c)  # stencil weights sum to 1 (exact in binary)
        out = np.ones(got.shape[1:], dtype=bool)
        out[sl] = False
        assert np.array_equal(got[n][out], c0[n][out])


def test_average_down_is_injection():
    lo, hi = BOXES[3]
    fine, g = _fine_field(lo, hi, 4)
    crse = np.zeros(3 * int(np.prod(NC + 1)))
    got = ora.amr_average_down(NC, lo, hi, crse, fine).reshape(3, NC[2] + 1, NC[1] + 1, NC[0] + 1)
    F = fine.reshape(3, g[2], g[1], g[0])
    sl = tuple(slice(lo[d] // 2, (hi[d] + 1) // 2 + 1) for d in (2, 1, 0))
    np.testing.assert_array_equal(got[(slice(None),) + sl], F[:, 1:-1:2, 1:-1:2, 1:-1:2])
    assert np.count_nonzero(got) == np.count_nonzero(F[:, 1:-1:2, 1:-1:2, 1:-1:2])


def test_fixup_residual_mask():
    mask = ora.amr_fine_mask(NC, [b[0] for b in BOXES], [b[1] for b in BOXES])
    res = np.random.default_rng(5).integers(0, 2, mask.size).astype(np.int32)
    got = ora.amr_fixup_residual_mask(mask, res)
    np.testing.assert_array_equal(got, np.where(mask == 1, 1, res))
