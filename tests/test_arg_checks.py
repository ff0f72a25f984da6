"""Device-field checks of the Python mirror (CPU only): a host tensor, another dtype, a
non-contiguous view or a short buffer raises AlamoError before anything reaches the library, so no
kernel ever indexes past a caller's buffer.  The handle is a stand-in (no GPU here): a check that
did not fire would reach the library with a null handle and fail with another message."""
import numpy as np
import pytest
import torch

from alamo_amd import problems as P
from alamo_amd._lib import AlamoError, dev_ptr
from alamo_amd.operator import Elastic, Linear, Krylov


def test_dev_ptr_rejects():
    with pytest.raises(AlamoError, match="not on the GPU"):
        dev_ptr(torch.zeros(10, dtype=torch.float64), 10, "x")
    with pytest.raises(AlamoError, match="expected a torch tensor"):
        dev_ptr(np.zeros(10), 10, "x")


def _fake_op(ncell=(4, 4, 4), sym=P.ISO):
    op = object.__new__(Elastic)
    op.h = None
    op.sym = sym
    op.ncell = ncell
    n = int(np.prod([c + 1 for c in ncell]))
    op._size = [n, int(np.prod([c // 2 + 1 for c in ncell]))]
    return op, n


def test_wrappers_reject_host_tensors():
    op, n = _fake_op()
    x = torch.zeros(3 * n, dtype=torch.float64)
    for call in (lambda: op.Fapply(0, 0, x, x), lambda: op.Fsmooth(0, 0, x, x),
                 lambda: op.correctionResidual(0, 0, x, x, x), lambda: op.NewtonRHS(x, x, None, x),
                 lambda: op.BCInit(x), lambda: op.Norms(0, x), lambda: op.SetModel(torch.zeros(2 * n)),
                 lambda: Linear(op).solve(x, x, 1e-10, 0.0), lambda: Krylov.solve(_kry(op), x, x, 1e-10, 0.0)):
        with pytest.raises(AlamoError, match="not on the GPU"):
            call()


def _kry(op):
    k = object.__new__(Krylov)
    k.op = op
    return k


def test_size_and_dtype_checks_fire_first(monkeypatch):
    """with the residency check passed, dtype, contiguity and size are checked in turn"""
    op, n = _fake_op()
    import alamo_amd._lib as L
    monkeypatch.setattr(torch.Tensor, "is_cuda", property(lambda self: True))
    with pytest.raises(AlamoError, match="dtype"):
        op.Fapply(0, 0, torch.zeros(3 * n, dtype=torch.float32), torch.zeros(3 * n, dtype=torch.float32))
    with pytest.raises(AlamoError, match="need at least %d" % (3 * n)):
        op.Fapply(0, 0, torch.zeros(3 * n - 1, dtype=torch.float64), torch.zeros(3 * n, dtype=torch.float64))
    with pytest.raises(AlamoError, match="contiguous"):
        op.Fapply(0, 0, torch.zeros(6 * n, dtype=torch.float64)[::2], torch.zeros(3 * n, dtype=torch.float64))
    # the coarse level of a restriction is sized by its own level
    with pytest.raises(AlamoError, match="restriction crse"):
        op.restriction(0, 1, torch.zeros(3 * op._size[1] - 3, dtype=torch.float64), torch.zeros(3 * n, dtype=torch.float64))
    with pytest.raises(AlamoError, match="mglev 2 out of range"):
        op.Fapply(0, 2, torch.zeros(3 * n, dtype=torch.float64), torch.zeros(3 * n, dtype=torch.float64))
    assert L is not None


def test_grain_table_shapes():
    op, n = _fake_op(sym=P.MAJORMINOR)
    op.LevelNodes = lambda mglev: n
    with pytest.raises(AlamoError, match="grain ids"):
        op.SetModelGrains(np.zeros(n - 1, dtype=np.uint16), np.zeros((2, 21)))
    with pytest.raises(AlamoError, match="table must be"):
        op.SetModelGrains(np.zeros(n, dtype=np.uint16), np.zeros(42))


def test_amr_create_rejects_bad_meshes():
    """alamo_amr_create validates the explicit mesh before touching the device"""
    from alamo_amd.amr import Composite
    n = 16
    bad = [
        [[((3, 4, 4), (11, 11, 11))]],                               # lo odd
        [[((0, 4, 4), (11, 11, 11))]],                               # no coarse cell of margin
        [[((4, 4, 4), (11, 11, 11))], [((8, 8, 8), (23, 23, 23))]],  # level 2 not nested with margin
        [[((4, 4, 4), (11, 11, 11))] * 65],                          # too many boxes
    ]
    for boxes in bad:
        with pytest.raises(AlamoError):
            Composite((n, n, n), (1.0 / n,) * 3, boxes)
