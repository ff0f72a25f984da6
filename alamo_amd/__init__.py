"""MI355X-native nodal elastic multigrid (Operator::Elastic / Solver::Nonlocal::Linear).

All arithmetic runs in libalamo_amd.so (hand-written HIP for gfx950); this package is
the host-side mirror of the reference's operator/solver interface.
"""
from . import problems  # noqa: F401
from ._lib import AlamoError, lib  # noqa: F401

__all__ = ["problems", "AlamoError", "lib"]
