"""Why the two-phase 1:1000 field is solved with an erf interface of a few cells: on a thresholded
field (width 0) the strong-form nodal operator (Elastic.cpp:312-493: C:grad grad u + grad C . grad u
by central differences) is indefinite -- D^-1 A has eigenvalues with negative real part, so no
Jacobi-type smoother or multigrid cycle built on it converges -- while an interface of a little
over one cell (10-90 width) already gives a spectrum in the right half-plane.  Dense eigenvalues of
the oracle's operator on 9^3 nodes."""
import numpy as np
import pytest

from alamo_amd import problems as P


def _spectrum(ora, width):
    n = 8
    pr = P.two_phase_problem(n=n, corr_cells=4, contrast=1000.0, seed=3, width=width)
    N = 3 * (n + 1) ** 3
    A = np.empty((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[j] = 1.0
        A[:, j] = ora.fapply(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], e)
        e[j] = 0.0
    return pr["interface"], np.linalg.eigvals(A / np.diag(A)[:, None])


def test_thresholded_field_is_indefinite(ora):
    itf, ev = _spectrum(ora, 0.0)
    assert itf["pure_phase_frac"] == 1.0
    assert (ev.real < 0).sum() >= 20 and ev.real.min() < -10.0


def test_one_cell_interface_is_positive(ora):
    itf, ev = _spectrum(ora, 0.5)
    assert 1.0 <= itf["interface_cells_10_90"] <= 1.6
    assert ev.real.min() > 0.0
