"""Why the two-phase 1:1000 field is solved with an erf interface of a few cells.

On a thresholded field (width 0) the strong-form nodal operator (Elastic.cpp:204-287:
C:grad grad u + grad C . grad u by central differences) is indefinite -- D^-1 A has eigenvalues
with negative real part, so no Jacobi-type smoother or multigrid cycle built on it converges --
while an interface of a little over one cell (10-90 width) already gives a spectrum in the right
half-plane.  Beyond the spectrum, the discrete solution itself is not the elastic one for
interfaces of about two cells and less: the central difference of C across the jump makes the
grad C . grad u term dominate the rows next to the interface, and the solution of the discrete
system carries strains hundreds of times the eigenstrain that loads it (a stiff inclusion
expanding freely would strain by the eigenstrain itself), with a condition number three orders
of magnitude above that of a diffuse field.  Dense operators of the oracle on 9^3 nodes."""
import functools

import numpy as np
import pytest

from alamo_amd import problems as P

LOAD = 1e-3  # the eigenstrain of two_phase_problem


@functools.lru_cache(maxsize=None)
def _dense_cached(width, n=8, corr=4, seed=3):
    import oracle_ctypes as ora
    pr = P.two_phase_problem(n=n, corr_cells=corr, contrast=1000.0, seed=seed, width=width)
    N = 3 * (n + 1) ** 3
    A = np.empty((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[j] = 1.0
        A[:, j] = ora.fapply(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["bct"], e)
        e[j] = 0.0
    return pr, A


def _dense(ora, width):
    return _dense_cached(width)


def _spectrum(ora, width):
    pr, A = _dense(ora, width)
    return pr["interface"], np.linalg.eigvals(A / np.diag(A)[:, None])


def test_thresholded_field_is_indefinite(ora):
    itf, ev = _spectrum(ora, 0.0)
    assert itf["pure_phase_frac"] == 1.0
    assert (ev.real < 0).sum() >= 20 and ev.real.min() < -10.0


def test_one_cell_interface_is_positive(ora):
    itf, ev = _spectrum(ora, 0.5)
    assert 1.0 <= itf["interface_cells_10_90"] <= 1.6
    assert ev.real.min() > 0.0


@functools.lru_cache(maxsize=None)
def _solution_cached(width):
    import oracle_ctypes as ora
    return _solution(ora, width)


def _solution(ora, width):
    pr, A = _dense(ora, width)
    n = pr["ncell"][0]
    N = A.shape[0]
    b = ora.newton_rhs(pr["ncell"], pr["dx"], P.ISO, pr["C"], pr["F0"], pr["bct"], np.zeros(N),
                       ora.bc_init(pr["ncell"], pr["bct"], pr["bcv"]))
    u = np.linalg.solve(A, b)
    assert np.abs(A @ u - b).max() <= 1e-9 * np.abs(b).max()
    U = u.reshape(3, n + 1, n + 1, n + 1)  # [component][k][j][i]
    strain = max(np.abs(np.diff(U[c], axis=2 - c)).max() * n for c in range(3))  # max |d_c u_c|
    return pr["interface"], strain, np.linalg.cond(A, 1)


@pytest.mark.parametrize("width", [0.0, 0.25])
def test_sharp_interface_solution_is_not_elastic(ora, width):
    # thresholded and a ~0.7-cell interface against a ~4-cell one on the same random field
    itf, strain, cond = _solution_cached(width)
    itf_d, strain_d, cond_d = _solution_cached(1.5)
    assert itf_d["interface_cells_10_90"] >= 3.5
    print("width", width, itf, "max strain %.3e cond %.2e; diffuse: max strain %.3e cond %.2e"
          % (strain, cond, strain_d, cond_d))
    assert strain_d <= 10 * LOAD          # a diffuse interface: strains of the eigenstrain's order
    assert strain >= 100 * LOAD           # sharp: hundreds of times the load
    assert cond >= 100 * cond_d
