"""Why the dense bottom ends the traction-free polycrystal's hierarchy at 4 cells: with traction-free
lateral faces the re-discretised operator of a 2-cell level (the reference's coefficient
restriction, Elastic.cpp:746-836, then the same rows) is singular -- the one-sided traction rows of
opposite faces meet -- while the 4-cell level is regular, and with Neumann sides the 2-cell level
is regular too.  Dense singular values of the oracle's level operators."""
import numpy as np

from alamo_amd import problems as P


def _cond(ora, lateral, level, n=8):
    pr = P.polycrystal_problem(n=n, ngrains=8, seed=3, lateral=lateral)
    C = pr["table"][pr["gid"].ravel()].ravel()
    Cl = ora.coeff_level(pr["ncell"], pr["dx"], P.MAJORMINOR, C, level)
    nc = [n >> level] * 3
    N = 3 * (nc[0] + 1) ** 3
    A = np.empty((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[j] = 1.0
        A[:, j] = ora.fapply(nc, [1.0 / nc[0]] * 3, P.MAJORMINOR, Cl, pr["bct"], e)
        e[j] = 0.0
    s = np.linalg.svd(A, compute_uv=False)
    return s[0] / s[-1]


def test_two_cell_level_singular_with_traction_free_sides(ora):
    assert _cond(ora, P.TRACTION, 2) > 1e12
    assert _cond(ora, P.TRACTION, 1) < 1e5
    assert _cond(ora, P.NEUMANN, 2) < 1e5
