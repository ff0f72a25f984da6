"""The reference's own 3D golden data: tests/Solid (4^3 cells, tension test uniaxial_stress,
cubic C11=168.3, C12=1.221, C44=0.757) checked with its metric, the time-integrated relative
L2 error of trac_xhi_x (tests/Solid/test:26-32), tolerance 1e-4.

Each step of the reference run is a static solve at the ramped boundary value; the model is
linear, so the series is t * trac(disp = 0.01) (+ trac(eigenstrain only) for the affine case).
"""
import os

import numpy as np

from alamo_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        head = f.readline().split()
    raw = np.loadtxt(os.path.join(HERE, "golden", name), skiprows=1)
    return {h: raw[:, i] for i, h in enumerate(head)}


def integrate(x, y):
    return float(np.sum(0.5 * (y[1:] + y[:-1]) * (x[1:] - x[:-1])))


def error(ref, series):
    diff = np.interp(ref["time"], ref["time"], ref["trac_xhi_x"]) - series
    return np.sqrt(integrate(ref["time"], diff ** 2)) / np.sqrt(integrate(ref["time"], ref["trac_xhi_x"] ** 2))


def trac(ora, disp, F0):
    n = 4
    ncell, dx = (n, n, n), (1.0 / n,) * 3
    N = (n + 1) ** 3
    C = np.tile(P.cubic_mm(168.3, 1.221, 0.757), N)
    bct, bcv = P.bc_tension_test("uniaxial_stress", disp=disp)
    F = None if F0 is None else np.tile(F0.ravel(), N)
    b = ora.bc_init(ncell, bct, bcv)
    rhs = ora.newton_rhs(ncell, dx, P.MAJORMINOR, C, F, bct, np.zeros(3 * N), b)
    x, it, *_ = ora.mg_solve(ncell, dx, P.MAJORMINOR, C, bct, rhs, tol_rel=1e-12)
    assert it > 0
    return ora.trac_xhi(ncell, dx, P.MAJORMINOR, C, F, x)[0]


def test_linear_cubic(ora):
    ref = load("solid_linear_cubic.dat")
    t1 = trac(ora, 0.01, None)            # disp ramp (0,1:0,0.01) at t = 1
    assert error(ref, ref["time"] * t1) < 1e-4


def test_affine_cubic(ora):
    ref = load("solid_affine_cubic.dat")
    F0 = 1e-3 * np.eye(3)
    t0 = trac(ora, 0.0, F0)
    t1 = trac(ora, 0.01, None)
    assert error(ref, t0 + ref["time"] * t1) < 1e-4
