"""CPU checks that pin the oracle of the learned correction network (oracle/conv_oracle.cpp).

The network has no counterpart in the reference, so nothing here is a golden vector; each
oracle function is pinned by an independent route instead:
  * the convolution against scipy.ndimage.correlate (a library routine),
  * the backward-data and weight-gradient passes by the adjoint identities that define them,
  * the prolongation against the multigrid interpolation of the operator oracle (itself pinned
    against the reference's nodal interpolation) and against exact reproduction of linear fields,
  * the whole backward pass by finite differences of the forward pass.
"""
import numpy as np
import pytest
import scipy.ndimage

import oracle_ctypes as ora


def _rng(seed):
    return np.random.default_rng(seed)


def test_conv3_matches_scipy_correlate():
    rng = _rng(1)
    N = (7, 5, 6)
    V = int(np.prod(N))
    cin, cout = 3, 2
    x = rng.standard_normal((V, cin))
    W = rng.standard_normal((27, cout, cin))
    b = rng.standard_normal(cout)
    got = ora.conv3(N, cin, cout, x, W, b)
    # arrays indexed [z][y][x]; offset o = (dz+1)*9 + (dy+1)*3 + (dx+1)
    for co in range(cout):
        ref = np.full((N[2], N[1], N[0]), b[co])
        for ci in range(cin):
            K = W[:, co, ci].reshape(3, 3, 3)
            ref += scipy.ndimage.correlate(x[:, ci].reshape(N[2], N[1], N[0]), K, mode="constant", cval=0.0)
        np.testing.assert_allclose(got[:, co], ref.ravel(), rtol=1e-12, atol=1e-12)


def test_conv3_adjoint_and_weight_gradient():
    rng = _rng(2)
    N = (6, 7, 5)
    V = int(np.prod(N))
    cin, cout = 4, 3
    x = rng.standard_normal((V, cin))
    g = rng.standard_normal((V, cout))
    W = rng.standard_normal((27, cout, cin))
    b = rng.standard_normal(cout)
    y = ora.conv3(N, cin, cout, x, W)
    # <conv(x), g> = <x, conv^T(g)>
    assert np.isclose(np.sum(y * g), np.sum(x * ora.conv3_T(N, cin, cout, g, W)), rtol=1e-12)
    # <conv(x; W, b), g> is linear in (W, b) with gradient (dW, dB)
    dW, dB = ora.conv3_wgrad(N, cin, cout, x, g)
    yb = ora.conv3(N, cin, cout, x, W, b)
    assert np.isclose(np.sum(yb * g), np.sum(W * dW) + np.sum(b * dB), rtol=1e-12)


def test_prolong_matches_multigrid_interpolation():
    rng = _rng(3)
    ncell_f = np.array([8, 12, 4])
    Nf = ncell_f + 1
    Nc = ncell_f // 2 + 1
    c = rng.standard_normal((int(np.prod(Nc)), 3))
    got = ora.net_prolong(Nf, 3, c)
    ref = ora.interp_add(ncell_f, np.ones(3), c.T.copy(), np.zeros((3, int(np.prod(Nf)))))
    np.testing.assert_allclose(got, ref.reshape(3, -1).T, rtol=0, atol=1e-14)


def test_prolong_reproduces_linear_fields():
    Nf = np.array([9, 5, 13])
    Nc = (Nf - 1) // 2 + 1
    Z, Y, X = np.meshgrid(np.arange(Nc[2]), np.arange(Nc[1]), np.arange(Nc[0]), indexing="ij")
    lin = lambda x, y, z: 0.3 + 1.5 * x - 0.7 * y + 2.0 * z
    c = lin(2.0 * X, 2.0 * Y, 2.0 * Z).reshape(-1, 1)
    z, y, x = np.meshgrid(np.arange(Nf[2]), np.arange(Nf[1]), np.arange(Nf[0]), indexing="ij")
    np.testing.assert_allclose(ora.net_prolong(Nf, 1, c)[:, 0], lin(x, y, z).ravel(), atol=1e-12)


def test_prolong_and_injection_adjoints():
    rng = _rng(4)
    Nf = np.array([9, 7, 5])
    Vf = int(np.prod(Nf))
    Vc = int(np.prod((Nf - 1) // 2 + 1))
    C = 5
    c = rng.standard_normal((Vc, C))
    u = rng.standard_normal((Vf, C))
    assert np.isclose(np.sum(ora.net_prolong(Nf, C, c) * u), np.sum(c * ora.net_prolong_T(Nf, C, u)), rtol=1e-12)
    assert np.isclose(np.sum(ora.net_inject(Nf, C, u) * c), np.sum(u * ora.net_inject_T(Nf, C, c)), rtol=1e-12)
    # injection picks the even nodes
    a = np.arange(Vf * C, dtype=float).reshape(Nf[2], Nf[1], Nf[0], C)
    np.testing.assert_array_equal(ora.net_inject(Nf, C, a.reshape(Vf, C)), a[::2, ::2, ::2].reshape(Vc, C))


def _params(seed, scale=0.3):
    rng = _rng(seed)
    return (scale * rng.standard_normal(ora.net_nparam())).astype(np.float32)


@pytest.mark.parametrize("N0", [(5, 5, 5), (9, 5, 9)])
def test_net_backward_finite_differences(N0):
    rng = _rng(5)
    V = int(np.prod(N0))
    p = _params(6)
    x = rng.standard_normal((V, 16)).astype(np.float32)
    dy = rng.standard_normal((V, 16)).astype(np.float32)
    loss = lambda pp, xx: float(np.sum(ora.net_forward(N0, pp, xx, bf16=False) * dy.astype(np.float64)))
    g, dx = ora.net_backward(N0, p, x, dy, bf16=False)
    # directional derivatives along random directions restricted to each layer's weights and
    # biases; the applied step is measured in float32 so the comparison is exact to O(eps^2)
    # as long as no ReLU changes state, which a step this small does not do here (the 1e-9
    # absolute term is the FP64 rounding of losses of order 1e3)
    eps = 1e-6
    for ws, _, bs in ora.net_layer_slices():
        for sl in (ws, bs):
            d = np.zeros_like(p)
            d[sl] = rng.standard_normal(sl.stop - sl.start).astype(np.float32)
            pp, pm = (p + eps * d).astype(np.float32), (p - eps * d).astype(np.float32)
            fd = loss(pp, x) - loss(pm, x)
            an = float(np.dot(g, pp.astype(np.float64) - pm.astype(np.float64)))
            assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, (sl, fd, an)
    d = rng.standard_normal(x.shape).astype(np.float32)
    xp, xm = (x + eps * d).astype(np.float32), (x - eps * d).astype(np.float32)
    fd = loss(p, xp) - loss(p, xm)
    an = float(np.sum(dx * (xp.astype(np.float64) - xm.astype(np.float64))))
    assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, (fd, an)


def test_net_bf16_mode_is_close_to_fp64():
    rng = _rng(7)
    N0 = (9, 9, 5)
    V = int(np.prod(N0))
    p = _params(8, 0.1)
    x = rng.standard_normal((V, 16)).astype(np.float32)
    y64 = ora.net_forward(N0, p, x, bf16=False)
    y16 = ora.net_forward(N0, p, x, bf16=True)
    assert np.linalg.norm(y16 - y64) <= 3e-2 * np.linalg.norm(y64)
    # bf16 rounding: round to nearest even on the upper 16 bits
    assert ora.bf16(1.0 + 2.0 ** -8) == 1.0
    assert ora.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert ora.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
