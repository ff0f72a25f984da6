"""ctypes view of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg as the checker.  The product package never imports
this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)

# Face enum of BC/Operator/Elastic/Elastic.H:35-43
FACES = ["XLO", "YLO", "ZLO", "XHI", "YHI", "ZHI",
         "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI",
         "ZLO_XLO", "ZLO_XHI", "ZHI_XLO", "ZHI_XHI",
         "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI",
         "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI",
         "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI", "INT"]
FACE = {n: i for i, n in enumerate(FACES)}
DISPLACEMENT, TRACTION, PERIODIC, NEUMANN, NONE = range(5)
ISO, MAJORMINOR, DIAGONAL = range(3)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def ncoef(sym):
    return {ISO: 2, MAJORMINOR: 21, DIAGONAL: 9}[sym]


def nodes(ncell):
    return tuple(int(n) + 1 for n in ncell)


def fapply(ncell, dx, sym, C, bct, u):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); u = _c(u)
    f = np.zeros_like(u)
    lib().ora_fapply(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(u), _d(f))
    return f


def diagonal(ncell, dx, sym, C, bct):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    N = np.prod(nodes(ncell))
    Dg = np.zeros(3 * N)
    lib().ora_diagonal(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(Dg))
    return Dg


def jacobi(ncell, dx, sym, C, bct, Dg, omega, sweeps, x, b):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    Dg = _c(Dg); x = _c(x).copy(); b = _c(b)
    lib().ora_jacobi(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(Dg), ctypes.c_double(omega), sweeps, _d(x), _d(b))
    return x


def l1_diag(ncell, dx, sym, C, bct):
    """l1-Jacobi smoother diagonal: sign(a_ii) sum_j |a_ij| per row"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    N = np.prod(nodes(ncell))
    d = np.zeros(3 * N)
    lib().ora_l1_diag(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(d))
    return d


def jacobi_l1(ncell, dx, sym, C, bct, d, sweeps, x, b):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    d = _c(d); x = _c(x).copy(); b = _c(b)
    lib().ora_jacobi_l1(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(d), sweeps, _d(x), _d(b))
    return x


def l1_diag_rows(ncell, dx, sym, C, bct):
    """the l1 smoother diagonal assembled row by row from the stencils (the CPU baseline's)"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    N = np.prod(nodes(ncell))
    d = np.zeros(3 * N)
    lib().ora_l1_diag_rows(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(d))
    return d


def krylov_solve(ncell, dx, sym, C, bct, b, x0=None, tol_rel=1e-10, tol_abs=0.0, restart=16, max_iter=400, nu1=1,
                 nu2=1, sweeps=3, sweeps_finest=1, smoother=1, omega=2.0 / 3.0, bottom_tol=1e-10, cycle=0,
                 max_level=-1):
    """FGMRES(restart) with one V-cycle (cycle 0) or F-cycle (cycle 1) as right preconditioner, over
    levels 0 .. max_level (-1: coarsen to 2 cells); returns (x, iterations, final |r|_inf)"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); b = _c(b)
    x = np.zeros_like(b) if x0 is None else _c(x0).copy()
    fr = ctypes.c_double(0)
    it = lib().ora_krylov_solve(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(x), _d(b), ctypes.c_double(tol_rel),
                                ctypes.c_double(tol_abs), restart, max_iter, nu1, nu2, sweeps, sweeps_finest, smoother,
                                ctypes.c_double(omega), ctypes.c_double(bottom_tol), ctypes.byref(fr), int(cycle),
                                int(max_level))
    return x, it, fr.value


def galerkin(ncell, dx, sym, C, bct, level):
    """Galerkin stencil R A P of multigrid level `level` (243 planes: 9 o + 3 q + p)"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32)
    nc = tuple(int(n) // (2 ** level) for n in ncell)
    out = np.zeros(243 * int(np.prod([n + 1 for n in nc])))
    lib().ora_galerkin(_i(ncell), _d(dx), sym, _d(C), _i(bct), level, _d(out))
    return out


def stencil_apply(ncell, dx, S, x):
    ncell = _c(ncell, np.int32); dx = _c(dx); S = _c(S); x = _c(x)
    y = np.zeros_like(x)
    lib().ora_stencil_apply(_i(ncell), _d(dx), _d(S), _d(x), _d(y))
    return y


def residual(ncell, dx, sym, C, bct, x, b):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); x = _c(x); b = _c(b)
    r = np.zeros_like(b)
    lib().ora_residual(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(x), _d(b), _d(r))
    return r


def restrict(ncell_fine, dx_fine, fine, w=1, planes=3):
    ncell_fine = _c(ncell_fine, np.int32); dx_fine = _c(dx_fine); fine = _c(fine)
    nc = tuple(int(n) // 2 + 1 for n in ncell_fine)
    N = int(np.prod(nc))
    out = np.zeros((planes if planes > 0 else 1) * N * w)
    lib().ora_restrict(_i(ncell_fine), _d(dx_fine), w, planes, _d(fine), _d(out))
    return out


def interp_add(ncell_fine, dx_fine, crse, fine):
    ncell_fine = _c(ncell_fine, np.int32); dx_fine = _c(dx_fine); crse = _c(crse); fine = _c(fine).copy()
    lib().ora_interp_add(_i(ncell_fine), _d(dx_fine), _d(crse), _d(fine))
    return fine


def mg_solve(ncell, dx, sym, C, bct, b, x0=None, tol_rel=1e-10, tol_abs=0.0, cycle=1, nu1=2, nu2=2, sweeps=2,
             omega=2.0 / 3.0, max_iter=200, max_coarsening=-1):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); b = _c(b)
    x = np.zeros_like(b) if x0 is None else _c(x0).copy()
    hist = np.zeros(max_iter + 1)
    fr = ctypes.c_double(0)
    nl = ctypes.c_int(0)
    it = lib().ora_mg_solve(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(x), _d(b), ctypes.c_double(tol_rel),
                            ctypes.c_double(tol_abs), cycle, nu1, nu2, sweeps, ctypes.c_double(omega), max_iter,
                            max_coarsening, _d(hist), max_iter + 1, ctypes.byref(fr), ctypes.byref(nl))
    return x, it, hist[:abs(it) + 1], fr.value, nl.value


def coeff_level(ncell, dx, sym, C, level):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C)
    nc = tuple(int(n) // (2 ** level) + 1 for n in ncell)
    out = np.zeros(int(np.prod(nc)) * ncoef(sym))
    lib().ora_coeff_level(_i(ncell), _d(dx), sym, _d(C), level, _d(out))
    return out


def newton_rhs(ncell, dx, sym, C, F0, bct, u, b):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); u = _c(u); b = _c(b)
    rhs = np.zeros_like(b)
    f0p = None if F0 is None else _c(F0)
    lib().ora_newton_rhs(_i(ncell), _d(dx), sym, _d(C), None if f0p is None else _d(f0p), _i(bct), _d(u), _d(b),
                         _d(rhs))
    return rhs


# psi variants (Operator/Elastic.cpp psi terms): psi is the level-0 cell field, prod(ncell) values
def fapply_psi(ncell, dx, sym, C, bct, psi, psi_small, u):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); u = _c(u); psi = _c(psi)
    f = np.zeros_like(u)
    lib().ora_fapply_psi(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(psi), ctypes.c_double(psi_small), _d(u), _d(f))
    return f


def diagonal_psi(ncell, dx, sym, C, bct, psi, psi_small):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); psi = _c(psi)
    Dg = np.zeros(3 * int(np.prod(nodes(ncell))))
    lib().ora_diagonal_psi(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(psi), ctypes.c_double(psi_small), _d(Dg))
    return Dg


def newton_rhs_psi(ncell, dx, sym, C, F0, bct, psi, u, b):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); u = _c(u); b = _c(b)
    psi = _c(psi)
    rhs = np.zeros_like(b)
    f0p = None if F0 is None else _c(F0)
    lib().ora_newton_rhs_psi(_i(ncell), _d(dx), sym, _d(C), None if f0p is None else _d(f0p), _i(bct), _d(psi),
                             _d(u), _d(b), _d(rhs))
    return rhs


def psi_level(ncell, dx, psi, level):
    ncell = _c(ncell, np.int32); dx = _c(dx); psi = _c(psi)
    nc = [int(n) // (2 ** level) for n in ncell]
    out = np.zeros(int(np.prod(nc)))
    lib().ora_psi_level(_i(ncell), _d(dx), _d(psi), level, _d(out))
    return out


def mg_solve_psi(ncell, dx, sym, C, bct, psi, psi_small, b, tol_rel=1e-10, tol_abs=0.0, cycle=1, nu1=2, nu2=2,
                 sweeps=2, omega=2.0 / 3.0, max_iter=200):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); b = _c(b); psi = _c(psi)
    x = np.zeros_like(b)
    fr = ctypes.c_double(0)
    it = lib().ora_mg_solve_psi(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(psi), ctypes.c_double(psi_small), _d(x),
                                _d(b), ctypes.c_double(tol_rel), ctypes.c_double(tol_abs), cycle, nu1, nu2, sweeps,
                                ctypes.c_double(omega), max_iter, ctypes.byref(fr))
    return x, it, fr.value


def post(ncell, dx, sym, C, u, kind, voigt=False, psi=None, psi_small=1e-8, energy=None):
    """kind 0 strain, 1 stress, 2 energy (accumulates into `energy`)"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); u = _c(u)
    N = int(np.prod(nodes(ncell)))
    if kind == 2:
        out = np.zeros(N) if energy is None else _c(energy).copy()
    else:
        out = np.zeros((6 if voigt else 9) * N)
    ps = None if psi is None else _c(psi)
    lib().ora_post(_i(ncell), _d(dx), sym, _d(C), None if ps is None else _d(ps), ctypes.c_double(psi_small), _d(u),
                   kind, int(voigt), _d(out))
    return out


def fapply_uniform(ncell, dx, sym, C, bct, u):
    """Fapply with Elastic::SetUniform(true): interior rows without grad(C) . grad(u)"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); bct = _c(bct, np.int32); u = _c(u)
    f = np.zeros_like(u)
    lib().ora_fapply_uniform(_i(ncell), _d(dx), sym, _d(C), _i(bct), _d(u), _d(f))
    return f


def model_dw(ncell, dx, sym, C, F0, u):
    """Newton::DW of an affine model: sig = C:(grad u - F0), 9 row-major planes"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); u = _c(u)
    out = np.zeros(9 * int(np.prod(nodes(ncell))))
    f0p = None if F0 is None else _c(F0)
    lib().ora_model_post(_i(ncell), _d(dx), sym, _d(C), None if f0p is None else _d(f0p), _d(u), 3, _d(out))
    return out


def model_w(ncell, dx, sym, C, F0, u):
    """Newton::W of an affine model: w = (grad u - F0):C:(grad u - F0) / 2 per node"""
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); u = _c(u)
    out = np.zeros(int(np.prod(nodes(ncell))))
    f0p = None if F0 is None else _c(F0)
    lib().ora_model_post(_i(ncell), _d(dx), sym, _d(C), None if f0p is None else _d(f0p), _d(u), 4, _d(out))
    return out


def blend(eta, table, aos=True):
    eta = _c(eta); table = _c(table)
    nmodels, nrec = table.shape
    nnodes = eta.size // nmodels
    out = np.zeros(nnodes * nrec)
    lib().ora_blend(ctypes.c_int64(nnodes), nmodels, nrec, _d(eta), _d(table), _d(out), int(aos))
    return out


def bc_init(ncell, bct, bcv, homogeneous=False, b=None):
    ncell = _c(ncell, np.int32); bct = _c(bct, np.int32); bcv = _c(bcv)
    N = int(np.prod(nodes(ncell)))
    b = np.zeros(3 * N) if b is None else _c(b).copy()
    lib().ora_bc_init(_i(ncell), _i(bct), _d(bcv), int(homogeneous), _d(b))
    return b


def trac_xhi(ncell, dx, sym, C, F0, u):
    ncell = _c(ncell, np.int32); dx = _c(dx); C = _c(C); u = _c(u)
    out = np.zeros(3)
    f0p = None if F0 is None else _c(F0)
    lib().ora_trac_xhi(_i(ncell), _d(dx), sym, _d(C), None if f0p is None else _d(f0p), _d(u), _d(out))
    return out


def cubic(C11, C12, C44, phi1=0.0, Phi=0.0, phi2=0.0):
    out = np.zeros(21)
    lib().ora_cubic(ctypes.c_double(C11), ctypes.c_double(C12), ctypes.c_double(C44), ctypes.c_double(phi1),
                    ctypes.c_double(Phi), ctypes.c_double(phi2), _d(out))
    return out


def euler_R(phi1, Phi, phi2):
    R = np.zeros(9)
    lib().ora_euler_R(ctypes.c_double(phi1), ctypes.c_double(Phi), ctypes.c_double(phi2), _d(R))
    return R.reshape(3, 3)


def mm_slot(i, j, k, l):
    return lib().ora_mm_slot(i, j, k, l)


def c4(sym, c, i, j, k, l):
    lib().ora_c4.restype = ctypes.c_double
    c = _c(c)
    return lib().ora_c4(sym, _d(c), i, j, k, l)


def classify(i, j, k, lo, hi):
    lo = _c(lo, np.int32); hi = _c(hi, np.int32)
    return lib().ora_classify(i, j, k, _i(lo), _i(hi))


def normal(face):
    n = np.zeros(3)
    lib().ora_normal(face, _d(n))
    return n


# ---------------------------------------------------------------------------
# learned correction network (oracle/conv_oracle.cpp)
# ---------------------------------------------------------------------------
NET_LAYERS = [(16, 16, 1), (16, 32, 1), (32, 64, 1), (64, 32, 0), (32, 16, 1), (16, 16, 0)]


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def net_nparam():
    f = lib().ocv_nparam
    f.restype = ctypes.c_int64
    return int(f())


def bf16(x):
    f = lib().ocv_bf16
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    return np.vectorize(lambda v: f(float(v)))(np.asarray(x, dtype=np.float64))


def conv3(N, cin, cout, x, W, bias=None):
    """out[v][co] = bias + sum_{o,ci} W[o][co][ci] x[v+off(o)][ci]; x [V][cin]"""
    N = _c(N, np.int32)
    x = _c(x)
    W = _c(W)
    out = np.zeros((int(np.prod(N)), cout))
    b = _c(bias) if bias is not None else None
    lib().ocv_conv3(_i(N), cin, cout, _d(x), _d(W), _d(b) if b is not None else None, _d(out))
    return out


def conv3_T(N, cin, cout, dout, W):
    N = _c(N, np.int32)
    dout = _c(dout)
    W = _c(W)
    out = np.zeros((int(np.prod(N)), cin))
    lib().ocv_conv3_T(_i(N), cin, cout, _d(dout), _d(W), _d(out))
    return out


def conv3_wgrad(N, cin, cout, x, dout):
    N = _c(N, np.int32)
    x = _c(x)
    dout = _c(dout)
    dW = np.zeros((27, cout, cin))
    dB = np.zeros(cout)
    lib().ocv_conv3_wgrad(_i(N), cin, cout, _d(x), _d(dout), _d(dW), _d(dB))
    return dW, dB


def _coarse(N):
    return np.array([(n - 1) // 2 + 1 for n in N], dtype=np.int32)


def net_prolong(Nf, C, c):
    Nf = _c(Nf, np.int32)
    c = _c(c)
    out = np.zeros((int(np.prod(Nf)), C))
    lib().ocv_prolong(_i(Nf), C, _d(c), _d(out))
    return out


def net_prolong_T(Nf, C, gu):
    Nf = _c(Nf, np.int32)
    gu = _c(gu)
    out = np.zeros((int(np.prod(_coarse(Nf))), C))
    lib().ocv_prolong_T(_i(Nf), C, _d(gu), _d(out))
    return out


def net_inject(Nf, C, a):
    Nf = _c(Nf, np.int32)
    a = _c(a)
    out = np.zeros((int(np.prod(_coarse(Nf))), C))
    lib().ocv_inject(_i(Nf), C, _d(a), _d(out))
    return out


def net_inject_T(Nf, C, gd):
    Nf = _c(Nf, np.int32)
    gd = _c(gd)
    out = np.zeros((int(np.prod(Nf)), C))
    lib().ocv_inject_T(_i(Nf), C, _d(gd), _d(out))
    return out


def net_forward(N0, params, x, bf16=True):
    """y [V][16] of the network with float32 params (library layout) on input x [V][16]"""
    N0 = _c(N0, np.int32)
    p = _c(params, np.float32)
    x = _c(x, np.float32)
    y = np.zeros((int(np.prod(N0)), 16))
    lib().ocv_net_forward(_i(N0), _f(p), _f(x), _d(y), int(bf16))
    return y


def net_backward(N0, params, x, dy, bf16=True, want_dx=True):
    """(grads [nparam], dx [V][16]) of <y(x), dy>"""
    N0 = _c(N0, np.int32)
    p = _c(params, np.float32)
    x = _c(x, np.float32)
    dy = _c(dy, np.float32)
    g = np.zeros(net_nparam())
    dx = np.zeros((int(np.prod(N0)), 16)) if want_dx else None
    lib().ocv_net_backward(_i(N0), _f(p), _f(x), _f(dy), _d(g), _d(dx) if want_dx else None, int(bf16))
    return g, dx


def net_layer_slices():
    """[(W slice, W shape, bias slice)] of each layer in the parameter vector"""
    out, off = [], 0
    for cin, cout, _ in NET_LAYERS:
        nw = 27 * cin * cout
        out.append((slice(off, off + nw), (27, cout, cin), slice(off + nw, off + nw + cout)))
        off += nw + cout
    return out


# ---------------------------------------------------------------------------
# two-level AMR helpers (oracle/amr_oracle.cpp); fine boxes are inclusive cell boxes in fine
# index space, fine fields live on the fine nodal box grown by one node
# ---------------------------------------------------------------------------
def _boxes(fine_lo, fine_hi):
    lo = _c(np.atleast_2d(fine_lo), np.int32)
    hi = _c(np.atleast_2d(fine_hi), np.int32)
    return lo, hi


def amr_fine_mask(ncell_crse, fine_lo, fine_hi):
    nc = _c(ncell_crse, np.int32)
    lo, hi = _boxes(fine_lo, fine_hi)
    mask = np.zeros(int(np.prod(nc + 1)), dtype=np.int32)
    lib().oamr_fine_mask(_i(nc), len(lo), _i(lo), _i(hi), _i(mask))
    return mask


def amr_grown_fine_nodes(flo, fhi):
    return [int(fhi[d] - flo[d] + 4) for d in range(3)]


def amr_reflux(ncell_crse, flo, fhi, mask, crse, fine):
    nc = _c(ncell_crse, np.int32)
    flo = _c(flo, np.int32); fhi = _c(fhi, np.int32)
    out = _c(crse).copy()
    lib().oamr_reflux(_i(nc), _i(flo), _i(fhi), _i(_c(mask, np.int32)), _d(out), _d(_c(fine)))
    return out


def amr_average_down(ncell_crse, flo, fhi, crse, fine):
    nc = _c(ncell_crse, np.int32)
    flo = _c(flo, np.int32); fhi = _c(fhi, np.int32)
    out = _c(crse).copy()
    lib().oamr_average_down(_i(nc), _i(flo), _i(fhi), _d(out), _d(_c(fine)))
    return out


def amr_fixup_residual_mask(mask, resmsk):
    out = _c(resmsk, np.int32).copy()
    f = lib().oamr_fixup_residual_mask
    f.argtypes = [ctypes.c_int64, I, I]
    f(len(out), _i(_c(mask, np.int32)), _i(out))
    return out


# ---------------------------------------------------------------------------
# Composite multi-level (AMR) solve: levels l = 0 .. nl-1 at n_cell 2^l, boxes (per level >= 1, a
# list of (lo, hi) inclusive cell boxes in the level's own index space); C, u, b: lists of per-level
# arrays (C AoS ncoef per node, u / b 3 SoA planes)
# ---------------------------------------------------------------------------
def _amr_boxes(boxes):
    nbox = np.array([len(bl) for bl in boxes], dtype=np.int32)
    lo = np.array([b[0] for bl in boxes for b in bl], dtype=np.int32).reshape(-1)
    hi = np.array([b[1] for bl in boxes for b in bl], dtype=np.int32).reshape(-1)
    return nbox, lo, hi


def amr_solve(ncell0, dx0, sym, Cs, bct, boxes, bs, us=None, tol_rel=1e-10, tol_abs=0.0, cycle=0, nu1=2, nu2=2,
              sweeps=2, smoother=0, omega=2.0 / 3.0, max_iter=400):
    nl = len(Cs)
    ncell0 = _c(ncell0, np.int32); dx0 = _c(dx0); bct = _c(bct, np.int32)
    nbox, lo, hi = _amr_boxes(boxes)
    C = _c(np.concatenate([np.asarray(c, dtype=np.float64).ravel() for c in Cs]))
    b = _c(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in bs]))
    u = np.zeros_like(b) if us is None else _c(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in us])).copy()
    fr = ctypes.c_double(0)
    it = lib().ora_amr_solve(_i(ncell0), _d(dx0), sym, _d(C), _i(bct), nl, _i(nbox), _i(lo), _i(hi), _d(u), _d(b),
                             ctypes.c_double(tol_rel), ctypes.c_double(tol_abs), cycle, nu1, nu2, sweeps, smoother,
                             ctypes.c_double(omega), max_iter, ctypes.byref(fr))
    sizes = [3 * int(np.prod([(int(n) << l) + 1 for n in ncell0])) for l in range(nl)]
    return np.split(u, np.cumsum(sizes)[:-1]), it, fr.value


def amr_residual(ncell0, dx0, sym, Cs, bct, boxes, us, bs):
    nl = len(Cs)
    ncell0 = _c(ncell0, np.int32); dx0 = _c(dx0); bct = _c(bct, np.int32)
    nbox, lo, hi = _amr_boxes(boxes)
    C = _c(np.concatenate([np.asarray(c, dtype=np.float64).ravel() for c in Cs]))
    u = _c(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in us]))
    b = _c(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in bs]))
    r = np.zeros_like(b)
    lib().ora_amr_residual(_i(ncell0), _d(dx0), sym, _d(C), _i(bct), nl, _i(nbox), _i(lo), _i(hi), _d(u), _d(b), _d(r))
    sizes = [3 * int(np.prod([(int(n) << l) + 1 for n in ncell0])) for l in range(nl)]
    return np.split(r, np.cumsum(sizes)[:-1])
