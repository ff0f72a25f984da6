"""Host-side mirror of the reference's operator/solver plugin surface over the C ABI.

Names, argument meaning and error behaviour follow

* ``Operator::Elastic<SYM>``       (src/Operator/Elastic.H:21-190, Elastic.cpp)
* ``Operator::Operator<Grid::Node>`` (src/Operator/Operator.H:27-125)
* ``Solver::Nonlocal::Linear``      (src/Solver/Nonlocal/Linear.H:18-389)

Device fields are torch tensors (float64, on the GPU) used only as memory: every
computation runs in libalamo_amd.so.  A reference ``Util::Abort`` becomes an
``AlamoError`` exception.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import KrylovOpts, MGDesc, SolveStats, check, dev_ptr as _dev, lib
from .problems import ISO, MAJORMINOR, DIAGONAL, NCOEF

LAYOUT_AOS, LAYOUT_SOA = 0, 1
PRECOND_NONE, PRECOND_MG, PRECOND_LEARNED = 0, 1, 2
LOSS_ERROR, LOSS_RESIDUAL = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
CYCLE_V, CYCLE_F = 0, 1


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


BOTTOM_SOLVERS = {"bicgstab": 0, "smoother": 1, "dense": 2}
SMOOTHERS = {"jacobi": 0, "l1": 1}
COARSE_OPS = {"restrict": 0, "galerkin": 1}


def coarse_op_code(name):
    """operators of the coarse levels: "restrict" (re-discretised from restricted coefficients, the
    reference, Elastic.cpp:746-836) or "galerkin" (R A P; no reference counterpart)."""
    if isinstance(name, int):
        if name not in COARSE_OPS.values():
            raise ValueError("unknown coarse_op %d" % name)
        return name
    if name not in COARSE_OPS:
        raise ValueError("coarse_op must be one of %s, not %r" % (sorted(COARSE_OPS), name))
    return COARSE_OPS[name]


def smoother_code(name):
    """smoother of the cycles: "jacobi" (the reference's damped point Jacobi, Operator.cpp:109-146) or
    "l1" (l1-Jacobi, x += (b - A x) / d with d the row's absolute sum; no reference counterpart)."""
    if isinstance(name, int):
        if name not in SMOOTHERS.values():
            raise ValueError("unknown smoother %d" % name)
        return name
    if name not in SMOOTHERS:
        raise ValueError("smoother must be one of %s, not %r" % (sorted(SMOOTHERS), name))
    return SMOOTHERS[name]


def bottom_solver_code(name):
    """solver.bottom_solver value (Linear.H:280-282) -> ALAMO_BOTTOM_*; the reference's "cg" is not
    built (the operator is not symmetric; the reference itself forces bicgstab, Linear.H:258)."""
    if isinstance(name, int):
        if name not in BOTTOM_SOLVERS.values():
            raise ValueError("unknown bottom_solver %d" % name)
        return name
    if name not in BOTTOM_SOLVERS:
        raise ValueError("bottom_solver must be one of %s, not %r" % (sorted(BOTTOM_SOLVERS), name))
    return BOTTOM_SOLVERS[name]


def make_desc(ncell, dx, sym=ISO, bc=None, precision=PREC_FP64, cycle=CYCLE_F, pre_smooth=2, post_smooth=2,
              sweeps_per_smooth=2, omega=2.0 / 3.0, max_iter=200, bottom_max_iter=200, bottom_tol_rel=1e-4,
              max_coarsening_level=-1, verbose=0, smoother="jacobi", fixed_iter=0, bottom_solver="bicgstab",
              sweeps_finest=0, coarse_op="restrict"):
    """alamo_mg_desc from the operator's options (defaults: Linear.H:183-244 and Elastic.H)"""
    d = MGDesc()
    check(lib().alamo_mg_desc_default(ctypes.byref(d)))
    for q in range(3):
        d.n_cell[q] = int(ncell[q])
        d.dx[q] = float(dx[q])
    d.sym = int(sym)
    if bc is not None:
        t, v = bc
        t = np.asarray(t, dtype=np.int32).reshape(26, 3)
        v = np.asarray(v, dtype=np.float64).reshape(26, 3)
        for f in range(26):
            for c in range(3):
                d.bc_type[f][c] = int(t[f, c])
                d.bc_val[f][c] = float(v[f, c])
    d.precision, d.cycle = precision, cycle
    d.pre_smooth, d.post_smooth, d.sweeps_per_smooth = pre_smooth, post_smooth, sweeps_per_smooth
    d.omega, d.max_iter, d.bottom_max_iter, d.bottom_tol_rel = omega, max_iter, bottom_max_iter, bottom_tol_rel
    d.max_coarsening_level, d.verbose = max_coarsening_level, verbose
    d.smoother = smoother_code(smoother)
    d.coarse_op = coarse_op_code(coarse_op)
    d.fixed_iter = int(fixed_iter)
    d.bottom_solver = bottom_solver_code(bottom_solver)
    d.sweeps_finest = int(sweeps_finest)
    return d


class Elastic:
    """Operator::Elastic<SYM> on one box covering a non-periodic nodal domain.

    ``bc`` is the pair (bc_type[26][3], bc_val[26][3]) of BC::Operator::Elastic::Constant."""

    def __init__(self, ncell, dx, sym=ISO, bc=None, **opts):
        d = make_desc(ncell, dx, sym, bc, **opts)
        self.desc = d
        self.sym = sym
        self.ncell = tuple(int(n) for n in ncell)
        self.dx = tuple(float(x) for x in dx)
        h = ctypes.c_void_p()
        check(lib().alamo_mg_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self._size = [self.LevelLayout(l)[2] for l in range(self.NMGLevels())]

    @classmethod
    def _view(cls, h, owner, ncell, dx, sym, desc):
        """an operator on a handle that `owner` holds and destroys (an AMR level of Composite)"""
        self = cls.__new__(cls)
        self.desc, self.sym, self._owner = desc, sym, owner
        self.ncell = tuple(int(n) for n in ncell)
        self.dx = tuple(float(x) for x in dx)
        self.h = h
        self._size = [self.LevelLayout(l)[2] for l in range(self.NMGLevels())]
        return self

    def _vec(self, t, mglev, what, ncomp=3):
        """checked pointer of a device field of ncomp planes on level mglev"""
        if not 0 <= mglev < len(self._size):
            raise _lib.AlamoError(1, "mglev %d out of range" % mglev)
        return _dev(t, ncomp * self._size[mglev], what)

    def __del__(self):
        h = getattr(self, "h", None)
        if getattr(self, "_owner", None) is not None:
            return
        if h is not None and h.value:
            try:
                lib().alamo_mg_destroy(h)
            except Exception:
                pass
            self.h = None

    # --- hierarchy -------------------------------------------------------------
    def NMGLevels(self):
        n = ctypes.c_int()
        check(lib().alamo_mg_num_levels(self.h, ctypes.byref(n)))
        return n.value

    def LevelNCell(self, mglev):
        a = (ctypes.c_int * 3)()
        check(lib().alamo_mg_level_ncell(self.h, mglev, a))
        return tuple(a)

    def LevelNodes(self, mglev):
        n = self.LevelNCell(mglev)
        return (n[0] + 1) * (n[1] + 1) * (n[2] + 1)

    def UpdateOpts(self, **kw):
        """Change solver options of the built operator (cycle, pre_smooth, post_smooth,
        sweeps_per_smooth, omega, max_iter, bottom_max_iter, bottom_tol_rel, fixed_iter, verbose,
        bottom_solver, max_fmg_iter, sweeps_finest, smoother, coarse_op)."""
        allowed = {"cycle", "pre_smooth", "post_smooth", "sweeps_per_smooth", "omega", "max_iter",
                   "bottom_max_iter", "bottom_tol_rel", "fixed_iter", "verbose", "bottom_solver", "max_fmg_iter",
                   "sweeps_finest", "smoother", "coarse_op"}
        for k, v in kw.items():
            if k not in allowed:
                raise ValueError("not a solver option: " + k)
            if k == "bottom_solver":
                v = bottom_solver_code(v)
            if k == "smoother":
                v = smoother_code(v)
            if k == "coarse_op":
                v = coarse_op_code(v)
            setattr(self.desc, k, v)
        check(lib().alamo_mg_update_opts(self.h, ctypes.byref(self.desc)))

    def SetOmega(self, omega):
        """Operator::SetOmega (Operator.H:57)"""
        self.UpdateOpts(omega=float(omega))

    def LevelLayout(self, mglev=0):
        """(row pitch, plane pitch, elements per component) of a device field on level mglev."""
        a = (ctypes.c_int64 * 3)()
        check(lib().alamo_mg_level_layout(self.h, mglev, a))
        return tuple(int(v) for v in a)

    def LevelSize(self, mglev=0):
        """elements per component of a device field on level mglev (its node count)"""
        return self.LevelLayout(mglev)[2]

    # --- coefficients ----------------------------------------------------------
    def SetModel(self, C, layout=LAYOUT_AOS, amrlev=0):
        """Elastic::SetModel(amrlev, MultiTab) (Elastic.cpp:87-116); C as numpy (host) or tensor (device)."""
        if isinstance(C, np.ndarray):
            C = np.ascontiguousarray(C, dtype=np.float64)
            expect = NCOEF[self.sym] * self.LevelNodes(0)
            if C.size != expect:
                raise _lib.AlamoError(1, "Inconsistent # of components: got %d values, need %d" % (C.size, expect))
            check(lib().alamo_mg_set_model_host(self.h, amrlev, C.ctypes.data_as(ctypes.c_void_p), layout))
        else:
            check(lib().alamo_mg_set_model(self.h, amrlev, self._vec(C, 0, "SetModel C", NCOEF[self.sym]), layout))

    def SetModelGrains(self, gid, table):
        """Per-grain MajorMinor model: gid one grain id per level-0 node (i fastest), table (ngrains, 21)."""
        gid = np.ascontiguousarray(gid, dtype=np.uint16).ravel()
        table = np.ascontiguousarray(table, dtype=np.float64)
        if gid.size != self.LevelNodes(0):
            raise _lib.AlamoError(1, "SetModelGrains: %d grain ids for %d nodes" % (gid.size, self.LevelNodes(0)))
        if table.ndim != 2 or table.shape[1] != 21 or table.shape[0] < 1:
            raise _lib.AlamoError(1, "SetModelGrains: table must be (ngrains, 21), got %s" % (table.shape,))
        check(lib().alamo_mg_set_model_grains_host(self.h, gid.ctypes.data_as(ctypes.c_void_p),
                                                   table.ctypes.data_as(ctypes.c_void_p), table.shape[0]))

    def SetPsi(self, psi, psi_small=1e-8, amrlev=0):
        """Elastic::SetPsi(amrlev, psi, psi_small) (Elastic.H:55-57, Elastic.cpp:118-137): the cell
        field psi (prod(n_cell) values, i fastest), numpy (host) or tensor (device)."""
        if isinstance(psi, np.ndarray):
            psi = np.ascontiguousarray(psi, dtype=np.float64).ravel()
            if psi.size != int(np.prod(self.ncell)):
                raise _lib.AlamoError(1, "psi needs one value per cell")
            check(lib().alamo_mg_set_psi_host(self.h, amrlev, psi.ctypes.data_as(ctypes.c_void_p), float(psi_small)))
        else:
            check(lib().alamo_mg_set_psi(self.h, amrlev, _dev(psi, int(np.prod(self.ncell)), "SetPsi psi"),
                                         float(psi_small)))

    def prepareForSolve(self):
        check(lib().alamo_mg_prepare(self.h))

    def GetCoeff(self, mglev, out):
        check(lib().alamo_mg_get_coeff(self.h, mglev, self._vec(out, mglev, "GetCoeff out", NCOEF[self.sym])))

    # --- operator virtuals -----------------------------------------------------
    def Fapply(self, amrlev, mglev, out, inp, stream=None):
        check(lib().alamo_mg_fapply(self.h, amrlev, mglev, self._vec(out, mglev, "Fapply out"),
                                     self._vec(inp, mglev, "Fapply in"), stream))

    def Apply(self, amrlev, mglev, out, inp):
        self.Fapply(amrlev, mglev, out, inp)

    def Diagonal(self, amrlev, mglev, diag, stream=None):
        check(lib().alamo_mg_diagonal(self.h, amrlev, mglev, self._vec(diag, mglev, "Diagonal"), stream))

    def normalize(self, amrlev, mglev, x, stream=None):
        """Operator<Grid::Node>::normalize (Operator.cpp:156-165): x /= D"""
        check(lib().alamo_mg_normalize(self.h, amrlev, mglev, self._vec(x, mglev, "normalize x"), stream))

    def Fsmooth(self, amrlev, mglev, x, b, stream=None):
        check(lib().alamo_mg_fsmooth(self.h, amrlev, mglev, self._vec(x, mglev, "Fsmooth x"),
                                      self._vec(b, mglev, "Fsmooth b"), stream))

    def FsmoothNoAlloc(self, amrlev, mglev, x, b, stream=None):
        """alamo_mg_fsmooth_noalloc: Fsmooth through the library's static device scratch (no
        allocation; the handle must be prepared; levels of at most 8388608 / 3 nodes)."""
        check(lib().alamo_mg_fsmooth_noalloc(self.h, amrlev, mglev, self._vec(x, mglev, "Fsmooth x"),
                                             self._vec(b, mglev, "Fsmooth b"), stream))

    def correctionResidual(self, amrlev, mglev, resid, x, b, stream=None):
        check(lib().alamo_mg_residual(self.h, amrlev, mglev, self._vec(resid, mglev, "residual r"),
                                       self._vec(x, mglev, "residual x"), self._vec(b, mglev, "residual b"), stream))

    def solutionResidual(self, amrlev, resid, x, b, stream=None):
        self.correctionResidual(amrlev, 0, resid, x, b, stream)

    def restriction(self, amrlev, cmglev, crse, fine, stream=None):
        check(lib().alamo_mg_restrict(self.h, amrlev, cmglev, self._vec(crse, cmglev, "restriction crse"),
                                       self._vec(fine, cmglev - 1, "restriction fine"), stream))

    def interpolation(self, amrlev, fmglev, fine, crse, stream=None):
        check(lib().alamo_mg_interpolate(self.h, amrlev, fmglev, self._vec(fine, fmglev, "interpolation fine"),
                                          self._vec(crse, fmglev + 1, "interpolation crse"), stream))

    def Norms(self, mglev, x, stream=None):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().alamo_mg_norms(self.h, mglev, self._vec(x, mglev, "Norms x"), ctypes.byref(a), ctypes.byref(b), stream))
        return a.value, b.value

    # --- post-processing (Elastic.cpp:430-603) -----------------------------------
    def Strain(self, amrlev, eps, u, voigt=False, stream=None):
        """eps = sym(grad u): 9 planes row-major, or 6 Voigt planes 00, 11, 22, 12, 20, 01"""
        check(lib().alamo_mg_strain(self.h, amrlev, self._vec(u, 0, "Strain u"),
                                     self._vec(eps, 0, "Strain eps", 6 if voigt else 9), int(voigt), stream))

    def Stress(self, amrlev, sigma, u, voigt=False, stream=None):
        """sigma = (C : grad u) psi_avg"""
        check(lib().alamo_mg_stress(self.h, amrlev, self._vec(u, 0, "Stress u"),
                                     self._vec(sigma, 0, "Stress sigma", 6 if voigt else 9), int(voigt), stream))

    def Energy(self, amrlev, energy, u, stream=None):
        """energy += (C : grad u) : sym(grad u) / 2"""
        check(lib().alamo_mg_energy(self.h, amrlev, self._vec(u, 0, "Energy u"), self._vec(energy, 0, "Energy", 1),
                                     stream))

    def TractionXhi(self, u, F0=None):
        """Base::Mechanics::Integrate trac_hi[0] (Mechanics.H:391-448): the x-hi face integral of
        sigma e_x with sigma = C : (grad u - F0); returns a 3-vector"""
        out = (ctypes.c_double * 3)()
        check(lib().alamo_mg_trac_xhi(self.h, self._vec(u, 0, "TractionXhi u"),
                                       None if F0 is None else self._vec(F0, 0, "TractionXhi F0", 9), out))
        return np.array(list(out))

    # --- right-hand sides ------------------------------------------------------
    def BCInit(self, rhs, homogeneous=False):
        """BC::Operator::Elastic::Constant::Init (Constant.H:197-286)."""
        check(lib().alamo_bc_init(self.h, self._vec(rhs, 0, "BCInit rhs"), int(homogeneous)))

    def NewtonRHS(self, u, b, F0, rhs):
        """Newton<T>::prepareForSolve residual form (Newton.H:144-231)."""
        check(lib().alamo_newton_rhs(self.h, self._vec(u, 0, "NewtonRHS u"), self._vec(b, 0, "NewtonRHS b"),
                                      None if F0 is None else self._vec(F0, 0, "NewtonRHS F0", 9),
                                      self._vec(rhs, 0, "NewtonRHS rhs")))

    # --- flags and queries of the operator (Elastic.H:48,104-107; Operator.H:57-59,96-101) -----
    def SetUniform(self, uniform):
        """Elastic::SetUniform: True drops grad(C) . grad(u) from the interior rows (Elastic.cpp:245)."""
        check(lib().alamo_mg_set_uniform(self.h, int(bool(uniform))))
        self.uniform = bool(uniform)

    def SetHomogeneous(self, homogeneous):
        """Elastic::SetHomogeneous: stored; the reference never reads m_homogeneous (Elastic.H:48,172)."""
        self.homogeneous = bool(homogeneous)

    def SetTesting(self, testing):
        """Elastic::SetTesting: stored; m_testing is not read by the operator (Elastic.H:104,170)."""
        self.testing = bool(testing)

    def SetAverageDownCoeffs(self, flag):
        """Elastic::SetAverageDownCoeffs: the coarse-AMR-level overwrite of Elastic.cpp:611; with the
        one AMR level of this handle it has nothing to overwrite."""
        self.average_down_coeffs = bool(flag)

    def SetNormalizeDDW(self, flag):
        """Operator::SetNormalizeDDW: stored; m_normalize_ddw is not read (Operator.H:59,124)."""
        self.normalize_ddw = bool(flag)

    def isSingular(self, amrlev):
        """Operator<Node>::isSingular (Operator.H:96): m_is_bottom_singular, always false."""
        return False

    def isBottomSingular(self):
        return False

    def getNGrow(self, amrlev=0, mglev=0):
        """Operator<Node>::getNGrow (Operator.H:101)"""
        return 2

    def getNComp(self):
        return 3

    def Reflux(self, fine_lo, fine_hi, mask, crse_res, fine_res, stream=None):
        """Operator<Grid::Node>::reflux (Operator.cpp:526-633) of one fine box onto this handle's
        level-0 residual (compact layout); mask from alamo_amd.amr.fine_mask."""
        lo = np.ascontiguousarray(np.asarray(fine_lo, dtype=np.int32).reshape(3))
        hi = np.ascontiguousarray(np.asarray(fine_hi, dtype=np.int32).reshape(3))
        nf = int(np.prod(hi - lo + 4))  # nodes of the fine box's grown nodal box [lo - 1, hi + 2]
        nm = int(np.prod([n + 1 for n in self.ncell]))
        check(lib().alamo_mg_reflux(self.h, lo.ctypes.data_as(ctypes.c_void_p), hi.ctypes.data_as(ctypes.c_void_p),
                                    _dev(mask, nm, "Reflux mask", "int32"), self._vec(crse_res, 0, "Reflux crse"),
                                    _dev(fine_res, 3 * nf, "Reflux fine"), _ptr(stream)))

    def NormInf(self, mglev, x, stream=None):
        v = ctypes.c_double()
        check(lib().alamo_mg_norm_inf(self.h, mglev, self._vec(x, mglev, "NormInf x"), ctypes.byref(v), _ptr(stream)))
        return v.value

    def NormL2(self, mglev, x, stream=None):
        v = ctypes.c_double()
        check(lib().alamo_mg_norm_l2(self.h, mglev, self._vec(x, mglev, "NormL2 x"), ctypes.byref(v), _ptr(stream)))
        return v.value

    def Error0x(self, amrlev, mglev, R0x, x, stream=None):
        """Elastic::Error0x (Elastic.cpp:382-404): R0x = x - A (x / D)."""
        check(lib().alamo_mg_error0x(self.h, amrlev, mglev, self._vec(R0x, mglev, "Error0x R0x"),
                                      self._vec(x, mglev, "Error0x x"), _ptr(stream)))


class Linear:
    """Solver::Nonlocal::Linear (Linear.H:18-389)."""

    def __init__(self, op=None):
        self.linop = None
        if op is not None:
            self.Define(op)

    def Define(self, op):
        if self.linop is not None:
            raise _lib.AlamoError(3, "Solver cannot be re-defined")
        self.linop = op

    def Clear(self):
        if self.linop is None:
            raise _lib.AlamoError(3, "Solver cannot be cleared if not defined")
        self.linop = None

    # Linear::setMaxIter / setBottomMaxIter / setFixedIter / setVerbose / setPreSmooth /
    # setPostSmooth (Linear.H:178-184)
    def _opt(self, **kw):
        if self.linop is None:
            raise _lib.AlamoError(3, "Solver not defined")
        self.linop.UpdateOpts(**kw)

    def setMaxIter(self, n):
        self._opt(max_iter=int(n))

    def setBottomMaxIter(self, n):
        self._opt(bottom_max_iter=int(n))

    def setFixedIter(self, n):
        self._opt(fixed_iter=int(n))

    def setVerbose(self, v):
        self._opt(verbose=int(v))

    def setPreSmooth(self, n):
        self._opt(pre_smooth=int(n))

    def setPostSmooth(self, n):
        self._opt(post_smooth=int(n))

    def setBottomSolver(self, name):
        """solver.bottom_solver (Linear.H:280-282): "bicgstab", "smoother" or "dense"."""
        self._opt(bottom_solver=name)

    def setMaxFmgIter(self, n):
        """amrex::MLMG::setMaxFmgIter: F-cycles for the first n iterations, then V-cycles (the
        reference's own setting is 1e8, Linear.H:261)."""
        self._opt(max_fmg_iter=int(n))

    # Linear::Parse (Linear.H:296-389): the solver.* keys of the reference's input file, given
    # here as a mapping (key -> value).  final_smooth, bottom_smooth, bottom_tol_abs and the
    # cg_* keys have no counterpart in this handle and are rejected; dump_on_fail /
    # abort_on_fail are kept as attributes (a failed solve raises unless raise_on_fail=False).
    _PARSE_KEYS = {"max_iter": int, "bottom_max_iter": int, "max_fmg_iter": int, "fixed_iter": int,
                   "verbose": int, "pre_smooth": int, "post_smooth": int, "bottom_solver": str,
                   "bottom_tol_rel": float, "omega": float}

    def Parse(self, pp):
        rest = dict(pp)
        opts = {}
        for k in list(rest):
            if k in self._PARSE_KEYS:
                v = self._PARSE_KEYS[k](rest.pop(k))
                if k == "omega":
                    self.linop.SetOmega(v)
                elif not (isinstance(v, int) and v < 0):  # -1: keep the default, as Linear.H:263-278
                    opts[k] = v
        for k in ("tol_rel", "tol_abs", "dump_on_fail", "abort_on_fail", "average_down_coeffs", "normalize_ddw"):
            if k in rest:
                setattr(self, k, rest.pop(k))
        if getattr(self, "average_down_coeffs", False):
            self.linop.SetAverageDownCoeffs(True)
        if getattr(self, "normalize_ddw", False):
            self.linop.SetNormalizeDDW(True)
        if rest:
            raise ValueError("solver keys not supported by this handle: %s" % sorted(rest))
        if opts:
            self._opt(**opts)
        return self

    def apply(self, rhs, sol):
        """Linear::apply (Linear.H:169-174): rhs = A sol on the finest level"""
        self.linop.Fapply(0, 0, rhs, sol)

    def solveaffine(self, sol, rhs, tol_rel, tol_abs, raise_on_fail=True):
        """Linear::solveaffine (Linear.H:53-130): the solve of rhs minus A 0 evaluated with
        inhomogeneous boundaries.  Elastic's homogeneity flag is stored but never read
        (Elastic.H:48,172), so A 0 = 0 on every node and this is the plain solve."""
        return self.solve(sol, rhs, tol_rel, tol_abs, raise_on_fail)

    def solve(self, sol, rhs, tol_rel, tol_abs, raise_on_fail=True):
        """Returns the stats dict; sol holds the initial guess on entry (Linear.H:131-147)."""
        if self.linop is None:
            raise _lib.AlamoError(3, "Solver not defined")
        st = SolveStats()
        rc = lib().alamo_linear_solve(self.linop.h, self.linop._vec(sol, 0, "solve sol"),
                                      self.linop._vec(rhs, 0, "solve rhs"), tol_rel, tol_abs, ctypes.byref(st))
        if rc != 0 and raise_on_fail:
            check(rc)
        out = st.as_dict()
        out["rc"] = rc
        return out


class Newton(Linear):
    """Solver::Nonlocal::Newton<T> (Newton.H:14-654) for an affine model: stiffness C (the model's
    DDW, any layout SetModel takes) and eigenstrain F0 (9 AoS doubles per node, or None), both
    device fields.  solve sets the model, then each Newton-Raphson iteration (alamo_newton_solve)
    forms rhs = b - div DW(grad u) (boundary rows b - BC(u, grad u, DW)), solves A du = rhs with
    the linear solver and adds du to u; du starts at zero once and each later linear solve starts
    from the previous du, as in the reference (Newton.H:279-358)."""

    def __init__(self, op=None):
        super().__init__(op)
        self.m_nriters = 1         # Newton.H:630
        self.m_nrtolerance = 0.0   # Newton.H:631
        self.m_psi = None

    def setNRIters(self, n):
        self.m_nriters = int(n)

    def setPsi(self, psi, psi_small=1e-8):
        """Newton::setPsi (Newton.H:46-50): the operator's psi field too."""
        self.m_psi = psi
        self.linop.SetPsi(psi, psi_small)

    def _prepare(self, u, b, C, F0, rhs, layout):
        # the operator keeps psi across SetModel, so the right-hand side takes the psi form
        # (Newton.H:236-266) once setPsi has been called
        self.linop.SetModel(C, layout)
        self.linop.NewtonRHS(u, b, F0, rhs)

    def solve(self, u, b, C, F0, tol_rel, tol_abs, layout=LAYOUT_AOS, raise_on_fail=True):
        """Newton::solve: returns the relative correction norm of the last iteration (0 when
        m_nrtolerance never stopped it early), u updated in place."""
        if self.linop is None:
            raise _lib.AlamoError(3, "Solver not defined")
        self.linop.SetModel(C, layout)
        st = SolveStats()
        rel = ctypes.c_double(0.0)
        op = self.linop
        rc = lib().alamo_newton_solve(op.h, op._vec(u, 0, "Newton u"), op._vec(b, 0, "Newton b"),
                                      None if F0 is None else op._vec(F0, 0, "Newton F0", 9), int(self.m_nriters),
                                      float(self.m_nrtolerance), tol_rel, tol_abs, ctypes.byref(st), ctypes.byref(rel))
        self.last_stats = st.as_dict()
        self.last_stats["rc"] = rc
        if rc != 0 and raise_on_fail:
            check(rc)
        return rel.value

    def compResidual(self, res, u, b, C, F0, layout=LAYOUT_AOS):
        """Newton::compResidual (Newton.H:445-500): res = b - div DW(grad u), boundary rows
        b - BC(u, grad u, DW)."""
        self._prepare(u, b, C, F0, res, layout)

    def compLinearSolverResidual(self, res, u, b):
        """Newton::compLinearSolverResidual (Newton.H:502-526): res = b - A u."""
        self.linop.solutionResidual(0, res, u, b)

    def W(self, w, u, F0):
        """Newton::W (Newton.H:528-569): w = (grad u - F0) : C : (grad u - F0) / 2 per node."""
        op = self.linop
        check(lib().alamo_newton_w(op.h, 0, op._vec(u, 0, "W u"), None if F0 is None else op._vec(F0, 0, "W F0", 9),
                                   op._vec(w, 0, "W w", 1), None))

    def DW(self, dw, u, F0):
        """Newton::DW (Newton.H:571-611): dw = C : (grad u - F0), 9 row-major planes."""
        op = self.linop
        check(lib().alamo_newton_dw(op.h, 0, op._vec(u, 0, "DW u"), None if F0 is None else op._vec(F0, 0, "DW F0", 9),
                                    op._vec(dw, 0, "DW dw", 9), None))

    def Parse(self, pp):
        """Newton::Parse (Newton.H:643-652): Linear::Parse, then nriters and nrtolerance."""
        rest = dict(pp)
        if "nriters" in rest:
            self.m_nriters = int(rest.pop("nriters"))
        if "nrtolerance" in rest:
            self.m_nrtolerance = float(rest.pop("nrtolerance"))
        return Linear.Parse(self, rest)


def blend_models(eta, table, out, layout=LAYOUT_AOS, stream=None):
    """Integrator::Mechanics::UpdateModel blending (Mechanics.H:242-254): out = sum_n eta_n table_n.
    eta: device tensor of nmodels planes; table: numpy (nmodels, nrec); out: device tensor."""
    table = np.ascontiguousarray(table, dtype=np.float64)
    nmodels, nrec = table.shape
    nnodes = eta.numel() // nmodels
    eta_p = _dev(eta, nmodels, "blend eta")
    check(lib().alamo_blend_models(ctypes.c_int64(nnodes), nmodels, nrec, eta_p,
                                   table.ctypes.data_as(ctypes.c_void_p), _dev(out, nrec * nnodes, "blend out"), layout,
                                   stream))


def precond_apply(op, kind, r, z):
    """z = P(r) for the operator's preconditioner kind (one cycle, optionally with the network)"""
    check(lib().alamo_mg_precond(op.h, int(kind), op._vec(r, 0, "precond r"), op._vec(z, 0, "precond z")))


def apply_transpose(op, r, out):
    """out = A^T r for the level-0 operator (no reference counterpart; no psi terms)"""
    check(lib().alamo_mg_apply_transpose(op.h, op._vec(r, 0, "apply_transpose r"),
                                         op._vec(out, 0, "apply_transpose out")))


class Krylov:
    """Flexible GMRES(restart) around the operator's multigrid cycle (no reference counterpart:
    an outer accelerator for the same MLMG-style cycle, stopping on the reference's max-norm test)."""

    def __init__(self, op, restart=16, precond=PRECOND_MG, max_iter=400, stagnation_ratio=0.9, verbose=0):
        self.op = op
        o = KrylovOpts()
        check(lib().alamo_krylov_opts_default(ctypes.byref(o)))
        o.restart, o.precond, o.max_iter, o.stagnation_ratio, o.verbose = (restart, precond, max_iter,
                                                                           stagnation_ratio, verbose)
        self.opts = o

    def solve(self, sol, rhs, tol_rel, tol_abs, raise_on_fail=True):
        st = SolveStats()
        rc = lib().alamo_krylov_solve(self.op.h, self.op._vec(sol, 0, "Krylov sol"), self.op._vec(rhs, 0, "Krylov rhs"),
                                      tol_rel, tol_abs, ctypes.byref(self.opts),
                                      ctypes.byref(st))
        if rc != 0 and raise_on_fail:
            check(rc)
        out = st.as_dict()
        out["rc"] = rc
        return out


def learned_solve(op, sol, rhs, tol_rel, tol_abs, restart=16, max_iter=400, stagnation_ratio=0.9,
                  raise_on_fail=True):
    """alamo_learned_solve: FGMRES with the learned preconditioner of the network attached to op."""
    o = KrylovOpts()
    check(lib().alamo_krylov_opts_default(ctypes.byref(o)))
    o.restart, o.max_iter, o.stagnation_ratio = restart, max_iter, stagnation_ratio
    st = SolveStats()
    rc = lib().alamo_learned_solve(op.h, op._vec(sol, 0, "learned sol"), op._vec(rhs, 0, "learned rhs"), tol_rel, tol_abs,
                                   ctypes.byref(o), ctypes.byref(st))
    if rc != 0 and raise_on_fail:
        check(rc)
    out = st.as_dict()
    out["rc"] = rc
    return out


class Net:
    """The learned correction network (no reference counterpart).  Parameters live on the GPU in
    libalamo_amd.so; host arrays here are float32 numpy arrays of the library's layout."""

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        check(lib().alamo_net_create(ctypes.byref(h), ctypes.c_uint64(seed)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if getattr(self, "_owner", None) is not None:
            return
        if h is not None and h.value:
            try:
                lib().alamo_net_destroy(h)
            except Exception:
                pass
            self.h = None

    def num_params(self):
        n = ctypes.c_int64()
        check(lib().alamo_net_num_params(self.h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().alamo_net_get_params(self.h, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("expected %d parameters" % self.num_params())
        check(lib().alamo_net_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p)))

    def get_grads(self):
        g = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().alamo_net_get_grads(self.h, g.ctypes.data_as(ctypes.c_void_p)))
        return g

    def zero_grads(self):
        check(lib().alamo_net_zero_grads(self.h))

    def forward(self, nnodes, x):
        nn = np.ascontiguousarray(nnodes, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((int(np.prod(nn)), 16), dtype=np.float32)
        check(lib().alamo_net_forward_host(self.h, nn.ctypes.data_as(ctypes.c_void_p), x.ctypes.data_as(ctypes.c_void_p),
                                           y.ctypes.data_as(ctypes.c_void_p)))
        return y

    def backward(self, nnodes, x, dy, want_dx=True):
        nn = np.ascontiguousarray(nnodes, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        dx = np.zeros((int(np.prod(nn)), 16), dtype=np.float32) if want_dx else None
        check(lib().alamo_net_backward_host(self.h, nn.ctypes.data_as(ctypes.c_void_p),
                                            x.ctypes.data_as(ctypes.c_void_p), dy.ctypes.data_as(ctypes.c_void_p),
                                            dx.ctypes.data_as(ctypes.c_void_p) if want_dx else None))
        return dx

    def adam_step(self, lr):
        check(lib().alamo_net_adam_step(self.h, float(lr)))

    def train(self, op, npairs=8, steps=200, lr=1e-3, seed=1, solve_max_iter=60, solve_tol=1e-6, verbose=0,
              loss="error"):
        """Adam on pairs generated with op: loss "error" (MSE against the cycle's remaining error, solved
        per pair) or "residual" (1 - cos^2(q, A y), the stage's squared residual ratio)"""
        losses = {"error": LOSS_ERROR, "residual": LOSS_RESIDUAL}
        if loss not in losses:
            raise _lib.AlamoError(1, "loss must be one of %s" % sorted(losses))
        o = _lib.TrainOpts(npairs, steps, lr, seed, solve_max_iter, solve_tol, verbose, losses[loss])
        st = _lib.TrainStats()
        check(lib().alamo_net_train(self.h, op.h, ctypes.byref(o), ctypes.byref(st)))
        return st.as_dict()

    def layer_times(self, op, reps=5):
        """mean forward-kernel milliseconds of each conv layer at op's level-0 size"""
        ms = np.zeros(6)
        check(lib().alamo_net_layer_times(self.h, op.h, int(reps), ms.ctypes.data_as(ctypes.c_void_p)))
        return ms

    def layer_times_bwd(self, op, reps=5):
        """mean milliseconds of each conv layer's backward-data convolution and weight gradient at op's
        level-0 size: (ms_data[6], ms_wgrad[6])"""
        md, mw = np.zeros(6), np.zeros(6)
        check(lib().alamo_net_layer_times_bwd(self.h, op.h, int(reps), md.ctypes.data_as(ctypes.c_void_p),
                                              mw.ctypes.data_as(ctypes.c_void_p)))
        return md, mw

    def save(self, path):
        check(lib().alamo_net_save(self.h, str(path).encode()))

    def load(self, path):
        check(lib().alamo_net_load(self.h, str(path).encode()))

    def attach(self, op):
        """use this network for op's ALAMO_PRECOND_LEARNED solves (keeps a reference to it)"""
        check(lib().alamo_mg_set_net(op.h, self.h))
        op._net = self
