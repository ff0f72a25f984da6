"""Boundary-condition tables and synthetic input fields for the elastic solves.

Host-side input generators (not timed).  Each restates the reference routine
it mirrors:

* ``bc_constant``      BC::Operator::Elastic::Constant defaults (BC/Operator/Elastic/Constant.H:36-45)
* ``bc_tension_test``  BC::Operator::Elastic::TensionTest::Parse (BC/Operator/Elastic/TensionTest.H:48-306)
* ``sphere_indicator`` IC::Ellipse::Add node-position indicator (IC/Ellipse.H:26-86), sharp or erf form
* ``cubic_mm``         Set::Matrix4<3,MajorMinor>::Cubic (Set/Matrix4_MajorMinor.H:518-556)
* ``voronoi_grains``   IC::Voronoi grain partition (IC/Voronoi.H:37-118) with this package's own seeded RNG
* ``two_phase_indicator`` Gaussian random field thresholded at 0 with an erf interface (no reference counterpart)

Random numbers come from SplitMix64 (published by Steele, Lea and Flood, 2014),
so every field is reproducible from its integer seed on any machine.
"""
import math

import numpy as np

# Face enum of BC/Operator/Elastic/Elastic.H:35-43 (the [26][3] table index)
FACES = ["XLO", "YLO", "ZLO", "XHI", "YHI", "ZHI",
         "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI",
         "ZLO_XLO", "ZLO_XHI", "ZHI_XLO", "ZHI_XHI",
         "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI",
         "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI",
         "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI"]
F = {n: i for i, n in enumerate(FACES)}
# BC/Operator/Elastic/Elastic.H:26
DISPLACEMENT, TRACTION, PERIODIC, NEUMANN, NONE = range(5)
# Set::Sym subset handled by the operator
ISO, MAJORMINOR, DIAGONAL = range(3)
NCOEF = {ISO: 2, MAJORMINOR: 21, DIAGONAL: 9}


# ---------------------------------------------------------------------------
# SplitMix64
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def splitmix_uniform_array(seed, n):
    """Vectorised SplitMix64 stream: element q is the (q+1)-th draw of SplitMix64(seed)."""
    q = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + q * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ---------------------------------------------------------------------------
# Boundary conditions
# ---------------------------------------------------------------------------
def bc_constant(types=None, vals=None):
    """All 26 regions Displacement with value 0 (Constant.H:36-45), then overrides.

    ``types``/``vals`` map a face name to a 3-list of type codes / values.
    Returns (bc_type int32[26,3], bc_val float64[26,3]).
    """
    t = np.full((26, 3), DISPLACEMENT, dtype=np.int32)
    v = np.zeros((26, 3), dtype=np.float64)
    for k, x in (types or {}).items():
        t[F[k]] = x
    for k, x in (vals or {}).items():
        v[F[k]] = x
    return t, v


def bc_tension_test(kind="uniaxial_stress", disp=None, trac=None):
    """BC::Operator::Elastic::TensionTest::Parse (TensionTest.H:48-306) at a fixed time."""
    t, v = bc_constant()
    if (disp is None) == (trac is None):
        raise ValueError("specify exactly one of disp or trac")
    mode_disp = disp is not None
    val = disp if mode_disp else trac
    xhi_all = ["XHI", "XHI_YLO", "XHI_YHI", "ZLO_XHI", "ZHI_XHI",
               "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI"]
    xlo_all = ["XLO", "XLO_YLO", "XLO_YHI", "ZLO_XLO", "ZHI_XLO",
               "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI"]
    if kind == "uniaxial_stress_clamp":
        for f in xhi_all:
            v[F[f], 0] = val
        for f in ["YLO", "YHI", "ZLO", "ZHI", "YLO_ZLO", "YHI_ZLO", "YLO_ZHI", "YHI_ZHI"]:
            t[F[f], :] = TRACTION
    elif kind == "uniaxial_kolsky":
        for f in ["XLO", "XLO_YLO", "XLO_YHI"]:
            v[F[f], 0] = val
        for f in ["YLO", "YHI"]:
            t[F[f], :] = TRACTION
    elif kind in ("uniaxial_stress", "uniaxial_strain"):
        t[:, :] = TRACTION
        for f in xlo_all:
            t[F[f], 0] = DISPLACEMENT
        if kind == "uniaxial_strain" and not mode_disp:
            raise ValueError("uniaxial_strain takes disp")
        # x load on the xhi face/edges/corners (TensionTest.H:136-186 / :236-256)
        for f in xhi_all:
            t[F[f], 0] = DISPLACEMENT if mode_disp else TRACTION
            v[F[f], 0] = val
        if kind == "uniaxial_stress":
            for f in ["YLO", "XLO_YLO", "XHI_YLO", "YLO_ZLO", "YLO_ZHI",
                      "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XHI_YLO_ZLO", "XHI_YLO_ZHI"]:
                t[F[f], 1] = DISPLACEMENT
            for f in ["ZLO", "YLO_ZLO", "YHI_ZLO", "ZLO_XLO", "ZLO_XHI",
                      "XLO_YLO_ZLO", "XLO_YHI_ZLO", "XHI_YLO_ZLO", "XHI_YHI_ZLO"]:
                t[F[f], 2] = DISPLACEMENT
        else:
            for f in ["YLO", "XLO_YLO", "XHI_YLO", "YHI", "XLO_YHI", "XHI_YHI",
                      "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI"] + FACES[18:26]:
                t[F[f], 1] = DISPLACEMENT
            for f in ["ZLO", "ZHI", "YLO_ZLO", "YHI_ZLO", "ZLO_XLO", "ZLO_XHI",
                      "YLO_ZHI", "YHI_ZHI", "ZHI_XLO", "ZHI_XHI"] + FACES[18:26]:
                t[F[f], 2] = DISPLACEMENT
    else:
        raise ValueError("unknown tension test type " + kind)
    return t, v


def bc_top_pull(uz, lateral=NEUMANN):
    """Boundary table of the polycrystal workload: z=0 clamped (u=0), z=1 gripped with
    u = (0, 0, uz); the four lateral faces and the lateral edges use `lateral`
    (Neumann grad(u).n = 0 by default, TRACTION for traction-free sides).  Regions touching a
    z face follow the z-face rule.

    With traction-free sides the re-discretised operator of a 2-cell multigrid level is
    singular (the one-sided traction rows of opposite faces meet); the dense bottom solver
    then ends the hierarchy at 4 cells, where the operator is regular."""
    t, v = bc_constant()
    t[:, :] = lateral
    for i, name in enumerate(FACES):
        parts = name.split("_")
        if "ZLO" in parts:
            t[i, :] = DISPLACEMENT
            v[i, :] = 0.0
        elif "ZHI" in parts:
            t[i, :] = DISPLACEMENT
            v[i, 2] = uz
    return t, v


# ---------------------------------------------------------------------------
# Stiffness fields
# ---------------------------------------------------------------------------
def node_coords(ncell, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0)):
    """Node positions prob_lo + i*dx (IC/Ellipse.H:44-47), arrays shaped [k, j, i]."""
    dx = [(prob_hi[d] - prob_lo[d]) / ncell[d] for d in range(3)]
    x = prob_lo[0] + np.arange(ncell[0] + 1) * dx[0]
    y = prob_lo[1] + np.arange(ncell[1] + 1) * dx[1]
    z = prob_lo[2] + np.arange(ncell[2] + 1) * dx[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z, dx


def sphere_indicator(ncell, center=(0.5, 0.5, 0.5), radius=0.25, eps=0.0, prob_lo=(0, 0, 0), prob_hi=(1, 1, 1)):
    """Inclusion indicator at node positions.

    eps == 0: sharp (1 inside |x-c| < r, else 0).  eps > 0: the reference's erf
    form 0.5 - 0.5 erf(((x-c)^T A (x-c) - 1) / eps / |A (x-c)|) with A = I/r^2
    (IC/Ellipse.H:59-61)."""
    X, Y, Z, dx = node_coords(ncell, prob_lo, prob_hi)
    rx, ry, rz = X - center[0], Y - center[1], Z - center[2]
    if eps == 0.0:
        return (rx * rx + ry * ry + rz * rz < radius * radius).astype(np.float64)
    a = 1.0 / radius ** 2
    q = a * (rx * rx + ry * ry + rz * rz) - 1.0
    norm = a * np.sqrt(rx * rx + ry * ry + rz * rz)
    norm = np.where(norm == 0.0, 1e-300, norm)
    from scipy.special import erf
    return 0.5 - 0.5 * erf(q / eps / norm)


def iso_field(lam, mu):
    """AoS isotropic stiffness (lambda, mu per node), the member order of Matrix4<3,Isotropic>."""
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    return np.stack([lam.ravel(), mu.ravel()], axis=1).ravel()


def eshelby_problem(n=64, lam_in=10.0, mu_in=10.0, lam_out=1.0, mu_out=1.0, eig=1e-3, radius=0.25, eps=0.1):
    """Matrix/inclusion blend model = eta*inclusion + (1-eta)*matrix, componentwise on
    (C, F0) (Integrator/Mechanics.H:233-254 with affine isotropic models).

    Returns dict(ncell, dx, C (AoS iso), F0 (9 per node), eta)."""
    ncell = (n, n, n)
    eta = sphere_indicator(ncell, radius=radius, eps=eps)
    lam = eta * lam_in + (1 - eta) * lam_out
    mu = eta * mu_in + (1 - eta) * mu_out
    C = iso_field(lam, mu)
    F0 = np.zeros((eta.size, 9))
    F0[:, 0] = F0[:, 4] = F0[:, 8] = (eta * eig).ravel()
    return dict(ncell=ncell, dx=(1.0 / n,) * 3, C=C, F0=F0.ravel(), eta=eta)


def euler_R(phi1, Phi, phi2):
    """R = Rx(phi2) Rz(Phi) Rx(phi1) (Set/Matrix4_MajorMinor.H:551-554)."""
    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])

    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return rx(phi2) @ rz(Phi) @ rx(phi1)


def _pair(a, b):
    a, b = min(a, b), max(a, b)
    return b if a == 0 else (2 + b if a == 1 else 5)


def mm_slot(i, j, k, l):
    r, s = _pair(i, j), _pair(k, l)
    r, s = min(r, s), max(r, s)
    return r * 6 - r * (r - 1) // 2 + (s - r)


def cubic_mm(C11, C12, C44, R=None):
    """21-slot MajorMinor record of the rotated cubic stiffness, written slot by slot in the
    reference's (p,q,s,t) loop order so the last permutation written wins
    (Set/Matrix4_MajorMinor.H:518-547)."""
    if R is None:
        R = np.eye(3)
    Ct = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    if i == j == k == l:
                        Ct[i, j, k, l] = C11
                    elif i == k and j == l:
                        Ct[i, j, k, l] = C44
                    elif i == j and k == l:
                        Ct[i, j, k, l] = C12
    full = np.einsum("pi,qj,sk,tl,ijkl->pqst", R, R, R, R, Ct)
    out = np.zeros(21)
    for p in range(3):
        for q in range(3):
            for s in range(3):
                for t in range(3):
                    out[mm_slot(p, q, s, t)] = full[p, q, s, t]
    return out


def voronoi_grains(ncell, ngrains, seed):
    """Grain id per node: nearest of `ngrains` seed points drawn uniformly in the unit cube
    (IC/Voronoi.H:37-118 partitions by nearest seed; the seeds here come from SplitMix64).
    Positions are node positions i*dx."""
    rng = SplitMix64(seed)
    pts = np.array([[rng.uniform() for _ in range(3)] for _ in range(ngrains)])
    X, Y, Z, dx = node_coords(ncell)
    gid = np.zeros(X.shape, dtype=np.int32)
    best = np.full(X.shape, np.inf)
    for g in range(ngrains):
        d2 = (X - pts[g, 0]) ** 2 + (Y - pts[g, 1]) ** 2 + (Z - pts[g, 2]) ** 2
        m = d2 < best
        best[m] = d2[m]
        gid[m] = g
    return gid, pts


def grain_table(ngrains, seed, C11=1.68, C12=1.21, C44=0.75):
    """Per-grain rotated cubic records; Euler angles uniform on [0, 2pi) as
    Model::Solid::Linear::Cubic::Random (Linear/Cubic.H:81-90), drawn from SplitMix64."""
    rng = SplitMix64(seed ^ 0xC0B1C)
    tab = np.zeros((ngrains, 21))
    for g in range(ngrains):
        a = [2.0 * math.pi * rng.uniform() for _ in range(3)]
        tab[g] = cubic_mm(C11, C12, C44, euler_R(a[0], a[1], a[2]))
    return tab


def polycrystal_problem(n=256, ngrains=64, seed=64, uz=1e-3, lateral=NEUMANN):
    ncell = (n, n, n)
    gid, _ = voronoi_grains(ncell, ngrains, seed)
    tab = grain_table(ngrains, seed)
    bct, bcv = bc_top_pull(uz, lateral)
    return dict(ncell=ncell, dx=(1.0 / n,) * 3, gid=gid, table=tab, bct=bct, bcv=bcv)


def grf(ncell, corr_cells, seed):
    """Standardised Gaussian random field on the nodes ([k, j, i] array) with Gaussian
    covariance of correlation length `corr_cells` cells.  White noise is a SplitMix64
    stream made Gaussian by Box-Muller; the covariance is applied as an FFT filter."""
    shape = tuple(int(n) + 1 for n in ncell)[::-1]
    N = int(np.prod(shape))
    u1 = splitmix_uniform_array(seed, N)
    u2 = splitmix_uniform_array(seed ^ 0x5DEECE66D, N)
    g = (np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)).reshape(shape)
    ks = [np.fft.fftfreq(s) for s in shape]
    KZ, KY, KX = np.meshgrid(*ks, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    filt = np.exp(-(math.pi ** 2) * (corr_cells ** 2) * k2)
    field = np.real(np.fft.ifftn(np.fft.fftn(g) * filt))
    return field / field.std()


def two_phase_indicator(ncell, corr_cells, seed, width=1.0):
    """Phase fraction eta in [0, 1]: the GRF thresholded at 0 with an erf interface,
    eta = (1 + erf(g / width)) / 2.  width -> 0 gives the sharp threshold."""
    from scipy.special import erf
    g = grf(ncell, corr_cells, seed)
    if width <= 0.0:
        return (g > 0.0).astype(np.float64)
    return 0.5 * (1.0 + erf(g / width))


def interface_stats(g, eta, width):
    """Shape of a two-phase field: the 10-90 % transition width of eta across the interface in
    cells (1.8124 width / mean |grad g| over the nodes with |g| < 0.05, per cell), the fraction of
    nodes within 5 % of a pure phase, and the fraction with 0.25 <= eta <= 0.75."""
    gx, gy, gz = np.gradient(g)
    near = np.abs(g) < 0.05
    grad = float(np.sqrt(gx[near] ** 2 + gy[near] ** 2 + gz[near] ** 2).mean()) if near.any() else float("nan")
    w1090 = 2.0 * 0.9061938 * width / grad if width > 0 else 0.0
    return {"interface_cells_10_90": w1090, "pure_phase_frac": float(((eta < 0.05) | (eta > 0.95)).mean()),
            "mid_frac": float(((eta >= 0.25) & (eta <= 0.75)).mean())}


def two_phase_problem(n=256, corr_cells=8, contrast=1000.0, seed=256, load=1e-3, width=0.5):
    """Two-phase isotropic field: lambda = mu = contrast**eta (1 in phase 0, `contrast` in
    phase 1, geometric blend across the interface) with an affine dilatational eigenstrain
    `load`*eta; all-Dirichlet (u = 0) box, so the forcing sits at the interfaces.

    eta = (1 + erf(g / width)) / 2 for the standardised Gaussian random field g.  The default
    width 0.5 puts the 10-90 % transition across about 4 cells at correlation length 8 (see
    interface_stats).  Sharper interfaces (width <= 0.25, two cells and less, and the thresholded
    width 0) make the strong-form rows next to the interface lose their diagonal: the operator
    has eigenvalues of D^-1 A in the left half plane and its discrete solution exceeds the
    eigenstrain's physical displacement scale many times, so no multigrid-preconditioned Krylov
    method converges on it."""
    ncell = (n, n, n)
    g = grf(ncell, corr_cells, seed)
    if width <= 0.0:
        eta = (g > 0.0).astype(np.float64)
    else:
        from scipy.special import erf
        eta = 0.5 * (1.0 + erf(g / width))
    stats = interface_stats(g, eta, width)
    del g
    lam = np.exp(eta * math.log(contrast))
    C = iso_field(lam, lam)
    F0 = np.zeros((eta.size, 9))
    F0[:, 0] = F0[:, 4] = F0[:, 8] = (eta * load).ravel()
    bct, bcv = bc_constant()
    return dict(ncell=ncell, dx=(1.0 / n,) * 3, C=C, F0=F0.ravel(), eta=eta, bct=bct, bcv=bcv, width=width,
                interface=stats)


# ---------------------------------------------------------------------------
# The reference's multi-level regression problems (tests/UniaxialTension, tests/TrigTest) and the
# AMR Eshelby configuration, on explicit meshes (explicitmesh.lo<l> / hi<l>, Integrator.cpp:121-139)
# ---------------------------------------------------------------------------
def _face_key(name):
    """the reference's lower-case region key ("xloylo", "zhixlo", ...) -> FACES name"""
    parts = [name[i:i + 3] for i in range(0, len(name), 3)]
    return "_".join(p.upper() for p in parts)


def explicit_mesh(n, nlevels=3):
    """tests/UniaxialTension/input:42-46 and tests/TrigTest/input:31-35 scaled to n base cells: level
    1 the middle half of the domain, level 2 the middle half of level 1 (level l box in level-l
    cells).  n = 32 gives the reference's lo1 16 16 16 / hi1 47 47 47, lo2 48 48 48 / hi2 79 79 79."""
    boxes = []
    for l in range(1, nlevels):
        N = n << l                      # level-l cells across the domain
        w = N >> l                      # the box spans 1/2^l of the domain
        lo = (N - w) // 2
        boxes.append([((lo,) * 3, (lo + w - 1,) * 3)])
    return boxes


def uniaxial_tension_problem(n=32, E=210.0, nu=0.3, disp=0.1):
    """tests/UniaxialTension/input: the box [-8, 8]^3 of n^3 cells, isotropic E = 210, nu = 0.3
    (Model/Solid/Linear/Isotropic.H:166-170: lambda = E nu / (1 + nu) / (1 - 2 nu), mu = E / 2 / (1 + nu)),
    u_x = 0 at x = -8, u_x = 0.1 at x = 8, u_y = 0 at y = -8, u_z = 0 at z = -8, traction-free otherwise
    (the input's 26-region table, :72-125).  Exact solution (tests/UniaxialTension/test:26-27):
    u = eps ((x + 8), -nu (y + 8), -nu (z + 8)) with eps = 0.1 / 16."""
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = E / 2.0 / (1.0 + nu)
    D, T = DISPLACEMENT, TRACTION
    table = {
        "xlo": (D, T, T, 0), "xhi": (D, T, T, disp), "ylo": (T, D, T, 0), "yhi": (T, T, T, 0),
        "zlo": (T, T, D, 0), "zhi": (T, T, T, 0),
        "xloylo": (D, T, T, 0), "xloyhi": (D, T, T, 0), "xhiylo": (D, T, T, disp), "xhiyhi": (D, T, T, disp),
        "zloxlo": (D, T, T, 0), "zhixlo": (D, T, T, 0), "zloxhi": (D, T, T, disp), "zhixhi": (D, T, T, disp),
        "ylozlo": (T, D, D, 0), "ylozhi": (T, D, T, 0), "yhizlo": (T, T, D, 0), "yhizhi": (T, T, T, 0),
        "xloylozlo": (D, D, D, 0), "xloylozhi": (D, D, T, 0), "xloyhizlo": (D, T, D, 0), "xloyhizhi": (D, T, T, 0),
        "xhiylozlo": (D, D, D, disp), "xhiylozhi": (D, D, T, disp), "xhiyhizlo": (D, T, D, disp),
        "xhiyhizhi": (D, T, T, disp),
    }
    t, v = bc_constant()
    for key, (a, b, c, vx) in table.items():
        f = F[_face_key(key)]
        t[f] = (a, b, c)
        v[f] = (vx, 0.0, 0.0)
    eps = disp / 16.0

    def exact(ncell):
        X, Y, Z, _ = node_coords(ncell, (-8, -8, -8), (8, 8, 8))
        return np.concatenate([(eps * (X + 8)).ravel(), (-nu * eps * (Y + 8)).ravel(), (-nu * eps * (Z + 8)).ravel()])

    def model(ncell):
        N = int(np.prod([c + 1 for c in ncell]))
        return iso_field(np.full(N, lam), np.full(N, mu))

    return dict(ncell=(n, n, n), dx=(16.0 / n,) * 3, bct=t, bcv=v, lam=lam, mu=mu, exact=exact, model=model)


def trig_problem(n=32):
    """tests/TrigTest/input: the unit box of n^3 cells, the Laplacian model with alpha = 1
    (Model/Solid/Linear/Laplacian.H:26-60: Matrix4<Diagonal> of ones, f_i = u_i,jj), u = 0 on all 26
    regions, body force b_x = sin(pi x) sin(pi y) sin(pi z) at the nodes (IC/Trig.H:112-158 with ni =
    1 1 1).  Exact solution (tests/TrigTest/test:24): u_x = -sin(pi x) sin(pi y) sin(pi z) / (3 pi^2)."""
    t, v = bc_constant()

    def rhs(ncell):
        X, Y, Z, _ = node_coords(ncell)
        b = np.zeros((3, X.size))
        b[0] = (np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)).ravel()
        return b.ravel()

    def exact(ncell):
        X, Y, Z, _ = node_coords(ncell)
        u = np.zeros((3, X.size))
        u[0] = (-np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z) / (3.0 * np.pi ** 2)).ravel()
        return u.ravel()

    def model(ncell):
        N = int(np.prod([c + 1 for c in ncell]))
        return np.ones(9 * N)

    return dict(ncell=(n, n, n), dx=(1.0 / n,) * 3, bct=t, bcv=v, rhs=rhs, exact=exact, model=model)


def eshelby_amr_mesh(n=64):
    """The AMR Eshelby configuration's explicit mesh (base n^3 cells; boxes of at most 32^3 cells):
    level 1 covers coarse cells [n/8, 7n/8) (the spherical interface r = 0.25 with its erf shell
    and a margin), level 2 the coarse cells [n/4, 3n/4) (the interface), each as a grid of 32^3
    boxes (n = 64: 27 and 64 boxes)."""
    def grid(lo, hi, s=32):
        out = []
        edges = list(range(lo, hi + 1, s))
        if edges[-1] != hi + 1:
            edges.append(hi + 1)
        for z0, z1 in zip(edges[:-1], edges[1:]):
            for y0, y1 in zip(edges[:-1], edges[1:]):
                for x0, x1 in zip(edges[:-1], edges[1:]):
                    out.append(((x0, y0, z0), (x1 - 1, y1 - 1, z1 - 1)))
        return out
    l1 = grid(2 * (n // 8), 2 * (7 * n // 8) - 1)
    l2 = grid(4 * (n // 4), 4 * (3 * n // 4) - 1)
    return [l1, l2]
