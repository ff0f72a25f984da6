// CPU oracle for the nodal elastic multigrid hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain FP64 restatement of the
// reference's algorithm, used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py as the checker.  The product (alamo_amd/) never
// links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the alamo source tree).  Layout conventions:
//   * a level has n[d] cells and n[d]+1 nodes per direction, node (i,j,k) lives
//     at flat index i + N0*(j + N1*k) (i fastest, as amrex::Array4);
//   * vector fields are SoA: component c occupies [c*Ntot, (c+1)*Ntot);
//   * stiffness fields are AoS, ncoef doubles per node, exactly the member
//     layout of Set::Matrix4<3,SYM> (Isotropic: lambda,mu; MajorMinor: the 21
//     slots of src/Set/Matrix4_MajorMinor.H:321-372; Diagonal: A[3][3]).
//
// Parity status of each part is stated next to it ("pinned by" = the test that
// checks it against something other than this file).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <functional>

namespace ora {

enum Sym { ISO = 0, MAJORMINOR = 1, DIAGONAL = 2 };
enum BCType { DISPLACEMENT = 0, TRACTION = 1, PERIODIC = 2, NEUMANN = 3, NONE = 4 };  // BC/Operator/Elastic/Elastic.H:26
enum Face {
    XLO, YLO, ZLO, XHI, YHI, ZHI,
    YLO_ZLO, YLO_ZHI, YHI_ZLO, YHI_ZHI,
    ZLO_XLO, ZLO_XHI, ZHI_XLO, ZHI_XHI,
    XLO_YLO, XLO_YHI, XHI_YLO, XHI_YHI,
    XLO_YLO_ZLO, XLO_YLO_ZHI, XLO_YHI_ZLO, XLO_YHI_ZHI,
    XHI_YLO_ZLO, XHI_YLO_ZHI, XHI_YHI_ZLO, XHI_YHI_ZHI,
    INT
};  // BC/Operator/Elastic/Elastic.H:35-43

static inline int ncoef(int sym) { return sym == ISO ? 2 : sym == MAJORMINOR ? 21 : 9; }

struct Level {
    int n[3];          // cells
    double dx[3];
    int N[3];          // nodes
    int64_t tot;
    Level() {}
    Level(const int* nc, const double* h) {
        for (int d = 0; d < 3; d++) { n[d] = nc[d]; dx[d] = h[d]; N[d] = nc[d] + 1; }
        tot = (int64_t)N[0] * N[1] * N[2];
    }
    inline int64_t id(int i, int j, int k) const { return i + (int64_t)N[0] * (j + (int64_t)N[1] * k); }
    inline bool bnd(int i, int j, int k) const {
        return i == 0 || j == 0 || k == 0 || i == n[0] || j == n[1] || k == n[2];
    }
    // Elastic::m_psi_mf of this level (cell field, n[0] n[1] n[2] values, i fastest) and
    // m_psi_small; psi == nullptr is the reference's m_psi_set == false
    const double* psi = nullptr;
    double psi_small = 1e-8;
    // Elastic::SetUniform(true) (Elastic.H:105): the interior rows drop grad(C) . grad(u)
    bool uniform = false;
};

// ---------------------------------------------------------------------------
// psi (Operator/Elastic.cpp:118-137,191-192,255-260; Numeric/Stencil.H:699-715,1385-1406).
// The operator fills its whole psi field, ghost cells included, with 1 at define time
// (Elastic.cpp:49-52) and SetPsi copies only the domain's cells (:122-133), so a cell outside
// the non-periodic domain reads 1.
// ---------------------------------------------------------------------------
static inline double psi_cell(const Level& L, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= L.n[0] || j >= L.n[1] || k >= L.n[2]) return 1.0;
    return L.psi[i + (int64_t)L.n[0] * (j + (int64_t)L.n[1] * k)];
}
// Interpolate::CellToNodeAverage: mean of the 8 cells around node (i,j,k)
static inline double psi_node_mean(const Level& L, int i, int j, int k) {
    double s = 0.0;
    for (int c = 0; c < 8; c++) s += psi_cell(L, i - (c & 1), j - ((c >> 1) & 1), k - (c >> 2));
    return s * 0.125;
}
// psi_avg = (1 - psi_small) * CellToNodeAverage(psi) + psi_small   (Elastic.cpp:191-192)
static inline double psi_avg(const Level& L, int i, int j, int k) {
    return L.psi ? (1.0 - L.psi_small) * psi_node_mean(L, i, j, k) + L.psi_small : 1.0;
}
// Numeric::CellGradientOnNode (Stencil.H:709-712): for d, the mean over the 4 cell pairs that
// straddle the node in d of (cell on the + side - cell on the - side) / dx_d
static inline void psi_grad(const Level& L, int i, int j, int k, double g[3]) {
    const int idx[3] = {i, j, k};
    for (int d = 0; d < 3; d++) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        double s = 0.0;
        for (int qa = 0; qa < 2; qa++)
            for (int qb = 0; qb < 2; qb++) {
                int p[3], q[3];
                p[d] = idx[d];
                q[d] = idx[d] - 1;
                p[a] = q[a] = idx[a] - qa;
                p[b] = q[b] = idx[b] - qb;
                s += psi_cell(L, p[0], p[1], p[2]) - psi_cell(L, q[0], q[1], q[2]);
            }
        g[d] = 0.25 * s / L.dx[d];
    }
}

// ---------------------------------------------------------------------------
// Set::Matrix4 element access.  Pinned by tests/test_oracle_tensor.py (dense
// 81-entry tensors built from the definitions, symmetry, the slot table).
// ---------------------------------------------------------------------------
// pair index of (a,b) in the order 00,01,02,11,12,22
static inline int pair_of(int a, int b) {
    if (a > b) std::swap(a, b);
    return a == 0 ? b : (a == 1 ? 2 + b : 5);
}
// MajorMinor slot of C_ijkl: upper triangle (row-major) of the 6x6 pair matrix.
// Restates the uid table of Set/Matrix4_MajorMinor.H:327-375.
static inline int mm_slot(int i, int j, int k, int l) {
    int r = pair_of(i, j), s = pair_of(k, l);
    if (r > s) std::swap(r, s);
    return r * 6 - r * (r - 1) / 2 + (s - r);
}
// C_ijkl for a node's coefficient record.
//   Isotropic: Set/Matrix4_Isotropic.H:52-59
//   Diagonal:  Set/Matrix4_Diagonal.H (operator(): A[i][j] if i==k && j==l)
static inline double C4(int sym, const double* c, int i, int j, int k, int l) {
    if (sym == ISO) {
        double r = 0.0;
        if (i == k && j == l) r += c[1];
        if (i == l && j == k) r += c[1];
        if (i == j && k == l) r += c[0];
        return r;
    } else if (sym == MAJORMINOR) {
        return c[mm_slot(i, j, k, l)];
    } else {
        return (i == k && j == l) ? c[3 * i + j] : 0.0;
    }
}

// Cubic stiffness rotated by R, restating Set/Matrix4_MajorMinor.H:518-547
// INCLUDING its storage behaviour: the 81 rotated entries are written into the
// 21 shared slots in (p,q,s,t) loop order, so the last permutation written
// wins.  (Ctmp is not minor-symmetric: Ctmp_ijji = 0 for i != j.)
static void cubic_mm(double C11, double C12, double C44, const double R[9], double* out) {
    double Ct[3][3][3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
        if (i == j && j == k && k == l) Ct[i][j][k][l] = C11;
        else if (i == k && j == l) Ct[i][j][k][l] = C44;
        else if (i == j && k == l) Ct[i][j][k][l] = C12;
        else Ct[i][j][k][l] = 0.0;
    }
    for (int p = 0; p < 3; p++) for (int q = 0; q < 3; q++)
    for (int s = 0; s < 3; s++) for (int t = 0; t < 3; t++) {
        double v = 0.0;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++)
            v += R[3 * p + i] * R[3 * s + k] * Ct[i][j][k][l] * R[3 * q + j] * R[3 * t + l];
        out[mm_slot(p, q, s, t)] = v;
    }
}
// R = Rx(phi2) * Rz(Phi) * Rx(phi1)   (Set/Matrix4_MajorMinor.H:548-556)
static void euler_R(double phi1, double Phi, double phi2, double* R) {
    auto mul = [](const double* a, const double* b, double* c) {
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            double s = 0; for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j]; c[3 * i + j] = s; }
    };
    double c1 = std::cos(phi1), s1 = std::sin(phi1), c = std::cos(Phi), s = std::sin(Phi),
           c2 = std::cos(phi2), s2 = std::sin(phi2);
    double Rx2[9] = {1, 0, 0, 0, c2, -s2, 0, s2, c2};
    double Rz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    double Rx1[9] = {1, 0, 0, 0, c1, -s1, 0, s1, c1};
    double T[9];
    mul(Rx2, Rz, T);
    mul(T, Rx1, R);
}

// sigma = C : A   (Matrix4 * Set::Matrix, e.g. Matrix4_Isotropic.H:137-157)
static inline void c_dot_mat(int sym, const double* c, const double A[3][3], double S[3][3]) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) s += C4(sym, c, i, j, k, l) * A[k][l];
        S[i][j] = s;
    }
}

// ---------------------------------------------------------------------------
// Boundary classification and rows.  Restates BC/Operator/Elastic/Constant.H:
//   region precedence :373-400 (same order as getType :309-336 and Init :231-258)
//   row values  set() :409-425, normals with the truncated SQRT3INV/SQRT2INV of
//   BC/Operator/Elastic/Elastic.H:79-80.
// Pinned by tests/test_oracle_bc.py (exhaustive 5^3 classification, hand rows).
// ---------------------------------------------------------------------------
static const double S3 = 0.57735026919, S2 = 0.70710678118;

static int classify(int i, int j, int k, const int* lo, const int* hi) {
    bool xl = i == lo[0], xh = i == hi[0], yl = j == lo[1], yh = j == hi[1], zl = k == lo[2], zh = k == hi[2];
    if (xl && yl && zl) return XLO_YLO_ZLO;
    if (xl && yl && zh) return XLO_YLO_ZHI;
    if (xl && yh && zl) return XLO_YHI_ZLO;
    if (xl && yh && zh) return XLO_YHI_ZHI;
    if (xh && yl && zl) return XHI_YLO_ZLO;
    if (xh && yl && zh) return XHI_YLO_ZHI;
    if (xh && yh && zl) return XHI_YHI_ZLO;
    if (xh && yh && zh) return XHI_YHI_ZHI;
    if (yl && zl) return YLO_ZLO;
    if (yl && zh) return YLO_ZHI;
    if (yh && zl) return YHI_ZLO;
    if (yh && zh) return YHI_ZHI;
    if (zl && xl) return ZLO_XLO;
    if (zl && xh) return ZLO_XHI;
    if (zh && xl) return ZHI_XLO;
    if (zh && xh) return ZHI_XHI;
    if (xl && yl) return XLO_YLO;
    if (xl && yh) return XLO_YHI;
    if (xh && yl) return XHI_YLO;
    if (xh && yh) return XHI_YHI;
    if (xl) return XLO;
    if (xh) return XHI;
    if (yl) return YLO;
    if (yh) return YHI;
    if (zl) return ZLO;
    if (zh) return ZHI;
    return INT;
}
static void normal_of(int f, double* n) {
    static const double tab[26][3] = {
        {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
        {0, -S2, -S2}, {0, -S2, S2}, {0, S2, -S2}, {0, S2, S2},
        {-S2, 0, -S2}, {S2, 0, -S2}, {-S2, 0, S2}, {S2, 0, S2},
        {-S2, -S2, 0}, {-S2, S2, 0}, {S2, -S2, 0}, {S2, S2, 0},
        {-S3, -S3, -S3}, {-S3, -S3, S3}, {-S3, S3, -S3}, {-S3, S3, S3},
        {S3, -S3, -S3}, {S3, -S3, S3}, {S3, S3, -S3}, {S3, S3, S3}};
    for (int d = 0; d < 3; d++) n[d] = tab[f][d];
}
static inline void bc_row(const int* type /*[26][3]*/, int face, const double u[3], const double G[3][3],
                          const double S[3][3], double f[3]) {
    double n[3];
    normal_of(face, n);
    for (int c = 0; c < 3; c++) {
        int t = type[3 * face + c];
        if (t == DISPLACEMENT) f[c] = u[c];
        else if (t == TRACTION) f[c] = S[c][0] * n[0] + S[c][1] * n[1] + S[c][2] * n[2];
        else if (t == NEUMANN) f[c] = G[c][0] * n[0] + G[c][1] * n[1] + G[c][2] * n[2];
        else f[c] = 0.0;
    }
}

// First derivative of component c along d with the stencil type chosen by
// Numeric::GetStencil (Numeric/Stencil.H:41-60, 70-191): forward at the lo
// node, backward at the hi node, central elsewhere.
static inline double d1(const Level& L, const double* f, int i, int j, int k, int d) {
    int idx[3] = {i, j, k};
    int64_t s = d == 0 ? 1 : d == 1 ? L.N[0] : (int64_t)L.N[0] * L.N[1];
    int64_t m = L.id(i, j, k);
    if (idx[d] == 0) return (f[m + s] - f[m]) / L.dx[d];
    if (idx[d] == L.n[d]) return (f[m] - f[m - s]) / L.dx[d];
    return (f[m + s] - f[m - s]) * 0.5 / L.dx[d];
}

static inline void grad_u(const Level& L, const double* u, int i, int j, int k, double G[3][3]) {
    for (int p = 0; p < 3; p++)
        for (int d = 0; d < 3; d++) G[p][d] = d1(L, u + p * L.tot, i, j, k, d);
}

// ---------------------------------------------------------------------------
// a1: Operator::Elastic<SYM>::Fapply (Operator/Elastic.cpp:139-292), single
// box covering the whole non-periodic domain, m_uniform = false
// (Integrator/Base/Mechanics.H:194); with psi set, the psi terms of :191-195,243-260.
//   interior f_i = C_ijkl u_k,lj + (dC/dx_d)_idkl u_k,l          (:211-254)
//   boundary f   = BC(u, grad u, C:grad u)                        (:195-203)
// Pinned by tests/test_oracle_fapply.py: patch test, quadratic fields with
// constant C (closed form C:grad grad u), linear C (closed form), brute-force
// dense assembly symmetry-free checks.
// ---------------------------------------------------------------------------
static void fapply_node(const Level& L, int sym, const double* C, const int* bct, const double* u,
                        int i, int j, int k, double f[3]) {
    const int nc = ncoef(sym);
    int64_t m = L.id(i, j, k);
    const double* c = C + m * nc;
    double G[3][3];
    grad_u(L, u, i, j, k, G);
    const double pa = psi_avg(L, i, j, k);
    if (L.bnd(i, j, k)) {
        int lo[3] = {0, 0, 0}, hi[3] = {L.n[0], L.n[1], L.n[2]};
        int face = classify(i, j, k, lo, hi);
        double S[3][3], uu[3];
        c_dot_mat(sym, c, G, S);
        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) S[a][b] *= pa;  // sig = (DDW grad u) psi_avg (:195)
        for (int p = 0; p < 3; p++) uu[p] = u[p * L.tot + m];
        bc_row(bct, face, uu, G, S, f);
        return;
    }
    const int64_t st[3] = {1, L.N[0], (int64_t)L.N[0] * L.N[1]};
    // gg[p][l][j] = u_{p,lj}
    double gg[3][3][3];
    for (int p = 0; p < 3; p++) {
        const double* up = u + p * L.tot;
        for (int a = 0; a < 3; a++) {
            gg[p][a][a] = (up[m + st[a]] - 2.0 * up[m] + up[m - st[a]]) / L.dx[a] / L.dx[a];  // Stencil.H:207
        }
        for (int a = 0; a < 3; a++) for (int b = a + 1; b < 3; b++) {
            double v = 0.5 * 0.5 * (up[m + st[a] + st[b]] + up[m - st[a] - st[b]] - up[m + st[a] - st[b]] -
                                    up[m - st[a] + st[b]]) / (L.dx[a] * L.dx[b]);                  // Stencil.H:261
            gg[p][a][b] = v;
            gg[p][b][a] = v;
        }
    }
    for (int ii = 0; ii < 3; ii++) {
        double s = 0.0;
        for (int jj = 0; jj < 3; jj++) for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++)
            s += C4(sym, c, ii, jj, kk, ll) * gg[kk][ll][jj];
        f[ii] = s;
    }
    // grad(C) . grad(u), with Stencil<MATRIX4,...>::D = (C(+) - C(-)) * 0.5 / dx  (Elastic.cpp:245-254),
    // skipped when m_uniform
    for (int d = 0; d < 3 && !L.uniform; d++) {
        const double* cp = C + (m + st[d]) * nc;
        const double* cm = C + (m - st[d]) * nc;
        double dc[21];
        for (int q = 0; q < nc; q++) dc[q] = (cp[q] - cm[q]) * 0.5 / L.dx[d];
        for (int ii = 0; ii < 3; ii++) {
            double s = 0.0;
            for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++) s += C4(sym, dc, ii, d, kk, ll) * G[kk][ll];
            f[ii] += s;
        }
    }
    if (L.psi) {
        // both terms above scale by psi_avg (:243,:253); then f += (DDW grad u) grad psi (1 - psi_small) (:255-260)
        double S[3][3], gp[3];
        c_dot_mat(sym, c, G, S);
        psi_grad(L, i, j, k, gp);
        for (int ii = 0; ii < 3; ii++) {
            double s = 0.0;
            for (int jj = 0; jj < 3; jj++) s += S[ii][jj] * gp[jj] * (1.0 - L.psi_small);
            f[ii] = f[ii] * pa + s;
        }
    }
}

static void fapply(const Level& L, int sym, const double* C, const int* bct, const double* u, double* f) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                double r[3];
                fapply_node(L, sym, C, bct, u, i, j, k, r);
                int64_t m = L.id(i, j, k);
                for (int p = 0; p < 3; p++) f[p * L.tot + m] = r[p];
            }
}

// ---------------------------------------------------------------------------
// a6: Operator::Elastic<SYM>::Diagonal (Operator/Elastic.cpp:296-379) with
// Numeric::Gradient_Diagonal (Numeric/Stencil.H:906-963).
// Pinned by tests/test_oracle_fapply.py::test_diagonal_brute_force
// (D(n,p) == (A e_{n,p})_{n,p}, probing fapply on small grids).
// ---------------------------------------------------------------------------
static void diagonal(const Level& L, int sym, const double* C, const int* bct, double* D) {
    const int nc = ncoef(sym);
    int lo[3] = {0, 0, 0}, hi[3] = {L.n[0], L.n[1], L.n[2]};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                int64_t m = L.id(i, j, k);
                const double* c = C + m * nc;
                int idx[3] = {i, j, k};
                if (L.bnd(i, j, k)) {
                    int face = classify(i, j, k, lo, hi);
                    double cd[3];
                    for (int d = 0; d < 3; d++) {
                        bool mn = idx[d] == 0, mx = idx[d] == L.n[d];
                        cd[d] = ((mx ? 1.0 : 0.0) - (mn ? 1.0 : 0.0)) / ((mn || mx ? 1.0 : 2.0) * L.dx[d]);
                    }
                    for (int p = 0; p < 3; p++) {
                        double G[3][3] = {{0}}, S[3][3], uu[3] = {0, 0, 0}, f[3];
                        for (int d = 0; d < 3; d++) G[p][d] = cd[d];
                        c_dot_mat(sym, c, G, S);
                        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) S[a][b] *= psi_avg(L, i, j, k);
                        uu[p] = 1.0;
                        bc_row(bct, face, uu, G, S, f);
                        D[p * L.tot + m] = f[p];
                    }
                } else {
                    for (int p = 0; p < 3; p++) {
                        double s = 0.0;
                        for (int jj = 0; jj < 3; jj++) s += C4(sym, c, p, jj, p, jj) * (-2.0 / L.dx[jj] / L.dx[jj]);
                        D[p * L.tot + m] = s * psi_avg(L, i, j, k);  // Elastic.cpp:363
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// a7: Operator<Grid::Node>::Fsmooth (Operator/Operator.cpp:88-154): `sweeps`
// damped-Jacobi sweeps  x <- (1-w) x + w (b - (Ax - D x)) / D  (:109-146);
// the reference does 2 sweeps per call (:107).
// Pinned by tests/test_oracle_mg.py (one sweep vs formula, Jacobi fixed point
// equals a dense solve).
// ---------------------------------------------------------------------------
static void jacobi(const Level& L, int sym, const double* C, const int* bct, const double* D, double omega,
                   int sweeps, double* x, const double* b) {
    std::vector<double> Ax(3 * L.tot);
    for (int s = 0; s < sweeps; s++) {
        fapply(L, sym, C, bct, x, Ax.data());
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < 3 * L.tot; q++) {
            double Rx = Ax[q] - D[q] * x[q];
            x[q] = (1.0 - omega) * x[q] + omega * (b[q] - Rx) / D[q];
        }
    }
}

// ---------------------------------------------------------------------------
// l1-Jacobi smoother (no reference counterpart; a preconditioner choice that leaves the operator
// unchanged): d_(n,q) = sign(a_nq,nq) sum_j |a_nq,j| over the row of node n, component q.  The
// rows are read by probing this file's fapply: the stencil reaches one node in each direction,
// so the probes of the 27 node classes (i mod 3, j mod 3, k mod 3) x 3 components each hit one
// column of every row.  x += (b - A x) / d per sweep (a zero row keeps x).
// Pinned by tests/test_oracle_l1.py (dense brute force of the row sums).
// ---------------------------------------------------------------------------
static void l1_diag(const Level& L, int sym, const double* C, const int* bct, double* d) {
    const int64_t n = 3 * L.tot;
    std::vector<double> e(n), y(n), D(n);
    diagonal(L, sym, C, bct, D.data());
    std::fill(d, d + n, 0.0);
    for (int c = 0; c < 27; c++)
        for (int p = 0; p < 3; p++) {
            std::fill(e.begin(), e.end(), 0.0);
            for (int k = c / 9; k < L.N[2]; k += 3)
                for (int j = (c / 3) % 3; j < L.N[1]; j += 3)
                    for (int i = c % 3; i < L.N[0]; i += 3) e[p * L.tot + L.id(i, j, k)] = 1.0;
            fapply(L, sym, C, bct, e.data(), y.data());
            for (int64_t q = 0; q < n; q++) d[q] += std::fabs(y[q]);
        }
    for (int64_t q = 0; q < n; q++)
        if (D[q] < 0.0) d[q] = -d[q];
}

static void jacobi_l1(const Level& L, int sym, const double* C, const int* bct, const double* d, int sweeps,
                      double* x, const double* b) {
    std::vector<double> Ax(3 * L.tot);
    for (int s = 0; s < sweeps; s++) {
        fapply(L, sym, C, bct, x, Ax.data());
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < 3 * L.tot; q++)
            if (d[q] != 0.0) x[q] += (b[q] - Ax[q]) / d[q];
    }
}

// The same l1 diagonal assembled row by row from the stencils of fapply_node (interior: second
// differences, four-corner mixed differences and central differences of u and C; boundary: the BC
// row with one-sided first derivatives), at the cost of one pass; the CPU baseline's smoother.
// Pinned against l1_diag (probing) by tests/test_oracle_l1.py.
static void l1_diag_rows(const Level& L, int sym, const double* C, const int* bct, double* d) {
    const int nc = ncoef(sym);
    int lo[3] = {0, 0, 0}, hi[3] = {L.n[0], L.n[1], L.n[2]};
    const int sl[3] = {1, 3, 9};
    const int64_t st[3] = {1, L.N[0], (int64_t)L.N[0] * L.N[1]};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                const int64_t m = L.id(i, j, k);
                const double* c = C + m * nc;
                const int idx[3] = {i, j, k};
                for (int q = 0; q < 3; q++) {
                    double acc[27][3] = {{0}};
                    if (L.bnd(i, j, k)) {
                        const int face = classify(i, j, k, lo, hi);
                        double nrm[3];
                        normal_of(face, nrm);
                        auto d1w = [&](int l, double coef, int comp) {
                            if (idx[l] == 0) { acc[13 + sl[l]][comp] += coef / L.dx[l]; acc[13][comp] -= coef / L.dx[l]; }
                            else if (idx[l] == L.n[l]) { acc[13][comp] += coef / L.dx[l]; acc[13 - sl[l]][comp] -= coef / L.dx[l]; }
                            else { acc[13 + sl[l]][comp] += 0.5 * coef / L.dx[l]; acc[13 - sl[l]][comp] -= 0.5 * coef / L.dx[l]; }
                        };
                        const int t = bct[3 * face + q];
                        if (t == 0) acc[13][q] = 1.0;  // Displacement
                        else if (t == 1) {             // Traction: (C : grad u) n
                            for (int kk = 0; kk < 3; kk++)
                                for (int l = 0; l < 3; l++) {
                                    double coef = 0.0;
                                    for (int e = 0; e < 3; e++) coef += nrm[e] * C4(sym, c, q, e, kk, l);
                                    d1w(l, coef, kk);
                                }
                        } else if (t == 3) {  // Neumann: grad u n
                            for (int l = 0; l < 3; l++) d1w(l, nrm[l], q);
                        }
                    } else {
                        for (int kk = 0; kk < 3; kk++)
                            for (int jj = 0; jj < 3; jj++)
                                for (int ll = 0; ll < 3; ll++) {
                                    const double cf = C4(sym, c, q, jj, kk, ll);
                                    if (jj == ll) {
                                        const double w = cf / (L.dx[jj] * L.dx[jj]);
                                        acc[13 + sl[jj]][kk] += w;
                                        acc[13 - sl[jj]][kk] += w;
                                        acc[13][kk] -= 2.0 * w;
                                    } else {
                                        const double w = 0.25 * cf / (L.dx[jj] * L.dx[ll]);
                                        acc[13 + sl[jj] + sl[ll]][kk] += w;
                                        acc[13 - sl[jj] - sl[ll]][kk] += w;
                                        acc[13 + sl[jj] - sl[ll]][kk] -= w;
                                        acc[13 - sl[jj] + sl[ll]][kk] -= w;
                                    }
                                    const double dc = (C4(sym, C + (m + st[jj]) * nc, q, jj, kk, ll) -
                                                       C4(sym, C + (m - st[jj]) * nc, q, jj, kk, ll)) * 0.5 / L.dx[jj];
                                    acc[13 + sl[ll]][kk] += dc * 0.5 / L.dx[ll];
                                    acc[13 - sl[ll]][kk] -= dc * 0.5 / L.dx[ll];
                                }
                    }
                    double sum = 0.0;
                    for (int o = 0; o < 27; o++)
                        for (int kk = 0; kk < 3; kk++) sum += std::fabs(acc[o][kk]);
                    d[q * L.tot + m] = acc[13][q] < 0.0 ? -sum : sum;
                }
            }
}

// a11: solutionResidual / correctionResidual  r = b - A x  (Operator.cpp:635-657)
static void residual(const Level& L, int sym, const double* C, const int* bct, const double* x, const double* b,
                     double* r) {
    fapply(L, sym, C, bct, x, r);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < 3 * L.tot; q++) r[q] = b[q] - r[q];
}

// ---------------------------------------------------------------------------
// a8: Operator<Grid::Node>::restriction (Operator.cpp:262-363): 27-point full
// weighting; on the coarse domain's faces/edges/corners the in-plane 9-point,
// 1-D 3-point and injection stencils (:297-338).  The same weights restrict
// the stiffness (Elastic.cpp:746-836, stride ncomp records of width w).
// Pinned by tests/test_oracle_mg.py (weights, constants, R = P^T/8 on the
// interior).
// ---------------------------------------------------------------------------
static void restrict_fw(const Level& Lc, const Level& Lf, int w, int ncomp_planes, const double* fine,
                        double* crse) {
    // field layout: if ncomp_planes > 0 SoA planes of size tot (w must be 1);
    // otherwise AoS records of width w.
    const int nrec = ncomp_planes > 0 ? ncomp_planes : 1;
    const int64_t sf[3] = {1, Lf.N[0], (int64_t)Lf.N[0] * Lf.N[1]};
#pragma omp parallel for schedule(static)
    for (int K = 0; K < Lc.N[2]; K++)
        for (int J = 0; J < Lc.N[1]; J++)
            for (int I = 0; I < Lc.N[0]; I++) {
                bool bx = I == 0 || I == Lc.n[0], by = J == 0 || J == Lc.n[1], bz = K == 0 || K == Lc.n[2];
                int64_t mf = Lf.id(2 * I, 2 * J, 2 * K), mc = Lc.id(I, J, K);
                for (int rr = 0; rr < nrec; rr++)
                    for (int q = 0; q < w; q++) {
                        const double* F;
                        int64_t stride;
                        if (ncomp_planes > 0) { F = fine + rr * Lf.tot; stride = 1; }
                        else { F = fine + q; stride = w; }
                        auto at = [&](int a, int b, int c) {
                            return F[(mf + a * sf[0] + b * sf[1] + c * sf[2]) * stride];
                        };
                        double v;
                        if (bx && by && bz) v = at(0, 0, 0);
                        else if (by && bz) v = 0.25 * at(-1, 0, 0) + 0.5 * at(0, 0, 0) + 0.25 * at(1, 0, 0);
                        else if (bz && bx) v = 0.25 * at(0, -1, 0) + 0.5 * at(0, 0, 0) + 0.25 * at(0, 1, 0);
                        else if (bx && by) v = 0.25 * at(0, 0, -1) + 0.5 * at(0, 0, 0) + 0.25 * at(0, 0, 1);
                        else if (bx)
                            v = (at(0, -1, -1) + 2.0 * at(0, 0, -1) + at(0, 1, -1) + 2.0 * at(0, -1, 0) +
                                 4.0 * at(0, 0, 0) + 2.0 * at(0, 1, 0) + at(0, -1, 1) + 2.0 * at(0, 0, 1) +
                                 at(0, 1, 1)) / 16.0;
                        else if (by)
                            v = (at(-1, 0, -1) + 2.0 * at(-1, 0, 0) + at(-1, 0, 1) + 2.0 * at(0, 0, -1) +
                                 4.0 * at(0, 0, 0) + 2.0 * at(0, 0, 1) + at(1, 0, -1) + 2.0 * at(1, 0, 0) +
                                 at(1, 0, 1)) / 16.0;
                        else if (bz)
                            v = (at(-1, -1, 0) + 2.0 * at(0, -1, 0) + at(1, -1, 0) + 2.0 * at(-1, 0, 0) +
                                 4.0 * at(0, 0, 0) + 2.0 * at(1, 0, 0) + at(-1, 1, 0) + 2.0 * at(0, 1, 0) +
                                 at(1, 1, 0)) / 16.0;
                        else {
                            double c8 = 0, c12 = 0, c6 = 0;
                            for (int a = -1; a <= 1; a++) for (int b = -1; b <= 1; b++) for (int c = -1; c <= 1; c++) {
                                int nz = (a != 0) + (b != 0) + (c != 0);
                                if (nz == 3) c8 += at(a, b, c);
                                else if (nz == 2) c12 += at(a, b, c);
                                else if (nz == 1) c6 += at(a, b, c);
                            }
                            v = c8 / 64.0 + c12 / 32.0 + c6 / 16.0 + at(0, 0, 0) / 8.0;
                        }
                        if (ncomp_planes > 0) crse[rr * Lc.tot + mc] = v;
                        else crse[mc * w + q] = v;
                    }
            }
}

// a9: Operator<Grid::Node>::interpolation (Operator.cpp:365-434): fine += I(crse)
// with trilinear nodal weights chosen by the parity of (i,j,k) (:399-425).
static void interp_add(const Level& Lc, const Level& Lf, const double* crse, double* fine) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Lf.N[2]; k++)
        for (int j = 0; j < Lf.N[1]; j++)
            for (int i = 0; i < Lf.N[0]; i++) {
                int I = i / 2, J = j / 2, K = k / 2;
                for (int c = 0; c < 3; c++) {
                    const double* C = crse + c * Lc.tot;
                    auto at = [&](int a, int b, int d) { return C[Lc.id(I + a, J + b, K + d)]; };
                    bool ei = i % 2 == 0, ej = j % 2 == 0, ek = k % 2 == 0;
                    double v;
                    if (ei && ej && ek) v = at(0, 0, 0);
                    else if (ej && ek) v = 0.5 * (at(0, 0, 0) + at(1, 0, 0));
                    else if (ek && ei) v = 0.5 * (at(0, 0, 0) + at(0, 1, 0));
                    else if (ei && ej) v = 0.5 * (at(0, 0, 0) + at(0, 0, 1));
                    else if (ei) v = 0.25 * (at(0, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1));
                    else if (ej) v = 0.25 * (at(0, 0, 0) + at(0, 0, 1) + at(1, 0, 0) + at(1, 0, 1));
                    else if (ek) v = 0.25 * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(1, 1, 0));
                    else
                        v = 0.125 * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1) +
                                     at(1, 0, 1) + at(1, 1, 0) + at(1, 1, 1));
                    fine[c * Lf.tot + Lf.id(i, j, k)] += v;
                }
            }
}

// max |x|; NaN if any entry is NaN, so that a diverged solve cannot look converged
static double norm_inf(const double* x, int64_t n) {
    double m = 0.0;
    int bad = 0;
#pragma omp parallel for reduction(max : m) reduction(| : bad)
    for (int64_t q = 0; q < n; q++) {
        if (x[q] != x[q]) bad = 1;
        else m = std::max(m, std::fabs(x[q]));
    }
    return bad ? std::nan("") : m;
}
static double dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (int64_t q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

// ---------------------------------------------------------------------------
// a16: multigrid driver, a restatement of amrex::MLMG (AMReX 26.06, absent
// here; configured by Solver/Nonlocal/Linear.H:255-290) from its documented
// algorithm: F-cycle or V-cycle, nu1/nu2 smoothing calls of Fsmooth, restriction
// of the correction residual, diagonally (normalize) preconditioned BiCGStab
// bottom solve, max-norm stopping test
//   |b - A x|_inf <= max(tol_abs, tol_rel * max(|b|_inf, |r0|_inf)).
// Parity: converged solutions only (the fixed point is A^-1 b whatever the
// cycle); iteration counts unpinned.
// ---------------------------------------------------------------------------
struct Hier {
    int sym;
    const int* bct;
    std::vector<Level> L;
    std::vector<std::vector<double>> C, D, psi, Dl1;
};

// Restriction of psi to the next multigrid level (Operator/Elastic.cpp:839-907): coarse cell
// (I,J,K) from fine cells around (2I,2J,2K) with the nodal full-weighting weights.  The
// branch bounds are the coarse level's nodal box [0, n], which a cell index reaches only on the
// lo side, so only I == 0 (J == 0, K == 0) takes the face / edge / corner stencils.
static void restrict_psi(const Level& Lc, const Level& Lf, const double* fine, double* crse) {
    auto F = [&](int i, int j, int k) { return fine[i + (int64_t)Lf.n[0] * (j + (int64_t)Lf.n[1] * k)]; };
    for (int K = 0; K < Lc.n[2]; K++)
        for (int J = 0; J < Lc.n[1]; J++)
            for (int I = 0; I < Lc.n[0]; I++) {
                const int i = 2 * I, j = 2 * J, k = 2 * K;
                const bool xl = I == 0, yl = J == 0, zl = K == 0;
                double v;
                if (xl && yl && zl) v = F(i, j, k);
                else if (yl && zl) v = F(i - 1, j, k) * 0.25 + F(i, j, k) * 0.5 + F(i + 1, j, k) * 0.25;
                else if (zl && xl) v = F(i, j - 1, k) * 0.25 + F(i, j, k) * 0.5 + F(i, j + 1, k) * 0.25;
                else if (xl && yl) v = F(i, j, k - 1) * 0.25 + F(i, j, k) * 0.5 + F(i, j, k + 1) * 0.25;
                else if (xl)
                    v = (F(i, j - 1, k - 1) + F(i, j, k - 1) * 2.0 + F(i, j + 1, k - 1) + F(i, j - 1, k) * 2.0 +
                         F(i, j, k) * 4.0 + F(i, j + 1, k) * 2.0 + F(i, j - 1, k + 1) + F(i, j, k + 1) * 2.0 +
                         F(i, j + 1, k + 1)) / 16.0;
                else if (yl)
                    v = (F(i - 1, j, k - 1) + F(i - 1, j, k) * 2.0 + F(i - 1, j, k + 1) + F(i, j, k - 1) * 2.0 +
                         F(i, j, k) * 4.0 + F(i, j, k + 1) * 2.0 + F(i + 1, j, k - 1) + F(i + 1, j, k) * 2.0 +
                         F(i + 1, j, k + 1)) / 16.0;
                else if (zl)
                    v = (F(i - 1, j - 1, k) + F(i, j - 1, k) * 2.0 + F(i + 1, j - 1, k) + F(i - 1, j, k) * 2.0 +
                         F(i, j, k) * 4.0 + F(i + 1, j, k) * 2.0 + F(i - 1, j + 1, k) + F(i, j + 1, k) * 2.0 +
                         F(i + 1, j + 1, k)) / 16.0;
                else {
                    double c8 = 0, c12 = 0, c6 = 0;
                    for (int a = -1; a <= 1; a++)
                        for (int b = -1; b <= 1; b++)
                            for (int c = -1; c <= 1; c++) {
                                const int nz = (a != 0) + (b != 0) + (c != 0);
                                if (nz == 3) c8 += F(i + a, j + b, k + c);
                                else if (nz == 2) c12 += F(i + a, j + b, k + c);
                                else if (nz == 1) c6 += F(i + a, j + b, k + c);
                            }
                    v = c8 / 64.0 + c12 / 32.0 + c6 / 16.0 + F(i, j, k) / 8.0;
                }
                crse[I + (int64_t)Lc.n[0] * (J + (int64_t)Lc.n[1] * K)] = v;
            }
}

static void build_hier(Hier& H, const int* ncell, const double* dx, int sym, const int* bct, const double* C0,
                       int max_coarsening, const double* psi0 = nullptr, double psi_small = 1e-8) {
    H.sym = sym;
    H.bct = bct;
    H.L.clear(); H.C.clear(); H.D.clear(); H.psi.clear();
    int n[3] = {ncell[0], ncell[1], ncell[2]};
    double h[3] = {dx[0], dx[1], dx[2]};
    H.L.push_back(Level(n, h));
    // coarsen while every direction stays >= 2 cells (AMReX mg_domain_min_width)
    while ((max_coarsening < 0 || (int)H.L.size() <= max_coarsening) && n[0] % 2 == 0 && n[1] % 2 == 0 &&
           n[2] % 2 == 0 && n[0] / 2 >= 2 && n[1] / 2 >= 2 && n[2] / 2 >= 2) {
        for (int d = 0; d < 3; d++) { n[d] /= 2; h[d] *= 2.0; }
        H.L.push_back(Level(n, h));
    }
    int nc = ncoef(sym);
    H.C.resize(H.L.size());
    H.D.resize(H.L.size());
    H.C[0].assign(C0, C0 + H.L[0].tot * nc);
    for (size_t l = 1; l < H.L.size(); l++) {
        H.C[l].resize(H.L[l].tot * nc);
        restrict_fw(H.L[l], H.L[l - 1], nc, 0, H.C[l - 1].data(), H.C[l].data());   // Elastic.cpp:746-836
    }
    if (psi0) {
        H.psi.resize(H.L.size());
        H.psi[0].assign(psi0, psi0 + (int64_t)H.L[0].n[0] * H.L[0].n[1] * H.L[0].n[2]);
        for (size_t l = 1; l < H.L.size(); l++) {
            H.psi[l].resize((int64_t)H.L[l].n[0] * H.L[l].n[1] * H.L[l].n[2]);
            restrict_psi(H.L[l], H.L[l - 1], H.psi[l - 1].data(), H.psi[l].data());
        }
        for (size_t l = 0; l < H.L.size(); l++) {
            H.L[l].psi = H.psi[l].data();
            H.L[l].psi_small = psi_small;
        }
    }
    for (size_t l = 0; l < H.L.size(); l++) {
        H.D[l].resize(3 * H.L[l].tot);
        diagonal(H.L[l], sym, H.C[l].data(), bct, H.D[l].data());
    }
}

struct MGOpts {
    int cycle;        // 0 = V, 1 = F
    int nu1, nu2;     // smoothing calls
    int sweeps;       // Jacobi sweeps per smoothing call (reference: 2)
    double omega;
    int max_iter;
    int bottom_max_iter;
    double bottom_tol;
    int smoother = 0;       // 0: damped point Jacobi (the reference), 1: l1-Jacobi (Hier::Dl1)
    int sweeps_finest = 0;  // > 0: sweeps per smoothing call on level 0
};

static void apply_l(Hier& H, int l, const double* x, double* y) {
    fapply(H.L[l], H.sym, H.C[l].data(), H.bct, x, y);
}

// diagonally preconditioned BiCGStab (the bottom solver amrex::MLCGSolver
// uses, with Operator<Node>::normalize as the preconditioner, Operator.cpp:156-165)
static int bicgstab(Hier& H, int l, double* x, const double* b, double rtol, int maxit) {
    const Level& L = H.L[l];
    int64_t n = 3 * L.tot;
    const double* D = H.D[l].data();
    std::vector<double> r(n), rh(n), p(n, 0.0), v(n, 0.0), s(n), t(n), ph(n), sh(n);
    apply_l(H, l, x, r.data());
    for (int64_t q = 0; q < n; q++) r[q] = b[q] - r[q];
    rh = r;
    double r0 = norm_inf(r.data(), n);
    if (r0 == 0.0) return 0;
    double rho_1 = 0, alpha = 0, omega = 0;
    int it;
    for (it = 1; it <= maxit; it++) {
        double rho = dot(rh.data(), r.data(), n);
        if (rho == 0.0) break;
        if (it == 1) p = r;
        else {
            double beta = (rho / rho_1) * (alpha / omega);
            for (int64_t q = 0; q < n; q++) p[q] = r[q] + beta * (p[q] - omega * v[q]);
        }
        for (int64_t q = 0; q < n; q++) ph[q] = p[q] / D[q];
        apply_l(H, l, ph.data(), v.data());
        double rv = dot(rh.data(), v.data(), n);
        if (rv == 0.0) break;
        alpha = rho / rv;
        for (int64_t q = 0; q < n; q++) { x[q] += alpha * ph[q]; s[q] = r[q] - alpha * v[q]; }
        if (norm_inf(s.data(), n) <= rtol * r0) { r = s; break; }
        for (int64_t q = 0; q < n; q++) sh[q] = s[q] / D[q];
        apply_l(H, l, sh.data(), t.data());
        double tt = dot(t.data(), t.data(), n);
        if (tt == 0.0) break;
        omega = dot(t.data(), s.data(), n) / tt;
        for (int64_t q = 0; q < n; q++) { x[q] += omega * sh[q]; r[q] = s[q] - omega * t[q]; }
        if (norm_inf(r.data(), n) <= rtol * r0) break;
        if (omega == 0.0) break;
        rho_1 = rho;
    }
    return it;
}

struct MGWork {
    std::vector<std::vector<double>> cor, res, rescor;
};

static void smooth(Hier& H, const MGOpts& o, int l, double* x, const double* b, int ncalls) {
    const int sw = (l == 0 && o.sweeps_finest > 0) ? o.sweeps_finest : o.sweeps;
    for (int c = 0; c < ncalls; c++) {
        if (o.smoother == 1) jacobi_l1(H.L[l], H.sym, H.C[l].data(), H.bct, H.Dl1[l].data(), sw, x, b);
        else jacobi(H.L[l], H.sym, H.C[l].data(), H.bct, H.D[l].data(), o.omega, sw, x, b);
    }
}

static void vcycle(Hier& H, const MGOpts& o, MGWork& W, int top) {
    int bot = (int)H.L.size() - 1;
    for (int l = top; l < bot; l++) {
        std::fill(W.cor[l].begin(), W.cor[l].end(), 0.0);
        smooth(H, o, l, W.cor[l].data(), W.res[l].data(), o.nu1);
        residual(H.L[l], H.sym, H.C[l].data(), H.bct, W.cor[l].data(), W.res[l].data(), W.rescor[l].data());
        restrict_fw(H.L[l + 1], H.L[l], 1, 3, W.rescor[l].data(), W.res[l + 1].data());
    }
    std::fill(W.cor[bot].begin(), W.cor[bot].end(), 0.0);
    bicgstab(H, bot, W.cor[bot].data(), W.res[bot].data(), o.bottom_tol, o.bottom_max_iter);
    for (int l = bot - 1; l >= top; l--) {
        interp_add(H.L[l + 1], H.L[l], W.cor[l + 1].data(), W.cor[l].data());
        smooth(H, o, l, W.cor[l].data(), W.res[l].data(), o.nu2);
    }
}

static void fcycle(Hier& H, const MGOpts& o, MGWork& W) {
    int bot = (int)H.L.size() - 1;
    for (int l = 1; l <= bot; l++) restrict_fw(H.L[l], H.L[l - 1], 1, 3, W.res[l - 1].data(), W.res[l].data());
    std::fill(W.cor[bot].begin(), W.cor[bot].end(), 0.0);
    bicgstab(H, bot, W.cor[bot].data(), W.res[bot].data(), o.bottom_tol, o.bottom_max_iter);
    for (int l = bot - 1; l >= 0; l--) {
        std::fill(W.cor[l].begin(), W.cor[l].end(), 0.0);
        interp_add(H.L[l + 1], H.L[l], W.cor[l + 1].data(), W.cor[l].data());
        // rescor = res - L(cor); res = rescor; hold cor; V-cycle; cor += hold
        residual(H.L[l], H.sym, H.C[l].data(), H.bct, W.cor[l].data(), W.res[l].data(), W.rescor[l].data());
        W.res[l] = W.rescor[l];
        std::vector<double> hold = W.cor[l];
        vcycle(H, o, W, l);
        for (size_t q = 0; q < hold.size(); q++) W.cor[l][q] += hold[q];
    }
}

static int mg_solve(Hier& H, const MGOpts& o, double* x, const double* b, double tol_rel, double tol_abs,
                    double* hist, int hist_cap, double* final_res) {
    const Level& L = H.L[0];
    int64_t n = 3 * L.tot;
    MGWork W;
    W.cor.resize(H.L.size()); W.res.resize(H.L.size()); W.rescor.resize(H.L.size());
    for (size_t l = 0; l < H.L.size(); l++) {
        W.cor[l].assign(3 * H.L[l].tot, 0.0);
        W.res[l].assign(3 * H.L[l].tot, 0.0);
        W.rescor[l].assign(3 * H.L[l].tot, 0.0);
    }
    residual(L, H.sym, H.C[0].data(), H.bct, x, b, W.res[0].data());
    double r0 = norm_inf(W.res[0].data(), n), bn = norm_inf(b, n);
    double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rn = r0;
    int it = 0;
    if (hist && hist_cap > 0) hist[0] = r0;
    while (rn > target && it < o.max_iter) {
        if (o.cycle == 1) fcycle(H, o, W);
        else vcycle(H, o, W, 0);
        for (int64_t q = 0; q < n; q++) x[q] += W.cor[0][q];
        residual(L, H.sym, H.C[0].data(), H.bct, x, b, W.res[0].data());
        rn = norm_inf(W.res[0].data(), n);
        it++;
        if (hist && it < hist_cap) hist[it] = rn;
    }
    if (final_res) *final_res = rn;
    return rn <= target ? it : -it;
}

// ---------------------------------------------------------------------------
// Flexible GMRES(m) right-preconditioned by one V-cycle from zero (no reference counterpart: the
// CPU baseline of the library's Krylov solver).  Modified Gram-Schmidt in FP64; a restart cycle
// ends early when its least-squares residual estimate, scaled by the max/2-norm ratio of the
// restart's true residual, reaches the target; convergence is decided by the true residual's max
// norm at the restarts, as in the library.
// ---------------------------------------------------------------------------
static int fgmres(Hier& H, const MGOpts& o, int m, double* x, const double* b, double tol_rel, double tol_abs,
                  double* final_res) {
    const Level& L = H.L[0];
    const int64_t n = 3 * L.tot;
    MGWork W;
    W.cor.resize(H.L.size()); W.res.resize(H.L.size()); W.rescor.resize(H.L.size());
    for (size_t l = 0; l < H.L.size(); l++) {
        W.cor[l].assign(3 * H.L[l].tot, 0.0);
        W.res[l].assign(3 * H.L[l].tot, 0.0);
        W.rescor[l].assign(3 * H.L[l].tot, 0.0);
    }
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), Z(m, std::vector<double>(n));
    std::vector<double> r(n);
    residual(L, H.sym, H.C[0].data(), H.bct, x, b, r.data());
    const double r0 = norm_inf(r.data(), n), bn = norm_inf(b, n);
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rinf = r0, beta = std::sqrt(dot(r.data(), r.data(), n));
    std::vector<std::vector<double>> Hm(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), gv(m + 1), y(m);
    int it = 0;
    while (rinf > target && it < o.max_iter && std::isfinite(rinf) && beta > 0.0) {
        const double ratio = rinf / beta;
        for (int64_t q = 0; q < n; q++) V[0][q] = r[q] / beta;
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int jend = 0;
        for (int j = 0; j < m; j++) {
            W.res[0] = V[j];
            if (o.cycle == 1) fcycle(H, o, W);
            else vcycle(H, o, W, 0);
            Z[j] = W.cor[0];
            apply_l(H, 0, Z[j].data(), V[j + 1].data());
            for (int i = 0; i <= j; i++) {
                const double hij = dot(V[j + 1].data(), V[i].data(), n);
                Hm[i][j] = hij;
#pragma omp parallel for schedule(static)
                for (int64_t q = 0; q < n; q++) V[j + 1][q] -= hij * V[i][q];
            }
            const double hn = std::sqrt(dot(V[j + 1].data(), V[j + 1].data(), n));
            Hm[j + 1][j] = hn;
            if (hn > 0.0)
#pragma omp parallel for schedule(static)
                for (int64_t q = 0; q < n; q++) V[j + 1][q] /= hn;
            for (int i = 0; i < j; i++) {
                const double a = cs[i] * Hm[i][j] + sn[i] * Hm[i + 1][j];
                Hm[i + 1][j] = -sn[i] * Hm[i][j] + cs[i] * Hm[i + 1][j];
                Hm[i][j] = a;
            }
            const double den = std::hypot(Hm[j][j], Hm[j + 1][j]);
            cs[j] = den > 0 ? Hm[j][j] / den : 1.0;
            sn[j] = den > 0 ? Hm[j + 1][j] / den : 0.0;
            Hm[j][j] = den;
            Hm[j + 1][j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            it++;
            jend = j + 1;
            if (std::fabs(gv[j + 1]) * ratio <= target || it >= o.max_iter || hn == 0.0 || !std::isfinite(hn)) break;
        }
        for (int i = jend - 1; i >= 0; i--) {
            double sum = gv[i];
            for (int k = i + 1; k < jend; k++) sum -= Hm[i][k] * y[k];
            y[i] = Hm[i][i] != 0.0 ? sum / Hm[i][i] : 0.0;
        }
        for (int i = 0; i < jend; i++)
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < n; q++) x[q] += y[i] * Z[i][q];
        residual(L, H.sym, H.C[0].data(), H.bct, x, b, r.data());
        rinf = norm_inf(r.data(), n);
        beta = std::sqrt(dot(r.data(), r.data(), n));
    }
    if (final_res) *final_res = rinf;
    return rinf <= target ? it : -it;
}

// ---------------------------------------------------------------------------
// a17: Solver::Nonlocal::Newton<T>::prepareForSolve, vector path
// (Solver/Nonlocal/Newton.H:144-231; with L.psi set, the psi variant :234-266) for a linear or affine model with
// DW = C : (grad u - F0) (Model/Solid/Affine/Isotropic.H:38-49,
// Model/Solid/Linear/Cubic.H:34-41):
//   boundary: rhs = b - BC(u, grad u, DW)      (:220-225)
//   interior: rhs = b - Divergence(DW)         (:226-230, Numeric/Stencil.H:588-657)
// F0 may be null (linear model).  Pinned by tests/test_oracle_newton.py.
// ---------------------------------------------------------------------------
static void newton_rhs(const Level& L, int sym, const double* C, const double* F0, const int* bct, const double* u,
                       const double* b, double* rhs, double* DWout) {
    const int nc = ncoef(sym);
    std::vector<double> DW(9 * L.tot);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                int64_t m = L.id(i, j, k);
                double G[3][3], S[3][3];
                grad_u(L, u, i, j, k, G);
                if (F0) for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) G[a][c] -= F0[9 * m + 3 * a + c];
                c_dot_mat(sym, C + m * nc, G, S);
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) DW[9 * m + 3 * a + c] = S[a][c];
            }
    int lo[3] = {0, 0, 0}, hi[3] = {L.n[0], L.n[1], L.n[2]};
    const int64_t st[3] = {1, L.N[0], (int64_t)L.N[0] * L.N[1]};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                int64_t m = L.id(i, j, k);
                double f[3];
                if (L.bnd(i, j, k)) {
                    // with psi the boundary rows keep psiavg = 1: the node is on the box's edge (:247-252)
                    double G[3][3], S[3][3], uu[3];
                    grad_u(L, u, i, j, k, G);
                    for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) S[a][c] = DW[9 * m + 3 * a + c];
                    for (int p = 0; p < 3; p++) uu[p] = u[p * L.tot + m];
                    bc_row(bct, classify(i, j, k, lo, hi), uu, G, S, f);
                } else {
                    for (int a = 0; a < 3; a++) {
                        double s = 0.0;
                        for (int d = 0; d < 3; d++)
                            s += (DW[9 * (m + st[d]) + 3 * a + d] - DW[9 * (m - st[d]) + 3 * a + d]) / 2. / L.dx[d];
                        f[a] = s;
                    }
                    if (L.psi) {
                        // divdw *= CellToNodeAverage(psi); divdw += dw grad psi  (Newton.H:258-263; no psi_small)
                        const double pm = psi_node_mean(L, i, j, k);
                        double gp[3];
                        psi_grad(L, i, j, k, gp);
                        for (int a = 0; a < 3; a++) {
                            double s = 0.0;
                            for (int d = 0; d < 3; d++) s += DW[9 * m + 3 * a + d] * gp[d];
                            f[a] = f[a] * pm + s;
                        }
                    }
                }
                for (int p = 0; p < 3; p++) rhs[p * L.tot + m] = b[p * L.tot + m] - f[p];
            }
    if (DWout) std::copy(DW.begin(), DW.end(), DWout);
}

// ---------------------------------------------------------------------------
// Elastic::Strain / Stress / Energy (Operator/Elastic.cpp:430-603) on every node with the
// GetStencil gradient: kind 0 eps = sym(grad u), 1 sig = (C : grad u) psi_avg, 2 energy +=
// (C : grad u) : eps / 2.  Tensors row-major (9 planes) or Voigt 00,11,22,12,20,01 (6 planes).
// Pinned by tests/test_oracle_post.py (linear fields, closed forms).
// ---------------------------------------------------------------------------
static void post(const Level& L, int sym, const double* C, const double* u, int kind, int voigt, double* out) {
    const int nc = ncoef(sym);
    const int va[6] = {0, 1, 2, 1, 2, 0}, vb[6] = {0, 1, 2, 2, 0, 1};
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                const int64_t m = L.id(i, j, k);
                double G[3][3], T[3][3], E[3][3];
                grad_u(L, u, i, j, k, G);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) E[a][b] = 0.5 * (G[a][b] + G[b][a]);
                if (kind == 0) {
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) T[a][b] = E[a][b];
                } else {
                    c_dot_mat(sym, C + m * nc, G, T);
                    if (kind == 2) {
                        double w = 0.0;
                        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) w += 0.5 * T[a][b] * E[a][b];
                        out[m] += w;
                        continue;
                    }
                    const double pa = psi_avg(L, i, j, k);
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) T[a][b] *= pa;
                }
                if (voigt) for (int q = 0; q < 6; q++) out[q * L.tot + m] = T[va[q]][vb[q]];
                else for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) out[(3 * a + b) * L.tot + m] = T[a][b];
            }
}

// Solver::Nonlocal::Newton::DW / W (Solver/Nonlocal/Newton.H:528-611) for an affine model
// (Model/Solid/Affine/Isotropic.H:38-45, kinematic variable grad u): sig = C:(grad u - F0),
// written as 9 values per node row-major (kind 3), and w = (1/2) tr((grad u - F0)^T sig)
// (kind 4), with the GetStencil gradients (one-sided on the faces).  F0 holds 9 values per node
// (may be null: F0 = 0).
static void model_post(const Level& L, int sym, const double* C, const double* F0, const double* u, int kind,
                       double* out) {
    const int nc = ncoef(sym);
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                const int64_t m = L.id(i, j, k);
                double G[3][3], S[3][3];
                grad_u(L, u, i, j, k, G);
                if (F0)
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) G[a][b] -= F0[9 * m + 3 * a + b];
                c_dot_mat(sym, C + m * nc, G, S);
                if (kind == 3) {
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) out[(3 * a + b) * L.tot + m] = S[a][b];
                } else {
                    double w = 0.0;
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) w += G[a][b] * S[a][b];
                    out[m] = 0.5 * w;
                }
            }
}

// BC::Operator::Elastic::Constant::Init (BC/Operator/Elastic/Constant.H:197-286),
// non-periodic: boundary nodes of b take the prescribed values.
static void bc_init(const Level& L, const int* bct, const double* bcv, int homogeneous, double* b) {
    int lo[3] = {0, 0, 0}, hi[3] = {L.n[0], L.n[1], L.n[2]};
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                int f = classify(i, j, k, lo, hi);
                if (f == INT) continue;
                for (int c = 0; c < 3; c++) {
                    double v = (homogeneous && bct[3 * f + c] == DISPLACEMENT) ? 0.0 : bcv[3 * f + c];
                    b[c * L.tot + L.id(i, j, k)] = v;
                }
            }
}

// Stress on nodes after the solve (Integrator/Base/Mechanics.H:211-248) and the
// xhi-face traction integral thermo value trac_xhi (:391-448).
static void trac_xhi(const Level& L, int sym, const double* C, const double* F0, const double* u, double* out3) {
    const int nc = ncoef(sym);
    double t[3] = {0, 0, 0};
    double da = L.dx[1] * L.dx[2];
    int i = L.n[0];
    auto sig = [&](int ii, int jj, int kk, double S[3][3]) {
        double G[3][3];
        grad_u(L, u, ii, jj, kk, G);
        int64_t m = L.id(ii, jj, kk);
        if (F0) for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) G[a][c] -= F0[9 * m + 3 * a + c];
        c_dot_mat(sym, C + m * nc, G, S);
    };
    for (int k = 0; k < L.n[2]; k++)
        for (int j = 0; j < L.n[1]; j++) {
            double S0[3][3], S1[3][3], S2_[3][3], S3_[3][3];
            sig(i, j, k, S0); sig(i, j + 1, k, S1); sig(i, j, k + 1, S2_); sig(i, j + 1, k + 1, S3_);
            for (int a = 0; a < 3; a++) t[a] += 0.25 * (S0[a][0] + S1[a][0] + S2_[a][0] + S3_[a][0]) * da;
        }
    for (int a = 0; a < 3; a++) out3[a] = t[a];
}

}  // namespace ora

// ===========================================================================
// C ABI for the tests (ctypes) and the bench's cpu_baseline leg.
// ===========================================================================
using namespace ora;
extern "C" {

int ora_ncoef(int sym) { return ncoef(sym); }

int ora_mm_slot(int i, int j, int k, int l) { return mm_slot(i, j, k, l); }
double ora_c4(int sym, const double* c, int i, int j, int k, int l) { return C4(sym, c, i, j, k, l); }
void ora_cubic(double C11, double C12, double C44, double phi1, double Phi, double phi2, double* out21) {
    double R[9];
    euler_R(phi1, Phi, phi2, R);
    cubic_mm(C11, C12, C44, R, out21);
}
void ora_cubic_R(double C11, double C12, double C44, const double* R, double* out21) { cubic_mm(C11, C12, C44, R, out21); }
void ora_euler_R(double phi1, double Phi, double phi2, double* R) { euler_R(phi1, Phi, phi2, R); }
int ora_classify(int i, int j, int k, const int* lo, const int* hi) { return classify(i, j, k, lo, hi); }
void ora_normal(int face, double* n) { normal_of(face, n); }

void ora_fapply(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* u,
                double* f) {
    Level L(ncell, dx);
    fapply(L, sym, C, bct, u, f);
}
void ora_diagonal(const int* ncell, const double* dx, int sym, const double* C, const int* bct, double* D) {
    Level L(ncell, dx);
    diagonal(L, sym, C, bct, D);
}
void ora_jacobi(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* D,
                double omega, int sweeps, double* x, const double* b) {
    Level L(ncell, dx);
    jacobi(L, sym, C, bct, D, omega, sweeps, x, b);
}
void ora_l1_diag(const int* ncell, const double* dx, int sym, const double* C, const int* bct, double* d) {
    Level L(ncell, dx);
    l1_diag(L, sym, C, bct, d);
}
void ora_jacobi_l1(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* d,
                   int sweeps, double* x, const double* b) {
    Level L(ncell, dx);
    jacobi_l1(L, sym, C, bct, d, sweeps, x, b);
}
void ora_residual(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* x,
                  const double* b, double* r) {
    Level L(ncell, dx);
    residual(L, sym, C, bct, x, b, r);
}
// restriction of a 3-plane SoA vector (w=1, planes=3) or of an AoS record field (w=ncoef, planes=0)
void ora_restrict(const int* ncell_fine, const double* dx_fine, int w, int planes, const double* fine, double* crse) {
    Level Lf(ncell_fine, dx_fine);
    int nc[3] = {ncell_fine[0] / 2, ncell_fine[1] / 2, ncell_fine[2] / 2};
    double h[3] = {2 * dx_fine[0], 2 * dx_fine[1], 2 * dx_fine[2]};
    Level Lc(nc, h);
    restrict_fw(Lc, Lf, w, planes, fine, crse);
}
void ora_interp_add(const int* ncell_fine, const double* dx_fine, const double* crse, double* fine) {
    Level Lf(ncell_fine, dx_fine);
    int nc[3] = {ncell_fine[0] / 2, ncell_fine[1] / 2, ncell_fine[2] / 2};
    double h[3] = {2 * dx_fine[0], 2 * dx_fine[1], 2 * dx_fine[2]};
    Level Lc(nc, h);
    interp_add(Lc, Lf, crse, fine);
}
// returns #iterations (negative if not converged); hist receives max-norm residuals
int ora_mg_solve(const int* ncell, const double* dx, int sym, const double* C, const int* bct, double* x,
                 const double* b, double tol_rel, double tol_abs, int cycle, int nu1, int nu2, int sweeps,
                 double omega, int max_iter, int max_coarsening, double* hist, int hist_cap, double* final_res,
                 int* nlevels) {
    Hier H;
    build_hier(H, ncell, dx, sym, bct, C, max_coarsening);
    if (nlevels) *nlevels = (int)H.L.size();
    MGOpts o{cycle, nu1, nu2, sweeps, omega, max_iter, 200, 1e-4};
    return mg_solve(H, o, x, b, tol_rel, tol_abs, hist, hist_cap, final_res);
}
// Galerkin coarse operators A_l = R A_{l-1} P (no reference counterpart; the library's coarse_op
// GALERKIN), read off by probing the composition restrict_fw(A_{l-1}(interp_add(e))): the image
// of a coarse unit vector reaches the coarse nodes within one node of it, so the 27 node classes
// (I mod 3, J mod 3, K mod 3) x 3 components give every entry.  Level 0's A is fapply; deeper
// levels apply the stencil of the level above.  Stencil layout: plane 9 o + 3 q + p (o = (dx+1) +
// 3 (dy+1) + 9 (dz+1)) of tot values, the coefficient of u_p at node + o in row q.
// Pinned by tests/test_oracle_galerkin.py (stencil products equal the composition on random
// vectors; interior rows of constant-coefficient operators annihilate linear fields).
static void stencil_apply(const Level& L, const double* S, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                const int64_t m = L.id(i, j, k);
                double f[3] = {0, 0, 0};
                for (int o = 0; o < 27; o++) {
                    const int ii = i + o % 3 - 1, jj = j + (o / 3) % 3 - 1, kk = k + o / 9 - 1;
                    if (ii < 0 || jj < 0 || kk < 0 || ii > L.n[0] || jj > L.n[1] || kk > L.n[2]) continue;
                    const int64_t mn = L.id(ii, jj, kk);
                    for (int q = 0; q < 3; q++)
                        for (int pp = 0; pp < 3; pp++) f[q] += S[(9 * o + 3 * q + pp) * L.tot + m] * x[pp * L.tot + mn];
                }
                for (int q = 0; q < 3; q++) y[q * L.tot + m] = f[q];
            }
}

static void galerkin_levels(Hier& H, int upto, std::vector<std::vector<double>>& S) {
    S.assign(H.L.size(), {});
    for (int l = 1; l <= upto && l < (int)H.L.size(); l++) {
        const Level& Lc = H.L[l];
        const Level& Lf = H.L[l - 1];
        S[l].assign((size_t)243 * Lc.tot, 0.0);
        std::vector<double> e(3 * Lc.tot), pf(3 * Lf.tot), af(3 * Lf.tot), y(3 * Lc.tot);
        for (int c = 0; c < 27; c++)
            for (int p = 0; p < 3; p++) {
                std::fill(e.begin(), e.end(), 0.0);
                for (int K = c / 9; K < Lc.N[2]; K += 3)
                    for (int J = (c / 3) % 3; J < Lc.N[1]; J += 3)
                        for (int I = c % 3; I < Lc.N[0]; I += 3) e[p * Lc.tot + Lc.id(I, J, K)] = 1.0;
                std::fill(pf.begin(), pf.end(), 0.0);
                interp_add(Lc, Lf, e.data(), pf.data());
                if (l == 1) fapply(Lf, H.sym, H.C[0].data(), H.bct, pf.data(), af.data());
                else stencil_apply(Lf, S[l - 1].data(), pf.data(), af.data());
                restrict_fw(Lc, Lf, 1, 3, af.data(), y.data());
                const int cc[3] = {c % 3, (c / 3) % 3, c / 9};
                for (int K = 0; K < Lc.N[2]; K++)
                    for (int J = 0; J < Lc.N[1]; J++)
                        for (int I = 0; I < Lc.N[0]; I++) {
                            const int idx[3] = {I, J, K};
                            int off[3];
                            bool ok = true;
                            for (int d = 0; d < 3; d++) {
                                // the node of class cc[d] among idx - 1 .. idx + 1
                                off[d] = -2;
                                for (int t = -1; t <= 1; t++)
                                    if (((idx[d] + t) % 3 + 3) % 3 == cc[d]) off[d] = t;
                                const int nb = idx[d] + off[d];
                                ok = ok && nb >= 0 && nb <= Lc.n[d];
                            }
                            if (!ok) continue;
                            const int o = (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1);
                            const int64_t m = Lc.id(I, J, K);
                            for (int q = 0; q < 3; q++) S[l][(size_t)(9 * o + 3 * q + p) * Lc.tot + m] = y[q * Lc.tot + m];
                        }
            }
    }
}

// ---------------------------------------------------------------------------
// Composite multi-level (AMR) system and solve, the library's alamo_amr_* (no single reference
// function: Linear::solve over several AMR levels, Linear.H:131-168, drives amrex::MLMG's composite
// cycle with Operator<Grid::Node>::reflux, Operator.cpp:526-633, and averageDownSolutionRHS,
// :435-447; AMReX is not available, so the C/F treatment is this restatement's own, identical
// to the library's).  Every level is a full-domain field at resolution n_cell 2^l:
//   * unknowns: level 0 and the nodes strictly inside each fine region (region mask 2); fine
//     nodes on or outside the region boundary follow the coarser level, u_l = P u_{l-1}; coarse
//     nodes covered by the next level (nodal fine mask 1, 2) hold the coincident fine values;
//   * level l rows (fapply) on level 0 and on each fine region's interior, the coarse/fine
//     interface nodes (fine mask 1) keeping their coarse rows;
//   * strictly covered coarse nodes (fine mask 2) take the full-weighting restriction of the finer
//     residual (the reflux of the correction scheme).
// The solve: composite V-cycles (masked smoothing and reflux down the levels, the level-0
// multigrid V-cycle, prolongation and smoothing up), then covered coarse nodes take the fine
// solution.  Pinned by tests/test_oracle_amr_composite.py (the reference's UniaxialTension
// closed form: linear fields are reproduced exactly; the TrigTest closed form to O(h^2)).
// ---------------------------------------------------------------------------
extern "C" void oamr_fine_mask(const int* nc, int nbox, const int* flo, const int* fhi, int* mask);

struct Composite {
    int nl = 0;
    std::vector<Hier> H;
    std::vector<std::vector<int>> region, fmask;
    MGOpts o{};
};

static void composite_build(Composite& A, const int* ncell0, const double* dx0, int sym, const int* bct, int nl,
                            const double* const* C, const int* nbox, const int* box_lo, const int* box_hi,
                            const MGOpts& o) {
    A.nl = nl;
    A.o = o;
    A.H.resize(nl);
    A.region.resize(nl);
    A.fmask.resize(nl);
    std::vector<int> off(nl + 1, 0);
    for (int l = 1; l < nl; l++) off[l + 1] = off[l] + nbox[l - 1];
    for (int l = 0; l < nl; l++) {
        int nc[3];
        double h[3];
        for (int d = 0; d < 3; d++) { nc[d] = ncell0[d] << l; h[d] = dx0[d] / (double)(1 << l); }
        build_hier(A.H[l], nc, h, sym, bct, C[l], l == 0 ? -1 : 0);
        if (o.smoother == 1) {
            A.H[l].Dl1.resize(A.H[l].L.size());
            for (size_t m = 0; m < A.H[l].L.size(); m++) {
                A.H[l].Dl1[m].resize(3 * A.H[l].L[m].tot);
                l1_diag_rows(A.H[l].L[m], sym, A.H[l].C[m].data(), bct, A.H[l].Dl1[m].data());
            }
        }
        const int64_t tot = A.H[l].L[0].tot;
        if (l > 0) {
            // the level's own boxes, written as the boxes of a virtual level above it
            std::vector<int> flo(3 * nbox[l - 1]), fhi(3 * nbox[l - 1]);
            for (int t = 0; t < nbox[l - 1]; t++)
                for (int d = 0; d < 3; d++) {
                    flo[3 * t + d] = 2 * box_lo[3 * (off[l] + t) + d];
                    fhi[3 * t + d] = 2 * box_hi[3 * (off[l] + t) + d] + 1;
                }
            A.region[l].resize(tot);
            oamr_fine_mask(nc, nbox[l - 1], flo.data(), fhi.data(), A.region[l].data());
        }
        if (l < nl - 1) {
            A.fmask[l].resize(tot);
            oamr_fine_mask(nc, nbox[l], box_lo + 3 * off[l + 1], box_hi + 3 * off[l + 1], A.fmask[l].data());
        }
    }
}

static void masked_copy(const std::vector<int>& mask, int lo, int hi, int64_t n, double* dst, const double* src) {
    for (int64_t q = 0; q < n; q++)
        if (mask[q] >= lo && mask[q] <= hi)
            for (int c = 0; c < 3; c++) dst[c * n + q] = src[c * n + q];
}
static void masked_zero(const std::vector<int>& mask, int lo, int hi, int64_t n, double* x) {
    for (int64_t q = 0; q < n; q++)
        if (mask[q] < lo || mask[q] > hi)
            for (int c = 0; c < 3; c++) x[c * n + q] = 0.0;
}

// covered coarse nodes (nodal fine mask 1, 2) take the coincident fine values, finest first
static void composite_inject(Composite& A, double* const* u) {
    for (int l = A.nl - 2; l >= 0; l--) {
        const Level& Lc = A.H[l].L[0];
        const Level& Lf = A.H[l + 1].L[0];
        for (int K = 0; K < Lc.N[2]; K++)
            for (int J = 0; J < Lc.N[1]; J++)
                for (int I = 0; I < Lc.N[0]; I++) {
                    const int64_t m = Lc.id(I, J, K);
                    if (A.fmask[l][m] < 1) continue;
                    const int64_t mf = Lf.id(2 * I, 2 * J, 2 * K);
                    for (int q = 0; q < 3; q++) u[l][q * Lc.tot + m] = u[l + 1][q * Lf.tot + mf];
                }
    }
}

// e~ = the fine field injected at the nodes strictly covered (fine mask 2), zero elsewhere
static void inject_strict(const Level& Lc, const Level& Lf, const std::vector<int>& fm, const double* fine, double* out) {
    std::fill(out, out + 3 * Lc.tot, 0.0);
    for (int K = 0; K < Lc.N[2]; K++)
        for (int J = 0; J < Lc.N[1]; J++)
            for (int I = 0; I < Lc.N[0]; I++) {
                const int64_t m = Lc.id(I, J, K);
                if (fm[m] != 2) continue;
                const int64_t mf = Lf.id(2 * I, 2 * J, 2 * K);
                for (int q = 0; q < 3; q++) out[q * Lc.tot + m] = fine[q * Lf.tot + mf];
            }
}

// covered coarse nodes first take the fine values; then r_l = b_l - A_l u_l on level 0 and on each
// fine region's interior (zero elsewhere), the strictly covered nodes replaced by the restriction of
// r_{l+1}
static void composite_residual(Composite& A, double* const* u, const double* const* b, double* const* r) {
    composite_inject(A, u);
    for (int l = A.nl - 1; l >= 0; l--) {
        const Level& L = A.H[l].L[0];
        residual(L, A.H[l].sym, A.H[l].C[0].data(), A.H[l].bct, u[l], b[l], r[l]);
        if (l > 0) masked_zero(A.region[l], 2, 2, L.tot, r[l]);
        if (l < A.nl - 1) {
            std::vector<double> t(3 * L.tot);
            restrict_fw(L, A.H[l + 1].L[0], 1, 3, r[l + 1], t.data());
            masked_copy(A.fmask[l], 2, 2, L.tot, r[l], t.data());
        }
    }
}

static void composite_smooth(Composite& A, int l, double* e, const double* c, int nsweeps) {
    const Level& L = A.H[l].L[0];
    std::vector<double> t(e, e + 3 * L.tot);
    for (int s = 0; s < nsweeps; s++) {
        if (A.o.smoother == 1) jacobi_l1(L, A.H[l].sym, A.H[l].C[0].data(), A.H[l].bct, A.H[l].Dl1[0].data(), 1, t.data(), c);
        else jacobi(L, A.H[l].sym, A.H[l].C[0].data(), A.H[l].bct, A.H[l].D[0].data(), A.o.omega, 1, t.data(), c);
        masked_copy(A.region[l], 2, 2, L.tot, e, t.data());
        std::copy(e, e + 3 * L.tot, t.begin());
    }
}

static int composite_solve(Composite& A, double* const* u, const double* const* b, double tol_rel, double tol_abs,
                           double* final_res) {
    const int nl = A.nl;
    std::vector<std::vector<double>> r(nl), c(nl), e(nl), rr(nl);
    std::vector<double*> rp(nl);
    for (int l = 0; l < nl; l++) {
        const int64_t n = 3 * A.H[l].L[0].tot;
        r[l].assign(n, 0.0); c[l].assign(n, 0.0); e[l].assign(n, 0.0); rr[l].assign(n, 0.0);
        rp[l] = r[l].data();
    }
    for (int l = 1; l < nl; l++) {
        const int64_t n = A.H[l].L[0].tot;
        std::vector<double> t(3 * n, 0.0);
        interp_add(A.H[l - 1].L[0], A.H[l].L[0], u[l - 1], t.data());
        masked_copy(A.region[l], 0, 1, n, u[l], t.data());
    }
    double bn = 0.0;
    for (int l = 0; l < nl; l++) {
        const int64_t n = A.H[l].L[0].tot;
        std::vector<double> t(b[l], b[l] + 3 * n);
        if (l > 0) masked_zero(A.region[l], 2, 2, n, t.data());
        bn = std::max(bn, norm_inf(t.data(), 3 * n));
    }
    auto rnorm = [&]() {
        double m = 0.0;
        for (int l = 0; l < nl; l++) m = std::max(m, norm_inf(r[l].data(), 3 * A.H[l].L[0].tot));
        return m;
    };
    composite_residual(A, u, b, rp.data());
    const double r0 = rnorm();
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * std::max(r0, bn));
    double rn = r0;
    int it = 0;
    MGWork W;
    W.cor.resize(A.H[0].L.size()); W.res.resize(A.H[0].L.size()); W.rescor.resize(A.H[0].L.size());
    for (size_t m = 0; m < A.H[0].L.size(); m++) {
        W.cor[m].assign(3 * A.H[0].L[m].tot, 0.0);
        W.res[m].assign(3 * A.H[0].L[m].tot, 0.0);
        W.rescor[m].assign(3 * A.H[0].L[m].tot, 0.0);
    }
    const int nsw = A.o.nu1 * A.o.sweeps, nsw2 = A.o.nu2 * A.o.sweeps;
    while (rn > target && it < A.o.max_iter && std::isfinite(rn)) {
        const int L = nl - 1;
        for (int l = L; l >= 0; l--) {
            const Level& Lv = A.H[l].L[0];
            c[l] = r[l];
            if (l < L) {
                // strictly covered nodes: the restricted residual of the finer correction equation;
                // coarse/fine nodes: their coarse rows less the effect of the finer correction
                // injected at the strictly covered nodes
                std::vector<double> t(3 * Lv.tot), et(3 * Lv.tot);
                restrict_fw(Lv, A.H[l + 1].L[0], 1, 3, rr[l + 1].data(), t.data());
                masked_copy(A.fmask[l], 2, 2, Lv.tot, c[l].data(), t.data());
                inject_strict(Lv, A.H[l + 1].L[0], A.fmask[l], e[l + 1].data(), et.data());
                residual(Lv, A.H[l].sym, A.H[l].C[0].data(), A.H[l].bct, et.data(), c[l].data(), t.data());
                masked_copy(A.fmask[l], 1, 1, Lv.tot, c[l].data(), t.data());
            }
            if (l == 0) break;
            std::fill(e[l].begin(), e[l].end(), 0.0);
            composite_smooth(A, l, e[l].data(), c[l].data(), nsw);
            residual(Lv, A.H[l].sym, A.H[l].C[0].data(), A.H[l].bct, e[l].data(), c[l].data(), rr[l].data());
            masked_zero(A.region[l], 2, 2, Lv.tot, rr[l].data());
        }
        W.res[0] = c[0];
        if (A.o.cycle == 1) fcycle(A.H[0], A.o, W);
        else vcycle(A.H[0], A.o, W, 0);
        e[0] = W.cor[0];
        for (int l = 1; l < nl; l++) {
            interp_add(A.H[l - 1].L[0], A.H[l].L[0], e[l - 1].data(), e[l].data());
            composite_smooth(A, l, e[l].data(), c[l].data(), nsw2);
        }
        for (int l = 0; l < nl; l++)
            for (size_t q = 0; q < e[l].size(); q++) u[l][q] += e[l][q];
        composite_residual(A, u, b, rp.data());
        rn = rnorm();
        it++;
    }
    composite_inject(A, u);
    if (final_res) *final_res = rn;
    return rn <= target ? it : -it;
}

// FGMRES(restart) with one V-cycle (nu1 / nu2 smoothing calls of `sweeps` sweeps, `sweeps_finest` on
// level 0 when > 0; smoother 0 damped point Jacobi, 1 l1-Jacobi; BiCGStab bottom to bottom_tol) as the
// right preconditioner.  Returns the iterations, negative when max_iter ended the solve.
int ora_krylov_solve(const int* ncell, const double* dx, int sym, const double* C, const int* bct, double* x,
                     const double* b, double tol_rel, double tol_abs, int restart, int max_iter, int nu1, int nu2,
                     int sweeps, int sweeps_finest, int smoother, double omega, double bottom_tol, double* final_res,
                     int cycle, int max_level) {
    Hier H;
    build_hier(H, ncell, dx, sym, bct, C, max_level);
    if (smoother == 1) {
        H.Dl1.resize(H.L.size());
        for (size_t l = 0; l < H.L.size(); l++) {
            H.Dl1[l].resize(3 * H.L[l].tot);
            l1_diag_rows(H.L[l], sym, H.C[l].data(), bct, H.Dl1[l].data());
        }
    }
    MGOpts o{cycle, nu1, nu2, sweeps, omega, max_iter, 200, bottom_tol};
    o.smoother = smoother;
    o.sweeps_finest = sweeps_finest;
    return fgmres(H, o, restart, x, b, tol_rel, tol_abs, final_res);
}
// the Galerkin stencil of multigrid level `level` >= 1 (243 planes of that level's nodes)
void ora_galerkin(const int* ncell, const double* dx, int sym, const double* C, const int* bct, int level, double* out) {
    Hier H;
    build_hier(H, ncell, dx, sym, bct, C, -1);
    std::vector<std::vector<double>> S;
    galerkin_levels(H, level, S);
    if (level >= 1 && level < (int)H.L.size()) std::copy(S[level].begin(), S[level].end(), out);
}
void ora_stencil_apply(const int* ncell, const double* dx, const double* S, const double* x, double* y) {
    Level L(ncell, dx);
    stencil_apply(L, S, x, y);
}
// Composite AMR solve / residual.  Level fields are concatenated: level l has 3 (n_cell 2^l + 1)^3
// values (C: ncoef per node, AoS); boxes as alamo_amr_create.
static void composite_args(int nl, const int* ncell0, int sym, const double* C, const double* v,
                           std::vector<const double*>& Cl, std::vector<double*>& vl, int nrec_v) {
    int64_t oc = 0, ov = 0;
    Cl.resize(nl);
    vl.resize(nl);
    for (int l = 0; l < nl; l++) {
        const int64_t tot = (int64_t)((ncell0[0] << l) + 1) * ((ncell0[1] << l) + 1) * ((ncell0[2] << l) + 1);
        Cl[l] = C + oc;
        vl[l] = (double*)v + ov;
        oc += tot * ncoef(sym);
        ov += tot * nrec_v;
    }
}
int ora_amr_solve(const int* ncell0, const double* dx0, int sym, const double* C, const int* bct, int nl,
                  const int* nbox, const int* box_lo, const int* box_hi, double* u, const double* b, double tol_rel,
                  double tol_abs, int cycle, int nu1, int nu2, int sweeps, int smoother, double omega, int max_iter,
                  double* final_res) {
    std::vector<const double*> Cl;
    std::vector<double*> ul, bl;
    composite_args(nl, ncell0, sym, C, u, Cl, ul, 3);
    std::vector<const double*> tmp;
    composite_args(nl, ncell0, sym, C, b, tmp, bl, 3);
    MGOpts o{cycle, nu1, nu2, sweeps, omega, max_iter, 200, 1e-4};
    o.smoother = smoother;
    Composite A;
    composite_build(A, ncell0, dx0, sym, bct, nl, Cl.data(), nbox, box_lo, box_hi, o);
    std::vector<const double*> bc(bl.begin(), bl.end());
    return composite_solve(A, ul.data(), bc.data(), tol_rel, tol_abs, final_res);
}
void ora_amr_residual(const int* ncell0, const double* dx0, int sym, const double* C, const int* bct, int nl,
                      const int* nbox, const int* box_lo, const int* box_hi, const double* u, const double* b,
                      double* r) {
    std::vector<const double*> Cl, tmp;
    std::vector<double*> ul, bl, rl;
    composite_args(nl, ncell0, sym, C, u, Cl, ul, 3);
    composite_args(nl, ncell0, sym, C, b, tmp, bl, 3);
    composite_args(nl, ncell0, sym, C, r, tmp, rl, 3);
    MGOpts o{0, 1, 1, 2, 2.0 / 3.0, 0, 200, 1e-4};
    Composite A;
    composite_build(A, ncell0, dx0, sym, bct, nl, Cl.data(), nbox, box_lo, box_hi, o);
    std::vector<const double*> bc(bl.begin(), bl.end());
    // the residual injects the fine values into covered coarse nodes: work on a copy of u
    int64_t nu = 0;
    for (int l = 0; l < nl; l++) nu += 3 * A.H[l].L[0].tot;
    std::vector<double> uc(u, u + nu);
    std::vector<double*> ucl;
    composite_args(nl, ncell0, sym, C, uc.data(), Cl, ucl, 3);
    composite_residual(A, ucl.data(), bc.data(), rl.data());
}
void ora_l1_diag_rows(const int* ncell, const double* dx, int sym, const double* C, const int* bct, double* d) {
    Level L(ncell, dx);
    l1_diag_rows(L, sym, C, bct, d);
}
// Multigrid hierarchy stiffness at level l (for parity of the coefficient restriction).
void ora_coeff_level(const int* ncell, const double* dx, int sym, const double* C, int level, double* out) {
    Hier H;
    int dummy[78] = {0};
    build_hier(H, ncell, dx, sym, dummy, C, level);
    if (level < (int)H.L.size()) std::copy(H.C[level].begin(), H.C[level].end(), out);
}
void ora_newton_rhs(const int* ncell, const double* dx, int sym, const double* C, const double* F0, const int* bct,
                    const double* u, const double* b, double* rhs) {
    Level L(ncell, dx);
    newton_rhs(L, sym, C, F0, bct, u, b, rhs, nullptr);
}
// psi variants: psi is the level-0 cell field (ncell[0] ncell[1] ncell[2], i fastest)
void ora_fapply_psi(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* psi,
                    double psi_small, const double* u, double* f) {
    Level L(ncell, dx);
    L.psi = psi;
    L.psi_small = psi_small;
    fapply(L, sym, C, bct, u, f);
}
void ora_diagonal_psi(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* psi,
                      double psi_small, double* D) {
    Level L(ncell, dx);
    L.psi = psi;
    L.psi_small = psi_small;
    diagonal(L, sym, C, bct, D);
}
void ora_newton_rhs_psi(const int* ncell, const double* dx, int sym, const double* C, const double* F0,
                        const int* bct, const double* psi, const double* u, const double* b, double* rhs) {
    Level L(ncell, dx);
    L.psi = psi;
    newton_rhs(L, sym, C, F0, bct, u, b, rhs, nullptr);
}
// restricted psi of multigrid level `level` (cells of that level)
void ora_psi_level(const int* ncell, const double* dx, const double* psi, int level, double* out) {
    Hier H;
    int dummy[78] = {0};
    std::vector<double> C((int64_t)(ncell[0] + 1) * (ncell[1] + 1) * (ncell[2] + 1) * 2, 1.0);
    build_hier(H, ncell, dx, 0, dummy, C.data(), level, psi, 1e-8);
    if (level < (int)H.L.size()) std::copy(H.psi[level].begin(), H.psi[level].end(), out);
}
int ora_mg_solve_psi(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* psi,
                     double psi_small, double* x, const double* b, double tol_rel, double tol_abs, int cycle, int nu1,
                     int nu2, int sweeps, double omega, int max_iter, double* final_res) {
    Hier H;
    build_hier(H, ncell, dx, sym, bct, C, -1, psi, psi_small);
    MGOpts o{cycle, nu1, nu2, sweeps, omega, max_iter, 200, 1e-4};
    return mg_solve(H, o, x, b, tol_rel, tol_abs, nullptr, 0, final_res);
}
void ora_post(const int* ncell, const double* dx, int sym, const double* C, const double* psi, double psi_small,
              const double* u, int kind, int voigt, double* out) {
    Level L(ncell, dx);
    L.psi = psi;
    L.psi_small = psi_small;
    post(L, sym, C, u, kind, voigt, out);
}
void ora_model_post(const int* ncell, const double* dx, int sym, const double* C, const double* F0, const double* u,
                    int kind, double* out) {
    Level L(ncell, dx);
    model_post(L, sym, C, F0, u, kind, out);
}
void ora_fapply_uniform(const int* ncell, const double* dx, int sym, const double* C, const int* bct, const double* u,
                        double* f) {
    Level L(ncell, dx);
    L.uniform = true;
    fapply(L, sym, C, bct, u, f);
}
// Integrator::Mechanics::UpdateModel blending (Integrator/Mechanics.H:242-254)
void ora_blend(int64_t nnodes, int nmodels, int nrec, const double* eta, const double* table, double* out, int aos) {
    for (int64_t v = 0; v < nnodes; v++)
        for (int q = 0; q < nrec; q++) {
            double s = 0.0;
            for (int n = 0; n < nmodels; n++) s += eta[n * nnodes + v] * table[n * nrec + q];
            out[aos ? v * nrec + q : q * nnodes + v] = s;
        }
}
void ora_bc_init(const int* ncell, const int* bct, const double* bcv, int homogeneous, double* b) {
    double dx[3] = {1, 1, 1};
    Level L(ncell, dx);
    bc_init(L, bct, bcv, homogeneous, b);
}
void ora_trac_xhi(const int* ncell, const double* dx, int sym, const double* C, const double* F0, const double* u,
                  double* out3) {
    Level L(ncell, dx);
    trac_xhi(L, sym, C, F0, u, out3);
}

}  // extern "C"
