// Per-node arithmetic of Operator::Elastic<SYM> shared by every kernel.
//
//   interior rows   f_i = C_ijkl u_k,lj + (dC_idkl/dx_d) u_k,l      (Operator/Elastic.cpp:204-254)
//   boundary rows   f   = BC(u, grad u, C : grad u)                 (Operator/Elastic.cpp:195-203,
//                                                                   BC/Operator/Elastic/Constant.H:344-425)
//   diagonal        Operator/Elastic.cpp:296-379 with Numeric/Stencil.H:906-963
//
// m_uniform is false, the configuration Integrator/Base/Mechanics.H:193-198 builds; with psi set
// (Elastic::SetPsi) the psi terms of Elastic.cpp:191-195,243-260,340-363 apply.
#pragma once
#include <type_traits>
#include "common.h"

namespace alamo {

template <typename T, int SYM>
__device__ __forceinline__ void load_coef(const T* __restrict__ C, int64_t tot, int64_t m, T* c) {
#pragma unroll
    for (int q = 0; q < NCoef<SYM>::value; q++) c[q] = C[q * tot + m];
}

// ---------------------------------------------------------------------------
// psi.  The operator fills its whole psi field with 1 (Elastic.cpp:49-52) and SetPsi copies the
// domain's cells only (:122-133), so a cell outside the non-periodic domain reads 1.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ const T* psi_field(const Geom& g);
template <> __device__ __forceinline__ const double* psi_field<double>(const Geom& g) { return g.psi; }
template <> __device__ __forceinline__ const float* psi_field<float>(const Geom& g) { return g.psif; }

template <typename T>
__device__ __forceinline__ T psi_cell(const T* __restrict__ ps, const Geom& g, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return T(1);
    return ps[i + (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k)];
}
// Interpolate::CellToNodeAverage (Numeric/Stencil.H:1385-1406): mean of the node's 8 cells
template <typename T>
__device__ __forceinline__ T psi_node_mean(const T* __restrict__ ps, const Geom& g, int i, int j, int k) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 8; c++) s += psi_cell(ps, g, i - (c & 1), j - ((c >> 1) & 1), k - (c >> 2));
    return s * T(0.125);
}
// psi_avg = (1 - psi_small) CellToNodeAverage(psi) + psi_small (Elastic.cpp:191-192); 1 unset
template <typename T>
__device__ __forceinline__ T psi_avg(const Geom& g, int i, int j, int k) {
    const T* ps = psi_field<T>(g);
    if (!ps) return T(1);
    return (T(1) - T(g.psi_small)) * psi_node_mean(ps, g, i, j, k) + T(g.psi_small);
}
// Numeric::CellGradientOnNode (Numeric/Stencil.H:709-712)
template <typename T>
__device__ __forceinline__ void psi_grad(const T* __restrict__ ps, const Geom& g, int i, int j, int k, T gp[3]) {
    const int idx[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        T s = T(0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int p[3], m[3];
            p[d] = idx[d];
            m[d] = idx[d] - 1;
            p[a] = m[a] = idx[a] - (q & 1);
            p[b] = m[b] = idx[b] - (q >> 1);
            s += psi_cell(ps, g, p[0], p[1], p[2]) - psi_cell(ps, g, m[0], m[1], m[2]);
        }
        gp[d] = T(0.25) * s / T(g.dx[d]);
    }
}

// First derivative with Numeric::GetStencil semantics (Numeric/Stencil.H:41-60, 70-191).
template <typename T>
__device__ __forceinline__ T d1(const T* __restrict__ f, int64_t m, int64_t s, int idx, int hi, T dx) {
    if (idx == 0) return (f[m + s] - f[m]) / dx;
    if (idx == hi) return (f[m] - f[m - s]) / dx;
    return (f[m + s] - f[m - s]) * T(0.5) / dx;
}

template <typename T>
__device__ __forceinline__ void bc_row(const BCTable& bc, int face, const T u[3], const T G[3][3], const T S[3][3],
                                       T f[3]) {
    double nd[3];
    region_normal(face, nd);
    const T n[3] = {T(nd[0]), T(nd[1]), T(nd[2])};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int t = bc.t[face][c];
        T v = T(0);
        if (t == DISPLACEMENT) v = u[c];
        else if (t == TRACTION) v = S[c][0] * n[0] + S[c][1] * n[1] + S[c][2] * n[2];
        else if (t == NEUMANN) v = G[c][0] * n[0] + G[c][1] * n[1] + G[c][2] * n[2];
        f[c] = v;
    }
}

// (A u) at a boundary node.
template <typename T, int SYM>
__device__ __forceinline__ void apply_boundary(const Geom& g, const BCTable& bc, const T* __restrict__ C,
                                               const T* __restrict__ u, int i, int j, int k, T f[3]) {
    const int64_t m = i + g.sy * j + g.sz * k;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    const int64_t st[3] = {1, g.sy, g.sz};
    T G[3][3], S[3][3], uu[3], c[NCoef<SYM>::value];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* up = u + p * g.tot;
        uu[p] = up[m];
#pragma unroll
        for (int d = 0; d < 3; d++) G[p][d] = d1(up, m, st[d], idx[d], hi[d], T(g.dx[d]));
    }
    load_coef<T, SYM>(C, g.tot, m, c);
    c_dot<SYM>(c, G, S);
    if (psi_field<T>(g)) {
        // sig = (DDW grad u) psi_avg (Elastic.cpp:195)
        const T pa = psi_avg<T>(g, i, j, k);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) S[a][b] *= pa;
    }
    bc_row(bc, classify(i, j, k, hi), uu, G, S, f);
}

// Diagonal of the boundary row for each component (Elastic.cpp:352-360 with
// Gradient_Diagonal<Set::Matrix>, Stencil.H:906-930).
template <typename T, int SYM>
__device__ __forceinline__ void diag_boundary(const Geom& g, const BCTable& bc, const T* __restrict__ C, int i, int j,
                                              int k, T D[3]) {
    const int64_t m = i + g.sy * j + g.sz * k;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    T c[NCoef<SYM>::value];
    load_coef<T, SYM>(C, g.tot, m, c);
    T cd[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        bool mn = idx[d] == 0, mx = idx[d] == hi[d];
        cd[d] = (T(mx ? 1 : 0) - T(mn ? 1 : 0)) / (T(mn || mx ? 1 : 2) * T(g.dx[d]));
    }
    const int face = classify(i, j, k, hi);
    const T pa = psi_avg<T>(g, i, j, k);  // Elastic.cpp:340-355
#pragma unroll
    for (int p = 0; p < 3; p++) {
        T G[3][3], S[3][3], uu[3] = {T(0), T(0), T(0)}, f[3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) G[a][d] = (a == p) ? cd[d] : T(0);
        c_dot<SYM>(c, G, S);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) S[a][b] *= pa;
        uu[p] = T(1);
        bc_row(bc, face, uu, G, S, f);
        D[p] = f[p];
    }
}

// Interior diagonal D_p = sum_j C_pjpj (-2/dx_j^2)   (Elastic.cpp:361-365, Stencil.H:932-963)
template <typename T, int SYM>
__device__ __forceinline__ void diag_interior(const Geom& g, const T* c, T D[3]) {
    const T w[3] = {T(-2.0 / (g.dx[0] * g.dx[0])), T(-2.0 / (g.dx[1] * g.dx[1])), T(-2.0 / (g.dx[2] * g.dx[2]))};
#pragma unroll
    for (int p = 0; p < 3; p++) {
        T s = T(0);
#pragma unroll
        for (int jj = 0; jj < 3; jj++) s += C4<SYM>(c, p, jj, p, jj) * w[jj];
        D[p] = s;
    }
}

// ---------------------------------------------------------------------------
// l1-Jacobi smoother diagonal: for each row component q, sign(a_qq) sum_j |a_qj| over the row's
// coefficients, assembled per (offset, component) from the same stencils as the rows above
// (interior: second differences, the four-corner mixed difference and central first
// differences of u and C; boundary: the BC row with GetStencil first derivatives).  Offsets are
// (dx, dy, dz) in {-1, 0, 1}^3, slot 13 + dx + 3 dy + 9 dz.  No psi terms (levels with psi keep
// the point smoother).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T l1_finish(const T acc[27][3], int q) {
    T s = T(0);
#pragma unroll
    for (int o = 0; o < 27; o++)
#pragma unroll
        for (int k = 0; k < 3; k++) s += fabs(acc[o][k]);
    return acc[13][q] < T(0) ? -s : s;
}

// The rows of the operator at node (i, j, k), no psi terms: sink(q, acc) for each row component q,
// acc[o][c] the weight of u_c at offset slot o.
template <typename T, int SYM, typename Sink>
__device__ void row_coeffs(const Geom& g, const BCTable& bc, const T* __restrict__ C, int i, int j, int k, Sink&& sink) {
    constexpr int NC = NCoef<SYM>::value;
    const int64_t m = i + g.sy * j + g.sz * k;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    const int sl[3] = {1, 3, 9};
    const T dx[3] = {T(g.dx[0]), T(g.dx[1]), T(g.dx[2])};
    T c[NC];
    load_coef<T, SYM>(C, g.tot, m, c);
    const bool bnd = i == 0 || j == 0 || k == 0 || i == hi[0] || j == hi[1] || k == hi[2];
    if (bnd) {
        const int face = classify(i, j, k, hi);
        double nd[3];
        region_normal(face, nd);
        for (int q = 0; q < 3; q++) {
            T acc[27][3] = {};
            const int t = bc.t[face][q];
            // first-derivative stencil along l at this node (Numeric::GetStencil): weight w at offset
            // sign s of slot 13 + s sl[l]
            auto d1w = [&](int l, T coef, int comp) {
                if (idx[l] == 0) {
                    acc[13 + sl[l]][comp] += coef / dx[l];
                    acc[13][comp] -= coef / dx[l];
                } else if (idx[l] == hi[l]) {
                    acc[13][comp] += coef / dx[l];
                    acc[13 - sl[l]][comp] -= coef / dx[l];
                } else {
                    acc[13 + sl[l]][comp] += coef * T(0.5) / dx[l];
                    acc[13 - sl[l]][comp] -= coef * T(0.5) / dx[l];
                }
            };
            if (t == DISPLACEMENT) {
                acc[13][q] = T(1);
            } else if (t == TRACTION) {
                // (C : grad u) n: coefficient of u_kk,l is sum_e n_e C_q e kk l
                for (int kk = 0; kk < 3; kk++)
                    for (int l = 0; l < 3; l++) {
                        T coef = T(0);
                        for (int e = 0; e < 3; e++) coef += T(nd[e]) * C4<SYM>(c, q, e, kk, l);
                        d1w(l, coef, kk);
                    }
            } else if (t == NEUMANN) {
                for (int l = 0; l < 3; l++) d1w(l, T(nd[l]), q);
            }
            sink(q, acc);
        }
        return;
    }
    T dc[3][NC];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int64_t st = d == 0 ? 1 : (d == 1 ? g.sy : g.sz);
#pragma unroll
        for (int p = 0; p < NC; p++) dc[d][p] = (C[p * g.tot + m + st] - C[p * g.tot + m - st]) * T(0.5) / dx[d];
    }
    for (int q = 0; q < 3; q++) {
        T acc[27][3] = {};
        for (int kk = 0; kk < 3; kk++)
            for (int jj = 0; jj < 3; jj++)
                for (int ll = 0; ll < 3; ll++) {
                    // C_q jj kk ll u_kk,ll jj
                    const T cf = C4<SYM>(c, q, jj, kk, ll);
                    if (jj == ll) {
                        const T w = cf / (dx[jj] * dx[jj]);
                        acc[13 + sl[jj]][kk] += w;
                        acc[13 - sl[jj]][kk] += w;
                        acc[13][kk] -= T(2) * w;
                    } else {
                        const T w = cf * T(0.25) / (dx[jj] * dx[ll]);
                        acc[13 + sl[jj] + sl[ll]][kk] += w;
                        acc[13 - sl[jj] - sl[ll]][kk] += w;
                        acc[13 + sl[jj] - sl[ll]][kk] -= w;
                        acc[13 - sl[jj] + sl[ll]][kk] -= w;
                    }
                    // (d_jj C_q jj kk ll) u_kk,ll
                    if (!g.uniform) {
                        const T w = C4<SYM>(dc[jj], q, jj, kk, ll) * T(0.5) / dx[ll];
                        acc[13 + sl[ll]][kk] += w;
                        acc[13 - sl[ll]][kk] -= w;
                    }
                }
        sink(q, acc);
    }
}

template <typename T, int SYM>
__device__ void l1_row(const Geom& g, const BCTable& bc, const T* __restrict__ C, int i, int j, int k, T Dl1[3]) {
    row_coeffs<T, SYM>(g, bc, C, i, j, k, [&](int q, const T(&acc)[27][3]) { Dl1[q] = l1_finish(acc, q); });
}

// (A u) at an interior node, generic form (any SYM), direct global loads.
template <typename T, int SYM>
__device__ __forceinline__ void apply_interior(const Geom& g, const T* __restrict__ C, const T* __restrict__ u,
                                               int64_t m, T f[3], T* cout /* may be null: receives C(m) */) {
    constexpr int NC = NCoef<SYM>::value;
    const int64_t st[3] = {1, g.sy, g.sz};
    const T dx[3] = {T(g.dx[0]), T(g.dx[1]), T(g.dx[2])};
    T gg[3][3][3], G[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* up = u + p * g.tot;
        const T u0 = up[m];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T up1 = up[m + st[a]], um1 = up[m - st[a]];
            gg[p][a][a] = (up1 - T(2) * u0 + um1) / dx[a] / dx[a];
            G[p][a] = (up1 - um1) * T(0.5) / dx[a];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = a + 1; b < 3; b++) {
                const T v = T(0.25) * (up[m + st[a] + st[b]] + up[m - st[a] - st[b]] - up[m + st[a] - st[b]] -
                                       up[m - st[a] + st[b]]) / (dx[a] * dx[b]);
                gg[p][a][b] = v;
                gg[p][b][a] = v;
            }
    }
    T c[NC];
    load_coef<T, SYM>(C, g.tot, m, c);
    if (cout) {
#pragma unroll
        for (int q = 0; q < NC; q++) cout[q] = c[q];
    }
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
        T s = T(0);
#pragma unroll
        for (int jj = 0; jj < 3; jj++)
#pragma unroll
            for (int kk = 0; kk < 3; kk++)
#pragma unroll
                for (int ll = 0; ll < 3; ll++) s += C4<SYM>(c, ii, jj, kk, ll) * gg[kk][ll][jj];
        f[ii] = s;
    }
    if (g.uniform) return;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T dc[NC];
#pragma unroll
        for (int q = 0; q < NC; q++) dc[q] = (C[q * g.tot + m + st[d]] - C[q * g.tot + m - st[d]]) * T(0.5) / dx[d];
#pragma unroll
        for (int ii = 0; ii < 3; ii++) {
            T s = T(0);
#pragma unroll
            for (int kk = 0; kk < 3; kk++)
#pragma unroll
                for (int ll = 0; ll < 3; ll++) s += C4<SYM>(dc, ii, d, kk, ll) * G[kk][ll];
            f[ii] += s;
        }
    }
}

// The 6 index pairs of MajorMinor storage (00, 01, 02, 11, 12, 22; common.h::pair_of order).
__host__ __device__ constexpr int mm_pa(int P) { return P < 3 ? 0 : (P < 5 ? 1 : 2); }
__host__ __device__ constexpr int mm_pb(int P) { return P == 0 ? 0 : P == 1 ? 1 : P == 2 ? 2 : P == 3 ? 1 : 2; }

// Terms of slot (pair I, pair J), I <= J: C_ijkl = c for every ordered (i, j) of one pair and
// ordered (k, l) of the other; each adds c u_k,lj + (d_j C) u_k,l to f_i (and c w_j to D_i on the
// diagonal slots).  All indices are template constants.
template <int I, int J, typename T, bool DC = true>
__device__ __forceinline__ void mm_slot_terms(T c, const T dc[3], const T gg[3][3][3], const T G[3][3], T f[3],
                                              T* D, const T w[3]) {
    constexpr int ia = mm_pa(I), ib = mm_pb(I), ka = mm_pa(J), kb = mm_pb(J);
    constexpr int ni = ia == ib ? 1 : 2, nk = ka == kb ? 1 : 2;
#pragma unroll
    for (int r = 0; r < ni; r++) {
        const int i = r == 0 ? ia : ib, j = r == 0 ? ib : ia;
#pragma unroll
        for (int t = 0; t < nk; t++) {
            const int k = t == 0 ? ka : kb, l = t == 0 ? kb : ka;
            if (DC) {
                f[i] += c * gg[k][l][j] + dc[j] * G[k][l];
                if (I != J) f[k] += c * gg[i][j][l] + dc[l] * G[i][j];
            } else {
                f[i] += c * gg[k][l][j];
                if (I != J) f[k] += c * gg[i][j][l];
            }
        }
        if (I == J && D) D[i] += c * w[j];
    }
}

// pairs (I, J) with I <= J in slot order (common.h::mm_slot)
template <int I, int J, typename F>
__device__ __forceinline__ void for_slots(F& f) {
    if constexpr (I < 6) {
        if constexpr (J < 6) {
            f(std::integral_constant<int, I>{}, std::integral_constant<int, J>{});
            for_slots<I, J + 1>(f);
        } else {
            for_slots<I + 1, I + 1>(f);
        }
    }
}

// MajorMinor interior row (Elastic.cpp:204-254 with Set/Matrix4_MajorMinor.H): the same sums as
// apply_interior, reordered slot-major.  The 21 coefficient slots are streamed one at a time, each
// with its six neighbours for the gradient term, so only the derivatives of u stay in registers.
// D (optional) receives the Jacobi diagonal.
// Coefficient sources of the MajorMinor row: SoA planes, or a grain table (ngrains x 21 values,
// staged in LDS) indexed by the node's grain id.  at(q, o) = slot q at node m + o, o one of the
// row's seven offsets (0, -+1 per direction; `dir` names it: 0 centre, 1 + 2d plus, 2 + 2d minus).
template <typename T>
struct CoefPlanes {
    const T* __restrict__ C;
    int64_t tot, m;
    int64_t off[7];
    __device__ __forceinline__ T at(int q, int dir) const { return C[(int64_t)q * tot + m + off[dir]]; }
};
template <typename T>
struct CoefGrains {
    const T* tab;  // LDS
    int row[7];    // 21 * grain id at the seven offsets
    __device__ __forceinline__ T at(int q, int dir) const { return tab[row[dir] + q]; }
};

template <typename T, typename CS>
__device__ __forceinline__ void apply_interior_mm_cs(const Geom& g, const CS& cs, const T* __restrict__ u, int64_t m,
                                                     T f[3], T* D);

template <typename T>
__device__ __forceinline__ void apply_interior_mm(const Geom& g, const T* __restrict__ C, const T* __restrict__ u,
                                                  int64_t m, T f[3], T* D) {
    CoefPlanes<T> cs{C, g.tot, m, {0, 1, -1, g.sy, -g.sy, g.sz, -g.sz}};
    apply_interior_mm_cs<T>(g, cs, u, m, f, D);
}

template <typename T, typename CS>
__device__ __forceinline__ void apply_interior_mm_cs(const Geom& g, const CS& cs, const T* __restrict__ u, int64_t m,
                                                     T f[3], T* D) {
    const int64_t st[3] = {1, g.sy, g.sz};
    const T ih[3] = {T(0.5 / g.dx[0]), T(0.5 / g.dx[1]), T(0.5 / g.dx[2])};  // central first difference
    const T i2[3] = {T(1.0 / (g.dx[0] * g.dx[0])), T(1.0 / (g.dx[1] * g.dx[1])), T(1.0 / (g.dx[2] * g.dx[2]))};
    const T w[3] = {T(-2) * i2[0], T(-2) * i2[1], T(-2) * i2[2]};
    T gg[3][3][3], G[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* up = u + p * g.tot;
        const T u0 = up[m];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T up1 = up[m + st[a]], um1 = up[m - st[a]];
            gg[p][a][a] = (up1 - T(2) * u0 + um1) * i2[a];
            G[p][a] = (up1 - um1) * ih[a];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = a + 1; b < 3; b++) {
                const T v = (up[m + st[a] + st[b]] + up[m - st[a] - st[b]] - up[m + st[a] - st[b]] -
                             up[m - st[a] + st[b]]) * (ih[a] * ih[b]);
                gg[p][a][b] = v;
                gg[p][b][a] = v;
            }
    }
    f[0] = f[1] = f[2] = T(0);
    if (D) D[0] = D[1] = D[2] = T(0);
    auto slot = [&](auto ti, auto tj) {
        constexpr int I = decltype(ti)::value, J = decltype(tj)::value;
        constexpr int q = mm_slot(mm_pa(I), mm_pb(I), mm_pa(J), mm_pb(J));
        const T c = cs.at(q, 0);
        T dc[3];
#pragma unroll
        for (int d = 0; d < 3; d++) dc[d] = (cs.at(q, 1 + 2 * d) - cs.at(q, 2 + 2 * d)) * ih[d];
        mm_slot_terms<I, J, T>(c, dc, gg, G, f, D, w);
    };
    for_slots<0, 0>(slot);
}

// The same MajorMinor interior row from node-local data only: the 21 slots at the node for
// C : grad grad u, and the prepared contraction D_ikl = sum_j d_j C_ijkl (Geom::mmd) for the
// gradient term sum_kl D_ikl u_k,l -- 39 coefficient loads per node instead of 21 x 7.
template <typename T>
__device__ __forceinline__ void apply_interior_mm_d(const Geom& g, const T* __restrict__ C,
                                                    const T* __restrict__ Dm, const T* __restrict__ u, int64_t m,
                                                    T f[3], T* D) {
    const int64_t st[3] = {1, g.sy, g.sz};
    const T ih[3] = {T(0.5 / g.dx[0]), T(0.5 / g.dx[1]), T(0.5 / g.dx[2])};
    const T i2[3] = {T(1.0 / (g.dx[0] * g.dx[0])), T(1.0 / (g.dx[1] * g.dx[1])), T(1.0 / (g.dx[2] * g.dx[2]))};
    const T w[3] = {T(-2) * i2[0], T(-2) * i2[1], T(-2) * i2[2]};
    T gg[3][3][3], G[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* up = u + p * g.tot;
        const T u0 = up[m];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T up1 = up[m + st[a]], um1 = up[m - st[a]];
            gg[p][a][a] = (up1 - T(2) * u0 + um1) * i2[a];
            G[p][a] = (up1 - um1) * ih[a];
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = a + 1; b < 3; b++) {
                const T v = (up[m + st[a] + st[b]] + up[m - st[a] - st[b]] - up[m + st[a] - st[b]] -
                             up[m - st[a] + st[b]]) * (ih[a] * ih[b]);
                gg[p][a][b] = v;
                gg[p][b][a] = v;
            }
    }
    f[0] = f[1] = f[2] = T(0);
    if (D) D[0] = D[1] = D[2] = T(0);
    const T dz[3] = {T(0), T(0), T(0)};
    auto slot = [&](auto ti, auto tj) {
        constexpr int I = decltype(ti)::value, J = decltype(tj)::value;
        constexpr int q = mm_slot(mm_pa(I), mm_pb(I), mm_pa(J), mm_pb(J));
        mm_slot_terms<I, J, T, false>(C[(int64_t)q * g.tot + m], dz, gg, G, f, D, w);
    };
    for_slots<0, 0>(slot);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        T s = T(0);
#pragma unroll
        for (int P = 0; P < 6; P++) {
            const int k = mm_pa(P), l = mm_pb(P);
            const T gs = k == l ? G[k][k] : G[k][l] + G[l][k];
            s += Dm[(int64_t)(6 * i + P) * g.tot + m] * gs;
        }
        f[i] += s;
    }
}

// D_ikl = sum_j d_j C_ijkl at an interior node (central differences of the 21 slot planes, the
// differences the rows above take): 18 values, out[6 i + pair(k, l)]
template <typename T>
__device__ __forceinline__ void mm_grad_contraction(const Geom& g, const T* __restrict__ C, int64_t m, T out[18]) {
    const int64_t st[3] = {1, g.sy, g.sz};
    const T ih[3] = {T(0.5 / g.dx[0]), T(0.5 / g.dx[1]), T(0.5 / g.dx[2])};
#pragma unroll
    for (int e = 0; e < 18; e++) out[e] = T(0);
#pragma unroll
    for (int j = 0; j < 3; j++) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int P = 0; P < 6; P++) {
                const int q = mm_slot(i, j, mm_pa(P), mm_pb(P));
                out[6 * i + P] += (C[(int64_t)q * g.tot + m + st[j]] - C[(int64_t)q * g.tot + m - st[j]]) * ih[j];
            }
    }
}

// psi terms of an interior row (Elastic.cpp:243-260): f = psi_avg f + (C : grad u) . grad psi (1 - s),
// applied to the row f of apply_interior; D (may be null) is scaled by psi_avg (:363)
template <typename T, int SYM>
__device__ __forceinline__ void psi_interior(const Geom& g, const T* __restrict__ C, const T* __restrict__ u, int i,
                                             int j, int k, T f[3], T* D) {
    const T* ps = psi_field<T>(g);
    if (!ps) return;
    const T pa = psi_avg<T>(g, i, j, k);
    if (D) {
#pragma unroll
        for (int p = 0; p < 3; p++) D[p] *= pa;
    }
    if (!u) return;
    const int64_t m = i + g.sy * j + g.sz * k;
    const int64_t st[3] = {1, g.sy, g.sz};
    T G[3][3], S[3][3], c[NCoef<SYM>::value], gp[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* up = u + p * g.tot;
#pragma unroll
        for (int d = 0; d < 3; d++) G[p][d] = (up[m + st[d]] - up[m - st[d]]) * T(0.5) / T(g.dx[d]);
    }
    load_coef<T, SYM>(C, g.tot, m, c);
    c_dot<SYM>(c, G, S);
    psi_grad(ps, g, i, j, k, gp);
    const T w = T(1) - T(g.psi_small);
#pragma unroll
    for (int a = 0; a < 3; a++) f[a] = f[a] * pa + w * (S[a][0] * gp[0] + S[a][1] * gp[1] + S[a][2] * gp[2]);
}

}  // namespace alamo
