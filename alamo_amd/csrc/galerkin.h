// Galerkin coarse operators A_c = R A_f P for the multigrid cycles (coarse_op GALERKIN; no
// reference counterpart: the reference re-discretises every coarse level from restricted
// coefficients, Operator/Elastic.cpp:746-836, which this library keeps as its default).  R is the
// restriction of Operator.cpp:262-363 (tensor weights 1/4 1/2 1/4 per direction, injection along
// a direction in which the coarse node lies on the domain boundary), P the trilinear prolongation
// of Operator.cpp:365-434.  Only the preconditioner changes: level 0 keeps the reference operator,
// so the converged solution is the same.
//
// Stencil layout of a coarse level: 27 offsets o = (dx+1) + 3 (dy+1) + 9 (dz+1) x 3 x 3 blocks,
// plane (9 o + 3 q + p) of `tot` values holds the coefficient of u_p at node + o in row q.
#pragma once
#include "kernels_generic.h"

namespace alamo {

constexpr int GAL_NPLANE = 243;

__host__ __device__ __forceinline__ int gal_plane(int o, int q, int p) { return 9 * o + 3 * q + p; }

// Coefficient block (row q, column u_p) of the level-0 row of node (i, j, k) for neighbour offset
// (ox, oy, oz): the stencils of apply_interior / apply_boundary written out.  c: the node's
// coefficients; dc[d]: their central differences along d (interior nodes).
template <typename T, int SYM>
__device__ __forceinline__ void row_block(const Geom& g, const BCTable& bc, const T* c, const T (*dc)[NCoef<SYM>::value],
                                          int i, int j, int k, int ox, int oy, int oz, T blk[3][3]) {
    const int idx[3] = {i, j, k};
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int o[3] = {ox, oy, oz};
    const T dx[3] = {T(g.dx[0]), T(g.dx[1]), T(g.dx[2])};
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++) blk[q][p] = T(0);
    const bool bnd = i == 0 || j == 0 || k == 0 || i == hi[0] || j == hi[1] || k == hi[2];
    if (bnd) {
        // first-derivative weight of offset o along l at this node (GetStencil: one-sided on the
        // boundary, central inside)
        auto d1w = [&](int l) -> T {
            for (int d = 0; d < 3; d++)
                if (d != l && o[d] != 0) return T(0);
            if (idx[l] == 0) return o[l] == 1 ? T(1) / dx[l] : (o[l] == 0 ? -T(1) / dx[l] : T(0));
            if (idx[l] == hi[l]) return o[l] == 0 ? T(1) / dx[l] : (o[l] == -1 ? -T(1) / dx[l] : T(0));
            return o[l] == 1 ? T(0.5) / dx[l] : (o[l] == -1 ? -T(0.5) / dx[l] : T(0));
        };
        const int face = classify(i, j, k, hi);
        double nd[3];
        region_normal(face, nd);
        for (int q = 0; q < 3; q++) {
            const int t = bc.t[face][q];
            if (t == DISPLACEMENT) {
                if (o[0] == 0 && o[1] == 0 && o[2] == 0) blk[q][q] = T(1);
            } else if (t == TRACTION) {
                for (int l = 0; l < 3; l++) {
                    const T w = d1w(l);
                    if (w == T(0)) continue;
                    for (int p = 0; p < 3; p++) {
                        T coef = T(0);
                        for (int e = 0; e < 3; e++) coef += T(nd[e]) * C4<SYM>(c, q, e, p, l);
                        blk[q][p] += coef * w;
                    }
                }
            } else if (t == NEUMANN) {
                for (int l = 0; l < 3; l++) blk[q][q] += T(nd[l]) * d1w(l);
            }
        }
        return;
    }
    const int nz = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
    if (nz == 3) return;
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++) {
            T s = T(0);
            // second and mixed differences: C_q jj p ll u_p,ll jj
            for (int jj = 0; jj < 3; jj++)
                for (int ll = 0; ll < 3; ll++) {
                    T w = T(0);
                    if (jj == ll) {
                        bool only = true;
                        for (int d = 0; d < 3; d++)
                            if (d != jj && o[d] != 0) only = false;
                        if (only) w = o[jj] == 0 ? T(-2) / (dx[jj] * dx[jj]) : T(1) / (dx[jj] * dx[jj]);
                    } else {
                        const int m3 = 3 - jj - ll;
                        if (o[jj] != 0 && o[ll] != 0 && o[m3] == 0)
                            w = T(o[jj] * o[ll]) * T(0.25) / (dx[jj] * dx[ll]);
                    }
                    if (w != T(0)) s += C4<SYM>(c, q, jj, p, ll) * w;
                }
            // (d_jj C_q jj p ll) u_p,ll
            if (!g.uniform && nz == 1)
                for (int ll = 0; ll < 3; ll++) {
                    if (o[ll] == 0) continue;
                    const T w = T(o[ll]) * T(0.5) / dx[ll];
                    for (int jj = 0; jj < 3; jj++) s += C4<SYM>(dc[jj], q, jj, p, ll) * w;
                }
            blk[q][p] = s;
        }
}

// Restriction weight of fine offset r (-1, 0, 1) along a direction in which coarse node I of a
// level with nc cells is (not) on the domain boundary.
__device__ __forceinline__ double gal_wr(int I, int nc, int r) {
    if (I == 0 || I == nc) return r == 0 ? 1.0 : 0.0;
    return r == 0 ? 0.5 : 0.25;
}

// Fine node (2I + r) along one direction, coarse target J = I + O: the offsets o of the fine row
// whose column 2I + r + o the prolongation maps to J, with weight 1 (coincident) or 1/2.
__device__ __forceinline__ double gal_wp(int r, int o, int O) {
    const int d = r + o - 2 * O;  // fine column - 2 J
    return d == 0 ? 1.0 : (d == 1 || d == -1 ? 0.5 : 0.0);
}

// Level 1 from the level-0 rows, one work-item per (coarse node, coarse offset), offset slowest.
template <int SYM>
__global__ void __launch_bounds__(256) k_rap0(Geom gc, Geom gf, BCTable bc, const double* __restrict__ C,
                                              double* __restrict__ S) {
    constexpr int NC = NCoef<SYM>::value;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 27 * gc.tot) return;
    const int O = (int)(t / gc.tot);
    const int64_t mc = t - (int64_t)O * gc.tot;
    int I[3];
    node_ijk(gc, mc, I[0], I[1], I[2]);
    const int Of[3] = {O % 3 - 1, (O / 3) % 3 - 1, O / 9 - 1};
    double acc[3][3] = {};
    bool inside = true;
    for (int d = 0; d < 3; d++) inside = inside && I[d] + Of[d] >= 0 && I[d] + Of[d] <= gc.n[d];
    if (inside) {
        for (int rz = -1; rz <= 1; rz++)
            for (int ry = -1; ry <= 1; ry++)
                for (int rx = -1; rx <= 1; rx++) {
                    const int r[3] = {rx, ry, rz};
                    double wr = 1.0;
                    for (int d = 0; d < 3; d++) wr *= gal_wr(I[d], gc.n[d], r[d]);
                    if (wr == 0.0) continue;
                    const int i = 2 * I[0] + rx, j = 2 * I[1] + ry, k = 2 * I[2] + rz;
                    const int64_t m = i + gf.sy * j + gf.sz * (int64_t)k;
                    double c[NC], dc[3][NC];
                    load_coef<double, SYM>(C, gf.tot, m, c);
                    const bool bnd = i == 0 || j == 0 || k == 0 || i == gf.n[0] || j == gf.n[1] || k == gf.n[2];
                    if (!bnd) {
                        const int64_t st[3] = {1, gf.sy, gf.sz};
                        for (int d = 0; d < 3; d++)
                            for (int q = 0; q < NC; q++)
                                dc[d][q] = (C[q * gf.tot + m + st[d]] - C[q * gf.tot + m - st[d]]) * 0.5 / gf.dx[d];
                    }
                    for (int oz = -1; oz <= 1; oz++) {
                        const double wz = gal_wp(rz, oz, Of[2]);
                        if (wz == 0.0 || k + oz < 0 || k + oz > gf.n[2]) continue;
                        for (int oy = -1; oy <= 1; oy++) {
                            const double wy = gal_wp(ry, oy, Of[1]);
                            if (wy == 0.0 || j + oy < 0 || j + oy > gf.n[1]) continue;
                            for (int ox = -1; ox <= 1; ox++) {
                                const double wx = gal_wp(rx, ox, Of[0]);
                                if (wx == 0.0 || i + ox < 0 || i + ox > gf.n[0]) continue;
                                double blk[3][3];
                                row_block<double, SYM>(gf, bc, c, dc, i, j, k, ox, oy, oz, blk);
                                const double w = wr * wz * wy * wx;
                                for (int q = 0; q < 3; q++)
                                    for (int p = 0; p < 3; p++) acc[q][p] += w * blk[q][p];
                            }
                        }
                    }
                }
    }
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++) S[(int64_t)gal_plane(O, q, p) * gc.tot + mc] = acc[q][p];
}

// Level l + 1 from the stencil of level l (same work split)
__global__ void __launch_bounds__(256) k_rap(Geom gc, Geom gf, const double* __restrict__ Sf, double* __restrict__ S) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 27 * gc.tot) return;
    const int O = (int)(t / gc.tot);
    const int64_t mc = t - (int64_t)O * gc.tot;
    int I[3];
    node_ijk(gc, mc, I[0], I[1], I[2]);
    const int Of[3] = {O % 3 - 1, (O / 3) % 3 - 1, O / 9 - 1};
    double acc[3][3] = {};
    bool inside = true;
    for (int d = 0; d < 3; d++) inside = inside && I[d] + Of[d] >= 0 && I[d] + Of[d] <= gc.n[d];
    if (inside) {
        for (int rz = -1; rz <= 1; rz++)
            for (int ry = -1; ry <= 1; ry++)
                for (int rx = -1; rx <= 1; rx++) {
                    const int r[3] = {rx, ry, rz};
                    double wr = 1.0;
                    for (int d = 0; d < 3; d++) wr *= gal_wr(I[d], gc.n[d], r[d]);
                    if (wr == 0.0) continue;
                    const int i = 2 * I[0] + rx, j = 2 * I[1] + ry, k = 2 * I[2] + rz;
                    const int64_t m = i + gf.sy * j + gf.sz * (int64_t)k;
                    for (int oz = -1; oz <= 1; oz++) {
                        const double wz = gal_wp(rz, oz, Of[2]);
                        if (wz == 0.0 || k + oz < 0 || k + oz > gf.n[2]) continue;
                        for (int oy = -1; oy <= 1; oy++) {
                            const double wy = gal_wp(ry, oy, Of[1]);
                            if (wy == 0.0 || j + oy < 0 || j + oy > gf.n[1]) continue;
                            for (int ox = -1; ox <= 1; ox++) {
                                const double wx = gal_wp(rx, ox, Of[0]);
                                if (wx == 0.0 || i + ox < 0 || i + ox > gf.n[0]) continue;
                                const int o = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
                                const double w = wr * wz * wy * wx;
                                for (int q = 0; q < 3; q++)
                                    for (int p = 0; p < 3; p++)
                                        acc[q][p] += w * Sf[(int64_t)gal_plane(o, q, p) * gf.tot + m];
                            }
                        }
                    }
                }
    }
    for (int q = 0; q < 3; q++)
        for (int p = 0; p < 3; p++) S[(int64_t)gal_plane(O, q, p) * gc.tot + mc] = acc[q][p];
}

// out = A u | b - A u | smoothing sweep | diagonal on a stencil level, one work-item per node.
// The point diagonal is the stencil's centre entry; with an l1 smoother diagonal on the level
// (g.dsm) the sweep is x += (b - A x) / dsm.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_stencil(Geom g, const T* __restrict__ S, const T* __restrict__ u,
                                                 const T* __restrict__ b, T* __restrict__ out, T omega) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    T f[3] = {T(0), T(0), T(0)};
    if (MODE != DIAG) {
        for (int oz = -1; oz <= 1; oz++) {
            if (k + oz < 0 || k + oz > g.n[2]) continue;
            for (int oy = -1; oy <= 1; oy++) {
                if (j + oy < 0 || j + oy > g.n[1]) continue;
                for (int ox = -1; ox <= 1; ox++) {
                    if (i + ox < 0 || i + ox > g.n[0]) continue;
                    const int o = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
                    const int64_t mn = m + ox + g.sy * oy + g.sz * oz;
                    const T u0 = u[mn], u1 = u[g.tot + mn], u2 = u[2 * g.tot + mn];
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const T* Sq = S + (int64_t)gal_plane(o, q, 0) * g.tot + m;
                        f[q] += Sq[0] * u0 + Sq[g.tot] * u1 + Sq[2 * g.tot] * u2;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int64_t qq = q * g.tot + m;
        const T D = S[(int64_t)gal_plane(13, q, q) * g.tot + m];
        if (MODE == APPLY) out[qq] = f[q];
        else if (MODE == RESID) out[qq] = b[qq] - f[q];
        else if (MODE == DIAG) out[qq] = D;
        else out[qq] = jacobi_update(g, qq, u[qq], b[qq], f[q], D, omega);
    }
}

// l1 smoother diagonal of a stencil level: sign(centre) sum |row|
template <typename T>
__global__ void __launch_bounds__(256) k_stencil_l1(Geom g, const T* __restrict__ S, T* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    for (int q = 0; q < 3; q++) {
        T s = T(0);
        for (int o = 0; o < 27; o++)
            for (int p = 0; p < 3; p++) s += fabs(S[(int64_t)gal_plane(o, q, p) * g.tot + m]);
        out[q * g.tot + m] = S[(int64_t)gal_plane(13, q, q) * g.tot + m] < T(0) ? -s : s;
    }
}

template <typename T>
inline void launch_stencil(int mode, const Geom& g, const T* S, const T* u, const T* b, T* out, T omega,
                           hipStream_t s) {
    const unsigned nb = (unsigned)((g.tot + 255) / 256);
    if (mode == APPLY) k_stencil<T, APPLY><<<nb, 256, 0, s>>>(g, S, u, b, out, omega);
    else if (mode == RESID) k_stencil<T, RESID><<<nb, 256, 0, s>>>(g, S, u, b, out, omega);
    else if (mode == JACOBI) k_stencil<T, JACOBI><<<nb, 256, 0, s>>>(g, S, u, b, out, omega);
    else k_stencil<T, DIAG><<<nb, 256, 0, s>>>(g, S, u, b, out, omega);
}

}  // namespace alamo
