// Post-processing of a solution and the model blending that feeds SetModel.
//
//   Elastic::Strain  eps = (grad u + grad u^T) / 2           (Operator/Elastic.cpp:430-484)
//   Elastic::Stress  sig = (C : grad u) psi_avg              (Operator/Elastic.cpp:487-550)
//   Elastic::Energy  W  += sig : eps / 2, sig = C : grad u   (Operator/Elastic.cpp:552-603)
// with the GetStencil gradients (one-sided on the domain's faces).  Tensors are written as 9
// SoA planes in row-major order, or as 6 Voigt planes 00, 11, 22, 12, 20, 01.
//
//   Newton::DW       sig = C : (grad u - F0)                  (Solver/Nonlocal/Newton.H:571-611)
//   Newton::W        w = (grad u - F0) : sig / 2              (Solver/Nonlocal/Newton.H:528-569)
// the last two for an affine model (Model/Solid/Affine/Isotropic.H:38-45, kinematic variable
// grad u), F0 9 values per node (may be null), DW as 9 row-major planes, no psi factor.
//
//   Integrator::Mechanics::UpdateModel  model = sum_n eta_n model_n   (Integrator/Mechanics.H:242-254)
#pragma once
#include "node_ops.h"

namespace alamo {

enum PostKind { POST_STRAIN = 0, POST_STRESS = 1, POST_ENERGY = 2, POST_DW = 3, POST_W = 4 };

template <int SYM, int KIND>
__global__ void __launch_bounds__(256) k_post(Geom g, const double* __restrict__ C, const double* __restrict__ u,
                                              double* __restrict__ out, int voigt, const double* __restrict__ F0) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int idx[3] = {i, j, k};
    const int64_t st[3] = {1, g.sy, g.sz};
    double G[3][3];
    for (int p = 0; p < 3; p++)
        for (int d = 0; d < 3; d++) G[p][d] = d1(u + p * g.tot, m, st[d], idx[d], hi[d], g.dx[d]);
    double T[3][3];
    if (KIND == POST_DW || KIND == POST_W) {
        if (F0)
            for (int p = 0; p < 3; p++)
                for (int d = 0; d < 3; d++) G[p][d] -= F0[9 * m + 3 * p + d];
        double c[NCoef<SYM>::value];
        load_coef<double, SYM>(C, g.tot, m, c);
        c_dot<SYM>(c, G, T);
        if (KIND == POST_W) {
            double w = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) w += G[a][b] * T[a][b];
            out[m] = 0.5 * w;
        } else {
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) out[(3 * a + b) * g.tot + m] = T[a][b];
        }
        return;
    }
    if (KIND == POST_STRAIN) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) T[a][b] = 0.5 * (G[a][b] + G[b][a]);
    } else {
        double c[NCoef<SYM>::value];
        load_coef<double, SYM>(C, g.tot, m, c);
        c_dot<SYM>(c, G, T);
        if (KIND == POST_STRESS) {
            const double pa = psi_avg<double>(g, i, j, k);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) T[a][b] *= pa;
        } else {
            double w = 0.0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) w += 0.5 * T[a][b] * 0.5 * (G[a][b] + G[b][a]);
            out[m] += w;
            return;
        }
    }
    if (voigt) {
        const int va[6] = {0, 1, 2, 1, 2, 0}, vb[6] = {0, 1, 2, 2, 0, 1};
        for (int q = 0; q < 6; q++) out[q * g.tot + m] = T[va[q]][vb[q]];
    } else {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) out[(3 * a + b) * g.tot + m] = T[a][b];
    }
}

// Integrator::Base::Mechanics::Integrate, trac_hi[0] (Integrator/Base/Mechanics.H:391-448, 3-D):
// sum over the x-hi face's cells of the mean of sigma e_x at the cell's 4 nodes, times dy dz, with
// sigma = C : (grad u - F0) (the model's DW; F0 may be null).  partial[c * gridDim.x + block] =
// the block's share of component c.
template <int SYM>
__global__ void __launch_bounds__(RED_BLOCK) k_trac_xhi(Geom g, const double* __restrict__ C,
                                                        const double* __restrict__ F0, const double* __restrict__ u,
                                                        double* __restrict__ partial) {
    const int hi[3] = {g.n[0], g.n[1], g.n[2]};
    const int64_t st[3] = {1, g.sy, g.sz};
    const int64_t ncell = (int64_t)g.n[1] * g.n[2];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += (int64_t)gridDim.x * blockDim.x) {
        const int j0 = (int)(q % g.n[1]), k0 = (int)(q / g.n[1]);
        for (int c = 0; c < 4; c++) {
            const int i = g.n[0], j = j0 + (c & 1), k = k0 + (c >> 1);
            const int idx[3] = {i, j, k};
            const int64_t m = i + g.sy * j + g.sz * k;
            double G[3][3], S[3][3], cf[NCoef<SYM>::value];
            for (int p = 0; p < 3; p++)
                for (int d = 0; d < 3; d++) {
                    G[p][d] = d1(u + p * g.tot, m, st[d], idx[d], hi[d], g.dx[d]);
                    if (F0) G[p][d] -= F0[9 * m + 3 * p + d];
                }
            load_coef<double, SYM>(C, g.tot, m, cf);
            c_dot<SYM>(cf, G, S);
            for (int a = 0; a < 3; a++) acc[a] += 0.25 * S[a][0] * g.dx[1] * g.dx[2];
        }
    }
    __shared__ double sh[3][RED_BLOCK / 64];
    for (int a = 0; a < 3; a++) {
        const double v = wave_sum(acc[a]);
        if (threadIdx.x % 64 == 0) sh[a][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < RED_BLOCK / 64; w++) t += sh[threadIdx.x][w];
        partial[threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// out[node] = sum_n eta[n][node] table[n][q] for nrec values per model; eta holds nmodels SoA
// planes of nnodes values; out is AoS (nrec per node) or SoA (nrec planes)
__global__ void __launch_bounds__(256) k_blend(int64_t nnodes, int nmodels, int nrec, const double* __restrict__ eta,
                                               const double* __restrict__ table, double* __restrict__ out, int aos) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nnodes; v += (int64_t)gridDim.x * blockDim.x) {
        for (int q = 0; q < nrec; q++) {
            double s = 0.0;
            for (int n = 0; n < nmodels; n++) s += eta[n * nnodes + v] * table[n * nrec + q];
            out[aos ? v * nrec + q : q * nnodes + v] = s;
        }
    }
}

}  // namespace alamo
