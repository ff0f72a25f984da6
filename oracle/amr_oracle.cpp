// Oracle (test infrastructure only): the two-level AMR helpers of the nodal operator.
//
// Boxes are cell boxes of the fine level in fine index space (inclusive lo/hi, refinement
// ratio 2, lo even and hi odd so they coarsen exactly).  Fine-level fields passed here live on
// the fine box's nodal box grown by one node, [flo - 1, fhi + 2]^3, 3 SoA planes.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace oamr {

// Nodal fine mask of the coarse level (the m_nd_fine_mask that Operator.cpp:236,553 reads; built
// by amrex::MLNodeLinOp::buildMasks, AMReX 26.06, not vendored: values of
// amrex::nodelap_detail, crse_node 0, crse_fine_node 1, fine_node 2).  A coarse cell is fine when a
// coarsened fine box covers it; a node is fine_node when all eight cells around it are fine,
// crse_fine_node when some are, crse_node otherwise.  Cells outside the non-periodic domain take
// the value of their mirror image inside.
static void fine_mask(const int* nc, int nbox, const int* flo, const int* fhi, int* mask) {
    const int64_t C0 = nc[0], C1 = nc[1], C2 = nc[2];
    std::vector<uint8_t> cov(C0 * C1 * C2, 0);
    for (int b = 0; b < nbox; b++) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = flo[3 * b + d] / 2;
            hi[d] = (fhi[3 * b + d] + 1) / 2 - 1;
        }
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) cov[i + C0 * (j + C1 * k)] = 1;
    }
    auto refl = [](int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); };
    const int64_t N0 = C0 + 1, N1 = C1 + 1;
    for (int K = 0; K <= nc[2]; K++)
        for (int J = 0; J <= nc[1]; J++)
            for (int I = 0; I <= nc[0]; I++) {
                int nf = 0;
                for (int c = 0; c < 8; c++) {
                    const int ci = refl(I - 1 + (c & 1), nc[0]), cj = refl(J - 1 + ((c >> 1) & 1), nc[1]),
                              ck = refl(K - 1 + (c >> 2), nc[2]);
                    nf += cov[ci + C0 * (cj + C1 * ck)];
                }
                mask[I + N0 * (J + N1 * K)] = nf == 8 ? 2 : (nf > 0 ? 1 : 0);
            }
}

// Operator<Grid::Node>::reflux (Operator.cpp:526-633) for one fine box: on the coarse nodes of the
// coarsened box whose mask is crse_fine_node or fine_node, the coarse residual becomes the fine
// residual restricted with the full-weighting stencil; on the lo/hi faces, edges and corners of the
// working box (the coarsened nodal box grown by one node, clipped to the domain) the 2-D 9-point,
// 1-D 3-point and injection stencils.  Only the coarsened box's own nodes are written back
// (the ParallelCopy of :622).
static void reflux(const int* nc, const int* flo, const int* fhi, const int* mask, double* crse, const double* fine) {
    int vlo[3], vhi[3], blo[3], bhi[3], gf[3];
    for (int d = 0; d < 3; d++) {
        vlo[d] = flo[d] / 2;
        vhi[d] = (fhi[d] + 1) / 2;
        blo[d] = std::max(vlo[d] - 1, 0);
        bhi[d] = std::min(vhi[d] + 1, nc[d]);
        gf[d] = fhi[d] - flo[d] + 4;  // grown fine nodal box
    }
    const int64_t N0 = nc[0] + 1, N1 = nc[1] + 1, Nc = N0 * N1 * (nc[2] + 1);
    const int64_t Vf = (int64_t)gf[0] * gf[1] * gf[2];
    for (int n = 0; n < 3; n++)
        for (int K = vlo[2]; K <= vhi[2]; K++)
            for (int J = vlo[1]; J <= vhi[1]; J++)
                for (int I = vlo[0]; I <= vhi[0]; I++) {
                    const int64_t mc = I + N0 * (J + N1 * K);
                    if (mask[mc] != 1 && mask[mc] != 2) continue;
                    const int i = 2 * I - (flo[0] - 1), j = 2 * J - (flo[1] - 1), k = 2 * K - (flo[2] - 1);
                    auto f = [&](int a, int b, int c) {
                        return fine[n * Vf + (i + a) + (int64_t)gf[0] * ((j + b) + (int64_t)gf[1] * (k + c))];
                    };
                    const bool xe = I == blo[0] || I == bhi[0], ye = J == blo[1] || J == bhi[1],
                               ze = K == blo[2] || K == bhi[2];
                    double v;
                    if (xe && ye && ze) v = f(0, 0, 0);
                    else if (ye && ze) v = 0.25 * f(-1, 0, 0) + 0.5 * f(0, 0, 0) + 0.25 * f(1, 0, 0);
                    else if (ze && xe) v = 0.25 * f(0, -1, 0) + 0.5 * f(0, 0, 0) + 0.25 * f(0, 1, 0);
                    else if (xe && ye) v = 0.25 * f(0, 0, -1) + 0.5 * f(0, 0, 0) + 0.25 * f(0, 0, 1);
                    else if (xe)
                        v = (f(0, -1, -1) + 2.0 * f(0, 0, -1) + f(0, 1, -1) + 2.0 * f(0, -1, 0) + 4.0 * f(0, 0, 0) +
                             2.0 * f(0, 1, 0) + f(0, -1, 1) + 2.0 * f(0, 0, 1) + f(0, 1, 1)) / 16.0;
                    else if (ye)
                        v = (f(-1, 0, -1) + 2.0 * f(-1, 0, 0) + f(-1, 0, 1) + 2.0 * f(0, 0, -1) + 4.0 * f(0, 0, 0) +
                             2.0 * f(0, 0, 1) + f(1, 0, -1) + 2.0 * f(1, 0, 0) + f(1, 0, 1)) / 16.0;
                    else if (ze)
                        v = (f(-1, -1, 0) + 2.0 * f(0, -1, 0) + f(1, -1, 0) + 2.0 * f(-1, 0, 0) + 4.0 * f(0, 0, 0) +
                             2.0 * f(1, 0, 0) + f(-1, 1, 0) + 2.0 * f(0, 1, 0) + f(1, 1, 0)) / 16.0;
                    else {
                        double c8 = 0, c12 = 0, c6 = 0;
                        for (int a = -1; a <= 1; a++)
                            for (int b = -1; b <= 1; b++)
                                for (int c = -1; c <= 1; c++) {
                                    const int nz = (a != 0) + (b != 0) + (c != 0);
                                    if (nz == 3) c8 += f(a, b, c);
                                    else if (nz == 2) c12 += f(a, b, c);
                                    else if (nz == 1) c6 += f(a, b, c);
                                }
                        v = c8 / 64.0 + c12 / 32.0 + c6 / 16.0 + f(0, 0, 0) / 8.0;
                    }
                    crse[n * Nc + mc] = v;
                }
}

// Operator<Grid::Node>::averageDownSolutionRHS (Operator.cpp:435-447): amrex::average_down of a
// nodal field with ratio 2 is injection, crse(I) = fine(2 I), on the coarsened box's nodes.
static void average_down(const int* nc, const int* flo, const int* fhi, double* crse, const double* fine) {
    int gf[3];
    for (int d = 0; d < 3; d++) gf[d] = fhi[d] - flo[d] + 4;
    const int64_t N0 = nc[0] + 1, N1 = nc[1] + 1, Nc = N0 * N1 * (nc[2] + 1);
    const int64_t Vf = (int64_t)gf[0] * gf[1] * gf[2];
    for (int n = 0; n < 3; n++)
        for (int K = flo[2] / 2; K <= (fhi[2] + 1) / 2; K++)
            for (int J = flo[1] / 2; J <= (fhi[1] + 1) / 2; J++)
                for (int I = flo[0] / 2; I <= (fhi[0] + 1) / 2; I++) {
                    const int i = 2 * I - (flo[0] - 1), j = 2 * J - (flo[1] - 1), k = 2 * K - (flo[2] - 1);
                    crse[n * Nc + I + N0 * (J + N1 * K)] = fine[n * Vf + i + (int64_t)gf[0] * (j + (int64_t)gf[1] * k)];
                }
}

// Operator<Grid::Node>::fixUpResidualMask (Operator.cpp:230-251): residual-mask entries of the
// crse_fine nodes become 1; the others keep their value.
static void fixup_residual_mask(int64_t nnodes, const int* mask, int* resmsk) {
    for (int64_t q = 0; q < nnodes; q++)
        if (mask[q] == 1) resmsk[q] = 1;
}

}  // namespace oamr

extern "C" {
void oamr_fine_mask(const int* nc, int nbox, const int* flo, const int* fhi, int* mask) {
    oamr::fine_mask(nc, nbox, flo, fhi, mask);
}
void oamr_reflux(const int* nc, const int* flo, const int* fhi, const int* mask, double* crse, const double* fine) {
    oamr::reflux(nc, flo, fhi, mask, crse, fine);
}
void oamr_average_down(const int* nc, const int* flo, const int* fhi, double* crse, const double* fine) {
    oamr::average_down(nc, flo, fhi, crse, fine);
}
void oamr_fixup_residual_mask(int64_t nnodes, const int* mask, int* resmsk) {
    oamr::fixup_residual_mask(nnodes, mask, resmsk);
}
}
