// Two-level AMR helpers of the nodal operator (Operator<Grid::Node>, src/Operator/Operator.cpp):
// the nodal fine mask, reflux of the fine residual onto the coarse level, average-down of the
// solution and the residual-mask fix-up.  The coarse level is one box covering the domain of
// nc cells; a fine box is an inclusive cell box [flo, fhi] in fine index space (ratio 2, flo
// even, fhi odd), and its fields live on the fine nodal box grown by one node, [flo - 1, fhi + 2]^3.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace alamo {
namespace amr {

constexpr int MAX_BOXES = 64;
struct Boxes {
    int n;
    int lo[MAX_BOXES][3], hi[MAX_BOXES][3];  // coarsened cell boxes (inclusive)
};
struct Dom {
    int nc[3];
    int64_t N0, N1, tot;  // coarse nodes per row / plane, total nodes
};
struct FBox {
    int vlo[3], vhi[3];  // coarse nodes of the coarsened box (its nodal box)
    int blo[3], bhi[3];  // working box: grown by one node, clipped to the domain
    int flo[3];          // fine cell lo (fine index space)
    int gf[3];           // grown fine nodal box extents
    int64_t vf;          // nodes of the grown fine box
};

// m_nd_fine_mask (read at Operator.cpp:236,553; built by amrex::MLNodeLinOp::buildMasks):
// 2 fine_node when the eight cells around the node are covered by the fine level, 1
// crse_fine_node when some are, 0 crse_node otherwise; cells outside the domain mirror inside.
__global__ void k_fine_mask(Dom d, Boxes b, int* __restrict__ mask) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.tot) return;
    const int I = (int)(q % d.N0), J = (int)((q / d.N0) % d.N1), K = (int)(q / (d.N0 * d.N1));
    int nf = 0;
    for (int c = 0; c < 8; c++) {
        int ci = I - 1 + (c & 1), cj = J - 1 + ((c >> 1) & 1), ck = K - 1 + (c >> 2);
        ci = ci < 0 ? 0 : (ci >= d.nc[0] ? d.nc[0] - 1 : ci);
        cj = cj < 0 ? 0 : (cj >= d.nc[1] ? d.nc[1] - 1 : cj);
        ck = ck < 0 ? 0 : (ck >= d.nc[2] ? d.nc[2] - 1 : ck);
        bool cov = false;
        for (int t = 0; t < b.n && !cov; t++)
            cov = ci >= b.lo[t][0] && ci <= b.hi[t][0] && cj >= b.lo[t][1] && cj <= b.hi[t][1] && ck >= b.lo[t][2] &&
                  ck <= b.hi[t][2];
        nf += cov;
    }
    mask[q] = nf == 8 ? 2 : (nf > 0 ? 1 : 0);
}

// Operator<Grid::Node>::reflux (Operator.cpp:526-633), one fine box: one work-item per coarse
// node of the coarsened box; stencil by position in the working box (corner: injection, edge:
// 1-D 3-point, face: 2-D 9-point, else 27-point full weighting).
__global__ void k_reflux(Dom d, FBox fb, const int* __restrict__ mask, double* __restrict__ crse,
                         const double* __restrict__ fine) {
    const int e0 = fb.vhi[0] - fb.vlo[0] + 1, e1 = fb.vhi[1] - fb.vlo[1] + 1, e2 = fb.vhi[2] - fb.vlo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)e0 * e1 * e2) return;
    const int I = fb.vlo[0] + (int)(q % e0), J = fb.vlo[1] + (int)((q / e0) % e1), K = fb.vlo[2] + (int)(q / ((int64_t)e0 * e1));
    const int64_t mc = I + d.N0 * (J + d.N1 * K);
    const int mk = mask[mc];
    if (mk != 1 && mk != 2) return;
    const int i = 2 * I - (fb.flo[0] - 1), j = 2 * J - (fb.flo[1] - 1), k = 2 * K - (fb.flo[2] - 1);
    const bool xe = I == fb.blo[0] || I == fb.bhi[0], ye = J == fb.blo[1] || J == fb.bhi[1],
               ze = K == fb.blo[2] || K == fb.bhi[2];
    const int64_t sy = fb.gf[0], sz = (int64_t)fb.gf[0] * fb.gf[1];
    for (int n = 0; n < 3; n++) {
        const double* F = fine + n * fb.vf + i + sy * j + sz * k;
        auto f = [&](int a, int b, int c) { return F[a + sy * b + sz * c]; };
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
#pragma unroll
            for (int c = -1; c <= 1; c++)
#pragma unroll
                for (int b = -1; b <= 1; b++)
#pragma unroll
                    for (int a = -1; a <= 1; a++) {
                        const int nz = (a != 0) + (b != 0) + (c != 0);
                        const double x = f(a, b, c);
                        if (nz == 3) c8 += x;
                        else if (nz == 2) c12 += x;
                        else if (nz == 1) c6 += x;
                    }
            v = c8 / 64.0 + c12 / 32.0 + c6 / 16.0 + f(0, 0, 0) / 8.0;
        }
        crse[n * d.tot + mc] = v;
    }
}

// averageDownSolutionRHS (Operator.cpp:435-447): nodal injection crse(I) = fine(2 I)
__global__ void k_average_down(Dom d, FBox fb, double* __restrict__ crse, const double* __restrict__ fine) {
    const int e0 = fb.vhi[0] - fb.vlo[0] + 1, e1 = fb.vhi[1] - fb.vlo[1] + 1, e2 = fb.vhi[2] - fb.vlo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)e0 * e1 * e2) return;
    const int I = fb.vlo[0] + (int)(q % e0), J = fb.vlo[1] + (int)((q / e0) % e1), K = fb.vlo[2] + (int)(q / ((int64_t)e0 * e1));
    const int i = 2 * I - (fb.flo[0] - 1), j = 2 * J - (fb.flo[1] - 1), k = 2 * K - (fb.flo[2] - 1);
    const int64_t mf = i + (int64_t)fb.gf[0] * (j + (int64_t)fb.gf[1] * k);
    const int64_t mc = I + d.N0 * (J + d.N1 * K);
    for (int n = 0; n < 3; n++) crse[n * d.tot + mc] = fine[n * fb.vf + mf];
}

// fixUpResidualMask (Operator.cpp:230-251): resmsk = 1 on crse_fine nodes
__global__ void k_fixup_resmask(int64_t n, const int* __restrict__ mask, int* __restrict__ resmsk) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && mask[q] == 1) resmsk[q] = 1;
}


// ---------------------------------------------------------------------------
// Composite multi-level solve (alamo_amr_*): every AMR level is one full-domain nodal array at
// its own resolution, its region given by a node mask (k_fine_mask of the level's own boxes:
// 2 region interior, 1 region boundary, 0 outside) and its coverage by the next level by a
// second mask (k_fine_mask of the next level's boxes, coarsened).  The kernels below move fields
// between the levels under those masks; mask arrays hold one int per node, fields 3 planes.
// ---------------------------------------------------------------------------
// dst = src on the nodes whose mask lies in [lo, hi] (3 planes of n nodes)
__global__ void k_masked_copy(int64_t n, const int* __restrict__ mask, int lo, int hi, double* __restrict__ dst,
                              const double* __restrict__ src) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int m = mask[q];
    if (m < lo || m > hi) return;
    for (int c = 0; c < 3; c++) dst[c * n + q] = src[c * n + q];
}
// x = 0 on the nodes whose mask lies outside [lo, hi]
__global__ void k_masked_zero(int64_t n, const int* __restrict__ mask, int lo, int hi, double* __restrict__ x) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int m = mask[q];
    if (m >= lo && m <= hi) return;
    for (int c = 0; c < 3; c++) x[c * n + q] = 0.0;
}
// nodal injection crse(I) = fine(2 I) on the coarse nodes the next level covers (fine mask >= 1):
// averageDownSolutionRHS (Operator.cpp:435-447) with full-domain level arrays
__global__ void k_inject_covered(int64_t nc, int Nc0, int Nc1, int Nf0, int Nf1, int64_t nf,
                                 const int* __restrict__ fmask, double* __restrict__ crse,
                                 const double* __restrict__ fine) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc || fmask[q] < 1) return;
    const int I = (int)(q % Nc0), J = (int)((q / Nc0) % Nc1), K = (int)(q / ((int64_t)Nc0 * Nc1));
    const int64_t mf = 2 * I + (int64_t)Nf0 * (2 * J + (int64_t)Nf1 * 2 * K);
    for (int c = 0; c < 3; c++) crse[c * nc + q] = fine[c * nf + mf];
}

// out(I) = fine(2 I) on the coarse nodes strictly covered (fine mask 2), zero elsewhere: the finer
// correction as the coarse rows at the coarse/fine interface read it
__global__ void k_inject_strict(int64_t nc, int Nc0, int Nc1, int Nf0, int Nf1, int64_t nf,
                                const int* __restrict__ fmask, double* __restrict__ out,
                                const double* __restrict__ fine) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    if (fmask[q] != 2) {
        for (int c = 0; c < 3; c++) out[c * nc + q] = 0.0;
        return;
    }
    const int I = (int)(q % Nc0), J = (int)((q / Nc0) % Nc1), K = (int)(q / ((int64_t)Nc0 * Nc1));
    const int64_t mf = 2 * I + (int64_t)Nf0 * (2 * J + (int64_t)Nf1 * 2 * K);
    for (int c = 0; c < 3; c++) out[c * nc + q] = fine[c * nf + mf];
}

}  // namespace amr
}  // namespace alamo
