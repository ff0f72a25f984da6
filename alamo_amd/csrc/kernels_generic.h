// Node-parallel kernels: one work-item per node, any symmetry, FP64 or FP32.
// They serve every multigrid level and are the reference the tiled kernels of
// kernels_iso.h are checked against.
#pragma once
#include "node_ops.h"

namespace alamo {

enum Mode { APPLY = 0, RESID = 1, JACOBI = 2, DIAG = 3 };

template <typename T> __device__ __forceinline__ const T* smoother_diag(const Geom& g);
template <> __device__ __forceinline__ const double* smoother_diag<double>(const Geom& g) { return g.dsm; }
template <> __device__ __forceinline__ const float* smoother_diag<float>(const Geom& g) { return g.dsmf; }

// One smoothing update of element q from its row result f = (A x)_q: the reference's damped point
// Jacobi x + w (b - A x) / a_qq written as (1-w) x + w (b - (f - D x)) / D (Operator/Operator.cpp:
// 109-146), or, when the level has an l1 smoother diagonal, x + (b - f) / dsm_q (a zero row is
// left as it is).
template <typename T>
__device__ __forceinline__ T jacobi_update(const Geom& g, int64_t q, T x, T b, T f, T D, T omega) {
    const T* ds = smoother_diag<T>(g);
    if (ds) {
        const T d = ds[q];
        return d != T(0) ? x + (b - f) / d : x;
    }
    return (T(1) - omega) * x + omega * (b - (f - D * x)) / D;
}

// (i, j, k) of element m of a level's array.
// Levels below 2^31 elements (every level up to 1290^3 nodes) divide in 32 bits: a 64-bit
// division is a long software sequence on the GPU, a 32-bit one a few instructions.
__device__ __forceinline__ void node_ijk(const Geom& g, int64_t m, int& i, int& j, int& k) {
    if (g.tot <= 0x7fffffff) {
        const uint32_t mm = (uint32_t)m, sz = (uint32_t)g.sz, sy = (uint32_t)g.sy;
        const uint32_t kk = mm / sz, r = mm - kk * sz, jj = r / sy;
        k = (int)kk;
        j = (int)jj;
        i = (int)(r - jj * sy);
        return;
    }
    k = (int)(m / g.sz);
    int64_t r = m - (int64_t)k * g.sz;
    j = (int)(r / g.sy);
    i = (int)(r - (int64_t)j * g.sy);
}
// element offset of compact node index v (i fastest over the N[0] x N[1] x N[2] nodes)
__device__ __forceinline__ int64_t node_offset(const Geom& g, int64_t v) {
    if (g.tot <= 0x7fffffff) {
        const uint32_t vv = (uint32_t)v, N0 = (uint32_t)g.N[0], N01 = N0 * (uint32_t)g.N[1];
        const uint32_t k = vv / N01, r = vv - k * N01, j = r / N0;
        return (int64_t)(r - j * N0) + g.sy * j + g.sz * k;
    }
    const int64_t N01 = (int64_t)g.N[0] * g.N[1];
    const int k = (int)(v / N01);
    const int64_t r = v - (int64_t)k * N01;
    const int j = (int)(r / g.N[0]);
    const int i = (int)(r - (int64_t)j * g.N[0]);
    return i + g.sy * j + g.sz * k;
}

// Row result of one node: Au (MODE APPLY/RESID/JACOBI) and/or the diagonal.
template <typename T, int SYM, bool NEED_DIAG>
__device__ __forceinline__ void node_row(const Geom& g, const BCTable& bc, const T* __restrict__ C,
                                         const T* __restrict__ u, int i, int j, int k, T f[3], T D[3]) {
    const bool bnd = i == 0 || j == 0 || k == 0 || i == g.n[0] || j == g.n[1] || k == g.n[2];
    if (bnd) {
        if (u) apply_boundary<T, SYM>(g, bc, C, u, i, j, k, f);
        if (NEED_DIAG) diag_boundary<T, SYM>(g, bc, C, i, j, k, D);
    } else {
        const int64_t m = i + g.sy * j + g.sz * k;
        T c[NCoef<SYM>::value];
        if (u) apply_interior<T, SYM>(g, C, u, m, f, c);
        else load_coef<T, SYM>(C, g.tot, m, c);
        if (NEED_DIAG) diag_interior<T, SYM>(g, c, D);
        psi_interior<T, SYM>(g, C, u, i, j, k, f, NEED_DIAG ? D : nullptr);
    }
}

// out = A u | b - A u | Jacobi(u) | D, over all nodes of the level.
// Jacobi: x_new = (1-w) x + w (b - (A x - D x)) / D   (Operator/Operator.cpp:109-146)
template <typename T, int SYM, int MODE>
__global__ void __launch_bounds__(256) k_node(Geom g, BCTable bc, const T* __restrict__ C, const T* __restrict__ u,
                                              const T* __restrict__ b, T* __restrict__ out, T omega) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    T f[3] = {T(0), T(0), T(0)}, D[3] = {T(1), T(1), T(1)};
    node_row<T, SYM, MODE == JACOBI || MODE == DIAG>(g, bc, C, MODE == DIAG ? nullptr : u, i, j, k, f, D);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int64_t q = p * g.tot + m;
        if (MODE == APPLY) out[q] = f[p];
        else if (MODE == RESID) out[q] = b[q] - f[p];
        else if (MODE == DIAG) out[q] = D[p];
        else out[q] = jacobi_update(g, q, u[q], b[q], f[p], D[p], omega);
    }
}

// l1-Jacobi smoother diagonal of every node of the level (3 planes)
template <typename T, int SYM>
__global__ void __launch_bounds__(256) k_l1_diag(Geom g, BCTable bc, const T* __restrict__ C, T* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    T D[3];
    l1_row<T, SYM>(g, bc, C, i, j, k, D);
#pragma unroll
    for (int p = 0; p < 3; p++) out[p * g.tot + m] = D[p];
}

// Explicit rows of the level's operator (no psi terms), 243 planes: plane (27 q + o) 3 + c holds the
// weight of u_c at offset slot o (13 + dx + 3 dy + 9 dz) in row component q, per node (neighbouring
// work-items of the gather below read neighbouring elements).
template <int SYM>
__global__ void __launch_bounds__(256) k_row_table(Geom g, BCTable bc, const double* __restrict__ C,
                                                   double* __restrict__ tab) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    row_coeffs<double, SYM>(g, bc, C, i, j, k, [&](int q, const double(&acc)[27][3]) {
        for (int o = 0; o < 27; o++)
#pragma unroll
            for (int c = 0; c < 3; c++) tab[((int64_t)(27 * q + o) * 3 + c) * g.tot + m] = acc[o][c];
    });
}

// out = A^T r from the row table: out_d(n) = sum over slots o and components q of the weight of
// u_d(n) in row q of node n - offset(o), times r_q(n - offset(o)).  The stencil reaches one node,
// so these 27 rows are all that hold a weight of u(n).
__global__ void __launch_bounds__(256) k_row_apply_t(Geom g, const double* __restrict__ tab,
                                                     const double* __restrict__ r, double* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.tot) return;
    int i, j, k;
    node_ijk(g, m, i, j, k);
    if (i >= g.N[0]) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 27; o++) {
        const int a = i - (o % 3 - 1), b = j - ((o / 3) % 3 - 1), c = k - (o / 9 - 1);
        if (a < 0 || b < 0 || c < 0 || a > g.n[0] || b > g.n[1] || c > g.n[2]) continue;
        const int64_t mr = a + g.sy * b + g.sz * c;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double rq = r[q * g.tot + mr];
#pragma unroll
            for (int d = 0; d < 3; d++) s[d] += tab[((int64_t)(27 * q + o) * 3 + d) * g.tot + mr] * rq;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) out[d * g.tot + m] = s[d];
}

// ---------------------------------------------------------------------------
// Reductions: per-block partial (max |x|, sum x^2) or (sum a*b), wavefront
// shuffles inside a 64-wide wavefront, LDS across the 4 wavefronts of a block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int RED_BLOCK = 256;

template <typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_norm_partial(const T* __restrict__ x, int64_t n,
                                                            double* __restrict__ partial) {
    double mx = 0.0, ss = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)x[q];
        // a NaN poisons the max so that a diverged solve cannot report a small norm
        mx = (v != v) ? v : fmax(mx, fabs(v));
        ss += v * v;
    }
    mx = wave_max(mx);
    ss = wave_sum(ss);
    __shared__ double smx[RED_BLOCK / 64], sss[RED_BLOCK / 64];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { smx[w] = mx; sss[w] = ss; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = smx[0], s = sss[0];
        for (int q = 1; q < RED_BLOCK / 64; q++) { a = (smx[q] != smx[q]) ? smx[q] : fmax(a, smx[q]); s += sss[q]; }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(RED_BLOCK) k_dot_partial(const T* __restrict__ a, const T* __restrict__ b,
                                                           int64_t n, double* __restrict__ partial) {
    double ss = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        ss += (double)a[q] * (double)b[q];
    ss = wave_sum(ss);
    __shared__ double sss[RED_BLOCK / 64];
    if (threadIdx.x % 64 == 0) sss[threadIdx.x / 64] = ss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) s += sss[q];
        partial[blockIdx.x] = s;
    }
}

// Final pass of the two-level norm reduction: out[0] = max, out[1] = sum of squares.
__global__ void __launch_bounds__(RED_BLOCK) k_norm_final(const double* __restrict__ partial, int nb,
                                                          double* __restrict__ out) {
    double mx = 0.0, ss = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
        const double a = partial[2 * q];
        mx = (a != a) ? a : fmax(mx, a);
        ss += partial[2 * q + 1];
    }
    mx = wave_max(mx);
    ss = wave_sum(ss);
    __shared__ double smx[RED_BLOCK / 64], sss[RED_BLOCK / 64];
    if (threadIdx.x % 64 == 0) { smx[threadIdx.x / 64] = mx; sss[threadIdx.x / 64] = ss; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = smx[0], s = sss[0];
        for (int q = 1; q < RED_BLOCK / 64; q++) { a = (smx[q] != smx[q]) ? smx[q] : fmax(a, smx[q]); s += sss[q]; }
        out[0] = a;
        out[1] = s;
    }
}

__global__ void __launch_bounds__(RED_BLOCK) k_sum_final(const double* __restrict__ partial, int nb,
                                                         double* __restrict__ out) {
    double ss = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) ss += partial[q];
    ss = wave_sum(ss);
    __shared__ double sss[RED_BLOCK / 64];
    if (threadIdx.x % 64 == 0) sss[threadIdx.x / 64] = ss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < RED_BLOCK / 64; q++) s += sss[q];
        out[0] = s;
    }
}

// ---------------------------------------------------------------------------
// Grid transfers.
// ---------------------------------------------------------------------------
// Full-weighting restriction, one work-item per coarse node, `np` SoA planes.
// Operator/Operator.cpp:294-352 (residual) and Operator/Elastic.cpp:786-829 (stiffness).
template <typename T>
// xo (may be null; 3 planes): also the first damped-Jacobi sweep of the coarse level from a zero
// guess, xo = omega crse / Dc (0 where Dc is 0), the value the V-cycle's pre-smoothing would
// compute from crse next
__global__ void __launch_bounds__(256) k_restrict(Geom gc, Geom gf, int np, const T* __restrict__ fine,
                                                  T* __restrict__ crse, T* __restrict__ xo = nullptr,
                                                  const T* __restrict__ Dc = nullptr, T omega = T(0)) {
    const int64_t mc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (mc >= gc.tot) return;
    int I, J, K;
    node_ijk(gc, mc, I, J, K);
    if (I >= gc.N[0]) return;
    const bool bx = I == 0 || I == gc.n[0], by = J == 0 || J == gc.n[1], bz = K == 0 || K == gc.n[2];
    const int64_t mf = 2 * I + gf.sy * (2 * J) + gf.sz * (2 * K);
    const int64_t sx = 1, sy = gf.sy, sz = gf.sz;
    for (int p = 0; p < np; p++) {
        const T* F = fine + p * gf.tot + mf;
        auto at = [&](int a, int b, int c) { return F[a * sx + b * sy + c * sz]; };
        T v;
        if (bx && by && bz) v = at(0, 0, 0);
        else if (by && bz) v = T(0.25) * at(-1, 0, 0) + T(0.5) * at(0, 0, 0) + T(0.25) * at(1, 0, 0);
        else if (bz && bx) v = T(0.25) * at(0, -1, 0) + T(0.5) * at(0, 0, 0) + T(0.25) * at(0, 1, 0);
        else if (bx && by) v = T(0.25) * at(0, 0, -1) + T(0.5) * at(0, 0, 0) + T(0.25) * at(0, 0, 1);
        else if (bx)
            v = (at(0, -1, -1) + T(2) * at(0, 0, -1) + at(0, 1, -1) + T(2) * at(0, -1, 0) + T(4) * at(0, 0, 0) +
                 T(2) * at(0, 1, 0) + at(0, -1, 1) + T(2) * at(0, 0, 1) + at(0, 1, 1)) / T(16);
        else if (by)
            v = (at(-1, 0, -1) + T(2) * at(-1, 0, 0) + at(-1, 0, 1) + T(2) * at(0, 0, -1) + T(4) * at(0, 0, 0) +
                 T(2) * at(0, 0, 1) + at(1, 0, -1) + T(2) * at(1, 0, 0) + at(1, 0, 1)) / T(16);
        else if (bz)
            v = (at(-1, -1, 0) + T(2) * at(0, -1, 0) + at(1, -1, 0) + T(2) * at(-1, 0, 0) + T(4) * at(0, 0, 0) +
                 T(2) * at(1, 0, 0) + at(-1, 1, 0) + T(2) * at(0, 1, 0) + at(1, 1, 0)) / T(16);
        else {
            T c8 = T(0), c12 = T(0), c6 = T(0);
#pragma unroll
            for (int a = -1; a <= 1; a++)
#pragma unroll
                for (int b = -1; b <= 1; b++)
#pragma unroll
                    for (int c = -1; c <= 1; c++) {
                        const int nz = (a != 0) + (b != 0) + (c != 0);
                        if (nz == 3) c8 += at(a, b, c);
                        else if (nz == 2) c12 += at(a, b, c);
                        else if (nz == 1) c6 += at(a, b, c);
                    }
            v = c8 / T(64) + c12 / T(32) + c6 / T(16) + at(0, 0, 0) / T(8);
        }
        crse[p * gc.tot + mc] = v;
        if (xo) {
            const T d = Dc[p * gc.tot + mc];
            xo[p * gc.tot + mc] = d != T(0) ? omega * v / d : T(0);
        }
    }
}

// Restriction of the psi cell field (Operator/Elastic.cpp:839-907): coarse cell (I,J,K) from the
// fine cells around (2I,2J,2K) with the nodal full-weighting weights; the branch bounds are the
// coarse nodal box [0, n], which a cell index reaches only on the lo side.
__global__ void __launch_bounds__(256) k_restrict_psi(Geom gc, Geom gf, const double* __restrict__ fine,
                                                      double* __restrict__ crse) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n01 = (int64_t)gc.n[0] * gc.n[1];
    if (q >= n01 * gc.n[2]) return;
    const int K = (int)(q / n01), J = (int)((q / gc.n[0]) % gc.n[1]), I = (int)(q % gc.n[0]);
    const int i = 2 * I, j = 2 * J, k = 2 * K;
    auto F = [&](int a, int b, int c) { return fine[(i + a) + (int64_t)gf.n[0] * ((j + b) + (int64_t)gf.n[1] * (k + c))]; };
    const bool xl = I == 0, yl = J == 0, zl = K == 0;
    double v;
    if (xl && yl && zl) v = F(0, 0, 0);
    else if (yl && zl) v = F(-1, 0, 0) * 0.25 + F(0, 0, 0) * 0.5 + F(1, 0, 0) * 0.25;
    else if (zl && xl) v = F(0, -1, 0) * 0.25 + F(0, 0, 0) * 0.5 + F(0, 1, 0) * 0.25;
    else if (xl && yl) v = F(0, 0, -1) * 0.25 + F(0, 0, 0) * 0.5 + F(0, 0, 1) * 0.25;
    else if (xl)
        v = (F(0, -1, -1) + F(0, 0, -1) * 2.0 + F(0, 1, -1) + F(0, -1, 0) * 2.0 + F(0, 0, 0) * 4.0 + F(0, 1, 0) * 2.0 +
             F(0, -1, 1) + F(0, 0, 1) * 2.0 + F(0, 1, 1)) / 16.0;
    else if (yl)
        v = (F(-1, 0, -1) + F(-1, 0, 0) * 2.0 + F(-1, 0, 1) + F(0, 0, -1) * 2.0 + F(0, 0, 0) * 4.0 + F(0, 0, 1) * 2.0 +
             F(1, 0, -1) + F(1, 0, 0) * 2.0 + F(1, 0, 1)) / 16.0;
    else if (zl)
        v = (F(-1, -1, 0) + F(0, -1, 0) * 2.0 + F(1, -1, 0) + F(-1, 0, 0) * 2.0 + F(0, 0, 0) * 4.0 + F(1, 0, 0) * 2.0 +
             F(-1, 1, 0) + F(0, 1, 0) * 2.0 + F(1, 1, 0)) / 16.0;
    else {
        double c8 = 0, c12 = 0, c6 = 0;
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int c = -1; c <= 1; c++) {
                    const int nz = (a != 0) + (b != 0) + (c != 0);
                    if (nz == 3) c8 += F(a, b, c);
                    else if (nz == 2) c12 += F(a, b, c);
                    else if (nz == 1) c6 += F(a, b, c);
                }
        v = c8 / 64.0 + c12 / 32.0 + c6 / 16.0 + F(0, 0, 0) / 8.0;
    }
    crse[q] = v;
}

// fine += I(crse), trilinear nodal prolongation by parity (Operator/Operator.cpp:399-428).
// One work-item per fine node, i fastest (coalesced loads and stores; a wavefront spans at most
// two rows, so the j, k parities diverge only where it crosses a row end).  The x parity is per
// lane, so every lane evaluates the even and the odd formula of its row case and keeps its own.  The odd formula's
// x + 1 coarse neighbour is read at offset `odd` (0 on even lanes), so no read leaves the level.
// Each node keeps its case's formula and term order.
// SET: fine = I(crse) instead (the F-cycle's initial guess of a level, without clearing it first)
template <typename T, bool SET = false>
__global__ void __launch_bounds__(256) k_interp_add(Geom gc, Geom gf, const T* __restrict__ crse,
                                                    T* __restrict__ fine) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nt = (int64_t)gf.N[0] * gf.N[1] * gf.N[2];
    if (t >= nt) return;
    int i, j, k;
    if (nt <= 0x7fffffff) {
        const uint32_t tt = (uint32_t)t, N0 = (uint32_t)gf.N[0], N1 = (uint32_t)gf.N[1];
        const uint32_t row = tt / N0, kk = row / N1;
        i = (int)(tt - row * N0);
        k = (int)kk;
        j = (int)(row - kk * N1);
    } else {
        const int64_t row = t / gf.N[0];
        i = (int)(t - row * gf.N[0]);
        k = (int)(row / gf.N[1]);
        j = (int)(row - (int64_t)k * gf.N[1]);
    }
    const int I = i >> 1, o = i & 1, J = j / 2, K = k / 2;
    const bool ej = (j & 1) == 0, ek = (k & 1) == 0;
    const int64_t mc = I + gc.sy * J + gc.sz * K;
    const int64_t mf = i + gf.sy * j + gf.sz * k;
    const int64_t sy = gc.sy, sz = gc.sz;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* Cp = crse + p * gc.tot + mc;
        // at(a, b, c) with a = 1 read at x offset o
        auto at = [&](int a, int b, int c) { return Cp[a * o + b * sy + c * sz]; };
        T ve, vo;
        if (ej && ek) {
            ve = at(0, 0, 0);
            vo = T(0.5) * (at(0, 0, 0) + at(1, 0, 0));
        } else if (ek) {
            ve = T(0.5) * (at(0, 0, 0) + at(0, 1, 0));
            vo = T(0.25) * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(1, 1, 0));
        } else if (ej) {
            ve = T(0.5) * (at(0, 0, 0) + at(0, 0, 1));
            vo = T(0.25) * (at(0, 0, 0) + at(0, 0, 1) + at(1, 0, 0) + at(1, 0, 1));
        } else {
            ve = T(0.25) * (at(0, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1));
            vo = T(0.125) * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1) + at(1, 0, 1) +
                             at(1, 1, 0) + at(1, 1, 1));
        }
        if (SET) fine[p * gf.tot + mf] = o ? vo : ve;
        else fine[p * gf.tot + mf] += o ? vo : ve;
    }
}

// The same prolongation by fine x-rows: a wavefront takes 64 coarse nodes I < N0 - 1 of one fine row
// (j, k) and writes its fine nodes 2I and 2I + 1 (the lane of I = N0 - 2 also the row's last node),
// so every coarse value is one coalesced load per needed coarse row (one to four by the parities of
// j and k; I + 1 by a second, shifted load) instead of up to eight loads per fine node, and a row of
// 2^k + 1 nodes takes 2^(k-7) wavefronts.  Same sums in the same order as k_interp_add.  Needs
// gf.n = 2 gc.n.
constexpr int INTERP_ROW_WAVES = 4;
template <typename T, bool SET = false>
__global__ void __launch_bounds__(64 * INTERP_ROW_WAVES) k_interp_rows(Geom gc, Geom gf, int nchunk,
                                                                       const T* __restrict__ crse,
                                                                       T* __restrict__ fine) {
    const int64_t w = (int64_t)blockIdx.x * INTERP_ROW_WAVES + (threadIdx.x >> 6);
    const int64_t row = w / nchunk;
    if (row >= (int64_t)gf.N[1] * gf.N[2]) return;
    const int j = (int)(row % gf.N[1]), k = (int)(row / gf.N[1]);
    const int I = (int)(w - row * nchunk) * 64 + (threadIdx.x & 63);
    if (I >= gc.N[0] - 1) return;
    const bool last = I + 2 == gc.N[0];  // also the row's last fine node 2I + 2, the even node of I + 1
    const int J = j >> 1, K = k >> 1;
    const bool ej = (j & 1) == 0, ek = (k & 1) == 0;
    const int64_t mc = I + gc.sy * J + gc.sz * (int64_t)K;
    const int64_t mf = 2 * I + gf.sy * j + gf.sz * (int64_t)k;
    const int64_t sy = gc.sy, sz = gc.sz;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* Cp = crse + p * gc.tot + mc;
        auto at = [&](int a, int b, int c) { return Cp[a + b * sy + c * sz]; };
        T ve, vo, vl;
        if (ej && ek) {
            ve = at(0, 0, 0);
            vo = T(0.5) * (at(0, 0, 0) + at(1, 0, 0));
            vl = at(1, 0, 0);
        } else if (ek) {
            ve = T(0.5) * (at(0, 0, 0) + at(0, 1, 0));
            vo = T(0.25) * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(1, 1, 0));
            vl = T(0.5) * (at(1, 0, 0) + at(1, 1, 0));
        } else if (ej) {
            ve = T(0.5) * (at(0, 0, 0) + at(0, 0, 1));
            vo = T(0.25) * (at(0, 0, 0) + at(0, 0, 1) + at(1, 0, 0) + at(1, 0, 1));
            vl = T(0.5) * (at(1, 0, 0) + at(1, 0, 1));
        } else {
            ve = T(0.25) * (at(0, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1));
            vo = T(0.125) * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1) + at(1, 0, 1) +
                             at(1, 1, 0) + at(1, 1, 1));
            vl = T(0.25) * (at(1, 0, 0) + at(1, 1, 0) + at(1, 0, 1) + at(1, 1, 1));
        }
        T* Fp = fine + p * gf.tot + mf;
        if (SET) {
            Fp[0] = ve;
            Fp[1] = vo;
            if (last) Fp[2] = vl;
        } else {
            Fp[0] += ve;
            Fp[1] += vo;
            if (last) Fp[2] += vl;
        }
    }
}

// launch shape of k_interp_add: one work-item per fine node
inline unsigned interp_blocks(const Geom& gc, const Geom& gf) {
    (void)gc;
    return (unsigned)(((int64_t)gf.N[0] * gf.N[1] * gf.N[2] + 255) / 256);
}

// fine (+)= P crse: the row kernel when the fine level is the coarse one refined by 2, else the
// per-node kernel
template <typename T, bool SET = false>
inline void launch_interp(const Geom& gc, const Geom& gf, const T* crse, T* fine, hipStream_t s) {
    if (gf.n[0] == 2 * gc.n[0] && gf.n[1] == 2 * gc.n[1] && gf.n[2] == 2 * gc.n[2]) {
        const int nchunk = (gc.N[0] - 1 + 63) / 64;
        const int64_t waves = (int64_t)gf.N[1] * gf.N[2] * nchunk;
        k_interp_rows<T, SET><<<(unsigned)((waves + INTERP_ROW_WAVES - 1) / INTERP_ROW_WAVES), 64 * INTERP_ROW_WAVES,
                                0, s>>>(gc, gf, nchunk, crse, fine);
    } else {
        k_interp_add<T, SET><<<interp_blocks(gc, gf), 256, 0, s>>>(gc, gf, crse, fine);
    }
}

// ---------------------------------------------------------------------------
// Vector kernels.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ x, int64_t n, T v) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        x[q] = v;
}
// FP32 value of a bfloat16 bit pattern
__device__ __forceinline__ float bf16_bits_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// bfloat16 bit pattern of 1 / d (round to nearest even), the FP32 sweeps' smoother reciprocal
__global__ void k_recip_bf16(const double* __restrict__ d, uint16_t* __restrict__ out, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = __builtin_bit_cast(uint32_t, (float)(1.0 / d[q]));
        out[q] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
}

template <typename Ti, typename To>
__global__ void k_convert(const Ti* __restrict__ a, To* __restrict__ b, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        b[q] = (To)a[q];
}
// y += alpha * x   (x may be a different precision)
template <typename Tx, typename Ty>
__global__ void k_axpy(Ty* __restrict__ y, const Tx* __restrict__ x, int64_t n, double alpha) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] = (Ty)((double)y[q] + alpha * (double)x[q]);
}
// MajorMinor interior rows, one work-item per interior node (boundary rows: k_boundary)
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_mm_interior(Geom g, const T* __restrict__ C, const T* __restrict__ u,
                                                     const T* __restrict__ b, T* __restrict__ out, T omega) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t I0 = g.N[0] - 2, I1 = g.N[1] - 2;
    if (q >= I0 * I1 * (g.N[2] - 2)) return;
    const int i = 1 + (int)(q % I0), j = 1 + (int)((q / I0) % I1), k = 1 + (int)(q / (I0 * I1));
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
    T f[3], D[3];
    apply_interior_mm<T>(g, C, u, m, f, MODE == JACOBI ? D : nullptr);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int64_t qq = p * g.tot + m;
        if (MODE == APPLY) out[qq] = f[p];
        else if (MODE == RESID) out[qq] = b[qq] - f[p];
        else out[qq] = jacobi_update(g, qq, u[qq], b[qq], f[p], D[p], omega);
    }
}

template <typename T> __device__ __forceinline__ const T* mm_contraction(const Geom& g);
template <> __device__ __forceinline__ const double* mm_contraction<double>(const Geom& g) { return g.mmd; }
template <> __device__ __forceinline__ const float* mm_contraction<float>(const Geom& g) { return g.mmdf; }

// MajorMinor interior rows from the prepared gradient contraction (node-local coefficients only)
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_mm_interior_d(Geom g, const T* __restrict__ C, const T* __restrict__ u,
                                                       const T* __restrict__ b, T* __restrict__ out, T omega) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t I0 = g.N[0] - 2, I1 = g.N[1] - 2;
    if (q >= I0 * I1 * (g.N[2] - 2)) return;
    const int i = 1 + (int)(q % I0), j = 1 + (int)((q / I0) % I1), k = 1 + (int)(q / (I0 * I1));
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
    T f[3], D[3];
    apply_interior_mm_d<T>(g, C, mm_contraction<T>(g), u, m, f, MODE == JACOBI ? D : nullptr);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int64_t qq = p * g.tot + m;
        if (MODE == APPLY) out[qq] = f[p];
        else if (MODE == RESID) out[qq] = b[qq] - f[p];
        else out[qq] = jacobi_update(g, qq, u[qq], b[qq], f[p], D[p], omega);
    }
}

// the 18 planes of D_ikl = sum_j d_j C_ijkl on a MajorMinor level's interior nodes (boundary
// nodes keep what the buffer holds; no row reads them there)
__global__ void __launch_bounds__(256) k_mm_dcoef(Geom g, const double* __restrict__ C, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t I0 = g.N[0] - 2, I1 = g.N[1] - 2;
    if (q >= I0 * I1 * (g.N[2] - 2)) return;
    const int i = 1 + (int)(q % I0), j = 1 + (int)((q / I0) % I1), k = 1 + (int)(q / (I0 * I1));
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
    double d[18];
    mm_grad_contraction<double>(g, C, m, d);
#pragma unroll
    for (int e = 0; e < 18; e++) out[(int64_t)e * g.tot + m] = d[e];
}

// The same rows for a model given per grain: the grain table is staged in LDS once per workgroup
// and each slot's seven reads come from it by the grain ids of the node and its six neighbours
// (2-byte loads) instead of from 21 coefficient planes.  Same arithmetic in the same order as
// k_mm_interior, so both give the same bits.
template <typename T> __device__ __forceinline__ const T* grain_table(const Geom& g);
template <> __device__ __forceinline__ const double* grain_table<double>(const Geom& g) { return g.gtab; }
template <> __device__ __forceinline__ const float* grain_table<float>(const Geom& g) { return g.gtabf; }

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_mm_interior_gid(Geom g, const T* __restrict__ u, const T* __restrict__ b,
                                                         T* __restrict__ out, T omega) {
    extern __shared__ unsigned char smem[];
    T* tab = (T*)smem;
    const T* gt = grain_table<T>(g);
    for (int e = threadIdx.x; e < 21 * g.ngrains; e += blockDim.x) tab[e] = gt[e];
    __syncthreads();
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t I0 = g.N[0] - 2, I1 = g.N[1] - 2;
    if (q >= I0 * I1 * (g.N[2] - 2)) return;
    const int i = 1 + (int)(q % I0), j = 1 + (int)((q / I0) % I1), k = 1 + (int)(q / (I0 * I1));
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
    const int64_t off[7] = {0, 1, -1, g.sy, -g.sy, g.sz, -g.sz};
    CoefGrains<T> cs;
    cs.tab = tab;
#pragma unroll
    for (int d = 0; d < 7; d++) cs.row[d] = 21 * (int)g.gid[m + off[d]];
    T f[3], D[3];
    apply_interior_mm_cs<T>(g, cs, u, m, f, MODE == JACOBI ? D : nullptr);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int64_t qq = p * g.tot + m;
        if (MODE == APPLY) out[qq] = f[p];
        else if (MODE == RESID) out[qq] = b[qq] - f[p];
        else out[qq] = jacobi_update(g, qq, u[qq], b[qq], f[p], D[p], omega);
    }
}

// compact node values (i fastest) -> the level's layout
__global__ void k_scatter_u16(Geom g, const uint16_t* __restrict__ src, uint16_t* __restrict__ dst) {
    const int64_t Vn = (int64_t)g.N[0] * g.N[1] * g.N[2];
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < Vn; v += (int64_t)gridDim.x * blockDim.x)
        dst[node_offset(g, v)] = src[v];
}

// x /= D (Operator<Grid::Node>::normalize)
template <typename T>
__global__ void k_divide(T* __restrict__ x, const T* __restrict__ D, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        if (D[q] != T(0)) x[q] /= D[q];
}
// y = a + b
template <typename T>
__global__ void k_add(T* __restrict__ y, const T* __restrict__ a, const T* __restrict__ b, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] = a[q] + b[q];
}
// compact nc-coefficient records (AoS [v][nc] or SoA [nc][v]) -> SoA planes of the level's layout
template <typename T>
__global__ void k_compact_to_soa(Geom g, const double* __restrict__ src, T* __restrict__ soa, int nc, int aos) {
    const int64_t Vn = (int64_t)g.N[0] * g.N[1] * g.N[2];
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < Vn; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = node_offset(g, v);
        for (int q = 0; q < nc; q++) soa[q * g.tot + m] = (T)(aos ? src[v * nc + q] : src[q * Vn + v]);
    }
}
// AoS (ncoef per node) -> SoA planes
template <typename T>
__global__ void k_aos_to_soa(const double* __restrict__ aos, T* __restrict__ soa, int64_t tot, int nc) {
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < tot; m += (int64_t)gridDim.x * blockDim.x)
        for (int q = 0; q < nc; q++) soa[q * tot + m] = (T)aos[m * nc + q];
}

// ---------------------------------------------------------------------------
// Bottom solve: Jacobi-preconditioned BiCGStab (the amrex::MLCGSolver choice the
// reference forces, Solver/Nonlocal/Linear.H:258, preconditioned by
// Operator<Node>::normalize, Operator/Operator.cpp:156-165), run entirely by one
// workgroup on the coarsest level.  Work vectors live in `w` (8 x 3*tot).
// Global-memory phases are separated by device-scope fences + barriers.
// ---------------------------------------------------------------------------
constexpr int BOTTOM_BLOCK = 1024;

template <typename T, int SYM>
__device__ void bottom_apply(const Geom& g, const BCTable& bc, const T* C, const T* x, T* y) {
    for (int64_t m = threadIdx.x; m < g.tot; m += blockDim.x) {
        int i, j, k;
        node_ijk(g, m, i, j, k);
        if (i >= g.N[0]) continue;
        T f[3], D[3];
        node_row<T, SYM, false>(g, bc, C, x, i, j, k, f, D);
        for (int p = 0; p < 3; p++) y[p * g.tot + m] = f[p];
    }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x / 64); q++) s += sh[q];
    return s;
}
__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x / 64); q++) s = (sh[q] != sh[q]) ? sh[q] : fmax(s, sh[q]);
    return s;
}

__device__ __forceinline__ void phase_sync() {
    __threadfence();
    __syncthreads();
}

template <typename T, int SYM>
__global__ void __launch_bounds__(BOTTOM_BLOCK) k_bottom_bicgstab(Geom g, BCTable bc, const T* __restrict__ C,
                                                                   const T* __restrict__ Dg, T* x,
                                                                   const T* __restrict__ b, T* w, double rtol,
                                                                   int maxit, int* iters) {
    __shared__ double sh[BOTTOM_BLOCK / 64];
    const int64_t n = 3 * g.tot;
    T *r = w, *rh = w + n, *p = w + 2 * n, *v = w + 3 * n, *s = w + 4 * n, *t = w + 5 * n, *ph = w + 6 * n,
      *sh_ = w + 7 * n;
    bottom_apply<T, SYM>(g, bc, C, x, r);
    phase_sync();
    double r0 = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
        r[q] = b[q] - r[q];
        rh[q] = r[q];
        p[q] = T(0);
        v[q] = T(0);
        r0 = fmax(r0, fabs((double)r[q]));
    }
    r0 = block_max(r0, sh);
    phase_sync();
    int it = 0;
    if (r0 > 0.0) {
        double rho_1 = 0.0, alpha = 0.0, omega = 0.0;
        for (it = 1; it <= maxit; it++) {
            double rho = 0.0;
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) rho += (double)rh[q] * (double)r[q];
            rho = block_sum(rho, sh);
            if (rho == 0.0) break;
            const double beta = (it == 1) ? 0.0 : (rho / rho_1) * (alpha / omega);
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
                const T pn = (it == 1) ? r[q] : T(r[q] + beta * (p[q] - omega * v[q]));
                p[q] = pn;
                ph[q] = pn / Dg[q];
            }
            phase_sync();
            bottom_apply<T, SYM>(g, bc, C, ph, v);
            phase_sync();
            double rv = 0.0;
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) rv += (double)rh[q] * (double)v[q];
            rv = block_sum(rv, sh);
            if (rv == 0.0) break;
            alpha = rho / rv;
            double smax = 0.0;
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
                x[q] = T(x[q] + alpha * ph[q]);
                s[q] = T(r[q] - alpha * v[q]);
                sh_[q] = s[q] / Dg[q];
                smax = fmax(smax, fabs((double)s[q]));
            }
            smax = block_max(smax, sh);
            if (smax <= rtol * r0) break;
            phase_sync();
            bottom_apply<T, SYM>(g, bc, C, sh_, t);
            phase_sync();
            double tt = 0.0, ts = 0.0;
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
                tt += (double)t[q] * (double)t[q];
                ts += (double)t[q] * (double)s[q];
            }
            tt = block_sum(tt, sh);
            ts = block_sum(ts, sh);
            if (tt == 0.0) break;
            omega = ts / tt;
            double rmax = 0.0;
            for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
                x[q] = T(x[q] + omega * sh_[q]);
                r[q] = T(s[q] - omega * t[q]);
                rmax = fmax(rmax, fabs((double)r[q]));
            }
            rmax = block_max(rmax, sh);
            phase_sync();
            if (rmax <= rtol * r0 || omega == 0.0) break;
            rho_1 = rho;
        }
    }
    if (threadIdx.x == 0 && iters) *iters = it;
}

// ---------------------------------------------------------------------------
// Dense bottom solve (solver.bottom_solver "dense"): the coarsest operator is assembled column
// by column at prepare (column j = A e_j: block j applies the level's operator to unit vector j,
// column j of E), inverted on the host, and each cycle's bottom solve is x = A^-1 b, one
// wavefront per row of the row-major inverse.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_unit_columns(double* __restrict__ E, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n * n; q += (int64_t)gridDim.x * blockDim.x)
        E[q] = (q / n == q % n) ? 1.0 : 0.0;
}

template <int SYM>
__global__ void __launch_bounds__(256) k_bottom_columns(Geom g, BCTable bc, const double* __restrict__ C,
                                                        const double* __restrict__ E, double* __restrict__ A) {
    const int64_t n = 3 * g.tot;
    bottom_apply<double, SYM>(g, bc, C, E + (int64_t)blockIdx.x * n, A + (int64_t)blockIdx.x * n);
}

// (b and x may be FP32 vectors: the products and the sum run in FP64, x is rounded once, as
// converting b, multiplying in FP64 and converting x would give)
template <typename Tb, typename Tx>
__global__ void __launch_bounds__(256) k_dense_matvec(const double* __restrict__ Ainv, const Tb* __restrict__ b,
                                                      Tx* __restrict__ x, int64_t n) {
    const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (row >= n) return;  // whole wavefronts leave together
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s += Ainv[row * n + j] * (double)b[j];
    s = wave_sum(s);
    if (lane == 0) x[row] = (Tx)s;
}

}  // namespace alamo
