// The last smoothed level L of an FP32 l1 cycle (at most TAIL_MAX nodes, all-Dirichlet isotropic
// rows) in two single-workgroup launches around the dense bottom product: the down part does L's
// pre-smoothing sweeps, its residual and the full-weighting restriction to the coarsest level; the
// up part the trilinear interpolation of the bottom correction and the post-smoothing sweeps.
// Every vector of L lives in LDS between the steps, which keep the arithmetic of the separate
// launches (k_iso_rows64 FP32 l1 sweeps and residuals, k_restrict, k_interp_rows); those take about
// 4.7 us each on such levels whatever their size, 8 of them per visit.  The dense product stays a
// launch over the whole chip (k_dense_matvec): one workgroup reading the FP64 inverse alone takes
// several times longer.
#pragma once

#include "kernels_generic.h"

namespace alamo {

constexpr int TAIL_MAX = 1000;   // nodes of level L
constexpr int TAIL_MAXC = 216;   // nodes of the coarsest level
constexpr int TAIL_NT = 1024;    // work-items

// The interior row f_v (v = 0, 1, 2) of the isotropic operator, read through U(v, dx, dy, dz),
// L(dx, dy, dz) (lambda) and M(dx, dy, dz) (mu): the terms and their order as in k_iso_rows64.
template <typename FU, typename FL, typename FM>
__device__ __forceinline__ void iso_row_f32(const FU& U, const FL& L, const FM& M, float hx, float hy, float hz,
                                            float ix2, float iy2, float iz2, float hxy, float f[3]) {
    float ux[3], uy[3], uz[3], lap[3];
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const float c0 = U(v, 0, 0, 0), e = U(v, 1, 0, 0), wv = U(v, -1, 0, 0);
        const float nn = U(v, 0, 1, 0), ss = U(v, 0, -1, 0), tp = U(v, 0, 0, 1), bt = U(v, 0, 0, -1);
        ux[v] = (e - wv) * hx;
        uy[v] = (nn - ss) * hy;
        uz[v] = (tp - bt) * hz;
        lap[v] = (e - 2.f * c0 + wv) * ix2 + (nn - 2.f * c0 + ss) * iy2 + (tp - 2.f * c0 + bt) * iz2;
    }
    const float uxx0 = (U(0, 1, 0, 0) - 2.f * U(0, 0, 0, 0) + U(0, -1, 0, 0)) * ix2;
    const float uyy1 = (U(1, 0, 1, 0) - 2.f * U(1, 0, 0, 0) + U(1, 0, -1, 0)) * iy2;
    const float uzz2 = (U(2, 0, 0, 1) - 2.f * U(2, 0, 0, 0) + U(2, 0, 0, -1)) * iz2;
    const float uxy0 = (U(0, 1, 1, 0) - U(0, -1, 1, 0) - U(0, 1, -1, 0) + U(0, -1, -1, 0)) * hxy;
    const float uxy1 = (U(1, 1, 1, 0) - U(1, -1, 1, 0) - U(1, 1, -1, 0) + U(1, -1, -1, 0)) * hxy;
    const float uxz0 = ((U(0, 1, 0, 1) - U(0, -1, 0, 1)) * hx - (U(0, 1, 0, -1) - U(0, -1, 0, -1)) * hx) * hz;
    const float uxz2 = ((U(2, 1, 0, 1) - U(2, -1, 0, 1)) * hx - (U(2, 1, 0, -1) - U(2, -1, 0, -1)) * hx) * hz;
    const float uyz1 = ((U(1, 0, 1, 1) - U(1, 0, -1, 1)) * hy - (U(1, 0, 1, -1) - U(1, 0, -1, -1)) * hy) * hz;
    const float uyz2 = ((U(2, 0, 1, 1) - U(2, 0, -1, 1)) * hy - (U(2, 0, 1, -1) - U(2, 0, -1, -1)) * hy) * hz;
    const float gd[3] = {uxx0 + uxy1 + uxz2, uxy0 + uyy1 + uyz2, uxz0 + uyz1 + uzz2};
    const float lam = L(0, 0, 0), mu = M(0, 0, 0);
    const float dl[3] = {(L(1, 0, 0) - L(-1, 0, 0)) * hx, (L(0, 1, 0) - L(0, -1, 0)) * hy,
                         (L(0, 0, 1) - L(0, 0, -1)) * hz};
    const float dm[3] = {(M(1, 0, 0) - M(-1, 0, 0)) * hx, (M(0, 1, 0) - M(0, -1, 0)) * hy,
                         (M(0, 0, 1) - M(0, 0, -1)) * hz};
    const float G[3][3] = {{ux[0], uy[0], uz[0]}, {ux[1], uy[1], uz[1]}, {ux[2], uy[2], uz[2]}};
    const float tr = G[0][0] + G[1][1] + G[2][2];
#pragma unroll
    for (int v = 0; v < 3; v++) {
        float fv = mu * (lap[v] + gd[v]) + lam * gd[v];
        fv += dl[v] * tr + dm[0] * (G[v][0] + G[0][v]) + dm[1] * (G[v][1] + G[1][v]) + dm[2] * (G[v][2] + G[2][v]);
        f[v] = fv;
    }
}

// PART 0 (down): init 0 -- cor holds the start (an interpolated guess, or the first sweep from zero
// written by the restriction into L); 1 -- the first sweep from zero here, cor = b / d (the prepared
// l1 diagonal, FP32); 2 -- cor = 0.  Then nsw sweeps, the residual and its restriction into resc;
// cor written back.  PART 1 (up): cor += P corc, then nsw sweeps; cor written back.
template <int PART>
__global__ void __launch_bounds__(TAIL_NT) k_cycle_tail(Geom g, Geom gc, const float* __restrict__ C,
                                                        const float* __restrict__ dsmf,
                                                        const uint16_t* __restrict__ dsmr,
                                                        const float* __restrict__ res, float* __restrict__ cor,
                                                        float* __restrict__ resc, const float* __restrict__ corc,
                                                        int init, int nsw) {
    __shared__ float su[2][3 * TAIL_MAX];  // ping-pong iterates of L
    __shared__ float sb[3 * TAIL_MAX];     // right-hand side of L
    __shared__ float sr[3 * TAIL_MAX];     // reciprocal l1 diagonal of L (rounded to BF16)
    __shared__ float sc[2 * TAIL_MAX];     // lambda, mu of L
    __shared__ float sxc[3 * TAIL_MAXC];   // bottom correction (up part)
    const int t = threadIdx.x;
    const int n = (int)g.tot, nc = (int)gc.tot;
    const int N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
    for (int q = t; q < 3 * n; q += TAIL_NT) {
        sb[q] = res[q];
        sr[q] = bf16_bits_to_f32(dsmr[q]);
        if (PART == 1 || init == 0) su[0][q] = cor[q];
        else if (init == 1) {
            const float d = dsmf[q];
            su[0][q] = d != 0.f ? res[q] / d : 0.f;
        } else su[0][q] = 0.f;
    }
    for (int q = t; q < 2 * n; q += TAIL_NT) sc[q] = C[q];
    if (PART == 1)
        for (int q = t; q < 3 * nc; q += TAIL_NT) sxc[q] = corc[q];
    __syncthreads();
    const float idx = float(1.0 / g.dx[0]), idy = float(1.0 / g.dx[1]), idz = float(1.0 / g.dx[2]);
    const float hx = 0.5f * idx, hy = 0.5f * idy, hz = 0.5f * idz;
    const float ix2 = idx * idx, iy2 = idy * idy, iz2 = idz * idz, hxy = hx * hy;
    const int sy = N0, sz = N0 * N1;
    int cur = 0;
    // f = A x at node q of iterate x (interior nodes)
    auto row = [&](const float* x, int q, float f[3]) {
        auto U = [&](int v, int dx, int dy, int dz) { return x[v * n + q + dx + dy * sy + dz * sz]; };
        auto L = [&](int dx, int dy, int dz) { return sc[q + dx + dy * sy + dz * sz]; };
        auto M = [&](int dx, int dy, int dz) { return sc[n + q + dx + dy * sy + dz * sz]; };
        iso_row_f32(U, L, M, hx, hy, hz, ix2, iy2, iz2, hxy, f);
    };
    auto boundary = [&](int q) {
        const int i = q % N0, j = (q / N0) % N1, k = q / sz;
        return i == 0 || i == N0 - 1 || j == 0 || j == N1 - 1 || k == 0 || k == N2 - 1;
    };
    auto sweeps = [&](int count) {
        for (int s = 0; s < count; s++) {
            const float* x = su[cur];
            float* y = su[cur ^ 1];
            for (int q = t; q < n; q += TAIL_NT) {
                if (boundary(q)) {
#pragma unroll
                    for (int v = 0; v < 3; v++) y[v * n + q] = sb[v * n + q];
                } else {
                    float f[3];
                    row(x, q, f);
#pragma unroll
                    for (int v = 0; v < 3; v++) y[v * n + q] = x[v * n + q] + (sb[v * n + q] - f[v]) * sr[v * n + q];
                }
            }
            __syncthreads();
            cur ^= 1;
        }
    };
    if (PART == 0) {
        sweeps(nsw);
        // residual of L into the free buffer
        {
            const float* x = su[cur];
            float* r = su[cur ^ 1];
            for (int q = t; q < n; q += TAIL_NT) {
                if (boundary(q)) {
#pragma unroll
                    for (int v = 0; v < 3; v++) r[v * n + q] = sb[v * n + q] - x[v * n + q];
                } else {
                    float f[3];
                    row(x, q, f);
#pragma unroll
                    for (int v = 0; v < 3; v++) r[v * n + q] = sb[v * n + q] - f[v];
                }
            }
        }
        __syncthreads();
        // full-weighting restriction (k_restrict)
        const float* r = su[cur ^ 1];
        const int C0 = gc.N[0], C1 = gc.N[1], c0 = gc.n[0], c1 = gc.n[1], c2 = gc.n[2];
        for (int e = t; e < 3 * nc; e += TAIL_NT) {
            const int p = e / nc, mc = e % nc;
            const int I = mc % C0, J = (mc / C0) % C1, K = mc / (C0 * C1);
            const bool bx = I == 0 || I == c0, by = J == 0 || J == c1, bz = K == 0 || K == c2;
            const float* F = r + p * n + 2 * I + sy * (2 * J) + sz * (2 * K);
            auto at = [&](int a, int b, int c) { return F[a + b * sy + c * sz]; };
            float v;
            if (bx && by && bz) v = at(0, 0, 0);
            else if (by && bz) v = 0.25f * at(-1, 0, 0) + 0.5f * at(0, 0, 0) + 0.25f * at(1, 0, 0);
            else if (bz && bx) v = 0.25f * at(0, -1, 0) + 0.5f * at(0, 0, 0) + 0.25f * at(0, 1, 0);
            else if (bx && by) v = 0.25f * at(0, 0, -1) + 0.5f * at(0, 0, 0) + 0.25f * at(0, 0, 1);
            else if (bx)
                v = (at(0, -1, -1) + 2.f * at(0, 0, -1) + at(0, 1, -1) + 2.f * at(0, -1, 0) + 4.f * at(0, 0, 0) +
                     2.f * at(0, 1, 0) + at(0, -1, 1) + 2.f * at(0, 0, 1) + at(0, 1, 1)) / 16.f;
            else if (by)
                v = (at(-1, 0, -1) + 2.f * at(-1, 0, 0) + at(-1, 0, 1) + 2.f * at(0, 0, -1) + 4.f * at(0, 0, 0) +
                     2.f * at(0, 0, 1) + at(1, 0, -1) + 2.f * at(1, 0, 0) + at(1, 0, 1)) / 16.f;
            else if (bz)
                v = (at(-1, -1, 0) + 2.f * at(0, -1, 0) + at(1, -1, 0) + 2.f * at(-1, 0, 0) + 4.f * at(0, 0, 0) +
                     2.f * at(1, 0, 0) + at(-1, 1, 0) + 2.f * at(0, 1, 0) + at(1, 1, 0)) / 16.f;
            else {
                float c8 = 0.f, c12 = 0.f, c6 = 0.f;
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
                v = c8 / 64.f + c12 / 32.f + c6 / 16.f + at(0, 0, 0) / 8.f;
            }
            resc[e] = v;
        }
    } else {
        // cor += P corc (the sums of k_interp_rows per fine node)
        float* x = su[cur];
        const int csy = gc.N[0], csz = gc.N[0] * gc.N[1];
        for (int e = t; e < 3 * n; e += TAIL_NT) {
            const int p = e / n, q = e % n;
            const int i = q % N0, j = (q / N0) % N1, k = q / sz;
            const int I = i >> 1, J = j >> 1, K = k >> 1;
            const bool oi = i & 1, oj = j & 1, ok = k & 1;
            const float* Cp = sxc + p * nc + I + csy * J + csz * K;
            auto at = [&](int a, int b, int c) { return Cp[a + b * csy + c * csz]; };
            float v;
            if (!oj && !ok) v = oi ? 0.5f * (at(0, 0, 0) + at(1, 0, 0)) : at(0, 0, 0);
            else if (!ok)
                v = oi ? 0.25f * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(1, 1, 0))
                       : 0.5f * (at(0, 0, 0) + at(0, 1, 0));
            else if (!oj)
                v = oi ? 0.25f * (at(0, 0, 0) + at(0, 0, 1) + at(1, 0, 0) + at(1, 0, 1))
                       : 0.5f * (at(0, 0, 0) + at(0, 0, 1));
            else
                v = oi ? 0.125f * (at(0, 0, 0) + at(1, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1) + at(1, 0, 1) +
                                   at(1, 1, 0) + at(1, 1, 1))
                       : 0.25f * (at(0, 0, 0) + at(0, 1, 0) + at(0, 0, 1) + at(0, 1, 1));
            x[e] += v;
        }
        __syncthreads();
        sweeps(nsw);
    }
    for (int q = t; q < 3 * n; q += TAIL_NT) cor[q] = su[cur][q];
}

}  // namespace alamo
