// LDS-tiled 2.5-D kernels for the isotropic operator on gfx950.
//
// Interior rows (Operator/Elastic.cpp:204-254 with Set/Matrix4_Isotropic.H:137-173) reduce to
//   f_i = mu (Lap u_i + d_i div u) + lambda d_i div u + d_i(lambda) div u + sum_d d_d(mu) (u_i,d + u_d,i)
// A workgroup owns a 64 x TY column tile of interior nodes and marches a chunk of z planes.
// Each step stages one z plane of (u_x, u_y, u_z, lambda, mu) with a one-node halo in LDS
// (double buffered, one barrier per plane); in-plane derivatives come from LDS, z
// derivatives from a three-plane register queue (x-z and y-z cross derivatives are the
// z-difference of in-plane first differences), so HBM sees each value about once.
// Boundary rows go to a separate kernel over the six faces.
#pragma once
#include "kernels_generic.h"

namespace alamo {

#ifndef ALAMO_ISO_TY
#define ALAMO_ISO_TY 4
#endif
#ifndef ALAMO_ISO_KCH
#define ALAMO_ISO_KCH 32
#endif
#ifndef ALAMO_ISO_XCD
#define ALAMO_ISO_XCD 0
#endif
constexpr int ISO_TX = 64;            // one wavefront per row segment
constexpr int ISO_TY = ALAMO_ISO_TY;  // rows per workgroup
constexpr int ISO_KCH = ALAMO_ISO_KCH;  // z planes marched per workgroup
constexpr int NUM_XCD = 8;            // MI355X accelerator complex dies, one L2 each

// Workgroups are dispatched round-robin over the XCDs; when enabled, give each XCD a
// contiguous range of tiles (x fastest, then y, then the z chunk) so neighbouring tiles
// share halo rows in the same L2.  The map is a bijection of [0, nblk).
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
    const int per = nblk / NUM_XCD, rem = nblk % NUM_XCD;
    const int x = bid % NUM_XCD, r = bid / NUM_XCD;
    return x * per + min(x, rem) + r;
}

template <typename T>
struct IsoPlane {
    T u[3], ux[3], uy[3], uxx[3], uyy[3], uxy[3];
    T lam, mu, lx, ly, mx, my;
};

// The tiled kernel marches z serially inside a workgroup, so it only pays on levels large
// enough to give every CU several workgroups; smaller levels use the node-parallel kernel.
inline bool iso_tiled_ok(const Geom& g) {
    return g.n[0] >= 16 && g.n[1] >= 16 && g.n[2] >= 16 && g.tot >= (int64_t)1 << 12;
}

template <typename T>
__device__ __forceinline__ void iso_stage_plane(T* __restrict__ s, const Geom& g, const T* __restrict__ u,
                                                const T* __restrict__ C, int i0, int j0, int p) {
    constexpr int W = ISO_TX + 2, H = ISO_TY + 2, A = W * H;
    const int tid = threadIdx.x + ISO_TX * threadIdx.y;
    const int64_t pz = (int64_t)p * g.sz;
    for (int q = tid; q < A; q += ISO_TX * ISO_TY) {
        const int a = q % W, b = q / W;
        const int gi = min(i0 - 1 + a, g.n[0]);
        const int gj = min(j0 - 1 + b, g.n[1]);
        const int64_t m = gi + g.sy * gj + pz;
        s[q] = u[m];
        s[A + q] = u[g.tot + m];
        s[2 * A + q] = u[2 * g.tot + m];
        s[3 * A + q] = C[m];
        s[4 * A + q] = C[g.tot + m];
    }
}

template <typename T>
__device__ __forceinline__ void iso_plane_rec(const T* __restrict__ s, int tx, int ty, const T idx, const T idy,
                                              IsoPlane<T>& r) {
    constexpr int W = ISO_TX + 2, H = ISO_TY + 2, A = W * H;
    const int c0 = (ty + 1) * W + (tx + 1);
    const T hx = T(0.5) * idx, hy = T(0.5) * idy;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const T* sp = s + p * A;
        const T v = sp[c0], e = sp[c0 + 1], w = sp[c0 - 1], n = sp[c0 + W], so = sp[c0 - W];
        const T ne = sp[c0 + W + 1], nw = sp[c0 + W - 1], se = sp[c0 - W + 1], sw = sp[c0 - W - 1];
        r.u[p] = v;
        r.ux[p] = (e - w) * hx;
        r.uy[p] = (n - so) * hy;
        r.uxx[p] = (e - T(2) * v + w) * idx * idx;
        r.uyy[p] = (n - T(2) * v + so) * idy * idy;
        r.uxy[p] = T(0.25) * (ne + sw - se - nw) * idx * idy;
    }
    const T* sl = s + 3 * A;
    const T* sm = s + 4 * A;
    r.lam = sl[c0];
    r.mu = sm[c0];
    r.lx = (sl[c0 + 1] - sl[c0 - 1]) * hx;
    r.ly = (sl[c0 + W] - sl[c0 - W]) * hy;
    r.mx = (sm[c0 + 1] - sm[c0 - 1]) * hx;
    r.my = (sm[c0 + W] - sm[c0 - W]) * hy;
}

template <typename T, int MODE>
__global__ void __launch_bounds__(ISO_TX* ISO_TY) k_iso_interior(Geom g, const T* __restrict__ C,
                                                                  const T* __restrict__ u, const T* __restrict__ b,
                                                                  T* __restrict__ out, T omega) {
    constexpr int A = (ISO_TX + 2) * (ISO_TY + 2);
    __shared__ T smem[2][5 * A];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (ALAMO_ISO_XCD) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int bid = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const int t = xcd_tile(bid, gx * gy * gridDim.z);
        bx = t % gx;
        by = (t / gx) % gy;
        bz = t / (gx * gy);
    }
    const int i0 = 1 + bx * ISO_TX, j0 = 1 + by * ISO_TY;
    const int k0 = 1 + bz * ISO_KCH;
    const int k1 = min(k0 + ISO_KCH, g.n[2]);  // exclusive; interior k <= n2 - 1
    const int i = i0 + tx, j = j0 + ty;
    const bool act = i < g.n[0] && j < g.n[1];
    const T idx = T(1.0 / g.dx[0]), idy = T(1.0 / g.dx[1]), idz = T(1.0 / g.dx[2]);
    const T hz = T(0.5) * idz;

    IsoPlane<T> prv, cur, nxt;
    iso_stage_plane<T>(smem[0], g, u, C, i0, j0, k0 - 1);
    iso_stage_plane<T>(smem[1], g, u, C, i0, j0, k0);
    __syncthreads();
    iso_plane_rec<T>(smem[0], tx, ty, idx, idy, prv);
    iso_plane_rec<T>(smem[1], tx, ty, idx, idy, cur);
    __syncthreads();
    int buf = 0;
    for (int k = k0; k < k1; k++) {
        // Plane k+1 goes into the buffer read two steps ago; every work-item has passed
        // the barrier of the previous step since, so one barrier per plane suffices.
        iso_stage_plane<T>(smem[buf], g, u, C, i0, j0, k + 1);
        __syncthreads();
        iso_plane_rec<T>(smem[buf], tx, ty, idx, idy, nxt);
        buf ^= 1;
        if (act) {
            T uz[3], uzz[3], uxz[3], uyz[3];
#pragma unroll
            for (int p = 0; p < 3; p++) {
                uz[p] = (nxt.u[p] - prv.u[p]) * hz;
                uzz[p] = (nxt.u[p] - T(2) * cur.u[p] + prv.u[p]) * idz * idz;
                uxz[p] = (nxt.ux[p] - prv.ux[p]) * hz;
                uyz[p] = (nxt.uy[p] - prv.uy[p]) * hz;
            }
            const T lz = (nxt.lam - prv.lam) * hz, mz = (nxt.mu - prv.mu) * hz;
            const T lam = cur.lam, mu = cur.mu;
            // grad u: G[p][d]
            const T G[3][3] = {{cur.ux[0], cur.uy[0], uz[0]}, {cur.ux[1], cur.uy[1], uz[1]},
                               {cur.ux[2], cur.uy[2], uz[2]}};
            const T tr = G[0][0] + G[1][1] + G[2][2];
            const T lap[3] = {cur.uxx[0] + cur.uyy[0] + uzz[0], cur.uxx[1] + cur.uyy[1] + uzz[1],
                              cur.uxx[2] + cur.uyy[2] + uzz[2]};
            const T gd[3] = {cur.uxx[0] + cur.uxy[1] + uxz[2], cur.uxy[0] + cur.uyy[1] + uyz[2],
                             uxz[0] + uyz[1] + uzz[2]};
            const T dl[3] = {cur.lx, cur.ly, lz}, dm[3] = {cur.mx, cur.my, mz};
            const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                T f = mu * (lap[p] + gd[p]) + lam * gd[p];
                f += dl[p] * tr + dm[0] * (G[p][0] + G[0][p]) + dm[1] * (G[p][1] + G[1][p]) +
                     dm[2] * (G[p][2] + G[2][p]);
                const int64_t q = p * g.tot + m;
                if (MODE == APPLY) out[q] = f;
                else if (MODE == RESID) out[q] = b[q] - f;
                else {
                    // D_p = -2 [ (lambda + 2 mu)/dx_p^2 + mu sum_{j != p} 1/dx_j^2 ]
                    const T w2[3] = {idx * idx, idy * idy, idz * idz};
                    const T D = T(-2) * ((lam + T(2) * mu) * w2[p] + mu * (w2[0] + w2[1] + w2[2] - w2[p]));
                    const T x = cur.u[p];
                    out[q] = jacobi_update(g, q, x, b[q], f, D, omega);
                }
            }
        }
        prv = cur;
        cur = nxt;
    }
}

// Boundary rows of any level: one work-item per node of the six faces.
__device__ __forceinline__ void boundary_ijk(const Geom& g, int64_t q, int& i, int& j, int& k) {
    const int64_t N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
    const int64_t nz = N0 * N1;            // one z face
    const int64_t ny = N0 * (N2 - 2);      // one y face without z edges
    if (q < 2 * nz) {
        k = q < nz ? 0 : g.n[2];
        const int64_t r = q % nz;
        j = (int)(r / N0);
        i = (int)(r % N0);
        return;
    }
    q -= 2 * nz;
    if (q < 2 * ny) {
        j = q < ny ? 0 : g.n[1];
        const int64_t r = q % ny;
        k = 1 + (int)(r / N0);
        i = (int)(r % N0);
        return;
    }
    q -= 2 * ny;
    const int64_t nx = (N1 - 2) * (N2 - 2);
    i = q < nx ? 0 : g.n[0];
    const int64_t r = q % nx;
    k = 1 + (int)(r / (N1 - 2));
    j = 1 + (int)(r % (N1 - 2));
}

inline int64_t boundary_count(const Geom& g) {
    const int64_t N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
    return 2 * N0 * N1 + 2 * N0 * (N2 - 2) + 2 * (N1 - 2) * (N2 - 2);
}

template <typename T, int SYM, int MODE>
__global__ void __launch_bounds__(256) k_boundary(Geom g, BCTable bc, const T* __restrict__ C,
                                                  const T* __restrict__ u, const T* __restrict__ b,
                                                  T* __restrict__ out, T omega, int64_t nb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    int i, j, k;
    boundary_ijk(g, q, i, j, k);
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
    T f[3], D[3];
    apply_boundary<T, SYM>(g, bc, C, u, i, j, k, f);
    if (MODE == JACOBI) diag_boundary<T, SYM>(g, bc, C, i, j, k, D);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int64_t qq = p * g.tot + m;
        if (MODE == APPLY) out[qq] = f[p];
        else if (MODE == RESID) out[qq] = b[qq] - f[p];
        else {
            const T x = u[qq];
            out[qq] = jacobi_update(g, qq, x, b[qq], f[p], D[p], omega);
        }
    }
}

template <typename T>
static void launch_iso_tiled(int mode, const Geom& g, const BCTable& bc, const T* C, const T* u, const T* b, T* out,
                             T omega, hipStream_t s) {
    dim3 blk(ISO_TX, ISO_TY);
    dim3 grid((g.n[0] - 1 + ISO_TX - 1) / ISO_TX, (g.n[1] - 1 + ISO_TY - 1) / ISO_TY,
              (g.n[2] - 1 + ISO_KCH - 1) / ISO_KCH);
    const int64_t nb = boundary_count(g);
    const unsigned gb = (unsigned)((nb + 255) / 256);
    if (mode == APPLY) {
        k_iso_interior<T, APPLY><<<grid, blk, 0, s>>>(g, C, u, b, out, omega);
        k_boundary<T, ISO, APPLY><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
    } else if (mode == RESID) {
        k_iso_interior<T, RESID><<<grid, blk, 0, s>>>(g, C, u, b, out, omega);
        k_boundary<T, ISO, RESID><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
    } else {
        k_iso_interior<T, JACOBI><<<grid, blk, 0, s>>>(g, C, u, b, out, omega);
        k_boundary<T, ISO, JACOBI><<<gb, 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
    }
}

}  // namespace alamo
