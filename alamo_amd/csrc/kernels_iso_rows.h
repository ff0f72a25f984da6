// Row-wavefront kernel for the interior rows of the isotropic operator on gfx950.
//
// Same interior formula as k_iso_interior (Operator/Elastic.cpp:204-254 with
// Set/Matrix4_Isotropic.H:137-173):
//   f_i = mu (Lap u_i + d_i div u) + lambda d_i div u + d_i(lambda) div u + sum_d d_d(mu) (u_i,d + u_d,i)
// Work split:
//   * one wavefront = one x row segment of 64 output nodes x0 .. x0 + 63 (x0 = 64 bx); x
//     neighbours come from the neighbouring lane through DPP wavefront shifts, the two edge
//     lanes' outer neighbours from LDS;
//   * a workgroup = RW such rows stacked in y; each plane's (u, lambda, mu) records of the rows
//     (plus one halo row each side) are staged in LDS, where the y neighbours read them.  Three
//     plane slots, one barrier per plane;
//   * z is marched serially over a chunk of planes; each lane keeps a plane register queue of u,
//     lambda, mu and the in-plane differences that the x-z and y-z cross derivatives need
//     (d_xz u = D_z(D_x u), exactly the four-corner formula).
// When both x faces are Dirichlet in every component, their face rows (f = u) are written here
// too, so the boundary kernel only visits the y and z faces.
#pragma once
#include <type_traits>

#include "kernels_iso.h"

namespace alamo {

#ifndef ALAMO_ROWS_RW
#define ALAMO_ROWS_RW 4
#endif
#ifndef ALAMO_ROWS_KCH
#define ALAMO_ROWS_KCH 16
#endif
constexpr int ROWS_RW = ALAMO_ROWS_RW;    // default wavefronts (rows) per workgroup
constexpr int ROWS_KCH = ALAMO_ROWS_KCH;  // default z planes per workgroup

#ifndef ALAMO_ROWS_XCD
#define ALAMO_ROWS_XCD 0
#endif
#ifndef ALAMO_ROWS_WPE
#define ALAMO_ROWS_WPE 1
#endif

// rows per workgroup and planes per chunk of the aligned-tile kernel below the large-level
// class, per precision.  Measured at 129^3 nodes in one process (4 x 16 in the hardware order in
// brackets), both in the XCD-contiguous order: FP32 4 x 8 Jacobi 21.6 us (29.1), residual
// 19.7 (28.5); FP64 4 x 16 Jacobi 39.7 (47.1), residual 36.4 (46.0).  Inside the 257^3 F-cycle
// FGMRES solve (whose 129^3 level this is) FP32 4 x 4 took 0.4026 s, 4 x 8 0.4048, 2 x 4 0.4032,
// 4 x 2 0.4086.
#ifndef ALAMO_ROWS64_RW_F32
#define ALAMO_ROWS64_RW_F32 4
#endif
#ifndef ALAMO_ROWS64_KCH_F32
#define ALAMO_ROWS64_KCH_F32 4
#endif
template <typename T> struct Rows64Cfg { static constexpr int RW = ROWS_RW, KCH = ROWS_KCH; };
template <> struct Rows64Cfg<float> { static constexpr int RW = ALAMO_ROWS64_RW_F32, KCH = ALAMO_ROWS64_KCH_F32; };
// Tile shape for levels of >= ROWS64_BIG cells in y and z, where taller or deeper tiles still
// give every CU several workgroups, launched in the XCD-contiguous tile order (neighbouring
// tiles share their halo rows in one XCD's L2).  Measured at 257^3 nodes in one process, the
// 4 x 16 shape in the hardware order in brackets: FP32 8 x 64: Jacobi 211 us (233), residual
// 192 (214), apply 151 (175); FP64 Jacobi and residual 4 x 128: 392 (414) and 365 (388); FP64
// apply 8 x 128: 281 (305).  Every shape and order writes the same values bit for bit.
constexpr int ROWS64_BIG = 192;
template <typename T, int MODE> struct Rows64Big { static constexpr int RW = 4, KCH = 128; };
template <> struct Rows64Big<double, 0> { static constexpr int RW = 8, KCH = 128; };
template <int MODE> struct Rows64Big<float, MODE> { static constexpr int RW = 8, KCH = 64; };
// The FP32 Jacobi sweep with the l1 smoother: when it formed the l1 diagonal itself it needed 141
// VGPRs (3 waves per SIMD), where an 8-row workgroup (8 waves) left a third of the wave slots of a
// CU empty and 4-row workgroups filled all 12 (at 257^3 inside the F-cycle FGMRES solve: 8 x 64
// 0.460 s, 4 x 32 0.447 s, 4 x 16 0.448 s, 2 x 32 0.449 s, 4 x 128 0.454 s, 4 x 64 0.462 s, 2 x 64
// 0.461 s).  Reading the prepared diagonal it needs 128 (4 waves), and the shapes even out: 4 x 32
// 0.3879 s, 8 x 32 0.3879, 8 x 64 0.3883, 4 x 16 0.3875, 4 x 64 0.3884, 2 x 32 0.3991.
template <> struct Rows64Big<float, 2> { static constexpr int RW = 4, KCH = 32; };
// levels of < ROWS64_MID cells in y or z (65^3 nodes and below): 2-row tiles of one output plane.
// These levels are latency-bound, and a workgroup's planes arrive one after another (the plane
// ring loads plane p + 1 while computing p - 1), so one output plane (three loads) beats two
// (four): the 257^3 F-cycle FGMRES solve, which visits them about 110 times per iteration, 0.4155
// -> 0.4068 s.  At 65^3 in FP32 the 2 x 2 tiles took Jacobi 6.3 us, residual 6.0 us (4 x 8: 13.1
// and 12.1; the node-parallel rows took 18 and 15)
constexpr int ROWS64_MID = 96;
struct Rows64Small { static constexpr int RW = 2, KCH = 1; };

constexpr int DPP_WAVE_SHL1 = 0x130;  // lane l receives lane l + 1 (0 past the last lane)
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane l receives lane l - 1 (0 before lane 0)

template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T from_right(T v) { return dpp_shift<DPP_WAVE_SHL1>(v); }
template <typename T>
__device__ __forceinline__ T from_left(T v) { return dpp_shift<DPP_WAVE_SHR1>(v); }

// per-lane record of one plane of the lane's own row
template <typename T>
struct RowPlane {
    T u[3], lam, mu;
    T dxux, dxuz, dyuy, dyuz;  // D_x u_x, D_x u_z, D_y u_y, D_y u_z
};

// levels the row kernel serves: every row, plane, halo and gather index it forms is clamped into
// [0, n], which holds for any n >= 2; the launch routing takes it from 4 cells per direction
inline bool iso_rows_ok(const Geom& g) { return g.n[0] >= 4 && g.n[1] >= 4 && g.n[2] >= 4; }

// 32-bit byte offset from a uniform base: lets the loads and stores use the scalar-base
// addressing form (the host checks that every offset fits)
template <typename T>
__device__ __forceinline__ T ldb(const T* __restrict__ base, uint32_t boff) {
    return *(const T*)((const char*)base + boff);
}
template <typename T>
__device__ __forceinline__ void stb(T* __restrict__ base, uint32_t boff, T v) {
    *(T*)((char*)base + boff) = v;
}

// ---------------------------------------------------------------------------
// x tiles: a wavefront row is the 64 output nodes x0 .. x0 + 63 with x0 = 64 bx.  The x
// neighbours of lanes 0 and 63 (x0 - 1, x0 + 64) sit at entries 0 and 65 of a 66-wide LDS row;
// the other lanes take their x neighbours by DPP.  The edge values of all staged rows (RW + 2
// rows x 2 edges x 5 arrays) are fetched by one gather of the first wavefront per plane, so the
// halo costs one load instruction per workgroup and plane.  With both x faces Dirichlet, a face
// node x = n0 just past the last tile is written by that tile's lane 63 from its halo value.
// ---------------------------------------------------------------------------
constexpr int R64_W = 66;  // LDS row: halo, 64 nodes, halo

// a y- or z-face node q (boundary_ijk order, q < boundary_count_yz) when every region is
// Dirichlet: f = u (diagonal 1)
template <typename T, int MODE>
__device__ __forceinline__ void dirichlet_yz_node(const Geom& g, const T* __restrict__ u, const T* __restrict__ b,
                                                  T* __restrict__ out, T omega, int64_t q) {
    int i, j, k;
    boundary_ijk(g, q, i, j, k);
    const int64_t m = i + g.sy * j + g.sz * (int64_t)k;
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const int64_t qq = v * g.tot + m;
        const T x = u[qq];
        if (MODE == APPLY) out[qq] = x;
        else if (MODE == RESID) out[qq] = b[qq] - x;
        else out[qq] = smoother_diag<T>(g) ? b[qq] : (T(1) - omega) * x + omega * b[qq];
    }
}

// The tile grid is gridDim.x x gridDim.y x gz_tiles; with FACES, the workgroups beyond it in z
// write the nfaces y- and z-face nodes of an all-Dirichlet level instead (one node per
// work-item) and leave before any LDS or barrier, so the faces need no launch of their own.
template <typename T, int MODE, int FOLD, int RW = Rows64Cfg<T>::RW, int KCH = Rows64Cfg<T>::KCH,
          bool XCD = (ALAMO_ROWS_XCD != 0), bool FACES = false>
__global__ void __launch_bounds__(64 * RW) __attribute__((amdgpu_waves_per_eu(ALAMO_ROWS_WPE)))
k_iso_rows64(Geom g, const T* __restrict__ C, const T* __restrict__ u, const T* __restrict__ b, T* __restrict__ out,
             T omega, int gz_tiles, int64_t nfaces) {
    static_assert(RW >= 2, "the first and the last wavefront stage one halo row each");
    constexpr int NR = RW + 2;  // staged rows: halo, RW rows, halo
    static_assert(NR * 10 <= 64 * RW, "the wavefronts' gathers cover the edge values of every staged row");
    if (FACES && (int)blockIdx.z >= gz_tiles) {
        const int64_t q = ((int64_t)(blockIdx.z - gz_tiles) * gridDim.y * gridDim.x +
                           (int64_t)blockIdx.y * gridDim.x + blockIdx.x) * (64 * RW) +
                          threadIdx.y * 64 + threadIdx.x;
        if (q < nfaces) dirichlet_yz_node<T, MODE>(g, u, b, out, omega, q);
        return;
    }
    constexpr int A = NR * R64_W;
    __shared__ T lds[3][5][A];  // [plane slot][u_x, u_y, u_z, lambda, mu][row * 66 + 1 + lane]
    const int lane = threadIdx.x, w = threadIdx.y;
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (XCD) {
        const int gx = gridDim.x, gy = gridDim.y;
        const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gz_tiles);
        bx = t % gx;
        by = (t / gx) % gy;
        bz = t / (gx * gy);
    }
    const int x0 = 64 * bx;  // lane l <-> node x0 + l
    const int y0 = 1 + RW * by;
    const int k0 = 1 + KCH * bz;
    const int k1 = min(k0 + KCH, g.n[2]);  // output planes k0 .. k1 - 1; loads k0 - 1 .. k1
    const int x = x0 + lane;
    const int xl = min(x, g.n[0]);
    const int y = y0 + w;
    const bool row_out = y <= g.n[1] - 1;
    const bool out_i = row_out && x >= 1 && x <= g.n[0] - 1;
    const bool out_f = FOLD >= 1 && row_out && (x == 0 || x == g.n[0]);
    const bool out_fh = FOLD >= 1 && row_out && lane == 63 && x0 + 64 == g.n[0];
    const int ee = lane == 0 ? 0 : R64_W - 1;  // LDS entry of an edge lane's halo node
    constexpr uint32_t E = sizeof(T);
    const uint32_t szb = (uint32_t)g.sz * E;
    const uint32_t colb = ((uint32_t)xl + (uint32_t)g.sy * (uint32_t)min(y, g.n[1])) * E;
    const uint32_t ecolb = ((uint32_t)min(x0 + 64, g.n[0]) + (uint32_t)g.sy * (uint32_t)min(y, g.n[1])) * E;
    const int hrow = w == 0 ? 0 : (w == RW - 1 ? NR - 1 : -1);
    const uint32_t hcolb = ((uint32_t)xl + (uint32_t)g.sy * (uint32_t)(w == 0 ? y0 - 1 : min(y0 + RW, g.n[1]))) * E;
    // edge gather, entry gi = 64 w + lane < 10 NR: staged row gr, edge ge, array ga
    const int gi = 64 * w + lane;
    const bool gl = gi < NR * 10;
    const int gr = gi / 10, ge = (gi / 5) & 1, ga = gi % 5;
    const int gy = gr == 0 ? y0 - 1 : min(y0 + gr - 1, g.n[1]);
    const int gx = ge == 0 ? max(x0 - 1, 0) : min(x0 + 64, g.n[0]);
    const T* __restrict__ gbase = ga < 3 ? u + (int64_t)ga * g.tot : C + (int64_t)(ga - 3) * g.tot;
    const uint32_t gcolb = ((uint32_t)gx + (uint32_t)g.sy * (uint32_t)gy) * E;
    const int gidx = gr * R64_W + (ge == 0 ? 0 : R64_W - 1);
    const T* __restrict__ U[3] = {u, u + g.tot, u + 2 * g.tot};
    const T* __restrict__ la = C;
    const T* __restrict__ ma = C + g.tot;
    const T* __restrict__ B[3] = {b, b + g.tot, b + 2 * g.tot};
    T* __restrict__ O[3] = {out, out + g.tot, out + 2 * g.tot};
    const T idx = T(1.0 / g.dx[0]), idy = T(1.0 / g.dx[1]), idz = T(1.0 / g.dx[2]);
    const bool l1 = MODE == JACOBI && smoother_diag<T>(g) != nullptr;  // l1-Jacobi smoother level
    // FP32 sweeps read the prepared l1 diagonal's BF16 reciprocal (6 more bytes per node, against
    // the VALU work of forming it here); FP64 sweeps form it from the row below
    constexpr bool L1_READ = sizeof(T) == 4;
    const T hx = T(0.5) * idx, hy = T(0.5) * idy, hz = T(0.5) * idz;
    const T ix2 = idx * idx, iy2 = idy * idy, iz2 = idz * idz, hxy = hx * hy;
    const int rs = w * R64_W + 1 + lane, rc = (w + 1) * R64_W + 1 + lane, rn = (w + 2) * R64_W + 1 + lane;
    const int rce = (w + 1) * R64_W + ee;

    // One-plane tiles (the small levels, whose launches are latency-bound) load their three planes
    // and the output plane's right-hand side up front, one memory latency instead of a chain of
    // three; deeper tiles load plane p + 1 while staging plane p.
    constexpr bool ONE = KCH == 1;
    T ph[5] = {}, pg = T(0);  // prefetched halo row and edge gather of the next plane
    T ph1[5] = {}, pg1 = T(0), ph2[5] = {}, pg2 = T(0);  // ONE: the halos of planes k0 and k0 + 1
    auto fetch_to = [&](int p, RowPlane<T>& d, T (&h)[5], T& gq) {
        const uint32_t pz = szb * (uint32_t)p;
        const uint32_t o = colb + pz;
#pragma unroll
        for (int v = 0; v < 3; v++) d.u[v] = ldb(U[v], o);
        d.lam = ldb(la, o);
        d.mu = ldb(ma, o);
        if (gl) gq = ldb(gbase, gcolb + pz);
        if (hrow >= 0) {
            const uint32_t oh = hcolb + pz;
#pragma unroll
            for (int v = 0; v < 3; v++) h[v] = ldb(U[v], oh);
            h[3] = ldb(la, oh);
            h[4] = ldb(ma, oh);
        }
    };
    auto fetch = [&](int p, RowPlane<T>& d) { fetch_to(p, d, ph, pg); };
    // right-hand side (and prepared l1 diagonal) of output plane p - 1
    auto load_b = [&](int p, T (&bq)[3], T (&bh)[3], T (&dq)[3]) {
        const uint32_t pc = szb * (uint32_t)(p - 1);
        const uint32_t mcb = colb + pc;
        if (out_i || out_f) {
#pragma unroll
            for (int v = 0; v < 3; v++) bq[v] = ldb(B[v], mcb);
        }
        if (L1_READ && l1 && out_i) {
            if (g.dsmr) {
                // the BF16 reciprocal: its bit pattern is the high half of an FP32 value
#pragma unroll
                for (int v = 0; v < 3; v++)
                    dq[v] = (T)__builtin_bit_cast(float, (uint32_t)ldb(g.dsmr + (int64_t)v * g.tot, mcb >> 1) << 16);
            } else {
#pragma unroll
                for (int v = 0; v < 3; v++) dq[v] = T(1) / ldb(smoother_diag<T>(g) + (int64_t)v * g.tot, mcb);
            }
        }
        if (out_fh) {
#pragma unroll
            for (int v = 0; v < 3; v++) bh[v] = ldb(B[v], ecolb + pc);
        }
    };
    T pbq[3] = {T(0), T(0), T(0)}, pbh[3] = {T(0), T(0), T(0)}, pdq[3] = {T(1), T(1), T(1)};  // ONE
    // x difference of a row value whose edge-lane neighbour is e (DPP in uniform flow)
    auto dx_of = [&](T v, T e) {
        const T r = from_right(v), l = from_left(v);
        return ((lane == 63 ? e : r) - (lane == 0 ? e : l)) * hx;
    };
    auto step = [&](int p, const RowPlane<T>& qm, const RowPlane<T>& qc, RowPlane<T>& qn, RowPlane<T>& qx) {
        const int sl = (p - k0 + 1) % 3;
        T(*sp)[A] = lds[sl];
#pragma unroll
        for (int v = 0; v < 3; v++) sp[v][rc] = qn.u[v];
        sp[3][rc] = qn.lam;
        sp[4][rc] = qn.mu;
        if (gl) sp[ga][gidx] = pg;
        if (hrow >= 0) {
#pragma unroll
            for (int v = 0; v < 5; v++) sp[v][hrow * R64_W + 1 + lane] = ph[v];
        }
        if (!ONE && p < k1) fetch(p + 1, qx);
        const uint32_t pc = szb * (uint32_t)(p - 1);
        const uint32_t mcb = colb + pc;
        T bq[3] = {T(0), T(0), T(0)}, bh[3] = {T(0), T(0), T(0)}, dq[3] = {T(1), T(1), T(1)};
        if (MODE != APPLY && p >= k0 + 1) {
            if (ONE) {
#pragma unroll
                for (int v = 0; v < 3; v++) {
                    bq[v] = pbq[v];
                    bh[v] = pbh[v];
                    dq[v] = pdq[v];
                }
            } else {
                load_b(p, bq, bh, dq);
            }
        }
        __syncthreads();
        qn.dxux = dx_of(qn.u[0], sp[0][rce]);
        qn.dxuz = dx_of(qn.u[2], sp[2][rce]);
        qn.dyuy = (sp[1][rn] - sp[1][rs]) * hy;
        qn.dyuz = (sp[2][rn] - sp[2][rs]) * hy;
        if (p < k0 + 1) return;
        // output plane c = p - 1
        const T(*sc)[A] = lds[sl == 0 ? 2 : sl - 1];
        T ce[5];  // the centre plane's edge-halo values of this lane's row
#pragma unroll
        for (int v = 0; v < 5; v++) ce[v] = sc[v][rce];
        T ue[3], uw[3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const T r = from_right(qc.u[v]), l = from_left(qc.u[v]);
            ue[v] = lane == 63 ? ce[v] : r;
            uw[v] = lane == 0 ? ce[v] : l;
        }
        const T lx = dx_of(qc.lam, ce[3]);
        const T mx = dx_of(qc.mu, ce[4]);
        if (out_i) {
            T ux[3], uy[3], uz[3], lap[3];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const T c0 = qc.u[v], e = ue[v], wv = uw[v];
                const T nn = sc[v][rn], ss = sc[v][rs];
                ux[v] = (e - wv) * hx;
                uy[v] = (nn - ss) * hy;
                uz[v] = (qn.u[v] - qm.u[v]) * hz;
                lap[v] = (e - T(2) * c0 + wv) * ix2 + (nn - T(2) * c0 + ss) * iy2 + (qn.u[v] - T(2) * c0 + qm.u[v]) * iz2;
            }
            const T uxx0 = (ue[0] - T(2) * qc.u[0] + uw[0]) * ix2;
            const T uyy1 = (sc[1][rn] - T(2) * qc.u[1] + sc[1][rs]) * iy2;
            const T uzz2 = (qn.u[2] - T(2) * qc.u[2] + qm.u[2]) * iz2;
            const T uxy0 = (sc[0][rn + 1] - sc[0][rn - 1] - sc[0][rs + 1] + sc[0][rs - 1]) * hxy;
            const T uxy1 = (sc[1][rn + 1] - sc[1][rn - 1] - sc[1][rs + 1] + sc[1][rs - 1]) * hxy;
            const T uxz0 = (qn.dxux - qm.dxux) * hz, uxz2 = (qn.dxuz - qm.dxuz) * hz;
            const T uyz1 = (qn.dyuy - qm.dyuy) * hz, uyz2 = (qn.dyuz - qm.dyuz) * hz;
            const T gd[3] = {uxx0 + uxy1 + uxz2, uxy0 + uyy1 + uyz2, uxz0 + uyz1 + uzz2};
            const T lam = qc.lam, mu = qc.mu;
            const T dl[3] = {lx, (sc[3][rn] - sc[3][rs]) * hy, (qn.lam - qm.lam) * hz};
            const T dm[3] = {mx, (sc[4][rn] - sc[4][rs]) * hy, (qn.mu - qm.mu) * hz};
            const T G[3][3] = {{ux[0], uy[0], uz[0]}, {ux[1], uy[1], uz[1]}, {ux[2], uy[2], uz[2]}};
            const T tr = G[0][0] + G[1][1] + G[2][2];
            // l1 smoother: d_v = sign(D) (|D| + the row's off-centre absolute sum).  The row's
            // coefficients: faces of u_v along a != v: mu / h_a^2 -+ d_a(mu) / 2h_a; faces of u_v
            // along v: (lambda + 2 mu) / h_v^2 -+ (d_v(lambda) + 2 d_v(mu)) / 2h_v; faces of u_a along
            // a: -+ d_v(lambda) / 2h_a; faces of u_a along v: -+ d_a(mu) / 2h_v; four corners of u_a in
            // the (v, a) plane: (lambda + mu) / 4 h_v h_a.  Using |p + q| + |p - q| = 2 max(|p|, |q|),
            // the parts shared by the three rows are formed once per node.
            const T ih[3] = {idx, idy, idz}, hh[3] = {hx, hy, hz}, w2[3] = {ix2, iy2, iz2};
            T Mf[3], am[3], SM = T(0), SA = T(0), SIH = T(0);
            if (l1 && !L1_READ) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    am[a] = fabs(dm[a]);
                    Mf[a] = T(2) * fmax(mu * w2[a], am[a] * hh[a]);
                    SM += Mf[a];
                    SA += am[a];
                    SIH += ih[a];
                }
            }
#pragma unroll
            for (int v = 0; v < 3; v++) {
                T f = mu * (lap[v] + gd[v]) + lam * gd[v];
                f += dl[v] * tr + dm[0] * (G[v][0] + G[0][v]) + dm[1] * (G[v][1] + G[1][v]) +
                     dm[2] * (G[v][2] + G[2][v]);
                if (MODE == APPLY) stb(O[v], mcb, f);
                else if (MODE == RESID) stb(O[v], mcb, bq[v] - f);
                else {
                    const T D = T(-2) * ((lam + T(2) * mu) * w2[v] + mu * (ix2 + iy2 + iz2 - w2[v]));
                    if (l1 && L1_READ) {
                        stb(O[v], mcb, qc.u[v] + (bq[v] - f) * dq[v]);
                    } else if (l1) {
                        const T io = SIH - ih[v];  // sum of 1/h_a over a != v
                        const T S = T(2) * fmax((lam + T(2) * mu) * w2[v], fabs(dl[v] + T(2) * dm[v]) * hh[v]) +
                                    (SM - Mf[v]) + fabs(dl[v]) * io + ih[v] * (SA - am[v]) +
                                    fabs(lam + mu) * ih[v] * io;
                        const T Dl = D < T(0) ? D - S : D + S;
                        stb(O[v], mcb, qc.u[v] + (bq[v] - f) / Dl);
                    } else {
                        stb(O[v], mcb, (T(1) - omega) * qc.u[v] + omega * (bq[v] - (f - D * qc.u[v])) / D);
                    }
                }
            }
        } else if (out_f) {
            // Dirichlet face row f = u (diagonal 1; the l1 smoother's sweep sets x = b)
#pragma unroll
            for (int v = 0; v < 3; v++) {
                if (MODE == APPLY) stb(O[v], mcb, qc.u[v]);
                else if (MODE == RESID) stb(O[v], mcb, bq[v] - qc.u[v]);
                else stb(O[v], mcb, l1 ? bq[v] : (T(1) - omega) * qc.u[v] + omega * bq[v]);
            }
        }
        if (out_fh) {
            const uint32_t fcb = ecolb + pc;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                if (MODE == APPLY) stb(O[v], fcb, ce[v]);
                else if (MODE == RESID) stb(O[v], fcb, bh[v] - ce[v]);
                else stb(O[v], fcb, l1 ? bh[v] : (T(1) - omega) * ce[v] + omega * bh[v]);
            }
        }
    };
    RowPlane<T> qa, qb, qc, qd;  // plane p lives in record (p - k0 + 1) % 4
    if constexpr (ONE) {
        // k1 = k0 + 1 <= n2: planes k0 - 1, k0, k0 + 1, output plane k0
        fetch_to(k0 - 1, qa, ph, pg);
        fetch_to(k0, qb, ph1, pg1);
        fetch_to(k0 + 1, qc, ph2, pg2);
        if (MODE != APPLY) load_b(k0 + 1, pbq, pbh, pdq);
        step(k0 - 1, qc, qd, qa, qb);
#pragma unroll
        for (int v = 0; v < 5; v++) ph[v] = ph1[v];
        pg = pg1;
        step(k0, qd, qa, qb, qc);
#pragma unroll
        for (int v = 0; v < 5; v++) ph[v] = ph2[v];
        pg = pg2;
        step(k0 + 1, qa, qb, qc, qd);
        return;
    }
    fetch(k0 - 1, qa);
    for (int p = k0 - 1; p <= k1; p += 4) {
        step(p, qc, qd, qa, qb);
        if (p + 1 <= k1) step(p + 1, qd, qa, qb, qc);
        if (p + 2 <= k1) step(p + 2, qa, qb, qc, qd);
        if (p + 3 <= k1) step(p + 3, qb, qc, qd, qa);
    }
}

// true when both x faces (without their edges) are Dirichlet in all three components
inline bool x_faces_dirichlet(const BCTable& bc) {
    for (int c = 0; c < 3; c++)
        if (bc.t[0][c] != DISPLACEMENT || bc.t[3][c] != DISPLACEMENT) return false;
    return true;
}

inline int64_t boundary_count_yz(const Geom& g) {
    const int64_t N0 = g.N[0], N1 = g.N[1], N2 = g.N[2];
    return 2 * N0 * N1 + 2 * N0 * (N2 - 2);
}

inline bool all_dirichlet(const BCTable& bc) {
    for (int f = 0; f < 26; f++)
        for (int c = 0; c < 3; c++)
            if (bc.t[f][c] != DISPLACEMENT) return false;
    return true;
}

// y- and z-face rows when every region is Dirichlet: f = u (diagonal 1), coalesced rows
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_dirichlet_yz(Geom g, const T* __restrict__ u, const T* __restrict__ b,
                                                      T* __restrict__ out, T omega, int64_t nb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    dirichlet_yz_node<T, MODE>(g, u, b, out, omega, q);
}

// row-kernel launch of one tile shape; nfaces > 0 appends the workgroups that write the y- and
// z-face nodes of an all-Dirichlet level
template <typename T, int MODE, int FOLD, int RW, int KCH>
static void launch_rows64(const Geom& g, const T* C, const T* u, const T* b, T* out, T omega, int64_t nfaces,
                          hipStream_t s) {
    const int gx = (g.n[0] + 63) / 64, gy = (g.n[1] - 1 + RW - 1) / RW, gz = (g.n[2] - 1 + KCH - 1) / KCH;
    if (nfaces > 0) {
        const int64_t slice = (int64_t)gx * gy * 64 * RW;
        const int fz = (int)((nfaces + slice - 1) / slice);
        k_iso_rows64<T, MODE, FOLD, RW, KCH, true, true><<<dim3(gx, gy, gz + fz), dim3(64, RW), 0, s>>>(
            g, C, u, b, out, omega, gz, nfaces);
    } else {
        k_iso_rows64<T, MODE, FOLD, RW, KCH, true, false><<<dim3(gx, gy, gz), dim3(64, RW), 0, s>>>(
            g, C, u, b, out, omega, gz, 0);
    }
}

// the row kernel with the tile shape of the level's size class; returns whether the launch also
// wrote the y- and z-face nodes (nfaces > 0)
template <typename T, int MODE, int FOLD>
static bool launch_rows_kernel(const Geom& g, const T* C, const T* u, const T* b, T* out, T omega, hipStream_t s,
                               int64_t nfaces = 0) {
    // tiles cover x = 0 .. 64 ceil(n0 / 64) - 1, every interior node and the x = 0 face
    if (g.n[1] >= ROWS64_BIG && g.n[2] >= ROWS64_BIG)
        launch_rows64<T, MODE, FOLD, Rows64Big<T, MODE>::RW, Rows64Big<T, MODE>::KCH>(g, C, u, b, out, omega, nfaces,
                                                                                   s);
    else if (g.n[1] < ROWS64_MID || g.n[2] < ROWS64_MID)
        launch_rows64<T, MODE, FOLD, Rows64Small::RW, Rows64Small::KCH>(g, C, u, b, out, omega, nfaces, s);
    else launch_rows64<T, MODE, FOLD, Rows64Cfg<T>::RW, Rows64Cfg<T>::KCH>(g, C, u, b, out, omega, nfaces, s);
    return nfaces > 0;
}

template <typename T, int MODE>
static void launch_iso_rows_mode(const Geom& g, const BCTable& bc, const T* C, const T* u, const T* b, T* out, T omega,
                                 hipStream_t s) {
    const bool xdir = x_faces_dirichlet(bc);
    const int64_t nb = xdir ? boundary_count_yz(g) : boundary_count(g);
    if (all_dirichlet(bc)) {
        if (!launch_rows_kernel<T, MODE, 1>(g, C, u, b, out, omega, s, nb))
            k_dirichlet_yz<T, MODE><<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(g, u, b, out, omega, nb);
        return;
    }
    if (xdir) launch_rows_kernel<T, MODE, 1>(g, C, u, b, out, omega, s);
    else launch_rows_kernel<T, MODE, 0>(g, C, u, b, out, omega, s);
    k_boundary<T, ISO, MODE><<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(g, bc, C, u, b, out, omega, nb);
}

template <typename T>
static void launch_iso_rows(int mode, const Geom& g, const BCTable& bc, const T* C, const T* u, const T* b, T* out,
                            T omega, hipStream_t s) {
    if (mode == APPLY) launch_iso_rows_mode<T, APPLY>(g, bc, C, u, b, out, omega, s);
    else if (mode == RESID) launch_iso_rows_mode<T, RESID>(g, bc, C, u, b, out, omega, s);
    else launch_iso_rows_mode<T, JACOBI>(g, bc, C, u, b, out, omega, s);
}

}  // namespace alamo
