// Learned correction for the multigrid preconditioner: a 3-level convolutional network
// (16/32/64 channels, 3x3x3 kernels, BF16 weights and activations, FP32 accumulation).
//
// No reference counterpart.  Layout: activations are channel-last BF16, act[v * C + c], on the
// nodes of one level (v = i + Nx (j + Ny k)); channel counts are padded to multiples of 16.
// Every convolution is an implicit GEMM on the gfx950 matrix cores
// (v_mfma_f32_16x16x16_bf16): M = 16 voxels along x, N = 16 output channels,
// K = 16 input channels of one of the 27 offsets.  Operand lane layout (checked by
// tests/test_gpu_learned.py::test_mfma_lane_layout):
//   A (16x16, M x K):  lane l holds A[l % 16][4 (l / 16) + 0..3]
//   B (16x16, K x N):  lane l holds B[4 (l / 16) + 0..3][l % 16]
//   D (16x16, M x N):  lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels_generic.h"

namespace alamo {
namespace net {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct Grid3 {
    int N[3];       // nodes per direction
    int64_t V;      // N0 N1 N2
};

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                           // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// offset index o = (dz+1)*9 + (dy+1)*3 + (dx+1)
__device__ __forceinline__ void off3(int o, int& dx, int& dy, int& dz) {
    dz = o / 9 - 1;
    dy = (o / 3) % 3 - 1;
    dx = o % 3 - 1;
}

// ---------------------------------------------------------------------------
// Forward convolution: out[v][co] = bias[co] + sum_{o, ci} W[o][co][ci] in[v + o][ci]
// (zero outside the domain), optional ReLU.  One wavefront owns a row of MT x 16
// consecutive voxels of an x row and all NT x 16 output channels: per (offset, 16-channel
// input chunk) it loads the NT weight fragments once and applies them to its MT voxel
// tiles (MT x NT accumulator tiles).  A whole source row outside the domain (y or z
// padding) is skipped in wave-uniform control flow.
// ---------------------------------------------------------------------------
constexpr int CONV_WAVES = 4;  // wavefronts per workgroup

// xb, tx: the wavefronts cover x in [xb, xb + tx * 16 MT) of every row (the launcher runs the
// bulk of a 4k+1-node row with MT = 4 and its last column with MT = 1)
template <int NT, int MT, bool OUT_F32>
__global__ void __launch_bounds__(64 * CONV_WAVES) k_conv3_fwd(Grid3 g, int Cin, int xb, int tx,
                                                               const uint16_t* __restrict__ in,
                                                               const uint16_t* __restrict__ W,
                                                               const float* __restrict__ bias, void* __restrict__ out,
                                                               int relu) {
    constexpr int Cout = 16 * NT;
    constexpr int ROW = 16 * MT;
    const int lane = threadIdx.x % 64;
    const int64_t tile = (int64_t)blockIdx.x * CONV_WAVES + threadIdx.x / 64;
    const int64_t ntile = (int64_t)tx * g.N[1] * g.N[2];
    if (tile >= ntile) return;
    const int x0 = xb + (int)(tile % tx) * ROW;
    const int y = (int)((tile / tx) % g.N[1]);
    const int z = (int)(tile / ((int64_t)tx * g.N[1]));
    const int m = lane % 16, kq = 4 * (lane / 16);
    f4 acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[a][t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int o = 0; o < 27; o++) {
        int dx, dy, dz;
        off3(o, dx, dy, dz);
        const int ys = y + dy, zs = z + dz;
        if (ys < 0 || ys >= g.N[1] || zs < 0 || zs >= g.N[2]) continue;  // uniform: zero padding row
        const int64_t rowv = (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs);
        for (int ci0 = 0; ci0 < Cin; ci0 += 16) {
            s4 bw[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) bw[t] = *(const s4*)(W + ((int64_t)o * Cout + t * 16 + m) * Cin + ci0 + kq);
#pragma unroll
            for (int a = 0; a < MT; a++) {
                const int xs = x0 + 16 * a + m + dx;
                s4 av = s4{0, 0, 0, 0};
                if (xs >= 0 && xs < g.N[0]) av = *(const s4*)(in + (rowv + xs) * Cin + ci0 + kq);
#pragma unroll
                for (int t = 0; t < NT; t++) acc[a][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bw[t], acc[a][t], 0, 0, 0);
            }
        }
    }
    const int64_t rowo = (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = t * 16 + m;
        const float bb = bias ? bias[co] : 0.f;
#pragma unroll
        for (int a = 0; a < MT; a++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int xv = x0 + 16 * a + kq + i;
                if (xv >= g.N[0]) continue;
                const int64_t v = rowo + xv;
                float r = acc[a][t][i] + bb;
                if (relu) r = fmaxf(r, 0.f);
                if (OUT_F32) ((float*)out)[v * Cout + co] = r;
                else ((uint16_t*)out)[v * Cout + co] = f2bf(r);
            }
    }
}

// ---------------------------------------------------------------------------
// Forward convolution for 16 or 32 input channels with the input staged in LDS: a workgroup of
// CL_TY wavefronts owns a 64 x CL_TY voxel tile of an xy plane and marches CL_KZ planes in z;
// a ring of three input planes (66 x (CL_TY + 2) voxels with their halo, zero outside the
// domain) lives in LDS, and each wavefront reads its 27 shifted A fragments from there instead
// of from L1/L2.  Same arithmetic and output as k_conv3_fwd.
// ---------------------------------------------------------------------------
constexpr int CL_TY = 4;
constexpr int CL_KZ = 8;
constexpr int CL_H = CL_TY + 2;

template <int CIN, int NT, int MT, bool OUT_F32>
__global__ void __launch_bounds__(64 * CL_TY) k_conv3_fwd_lds(Grid3 g, int xb, const uint16_t* __restrict__ in,
                                                            const uint16_t* __restrict__ W,
                                                            const float* __restrict__ bias, void* __restrict__ out,
                                                            int relu) {
    constexpr int Cout = 16 * NT;
    constexpr int CW = 16 * MT + 2;                          // staged x extent
    constexpr int CPV = CIN / 8;                             // 16-byte chunks per voxel
    constexpr int NCH = CL_H * CW * CPV;                     // 16-byte chunks per staged plane
    constexpr int PER = (NCH + 64 * CL_TY - 1) / (64 * CL_TY);  // chunks per work-item
    __shared__ __align__(16) uint16_t tile[3][CL_H][CW][CIN];
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int x0 = xb + blockIdx.x * 16 * MT, y0 = blockIdx.y * CL_TY, z0 = blockIdx.z * CL_KZ;
    const int z1 = min(z0 + CL_KZ, g.N[2]);
    const int m = lane % 16, kq = 4 * (lane / 16);
    auto slot = [](int p) { return (p + 3) % 3; };
    uint4 pre[PER];
    auto fetch = [&](int zs) {
#pragma unroll
        for (int c = 0; c < PER; c++) {
            const int q = threadIdx.x + 64 * CL_TY * c;
            const int part = q % CPV, vox = q / CPV;
            const int xx = vox % CW, yy = vox / CW;
            const int xs = x0 - 1 + xx, ys = y0 - 1 + yy;
            pre[c] = make_uint4(0, 0, 0, 0);
            if (q < NCH && zs >= 0 && zs < g.N[2] && xs >= 0 && xs < g.N[0] && ys >= 0 && ys < g.N[1])
                pre[c] = *(const uint4*)(in + (xs + (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs)) * CIN + 8 * part);
        }
    };
    auto put = [&](int zs) {
        const int sl = slot(zs);
#pragma unroll
        for (int c = 0; c < PER; c++) {
            const int q = threadIdx.x + 64 * CL_TY * c;
            if (q < NCH) {
                const int part = q % CPV, vox = q / CPV;
                *(uint4*)&tile[sl][vox / CW][vox % CW][8 * part] = pre[c];
            }
        }
    };
    fetch(z0 - 1);
    put(z0 - 1);
    fetch(z0);
    put(z0);
    fetch(z0 + 1);
    const int y = y0 + w;
    for (int z = z0; z < z1; z++) {
        // plane z + 1 replaces plane z - 2, which the previous plane's outputs were the last to read
        __syncthreads();
        put(z + 1);
        __syncthreads();
        if (z + 2 <= z1) fetch(z + 2);  // in flight while this plane computes
        if (y >= g.N[1]) continue;      // uniform per wavefront; the barriers above are reached by all
        f4 acc[MT][NT];
#pragma unroll
        for (int a = 0; a < MT; a++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[a][t] = f4{0.f, 0.f, 0.f, 0.f};
        for (int o = 0; o < 27; o++) {
            int dx, dy, dz;
            off3(o, dx, dy, dz);
            const uint16_t(*pl)[CW][CIN] = tile[slot(z + dz)];
#pragma unroll
            for (int ci0 = 0; ci0 < CIN; ci0 += 16) {
                s4 bw[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) bw[t] = *(const s4*)(W + ((int64_t)o * Cout + t * 16 + m) * CIN + ci0 + kq);
#pragma unroll
                for (int a = 0; a < MT; a++) {
                    const s4 av = *(const s4*)&pl[w + 1 + dy][16 * a + m + dx + 1][ci0 + kq];
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[a][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bw[t], acc[a][t], 0, 0, 0);
                }
            }
        }
        const int64_t rowo = (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = t * 16 + m;
            const float bb = bias ? bias[co] : 0.f;
#pragma unroll
            for (int a = 0; a < MT; a++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int xv = x0 + 16 * a + kq + i;
                    if (xv >= g.N[0]) continue;
                    const int64_t v = rowo + xv;
                    float r = acc[a][t][i] + bb;
                    if (relu) r = fmaxf(r, 0.f);
                    if (OUT_F32) ((float*)out)[v * Cout + co] = r;
                    else ((uint16_t*)out)[v * Cout + co] = f2bf(r);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[o][co][ci] += sum_v dout[v][co] in[v + o][ci].  GEMM with
// M = co, N = ci, K = voxels; one wavefront owns one (o, co tile, ci tile) and a chunk
// of voxels, and adds its 16x16 result into dW with float atomics.
// ---------------------------------------------------------------------------
constexpr int WGRAD_CHUNK = 2048;  // voxels per wavefront

__global__ void __launch_bounds__(256) k_conv3_wgrad(Grid3 g, int Cin, int Cout, const uint16_t* __restrict__ in,
                                                     const uint16_t* __restrict__ dout, float* __restrict__ dW) {
    const int lane = threadIdx.x % 64;
    const int64_t wid = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int NCO = Cout / 16, NCI = Cin / 16;
    const int64_t nchunk = (g.V + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
    const int64_t njob = (int64_t)27 * NCO * NCI * nchunk;
    if (wid >= njob) return;
    const int64_t chunk = wid % nchunk;
    int64_t r = wid / nchunk;
    const int cit = (int)(r % NCI);
    r /= NCI;
    const int cot = (int)(r % NCO);
    const int o = (int)(r / NCO);
    int dx, dy, dz;
    off3(o, dx, dy, dz);
    const int m = lane % 16, kq = 4 * (lane / 16);
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    const int64_t v0 = chunk * WGRAD_CHUNK;
    const int64_t v1 = v0 + WGRAD_CHUNK < g.V ? v0 + WGRAD_CHUNK : g.V;
    for (int64_t vb = v0; vb < v1; vb += 16) {
        s4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int64_t v = vb + kq + q;
            short av = 0, bv = 0;
            if (v < v1) {
                av = (short)dout[v * Cout + cot * 16 + m];
                const int x = (int)(v % g.N[0]);
                const int y = (int)((v / g.N[0]) % g.N[1]);
                const int z = (int)(v / ((int64_t)g.N[0] * g.N[1]));
                const int xs = x + dx, ys = y + dy, zs = z + dz;
                if (xs >= 0 && xs < g.N[0] && ys >= 0 && ys < g.N[1] && zs >= 0 && zs < g.N[2])
                    bv = (short)in[(xs + (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs)) * Cin + cit * 16 + m];
            }
            a[q] = av;
            b[q] = bv;
        }
        // A[co = m][k = voxel kq+q], B[k = voxel kq+q][ci = m]
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // D[co = kq + i][ci = m]
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int co = cot * 16 + kq + i, ci = cit * 16 + m;
        atomicAdd(&dW[((int64_t)o * Cout + co) * Cin + ci], acc[i]);
    }
}

// dB[co] += sum_v dout[v][co]; Cout divides 256: thread t reads channel t % Cout of voxel
// t / Cout (coalesced rows), then the block folds its 256 / Cout partial sums per channel
__global__ void __launch_bounds__(256) k_bias_grad(int64_t V, int Cout, const uint16_t* __restrict__ dout,
                                                   float* __restrict__ dB) {
    __shared__ float sh[256];
    const int t = threadIdx.x, per = 256 / Cout;
    const int c = t % Cout;
    float s = 0.f;
    for (int64_t v = (int64_t)blockIdx.x * per + t / Cout; v < V; v += (int64_t)gridDim.x * per)
        s += bf2f(dout[v * Cout + c]);
    sh[t] = s;
    __syncthreads();
    if (t < Cout) {
        float a = 0.f;
        for (int q = 0; q < per; q++) a += sh[t + q * Cout];
        atomicAdd(&dB[t], a);
    }
}

// Weights of the backward-data convolution: Wb[26 - o][ci][co] = W[o][co][ci]
__global__ void k_wflip(int Cin, int Cout, const float* __restrict__ W, uint16_t* __restrict__ Wb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)27 * Cin * Cout) return;
    const int ci = (int)(q % Cin);
    const int co = (int)((q / Cin) % Cout);
    const int o = (int)(q / ((int64_t)Cin * Cout));
    Wb[((int64_t)(26 - o) * Cin + ci) * Cout + co] = f2bf(W[q]);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        b[q] = f2bf(a[q]);
}

// g (f32) -> bf16, masked by the ReLU of the forward activation a (bf16): g * (a > 0)
__global__ void k_relu_bwd(const float* __restrict__ g, const uint16_t* __restrict__ a, uint16_t* __restrict__ out,
                           int64_t n, int use_mask) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        float v = g[q];
        if (use_mask && !(bf2f(a[q]) > 0.f)) v = 0.f;
        out[q] = f2bf(v);
    }
}

// down: d[I] = a[2I] (node injection); C channels
__global__ void k_inject(Grid3 gf, Grid3 gc, int C, const uint16_t* __restrict__ a, uint16_t* __restrict__ d) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= gc.V * C) return;
    const int c = (int)(q % C);
    const int64_t I = q / C;
    const int X = (int)(I % gc.N[0]), Y = (int)((I / gc.N[0]) % gc.N[1]), Z = (int)(I / ((int64_t)gc.N[0] * gc.N[1]));
    d[q] = a[((2 * X) + (int64_t)gf.N[0] * ((2 * Y) + (int64_t)gf.N[1] * (2 * Z))) * C + c];
}
// its adjoint: ga[2I] += gd[I]  (f32 gradients)
__global__ void k_inject_bwd(Grid3 gf, Grid3 gc, int C, const float* __restrict__ gd, float* __restrict__ ga) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= gc.V * C) return;
    const int c = (int)(q % C);
    const int64_t I = q / C;
    const int X = (int)(I % gc.N[0]), Y = (int)((I / gc.N[0]) % gc.N[1]), Z = (int)(I / ((int64_t)gc.N[0] * gc.N[1]));
    ga[((2 * X) + (int64_t)gf.N[0] * ((2 * Y) + (int64_t)gf.N[1] * (2 * Z))) * C + c] += gd[q];
}

// up: u[i] = (P c)[i] + skip[i], trilinear nodal prolongation (the weights of the multigrid
// interpolation); c is f32 or bf16 channel-last on the coarse level, skip bf16; u bf16.
__device__ __forceinline__ float prol_w(int i, int I) {
    // 1-D weight of coarse node I for fine node i: 1 if i == 2I, 1/2 if |i - 2I| == 1
    const int d = i - 2 * I;
    return d == 0 ? 1.f : (d == 1 || d == -1) ? 0.5f : 0.f;
}
template <typename TC>
__global__ void k_prolong_add(Grid3 gf, Grid3 gc, int C, const TC* __restrict__ c, const uint16_t* __restrict__ skip,
                              uint16_t* __restrict__ u) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= gf.V * C) return;
    const int ch = (int)(q % C);
    const int64_t v = q / C;
    const int x = (int)(v % gf.N[0]), y = (int)((v / gf.N[0]) % gf.N[1]), z = (int)(v / ((int64_t)gf.N[0] * gf.N[1]));
    float s = 0.f;
    const int X0 = x / 2, Y0 = y / 2, Z0 = z / 2;
    for (int dz = 0; dz <= (z & 1); dz++)
        for (int dy = 0; dy <= (y & 1); dy++)
            for (int dxx = 0; dxx <= (x & 1); dxx++) {
                const int X = X0 + dxx, Y = Y0 + dy, Z = Z0 + dz;
                const float w = prol_w(x, X) * prol_w(y, Y) * prol_w(z, Z);
                const int64_t I = X + (int64_t)gc.N[0] * (Y + (int64_t)gc.N[1] * Z);
                float cv;
                if constexpr (sizeof(TC) == 4) cv = (float)c[I * C + ch];
                else cv = bf2f((uint16_t)c[I * C + ch]);
                s += w * cv;
            }
    u[q] = f2bf(s + bf2f(skip[q]));
}
// adjoint of the prolongation: gc[I] = sum_i w(i, I) gu[i] over fine nodes i within one of 2I
__global__ void k_prolong_T(Grid3 gf, Grid3 gc, int C, const float* __restrict__ gu, float* __restrict__ gcout) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= gc.V * C) return;
    const int ch = (int)(q % C);
    const int64_t I = q / C;
    const int X = (int)(I % gc.N[0]), Y = (int)((I / gc.N[0]) % gc.N[1]), Z = (int)(I / ((int64_t)gc.N[0] * gc.N[1]));
    float s = 0.f;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dxx = -1; dxx <= 1; dxx++) {
                const int x = 2 * X + dxx, y = 2 * Y + dy, z = 2 * Z + dz;
                if (x < 0 || y < 0 || z < 0 || x >= gf.N[0] || y >= gf.N[1] || z >= gf.N[2]) continue;
                const float w = (dxx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
                s += w * gu[(x + (int64_t)gf.N[0] * (y + (int64_t)gf.N[1] * z)) * C + ch];
            }
    gcout[q] = s;
}

// Voxel v of the network (V = N0 N1 N2 voxels) is node node_offset(g, v) of the operator level g.
// Network input: channels 0-2 = r / s (r FP64, 3 SoA planes), 3-4 = coefficient features,
// 5-15 = 0.  feat holds 2 f32 per voxel.
__global__ void k_pack_input(Geom g, int64_t V, const double* __restrict__ r, double inv_s,
                             const float* __restrict__ feat, uint16_t* __restrict__ x) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
    uint16_t* xv = x + v * 16;
    xv[0] = f2bf((float)(r[m] * inv_s));
    xv[1] = f2bf((float)(r[g.tot + m] * inv_s));
    xv[2] = f2bf((float)(r[2 * g.tot + m] * inv_s));
    xv[3] = f2bf(feat[2 * v]);
    xv[4] = f2bf(feat[2 * v + 1]);
#pragma unroll
    for (int c = 5; c < 16; c++) xv[c] = 0;
}
// z (FP64 SoA) = s * y[v][0..2]
__global__ void k_unpack_output(Geom g, int64_t V, const float* __restrict__ y, double s, double* __restrict__ z,
                                int add) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
    for (int c = 0; c < 3; c++) {
        const double val = s * (double)y[v * 16 + c];
        z[c * g.tot + m] = add ? z[c * g.tot + m] + val : val;
    }
}
// feat = (log lambda, log mu) / fscale from the isotropic stiffness planes
__global__ void k_coef_features(Geom g, int64_t V, const double* __restrict__ C, double fscale,
                                float* __restrict__ feat) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t m = node_offset(g, v);
    feat[2 * v] = (float)(log(fmax(C[m], 1e-300)) * fscale);
    feat[2 * v + 1] = (float)(log(fmax(C[g.tot + m], 1e-300)) * fscale);
}

// Loss 0.5 * mean over voxels of |y - t|^2 (3 channels of y = [V][16], t = [V][3] f32) and its
// gradient dy = (y - t) / (3 V) in the first 3 channels.
__global__ void __launch_bounds__(256) k_mse(int64_t V, const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ dy, float* __restrict__ loss) {
    float s = 0.f;
    const float sc = 1.f / (3.f * (float)V);
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            float d = 0.f;
            if (c < 3) {
                d = y[v * 16 + c] - t[v * 3 + c];
                s += 0.5f * d * d * sc;
            }
            dy[v * 16 + c] = d * sc;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x % 64 == 0) atomicAdd(loss, s);
}

// Adam on the FP32 master parameters (Kingma & Ba 2015)
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.f - b1) * gq;
        const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    }
}

// Probe of the MFMA operand layout: D = A B for one 16x16x16 tile (row-major bf16 A, B).
__global__ void k_mfma_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, float* __restrict__ D) {
    const int l = threadIdx.x;
    s4 a, b;
    for (int i = 0; i < 4; i++) {
        a[i] = (short)A[(l % 16) * 16 + 4 * (l / 16) + i];
        b[i] = (short)B[(4 * (l / 16) + i) * 16 + l % 16];
    }
    f4 c = f4{0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    for (int i = 0; i < 4; i++) D[(4 * (l / 16) + i) * 16 + l % 16] = c[i];
}

}  // namespace net
}  // namespace alamo
