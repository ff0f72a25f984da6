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
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

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
// Forward convolution for 16, 32 or 64 input channels with the input staged in LDS: a workgroup of
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
        // gfx950's 16x16x32 products (operand layout at k_conv3_wgrad_lds): K = 32 is two offsets of
        // the 16 channels (lanes 0-31 the even offset of a pair, lanes 32-63 the odd one; the 27th
        // offset pairs with zero weights), or the 32 channels of one offset
        // (64 channels: two products per offset, channels 0-31 and 32-63)
        constexpr int NP = CIN == 16 ? 14 : 27 * (CIN / 32);
        const int half = lane >> 5;
        for (int pp = 0; pp < NP; pp++) {
            const int o = CIN == 16 ? 2 * pp + half : (CIN == 32 ? pp : pp >> 1);
            const int c0 = CIN == 16 ? 8 * ((lane >> 4) & 1) : (CIN == 32 ? 0 : 32 * (pp & 1)) + 8 * (lane >> 4);
            const bool live = o < 27;
            int dx, dy, dz;
            off3(live ? o : 26, dx, dy, dz);
            const uint16_t(*pl)[CW][CIN] = tile[slot(z + dz)];
            bf8 bw[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                uint4 wv = make_uint4(0, 0, 0, 0);
                if (live) wv = *(const uint4*)(W + ((int64_t)o * Cout + t * 16 + m) * CIN + c0);
                bw[t] = __builtin_bit_cast(bf8, wv);
            }
#pragma unroll
            for (int a = 0; a < MT; a++) {
                const bf8 av = __builtin_bit_cast(bf8, *(const uint4*)&pl[w + 1 + dy][16 * a + m + dx + 1][c0]);
#pragma unroll
                for (int t = 0; t < NT; t++) acc[a][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[t], acc[a][t], 0, 0, 0);
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
// Weight gradient as an LDS-staged implicit GEMM on v_mfma_f32_16x16x32_bf16 (gfx950), two
// passes, no float atomics:
//   pass 1 (k_conv3_wgrad_lds): workgroup b accumulates part[b][o][co][ci] = sum over the voxels of
//     its tiles of dout[v][co] in[v + o][ci];
//   pass 2 (k_wgrad_reduce): dW[i] += sum_b part[b][i] in a fixed order (deterministic).
// A tile is 32 voxels in x (one MFMA K chunk) x WG_TY rows x WG_TZ planes; the workgroup stages
// dout of the tile (all Cout channels) and the 16-channel input slice ci0 .. ci0 + 15 of the tile
// plus its one-voxel halo in LDS, channel-major (voxels of a row contiguous), so that a lane's 8
// consecutive voxels of one channel -- its MFMA operand -- are one 16-byte LDS read; the x shift
// of an offset (dx = -1, +1) reads two aligned 8-voxel groups and funnels them by one element.
// Operand layout of the 16x16x32 product (checked by tests/test_gpu_learned.py):
//   A (16 x 32, M x K): lane l holds A[l % 16][8 (l / 16) + 0..7]
//   B (32 x 16, K x N): lane l holds B[8 (l / 16) + 0..7][l % 16]
//   D (16 x 16):        lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// Here M = co (A = dout^T of a 32-voxel row), N = ci (B = the row's input shifted by o), K = voxels.
// Wavefront w of the 4 owns the offsets o = w, w + 4, ... (7 or 6 of the 27) for every co tile.
// ---------------------------------------------------------------------------
constexpr int WG_TY = 4, WG_TZ = 2;
constexpr int WG_ROWS = WG_TY * WG_TZ;                  // K chunks (rows of 32 voxels) per tile
constexpr int WG_SD = WG_ROWS * 32 + 8;                 // dout: elements per channel (+8: bank spread)
constexpr int WG_HR = (WG_TY + 2) * (WG_TZ + 2);        // input rows of a tile with its halo
constexpr int WG_SI = WG_HR * 48 + 8;                   // input: elements per channel; x at 8 + (x - x0)
constexpr int WG_NWG = 256;                              // pass-1 workgroups at most

__device__ __forceinline__ uint4 funnel1(uint4 lo, uint4 hi, bool up) {
    // the 8 elements starting one element after (up) or before (!up) group `up ? lo : hi`
    if (up)
        return make_uint4(__builtin_amdgcn_alignbyte(lo.y, lo.x, 2), __builtin_amdgcn_alignbyte(lo.z, lo.y, 2),
                          __builtin_amdgcn_alignbyte(lo.w, lo.z, 2), __builtin_amdgcn_alignbyte(hi.x, lo.w, 2));
    return make_uint4(__builtin_amdgcn_alignbyte(hi.x, lo.w, 2), __builtin_amdgcn_alignbyte(hi.y, hi.x, 2),
                      __builtin_amdgcn_alignbyte(hi.z, hi.y, 2), __builtin_amdgcn_alignbyte(hi.w, hi.z, 2));
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv3_wgrad_lds(Grid3 g, const uint16_t* __restrict__ in,
                                                         const uint16_t* __restrict__ dout, float* __restrict__ part) {
    constexpr int NCO = COUT / 16, NCI = CIN / 16, NJ = 7;
    __shared__ __align__(16) uint16_t sd[COUT * WG_SD];
    __shared__ __align__(16) uint16_t si[16 * WG_SI];
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    const int m = lane % 16, q = lane / 16;
    const int tx = (g.N[0] + 31) / 32, ty = (g.N[1] + WG_TY - 1) / WG_TY, tz = (g.N[2] + WG_TZ - 1) / WG_TZ;
    const int64_t ntile = (int64_t)tx * ty * tz;
    float* pb = part + (int64_t)blockIdx.x * 27 * CIN * COUT;
    for (int cit = 0; cit < NCI; cit++) {
        f4 acc[NJ][NCO];
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int c = 0; c < NCO; c++) acc[j][c] = f4{0.f, 0.f, 0.f, 0.f};
        for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
            const int x0 = 32 * (int)(tile % tx);
            const int y0 = WG_TY * (int)((tile / tx) % ty);
            const int z0 = WG_TZ * (int)(tile / ((int64_t)tx * ty));
            __syncthreads();  // the previous tile's reads of sd / si are done
            // dout of the tile: item = (voxel of the 32 x WG_ROWS tile, 8-channel chunk)
            for (int it = t; it < 32 * WG_ROWS * (COUT / 8); it += 256) {
                const int ch = it % (COUT / 8), vox = it / (COUT / 8);
                const int xx = vox % 32, r = vox / 32;
                const int x = x0 + xx, y = y0 + r % WG_TY, z = z0 + r / WG_TY;
                uint4 d = make_uint4(0, 0, 0, 0);
                if (x < g.N[0] && y < g.N[1] && z < g.N[2])
                    d = *(const uint4*)(dout + (x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z)) * COUT + 8 * ch);
                const uint32_t dd[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
                for (int e = 0; e < 8; e++)
                    sd[(8 * ch + e) * WG_SD + r * 32 + xx] = (uint16_t)(dd[e / 2] >> (16 * (e % 2)));
            }
            // input slice of the tile with its halo: x0 - 1 .. x0 + 32 at LDS x = 7 .. 40
            for (int it = t; it < 34 * WG_HR * 2; it += 256) {
                const int ch = it % 2, vox = it / 2;
                const int xx = vox % 34, hr = vox / 34;
                const int x = x0 - 1 + xx, y = y0 - 1 + hr % (WG_TY + 2), z = z0 - 1 + hr / (WG_TY + 2);
                uint4 d = make_uint4(0, 0, 0, 0);
                if (x >= 0 && x < g.N[0] && y >= 0 && y < g.N[1] && z >= 0 && z < g.N[2])
                    d = *(const uint4*)(in + (x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z)) * CIN + 16 * cit +
                                        8 * ch);
                const uint32_t dd[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
                for (int e = 0; e < 8; e++)
                    si[(8 * ch + e) * WG_SI + hr * 48 + 7 + xx] = (uint16_t)(dd[e / 2] >> (16 * (e % 2)));
            }
            __syncthreads();
            for (int r = 0; r < WG_ROWS; r++) {
                const int ry = r % WG_TY, rz = r / WG_TY;
                bf8 a[NCO];
#pragma unroll
                for (int c = 0; c < NCO; c++)
                    a[c] = __builtin_bit_cast(bf8, *(const uint4*)&sd[(16 * c + m) * WG_SD + r * 32 + 8 * q]);
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int o = w + 4 * j;
                    if (o >= 27) break;  // uniform per wavefront
                    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
                    const int hr = (ry + 1 + dy) + (WG_TY + 2) * (rz + 1 + dz);
                    const uint16_t* sb = &si[m * WG_SI + hr * 48 + 8 * q];
                    uint4 bv;
                    if (dx == 0) bv = *(const uint4*)(sb + 8);
                    else if (dx > 0) bv = funnel1(*(const uint4*)(sb + 8), *(const uint4*)(sb + 16), true);
                    else bv = funnel1(*(const uint4*)sb, *(const uint4*)(sb + 8), false);
                    const bf8 b = __builtin_bit_cast(bf8, bv);
#pragma unroll
                    for (int c = 0; c < NCO; c++) acc[j][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b, acc[j][c], 0, 0, 0);
                }
            }
        }
        // D[co = 16 c + 4 q + i][ci = 16 cit + m] of offset o
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int o = w + 4 * j;
            if (o >= 27) break;
#pragma unroll
            for (int c = 0; c < NCO; c++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    pb[((int64_t)o * COUT + 16 * c + 4 * q + i) * CIN + 16 * cit + m] = acc[j][c][i];
        }
    }
}

// dW[i] += sum over the nwg partials (fixed order)
__global__ void __launch_bounds__(256) k_wgrad_reduce(int nwg, int64_t n, const float* __restrict__ part,
                                                      float* __restrict__ dW) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int b = 0; b < nwg; b++) s += part[(int64_t)b * n + i];
    dW[i] += s;
}

// Probe of the 16x16x32 operand layout: D = A B for row-major bf16 A (16 x 32) and B (32 x 16).
__global__ void k_mfma_probe32(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, float* __restrict__ D) {
    const int l = threadIdx.x;
    uint16_t a[8], b[8];
    for (int i = 0; i < 8; i++) {
        a[i] = A[(l % 16) * 32 + 8 * (l / 16) + i];
        b[i] = B[(8 * (l / 16) + i) * 16 + l % 16];
    }
    f4 c = f4{0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
    for (int i = 0; i < 4; i++) D[(4 * (l / 16) + i) * 16 + l % 16] = c[i];
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
