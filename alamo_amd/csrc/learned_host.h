// Host side of the learned correction network: parameters, buffers, forward / backward passes,
// Adam, checkpoints, training-pair generation and the preconditioner hook.  Included by capi.hip.
//
// Network (channel counts padded to 16; all convolutions 3x3x3, zero padding):
//   level 0 (N):    a0 = relu(conv0(x; 16->16))          x = (r/s, log lambda, log mu, 0...)
//   level 1 (N/2):  a1 = relu(conv1(inject(a0); 16->32))
//   level 2 (N/4):  a2 = relu(conv2(inject(a1); 32->64)),  b2 = conv3(a2; 64->32)
//   level 1:        u1 = P b2 + a1,  b1 = relu(conv4(u1; 32->16))
//   level 0:        u0 = P b1 + a0,  y  = conv5(u0; 16->16)    (channels 0-2 used)
// As a preconditioner stage: z = V(r); q = r - A z; y = net(q / rms(q)); z += alpha y with
// the residual-minimising alpha.  Trained on targets e / rms(e), e = A^-1 q.
#pragma once
#include <cstdio>
#include <random>
#include <vector>

#include "learned.h"

// internal status of the network stage: its output was not finite (z still holds the plain cycle)
constexpr int NET_NONFINITE = 100;

struct NetLayer {
    int cin, cout, cin_real, level, relu;
    int64_t woff, boff;  // offsets into the parameter vector
};

struct NetBufs {
    int N0[3] = {0, 0, 0};
    net::Grid3 g[3];
    // activations (bf16, channel-last)
    uint16_t *x = nullptr, *a0 = nullptr, *u0 = nullptr, *d1 = nullptr, *a1 = nullptr, *u1 = nullptr, *b1 = nullptr,
             *d2 = nullptr, *a2 = nullptr, *b2 = nullptr;
    float* y = nullptr;
    // backward
    float *dy = nullptr, *du0 = nullptr, *da0 = nullptr, *db1 = nullptr, *du1 = nullptr, *da1 = nullptr,
          *db2 = nullptr, *da2 = nullptr, *dd2 = nullptr, *dd1 = nullptr, *dx = nullptr;
    uint16_t* gb = nullptr;  // bf16 staging of a gradient fed to a convolution (largest: 32 ch at level 0 size)
    float* feat = nullptr;   // coefficient features at level 0
    float* loss = nullptr;
    float* wpart = nullptr;  // per-workgroup weight-gradient partials (net::WG_NWG x 27 x 64 x 32)
    std::vector<void*> allocs;
};

struct alamo_net {
    std::vector<NetLayer> L;
    int64_t nparam = 0;
    float *w = nullptr, *gw = nullptr, *m = nullptr, *v = nullptr;  // FP32 master, grads, Adam moments
    std::vector<uint16_t*> wf, wb;                                   // bf16 forward / backward-data weights
    NetBufs B;
    hipStream_t s = nullptr;
    int adam_t = 0;
    uint64_t seed = 0, train_seed = 0;
    bool wdirty = true;
    const alamo_mg* feat_owner = nullptr;  // handle and model version the features were built from
    uint64_t feat_version = 0;
    std::vector<void*> allocs;
};

// the LDS-staged convolution for 16 and 32 input channels (ALAMO_CONV_LDS=0 selects the plain kernel)
static const bool g_conv_lds = [] {
    const char* e = getenv("ALAMO_CONV_LDS");
    return !(e && strcmp(e, "0") == 0);
}();

static int net_alloc(std::vector<void*>& list, void** p, size_t bytes) {
    HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 16)));
    list.push_back(*p);
    HIPCHK(hipMemset(*p, 0, std::max<size_t>(bytes, 16)));
    HIPCHK(hipDeviceSynchronize());
    return ALAMO_OK;
}
#define NALLOC(list, ptr, bytes) TRY(net_alloc(list, (void**)&(ptr), (size_t)(bytes)))

static void net_free_bufs(NetBufs& B) {
    for (void* p : B.allocs) (void)hipFree(p);
    B = NetBufs();
}

static int net_bufs(alamo_net* n, const int* N0) {
    NetBufs& B = n->B;
    if (B.N0[0] == N0[0] && B.N0[1] == N0[1] && B.N0[2] == N0[2]) return ALAMO_OK;
    for (int d = 0; d < 3; d++)
        if ((N0[d] - 1) % 4 != 0 || N0[d] < 5) FAIL(ALAMO_ERR_ARG, "network grids need (N-1) divisible by 4, N >= 5");
    HIPCHK(hipStreamSynchronize(n->s));
    net_free_bufs(B);
    for (int d = 0; d < 3; d++) B.N0[d] = N0[d];
    for (int l = 0; l < 3; l++) {
        int64_t V = 1;
        for (int d = 0; d < 3; d++) {
            B.g[l].N[d] = (N0[d] - 1) / (1 << l) + 1;
            V *= B.g[l].N[d];
        }
        B.g[l].V = V;
    }
    const int64_t V0 = B.g[0].V, V1 = B.g[1].V, V2 = B.g[2].V;
    auto& A = B.allocs;
    NALLOC(A, B.x, V0 * 16 * 2);
    NALLOC(A, B.a0, V0 * 16 * 2);
    NALLOC(A, B.u0, V0 * 16 * 2);
    NALLOC(A, B.d1, V1 * 16 * 2);
    NALLOC(A, B.a1, V1 * 32 * 2);
    NALLOC(A, B.u1, V1 * 32 * 2);
    NALLOC(A, B.b1, V1 * 16 * 2);
    NALLOC(A, B.d2, V2 * 32 * 2);
    NALLOC(A, B.a2, V2 * 64 * 2);
    NALLOC(A, B.b2, V2 * 32 * 2);
    NALLOC(A, B.y, V0 * 16 * 4);
    NALLOC(A, B.dy, V0 * 16 * 4);
    NALLOC(A, B.du0, V0 * 16 * 4);
    NALLOC(A, B.da0, V0 * 16 * 4);
    NALLOC(A, B.db1, V1 * 16 * 4);
    NALLOC(A, B.du1, V1 * 32 * 4);
    NALLOC(A, B.da1, V1 * 32 * 4);
    NALLOC(A, B.db2, V2 * 32 * 4);
    NALLOC(A, B.da2, V2 * 64 * 4);
    NALLOC(A, B.dd2, V2 * 32 * 4);
    NALLOC(A, B.dd1, V1 * 16 * 4);
    NALLOC(A, B.dx, V0 * 16 * 4);
    NALLOC(A, B.gb, std::max(V0 * 16, V1 * 64) * 2);
    NALLOC(A, B.feat, V0 * 2 * 4);
    NALLOC(A, B.loss, 16);
    NALLOC(A, B.wpart, (int64_t)net::WG_NWG * 27 * 64 * 32 * 4);
    return ALAMO_OK;
}

static inline unsigned nb256(int64_t n) { return (unsigned)std::min<int64_t>(65535 * 8, (n + 255) / 256); }

// The bulk of a row runs in 64-voxel tiles; when the row is 64 k + r nodes long the last r
// nodes run in 16-voxel tiles, so 4k+1-node rows waste one 16-voxel tile, not a 64-voxel one.
template <int NT, int MT>
static void conv_pass(alamo_net* n, const net::Grid3& g, int cin, int xb, int xe, const uint16_t* in,
                      const uint16_t* W, const float* bias, void* out, bool out_f32, int relu) {
    const int tx = (xe - xb + 16 * MT - 1) / (16 * MT);
    if (tx <= 0) return;
    // LDS staging pays for 16 input channels and for 32 or 64 with at most 32 outputs (with 64
    // outputs its register footprint leaves one wavefront per SIMD); 64 input channels take tiles
    // of at most 32 voxels (a 64-voxel tile's three planes would fill the LDS of a CU)
    const bool lds64 = cin == 64 && NT <= 2 && MT <= 2;
    if ((cin == 16 || (cin == 32 && NT <= 2) || lds64) && g_conv_lds) {
        dim3 grid(tx, (g.N[1] + net::CL_TY - 1) / net::CL_TY, (g.N[2] + net::CL_KZ - 1) / net::CL_KZ);
        const unsigned blk = 64 * net::CL_TY;
        if (cin == 16) {
            if (out_f32) net::k_conv3_fwd_lds<16, NT, MT, true><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
            else net::k_conv3_fwd_lds<16, NT, MT, false><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
        } else if (cin == 64) {
            if constexpr (MT <= 2 && NT <= 2) {
                if (out_f32)
                    net::k_conv3_fwd_lds<64, NT, MT, true><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
                else
                    net::k_conv3_fwd_lds<64, NT, MT, false><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
            }
        } else {
            if (out_f32) net::k_conv3_fwd_lds<32, NT, MT, true><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
            else net::k_conv3_fwd_lds<32, NT, MT, false><<<grid, blk, 0, n->s>>>(g, xb, in, W, bias, out, relu);
        }
        return;
    }
    const int64_t tiles = (int64_t)tx * g.N[1] * g.N[2];
    const unsigned grid = (unsigned)((tiles + net::CONV_WAVES - 1) / net::CONV_WAVES);
    if (out_f32)
        net::k_conv3_fwd<NT, MT, true><<<grid, 64 * net::CONV_WAVES, 0, n->s>>>(g, cin, xb, tx, in, W, bias, out, relu);
    else
        net::k_conv3_fwd<NT, MT, false><<<grid, 64 * net::CONV_WAVES, 0, n->s>>>(g, cin, xb, tx, in, W, bias, out, relu);
}

template <int NT>
static void conv_launch_nt(alamo_net* n, const net::Grid3& g, int cin, const uint16_t* in, const uint16_t* W,
                           const float* bias, void* out, bool out_f32, int relu) {
    if (cin == 64 && NT <= 2 && g_conv_lds) {
        const int xm = (g.N[0] / 32) * 32;  // 32-voxel tiles over [0, xm), 16-voxel tiles over [xm, N0)
        conv_pass<NT, 2>(n, g, cin, 0, xm, in, W, bias, out, out_f32, relu);
        conv_pass<NT, 1>(n, g, cin, xm, g.N[0], in, W, bias, out, out_f32, relu);
        return;
    }
    const int xm = (g.N[0] / 64) * 64;  // 64-voxel tiles over [0, xm), 16-voxel tiles over [xm, N0)
    conv_pass<NT, 4>(n, g, cin, 0, xm, in, W, bias, out, out_f32, relu);
    conv_pass<NT, 1>(n, g, cin, xm, g.N[0], in, W, bias, out, out_f32, relu);
}

static void conv_launch(alamo_net* n, const net::Grid3& g, int cin, int cout, const uint16_t* in, const uint16_t* W,
                        const float* bias, void* out, bool out_f32, int relu) {
    if (cout == 16) conv_launch_nt<1>(n, g, cin, in, W, bias, out, out_f32, relu);
    else if (cout == 32) conv_launch_nt<2>(n, g, cin, in, W, bias, out, out_f32, relu);
    else conv_launch_nt<4>(n, g, cin, in, W, bias, out, out_f32, relu);
}

template <int CIN, int COUT>
static void wgrad_pass(alamo_net* n, const net::Grid3& g, const uint16_t* in, const uint16_t* dout, float* dW) {
    const int64_t ntile = (int64_t)((g.N[0] + 31) / 32) * ((g.N[1] + net::WG_TY - 1) / net::WG_TY) *
                          ((g.N[2] + net::WG_TZ - 1) / net::WG_TZ);
    const int nwg = (int)std::min<int64_t>(ntile, net::WG_NWG);
    net::k_conv3_wgrad_lds<CIN, COUT><<<nwg, 256, 0, n->s>>>(g, in, dout, n->B.wpart);
    const int64_t nw = (int64_t)27 * CIN * COUT;
    net::k_wgrad_reduce<<<(unsigned)((nw + 255) / 256), 256, 0, n->s>>>(nwg, nw, n->B.wpart, dW);
}

// dW += weight gradient, dB += bias gradient of one convolution layer (in: its forward input,
// dout: the gradient of its output, both bf16 channel-last)
static void wgrad_launch(alamo_net* n, const net::Grid3& g, int cin, int cout, const uint16_t* in,
                         const uint16_t* dout, float* dW, float* dB) {
    if (cin == 16 && cout == 16) wgrad_pass<16, 16>(n, g, in, dout, dW);
    else if (cin == 16 && cout == 32) wgrad_pass<16, 32>(n, g, in, dout, dW);
    else if (cin == 32 && cout == 16) wgrad_pass<32, 16>(n, g, in, dout, dW);
    else if (cin == 32 && cout == 64) wgrad_pass<32, 64>(n, g, in, dout, dW);
    else wgrad_pass<64, 32>(n, g, in, dout, dW);
    if (dB)
        net::k_bias_grad<<<(unsigned)std::min<int64_t>(2048, (g.V * cout + 255) / 256), 256, 0, n->s>>>(g.V, cout, dout,
                                                                                                         dB);
}

static int net_sync_weights(alamo_net* n) {
    if (!n->wdirty) return ALAMO_OK;
    for (size_t l = 0; l < n->L.size(); l++) {
        const NetLayer& ly = n->L[l];
        const int64_t nw = (int64_t)27 * ly.cin * ly.cout;
        net::k_f32_to_bf16<<<nb256(nw), 256, 0, n->s>>>(n->w + ly.woff, n->wf[l], nw);
        net::k_wflip<<<nb256(nw), 256, 0, n->s>>>(ly.cin, ly.cout, n->w + ly.woff, n->wb[l]);
    }
    HIPCHK(hipGetLastError());
    n->wdirty = false;
    return ALAMO_OK;
}

// forward from B.x (bf16 level-0 input) to B.y (f32)
static int net_forward(alamo_net* n) {
    TRY(net_sync_weights(n));
    NetBufs& B = n->B;
    auto& L = n->L;
    auto W = [&](int l) { return (const uint16_t*)n->wf[l]; };
    auto bias = [&](int l) { return (const float*)(n->w + L[l].boff); };
    conv_launch(n, B.g[0], 16, 16, B.x, W(0), bias(0), B.a0, false, 1);
    net::k_inject<<<nb256(B.g[1].V * 16), 256, 0, n->s>>>(B.g[0], B.g[1], 16, B.a0, B.d1);
    conv_launch(n, B.g[1], 16, 32, B.d1, W(1), bias(1), B.a1, false, 1);
    net::k_inject<<<nb256(B.g[2].V * 32), 256, 0, n->s>>>(B.g[1], B.g[2], 32, B.a1, B.d2);
    conv_launch(n, B.g[2], 32, 64, B.d2, W(2), bias(2), B.a2, false, 1);
    conv_launch(n, B.g[2], 64, 32, B.a2, W(3), bias(3), B.b2, false, 0);
    net::k_prolong_add<uint16_t><<<nb256(B.g[1].V * 32), 256, 0, n->s>>>(B.g[1], B.g[2], 32, B.b2, B.a1, B.u1);
    conv_launch(n, B.g[1], 32, 16, B.u1, W(4), bias(4), B.b1, false, 1);
    net::k_prolong_add<uint16_t><<<nb256(B.g[0].V * 16), 256, 0, n->s>>>(B.g[0], B.g[1], 16, B.b1, B.a0, B.u0);
    conv_launch(n, B.g[0], 16, 16, B.u0, W(5), bias(5), B.y, true, 0);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

// backward from B.dy (f32 [V0][16]); accumulates into n->gw; optional dx
static int net_backward(alamo_net* n, bool want_dx) {
    NetBufs& B = n->B;
    auto& L = n->L;
    auto gW = [&](int l) { return n->gw + L[l].woff; };
    auto gB = [&](int l) { return n->gw + L[l].boff; };
    const int64_t V0 = B.g[0].V, V1 = B.g[1].V, V2 = B.g[2].V;
    // conv5 (no ReLU): y = conv5(u0)
    net::k_relu_bwd<<<nb256(V0 * 16), 256, 0, n->s>>>(B.dy, nullptr, B.gb, V0 * 16, 0);
    wgrad_launch(n, B.g[0], 16, 16, B.u0, B.gb, gW(5), gB(5));
    conv_launch(n, B.g[0], 16, 16, B.gb, n->wb[5], nullptr, B.du0, true, 0);
    // u0 = P b1 + a0
    net::k_prolong_T<<<nb256(V1 * 16), 256, 0, n->s>>>(B.g[0], B.g[1], 16, B.du0, B.db1);
    HIPCHK(hipMemcpyAsync(B.da0, B.du0, V0 * 16 * 4, hipMemcpyDeviceToDevice, n->s));
    // b1 = relu(conv4(u1))
    net::k_relu_bwd<<<nb256(V1 * 16), 256, 0, n->s>>>(B.db1, B.b1, B.gb, V1 * 16, 1);
    wgrad_launch(n, B.g[1], 32, 16, B.u1, B.gb, gW(4), gB(4));
    conv_launch(n, B.g[1], 16, 32, B.gb, n->wb[4], nullptr, B.du1, true, 0);
    // u1 = P b2 + a1
    net::k_prolong_T<<<nb256(V2 * 32), 256, 0, n->s>>>(B.g[1], B.g[2], 32, B.du1, B.db2);
    HIPCHK(hipMemcpyAsync(B.da1, B.du1, V1 * 32 * 4, hipMemcpyDeviceToDevice, n->s));
    // b2 = conv3(a2)
    net::k_relu_bwd<<<nb256(V2 * 32), 256, 0, n->s>>>(B.db2, nullptr, B.gb, V2 * 32, 0);
    wgrad_launch(n, B.g[2], 64, 32, B.a2, B.gb, gW(3), gB(3));
    conv_launch(n, B.g[2], 32, 64, B.gb, n->wb[3], nullptr, B.da2, true, 0);
    // a2 = relu(conv2(d2))
    net::k_relu_bwd<<<nb256(V2 * 64), 256, 0, n->s>>>(B.da2, B.a2, B.gb, V2 * 64, 1);
    wgrad_launch(n, B.g[2], 32, 64, B.d2, B.gb, gW(2), gB(2));
    conv_launch(n, B.g[2], 64, 32, B.gb, n->wb[2], nullptr, B.dd2, true, 0);
    // d2 = inject(a1)
    net::k_inject_bwd<<<nb256(V2 * 32), 256, 0, n->s>>>(B.g[1], B.g[2], 32, B.dd2, B.da1);
    // a1 = relu(conv1(d1))
    net::k_relu_bwd<<<nb256(V1 * 32), 256, 0, n->s>>>(B.da1, B.a1, B.gb, V1 * 32, 1);
    wgrad_launch(n, B.g[1], 16, 32, B.d1, B.gb, gW(1), gB(1));
    conv_launch(n, B.g[1], 32, 16, B.gb, n->wb[1], nullptr, B.dd1, true, 0);
    // d1 = inject(a0)
    net::k_inject_bwd<<<nb256(V1 * 16), 256, 0, n->s>>>(B.g[0], B.g[1], 16, B.dd1, B.da0);
    // a0 = relu(conv0(x))
    net::k_relu_bwd<<<nb256(V0 * 16), 256, 0, n->s>>>(B.da0, B.a0, B.gb, V0 * 16, 1);
    wgrad_launch(n, B.g[0], 16, 16, B.x, B.gb, gW(0), gB(0));
    if (want_dx) conv_launch(n, B.g[0], 16, 16, B.gb, n->wb[0], nullptr, B.dx, true, 0);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

static int net_adam(alamo_net* n, float lr) {
    n->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
    net::k_adam<<<nb256(n->nparam), 256, 0, n->s>>>(n->nparam, n->w, n->gw, n->m, n->v, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipMemsetAsync(n->gw, 0, n->nparam * 4, n->s));
    n->wdirty = true;
    return ALAMO_OK;
}

static int net_create(alamo_net** out, uint64_t seed) {
    alamo_net* n = new alamo_net();
    n->seed = seed;
    const int spec[6][5] = {// cin, cout, cin_real, level, relu
                            {16, 16, 5, 0, 1}, {16, 32, 16, 1, 1}, {32, 64, 32, 2, 1},
                            {64, 32, 64, 2, 0}, {32, 16, 32, 1, 1}, {16, 16, 16, 0, 0}};
    int64_t off = 0;
    for (auto& s : spec) {
        NetLayer ly{s[0], s[1], s[2], s[3], s[4], 0, 0};
        ly.woff = off;
        off += (int64_t)27 * ly.cin * ly.cout;
        ly.boff = off;
        off += ly.cout;
        n->L.push_back(ly);
    }
    n->nparam = off;
    // He-normal initial weights from SplitMix64 + Box-Muller; input channels beyond the real
    // ones and the unused output channels of the last layer start at zero; the last layer is
    // scaled by 0.1 so the correction starts small.
    std::vector<float> w(off, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t l = 0; l < n->L.size(); l++) {
        const NetLayer& ly = n->L[l];
        const double sd = std::sqrt(2.0 / (27.0 * ly.cin_real)) * (l + 1 == n->L.size() ? 0.1 : 1.0);
        for (int o = 0; o < 27; o++)
            for (int co = 0; co < ly.cout; co++)
                for (int ci = 0; ci < ly.cin; ci++) {
                    const double u1 = next(), u2 = next();
                    double g = std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * M_PI * u2);
                    if (ci >= ly.cin_real || (l + 1 == n->L.size() && co >= 3)) g = 0.0;
                    w[ly.woff + ((int64_t)o * ly.cout + co) * ly.cin + ci] = (float)(g * sd);
                }
    }
    int rc = ALAMO_OK;
    auto al = [&](void** p, size_t b) { if (rc == ALAMO_OK) rc = net_alloc(n->allocs, p, b); };
    al((void**)&n->w, off * 4);
    al((void**)&n->gw, off * 4);
    al((void**)&n->m, off * 4);
    al((void**)&n->v, off * 4);
    for (auto& ly : n->L) {
        uint16_t *a = nullptr, *b = nullptr;
        al((void**)&a, (size_t)27 * ly.cin * ly.cout * 2);
        al((void**)&b, (size_t)27 * ly.cin * ly.cout * 2);
        n->wf.push_back(a);
        n->wb.push_back(b);
    }
    if (rc != ALAMO_OK) {
        for (void* p : n->allocs) (void)hipFree(p);
        delete n;
        return rc;
    }
    if (hipMemcpy(n->w, w.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->gw, 0, off * 4) != hipSuccess || hipMemset(n->m, 0, off * 4) != hipSuccess ||
        hipMemset(n->v, 0, off * 4) != hipSuccess) {
        for (void* p : n->allocs) (void)hipFree(p);
        delete n;
        FAIL(ALAMO_ERR_HIP, "network parameter upload failed");
    }
    *out = n;
    return ALAMO_OK;
}

// ---------------------------------------------------------------------------
// Preconditioner stage and training data
// ---------------------------------------------------------------------------
__global__ void k_noise(int64_t n, uint64_t seed, double* __restrict__ x) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)(q + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[q] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

static int net_features(alamo_net* n, alamo_mg* h) {
    const Geom& g = h->L[0].g;
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    if (h->sym != ISO) FAIL(ALAMO_ERR_ARG, "the learned preconditioner needs an isotropic operator");
    net::k_coef_features<<<nb256(V), 256, 0, n->s>>>(g, V, h->L[0].C, 1.0 / 7.0, n->B.feat);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}

static int net_level0(alamo_net* n, alamo_mg* h) {
    const int N0[3] = {h->L[0].g.N[0], h->L[0].g.N[1], h->L[0].g.N[2]};
    const bool fresh = !(n->B.N0[0] == N0[0] && n->B.N0[1] == N0[1] && n->B.N0[2] == N0[2]) ||
                       n->feat_owner != h || n->feat_version != h->model_version;
    TRY(net_bufs(n, N0));
    n->s = h->stream;
    if (fresh) {
        TRY(net_features(n, h));
        n->feat_owner = h;
        n->feat_version = h->model_version;
    }
    return ALAMO_OK;
}

// z += alpha y with q = r - A z, y = net(q / rms(q)) and alpha = <q, A y> / <A y, A y>, the
// step that minimises |q - alpha A y|_2, so the stage never raises the residual.  w holds three
// FP64 vectors: q, A y, y.
static int net_correct(alamo_net* n, alamo_mg* h, const double* r, double* z, double* w) {
    LevelBuf& l0 = h->L[0];
    // V network voxels; n3 elements of a level-0 vector
    const int64_t V = (int64_t)l0.g.N[0] * l0.g.N[1] * l0.g.N[2], n3 = 3 * l0.g.tot;
    double *q = w, *Ay = w + n3, *y = w + 2 * n3;
    TRY(net_level0(n, h));
    launch_node<double>(h->sym, RESID, l0.g, h->bc, l0.C, z, r, q, 0.0, h->stream);
    double ql2 = 0;
    TRY(norms_dev<double>(h, q, n3, nullptr, &ql2));
    const double s = ql2 / std::sqrt(3.0 * (double)V);
    if (!(s > 0.0) || !std::isfinite(s)) return ALAMO_OK;
    net::k_pack_input<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, q, 1.0 / s, n->B.feat, n->B.x);
    TRY(net_forward(n));
    net::k_unpack_output<<<nb256(V), 256, 0, h->stream>>>(l0.g, V, n->B.y, 1.0, y, 0);
    launch_node<double>(h->sym, APPLY, l0.g, h->bc, l0.C, y, nullptr, Ay, 0.0, h->stream);
    double d[2];
    TRY(multidot(h, Ay, q, n3, 2, d));  // <Ay, q>, <Ay, Ay>
    if (!std::isfinite(d[0]) || !std::isfinite(d[1])) return NET_NONFINITE;  // z keeps the plain cycle
    if (!(d[1] > 0.0)) return ALAMO_OK;                                       // y = 0: nothing to add
    const double alpha = d[0] / d[1];
    k_axpy<double, double><<<gs_blocks(n3), 256, 0, h->stream>>>(z, y, n3, alpha);
    HIPCHK(hipGetLastError());
    return ALAMO_OK;
}
