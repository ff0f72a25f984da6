// Oracle (test infrastructure only) for the learned correction network: a plain FP64
// restatement of the network's forward and backward passes.  No reference counterpart: the
// reference has no learned preconditioner; this file states the network the library defines
// (3 levels, 16/32/64 channels, 3x3x3 zero-padded convolutions, node injection down,
// trilinear nodal prolongation up, additive skips).
//
// With `bf16` set, values are rounded to bfloat16 (round to nearest even) at exactly the points
// where the device stores them in bfloat16 (inputs, weights, activations, the gradients fed to
// convolutions), so the device result differs from this one only by FP32 accumulation order.
// With `bf16` clear everything is FP64, which is what the finite-difference gradient checks use.
//
// Layouts: activations channel-last [v][c], v = x + Nx (y + Ny z); weights [27][cout][cin] with
// offset o = (dz+1)*9 + (dy+1)*3 + (dx+1) and out[v] = sum W[o] in[v + offset(o)]; parameters
// are the concatenation over layers of W then bias[cout].
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace ocv {

static double bf(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct G {
    int N[3];
    int64_t V;
};
static G grid(const int* N) { return G{{N[0], N[1], N[2]}, (int64_t)N[0] * N[1] * N[2]}; }
static G coarse(const G& f) {
    int N[3] = {(f.N[0] - 1) / 2 + 1, (f.N[1] - 1) / 2 + 1, (f.N[2] - 1) / 2 + 1};
    return grid(N);
}

// out[v][co] = bias[co] + sum_{o, ci} W[o][co][ci] in[v + off(o)][ci], zero padding
static void conv(const G& g, int cin, int cout, const double* in, const double* W, const double* bias, double* out) {
    for (int z = 0; z < g.N[2]; z++)
        for (int y = 0; y < g.N[1]; y++)
            for (int x = 0; x < g.N[0]; x++) {
                const int64_t v = x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
                double* ov = out + v * cout;
                for (int co = 0; co < cout; co++) ov[co] = bias ? bias[co] : 0.0;
                for (int o = 0; o < 27; o++) {
                    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
                    const int xs = x + dx, ys = y + dy, zs = z + dz;
                    if (xs < 0 || ys < 0 || zs < 0 || xs >= g.N[0] || ys >= g.N[1] || zs >= g.N[2]) continue;
                    const double* iv = in + (xs + (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs)) * cin;
                    for (int co = 0; co < cout; co++) {
                        const double* w = W + ((int64_t)o * cout + co) * cin;
                        double s = 0;
                        for (int ci = 0; ci < cin; ci++) s += w[ci] * iv[ci];
                        ov[co] += s;
                    }
                }
            }
}

// adjoint of conv with respect to its input: din[u][ci] = sum_{o, co} W[o][co][ci] dout[u - off(o)][co]
static void conv_T(const G& g, int cin, int cout, const double* dout, const double* W, double* din) {
    std::fill(din, din + g.V * cin, 0.0);
    for (int z = 0; z < g.N[2]; z++)
        for (int y = 0; y < g.N[1]; y++)
            for (int x = 0; x < g.N[0]; x++) {
                const int64_t v = x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
                const double* dv = dout + v * cout;
                for (int o = 0; o < 27; o++) {
                    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
                    const int xs = x + dx, ys = y + dy, zs = z + dz;
                    if (xs < 0 || ys < 0 || zs < 0 || xs >= g.N[0] || ys >= g.N[1] || zs >= g.N[2]) continue;
                    double* iv = din + (xs + (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs)) * cin;
                    for (int co = 0; co < cout; co++) {
                        const double* w = W + ((int64_t)o * cout + co) * cin;
                        for (int ci = 0; ci < cin; ci++) iv[ci] += w[ci] * dv[co];
                    }
                }
            }
}

// gradients: dW[o][co][ci] += sum_v dout[v][co] in[v + off(o)][ci], dB[co] += sum_v dout[v][co]
static void conv_wgrad(const G& g, int cin, int cout, const double* in, const double* dout, double* dW, double* dB) {
    for (int z = 0; z < g.N[2]; z++)
        for (int y = 0; y < g.N[1]; y++)
            for (int x = 0; x < g.N[0]; x++) {
                const int64_t v = x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
                const double* dv = dout + v * cout;
                for (int co = 0; co < cout; co++) dB[co] += dv[co];
                for (int o = 0; o < 27; o++) {
                    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
                    const int xs = x + dx, ys = y + dy, zs = z + dz;
                    if (xs < 0 || ys < 0 || zs < 0 || xs >= g.N[0] || ys >= g.N[1] || zs >= g.N[2]) continue;
                    const double* iv = in + (xs + (int64_t)g.N[0] * (ys + (int64_t)g.N[1] * zs)) * cin;
                    for (int co = 0; co < cout; co++) {
                        double* w = dW + ((int64_t)o * cout + co) * cin;
                        for (int ci = 0; ci < cin; ci++) w[ci] += dv[co] * iv[ci];
                    }
                }
            }
}

// node injection d[I] = a[2I] and its adjoint
static void inject(const G& gf, const G& gc, int C, const double* a, double* d) {
    for (int Z = 0; Z < gc.N[2]; Z++)
        for (int Y = 0; Y < gc.N[1]; Y++)
            for (int X = 0; X < gc.N[0]; X++) {
                const int64_t I = X + (int64_t)gc.N[0] * (Y + (int64_t)gc.N[1] * Z);
                const int64_t i = 2 * X + (int64_t)gf.N[0] * (2 * Y + (int64_t)gf.N[1] * 2 * Z);
                for (int c = 0; c < C; c++) d[I * C + c] = a[i * C + c];
            }
}
static void inject_T(const G& gf, const G& gc, int C, const double* gd, double* ga) {
    for (int Z = 0; Z < gc.N[2]; Z++)
        for (int Y = 0; Y < gc.N[1]; Y++)
            for (int X = 0; X < gc.N[0]; X++) {
                const int64_t I = X + (int64_t)gc.N[0] * (Y + (int64_t)gc.N[1] * Z);
                const int64_t i = 2 * X + (int64_t)gf.N[0] * (2 * Y + (int64_t)gf.N[1] * 2 * Z);
                for (int c = 0; c < C; c++) ga[i * C + c] += gd[I * C + c];
            }
}

// trilinear nodal prolongation (out = P c) and its adjoint (gc = P^T gu)
static double w1(int i, int I) {
    const int d = i - 2 * I;
    return d == 0 ? 1.0 : (d == 1 || d == -1) ? 0.5 : 0.0;
}
static void prolong(const G& gf, const G& gc, int C, const double* c, double* out) {
    for (int z = 0; z < gf.N[2]; z++)
        for (int y = 0; y < gf.N[1]; y++)
            for (int x = 0; x < gf.N[0]; x++) {
                const int64_t v = x + (int64_t)gf.N[0] * (y + (int64_t)gf.N[1] * z);
                for (int ch = 0; ch < C; ch++) out[v * C + ch] = 0.0;
                for (int Z = z / 2; Z <= (z + 1) / 2; Z++)
                    for (int Y = y / 2; Y <= (y + 1) / 2; Y++)
                        for (int X = x / 2; X <= (x + 1) / 2; X++) {
                            const double w = w1(x, X) * w1(y, Y) * w1(z, Z);
                            const int64_t I = X + (int64_t)gc.N[0] * (Y + (int64_t)gc.N[1] * Z);
                            for (int ch = 0; ch < C; ch++) out[v * C + ch] += w * c[I * C + ch];
                        }
            }
}
static void prolong_T(const G& gf, const G& gc, int C, const double* gu, double* gcv) {
    std::fill(gcv, gcv + gc.V * C, 0.0);
    for (int z = 0; z < gf.N[2]; z++)
        for (int y = 0; y < gf.N[1]; y++)
            for (int x = 0; x < gf.N[0]; x++) {
                const int64_t v = x + (int64_t)gf.N[0] * (y + (int64_t)gf.N[1] * z);
                for (int Z = z / 2; Z <= (z + 1) / 2; Z++)
                    for (int Y = y / 2; Y <= (y + 1) / 2; Y++)
                        for (int X = x / 2; X <= (x + 1) / 2; X++) {
                            const double w = w1(x, X) * w1(y, Y) * w1(z, Z);
                            const int64_t I = X + (int64_t)gc.N[0] * (Y + (int64_t)gc.N[1] * Z);
                            for (int ch = 0; ch < C; ch++) gcv[I * C + ch] += w * gu[v * C + ch];
                        }
            }
}

struct Layer {
    int cin, cout, relu;
};
static const Layer LAYERS[6] = {{16, 16, 1}, {16, 32, 1}, {32, 64, 1}, {64, 32, 0}, {32, 16, 1}, {16, 16, 0}};

static int64_t nparam() {
    int64_t n = 0;
    for (auto& l : LAYERS) n += (int64_t)27 * l.cin * l.cout + l.cout;
    return n;
}

struct Net {
    G g0, g1, g2;
    bool bf16;
    std::vector<double> W[6], B[6];
    // forward state
    std::vector<double> x, a0, d1, a1, d2, a2, b2, u1, b1, u0, y;

    Net(const int* N0, const float* params, bool rb) : bf16(rb) {
        g0 = grid(N0);
        g1 = coarse(g0);
        g2 = coarse(g1);
        int64_t off = 0;
        for (int l = 0; l < 6; l++) {
            const int64_t nw = (int64_t)27 * LAYERS[l].cin * LAYERS[l].cout;
            W[l].resize(nw);
            for (int64_t q = 0; q < nw; q++) W[l][q] = r(params[off + q]);
            off += nw;
            B[l].assign(params + off, params + off + LAYERS[l].cout);  // biases stay FP32
            off += LAYERS[l].cout;
        }
    }
    double r(double v) const { return bf16 ? bf(v) : v; }
    void round(std::vector<double>& a) const {
        if (bf16)
            for (auto& v : a) v = bf(v);
    }
    void act(std::vector<double>& a, bool relu) const {
        if (relu)
            for (auto& v : a) v = v > 0 ? v : 0.0;
        round(a);
    }

    void forward(const float* xin) {
        x.assign(xin, xin + g0.V * 16);
        round(x);
        a0.resize(g0.V * 16);
        conv(g0, 16, 16, x.data(), W[0].data(), B[0].data(), a0.data());
        act(a0, true);
        d1.resize(g1.V * 16);
        inject(g0, g1, 16, a0.data(), d1.data());
        a1.resize(g1.V * 32);
        conv(g1, 16, 32, d1.data(), W[1].data(), B[1].data(), a1.data());
        act(a1, true);
        d2.resize(g2.V * 32);
        inject(g1, g2, 32, a1.data(), d2.data());
        a2.resize(g2.V * 64);
        conv(g2, 32, 64, d2.data(), W[2].data(), B[2].data(), a2.data());
        act(a2, true);
        b2.resize(g2.V * 32);
        conv(g2, 64, 32, a2.data(), W[3].data(), B[3].data(), b2.data());
        act(b2, false);
        u1.resize(g1.V * 32);
        prolong(g1, g2, 32, b2.data(), u1.data());
        for (int64_t q = 0; q < g1.V * 32; q++) u1[q] += a1[q];
        round(u1);
        b1.resize(g1.V * 16);
        conv(g1, 32, 16, u1.data(), W[4].data(), B[4].data(), b1.data());
        act(b1, true);
        u0.resize(g0.V * 16);
        prolong(g0, g1, 16, b1.data(), u0.data());
        for (int64_t q = 0; q < g0.V * 16; q++) u0[q] += a0[q];
        round(u0);
        y.resize(g0.V * 16);
        conv(g0, 16, 16, u0.data(), W[5].data(), B[5].data(), y.data());
    }

    // gradient fed to a convolution's backward: g * (a > 0) (mask optional), then rounded
    std::vector<double> gin(const std::vector<double>& g, const std::vector<double>* a) const {
        std::vector<double> o(g);
        if (a)
            for (size_t q = 0; q < o.size(); q++)
                if (!((*a)[q] > 0)) o[q] = 0.0;
        std::vector<double> t(o);
        round(t);
        return t;
    }

    void backward(const float* dyin, double* grads, double* dx) {
        std::vector<double> dW[6], dB[6];
        for (int l = 0; l < 6; l++) {
            dW[l].assign(W[l].size(), 0.0);
            dB[l].assign(B[l].size(), 0.0);
        }
        std::vector<double> dy(dyin, dyin + g0.V * 16);
        auto g5 = gin(dy, nullptr);
        conv_wgrad(g0, 16, 16, u0.data(), g5.data(), dW[5].data(), dB[5].data());
        std::vector<double> du0(g0.V * 16);
        conv_T(g0, 16, 16, g5.data(), W[5].data(), du0.data());
        std::vector<double> db1(g1.V * 16);
        prolong_T(g0, g1, 16, du0.data(), db1.data());
        std::vector<double> da0(du0);
        auto g4 = gin(db1, &b1);
        conv_wgrad(g1, 32, 16, u1.data(), g4.data(), dW[4].data(), dB[4].data());
        std::vector<double> du1(g1.V * 32);
        conv_T(g1, 32, 16, g4.data(), W[4].data(), du1.data());
        std::vector<double> db2(g2.V * 32);
        prolong_T(g1, g2, 32, du1.data(), db2.data());
        std::vector<double> da1(du1);
        auto g3 = gin(db2, nullptr);
        conv_wgrad(g2, 64, 32, a2.data(), g3.data(), dW[3].data(), dB[3].data());
        std::vector<double> da2(g2.V * 64);
        conv_T(g2, 64, 32, g3.data(), W[3].data(), da2.data());
        auto g2g = gin(da2, &a2);
        conv_wgrad(g2, 32, 64, d2.data(), g2g.data(), dW[2].data(), dB[2].data());
        std::vector<double> dd2(g2.V * 32);
        conv_T(g2, 32, 64, g2g.data(), W[2].data(), dd2.data());
        inject_T(g1, g2, 32, dd2.data(), da1.data());
        auto g1g = gin(da1, &a1);
        conv_wgrad(g1, 16, 32, d1.data(), g1g.data(), dW[1].data(), dB[1].data());
        std::vector<double> dd1(g1.V * 16);
        conv_T(g1, 16, 32, g1g.data(), W[1].data(), dd1.data());
        inject_T(g0, g1, 16, dd1.data(), da0.data());
        auto g0g = gin(da0, &a0);
        conv_wgrad(g0, 16, 16, x.data(), g0g.data(), dW[0].data(), dB[0].data());
        if (dx) conv_T(g0, 16, 16, g0g.data(), W[0].data(), dx);
        int64_t off = 0;
        for (int l = 0; l < 6; l++) {
            for (size_t q = 0; q < dW[l].size(); q++) grads[off + q] = dW[l][q];
            off += dW[l].size();
            for (size_t q = 0; q < dB[l].size(); q++) grads[off + q] = dB[l][q];
            off += dB[l].size();
        }
    }
};

}  // namespace ocv

extern "C" {
int64_t ocv_nparam() { return ocv::nparam(); }
double ocv_bf16(double x) { return ocv::bf(x); }
void ocv_conv3(const int* N, int cin, int cout, const double* in, const double* W, const double* bias, double* out) {
    ocv::conv(ocv::grid(N), cin, cout, in, W, bias, out);
}
void ocv_conv3_T(const int* N, int cin, int cout, const double* dout, const double* W, double* din) {
    ocv::conv_T(ocv::grid(N), cin, cout, dout, W, din);
}
void ocv_conv3_wgrad(const int* N, int cin, int cout, const double* in, const double* dout, double* dW, double* dB) {
    ocv::conv_wgrad(ocv::grid(N), cin, cout, in, dout, dW, dB);
}
void ocv_prolong(const int* Nf, int C, const double* c, double* out) {
    const ocv::G gf = ocv::grid(Nf);
    ocv::prolong(gf, ocv::coarse(gf), C, c, out);
}
void ocv_prolong_T(const int* Nf, int C, const double* gu, double* gc) {
    const ocv::G gf = ocv::grid(Nf);
    ocv::prolong_T(gf, ocv::coarse(gf), C, gu, gc);
}
void ocv_inject(const int* Nf, int C, const double* a, double* d) {
    const ocv::G gf = ocv::grid(Nf);
    ocv::inject(gf, ocv::coarse(gf), C, a, d);
}
void ocv_inject_T(const int* Nf, int C, const double* gd, double* ga) {
    const ocv::G gf = ocv::grid(Nf);
    ocv::inject_T(gf, ocv::coarse(gf), C, gd, ga);
}
// y [V][16]
void ocv_net_forward(const int* N0, const float* params, const float* x, double* y, int bf16) {
    ocv::Net n(N0, params, bf16 != 0);
    n.forward(x);
    std::memcpy(y, n.y.data(), n.y.size() * sizeof(double));
}
// grads [nparam] (overwritten), dx [V][16] (may be null)
void ocv_net_backward(const int* N0, const float* params, const float* x, const float* dy, double* grads, double* dx,
                      int bf16) {
    ocv::Net n(N0, params, bf16 != 0);
    n.forward(x);
    n.backward(dy, grads, dx);
}
}
