// CPU restatement of the learned warm start's network (test infrastructure only: the checker of
// the GPU forward, backward and training kernels; nothing in the product calls it).
//
// The reference has no learned warm start; this network is defined by the product's C ABI
// (AGX_MLP_* in include/agx.h): 24 -> 128 -> 128 -> 2, ReLU after the two hidden layers,
// loss (1/n) sum_rows sum_outputs (y - t)^2. Parameters in host layout: W1[128][24], b1[128],
// W2[128][128], b2[128], W3[2][128], b3[2].
//
// Two modes:
//   rounding = 0: exact arithmetic in double (the analytic gradient that finite differences pin);
//   rounding = 1: the GPU's rounding points. Features and weight matrices are rounded to bf16
//     (round to nearest even); every sum is exact and then rounded to fp32 once; hidden
//     activations are rounded to bf16 after bias and ReLU; in the backward pass dZ2 and dZ1 are
//     rounded to bf16 before they enter the weight-gradient and dH1 products, while the bias
//     gradients sum the fp32 values.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 24, H = 128, OUT = 2;
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
                 OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + OUT * H, PARAMS = OFF_B3 + OUT;
static_assert(PARAMS == 19970, "host parameter count");

double bf16(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return f;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}
double f32(double v) { return (double)(float)v; }

struct Net {
    bool round;
    std::vector<double> w1, b1, w2, b2, w3, b3;   // as used by the products (bf16 when rounding)
    Net(const float* p, bool r) : round(r) {
        auto take = [&](size_t off, size_t n, bool weight) {
            std::vector<double> v(n);
            for (size_t k = 0; k < n; k++) v[k] = (r && weight) ? bf16(p[off + k]) : (double)p[off + k];
            return v;
        };
        w1 = take(OFF_W1, H * IN, true); b1 = take(OFF_B1, H, false);
        w2 = take(OFF_W2, H * H, true); b2 = take(OFF_B2, H, false);
        w3 = take(OFF_W3, OUT * H, true); b3 = take(OFF_B3, OUT, false);
    }
    double act(double v) const { return round ? bf16(v) : v; }    // hidden activation storage
    double acc(double v) const { return round ? f32(v) : v; }     // an fp32 accumulator's result
};

// One row: x (24) -> h1, h2 (128 each, stored form) and y (2).
void forwardRow(const Net& n, const float* xr, double* xs, double* h1, double* h2, double* y) {
    for (int k = 0; k < IN; k++) xs[k] = n.round ? bf16(xr[k]) : (double)xr[k];
    for (int o = 0; o < H; o++) {
        double s = 0.0;
        for (int k = 0; k < IN; k++) s += n.w1[o * IN + k] * xs[k];
        const double z = n.acc(n.acc(s) + n.b1[o]);
        h1[o] = n.act(z > 0.0 ? z : 0.0);
    }
    for (int o = 0; o < H; o++) {
        double s = 0.0;
        for (int k = 0; k < H; k++) s += n.w2[o * H + k] * h1[k];
        const double z = n.acc(n.acc(s) + n.b2[o]);
        h2[o] = n.act(z > 0.0 ? z : 0.0);
    }
    for (int o = 0; o < OUT; o++) {
        double s = 0.0;
        for (int k = 0; k < H; k++) s += n.w3[o * H + k] * h2[k];
        y[o] = n.acc(n.acc(s) + n.b3[o]);
    }
}

}  // namespace

extern "C" {

void orc_mlp_forward(const float* params, const float* x, uint64_t rows, float* y, int rounding) {
    const Net n(params, rounding != 0);
    double xs[IN], h1[H], h2[H], yr[OUT];
    for (uint64_t r = 0; r < rows; r++) {
        forwardRow(n, x + r * IN, xs, h1, h2, yr);
        for (int o = 0; o < OUT; o++) y[r * OUT + o] = (float)yr[o];
    }
}

// Loss of one batch and its gradient (host parameter layout) into grad.
double orc_mlp_gradients(const float* params, const float* x, const float* t, uint64_t rows, float* grad, int rounding) {
    const Net n(params, rounding != 0);
    const bool rd = rounding != 0;
    std::vector<double> xs(rows * IN), h1(rows * H), h2(rows * H), dz2(rows * H), dz1(rows * H);
    std::vector<double> gb3(OUT, 0.0), gb2(H, 0.0), gb1(H, 0.0);
    std::vector<double> gw3(OUT * H, 0.0), gw2(H * H, 0.0), gw1(H * IN, 0.0);
    double loss = 0.0;
    const double inv = 2.0 / (double)rows;
    for (uint64_t r = 0; r < rows; r++) {
        double y[OUT];
        forwardRow(n, x + r * IN, &xs[r * IN], &h1[r * H], &h2[r * H], y);
        double dy[OUT];
        for (int o = 0; o < OUT; o++) {
            const double diff = rd ? f32(y[o] - (double)t[r * OUT + o]) : y[o] - (double)t[r * OUT + o];
            loss += diff * diff / (double)rows;
            dy[o] = rd ? f32(f32(inv) * diff) : inv * diff;
            gb3[o] += dy[o];
        }
        for (int j = 0; j < H; j++) {
            const double hv = h2[r * H + j];
            double dh = 0.0;
            for (int o = 0; o < OUT; o++) {
                gw3[o * H + j] += dy[o] * hv;
                dh += dy[o] * n.w3[o * H + j];
            }
            const double dz = hv > 0.0 ? n.acc(dh) : 0.0;
            gb2[j] += dz;
            dz2[r * H + j] = rd ? bf16(dz) : dz;
        }
    }
    // dW2[i][j] = sum_r dZ2[r][i] H1[r][j];  dH1[r][j] = sum_i dZ2[r][i] W2[i][j]
    for (uint64_t r = 0; r < rows; r++) {
        const double* a = &dz2[r * H];
        const double* b = &h1[r * H];
        for (int i = 0; i < H; i++) {
            if (a[i] == 0.0) continue;
            for (int j = 0; j < H; j++) gw2[i * H + j] += a[i] * b[j];
        }
        for (int j = 0; j < H; j++) {
            double s = 0.0;
            for (int i = 0; i < H; i++) s += a[i] * n.w2[i * H + j];
            const double dh = n.acc(s);
            const double dz = h1[r * H + j] > 0.0 ? dh : 0.0;
            gb1[j] += dz;
            dz1[r * H + j] = rd ? bf16(dz) : dz;
        }
    }
    for (uint64_t r = 0; r < rows; r++)
        for (int i = 0; i < H; i++) {
            const double a = dz1[r * H + i];
            if (a == 0.0) continue;
            for (int k = 0; k < IN; k++) gw1[i * IN + k] += a * xs[r * IN + k];
        }
    auto put = [&](size_t off, const std::vector<double>& v) {
        for (size_t k = 0; k < v.size(); k++) grad[off + k] = (float)v[k];
    };
    put(OFF_W1, gw1); put(OFF_B1, gb1); put(OFF_W2, gw2); put(OFF_B2, gb2); put(OFF_W3, gw3); put(OFF_B3, gb3);
    return loss;
}

// Loss only, in double without rounding (the finite-difference probe of the exact mode).
double orc_mlp_loss(const float* params, const double* delta, const float* x, const float* t, uint64_t rows) {
    std::vector<float> p(params, params + PARAMS);
    // delta is applied in double: the probe needs steps below fp32 resolution of large weights
    Net n(p.data(), false);
    if (delta) {
        size_t k = 0;
        for (double& v : n.w1) v += delta[OFF_W1 + k++];
        k = 0; for (double& v : n.b1) v += delta[OFF_B1 + k++];
        k = 0; for (double& v : n.w2) v += delta[OFF_W2 + k++];
        k = 0; for (double& v : n.b2) v += delta[OFF_B2 + k++];
        k = 0; for (double& v : n.w3) v += delta[OFF_W3 + k++];
        k = 0; for (double& v : n.b3) v += delta[OFF_B3 + k++];
    }
    double xs[IN], h1[H], h2[H], y[OUT], loss = 0.0;
    for (uint64_t r = 0; r < rows; r++) {
        forwardRow(n, x + r * IN, xs, h1, h2, y);
        for (int o = 0; o < OUT; o++) {
            const double d = y[o] - (double)t[r * OUT + o];
            loss += d * d / (double)rows;
        }
    }
    return loss;
}

}  // extern "C"
