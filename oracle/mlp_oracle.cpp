// mlp_oracle.cpp -- CPU restatement of the learned-closure network.
//
// TEST INFRASTRUCTURE ONLY (never linked into or called by the product
// library).  The reference has no subgrid model; this file states the
// network the GPU kernels evaluate and train, in FP64:
//
//   x (8 inputs) -> h1 = relu(W1 x + b1)   64
//                -> h2 = relu(W2 h1 + b2)  64
//                -> h3 = relu(W3 h2 + b3)  64
//                -> y  = W4 . h3 + b4      1
//   loss = (1/n) sum_c w_c (y_c - t_c)^2   (w_c = 1 unless weights are given)
//
// Parameters are one float array in this order: W1[64][8] (row = output
// unit), b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4 -- 8961
// values.
//
// With `quant` set, the quantization points of the BF16 matrix-core kernels
// are reproduced (round to nearest even): W1, W2, W3, the inputs and the
// hidden activations h1, h2 that feed a matrix product, and the backward
// deltas dz3, dz2 that feed the transposed products and every weight-
// gradient product (dz1..dz3 and their right-hand activations x, h1, h2).
// Everything else (biases, W4, h3, bias gradients, dW4, accumulation) is
// unrounded.  Without `quant` the same function is the plain FP64 network,
// which the finite-difference tests pin.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 8, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H, OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H,
              NPARAM = OFF_B4 + 1;
static_assert(NPARAM == 8961, "parameter count");

double bf16_round(double v) {
    const float f = float(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;  // inf / nan pass through
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return double(r);
}

struct Net {
    const double *p;
    bool quant;
    std::vector<double> w1, w2, w3;  // possibly rounded copies
    Net(const double *params, bool q) : p(params), quant(q) {
        auto cp = [&](int off, int n, std::vector<double> &dst) {
            dst.assign(p + off, p + off + n);
            if (quant)
                for (double &v : dst) v = bf16_round(v);
        };
        cp(OFF_W1, H * IN, w1);
        cp(OFF_W2, H * H, w2);
        cp(OFF_W3, H * H, w3);
    }
    double q(double v) const { return quant ? bf16_round(v) : v; }

    struct Act {
        double x[IN], z1[H], h1[H], z2[H], h2[H], z3[H], h3[H], y;
    };
    void forward(const double *xin, Act &a) const {
        for (int j = 0; j < IN; ++j) a.x[j] = q(xin[j]);
        for (int u = 0; u < H; ++u) {
            double s = p[OFF_B1 + u];
            for (int j = 0; j < IN; ++j) s += w1[u * IN + j] * a.x[j];
            a.z1[u] = s;
            a.h1[u] = q(s > 0 ? s : 0.0);
        }
        for (int u = 0; u < H; ++u) {
            double s = p[OFF_B2 + u];
            for (int j = 0; j < H; ++j) s += w2[u * H + j] * a.h1[j];
            a.z2[u] = s;
            a.h2[u] = q(s > 0 ? s : 0.0);
        }
        double y = p[OFF_B4];
        for (int u = 0; u < H; ++u) {
            double s = p[OFF_B3 + u];
            for (int j = 0; j < H; ++j) s += w3[u * H + j] * a.h2[j];
            a.z3[u] = s;
            a.h3[u] = s > 0 ? s : 0.0;
            y += p[OFF_W4 + u] * a.h3[u];
        }
        a.y = y;
    }
    // Adds dL/dparams of one sample with dL/dy = dy into g.
    void backward(const Act &a, double dy, double *g) const {
        double dz3[H], dz2[H], dz1[H];
        for (int u = 0; u < H; ++u) {
            g[OFF_W4 + u] += dy * a.h3[u];
            dz3[u] = a.z3[u] > 0 ? dy * p[OFF_W4 + u] : 0.0;
        }
        g[OFF_B4] += dy;
        // layer 3: bias gradient from the delta that feeds the products
        for (int u = 0; u < H; ++u) {
            const double d = q(dz3[u]);
            g[OFF_B3 + u] += d;
            for (int j = 0; j < H; ++j) g[OFF_W3 + u * H + j] += d * a.h2[j];
        }
        for (int j = 0; j < H; ++j) {
            double s = 0;
            for (int u = 0; u < H; ++u) s += w3[u * H + j] * q(dz3[u]);
            dz2[j] = a.z2[j] > 0 ? s : 0.0;
        }
        for (int u = 0; u < H; ++u) {
            const double d = q(dz2[u]);
            g[OFF_B2 + u] += d;
            for (int j = 0; j < H; ++j) g[OFF_W2 + u * H + j] += d * a.h1[j];
        }
        for (int j = 0; j < H; ++j) {
            double s = 0;
            for (int u = 0; u < H; ++u) s += w2[u * H + j] * q(dz2[u]);
            dz1[j] = a.z1[j] > 0 ? s : 0.0;
        }
        for (int u = 0; u < H; ++u) {
            const double d = q(dz1[u]);
            g[OFF_B1 + u] += d;
            for (int j = 0; j < IN; ++j) g[OFF_W1 + u * IN + j] += d * a.x[j];
        }
    }
};

double bf16_bits_to_double(uint16_t b) {
    const uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return double(f);
}

}  // namespace

extern "C" {

int oracle_mlp_params(void) { return NPARAM; }

// y[c] for n samples; x is n x 8 doubles.
void oracle_mlp_forward(const double *params, const double *x, int64_t n, int quant, double *y) {
    Net net(params, quant != 0);
#pragma omp parallel for
    for (int64_t c = 0; c < n; ++c) {
        Net::Act a;
        net.forward(x + 8 * c, a);
        y[c] = a.y;
    }
}

// Gradient of the weighted mean squared error (1/n) sum w (y - t)^2 over n
// samples (w = NULL: weights 1; grad has NPARAM entries) and the loss itself.
void oracle_mlp_grad(const double *params, const double *x, const double *t, const double *w, int64_t n, int quant,
                     double *grad, double *loss) {
    Net net(params, quant != 0);
    std::memset(grad, 0, sizeof(double) * NPARAM);
    double l = 0;
    for (int64_t c = 0; c < n; ++c) {
        Net::Act a;
        net.forward(x + 8 * c, a);
        const double e = a.y - t[c], we = w ? w[c] * e : e;
        l += we * e;
        net.backward(a, 2.0 * we / double(n), grad);
    }
    *loss = l / double(n);
}

// bf16 bit patterns -> doubles (n values).
void oracle_bf16_decode(const uint16_t *b, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = bf16_bits_to_double(b[i]);
}

// Adam step in double, the update rule the GPU kernel follows (Kingma & Ba
// 2015, bias-corrected moments).
void oracle_adam(double *params, const double *grad, double *m, double *v, int64_t n, int64_t step, double lr,
                 double b1, double b2, double eps) {
    const double c1 = 1.0 - std::pow(b1, double(step)), c2 = 1.0 - std::pow(b2, double(step));
    for (int64_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1 - b1) * grad[i];
        v[i] = b2 * v[i] + (1 - b2) * grad[i] * grad[i];
        params[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
    }
}

}  // extern "C"
