// knc_mlp_oracle.cpp -- FP64 CPU reference of the learned c2p-guess MLP,
// used ONLY as test infrastructure (gradient and forward checks of the HIP
// training and inference kernels).
//
// Network: 4 -> 64 -> 64 -> 1, tanh hidden activations, linear output.
// Parameter vector (4545 values): W1[64][4] | b1[64] | W2[64][64] | b2[64] |
// W3[64] | b3.  Loss: mean squared error over the batch.  Gradients by
// reverse-mode differentiation written out layer by layer.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 4, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;

struct Acts {
    double h1[H], h2[H], y;
};

void forward(const double* P, const double* x, Acts& a)
{
    for (int i = 0; i < H; ++i) {
        double z = P[OFF_B1 + i];
        for (int m = 0; m < IN; ++m) z += P[OFF_W1 + i * IN + m] * x[m];
        a.h1[i] = std::tanh(z);
    }
    for (int k = 0; k < H; ++k) {
        double z = P[OFF_B2 + k];
        for (int i = 0; i < H; ++i) z += P[OFF_W2 + k * H + i] * a.h1[i];
        a.h2[k] = std::tanh(z);
    }
    double y = P[OFF_B3];
    for (int k = 0; k < H; ++k) y += P[OFF_W3 + k] * a.h2[k];
    a.y = y;
}

}  // namespace

extern "C" {

int ko_mlp_num_params(void) { return NPARAM; }

int ko_mlp_forward(const double* params, int64_t n, const double* x, double* y)
{
    Acts a;
    for (int64_t s = 0; s < n; ++s) {
        forward(params, x + 4 * s, a);
        y[s] = a.y;
    }
    return 0;
}

// loss = (1/n) sum (y - t)^2 and its gradient with respect to every parameter.
int ko_mlp_loss_grad(const double* params, int64_t n, const double* x, const double* t, double* loss, double* grad)
{
    std::memset(grad, 0, sizeof(double) * NPARAM);
    double L = 0.0;
    Acts a;
    double d2[H];
    for (int64_t s = 0; s < n; ++s) {
        const double* xs = x + 4 * s;
        forward(params, xs, a);
        const double e = a.y - t[s];
        L += e * e;
        const double dy = 2.0 * e / (double)n;
        grad[OFF_B3] += dy;
        for (int k = 0; k < H; ++k) {
            grad[OFF_W3 + k] += dy * a.h2[k];
            d2[k] = dy * params[OFF_W3 + k] * (1.0 - a.h2[k] * a.h2[k]);
            grad[OFF_B2 + k] += d2[k];
            for (int i = 0; i < H; ++i) grad[OFF_W2 + k * H + i] += d2[k] * a.h1[i];
        }
        for (int i = 0; i < H; ++i) {
            double d1 = 0.0;
            for (int k = 0; k < H; ++k) d1 += d2[k] * params[OFF_W2 + k * H + i];
            d1 *= 1.0 - a.h1[i] * a.h1[i];
            grad[OFF_B1 + i] += d1;
            for (int m = 0; m < IN; ++m) grad[OFF_W1 + i * IN + m] += d1 * xs[m];
        }
    }
    *loss = L / (double)n;
    return 0;
}

}  // extern "C"
