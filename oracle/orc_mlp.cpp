// oracle/orc_mlp.cpp -- CPU restatement of the warm-start network (TEST INFRASTRUCTURE ONLY: the
// checker of the HIP network kernels, never linked into the product).
//
// The network: 24 -> 128 -> 128 -> 3, GELU in its tanh form 0.5 x (1 + tanh(c (x + a x^3))),
// c = sqrt(2/pi), a = 0.044715; loss = mean over rows of the squared error summed over the three
// outputs, outputs compared with targets / y_scale. Parameters in the layout of the C ABI
// (W1[24][128] b1 W2[128][128] b2 W3[128][3] b3 x_mean[24] x_scale[24] y_scale[3]), in double.
//
// emulate_bf16 = 1 rounds to BF16 exactly where the GPU kernels do (the normalised inputs, the
// weights, both hidden activations, dL/dY, dL/dZ2 and dL/dZ1) and keeps every sum in double;
// emulate_bf16 = 0 is the exact function in double, which the finite-difference checks use.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int F = 24, H = 128, O = 3;
constexpr int OW1 = 0, OB1 = OW1 + F * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * O;
constexpr int NT = OB3 + O, OXM = NT, OXS = OXM + F, OYS = OXS + F;

double bf(double v, int emulate) {
  if (!emulate) return v;
  float f = float(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return double(f);
}

const double kC = 0.7978845608028654, kA = 0.044715;
double gelu(double x) { return 0.5 * x * (1.0 + std::tanh(kC * (x + kA * x * x * x))); }
double gelu_grad(double x) {
  const double t = std::tanh(kC * (x + kA * x * x * x));
  return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * kC * (1.0 + 3.0 * kA * x * x);
}

struct Act {
  double a0[F], z1[H], h1[H], z2[H], h2[H], y[O];
};

void forward_row(const double* p, const float* x, int em, Act& a) {
  for (int k = 0; k < F; ++k) {
    // the GPU normalises in FP32: (x - mean) / scale
    const float v = (x[k] - float(p[OXM + k])) / float(p[OXS + k]);
    a.a0[k] = bf(v, em);
  }
  for (int n = 0; n < H; ++n) {
    double s = 0.0;
    for (int k = 0; k < F; ++k) s += a.a0[k] * bf(p[OW1 + k * H + n], em);
    a.z1[n] = s + p[OB1 + n];
    a.h1[n] = bf(gelu(a.z1[n]), em);
  }
  for (int n = 0; n < H; ++n) {
    double s = 0.0;
    for (int k = 0; k < H; ++k) s += a.h1[k] * bf(p[OW2 + k * H + n], em);
    a.z2[n] = s + p[OB2 + n];
    a.h2[n] = bf(gelu(a.z2[n]), em);
  }
  for (int j = 0; j < O; ++j) {
    double s = 0.0;
    for (int k = 0; k < H; ++k) s += a.h2[k] * bf(p[OW3 + k * O + j], em);
    a.y[j] = s + p[OB3 + j];
  }
}

}  // namespace

extern "C" {

int orc_mlp_num_params() { return OYS + O; }

// out[n][3]: predicted impulses (outputs times y_scale).
int orc_mlp_forward(const double* p, const float* X, int64_t n, int32_t emulate_bf16, double* out) {
  Act a;
  for (int64_t r = 0; r < n; ++r) {
    forward_row(p, X + r * F, emulate_bf16, a);
    for (int j = 0; j < O; ++j) out[r * O + j] = a.y[j] * p[OYS + j];
  }
  return 0;
}

// grad[20099] of the batch loss; *loss = the loss.
int orc_mlp_gradients(const double* p, const float* X, const float* T, int64_t n, int32_t emulate_bf16,
                      double* grad, double* loss) {
  const int em = emulate_bf16;
  std::memset(grad, 0, sizeof(double) * NT);
  double L = 0.0;
  Act a;
  double d3[O], dz2[H], dz1[H];
  for (int64_t r = 0; r < n; ++r) {
    forward_row(p, X + r * F, em, a);
    for (int j = 0; j < O; ++j) {
      const double e = a.y[j] - double(float(T[r * O + j]) / float(p[OYS + j]));
      L += e * e;
      d3[j] = bf(2.0 * e / double(n), em);
    }
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int j = 0; j < O; ++j) s += d3[j] * bf(p[OW3 + k * O + j], em);
      dz2[k] = bf(s * gelu_grad(a.z2[k]), em);
    }
    for (int k = 0; k < H; ++k) {
      double s = 0.0;
      for (int j = 0; j < H; ++j) s += dz2[j] * bf(p[OW2 + k * H + j], em);
      dz1[k] = bf(s * gelu_grad(a.z1[k]), em);
    }
    for (int k = 0; k < H; ++k)
      for (int j = 0; j < O; ++j) grad[OW3 + k * O + j] += a.h2[k] * d3[j];
    for (int j = 0; j < O; ++j) grad[OB3 + j] += d3[j];
    for (int k = 0; k < H; ++k)
      for (int j = 0; j < H; ++j) grad[OW2 + k * H + j] += a.h1[k] * dz2[j];
    for (int j = 0; j < H; ++j) grad[OB2 + j] += dz2[j];
    for (int k = 0; k < F; ++k)
      for (int j = 0; j < H; ++j) grad[OW1 + k * H + j] += a.a0[k] * dz1[j];
    for (int j = 0; j < H; ++j) grad[OB1 + j] += dz1[j];
  }
  *loss = L / double(n);
  return 0;
}

}  // extern "C"
