// mlp.cpp — TEST INFRASTRUCTURE, not product code.
//
// CPU restatement of the transport surrogate's network (no counterpart in
// the reference): a 3 x 64 SiLU MLP, forward pass and the gradient of the
// mean squared loss by backpropagation (Rumelhart, Hinton & Williams 1986),
// all in double.  With emulate != 0 the operands of every matrix product are
// rounded to bfloat16 (round to nearest even, from the float value) exactly
// where the GPU kernels round them: weights, the network inputs, the
// activations after SiLU, the output error and the pre-activation errors of
// the backward pass.  With emulate == 0 it is plain double arithmetic, which
// the tests pin with central finite differences.
//
// The forward pass of inference (orc_mlp_forward) follows the GPU's
// inference operands: the factor -log2(e) of the sigmoid's exponent is folded
// into the layer-1 weights and the biases of layers 1-3, so that the
// pre-activations are u = -log2(e) z and the activations carry the same
// factor, h' = u / (1 + 2^u) = -log2(e) SiLU(z); the output weights carry
// -ln(2) to undo it.  The scaled parameters are float products, as on the
// GPU.  emulate = 2 folds without rounding (equal to emulate = 0 up to the
// float scaling of the parameters).
//
// Parameters (float, flat): W1[64][nin] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[nout][64] b4[nout].  Samples: rows of 16 floats (zero padded).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int H = 64, IO = 16;

double bf16(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return f;  // inf / nan unchanged
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

double silu(double z) { return z / (1.0 + std::exp(-z)); }
double dsilu(double z) {
  const double s = 1.0 / (1.0 + std::exp(-z));
  return s * (1.0 + z * (1.0 - s));
}

constexpr float kFold = -1.4426950408889634f;    // -log2(e), as the GPU's float
constexpr float kUnfold = -0.69314718055994531f;  // -ln(2)

struct Net {
  int nin, nout;
  bool folded = false;
  std::vector<double> W1, b1, W2, b2, W3, b3, W4, b4;  // rounded if emulating
  int o_w1, o_b1, o_w2, o_b2, o_w3, o_b3, o_w4, o_b4, count;
  Net(const float *p, int nin_, int nout_, bool em, bool fold = false) : nin(nin_), nout(nout_), folded(fold) {
    o_w1 = 0;
    o_b1 = H * nin;
    o_w2 = o_b1 + H;
    o_b2 = o_w2 + H * H;
    o_w3 = o_b2 + H;
    o_b3 = o_w3 + H * H;
    o_w4 = o_b3 + H;
    o_b4 = o_w4 + nout * H;
    count = o_b4 + nout;
    auto take = [&](int off, int n, bool round, float scale) {
      std::vector<double> v(n);
      for (int i = 0; i < n; ++i) {
        const float f = scale == 1.0f ? p[off + i] : p[off + i] * scale;
        v[i] = round ? bf16(f) : (double)f;
      }
      return v;
    };
    const float s = fold ? kFold : 1.0f;
    W1 = take(o_w1, H * nin, em, s);
    b1 = take(o_b1, H, false, s);
    W2 = take(o_w2, H * H, em, 1.0f);
    b2 = take(o_b2, H, false, s);
    W3 = take(o_w3, H * H, em, 1.0f);
    b3 = take(o_b3, H, false, s);
    W4 = take(o_w4, nout * H, em, fold ? kUnfold : 1.0f);
    b4 = take(o_b4, nout, false, 1.0f);
  }
  // activation of pre-activation s: SiLU(s), or u / (1 + 2^u) when folded
  double act(double v) const { return folded ? v / (1.0 + std::exp2(v)) : silu(v); }
};

struct Acts {
  double h0[IO], z1[H], h1[H], z2[H], h2[H], z3[H], h3[H], y[IO];
};

void forward(const Net &n, const float *x, bool em, Acts &a) {
  auto r = [em](double v) { return em ? bf16(v) : v; };
  for (int i = 0; i < n.nin; ++i) a.h0[i] = r(x[i]);
  for (int o = 0; o < H; ++o) {
    double s = n.b1[o];
    for (int i = 0; i < n.nin; ++i) s += n.W1[o * n.nin + i] * a.h0[i];
    a.z1[o] = s;
    a.h1[o] = r(n.act(s));
  }
  for (int o = 0; o < H; ++o) {
    double s = n.b2[o];
    for (int i = 0; i < H; ++i) s += n.W2[o * H + i] * a.h1[i];
    a.z2[o] = s;
    a.h2[o] = r(n.act(s));
  }
  for (int o = 0; o < H; ++o) {
    double s = n.b3[o];
    for (int i = 0; i < H; ++i) s += n.W3[o * H + i] * a.h2[i];
    a.z3[o] = s;
    a.h3[o] = r(n.act(s));
  }
  for (int o = 0; o < n.nout; ++o) {
    double s = n.b4[o];
    for (int i = 0; i < H; ++i) s += n.W4[o * H + i] * a.h3[i];
    a.y[o] = s;
  }
}

}  // namespace

extern "C" {

// y[n][16]: network outputs (zero beyond nout); emulate: 0 exact, 1 the GPU's
// inference (folded, BF16 operands), 2 folded without rounding
int orc_mlp_forward(const float *params, int nin, int nout, const float *x, int n, int emulate, double *y) {
  if (nin < 1 || nin > IO || nout < 1 || nout > IO || n < 0) return -1;
  if (emulate < 0 || emulate > 2) return -1;
  const bool em = emulate == 1;
  const Net net(params, nin, nout, em, emulate != 0);
  for (int s = 0; s < n; ++s) {
    Acts a;
    forward(net, x + (size_t)s * IO, em, a);
    for (int o = 0; o < IO; ++o) y[(size_t)s * IO + o] = o < nout ? a.y[o] : 0.0;
  }
  return 0;
}

// loss = sum over samples and outputs (y - t)^2 / (n nout); grad in the
// parameter layout.
int orc_mlp_loss_grad(const float *params, int nin, int nout, const float *x, const float *t, int n, int emulate,
                      double *grad, double *loss) {
  if (nin < 1 || nin > IO || nout < 1 || nout > IO || n <= 0) return -1;
  const bool em = emulate != 0;
  auto r = [em](double v) { return em ? bf16(v) : v; };
  const Net net(params, nin, nout, em);
  std::vector<double> g(net.count, 0.0);
  const double scale = 1.0 / ((double)n * nout);
  double L = 0.0;
  for (int s = 0; s < n; ++s) {
    Acts a;
    forward(net, x + (size_t)s * IO, em, a);
    double dy[IO], dz3[H], dz2[H], dz1[H];
    for (int o = 0; o < nout; ++o) {
      const double d = a.y[o] - (double)t[(size_t)s * IO + o];
      L += scale * d * d;
      dy[o] = r(2.0 * scale * d);
    }
    for (int o = 0; o < nout; ++o) {
      g[net.o_b4 + o] += dy[o];
      for (int i = 0; i < H; ++i) g[net.o_w4 + o * H + i] += dy[o] * a.h3[i];
    }
    for (int i = 0; i < H; ++i) {
      double s3 = 0.0;
      for (int o = 0; o < nout; ++o) s3 += net.W4[o * H + i] * dy[o];
      dz3[i] = r(s3 * dsilu(a.z3[i]));
    }
    for (int o = 0; o < H; ++o) {
      g[net.o_b3 + o] += dz3[o];
      for (int i = 0; i < H; ++i) g[net.o_w3 + o * H + i] += dz3[o] * a.h2[i];
    }
    for (int i = 0; i < H; ++i) {
      double s2 = 0.0;
      for (int o = 0; o < H; ++o) s2 += net.W3[o * H + i] * dz3[o];
      dz2[i] = r(s2 * dsilu(a.z2[i]));
    }
    for (int o = 0; o < H; ++o) {
      g[net.o_b2 + o] += dz2[o];
      for (int i = 0; i < H; ++i) g[net.o_w2 + o * H + i] += dz2[o] * a.h1[i];
    }
    for (int i = 0; i < H; ++i) {
      double s1 = 0.0;
      for (int o = 0; o < H; ++o) s1 += net.W2[o * H + i] * dz2[o];
      dz1[i] = r(s1 * dsilu(a.z1[i]));
    }
    for (int o = 0; o < H; ++o) {
      g[net.o_b1 + o] += dz1[o];
      for (int i = 0; i < nin; ++i) g[net.o_w1 + o * nin + i] += dz1[o] * a.h0[i];
    }
  }
  for (int i = 0; i < net.count; ++i) grad[i] = g[i];
  if (loss) *loss = L;
  return 0;
}

}  // extern "C"
