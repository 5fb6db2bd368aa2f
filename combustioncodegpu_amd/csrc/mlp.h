// mlp.h — the learned transport surrogate: a 3 x 64 SiLU MLP mapping
// (T, log p, Y_1..Y_ns) to (log mu, log kappa, log rho*D_1..ns), normalised
// on both sides, evaluated with BF16 MFMA (v_mfma_f32_16x16x32_bf16) and
// FP32 accumulation.  No counterpart in the reference, whose transport is
// Transport.cu:43-96; that routine generates the training samples and is
// the fallback.
//
// Layout.  A wavefront evaluates 16 points per MFMA in the transposed form
// Z^T[out x pts] = W[out x in] H^T[in x pts]: W is the A operand (rows = output
// features, 16 per M-block), H^T the B operand (columns = the 16 points).  The
// accumulator of an M-block holds, in lane L, rows 4(L/16)+r (r = 0..3) of
// point L%16.  The 32 k-slots of an MFMA are assigned so that those four
// accumulator rows of M-blocks 2kb and 2kb+1 are exactly the eight B values
// lane L supplies for K-block kb of the next layer:
//     slot j of lane group g = L/16  ->  feature 32 kb + kperm(g, j),
//     kperm(g, j) = j < 4 ? 4g + j : 16 + 4g + (j - 4),
// so activations pass from layer to layer in registers (the weights are
// packed with the same permutation).  Weight gradients contract over points
// and go through an LDS transpose.
//
// Inference operands are packed separately with the factor -log2(e) of the
// sigmoid's exponent folded in (layer-1 weights, layer 1-3 biases; the output
// weights carry -ln 2): the pre-activations are then u = -log2(e) z and each
// activation is u / (1 + 2^u) = -log2(e) SiLU(z), one multiply fewer per
// activation than SiLU from z.
#pragma once

#include "device_physics.h"

namespace cfx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpH = 64;     // hidden width
constexpr int kMlpIO = 16;    // padded input/output width (ns + 2 <= 16)
constexpr int kMlpMaxSp = kMlpIO - 2;
constexpr float kMlpPureX = 0.98f;  // largest mole fraction the trained surrogate answers

// Packed A operands, in blocks of 64 lanes x 8 bf16 (one MFMA's A operand):
//   forward  W1: 4 (M)            W2, W3: 4 x 2 (M x K)     W4: 2 (K)
//   backward W4^T: 4 (M)          W3^T, W2^T: 4 x 2
enum MlpBlock : int {
  kF1 = 0, kF2 = 4, kF3 = 12, kF4 = 20, kB4 = 22, kB3 = 26, kB2 = 34, kNBlocks = 42
};
constexpr int kOpBlock = 64 * 8;

// FP32 master parameters, flat: W1[64][nin] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[nout][64] b4[nout].
struct MlpShape {
  int nin, nout;
  __host__ __device__ int w1() const { return 0; }
  __host__ __device__ int b1() const { return kMlpH * nin; }
  __host__ __device__ int w2() const { return b1() + kMlpH; }
  __host__ __device__ int b2() const { return w2() + kMlpH * kMlpH; }
  __host__ __device__ int w3() const { return b2() + kMlpH; }
  __host__ __device__ int b3() const { return w3() + kMlpH * kMlpH; }
  __host__ __device__ int w4() const { return b3() + kMlpH; }
  __host__ __device__ int b4() const { return w4() + nout * kMlpH; }
  __host__ __device__ int count() const { return b4() + nout; }
};

__host__ __device__ inline int mlp_kperm(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

// Normalisation, training envelope and fallback check (device resident).
struct MlpDev {
  int nin, nout;
  float x_mean[kMlpIO], x_inv_std[kMlpIO];  // inputs: T, log p, Y_l
  float t_mean[kMlpIO], t_std[kMlpIO];      // outputs: log mu, log kappa, log rho*D_l
  // envelope: box on T and log p, and the Mahalanobis distance of the
  // normalised input in the principal axes of the training inputs
  float t_lo, t_hi, lp_lo, lp_hi;
  float pca[kMlpIO][kMlpIO];  // rows: principal axes
  float pca_inv_var[kMlpIO];  // 1/eigenvalue, 0 for degenerate axes
  // the same distance as a quadratic form, d2 = sum_{i<=j} quad[i][j] x_i x_j
  // (quad = P^T diag(inv_var) P with the off-diagonal terms doubled)
  float quad[kMlpIO][kMlpIO];
  float d2_max;
  // near-pure mixtures: the transport routine switches rho*D_l to mu/Sc at
  // |X_l - 1| < 1e-3 (Transport.cu:88-91), a jump by the ratio of the other
  // species' mean molecular weight to W_l that a smooth network cannot
  // follow; points with a mole fraction >= x_pure take the Wilke routine
  float x_pure;
  float check_tol;            // relative error bound of the sampled Wilke check
  int check_every;            // one wavefront in check_every is checked
  int ready;
};

// Normalised network inputs of one point (x[f] = 0 beyond nin) and whether
// the point lies outside the training envelope: T or log p outside the
// sampled box, or a squared Mahalanobis distance in the principal axes of the
// training inputs beyond the largest one seen in training.
template <int NSP>
__device__ __forceinline__ bool mlp_inputs(const MlpDev *__restrict__ m, float t, float lp, const float *Y,
                                           float x[kMlpIO], float x_max = 0.0f) {
  x[0] = (t - m->x_mean[0]) * m->x_inv_std[0];
  x[1] = (lp - m->x_mean[1]) * m->x_inv_std[1];
#pragma unroll
  for (int f = 2; f < kMlpIO; ++f) x[f] = f < NSP + 2 ? (Y[f - 2] - m->x_mean[f]) * m->x_inv_std[f] : 0.0f;
  float d2 = 0.0f;
#pragma unroll
  for (int i = 0; i < NSP + 2; ++i) {
    float s = 0.0f;
#pragma unroll
    for (int j = i; j < NSP + 2; ++j) s += m->quad[i][j] * x[j];
    d2 += x[i] * s;
  }
  const bool box = t >= m->t_lo && t <= m->t_hi && lp >= m->lp_lo && lp <= m->lp_hi;
  return !box || !(d2 <= m->d2_max) || !(x_max < m->x_pure);
}

// ----------------------------------------------------------- activations ---
// sigma(z) = 1 / (1 + 2^(-z log2 e)) with the hardware exp2 and reciprocal
// (v_exp_f32, v_rcp_f32, 1 ulp); an overflow of 2^x to +inf gives sigma = 0.
__device__ __forceinline__ float mlp_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sigmoid(z); }
// inference form: u = -log2(e) z  ->  -log2(e) SiLU(z)
__device__ __forceinline__ float mlp_silu_folded(float u) {
  return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}
__device__ __forceinline__ float mlp_dsilu(float z) {
  const float s = mlp_sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 pack_pair(f32x4 lo, f32x4 hi) {
  bf16x8 r;
  r[0] = (__bf16)lo[0]; r[1] = (__bf16)lo[1]; r[2] = (__bf16)lo[2]; r[3] = (__bf16)lo[3];
  r[4] = (__bf16)hi[0]; r[5] = (__bf16)hi[1]; r[6] = (__bf16)hi[2]; r[7] = (__bf16)hi[3];
  return r;
}

__device__ __forceinline__ f32x4 silu4(f32x4 z) {
  f32x4 h;
  h[0] = mlp_silu(z[0]); h[1] = mlp_silu(z[1]); h[2] = mlp_silu(z[2]); h[3] = mlp_silu(z[3]);
  return h;
}
__device__ __forceinline__ f32x4 silu4_folded(f32x4 u) {
  f32x4 h;
  h[0] = mlp_silu_folded(u[0]); h[1] = mlp_silu_folded(u[1]);
  h[2] = mlp_silu_folded(u[2]); h[3] = mlp_silu_folded(u[3]);
  return h;
}

// operands and biases in global memory or (lds_*) staged in LDS
template <class T>
using lds_ptr = __attribute__((address_space(3))) T *;

__device__ __forceinline__ bf16x8 load_op(const bf16x8 *__restrict__ ops, int block, int lane) {
  return ops[block * 64 + lane];
}
__device__ __forceinline__ bf16x8 load_op(lds_ptr<const bf16x8> ops, int block, int lane) {
  return ops[block * 64 + lane];
}

// bias rows 4g..4g+3 of M-block mb of layer l (biases padded to 64 per layer)
__device__ __forceinline__ f32x4 load_bias(const float *__restrict__ bias, int layer, int mb, int lane) {
  return *reinterpret_cast<const f32x4 *>(bias + layer * kMlpH + 16 * mb + 4 * (lane >> 4));
}
__device__ __forceinline__ f32x4 load_bias(lds_ptr<const f32x4> bias, int layer, int mb, int lane) {
  return bias[(layer * kMlpH + 16 * mb) / 4 + (lane >> 4)];
}

// Activations of one 16-point tile kept for the backward pass.
struct MlpTile {
  f32x4 z1[4], z2[4], z3[4];
  bf16x8 h0, h1[2], h2[2], h3[2];
  f32x4 y;
};

// Forward pass of one tile: x is the B operand of the normalised inputs.
__device__ __forceinline__ void mlp_forward(const bf16x8 *__restrict__ ops, const float *__restrict__ bias,
                                            bf16x8 x, MlpTile &t) {
  const int lane = threadIdx.x & 63;
  t.h0 = x;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) t.z1[mb] = mfma_bf16(load_op(ops, kF1 + mb, lane), x, load_bias(bias, 0, mb, lane));
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) t.h1[kb] = pack_pair(silu4(t.z1[2 * kb]), silu4(t.z1[2 * kb + 1]));
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 a = load_bias(bias, 1, mb, lane);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) a = mfma_bf16(load_op(ops, kF2 + 2 * mb + kb, lane), t.h1[kb], a);
    t.z2[mb] = a;
  }
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) t.h2[kb] = pack_pair(silu4(t.z2[2 * kb]), silu4(t.z2[2 * kb + 1]));
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 a = load_bias(bias, 2, mb, lane);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) a = mfma_bf16(load_op(ops, kF3 + 2 * mb + kb, lane), t.h2[kb], a);
    t.z3[mb] = a;
  }
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) t.h3[kb] = pack_pair(silu4(t.z3[2 * kb]), silu4(t.z3[2 * kb + 1]));
  f32x4 y = load_bias(bias, 3, 0, lane);
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) y = mfma_bf16(load_op(ops, kF4 + kb, lane), t.h3[kb], y);
  t.y = y;
}

// Forward pass for inference only (no activations kept), on the folded
// inference operands.
__device__ __forceinline__ f32x4 mlp_infer_tile(const bf16x8 *__restrict__ ops, const float *__restrict__ bias,
                                                bf16x8 x) {
  const int lane = threadIdx.x & 63;
  f32x4 z[4];
  bf16x8 h[2];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) z[mb] = mfma_bf16(load_op(ops, kF1 + mb, lane), x, load_bias(bias, 0, mb, lane));
#pragma unroll
  for (int layer = 1; layer < 3; ++layer) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) h[kb] = pack_pair(silu4_folded(z[2 * kb]), silu4_folded(z[2 * kb + 1]));
    const int base = layer == 1 ? kF2 : kF3;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 a = load_bias(bias, layer, mb, lane);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) a = mfma_bf16(load_op(ops, base + 2 * mb + kb, lane), h[kb], a);
      z[mb] = a;
    }
  }
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) h[kb] = pack_pair(silu4_folded(z[2 * kb]), silu4_folded(z[2 * kb + 1]));
  f32x4 y = load_bias(bias, 3, 0, lane);
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) y = mfma_bf16(load_op(ops, kF4 + kb, lane), h[kb], y);
  return y;
}

// Inference of NTL tiles at once on the folded operands, weight-stationary:
// every A operand is fetched once and applied to all tiles.
template <int NTL, class OpPtr, class BiasPtr>
__device__ __forceinline__ void mlp_infer_tiles(OpPtr ops, BiasPtr bias, const bf16x8 *x, f32x4 *y) {
  const int lane = threadIdx.x & 63;
  f32x4 z[NTL][4];
  bf16x8 h[NTL][2];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const bf16x8 a = load_op(ops, kF1 + mb, lane);
    const f32x4 b0 = load_bias(bias, 0, mb, lane);
#pragma unroll
    for (int t = 0; t < NTL; ++t) z[t][mb] = mfma_bf16(a, x[t], b0);
  }
#pragma unroll
  for (int layer = 1; layer < 3; ++layer) {
#pragma unroll
    for (int t = 0; t < NTL; ++t)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) h[t][kb] = pack_pair(silu4_folded(z[t][2 * kb]), silu4_folded(z[t][2 * kb + 1]));
    const int base = layer == 1 ? kF2 : kF3;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const f32x4 b0 = load_bias(bias, layer, mb, lane);
#pragma unroll
      for (int t = 0; t < NTL; ++t) z[t][mb] = b0;
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        const bf16x8 a = load_op(ops, base + 2 * mb + kb, lane);
#pragma unroll
        for (int t = 0; t < NTL; ++t) z[t][mb] = mfma_bf16(a, h[t][kb], z[t][mb]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NTL; ++t)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) h[t][kb] = pack_pair(silu4_folded(z[t][2 * kb]), silu4_folded(z[t][2 * kb + 1]));
  const f32x4 b3 = load_bias(bias, 3, 0, lane);
#pragma unroll
  for (int t = 0; t < NTL; ++t) y[t] = b3;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    const bf16x8 a = load_op(ops, kF4 + kb, lane);
#pragma unroll
    for (int t = 0; t < NTL; ++t) y[t] = mfma_bf16(a, h[t][kb], y[t]);
  }
}

}  // namespace cfx
