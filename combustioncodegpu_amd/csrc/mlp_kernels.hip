// mlp_kernels.hip — training and inference kernels of the transport
// surrogate that do not depend on the species count (the network sizes are
// run-time: nin = nout = n_spec + 2 <= 16).  See mlp.h for the layout.
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "mlp_kernels.h"

namespace cfx {

namespace {

constexpr int kTS = 40;  // LDS row stride (bf16) of the 32-point transposes

__device__ __forceinline__ bf16x8 x_operand(const float *__restrict__ xs, long long s, int g) {
  // slots j < 4 carry features 4g..4g+3, slots j >= 4 features 16+ (zero)
  const f32x4 v = *reinterpret_cast<const f32x4 *>(xs + s * kMlpIO + 4 * g);
  bf16x8 r;
  r[0] = (__bf16)v[0]; r[1] = (__bf16)v[1]; r[2] = (__bf16)v[2]; r[3] = (__bf16)v[3];
  r[4] = (__bf16)0.0f; r[5] = (__bf16)0.0f; r[6] = (__bf16)0.0f; r[7] = (__bf16)0.0f;
  return r;
}

__device__ __forceinline__ bf16x8 lds_operand(const __bf16 (*t)[kTS], int row, int g) {
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = t[row][4 * g + j];
    r[4 + j] = t[row][16 + 4 * g + j];
  }
  return r;
}

__device__ __forceinline__ long long sample_of(long long k, long long offset, long long mult, long long n) {
  return (long long)(((unsigned long long)(k + offset) * (unsigned long long)mult) % (unsigned long long)n);
}

// One wavefront = 32 samples (two 16-point tiles): forward, loss, backward;
// the weight and bias gradients of the 32 samples go to partial[w * P + i].
// loss = scale * sum over samples and outputs of (y - t)^2.
__global__ void __launch_bounds__(64) mlp_train_kernel(const float *__restrict__ xs, const float *__restrict__ ts,
                                                       long long n_train, long long first, long long offset,
                                                       long long mult, const bf16x8 *__restrict__ ops,
                                                       const float *__restrict__ bias, MlpShape shp, float scale,
                                                       float *__restrict__ partial, float *__restrict__ loss_out) {
  __shared__ __bf16 tD[208][kTS];  // dZ rows: layer 1 [0,64), 2 [64,128), 3 [128,192), output [192,208)
  __shared__ __bf16 tH[208][kTS];  // H rows: inputs [0,16), H1 [16,80), H2 [80,144), H3 [144,208)
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int w = blockIdx.x;
  const int P = shp.count();
  float lsum = 0.0f;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const long long s = sample_of(first + 32LL * w + 16 * t + c, offset, mult, n_train);
    MlpTile T;
    mlp_forward(ops, bias, x_operand(xs, s, g), T);
    const f32x4 tg = *reinterpret_cast<const f32x4 *>(ts + s * kMlpIO + 4 * g);
    f32x4 dy;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float d = (4 * g + r < shp.nout) ? T.y[r] - tg[r] : 0.0f;
      lsum += d * d;
      dy[r] = 2.0f * scale * d;
    }
    const int col = 16 * t + c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      tH[4 * g + r][col] = T.h0[r];
      tD[192 + 4 * g + r][col] = (__bf16)dy[r];
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int f = 32 * kb + mlp_kperm(g, j);
        tH[16 + f][col] = T.h1[kb][j];
        tH[80 + f][col] = T.h2[kb][j];
        tH[144 + f][col] = T.h3[kb][j];
      }
    // backward through the output layer: dH3 = W4^T dY
    bf16x8 dyop;
#pragma unroll
    for (int r = 0; r < 4; ++r) dyop[r] = (__bf16)dy[r], dyop[4 + r] = (__bf16)0.0f;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 dz[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const f32x4 dh = mfma_bf16(load_op(ops, kB4 + mb, lane), dyop, zero);
#pragma unroll
      for (int r = 0; r < 4; ++r) dz[mb][r] = dh[r] * mlp_dsilu(T.z3[mb][r]);
    }
    // layers 3 and 2: store dZ_l, then dH_{l-1} = W_l^T dZ_l
#pragma unroll
    for (int layer = 3; layer >= 2; --layer) {
      const int drow = layer == 3 ? 128 : 64;
      bf16x8 op[2];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) op[kb] = pack_pair(dz[2 * kb], dz[2 * kb + 1]);
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) tD[drow + 32 * kb + mlp_kperm(g, j)][col] = op[kb][j];
      const int base = layer == 3 ? kB3 : kB2;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        f32x4 dh = zero;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) dh = mfma_bf16(load_op(ops, base + 2 * mb + kb, lane), op[kb], dh);
        const f32x4 &z = layer == 3 ? T.z2[mb] : T.z1[mb];
#pragma unroll
        for (int r = 0; r < 4; ++r) dz[mb][r] = dh[r] * mlp_dsilu(z[r]);
      }
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int r = 0; r < 4; ++r) tD[16 * mb + 4 * g + r][col] = (__bf16)dz[mb][r];
  }
  __syncthreads();
  // weight gradients dW = sum over the 32 points of dZ H^T; bias gradients
  float *__restrict__ out = partial + (long long)w * P;
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int layer = 0; layer < 4; ++layer) {
    const int drow = layer == 0 ? 0 : (layer == 1 ? 64 : (layer == 2 ? 128 : 192));
    const int hrow = layer == 0 ? 0 : (layer == 1 ? 16 : (layer == 2 ? 80 : 144));
    const int nmb = layer == 3 ? 1 : 4, nnb = layer == 0 ? 1 : 4;
    const int n_out = layer == 3 ? shp.nout : kMlpH, n_in = layer == 0 ? shp.nin : kMlpH;
    const int woff = layer == 0 ? shp.w1() : (layer == 1 ? shp.w2() : (layer == 2 ? shp.w3() : shp.w4()));
    const int boff = layer == 0 ? shp.b1() : (layer == 1 ? shp.b2() : (layer == 2 ? shp.b3() : shp.b4()));
    for (int mb = 0; mb < nmb; ++mb) {
      const bf16x8 a = lds_operand(tD, drow + 16 * mb + c, g);
      for (int nb = 0; nb < nnb; ++nb) {
        const bf16x8 b = lds_operand(tH, hrow + 16 * nb + c, g);
        const f32x4 d = mfma_bf16(a, b, zero);
        const int in = 16 * nb + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 16 * mb + 4 * g + r;
          if (o < n_out && in < n_in) out[woff + o * n_in + in] = d[r];
        }
      }
      // bias: sum of the bf16 dZ of the row over the 32 points
      const int row = 16 * mb + c;
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) s += (float)tD[drow + row][4 * g + j] + (float)tD[drow + row][16 + 4 * g + j];
      s += __shfl_xor(s, 16, 64);
      s += __shfl_xor(s, 32, 64);
      if (g == 0 && row < n_out) out[boff + row] = s;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if (lane == 0) loss_out[w] = lsum * scale;
}

// Sum of the per-wavefront gradients and an Adam step on the FP32 master
// parameters.  A workgroup takes 64 parameters; its four wavefronts sum
// quarters of the partials (each in increasing order, loads coalesced over
// the parameters) and the quarter sums are added in order, so the result is
// deterministic.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float *__restrict__ params, float *__restrict__ m,
                                                       float *__restrict__ v, const float *__restrict__ partial,
                                                       int n_partial, int P, float lr, float b1, float b2,
                                                       float eps, float c1, float c2, float *__restrict__ grad_out) {
  __shared__ float quarter[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int w0 = (int)((long long)n_partial * wv / 4), w1 = (int)((long long)n_partial * (wv + 1) / 4);
  float part = 0.0f;
  if (i < P)
    for (int w = w0; w < w1; ++w) part += partial[(long long)w * P + i];
  quarter[wv][lane] = part;
  __syncthreads();
  if (wv != 0 || i >= P) return;
  const float gsum = ((quarter[0][lane] + quarter[1][lane]) + quarter[2][lane]) + quarter[3][lane];
  if (grad_out) grad_out[i] = gsum;
  if (lr <= 0.0f) return;
  const float mi = b1 * m[i] + (1.0f - b1) * gsum;
  const float vi = b2 * v[i] + (1.0f - b2) * gsum * gsum;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 A operands of the forward and backward products, and the biases
// padded to 64 per layer, from the FP32 master parameters.  fold: the
// inference operands, with -log2(e) on the layer-1 weights and the layer 1-3
// biases and -ln(2) on the output weights (float products, then BF16).
constexpr float kFold = -1.4426950408889634f, kUnfold = -0.69314718055994531f;

__global__ void __launch_bounds__(256) mlp_pack_kernel(const float *__restrict__ p, MlpShape shp,
                                                       __bf16 *__restrict__ ops, float *__restrict__ bias,
                                                       int fold) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 4 * kMlpH) {
    const int layer = e / kMlpH, row = e % kMlpH;
    float b = 0.0f;
    if (layer == 0) b = p[shp.b1() + row];
    else if (layer == 1) b = p[shp.b2() + row];
    else if (layer == 2) b = p[shp.b3() + row];
    else if (row < shp.nout) b = p[shp.b4() + row];
    bias[e] = fold && layer < 3 ? b * kFold : b;
  }
  if (e >= kNBlocks * kOpBlock) return;
  const int blk = e / kOpBlock, lane = (e % kOpBlock) / 8, j = e % 8;
  const int g = lane >> 4, c = lane & 15, k = mlp_kperm(g, j);
  float val = 0.0f;
  if (blk < kF2) {  // W1, M-block blk
    const int row = 16 * blk + c;
    if (k < shp.nin) val = p[shp.w1() + row * shp.nin + k];
    if (fold) val *= kFold;
  } else if (blk < kF4) {  // W2 / W3
    const int q = blk < kF3 ? blk - kF2 : blk - kF3, mb = q / 2, kb = q % 2;
    const int w = blk < kF3 ? shp.w2() : shp.w3();
    val = p[w + (16 * mb + c) * kMlpH + 32 * kb + k];
  } else if (blk < kB4) {  // W4, K-block
    const int kb = blk - kF4;
    if (c < shp.nout) val = p[shp.w4() + c * kMlpH + 32 * kb + k];
    if (fold) val *= kUnfold;
  } else if (blk < kB3) {  // W4^T: rows = hidden units, k = outputs
    const int mb = blk - kB4;
    if (k < shp.nout) val = p[shp.w4() + k * kMlpH + 16 * mb + c];
  } else {  // W3^T / W2^T
    const int q = blk < kB2 ? blk - kB3 : blk - kB2, mb = q / 2, kb = q % 2;
    const int w = blk < kB2 ? shp.w3() : shp.w2();
    val = p[w + (32 * kb + k) * kMlpH + 16 * mb + c];
  }
  ops[e] = (__bf16)val;
}

// Normalised predictions of normalised inputs (validation), on the folded
// inference operands: n a multiple of 16 is not required; the last tile
// repeats a valid sample.
__global__ void __launch_bounds__(64) mlp_predict_kernel(const float *__restrict__ xs, long long n,
                                                         const bf16x8 *__restrict__ ops,
                                                         const float *__restrict__ bias, int nout,
                                                         float *__restrict__ ys) {
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const long long s = (long long)blockIdx.x * 16 + c;
  const long long sv = s < n ? s : n - 1;
  const f32x4 y = mlp_infer_tile(ops, bias, x_operand(xs, sv, g));
  if (s < n)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (4 * g + r < nout) ys[s * kMlpIO + 4 * g + r] = y[r];
}

// Per-workgroup sums of the features, their squares and (inputs) the
// cross products, for normalisation and the envelope (double).
__global__ void __launch_bounds__(256) mlp_moments_kernel(const double *__restrict__ x, long long n, int nf,
                                                          int cross, double *__restrict__ out) {
  // out per workgroup: nf sums, nf squares, nf*nf cross (if cross)
  __shared__ double red[256];
  const int nval = 2 * nf + (cross ? nf * nf : 0);
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long s0 = (long long)blockIdx.x * per, s1 = s0 + per < n ? s0 + per : n;
  for (int q = 0; q < nval; ++q) {
    double acc = 0.0;
    for (long long s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
      const double *r = x + s * nf;
      if (q < nf) acc += r[q];
      else if (q < 2 * nf) acc += r[q - nf] * r[q - nf];
      else acc += r[(q - 2 * nf) / nf] * r[(q - 2 * nf) % nf];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(long long)blockIdx.x * nval + q] = red[0];
    __syncthreads();
  }
}

// x_norm = (x - mean) * inv_std into rows of 16 floats (zero padded).
__global__ void __launch_bounds__(256) mlp_normalize_kernel(const double *__restrict__ x, long long n, int nf,
                                                            const float *__restrict__ mean,
                                                            const float *__restrict__ inv_std,
                                                            float *__restrict__ xs) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * kMlpIO) return;
  const long long s = e / kMlpIO;
  const int f = (int)(e % kMlpIO);
  xs[e] = f < nf ? (float)((x[s * nf + f] - (double)mean[f]) * (double)inv_std[f]) : 0.0f;
}

}  // namespace

int launch_mlp_train(const float *xs, const float *ts, long long n_train, long long first, long long offset,
                     long long mult, int waves, const void *ops, const float *bias, const MlpShape &shp,
                     float scale, float *partial, float *loss, hipStream_t st) {
  mlp_train_kernel<<<waves, 64, 0, st>>>(xs, ts, n_train, first, offset, mult, (const bf16x8 *)ops, bias, shp,
                                         scale, partial, loss);
  return (int)hipGetLastError();
}

int launch_mlp_adam(float *params, float *m, float *v, const float *partial, int n_partial, int P, float lr,
                    float b1, float b2, float eps, float c1, float c2, float *grad_out, hipStream_t st) {
  mlp_adam_kernel<<<(P + 63) / 64, 256, 0, st>>>(params, m, v, partial, n_partial, P, lr, b1, b2, eps, c1, c2,
                                                 grad_out);
  return (int)hipGetLastError();
}

int launch_mlp_pack(const float *params, const MlpShape &shp, void *ops, float *bias, int fold, hipStream_t st) {
  mlp_pack_kernel<<<(kNBlocks * kOpBlock + 255) / 256, 256, 0, st>>>(params, shp, (__bf16 *)ops, bias, fold);
  return (int)hipGetLastError();
}

int launch_mlp_predict(const float *xs, long long n, const void *ops, const float *bias, int nout, float *ys,
                       hipStream_t st) {
  if (n <= 0) return 0;
  mlp_predict_kernel<<<(int)((n + 15) / 16), 64, 0, st>>>(xs, n, (const bf16x8 *)ops, bias, nout, ys);
  return (int)hipGetLastError();
}

int launch_mlp_moments(const double *x, long long n, int nf, int cross, int groups, double *out, hipStream_t st) {
  mlp_moments_kernel<<<groups, 256, 0, st>>>(x, n, nf, cross, out);
  return (int)hipGetLastError();
}

int launch_mlp_normalize(const double *x, long long n, int nf, const float *mean, const float *inv_std, float *xs,
                         hipStream_t st) {
  const long long e = n * kMlpIO;
  mlp_normalize_kernel<<<(int)((e + 255) / 256), 256, 0, st>>>(x, n, nf, mean, inv_std, xs);
  return (int)hipGetLastError();
}

}  // namespace cfx
