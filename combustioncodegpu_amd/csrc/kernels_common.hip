// kernels_common.hip — the kernels that do not depend on the species count:
// CFL time step with its min-reduction, last-step snapshot, residual
// partial sums, 2-D k-gradient elimination.  Also defines the registries the
// species-count translation units fill.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "mesh.h"

namespace cfx {

KernelSet64 g_kernels64[kMaxSp + 1];
KernelSet32 g_kernels32[kMaxSp + 1];

namespace {

__device__ __forceinline__ long long pidx(const GView &B, int i, int j, int k) {
  return (long long)(k + kNgg) * B.sz + (long long)(j + kNgg) * B.sy + (i + kNgg);
}

// padded index of interior point q (i fastest)
__device__ __forceinline__ long long interior_pidx(const GView &B, long long q) {
  const unsigned u = (unsigned)q, mx = (unsigned)B.mx, my = (unsigned)B.my;
  const unsigned i = u % mx, r = u / mx;
  return pidx(B, (int)i, (int)(r % my), (int)(r / my));
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------- time step ---
// SteadyTemporalScheme::compute_time_step (TemporalScheme.cu:10-43) over
// kDtIter chunks of the block's interior points, then a wavefront
// min-reduction of the CFL time step into *dt_bits (positive doubles order
// like their bit patterns).
__global__ void __launch_bounds__(kTpb) dt_kernel(const BlockView *__restrict__ views, const int *__restrict__ off,
                                                  int nb, SchemeDev S, int ns, int store_local,
                                                  unsigned long long *dt_bits) {
  __shared__ double red[kTpb / 64];
  const int w = blockIdx.x + off[0];
  const int blk = find_seg(off, nb, w);
  const GView B = gview(views[blk]);
  const long long np = B.np;
  const long long base = (long long)(w - off[blk]) * (kTpb * kDtIter);
  double dmin = 1e300;
  for (int it = 0; it < kDtIter; ++it) {
    const long long q = base + it * kTpb + threadIdx.x;
    if (q >= B.ni) break;
    const long long p = interior_pidx(B, q);
    double mm[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) mm[c] = B.metric[c * np + p];
    const double gx = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
    const double gy = sqrt(mm[3] * mm[3] + mm[4] * mm[4] + mm[5] * mm[5]);
    const double gz = sqrt(mm[6] * mm[6] + mm[7] * mm[7] + mm[8] * mm[8]);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], ww = B.bv[3 * np + p];
    const double U = u * mm[0] + v * mm[1] + ww * mm[2];
    const double V = u * mm[3] + v * mm[4] + ww * mm[5];
    const double W = u * mm[6] + v * mm[7] + ww * mm[8];
    const double a = B.mixed ? (double)B.props32[np + p] : B.props[np + p];
    double sr = fabs(U) + fabs(V) + a * (gx + gy);
    double sv = gx * gx + gy * gy;
    if (S.dim == 3) {
      sr += fabs(W) + a * gz;
      sv += gz * gz;
    }
    const double gam = ns > 0 ? (B.mixed ? (double)B.props32[p] : B.props[p]) : S.gamma_air;
    const double mu = B.mixed ? (double)B.props32[2 * np + p] : B.props[2 * np + p];
    sv *= fmax(gam, 4.0 / 3.0) * (mu / B.bv[p] * S.inv_pr);
    const double dt = S.cfl / (sr + sv);
    if (store_local) B.dt_local[q] = dt;
    dmin = fmin(dmin, dt);
  }
  dmin = wave_min(dmin);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = dmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int k = 1; k < kTpb / 64; ++k) r = fmin(r, red[k]);
    atomicMin(dt_bits, (unsigned long long)__double_as_longlong(r));
  }
}

__global__ void set_u64_kernel(unsigned long long *dst, unsigned long long v) { *dst = v; }

// store_last_step (TimeAdvanceFunc.cu:13-24)
__global__ void __launch_bounds__(kTpb) store_last_kernel(const BlockView *__restrict__ views,
                                                          const int *__restrict__ off, int nb) {
  const int w = blockIdx.x + off[0];
  const int blk = find_seg(off, nb, w);
  const GView B = gview(views[blk]);
  const long long np = B.np, ni = B.ni;
  const long long base = (long long)(w - off[blk]) * (kTpb * kLastIter);
  for (int it = 0; it < kLastIter; ++it) {
    const long long q = base + it * kTpb + threadIdx.x;
    if (q >= ni) break;
    const long long p = interior_pidx(B, q);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], ww = B.bv[3 * np + p];
    B.bv_last[q] = B.bv[p];
    B.bv_last[ni + q] = sqrt(u * u + v * v + ww * ww);
    B.bv_last[2 * ni + q] = B.bv[4 * np + p];
    B.bv_last[3 * ni + q] = B.bv[5 * np + p];
  }
}

// compute_square_of_dbv (TimeAdvanceFunc.cu:292-303) + reduction_of_dbv_squared
// (Driver.cu:329-366): one partial sum per workgroup, summed in a fixed order.
__global__ void __launch_bounds__(kTpb) residual_partial_kernel(const BlockView *__restrict__ views,
                                                                const int *__restrict__ off, int nb,
                                                                double *partial) {
  __shared__ double red[kTpb / 64][4];
  const int w = blockIdx.x + off[0];
  const int blk = find_seg(off, nb, w);
  const GView B = gview(views[blk]);
  const long long np = B.np, ni = B.ni;
  const long long base = (long long)(w - off[blk]) * (kTpb * kResIter);
  double s[4] = {0, 0, 0, 0};
  for (int it = 0; it < kResIter; ++it) {
    const long long q = base + it * kTpb + threadIdx.x;
    if (q >= ni) break;
    const long long p = interior_pidx(B, q);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], ww = B.bv[3 * np + p];
    const double cur[4] = {B.bv[p], sqrt(u * u + v * v + ww * ww), B.bv[4 * np + p], B.bv[5 * np + p]};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double dd = cur[l] - B.bv_last[l * ni + q];
      s[l] += dd * dd;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double r = wave_sum(s[l]);
    if (lane == 0) red[wv][l] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0;
    for (int k = 0; k < kTpb / 64; ++k) r += red[k][threadIdx.x];
    partial[(long long)w * 4 + threadIdx.x] = r;
  }
}

__global__ void __launch_bounds__(kTpb) residual_final_kernel(const double *partial, int n, double *out) {
  __shared__ double red[kTpb / 64][4];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < n; b += blockDim.x)
#pragma unroll
    for (int l = 0; l < 4; ++l) s[l] += partial[b * 4 + l];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double r = wave_sum(s[l]);
    if (lane == 0) red[wv][l] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0;
    for (int k = 0; k < kTpb / 64; ++k) r += red[k][threadIdx.x];
    out[threadIdx.x] = r;
  }
}

// eliminate_k_gradient (Field.cu:268-292), 2-D only
__global__ void eliminate_k_kernel(BlockView view, int ns) {
  const GView B = gview(view);
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - kNgg;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - kNgg;
  if (i >= B.mx + kNgg || j >= B.my + kNgg) return;
  const long long np = B.np;
  const long long p0 = pidx(B, i, j, 0);
  for (int k = 1; k <= kNgg; ++k) {
    const long long pp = pidx(B, i, j, k), pm = pidx(B, i, j, -k);
    for (int l = 0; l < 6; ++l) {
      B.bv[l * np + pp] = B.bv[l * np + p0];
      B.bv[l * np + pm] = B.bv[l * np + p0];
    }
    if (B.mixed) {  // the mixed path derives Y from cv: copy rho and rho*Y instead
      for (int l = 0; l < 5 + ns; ++l) {
        if (l >= 1 && l < 5) continue;
        B.cv[l * np + pp] = B.cv[l * np + p0];
        B.cv[l * np + pm] = B.cv[l * np + p0];
      }
    } else {
      for (int l = 0; l < ns; ++l) {
        B.yk[l * np + pp] = B.yk[l * np + p0];
        B.yk[l * np + pm] = B.yk[l * np + p0];
      }
    }
  }
}

}  // namespace

int launch_dt_reset(unsigned long long *dt_bits, hipStream_t st) {
  set_u64_kernel<<<1, 1, 0, st>>>(dt_bits, 0x7FF0000000000000ull);
  return (int)hipGetLastError();
}

int launch_dt(const Batch &b, const SchemeDev &S, int ns, int store_local, unsigned long long *dt_bits,
              hipStream_t st) {
  if (b.total <= 0) return 0;
  dt_kernel<<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, S, ns, store_local, dt_bits);
  return (int)hipGetLastError();
}

// Jacobian and metrics of one block (Block::compute_jac_metric,
// Mesh.cpp:226-488): one thread per point of the ngg-padded range, reading the
// vertex coordinates over ngg+1 ghost layers (x, y, z planes of gx*gy*gz).
// FP32 copies are written when metric32 is set.  A point whose Jacobian is
// not positive is counted in *bad and left unwritten.
__global__ void __launch_bounds__(kTpb) metric_kernel(const double *__restrict__ xyz, int mx, int my, int mz,
                                                      double *__restrict__ metric, double *__restrict__ jac,
                                                      float *__restrict__ metric32, float *__restrict__ jac32,
                                                      int *bad) {
  const int px = mx + 2 * kNgg, py = my + 2 * kNgg, pz = mz + 2 * kNgg;
  const long long np = (long long)px * py * pz;
  const long long p = (long long)blockIdx.x * kTpb + threadIdx.x;
  if (p >= np) return;
  const int i = (int)(p % px) - kNgg;
  const long long r = p / px;
  const int j = (int)(r % py) - kNgg, k = (int)(r / py) - kNgg;
  const int G = kNgg + 1, gx = mx + 2 * G, gy = my + 2 * G;
  const long long ng = (long long)gx * gy * (mz + 2 * G);
  // i + di in [-ngg-1, m+ngg]: inside the ngg+1 ghost layers
  const auto X = [&](int cc, int di, int dj, int dk) {
    return xyz[cc * ng + ((long long)(k + dk + G) * gy + (j + dj + G)) * gx + (i + di + G)];
  };
  double J, m[9];
  if (!metric_point(X, J, m)) {
    atomicAdd(bad, 1);
    return;
  }
  jac[p] = J;
#pragma unroll
  for (int c = 0; c < 9; ++c) metric[c * np + p] = m[c];
  if (metric32) {
    jac32[p] = (float)J;
#pragma unroll
    for (int c = 0; c < 9; ++c) metric32[c * np + p] = (float)m[c];
  }
}

int launch_metrics(const double *xyz, int mx, int my, int mz, double *metric, double *jac, float *metric32,
                   float *jac32, int *bad, hipStream_t st) {
  const long long np = (long long)(mx + 2 * kNgg) * (my + 2 * kNgg) * (mz + 2 * kNgg);
  const long long nblk = (np + kTpb - 1) / kTpb;
  if (nblk <= 0 || nblk > 0x7fffffffLL) return (int)hipErrorInvalidValue;
  metric_kernel<<<(unsigned)nblk, kTpb, 0, st>>>(xyz, mx, my, mz, metric, jac, metric32, jac32, bad);
  return (int)hipGetLastError();
}

int launch_store_last(const Batch &b, hipStream_t st) {
  if (b.total <= 0) return 0;
  store_last_kernel<<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb);
  return (int)hipGetLastError();
}

int launch_residual_partial(const Batch &b, double *partial, hipStream_t st) {
  if (b.total <= 0) return 0;
  residual_partial_kernel<<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, partial);
  return (int)hipGetLastError();
}

int launch_residual_final(const double *partial, int n, double *out, hipStream_t st) {
  residual_final_kernel<<<1, kTpb, 0, st>>>(partial, n, out);
  return (int)hipGetLastError();
}

int launch_eliminate_k(const BlockView &B, int ns, hipStream_t st) {
  const dim3 tpb(16, 16, 1);
  const dim3 g((B.px + 15) / 16, (B.py + 15) / 16, 1);
  eliminate_k_kernel<<<g, tpb, 0, st>>>(B, ns);
  return (int)hipGetLastError();
}

}  // namespace cfx
