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

// ------------------------------------------------------------ ghost grid ---
// The (ngg+1)-layer coordinate arrays of Mesh::init_ghost_grid
// (Mesh.cpp:718-922) built on the device, each step in the host restatement's
// arithmetic and, where one value feeds another, in its order: a thread owns
// a whole line or copy point, so the result is bit-identical to the host path.
// Layout: 3 planes (x, y, z) of gx * gy * gz, gx = mx + 2(ngg+1), i fastest.
__device__ __forceinline__ long long gidx(int i, int j, int k, int gx, int gy) {
  constexpr int G = kNgg + 1;
  return ((long long)(k + G) * gy + (j + G)) * gx + (i + G);
}

// interior vertices (3 x mx*my*mz, i fastest) into the padded array
__global__ void __launch_bounds__(kTpb) coord_scatter_kernel(const double *__restrict__ raw, int mx, int my, int mz,
                                                              double *__restrict__ xyz) {
  constexpr int G = kNgg + 1;
  const long long ni = (long long)mx * my * mz;
  const long long s = (long long)blockIdx.x * kTpb + threadIdx.x;
  if (s >= ni) return;
  const int gx = mx + 2 * G, gy = my + 2 * G;
  const long long ng = (long long)gx * gy * (mz + 2 * G);
  const int i = (int)(s % mx);
  const long long r = s / mx;
  const int j = (int)(r % my), k = (int)(r / my);
  const long long g = gidx(i, j, k, gx, gy);
#pragma unroll
  for (int c = 0; c < 3; ++c) xyz[c * ng + g] = raw[c * ni + s];
}

// Mirror extrapolation along each direction a over the interior of the other
// two (Mesh.cpp:720-768): one thread per line, layers g = 1..ngg+1 in order.
__global__ void __launch_bounds__(kTpb) ghost_extrapolate_kernel(double *__restrict__ xyz, int mx, int my, int mz,
                                                                 int dim) {
  constexpr int G = kNgg + 1;
  const int m[3] = {mx, my, mz};
  const int gx = mx + 2 * G, gy = my + 2 * G;
  const long long ng = (long long)gx * gy * (mz + 2 * G);
  long long l = (long long)blockIdx.x * kTpb + threadIdx.x;
  int a = 0;
  for (; a < dim; ++a) {
    const long long nl = (long long)m[(a + 1) % 3] * m[(a + 2) % 3];
    if (l < nl) break;
    l -= nl;
  }
  if (a >= dim) return;
  const int b = (a + 1) % 3, c = (a + 2) % 3;
  const int u = (int)(l / m[c]), v = (int)(l % m[c]);
  int p[3];
  p[b] = u;
  p[c] = v;
  auto at = [&](int cc, int s) -> double & {
    p[a] = s;
    return xyz[cc * ng + gidx(p[0], p[1], p[2], gx, gy)];
  };
  for (int g = 1; g <= G; ++g)
    for (int cc = 0; cc < 3; ++cc) {
      at(cc, -g) = 2 * at(cc, 0) - at(cc, g);
      at(cc, m[a] - 1 + g) = 2 * at(cc, m[a] - 1) - at(cc, m[a] - 1 - g);
    }
}

// 2-D: unit spacing in z (Mesh.cpp:769-785), one thread per (i, j)
__global__ void __launch_bounds__(kTpb) ghost_z2d_kernel(double *__restrict__ xyz, int mx, int my) {
  constexpr int G = kNgg + 1;
  const int gx = mx + 2 * G, gy = my + 2 * G, mz = 1;
  const long long ng = (long long)gx * gy * (mz + 2 * G);
  const long long s = (long long)blockIdx.x * kTpb + threadIdx.x;
  if (s >= (long long)mx * my) return;
  const int i = (int)(s % mx), j = (int)(s / mx);
  auto X = [&](int cc, int k) -> double & { return xyz[cc * ng + gidx(i, j, k, gx, gy)]; };
  for (int g = 1; g <= G; ++g) {
    X(0, -g) = X(0, 0);
    X(1, -g) = X(1, 0);
    X(2, -g) = X(2, 0) - g;
    X(0, mz - 1 + g) = X(0, mz - 1);
    X(1, mz - 1 + g) = X(1, mz - 1);
    X(2, mz - 1 + g) = X(2, mz - 1) + g;
  }
}

// One inter-block copy (Mesh.cpp:874-922): dst point lo + step * d from src
// point t0 + tstep * d[src_tar]; one thread per point.
struct GhostCopyArgs {
  int lo[3], n[3], step[3], t0[3], tstep[3], src_tar[3];
  int dgx, dgy, sgx, sgy;
  long long dng, sng;
};
__global__ void __launch_bounds__(kTpb) ghost_inner_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                                           GhostCopyArgs A) {
  const long long t = (long long)blockIdx.x * kTpb + threadIdx.x;
  if (t >= (long long)A.n[0] * A.n[1] * A.n[2]) return;
  const int c = (int)(t % A.n[2]);
  const long long r = t / A.n[2];
  const int b = (int)(r % A.n[1]), a = (int)(r / A.n[1]);
  const int di[3] = {A.step[0] * a, A.step[1] * b, A.step[2] * c};
  const int ts[3] = {A.t0[0] + A.tstep[0] * di[A.src_tar[0]], A.t0[1] + A.tstep[1] * di[A.src_tar[1]],
                     A.t0[2] + A.tstep[2] * di[A.src_tar[2]]};
  const long long gd = gidx(A.lo[0] + di[0], A.lo[1] + di[1], A.lo[2] + di[2], A.dgx, A.dgy);
  const long long gs = gidx(ts[0], ts[1], ts[2], A.sgx, A.sgy);
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) dst[cc * A.dng + gd] = src[cc * A.sng + gs];
}

// Edges and corners by the parallelogram rule (Mesh.cpp:793-867), three
// phases; every point a phase reads was written by an earlier step.
template <int PHASE>
__global__ void __launch_bounds__(kTpb) ghost_corner_kernel(double *__restrict__ xyz, int mx, int my, int mz) {
  constexpr int G = kNgg + 1;
  const int gx = mx + 2 * G, gy = my + 2 * G;
  const long long ng = (long long)gx * gy * (mz + 2 * G);
  const long long t = (long long)blockIdx.x * kTpb + threadIdx.x;
  // PHASE 0: (s, t, k); 1: (s, j, t); 2: (i, s, t)
  const long long n0 = PHASE == 2 ? (long long)(mx + kNgg + G + 1) : G;
  const long long n1 = PHASE == 1 ? my : G;
  const long long n2 = PHASE == 0 ? mz : G;
  if (t >= n0 * n1 * n2) return;
  const int x2 = (int)(t % n2);
  const long long r = t / n2;
  const int x1 = (int)(r % n1), x0 = (int)(r / n1);
  for (int cc = 0; cc < 3; ++cc) {
    auto X = [&](int i, int j, int k) -> double & { return xyz[cc * ng + gidx(i, j, k, gx, gy)]; };
    if constexpr (PHASE == 0) {
      const int s = x0 + 1, tt = x1 + 1, k = x2;
      const int i1 = -s, j1 = -tt, i2 = mx - 1 + s, j2 = my - 1 + tt;
      X(i1, j1, k) = X(i1, 0, k) + X(0, j1, k) - X(0, 0, k);
      X(i2, j2, k) = X(i2, my - 1, k) + X(mx - 1, j2, k) - X(mx - 1, my - 1, k);
      X(i1, j2, k) = X(i1, my - 1, k) + X(0, j2, k) - X(0, my - 1, k);
      X(i2, j1, k) = X(i2, 0, k) + X(mx - 1, j1, k) - X(mx - 1, 0, k);
    } else if constexpr (PHASE == 1) {
      const int s = x0 + 1, j = x1, tt = x2 + 1;
      const int i1 = -s, k1 = -tt, i2 = mx - 1 + s, k2 = mz - 1 + tt;
      X(i1, j, k1) = X(i1, j, 0) + X(0, j, k1) - X(0, j, 0);
      X(i2, j, k2) = X(i2, j, mz - 1) + X(mx - 1, j, k2) - X(mx - 1, j, mz - 1);
      X(i1, j, k2) = X(i1, j, mz - 1) + X(0, j, k2) - X(0, j, mz - 1);
      X(i2, j, k1) = X(i2, j, 0) + X(mx - 1, j, k1) - X(mx - 1, j, 0);
    } else {
      const int i = x0 - G, s = x1 + 1, tt = x2 + 1;
      const int j1 = -s, k1 = -tt, j2 = my - 1 + s, k2 = mz - 1 + tt;
      X(i, j1, k1) = X(i, j1, 0) + X(i, 0, k1) - X(i, 0, 0);
      X(i, j2, k2) = X(i, j2, mz - 1) + X(i, my - 1, k2) - X(i, my - 1, mz - 1);
      X(i, j1, k2) = X(i, j1, mz - 1) + X(i, 0, k2) - X(i, 0, mz - 1);
      X(i, j2, k1) = X(i, j2, 0) + X(i, my - 1, k1) - X(i, my - 1, 0);
    }
  }
}

namespace {
unsigned grid_of(long long n) { return (unsigned)((n + kTpb - 1) / kTpb); }
}  // namespace

int launch_coord_scatter(const double *raw, int mx, int my, int mz, double *xyz, hipStream_t st) {
  const long long ni = (long long)mx * my * mz;
  if (ni <= 0) return 0;
  coord_scatter_kernel<<<grid_of(ni), kTpb, 0, st>>>(raw, mx, my, mz, xyz);
  return (int)hipGetLastError();
}

int launch_ghost_extrapolate(double *xyz, int mx, int my, int mz, int dim, hipStream_t st) {
  const long long nl = (long long)my * mz + (long long)mz * mx + (dim == 3 ? (long long)mx * my : 0);
  ghost_extrapolate_kernel<<<grid_of(nl), kTpb, 0, st>>>(xyz, mx, my, mz, dim);
  if (hipError_t e = hipGetLastError()) return (int)e;
  if (dim == 2) {
    ghost_z2d_kernel<<<grid_of((long long)mx * my), kTpb, 0, st>>>(xyz, mx, my);
    return (int)hipGetLastError();
  }
  return 0;
}

int launch_ghost_inner(const GhostCopy &g, double *dst, int dmx, int dmy, int dmz, const double *src, int smx,
                       int smy, int smz, hipStream_t st) {
  constexpr int G = kNgg + 1;
  GhostCopyArgs A{};
  for (int q = 0; q < 3; ++q) {
    A.lo[q] = g.lo[q];
    A.n[q] = g.n[q];
    A.step[q] = g.step[q];
    A.t0[q] = g.t0[q];
    A.tstep[q] = g.tstep[q];
    A.src_tar[q] = g.src_tar[q];
  }
  A.dgx = dmx + 2 * G, A.dgy = dmy + 2 * G, A.dng = (long long)A.dgx * A.dgy * (dmz + 2 * G);
  A.sgx = smx + 2 * G, A.sgy = smy + 2 * G, A.sng = (long long)A.sgx * A.sgy * (smz + 2 * G);
  const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
  if (n <= 0) return 0;
  ghost_inner_kernel<<<grid_of(n), kTpb, 0, st>>>(dst, src, A);
  return (int)hipGetLastError();
}

int launch_ghost_corners(double *xyz, int mx, int my, int mz, hipStream_t st) {
  constexpr int G = kNgg + 1;
  ghost_corner_kernel<0><<<grid_of((long long)G * G * mz), kTpb, 0, st>>>(xyz, mx, my, mz);
  if (hipError_t e = hipGetLastError()) return (int)e;
  ghost_corner_kernel<1><<<grid_of((long long)G * my * G), kTpb, 0, st>>>(xyz, mx, my, mz);
  if (hipError_t e = hipGetLastError()) return (int)e;
  ghost_corner_kernel<2><<<grid_of((long long)(mx + kNgg + G + 1) * G * G), kTpb, 0, st>>>(xyz, mx, my, mz);
  return (int)hipGetLastError();
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
