// kernels.hip — HIP kernels of the explicit step for gfx950 and their host
// launchers.  Each kernel names the reference routine it replaces.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "device_physics.h"
#include "kernels.h"
#include "rhs_pencil.h"

namespace cfx {

namespace {

__device__ __forceinline__ long long pidx(const BlockView &B, int i, int j, int k) {
  return (long long)(k + kNgg) * B.sz + (long long)(j + kNgg) * B.sy + (i + kNgg);
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

// Blocks are dealt to the 8 XCDs round-robin; hand each XCD a contiguous
// run of tiles so that stencil neighbours share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nt) {
  if (nt % 8 != 0) return b;
  return (b % 8) * (nt / 8) + b / 8;
}

// ---------------------------------------------------------- total energy ---
// compute_total_energy (Thermo.cu:85-105) from bv/yk/cv at padded point p.
template <int NS>
__device__ __forceinline__ void total_energy(const BlockView &B, SpeciesPtrT<double> sp,
                                             const SchemeDev &S, long long p) {
  const long long np = B.np;
  const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
  double e = 0.5 * B.bv[p] * (u * u + v * v + w * w);
  if constexpr (NS > 0) {
    const double t = B.bv[5 * np + p];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      double h, cp;
      species_h_cp<double>(sp, l, t, h, cp);
      e += h * B.cv[(5 + l) * np + p];
    }
    e -= B.bv[4 * np + p];
  } else {
    e += B.bv[4 * np + p] / (S.gamma_air - 1);
  }
  B.cv[4 * np + p] = e;
}

// ------------------------------------------------------------ properties ---
// update_physical_properties (Field.cu:270-302) and
// compute_transport_property (Transport.cu:206-259) over the padded range.
// The Wilke sums use s_l = sqrt(mu_l) so that sqrt(mu_m/mu_n) = s_m/s_n
// (NS square roots instead of NS^2) and T*^-0.145 = exp(-0.145 log T*) with
// one shared log T.
template <int NS, typename R>
__global__ void __launch_bounds__(256) props_kernel(BlockView B, const SpeciesDev *__restrict__ spg, SchemeDev S) {
  const SpeciesPtrT<R> sp = const_table<R>(spg);
  const long long np = B.np;
  constexpr bool f32 = sizeof(R) == 4;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
    const R t = R(B.bv[5 * np + p]);
    R out[4 + (NS > 0 ? NS : 0)];
    if constexpr (NS == 0) {
      const R c_temp = R(S.gamma_air * S.r_air);
      out[0] = R(S.gamma_air);
      out[1] = sqrt(c_temp * t);
      const R mu = R(1.716e-5) * pow(t / R(273), R(1.5)) * R(273 + 111) / (t + R(111));
      out[2] = mu;
      out[3] = mu * c_temp / R(S.gamma_air - 1) * R(S.inv_pr);
    } else {
      R Y[NS], cp[NS];
      R winv = 0, cpt = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        Y[l] = R(B.yk[l * np + p]);
        cp[l] = species_cp<R>(sp, l, t);
        winv += Y[l] * R(sp->inv_mw[l]);
        cpt += Y[l] * cp[l];
      }
      const R Ru = R(kRu);
      const R gam = cpt / (cpt - Ru * winv);
      out[0] = gam;
      out[1] = sqrt(gam * Ru * t * winv);
      // Chapman-Enskog species viscosity (Transport.cu:215-220)
      const R lt = log(t), st = sqrt(t);
      R mu_s[NS], X[NS], s[NS], rs[NS];
      const R imw = R(1) / winv;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        X[l] = Y[l] * imw * R(sp->inv_mw[l]);
        const R tdl = t * R(sp->lj_inv[l]);
        const R a = tdl + R(0.5);
        const R omega = R(1.147) * exp(R(-0.145) * (lt + R(sp->log_lj_inv[l]))) + R(1) / (a * a);
        mu_s[l] = R(sp->vis_coeff[l]) * st / omega;
        s[l] = sqrt(mu_s[l]);
        rs[l] = R(1) / s[l];
      }
      // Wilke mixing and Mason-Saxena conductivity (Transport.cu:221-246)
      R mu = 0, kappa = 0;
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        R vt = X[m];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
          if (n == m) continue;
          const R q = R(1) + s[m] * rs[n] * R(sp->wj4[m][n]);
          vt += q * q * R(sp->sqw8[m][n]) * X[n];
        }
        const R lam = mu_s[m] * (cp[m] + R(sp->cond_add[m]));
        mu += mu_s[m] * X[m] / vt;
        kappa += lam * X[m] / (R(1.065) * vt - R(0.065) * X[m]);
      }
      out[2] = mu;
      out[3] = kappa;
      // constant-Schmidt diffusivity (Transport.cu:249-255, X[l] in the test)
      const R mus = mu * R(S.inv_sc);
#pragma unroll
      for (int l = 0; l < NS; ++l)
        out[4 + l] = fabs(X[l] - R(1)) < R(1e-3) ? mus : (R(1) - Y[l]) * mu / ((R(1) - X[l]) * R(S.sc));
    }
    if constexpr (f32) {
#pragma unroll
      for (int v = 0; v < 4 + NS; ++v) B.props32[v * np + p] = out[v];
      // FP32 copies of the primitives for the FP32 RHS
#pragma unroll
      for (int v = 0; v < 6; ++v) B.bv32[v * np + p] = float(B.bv[v * np + p]);
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk32[l * np + p] = float(B.yk[l * np + p]);
    } else {
#pragma unroll
      for (int v = 0; v < 4 + NS; ++v) B.props[v * np + p] = out[v];
    }
  }
}

// ------------------------------------------------------------------- RHS ---
// compute_inviscid_flux + compute_viscous_flux (TimeAdvanceFunc.cu:84-289)
// fused: one workgroup owns a TX x TY x TZ tile of cells; for each direction
// the face fluxes F = F_inviscid - F_viscous of the tile's (T+1) faces are
// computed by a flat loop over the workgroup (each face once), staged in LDS,
// and differenced into per-thread accumulators; dq is written once.
template <int NS, int TX, int TY, int TZ, int WPE>
__global__ void __launch_bounds__(TX *TY *TZ) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) rhs_kernel(BlockView B, const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                         int ntx, int nty, int ntz) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  constexpr int NV = 5 + NS;
  constexpr int NT = TX * TY * TZ;
  extern __shared__ double sf[];
  const int nt = ntx * nty * ntz;
  const int lid = xcd_tile(blockIdx.x, nt);
  const int i0 = (lid % ntx) * TX, j0 = ((lid / ntx) % nty) * TY, k0 = (lid / (ntx * nty)) * TZ;
  const int tid = threadIdx.x;
  const int ci = tid % TX, cj = (tid / TX) % TY, ck = tid / (TX * TY);
  const bool mine = (i0 + ci < B.mx) && (j0 + cj < B.my) && (k0 + ck < B.mz);
  const long long st[3] = {1, B.sy, B.sz};
  const int m[3] = {B.mx, B.my, B.mz};
  double acc[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) acc[l] = 0.0;
  for (int d = 0; d < S.dim; ++d) {
    const int fx = TX + (d == 0), fy = TY + (d == 1), fz = TZ + (d == 2);
    const int nf = fx * fy * fz;
    const int flo = ci + fx * (cj + fy * ck);
    const int fhi = flo + (d == 0 ? 1 : (d == 1 ? fx : fx * fy));
    // phase 0: inviscid faces, phase 1: viscous faces (separate live sets)
    for (int ph = 0; ph < (S.viscous == 2 ? 2 : 1); ++ph) {
      for (int f = tid; f < nf; f += NT) {
        const int c[3] = {i0 + f % fx, j0 + (f / fx) % fy, k0 + f / (fx * fy)};
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) ok = ok && (a == d ? c[a] <= m[a] : c[a] < m[a]);
        if (!ok) continue;
        const long long p = pidx(B, c[0] - (d == 0), c[1] - (d == 1), c[2] - (d == 2));
        double F[NV];
        if (ph == 0)
          inviscid_flux<NS, double>(B, sp, S, d, p, st[d], F);
        else
          viscous_flux<NS, double>(B, sp, S, d, p, st, F);
#pragma unroll
        for (int l = 0; l < NV; ++l) sf[l * nf + f] = F[l];
      }
      __syncthreads();
      if (mine) {
        const double sgn = ph == 0 ? -1.0 : 1.0;
#pragma unroll
        for (int l = 0; l < NV; ++l) acc[l] += sgn * (sf[l * nf + fhi] - sf[l * nf + flo]);
      }
      __syncthreads();
    }
  }
  if (mine) {
    const long long q = ((long long)(k0 + ck) * B.my + (j0 + cj)) * B.mx + (i0 + ci);
#pragma unroll
    for (int l = 0; l < NV; ++l) B.dq[l * B.ni + q] = acc[l];
  }
}

// ------------------------------------------------------------- time step ---
// SteadyTemporalScheme::compute_time_step (TemporalScheme.cu:10-43) and a
// wavefront min-reduction of the CFL time step into *dt_bits (positive
// doubles order like their bit patterns).
__global__ void __launch_bounds__(256) dt_kernel(BlockView B, SchemeDev S, int ns, int store_local,
                                                 unsigned long long *dt_bits) {
  __shared__ double red[4];
  const long long np = B.np;
  double dmin = 1e300;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < B.ni;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % B.mx);
    const int j = (int)((q / B.mx) % B.my);
    const int k = (int)(q / ((long long)B.mx * B.my));
    const long long p = pidx(B, i, j, k);
    double mm[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) mm[c] = B.metric[c * np + p];
    const double gx = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
    const double gy = sqrt(mm[3] * mm[3] + mm[4] * mm[4] + mm[5] * mm[5]);
    const double gz = sqrt(mm[6] * mm[6] + mm[7] * mm[7] + mm[8] * mm[8]);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
    const double U = u * mm[0] + v * mm[1] + w * mm[2];
    const double V = u * mm[3] + v * mm[4] + w * mm[5];
    const double W = u * mm[6] + v * mm[7] + w * mm[8];
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
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmin(r, red[w]);
    atomicMin(dt_bits, (unsigned long long)__double_as_longlong(r));
  }
}

__global__ void set_u64_kernel(unsigned long long *dst, unsigned long long v) { *dst = v; }

// --------------------------------------------------------------- update ---
// update_cv_and_bv (TimeAdvanceFunc.cu:301-336) with the RK stage
// combination q = a*qn + b*(q + dt*dq/J) and the Newton inversion of T
// (Thermo.cu:108-144, stop |dT| <= 1e-10 K unless CFX_Q_NEWTON_REL).
template <int NS>
__global__ void __launch_bounds__(256) update_kernel(BlockView B, const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                     int stage, const double *__restrict__ dt_dev, int *flags,
                                                     unsigned long long *newton_fail) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  constexpr int NV = 5 + NS;
  const long long np = B.np, ni = B.ni;
  const double A = stage == 1 ? 0.75 : (stage == 2 ? 1.0 / 3.0 : 0.0);
  const double Bc = stage == 1 ? 0.25 : (stage == 2 ? 2.0 / 3.0 : 1.0);
  const double dt = *dt_dev;
  const bool rel = (S.quirks & CFX_Q_NEWTON_REL) != 0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < ni; q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % B.mx);
    const int j = (int)((q / B.mx) % B.my);
    const int k = (int)(q / ((long long)B.mx * B.my));
    const long long p = pidx(B, i, j, k);
    const double dtj = (S.rk == 3 ? dt : B.dt_local[q]) / B.jac[p];
    double c[NV];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const double cur = B.cv[l * np + p];
      const double dq = B.mixed ? (double)B.dq32[l * ni + q] : B.dq[l * ni + q];
      if (S.rk == 3) {
        double qn;
        if (stage == 0) {
          qn = cur;
          B.qn[l * ni + q] = cur;
        } else {
          qn = B.qn[l * ni + q];
        }
        c[l] = A * qn + Bc * (cur + dq * dtj);
      } else {
        c[l] = cur + dq * dtj;
      }
    }
    if (B.mz == 1) c[3] = 0.0;
#pragma unroll
    for (int l = 0; l < NV; ++l) B.cv[l * np + p] = c[l];
    const double rho = c[0];
    const double ir = 1.0 / rho;
    const double u = c[1] * ir, v = c[2] * ir, w = c[3] * ir;
    const double v2 = u * u + v * v + w * w;
    double T, pr;
    if constexpr (NS > 0) {
      double Y[NS];
      double winv = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        Y[l] = c[5 + l] * ir;
        B.yk[l * np + p] = Y[l];
        winv += Y[l] * sp->inv_mw[l];
      }
      const double R = kRu * winv;
      const double e = c[4] / rho - 0.5 * v2;
      T = B.bv[5 * np + p];
      double err = 1e30;
      int it = 0;
      const double eps = rel ? 1e-3 : 1e-10;
      while (err > eps && it < 1000) {
        ++it;
        double h = 0, cp = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          double hl, cpl;
          species_h_cp<double>(sp, l, T, hl, cpl);
          h += hl * Y[l];
          cp += cpl * Y[l];
        }
        const double T1 = T - (h - R * T - e) / (cp - R);
        err = rel ? fabs(1 - T1 / T) : fabs(T1 - T);
        T = T1;
      }
      if (err > eps) atomicAdd(newton_fail, 1ull);
      pr = rho * T * R;
    } else {
      pr = (S.gamma_air - 1) * (c[4] - 0.5 * rho * v2);
      T = pr * ir / S.r_air;
    }
    B.bv[p] = rho;
    B.bv[np + p] = u;
    B.bv[2 * np + p] = v;
    B.bv[3 * np + p] = w;
    B.bv[4 * np + p] = pr;
    B.bv[5 * np + p] = T;
    if (!(rho > 0.0) || !(T > 0.0) || !(pr > 0.0) || !isfinite(rho) || !isfinite(T)) atomicOr(flags, 1);
  }
}

// ------------------------------------------------------------ init / last ---
// compute_cv_from_bv (Field.cu:242-268) over the padded range
template <int NS>
__global__ void cv_from_bv_kernel(BlockView B, const SpeciesDev *__restrict__ spg, SchemeDev S) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  const long long np = B.np;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
    const double rho = B.bv[p];
    B.cv[p] = rho;
    B.cv[np + p] = rho * B.bv[np + p];
    B.cv[2 * np + p] = rho * B.bv[2 * np + p];
    B.cv[3 * np + p] = rho * B.bv[3 * np + p];
#pragma unroll
    for (int l = 0; l < NS; ++l) B.cv[(5 + l) * np + p] = rho * B.yk[l * np + p];
    total_energy<NS>(B, sp, S, p);
  }
}

// store_last_step (TimeAdvanceFunc.cu:59-70)
__global__ void store_last_kernel(BlockView B) {
  const long long np = B.np, ni = B.ni;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < ni; q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % B.mx);
    const int j = (int)((q / B.mx) % B.my);
    const int k = (int)(q / ((long long)B.mx * B.my));
    const long long p = pidx(B, i, j, k);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
    B.bv_last[q] = B.bv[p];
    B.bv_last[ni + q] = sqrt(u * u + v * v + w * w);
    B.bv_last[2 * ni + q] = B.bv[4 * np + p];
    B.bv_last[3 * ni + q] = B.bv[5 * np + p];
  }
}

// compute_square_of_dbv + reduction_of_dbv_squared (TimeAdvanceFunc.cu:338-349,
// Driver.cu:383-420): per-workgroup partial sums, deterministic order.
__global__ void __launch_bounds__(256) residual_partial_kernel(BlockView B, double *partial) {
  __shared__ double red[4][4];
  const long long np = B.np, ni = B.ni;
  double s[4] = {0, 0, 0, 0};
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < ni; q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % B.mx);
    const int j = (int)((q / B.mx) % B.my);
    const int k = (int)(q / ((long long)B.mx * B.my));
    const long long p = pidx(B, i, j, k);
    const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
    const double cur[4] = {B.bv[p], sqrt(u * u + v * v + w * w), B.bv[4 * np + p], B.bv[5 * np + p]};
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
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w][threadIdx.x];
    partial[blockIdx.x * 4 + threadIdx.x] = r;
  }
}

__global__ void __launch_bounds__(256) residual_final_kernel(const double *partial, int n, double *out) {
  __shared__ double red[4][4];
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
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w][threadIdx.x];
    out[threadIdx.x] = r;
  }
}

// ------------------------------------------------------------ boundaries ---
__device__ __forceinline__ bool face_point(const FaceRange &f, int &i, int &j, int &k) {
  i = f.rs[0] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  j = f.rs[1] + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  k = f.rs[2] + (int)(blockIdx.z * blockDim.z + threadIdx.z);
  return i <= f.re[0] && j <= f.re[1] && k <= f.re[2];
}

// apply_outflow (BoundCond.cu:455-487)
__global__ void bc_outflow_kernel(BlockView B, FaceRange f, int nv, int ns) {
  int i, j, k;
  if (!face_point(f, i, j, k)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  for (int g = 1; g <= kNgg; ++g) {
    const long long q = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
    for (int l = 0; l < 6; ++l) B.bv[l * np + q] = B.bv[l * np + p];
    for (int l = 0; l < nv; ++l) B.cv[l * np + q] = B.cv[l * np + p];
    for (int l = 0; l < ns; ++l) B.yk[l * np + q] = B.yk[l * np + p];
  }
}

// apply_inflow (BoundCond.cu:489-527); the state's total energy is
// precomputed on the host with the same formula as compute_total_energy.
__global__ void bc_inflow_kernel(BlockView B, FaceRange f, int ns, const InflowDev *__restrict__ in) {
  int i, j, k;
  if (!face_point(f, i, j, k)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const double rho = in->st[0];
  for (int g = 1; g <= kNgg; ++g) {
    const long long q = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
    for (int l = 0; l < 6; ++l) B.bv[l * np + q] = in->st[l];
    for (int l = 0; l < ns; ++l) {
      B.yk[l * np + q] = in->y[l];
      B.cv[(5 + l) * np + q] = rho * in->y[l];
    }
    B.cv[q] = rho;
    B.cv[np + q] = rho * in->st[1];
    B.cv[2 * np + q] = rho * in->st[2];
    B.cv[3 * np + q] = rho * in->st[3];
    B.cv[4 * np + q] = in->st[6];
  }
}

template <int NS>
__device__ __forceinline__ double mix_mw(const BlockView &B, SpeciesPtrT<double> sp, const SchemeDev &S,
                                         long long p) {
  if constexpr (NS == 0) {
    return kRu / S.r_air;
  } else {
    double w = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) w += B.yk[l * B.np + p] * sp->inv_mw[l];
    return 1.0 / w;
  }
}

// apply_wall (BoundCond.cu:529-629) over the face's own vertex range
// (the registered range shrunk back by ngg, BoundCond.cu:229-239).
template <int NS>
__global__ void bc_wall_kernel(BlockView B, FaceRange f, const SpeciesDev *__restrict__ spg, SchemeDev S,
                               WallDev wl) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  int i, j, k;
  if (!face_point(f, i, j, k)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  const long long pin = pidx(B, i - dir[0], j - dir[1], k - dir[2]);
  double t_wall = wl.temperature;
  if (wl.adiabatic) t_wall = B.bv[5 * np + pin];
  const double pw = B.bv[4 * np + pin];
#pragma unroll
  for (int l = 0; l < NS; ++l) B.yk[l * np + p] = B.yk[l * np + pin];
  double mw = mix_mw<NS>(B, sp, S, p);
  B.bv[p] = pw * mw / (t_wall * kRu);
  B.bv[np + p] = 0;
  B.bv[2 * np + p] = 0;
  B.bv[3 * np + p] = 0;
  B.bv[4 * np + p] = pw;
  B.bv[5 * np + p] = t_wall;
  B.cv[p] = B.bv[p];
  B.cv[np + p] = 0;
  B.cv[2 * np + p] = 0;
  B.cv[3 * np + p] = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) B.cv[(l + 5) * np + p] = B.yk[l * np + p] * B.bv[p];
  total_energy<NS>(B, sp, S, p);
  for (int g = 1; g <= kNgg; ++g) {
    const long long qi = pidx(B, i - g * dir[0], j - g * dir[1], k - g * dir[2]);
    const long long qg = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
#pragma unroll
    for (int l = 0; l < NS; ++l) B.yk[l * np + qg] = B.yk[l * np + qi];
    mw = mix_mw<NS>(B, sp, S, qg);
    const double u_i = B.bv[np + qi], v_i = B.bv[2 * np + qi], w_i = B.bv[3 * np + qi];
    const double p_i = B.bv[4 * np + qi], t_i = B.bv[5 * np + qi];
    double t_g = t_i;
    if (!wl.adiabatic) {
      t_g = 2 * t_wall - t_i;
      if (t_g <= 0.1 * t_wall) t_g = t_wall;
    }
    const double rho_g = p_i * mw / (t_g * kRu);
    B.bv[qg] = rho_g;
    B.bv[np + qg] = -u_i;
    B.bv[2 * np + qg] = -v_i;
    B.bv[3 * np + qg] = -w_i;
    B.bv[4 * np + qg] = p_i;
    B.bv[5 * np + qg] = t_g;
    B.cv[qg] = rho_g;
    B.cv[np + qg] = -rho_g * u_i;
    B.cv[2 * np + qg] = -rho_g * v_i;
    B.cv[3 * np + qg] = -rho_g * w_i;
#pragma unroll
    for (int l = 0; l < NS; ++l) B.cv[(l + 5) * np + qg] = B.yk[l * np + qg] * rho_g;
    total_energy<NS>(B, sp, S, qg);
  }
}

// Slip wall: the boundary point loses its velocity along the face normal
// n = grad(xi_face)/|grad(xi_face)|; ghost g mirrors interior point g with
// the normal velocity negated and rho, p, T, Y copied.
template <int NS>
__global__ void bc_symmetry_kernel(BlockView B, FaceRange f, const SpeciesDev *__restrict__ spg, SchemeDev S) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  int i, j, k;
  if (!face_point(f, i, j, k)) return;
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  double n[3];
  for (int c = 0; c < 3; ++c) n[c] = B.metric[(f.face * 3 + c) * np + p];
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= nn;
  n[1] /= nn;
  n[2] /= nn;
  {
    const double vn = B.bv[np + p] * n[0] + B.bv[2 * np + p] * n[1] + B.bv[3 * np + p] * n[2];
    for (int c = 0; c < 3; ++c) B.bv[(1 + c) * np + p] -= vn * n[c];
    for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + p] = B.bv[p] * B.bv[(1 + c) * np + p];
    total_energy<NS>(B, sp, S, p);
  }
  for (int g = 1; g <= kNgg; ++g) {
    const long long qi = pidx(B, i - g * dir[0], j - g * dir[1], k - g * dir[2]);
    const long long qg = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
    const double vn = B.bv[np + qi] * n[0] + B.bv[2 * np + qi] * n[1] + B.bv[3 * np + qi] * n[2];
    B.bv[qg] = B.bv[qi];
    for (int c = 0; c < 3; ++c) B.bv[(1 + c) * np + qg] = B.bv[(1 + c) * np + qi] - 2 * vn * n[c];
    B.bv[4 * np + qg] = B.bv[4 * np + qi];
    B.bv[5 * np + qg] = B.bv[5 * np + qi];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      B.yk[l * np + qg] = B.yk[l * np + qi];
      B.cv[(5 + l) * np + qg] = B.bv[qg] * B.yk[l * np + qg];
    }
    B.cv[qg] = B.bv[qg];
    for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + qg] = B.bv[qg] * B.bv[(1 + c) * np + qg];
    total_energy<NS>(B, sp, S, qg);
  }
}

// inner_communication (Field.cu:305-358) with the face loop of
// Mesh.cpp:905-917: shared-face points are averaged between the two blocks,
// ghost layers take the neighbour's values.
__global__ void exchange_kernel(BlockView B, BlockView T, InnerDev f, int nv, int ns) {
  const int n[3] = {(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(blockIdx.y * blockDim.y + threadIdx.y),
                    (int)(blockIdx.z * blockDim.z + threadIdx.z)};
  if (n[0] >= f.n_point[0] || n[1] >= f.n_point[1] || n[2] >= f.n_point[2]) return;
  int idx[3], d_idx[3], it[3];
  for (int q = 0; q < 3; ++q) {
    d_idx[q] = f.loop_dir[q] * n[q];
    idx[q] = f.rs[q] + d_idx[q];
  }
  for (int q = 0; q < 3; ++q) it[q] = f.ts[q] + f.target_loop_dir[q] * d_idx[f.src_tar[q]];
  const long long p = pidx(B, idx[0], idx[1], idx[2]);
  const long long t = pidx(T, it[0], it[1], it[2]);
  const long long np = B.np, tnp = T.np;
  const int face_layer = f.direction > 0 ? f.rs[f.face] : f.re[f.face];
  if (idx[f.face] == face_layer) {
    for (int l = 0; l < 6; ++l) {
      const double a = 0.5 * (T.bv[l * tnp + t] + B.bv[l * np + p]);
      B.bv[l * np + p] = a;
      T.bv[l * tnp + t] = a;
    }
    for (int l = 0; l < nv; ++l) {
      const double a = 0.5 * (T.cv[l * tnp + t] + B.cv[l * np + p]);
      B.cv[l * np + p] = a;
      T.cv[l * tnp + t] = a;
    }
    for (int l = 0; l < ns; ++l) {
      const double a = 0.5 * (T.yk[l * tnp + t] + B.yk[l * np + p]);
      B.yk[l * np + p] = a;
      T.yk[l * tnp + t] = a;
    }
  } else {
    for (int l = 0; l < 6; ++l) B.bv[l * np + p] = T.bv[l * tnp + t];
    for (int l = 0; l < 5; ++l) B.cv[l * np + p] = T.cv[l * tnp + t];
    for (int l = 0; l < ns; ++l) {
      B.yk[l * np + p] = T.yk[l * tnp + t];
      B.cv[(5 + l) * np + p] = T.cv[(5 + l) * tnp + t];
    }
  }
}

// eliminate_k_gradient (Field.cu:360-384), 2-D only
__global__ void eliminate_k_kernel(BlockView B, int ns) {
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
    for (int l = 0; l < ns; ++l) {
      B.yk[l * np + pp] = B.yk[l * np + p0];
      B.yk[l * np + pm] = B.yk[l * np + p0];
    }
  }
}

int grid_for(long long n, int tpb) {
  long long b = (n + tpb - 1) / tpb;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return (int)(b < 1 ? 1 : b);
}

dim3 face_grid(const FaceRange &f, dim3 &tpb) {
  unsigned t[3], g[3];
  for (int q = 0; q < 3; ++q) {
    const int n = f.re[q] - f.rs[q] + 1;
    t[q] = n <= 2 * kNgg + 1 ? 1 : 16;
    g[q] = (unsigned)((n - 1) / (int)t[q] + 1);
  }
  if (t[0] * t[1] * t[2] > 256) t[2] = 1, g[2] = (unsigned)(f.re[2] - f.rs[2] + 1);
  tpb = dim3(t[0], t[1], t[2]);
  return dim3(g[0], g[1], g[2]);
}

}  // namespace

// ------------------------------------------------------------- launchers ---
#define CFX_SWITCH_NS(ns, ...)                             \
  switch (ns) {                                            \
    case 0: { constexpr int NS_ = 0; __VA_ARGS__; } break; \
    case 1: { constexpr int NS_ = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int NS_ = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int NS_ = 3; __VA_ARGS__; } break; \
    case 5: { constexpr int NS_ = 5; __VA_ARGS__; } break; \
    case 9: { constexpr int NS_ = 9; __VA_ARGS__; } break; \
    default: return CFX_E_UNSUPPORTED;                     \
  }

bool ns_supported(int ns) { return ns == 0 || ns == 1 || ns == 2 || ns == 3 || ns == 5 || ns == 9; }

int launch_props(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, hipStream_t st) {
  const int g = grid_for(B.np, 256);
  if (B.mixed) {
    CFX_SWITCH_NS(ns, props_kernel<NS_, float><<<g, 256, 0, st>>>(B, sp, S));
  } else {
    CFX_SWITCH_NS(ns, props_kernel<NS_, double><<<g, 256, 0, st>>>(B, sp, S));
  }
  return (int)hipGetLastError();
}

// RHS variants for the autotuner.  Pencil variants (one LDS-staged kernel
// per face direction, tile TA x TB for each direction) come first; the fused
// all-direction kernel with its tile shape and target occupancy follows.
struct FusedVariant {
  int tx, ty, tz, wpe;
};
// pencil variants: per face direction d the tile ta[d] x tb[d] x tc[d]
struct PencilVariant {
  int ta[3], tb[3], tc[3], wpe;
};
// (variant 0 and 2/3 visit tiles in 4x4 tangential groups, variant 1 in
// plain order)
static const PencilVariant kPencil3[] = {{{32, 16, 16}, {4, 8, 8}, {2, 2, 2}, 2},
                                         {{32, 16, 16}, {4, 8, 8}, {2, 2, 2}, 2},
                                         {{32, 16, 16}, {4, 4, 4}, {4, 4, 4}, 2},
                                         {{64, 32, 32}, {2, 4, 4}, {2, 2, 2}, 2}};
static const PencilVariant kPencil2[] = {{{32, 16, 0}, {8, 16, 0}, {1, 1, 0}, 2},
                                         {{64, 32, 0}, {4, 8, 0}, {1, 1, 0}, 2},
                                         {{32, 16, 0}, {8, 16, 0}, {1, 1, 0}, 2},
                                         {{32, 16, 0}, {8, 16, 0}, {1, 1, 0}, 1}};
static constexpr int kNPencil = 4;  // slots reserved for pencil variants in both dimensions
static const FusedVariant kRhs3[] = {{16, 4, 4, 1}, {8, 8, 4, 1}, {32, 4, 2, 1}, {16, 8, 2, 1}, {64, 2, 2, 1},
                                     {16, 4, 4, 2}, {8, 8, 4, 2}, {32, 4, 2, 2}, {16, 4, 4, 3}, {8, 8, 2, 2}};
static const FusedVariant kRhs2[] = {{32, 8, 1, 1}, {16, 16, 1, 1}, {64, 4, 1, 1}, {32, 8, 1, 2}, {16, 16, 1, 2}};

int n_rhs_tiles(int dim) {
  return kNPencil + (dim == 3 ? (int)(sizeof(kRhs3) / sizeof(kRhs3[0])) : (int)(sizeof(kRhs2) / sizeof(kRhs2[0])));
}

int rhs_tile(int dim, int tile, int &tx, int &ty, int &tz) {
  if (tile < 0 || tile >= n_rhs_tiles(dim)) return -1;
  if (tile < kNPencil) {
    const PencilVariant &v = dim == 3 ? kPencil3[tile] : kPencil2[tile];
    tx = v.ta[0], ty = v.tb[0], tz = v.tc[0];
    return 0;
  }
  const FusedVariant &v = dim == 3 ? kRhs3[tile - kNPencil] : kRhs2[tile - kNPencil];
  tx = v.tx, ty = v.ty, tz = v.tz;
  return 0;
}

template <int NS, int TX, int TY, int TZ, int WPE>
static int launch_rhs_t(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st) {
  const int ntx = (B.mx + TX - 1) / TX, nty = (B.my + TY - 1) / TY, ntz = (B.mz + TZ - 1) / TZ;
  constexpr int fxn = (TX + 1) * TY * TZ, fyn = TX * (TY + 1) * TZ, fzn = TX * TY * (TZ + 1);
  constexpr int maxf = fxn > fyn ? (fxn > fzn ? fxn : fzn) : (fyn > fzn ? fyn : fzn);
  const size_t lds = sizeof(double) * (5 + NS) * maxf;
  rhs_kernel<NS, TX, TY, TZ, WPE><<<ntx * nty * ntz, TX * TY * TZ, lds, st>>>(B, sp, S, ntx, nty, ntz);
  return (int)hipGetLastError();
}

template <int NS, int DIR, int TA, int TB, int TC, int WPE, typename R, int G>
static int launch_pencil_t(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int accumulate,
                           hipStream_t st) {
  constexpr int A = Axes<DIR>::a, Bx = Axes<DIR>::b, Cx = Axes<DIR>::c;
  const int m[3] = {B.mx, B.my, B.mz};
  const int nta = (m[A] + TA - 2) / (TA - 1);  // TA faces cover TA-1 cells
  const int ntb = (m[Bx] + TB - 1) / TB;
  const int ntc = (m[Cx] + TC - 1) / TC;
  const int ngrid = nta * G * G * ((ntb + G - 1) / G) * ((ntc + G - 1) / G);
  rhs_pencil_kernel<NS, DIR, TA, TB, TC, WPE, R, G>
      <<<ngrid, TA * TB * TC, 0, st>>>(B, sp, S, accumulate, nta, ntb, ntc);
  return (int)hipGetLastError();
}

template <int NS, int TA0, int TB0, int TC0, int TA1, int TB1, int TC1, int TA2, int TB2, int TC2, int WPE,
          typename R, int G>
static int launch_pencils_r(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st) {
  int rc = launch_pencil_t<NS, 0, TA0, TB0, TC0, WPE, R, G>(B, sp, S, 0, st);
  if (rc) return rc;
  rc = launch_pencil_t<NS, 1, TA1, TB1, TC1, WPE, R, G>(B, sp, S, 1, st);
  if (rc || S.dim == 2) return rc;
  return launch_pencil_t<NS, 2, TA2, TB2, TC2, WPE, R, G>(B, sp, S, 1, st);
}

// FP64 with the variant's target occupancy; FP32 (mixed precision) needs
// half the registers and targets one more wave per SIMD
template <int NS, int TA0, int TB0, int TC0, int TA1, int TB1, int TC1, int TA2, int TB2, int TC2, int WPE,
          int G = 1>
static int launch_pencils(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st) {
  if (B.mixed)
    return launch_pencils_r<NS, TA0, TB0, TC0, TA1, TB1, TC1, TA2, TB2, TC2, WPE + 1, float, G>(B, sp, S, st);
  return launch_pencils_r<NS, TA0, TB0, TC0, TA1, TB1, TC1, TA2, TB2, TC2, WPE, double, G>(B, sp, S, st);
}

template <int NS>
static int launch_rhs_ns(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int tile, hipStream_t st) {
  if (S.dim == 3) {
    switch (tile) {
      case 0: return launch_pencils<NS, 32, 4, 2, 16, 8, 2, 16, 8, 2, 2, 4>(B, sp, S, st);
      case 1: return launch_pencils<NS, 32, 4, 2, 16, 8, 2, 16, 8, 2, 2, 1>(B, sp, S, st);
      case 2: return launch_pencils<NS, 32, 4, 4, 16, 4, 4, 16, 4, 4, 2, 4>(B, sp, S, st);
      case 3: return launch_pencils<NS, 64, 2, 2, 32, 4, 2, 32, 4, 2, 2, 4>(B, sp, S, st);
      default: break;
    }
  } else {
    switch (tile) {
      case 0: return launch_pencils<NS, 32, 8, 1, 16, 16, 1, 16, 16, 1, 2, 4>(B, sp, S, st);
      case 1: return launch_pencils<NS, 64, 4, 1, 32, 8, 1, 32, 8, 1, 2, 4>(B, sp, S, st);
      case 2: return launch_pencils<NS, 32, 8, 1, 16, 16, 1, 16, 16, 1, 2, 1>(B, sp, S, st);
      case 3: return launch_pencils<NS, 32, 8, 1, 16, 16, 1, 16, 16, 1, 1, 4>(B, sp, S, st);
      default: break;
    }
  }
  tile -= kNPencil;
  if (B.mixed) return CFX_E_UNSUPPORTED;  // the fused kernel is FP64 only
  if (S.dim == 3) {
    switch (tile) {
      case 0: return launch_rhs_t<NS, 16, 4, 4, 1>(B, sp, S, st);
      case 1: return launch_rhs_t<NS, 8, 8, 4, 1>(B, sp, S, st);
      case 2: return launch_rhs_t<NS, 32, 4, 2, 1>(B, sp, S, st);
      case 3: return launch_rhs_t<NS, 16, 8, 2, 1>(B, sp, S, st);
      case 4: return launch_rhs_t<NS, 64, 2, 2, 1>(B, sp, S, st);
      case 5: return launch_rhs_t<NS, 16, 4, 4, 2>(B, sp, S, st);
      case 6: return launch_rhs_t<NS, 8, 8, 4, 2>(B, sp, S, st);
      case 7: return launch_rhs_t<NS, 32, 4, 2, 2>(B, sp, S, st);
      case 8: return launch_rhs_t<NS, 16, 4, 4, 3>(B, sp, S, st);
      case 9: return launch_rhs_t<NS, 8, 8, 2, 2>(B, sp, S, st);
      default: return CFX_E_ARG;
    }
  }
  switch (tile) {
    case 0: return launch_rhs_t<NS, 32, 8, 1, 1>(B, sp, S, st);
    case 1: return launch_rhs_t<NS, 16, 16, 1, 1>(B, sp, S, st);
    case 2: return launch_rhs_t<NS, 64, 4, 1, 1>(B, sp, S, st);
    case 3: return launch_rhs_t<NS, 32, 8, 1, 2>(B, sp, S, st);
    case 4: return launch_rhs_t<NS, 16, 16, 1, 2>(B, sp, S, st);
    default: return CFX_E_ARG;
  }
}

int launch_rhs(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, int tile, hipStream_t st) {
  CFX_SWITCH_NS(ns, return launch_rhs_ns<NS_>(B, sp, S, tile, st));
  return 0;
}

int launch_dt_reset(unsigned long long *dt_bits, hipStream_t st) {
  set_u64_kernel<<<1, 1, 0, st>>>(dt_bits, 0x7FF0000000000000ull);
  return (int)hipGetLastError();
}

int launch_dt(const BlockView &B, const SchemeDev &S, int ns, int store_local, unsigned long long *dt_bits,
              hipStream_t st) {
  const int g = grid_for(B.ni, 256) > 2048 ? 2048 : grid_for(B.ni, 256);
  dt_kernel<<<g, 256, 0, st>>>(B, S, ns, store_local, dt_bits);
  return (int)hipGetLastError();
}

int launch_update(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, int stage, const double *dt,
                  int *flags, unsigned long long *newton_fail, hipStream_t st) {
  const int g = grid_for(B.ni, 256);
  CFX_SWITCH_NS(ns, update_kernel<NS_><<<g, 256, 0, st>>>(B, sp, S, stage, dt, flags, newton_fail));
  return (int)hipGetLastError();
}

int launch_cv_from_bv(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, hipStream_t st) {
  const int g = grid_for(B.np, 256);
  CFX_SWITCH_NS(ns, cv_from_bv_kernel<NS_><<<g, 256, 0, st>>>(B, sp, S));
  return (int)hipGetLastError();
}

int launch_store_last(const BlockView &B, hipStream_t st) {
  store_last_kernel<<<grid_for(B.ni, 256), 256, 0, st>>>(B);
  return (int)hipGetLastError();
}

int residual_partials(const BlockView &B) {
  const int g = grid_for(B.ni, 256);
  return g > 1024 ? 1024 : g;
}

int launch_residual_partial(const BlockView &B, double *partial, hipStream_t st) {
  residual_partial_kernel<<<residual_partials(B), 256, 0, st>>>(B, partial);
  return (int)hipGetLastError();
}

int launch_residual_final(const double *partial, int n, double *out, hipStream_t st) {
  residual_final_kernel<<<1, 256, 0, st>>>(partial, n, out);
  return (int)hipGetLastError();
}

int launch_bc(int type, const BlockView &B, const FaceRange &f, const SpeciesDev *sp, const SchemeDev &S, int ns,
              const InflowDev *inflow, const WallDev &wall, hipStream_t st) {
  dim3 tpb;
  FaceRange r = f;
  if (type == CFX_BC_WALL) {
    for (int q = 0; q < 3; ++q) {
      if (q == f.face) continue;
      r.rs[q] += kNgg;
      r.re[q] -= kNgg;
    }
    if (r.rs[0] > r.re[0] || r.rs[1] > r.re[1] || r.rs[2] > r.re[2]) return 0;
  }
  const dim3 g = face_grid(r, tpb);
  switch (type) {
    case CFX_BC_OUTFLOW: bc_outflow_kernel<<<g, tpb, 0, st>>>(B, r, 5 + ns, ns); break;
    case CFX_BC_INFLOW: bc_inflow_kernel<<<g, tpb, 0, st>>>(B, r, ns, inflow); break;
    case CFX_BC_WALL: CFX_SWITCH_NS(ns, bc_wall_kernel<NS_><<<g, tpb, 0, st>>>(B, r, sp, S, wall)); break;
    case CFX_BC_SYMMETRY: CFX_SWITCH_NS(ns, bc_symmetry_kernel<NS_><<<g, tpb, 0, st>>>(B, r, sp, S)); break;
    default: return CFX_E_ARG;
  }
  return (int)hipGetLastError();
}

int launch_exchange(const BlockView &B, const BlockView &T, const InnerDev &f, int ns, hipStream_t st) {
  unsigned t[3], g[3];
  for (int q = 0; q < 3; ++q) {
    t[q] = f.n_point[q] <= 2 * kNgg + 1 ? 1 : 16;
    g[q] = (unsigned)((f.n_point[q] - 1) / (int)t[q] + 1);
  }
  exchange_kernel<<<dim3(g[0], g[1], g[2]), dim3(t[0], t[1], t[2]), 0, st>>>(B, T, f, 5 + ns, ns);
  return (int)hipGetLastError();
}

int launch_eliminate_k(const BlockView &B, int ns, hipStream_t st) {
  const dim3 tpb(16, 16, 1);
  const dim3 g((B.px + 15) / 16, (B.py + 15) / 16, 1);
  eliminate_k_kernel<<<g, tpb, 0, st>>>(B, ns);
  return (int)hipGetLastError();
}

}  // namespace cfx
