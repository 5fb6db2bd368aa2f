// kernels_ns.hip — the kernels templated on the species count NS, compiled
// once per supported count (-DCFX_NS=n) and precision of the flux work
// (-DCFX_PREC=64: the FP64 set with the update, conversions and ghost fills;
// -DCFX_PREC=32: the FP32 properties and RHS of the mixed-precision path).
// Each build registers its launchers in g_kernels64[n] / g_kernels32[n].
#include <hip/hip_runtime.h>

#include <atomic>

#include "device_physics.h"
#include "kernels.h"
#include "mlp.h"
#include "rhs_pencil.h"

#ifndef CFX_NS
#error "CFX_NS (species count) must be defined"
#endif
#ifndef CFX_PREC
#error "CFX_PREC (64 or 32) must be defined"
#endif

namespace cfx {

namespace {

constexpr int NS = CFX_NS;
constexpr int NSA = NS > 0 ? NS : 1;  // extent of the per-species arrays
static_assert(NS >= 0 && NS <= kMaxSp, "species count out of range");

constexpr PencilTile tile_of(bool d3, int dir, int v) {
  return d3 ? kTiles3[dir][v] : kTiles2[dir < 2 ? dir : 0][v < kNVar2 ? v : 0];
}

__device__ __forceinline__ long long pidx(const GView &B, int i, int j, int k) {
  return (long long)(k + kNgg) * B.sz + (long long)(j + kNgg) * B.sy + (i + kNgg);
}

__device__ __forceinline__ long long interior_pidx(const GView &B, long long q) {
  const unsigned u = (unsigned)q, mx = (unsigned)B.mx, my = (unsigned)B.my;
  const unsigned i = u % mx, r = u / mx;
  return pidx(B, (int)i, (int)(r % my), (int)(r / my));
}

// Mass fraction l at padded point p.  The mixed-precision path keeps no FP64
// copy of Y after the initial upload: there Y_l = (rho Y_l) * (1 / rho) from
// the conserved state, in the update kernel's operation order, which gives
// at every updated point the value the update would have stored.
__device__ __forceinline__ double mass_fraction(const GView &B, long long p, int l) {
  return B.mixed ? B.cv[(5 + l) * B.np + p] * (1.0 / B.cv[p]) : B.yk[l * B.np + p];
}

// Point kernels: workgroup -> (block, first point of its kTpb chunk).
struct PointWork {
  int blk;
  long long base;
};
__device__ __forceinline__ PointWork point_work(const int *__restrict__ off, int nb) {
  const int w = blockIdx.x + off[0];
  const int blk = find_seg(off, nb, w);
  return {blk, (long long)(w - off[blk]) * kTpb};
}

// ------------------------------------------------------------ properties ---
// update_physical_properties (Field.cu:178-210): mixture cp, gamma and speed
// of sound at temperature t and mass fractions Y; cp_l and sum Y_l/W_l are
// returned for the transport.
template <typename R>
__device__ __forceinline__ void mixture_gamma_c(SpeciesPtrT<R> sp, R t, const R *Y, R *cp, R &winv, R *out) {
  R cpt = 0;
  winv = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    cp[l] = species_cp<R>(sp, l, t);
    winv += Y[l] * R(sp->inv_mw[l]);
    cpt += Y[l] * cp[l];
  }
  const R Ru = R(kRu);
  const R gam = fdiv<R>(cpt, cpt - Ru * winv);
  out[0] = gam;
  out[1] = fsqrt<R>(gam * Ru * t * winv);
}

// compute_transport_property (Transport.cu:43-96): out = mu, kappa,
// rho*D_1..ns.  The Wilke sums use s_l = sqrt(mu_l) so that sqrt(mu_m/mu_n)
// = s_m/s_n (NS square roots instead of NS^2), and T*^-0.145 =
// T^-0.145 (k/eps_l)^-0.145: one exponential per point, the species factor
// from the table.
template <typename R>
__device__ __forceinline__ void wilke_point(SpeciesPtrT<R> sp, const SchemeDev &S, R t, const R *Y, const R *cp,
                                            R winv, R *out) {
  // Chapman-Enskog species viscosity (Transport.cu:52-57)
  const R tm = fexp<R>(R(-0.145) * flog<R>(t)), st = fsqrt<R>(t);
  R mu_s[NSA], X[NSA], s[NSA], rs[NSA];
  const R imw = frcp<R>(winv);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    X[l] = Y[l] * imw * R(sp->inv_mw[l]);
    const R tdl = t * R(sp->lj_inv[l]);
    const R a = tdl + R(0.5);
    const R omega = R(1.147) * (tm * R(sp->om145[l])) + frcp<R>(a * a);
    mu_s[l] = fdiv<R>(R(sp->vis_coeff[l]) * st, omega);
    s[l] = fsqrt<R>(mu_s[l]);
    rs[l] = frcp<R>(s[l]);
  }
  // Wilke mixing and Mason-Saxena conductivity (Transport.cu:58-83)
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
    mu += fdiv<R>(mu_s[m] * X[m], vt);
    kappa += fdiv<R>(lam * X[m], R(1.065) * vt - R(0.065) * X[m]);
  }
  out[0] = mu;
  out[1] = kappa;
  // constant-Schmidt diffusivity (Transport.cu:85-92, X[l] in the test)
  const R mus = mu * R(S.inv_sc);
#pragma unroll
  for (int l = 0; l < NS; ++l)
    out[2 + l] = fabs(X[l] - R(1)) < R(1e-3) ? mus : fdiv<R>((R(1) - Y[l]) * mu, (R(1) - X[l]) * R(S.sc));
}

// Properties over the padded range.  In R = float (mixed precision) the
// kernel also writes the FP32 copies of the primitives the FP32 RHS reads.
template <typename R>
__global__ void __launch_bounds__(kTpb) props_kernel(const BlockView *__restrict__ views, const int *__restrict__ off,
                                                     int nb, const SpeciesDev *__restrict__ spg, SchemeDev S) {
  const SpeciesPtrT<R> sp = const_table<R>(spg);
  const PointWork pw = point_work(off, nb);
  const GView B = gview(views[pw.blk]);
  const long long np = B.np;
  constexpr bool f32 = sizeof(R) == 4;
  const long long p = pw.base + threadIdx.x;
  if (p >= np) return;
  const R t = R(B.bv[5 * np + p]);
  R out[4 + (NS > 0 ? NS : 0)];
  R Y[NSA];
  if constexpr (NS == 0) {
    const R c_temp = R(S.gamma_air * S.r_air);
    out[0] = R(S.gamma_air);
    out[1] = sqrt(c_temp * t);
    const R mu = R(1.716e-5) * pow(t / R(273), R(1.5)) * R(273 + 111) / (t + R(111));
    out[2] = mu;
    out[3] = mu * c_temp / R(S.gamma_air - 1) * R(S.inv_pr);
  } else {
    R cp[NSA], winv;
    if constexpr (f32) {
      const double ir = 1.0 / B.cv[p];
#pragma unroll
      for (int l = 0; l < NS; ++l) Y[l] = R(B.cv[(5 + l) * np + p] * ir);
    } else {
#pragma unroll
      for (int l = 0; l < NS; ++l) Y[l] = R(B.yk[l * np + p]);
    }
    mixture_gamma_c<R>(sp, t, Y, cp, winv, out);
    wilke_point<R>(sp, S, t, Y, cp, winv, out + 2);
  }
  if constexpr (f32) {
    // loads first (T and Y are already held), then the stores
    float b32[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) b32[v] = float(B.bv[v * np + p]);
#pragma unroll
    for (int v = 0; v < 4 + NS; ++v) B.props32[v * np + p] = out[v];
#pragma unroll
    for (int v = 0; v < 6; ++v) B.bv32[v * np + p] = b32[v];
    if constexpr (NS > 0) {
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk32[l * np + p] = float(Y[l]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4 + NS; ++v) B.props[v * np + p] = out[v];
  }
}

template <typename R>
int launch_props(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st) {
  if (b.total <= 0) return 0;
  props_kernel<R><<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, sp, S);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------------- RHS ---
// FP64 pencils at 2 waves per SIMD; FP32 needs half the registers and
// targets 3.
// grid < padded tile count: the persistent kernel strides through the tiles
template <int DIR, int V, typename R, bool D3>
int launch_rhs_dir(const Batch &b, int grid, const SpeciesDev *sp, const SchemeDev &S, int accumulate,
                   hipStream_t st) {
  constexpr PencilTile s = tile_of(D3, DIR, V);
  constexpr int WPE = sizeof(R) == 8 ? 2 : s.w32;
  if (b.total <= 0) return 0;
  const int padded = (b.total + 7) / 8 * 8;
  if (grid < padded)
    rhs_pencil_kernel<NS, DIR, s.ta, s.tb, s.tc, WPE, R, s.g, true>
        <<<grid, s.ta * s.tb * s.tc, 0, st>>>(b.views, b.off, b.nb, padded, sp, S, accumulate);
  else
    rhs_pencil_kernel<NS, DIR, s.ta, s.tb, s.tc, WPE, R, s.g, false>
        <<<padded, s.ta * s.tb * s.tc, 0, st>>>(b.views, b.off, b.nb, padded, sp, S, accumulate);
  return (int)hipGetLastError();
}

template <int DIR, int V, typename R, bool D3>
int occupancy_rhs_dir() {
  constexpr PencilTile s = tile_of(D3, DIR, V);
  constexpr int WPE = sizeof(R) == 8 ? 2 : s.w32;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &n, rhs_pencil_kernel<NS, DIR, s.ta, s.tb, s.tc, WPE, R, s.g, true>, s.ta * s.tb * s.tc, 0) != hipSuccess)
    return 0;
  return n;
}

template <typename R, typename Set>
void register_rhs(Set &k) {
  k.rhs3[0][0] = {launch_rhs_dir<0, 0, R, true>, occupancy_rhs_dir<0, 0, R, true>};
  k.rhs3[0][1] = {launch_rhs_dir<0, 1, R, true>, occupancy_rhs_dir<0, 1, R, true>};
  k.rhs3[0][2] = {launch_rhs_dir<0, 2, R, true>, occupancy_rhs_dir<0, 2, R, true>};
  k.rhs3[0][3] = {launch_rhs_dir<0, 3, R, true>, occupancy_rhs_dir<0, 3, R, true>};
  k.rhs3[0][4] = {launch_rhs_dir<0, 4, R, true>, occupancy_rhs_dir<0, 4, R, true>};
  k.rhs3[0][5] = {launch_rhs_dir<0, 5, R, true>, occupancy_rhs_dir<0, 5, R, true>};
  k.rhs3[1][0] = {launch_rhs_dir<1, 0, R, true>, occupancy_rhs_dir<1, 0, R, true>};
  k.rhs3[1][1] = {launch_rhs_dir<1, 1, R, true>, occupancy_rhs_dir<1, 1, R, true>};
  k.rhs3[1][2] = {launch_rhs_dir<1, 2, R, true>, occupancy_rhs_dir<1, 2, R, true>};
  k.rhs3[1][3] = {launch_rhs_dir<1, 3, R, true>, occupancy_rhs_dir<1, 3, R, true>};
  k.rhs3[1][4] = {launch_rhs_dir<1, 4, R, true>, occupancy_rhs_dir<1, 4, R, true>};
  k.rhs3[1][5] = {launch_rhs_dir<1, 5, R, true>, occupancy_rhs_dir<1, 5, R, true>};
  k.rhs3[2][0] = {launch_rhs_dir<2, 0, R, true>, occupancy_rhs_dir<2, 0, R, true>};
  k.rhs3[2][1] = {launch_rhs_dir<2, 1, R, true>, occupancy_rhs_dir<2, 1, R, true>};
  k.rhs3[2][2] = {launch_rhs_dir<2, 2, R, true>, occupancy_rhs_dir<2, 2, R, true>};
  k.rhs3[2][3] = {launch_rhs_dir<2, 3, R, true>, occupancy_rhs_dir<2, 3, R, true>};
  k.rhs3[2][4] = {launch_rhs_dir<2, 4, R, true>, occupancy_rhs_dir<2, 4, R, true>};
  k.rhs3[2][5] = {launch_rhs_dir<2, 5, R, true>, occupancy_rhs_dir<2, 5, R, true>};
  k.rhs2[0][0] = {launch_rhs_dir<0, 0, R, false>, occupancy_rhs_dir<0, 0, R, false>};
  k.rhs2[0][1] = {launch_rhs_dir<0, 1, R, false>, occupancy_rhs_dir<0, 1, R, false>};
  k.rhs2[1][0] = {launch_rhs_dir<1, 0, R, false>, occupancy_rhs_dir<1, 0, R, false>};
  k.rhs2[1][1] = {launch_rhs_dir<1, 1, R, false>, occupancy_rhs_dir<1, 1, R, false>};
}

#if CFX_PREC == 64
// ---------------------------------------------------------- total energy ---
// compute_total_energy (Thermo.cu:66-86) from bv/yk/cv at padded point p.
__device__ __forceinline__ void total_energy(const GView &B, SpeciesPtrT<double> sp, const SchemeDev &S,
                                             long long p) {
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

// The same total energy from values held in registers (rhoY: the partial
// densities), in total_energy's operation order.
__device__ __forceinline__ double total_energy_of(SpeciesPtrT<double> sp, const SchemeDev &S, double rho, double u,
                                                  double v, double w, double t, double pr, const double *rhoY) {
  double e = 0.5 * rho * (u * u + v * v + w * w);
  if constexpr (NS > 0) {
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      double h, cp;
      species_h_cp<double>(sp, l, t, h, cp);
      e += h * rhoY[l];
    }
    e -= pr;
  } else {
    e += pr / (S.gamma_air - 1);
  }
  return e;
}

// --------------------------------------------------------------- update ---
// update_cv_and_bv (TimeAdvanceFunc.cu:255-290) with the RK stage
// combination q = a*qn + b*(q + dt*dq/J) and the Newton inversion of T
// (Thermo.cu:89-125, stop |dT| <= 1e-10 K unless CFX_Q_NEWTON_REL).
// MIXED selects the FP32 dq.  The loads are issued before the first store:
// the stores go through pointers the compiler must assume alias the loaded
// arrays.
template <bool MIXED>
__global__ void __launch_bounds__(kTpb) update_kernel(const BlockView *__restrict__ views, const int *__restrict__ off,
                                                      int nb, const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                      int stage, const double *__restrict__ dt_dev, int *flags,
                                                      unsigned long long *newton_fail) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  constexpr int NV = 5 + NS;
  const PointWork pw = point_work(off, nb);
  const GView B = gview(views[pw.blk]);
  const long long np = B.np, ni = B.ni;
  const long long q = pw.base + threadIdx.x;
  if (q >= ni) return;
  const double A = stage == 1 ? 0.75 : (stage == 2 ? 1.0 / 3.0 : 0.0);
  const double Bc = stage == 1 ? 0.25 : (stage == 2 ? 2.0 / 3.0 : 1.0);
  const bool rel = (S.quirks & CFX_Q_NEWTON_REL) != 0;
  const bool rk3 = S.rk == 3;
  const long long p = interior_pidx(B, q);
  double cur[NV], dqv[NV], qnv[NV];
  // dq and the stage base qn are streamed once per stage: non-temporal
  // accesses keep them from displacing the state in the caches
#pragma unroll
  for (int l = 0; l < NV; ++l) cur[l] = B.cv[l * np + p];
#pragma unroll
  for (int l = 0; l < NV; ++l)
    dqv[l] = MIXED ? (double)__builtin_nontemporal_load(&B.dq32[l * ni + q]) : __builtin_nontemporal_load(&B.dq[l * ni + q]);
  const bool use_qn = rk3 && stage > 0;
  if (use_qn) {
#pragma unroll
    for (int l = 0; l < NV; ++l) qnv[l] = __builtin_nontemporal_load(&B.qn[l * ni + q]);
  }
  const double t_old = B.bv[5 * np + p];
  // LU-SGS: dq already holds the increment
  const double dtj = S.lusgs ? 1.0 : (rk3 ? *dt_dev : B.dt_local[q]) / B.jac[p];
  double c[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double base = use_qn ? qnv[l] : cur[l];
    c[l] = rk3 ? A * base + Bc * (cur[l] + dqv[l] * dtj) : cur[l] + dqv[l] * dtj;
  }
  if (B.mz == 1) c[3] = 0.0;
  // the stage base goes out now (every load of the point has been issued and
  // the Newton iteration reads only scalar memory), which frees cur; the
  // state's stores come after the iteration
  if (rk3 && stage == 0) {
#pragma unroll
    for (int l = 0; l < NV; ++l) __builtin_nontemporal_store(cur[l], &B.qn[l * ni + q]);
  }
  const double rho = c[0];
  const double ir = 1.0 / rho;
  const double u = c[1] * ir, v = c[2] * ir, w = c[3] * ir;
  const double v2 = u * u + v * v + w * w;
  double T, pr;
  double Y[NSA];
  if constexpr (NS > 0) {
    double winv = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      Y[l] = c[5 + l] * ir;
      winv += Y[l] * sp->inv_mw[l];
    }
    const double R = kRu * winv;
    const double e = c[4] / rho - 0.5 * v2;
    T = t_old;
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
#pragma unroll
  for (int l = 0; l < NV; ++l) B.cv[l * np + p] = c[l];
  if constexpr (NS > 0 && !MIXED) {  // the mixed path derives Y from cv where it needs it
#pragma unroll
    for (int l = 0; l < NS; ++l) B.yk[l * np + p] = Y[l];
  }
  B.bv[p] = rho;
  B.bv[np + p] = u;
  B.bv[2 * np + p] = v;
  B.bv[3 * np + p] = w;
  B.bv[4 * np + p] = pr;
  B.bv[5 * np + p] = T;
  if (!(rho > 0.0) || !(T > 0.0) || !(pr > 0.0) || !isfinite(rho) || !isfinite(T)) atomicOr(flags, 1);
  // apply_outflow (BoundCond.cu:455-487) of the faces folded into the
  // update: a point of the face layer writes its new state into the ghost
  // layers behind it (the same cache lines as its own stores for x faces)
  if (B.fuse_any) {
    const unsigned uq = (unsigned)q, mx = (unsigned)B.mx, my = (unsigned)B.my;
    const int ijk[3] = {(int)(uq % mx), (int)((uq / mx) % my), (int)(uq / (mx * my))};
    const int mm[3] = {B.mx, B.my, B.mz};
    const long long stride[3] = {1, B.sy, B.sz};
    const double bvn[6] = {rho, u, v, w, pr, T};
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        if (!B.fuse[2 * a + side] || ijk[a] != (side ? mm[a] - 1 : 0)) continue;
        const long long step = side ? stride[a] : -stride[a];
        for (int g = 1; g <= kNgg; ++g) {
          const long long qg = p + g * step;
#pragma unroll
          for (int l = 0; l < 6; ++l) B.bv[l * np + qg] = bvn[l];
#pragma unroll
          for (int l = 0; l < NV; ++l) B.cv[l * np + qg] = c[l];
          if constexpr (NS > 0 && !MIXED) {
#pragma unroll
            for (int l = 0; l < NS; ++l) B.yk[l * np + qg] = Y[l];
          }
        }
      }
  }
}

// compute_cv_from_bv (Field.cu:150-176) over the padded range
__global__ void __launch_bounds__(kTpb) cv_from_bv_kernel(const BlockView *__restrict__ views,
                                                          const int *__restrict__ off, int nb,
                                                          const SpeciesDev *__restrict__ spg, SchemeDev S) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  const PointWork pw = point_work(off, nb);
  const GView B = gview(views[pw.blk]);
  const long long np = B.np;
  const long long p = pw.base + threadIdx.x;
  if (p >= np) return;
  const double rho = B.bv[p];
  B.cv[p] = rho;
  B.cv[np + p] = rho * B.bv[np + p];
  B.cv[2 * np + p] = rho * B.bv[2 * np + p];
  B.cv[3 * np + p] = rho * B.bv[3 * np + p];
#pragma unroll
  for (int l = 0; l < NS; ++l) B.cv[(5 + l) * np + p] = rho * B.yk[l * np + p];
  total_energy(B, sp, S, p);
}

// ------------------------------------------------------------ ghost fill ---
// apply_outflow (BoundCond.cu:455-487) at face point (i, j, k)
__device__ __forceinline__ void bc_outflow_point(const GView &B, const FaceRange &f, int i, int j, int k) {
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  // every value of the face point is read before the first store: the
  // stores go to the same arrays, so the compiler could not move a later
  // load above an earlier store and each field would cost a memory round trip
  double bvp[6], cvp[5 + NS], ykp[NSA];
#pragma unroll
  for (int l = 0; l < 6; ++l) bvp[l] = B.bv[l * np + p];
#pragma unroll
  for (int l = 0; l < 5 + NS; ++l) cvp[l] = B.cv[l * np + p];
  if (!B.mixed) {
#pragma unroll
    for (int l = 0; l < NS; ++l) ykp[l] = B.yk[l * np + p];
  }
  for (int g = 1; g <= kNgg; ++g) {
    const long long q = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[l * np + q] = bvp[l];
#pragma unroll
    for (int l = 0; l < 5 + NS; ++l) B.cv[l * np + q] = cvp[l];
    if (!B.mixed) {
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk[l * np + q] = ykp[l];
    }
  }
}

// apply_inflow (BoundCond.cu:489-527); the state's total energy is
// precomputed on the host with the same formula as compute_total_energy.
__device__ __forceinline__ void bc_inflow_point(const GView &B, const FaceRange &f, int i, int j, int k,
                                                const InflowDev *__restrict__ in) {
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const double rho = in->st[0];
  for (int g = 1; g <= kNgg; ++g) {
    const long long q = pidx(B, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[l * np + q] = in->st[l];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if (!B.mixed) B.yk[l * np + q] = in->y[l];
      B.cv[(5 + l) * np + q] = rho * in->y[l];
    }
    B.cv[q] = rho;
    B.cv[np + q] = rho * in->st[1];
    B.cv[2 * np + q] = rho * in->st[2];
    B.cv[3 * np + q] = rho * in->st[3];
    B.cv[4 * np + q] = in->st[6];
  }
}

__device__ __forceinline__ double mix_mw(SpeciesPtrT<double> sp, const SchemeDev &S, const double *Y) {
  if constexpr (NS == 0) {
    return kRu / S.r_air;
  } else {
    double w = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) w += Y[l] * sp->inv_mw[l];
    return 1.0 / w;
  }
}

// apply_wall (BoundCond.cu:529-629) over the face's own vertex range (the
// registered range shrunk back by ngg, BoundCond.cu:229-239).
__device__ __forceinline__ void bc_wall_point(const GView &B, const FaceRange &f, int i, int j, int k,
                                              SpeciesPtrT<double> sp, const SchemeDev &S, const WallDev &wl) {
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  // every interior value is read before the first store (the stores could
  // alias later loads as far as the compiler knows; one round trip each);
  // interior point g = 1 is the one next to the wall
  long long qi[kNgg], qg[kNgg];
  double Yi[kNgg][NSA], ui[kNgg], vi[kNgg], wi[kNgg], pi[kNgg], ti[kNgg];
#pragma unroll
  for (int g = 0; g < kNgg; ++g) {
    qi[g] = pidx(B, i - (g + 1) * dir[0], j - (g + 1) * dir[1], k - (g + 1) * dir[2]);
    qg[g] = pidx(B, i + (g + 1) * dir[0], j + (g + 1) * dir[1], k + (g + 1) * dir[2]);
#pragma unroll
    for (int l = 0; l < NS; ++l) Yi[g][l] = mass_fraction(B, qi[g], l);
    ui[g] = B.bv[np + qi[g]], vi[g] = B.bv[2 * np + qi[g]], wi[g] = B.bv[3 * np + qi[g]];
    pi[g] = B.bv[4 * np + qi[g]], ti[g] = B.bv[5 * np + qi[g]];
  }
  double t_wall = wl.temperature;
  if (wl.adiabatic) t_wall = ti[0];
  const double pw = pi[0];
  double rY[NSA];
  {
    const double rho = pw * mix_mw(sp, S, Yi[0]) / (t_wall * kRu);
    if (!B.mixed) {
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk[l * np + p] = Yi[0][l];
    }
    B.bv[p] = rho;
    B.bv[np + p] = 0;
    B.bv[2 * np + p] = 0;
    B.bv[3 * np + p] = 0;
    B.bv[4 * np + p] = pw;
    B.bv[5 * np + p] = t_wall;
    B.cv[p] = rho;
    B.cv[np + p] = 0;
    B.cv[2 * np + p] = 0;
    B.cv[3 * np + p] = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      rY[l] = Yi[0][l] * rho;
      B.cv[(l + 5) * np + p] = rY[l];
    }
    B.cv[4 * np + p] = total_energy_of(sp, S, rho, 0.0, 0.0, 0.0, t_wall, pw, rY);
  }
#pragma unroll
  for (int g = 0; g < kNgg; ++g) {
    if (!B.mixed) {
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk[l * np + qg[g]] = Yi[g][l];
    }
    const double mw = mix_mw(sp, S, Yi[g]);
    double t_g = ti[g];
    if (!wl.adiabatic) {
      t_g = 2 * t_wall - ti[g];
      if (t_g <= 0.1 * t_wall) t_g = t_wall;
    }
    const double rho_g = pi[g] * mw / (t_g * kRu);
    B.bv[qg[g]] = rho_g;
    B.bv[np + qg[g]] = -ui[g];
    B.bv[2 * np + qg[g]] = -vi[g];
    B.bv[3 * np + qg[g]] = -wi[g];
    B.bv[4 * np + qg[g]] = pi[g];
    B.bv[5 * np + qg[g]] = t_g;
    B.cv[qg[g]] = rho_g;
    B.cv[np + qg[g]] = -rho_g * ui[g];
    B.cv[2 * np + qg[g]] = -rho_g * vi[g];
    B.cv[3 * np + qg[g]] = -rho_g * wi[g];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      rY[l] = Yi[g][l] * rho_g;
      B.cv[(l + 5) * np + qg[g]] = rY[l];
    }
    B.cv[4 * np + qg[g]] = total_energy_of(sp, S, rho_g, -ui[g], -vi[g], -wi[g], t_g, pi[g], rY);
  }
}

// Slip wall: the boundary point loses its velocity along the face normal
// n = grad(xi_face)/|grad(xi_face)|; ghost g mirrors interior point g with
// the normal velocity negated and rho, p, T, Y copied.
__device__ __forceinline__ void bc_symmetry_point(const GView &B, const FaceRange &f, int i, int j, int k,
                                                  SpeciesPtrT<double> sp, const SchemeDev &S) {
  int dir[3] = {0, 0, 0};
  dir[f.face] = f.direction;
  const long long np = B.np;
  const long long p = pidx(B, i, j, k);
  // every value is read before the first store (as in the wall condition)
  double n[3];
  for (int c = 0; c < 3; ++c) n[c] = B.metric[(f.face * 3 + c) * np + p];
  double bp[6], rYp[NSA];
#pragma unroll
  for (int l = 0; l < 6; ++l) bp[l] = B.bv[l * np + p];
#pragma unroll
  for (int l = 0; l < NS; ++l) rYp[l] = B.cv[(5 + l) * np + p];
  long long qi[kNgg], qg[kNgg];
  double bi[kNgg][6], Yi[kNgg][NSA];
#pragma unroll
  for (int g = 0; g < kNgg; ++g) {
    qi[g] = pidx(B, i - (g + 1) * dir[0], j - (g + 1) * dir[1], k - (g + 1) * dir[2]);
    qg[g] = pidx(B, i + (g + 1) * dir[0], j + (g + 1) * dir[1], k + (g + 1) * dir[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) bi[g][l] = B.bv[l * np + qi[g]];
#pragma unroll
    for (int l = 0; l < NS; ++l) Yi[g][l] = mass_fraction(B, qi[g], l);
  }
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= nn;
  n[1] /= nn;
  n[2] /= nn;
  {
    const double vn = bp[1] * n[0] + bp[2] * n[1] + bp[3] * n[2];
    for (int c = 0; c < 3; ++c) {
      bp[1 + c] -= vn * n[c];
      B.bv[(1 + c) * np + p] = bp[1 + c];
    }
    for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + p] = bp[0] * bp[1 + c];
    B.cv[4 * np + p] = total_energy_of(sp, S, bp[0], bp[1], bp[2], bp[3], bp[5], bp[4], rYp);
  }
#pragma unroll
  for (int g = 0; g < kNgg; ++g) {
    const double *b = bi[g];
    const double vn = b[1] * n[0] + b[2] * n[1] + b[3] * n[2];
    double u[3], rY[NSA];
    for (int c = 0; c < 3; ++c) u[c] = b[1 + c] - 2 * vn * n[c];
    B.bv[qg[g]] = b[0];
    for (int c = 0; c < 3; ++c) B.bv[(1 + c) * np + qg[g]] = u[c];
    B.bv[4 * np + qg[g]] = b[4];
    B.bv[5 * np + qg[g]] = b[5];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if (!B.mixed) B.yk[l * np + qg[g]] = Yi[g][l];
      rY[l] = b[0] * Yi[g][l];
      B.cv[(5 + l) * np + qg[g]] = rY[l];
    }
    B.cv[qg[g]] = b[0];
    for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + qg[g]] = b[0] * u[c];
    B.cv[4 * np + qg[g]] = total_energy_of(sp, S, b[0], u[0], u[1], u[2], b[5], b[4], rY);
  }
}

// inner_communication (Field.cu:213-266) with the face loop of
// Mesh.cpp:905-917 at loop point n: shared-face points are averaged between
// the two blocks, ghost layers take the neighbour's values.
__device__ __forceinline__ void exchange_point(const GView &B, const GView &T, const InnerDev &f,
                                               const int n[3]) {
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
  // Loads are grouped ahead of the stores (the two blocks' arrays may alias
  // as far as the compiler knows, so it keeps every store before any later
  // load: field by field, each would cost a memory round trip).
  if (idx[f.face] == face_layer) {
    // the shared layer: both sides of a group of fields, then the averages
    auto average = [&](gptr<double> bb, gptr<double> tt, auto kc) {
      constexpr int K = decltype(kc)::value;
      double x[K > 0 ? K : 1], y[K > 0 ? K : 1];
#pragma unroll
      for (int l = 0; l < K; ++l) x[l] = tt[l * tnp + t], y[l] = bb[l * np + p];
#pragma unroll
      for (int l = 0; l < K; ++l) {
        const double a = 0.5 * (x[l] + y[l]);
        bb[l * np + p] = a;
        tt[l * tnp + t] = a;
      }
    };
    average(B.bv, T.bv, std::integral_constant<int, 6>{});
    average(B.cv, T.cv, std::integral_constant<int, 5 + NS>{});
    if constexpr (NS > 0)
      if (!B.mixed) average(B.yk, T.yk, std::integral_constant<int, NS>{});
  } else {
    // a ghost point: the neighbour's values, all read before any store
    double bvt[6], cvt[5 + NS], ykt[NSA];
#pragma unroll
    for (int l = 0; l < 6; ++l) bvt[l] = T.bv[l * tnp + t];
#pragma unroll
    for (int l = 0; l < 5 + NS; ++l) cvt[l] = T.cv[l * tnp + t];
    if (!B.mixed) {
#pragma unroll
      for (int l = 0; l < NS; ++l) ykt[l] = T.yk[l * tnp + t];
    }
#pragma unroll
    for (int l = 0; l < 6; ++l) B.bv[l * np + p] = bvt[l];
#pragma unroll
    for (int l = 0; l < 5 + NS; ++l) B.cv[l * np + p] = cvt[l];
    if (!B.mixed) {
#pragma unroll
      for (int l = 0; l < NS; ++l) B.yk[l * np + p] = ykt[l];
    }
  }
}

// One launch runs a set of ghost-fill ops that touch disjoint data (one
// level of the op schedule, or a single op): workgroup -> op by the prefix
// off[], then one point of the op per thread.
__global__ void __launch_bounds__(kTpb) face_ops_kernel(const BlockView *__restrict__ views,
                                                        const FaceOp *__restrict__ ops, const int *__restrict__ off,
                                                        int nops, const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                        const InflowDev *__restrict__ inflow) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  const int w = blockIdx.x + off[0];
  const int s = find_seg(off, nops, w);
  const FaceOp &op = ops[s];
  const long long n = (long long)(w - off[s]) * kTpb + threadIdx.x;
  if (n >= op.npts) return;
  const int e0 = op.ext[0], e1 = op.ext[1];
  const int nn[3] = {(int)(n % e0), (int)((n / e0) % e1), (int)(n / ((long long)e0 * e1))};
  const GView B = gview(views[op.blk]);
  const int i = op.r.rs[0] + nn[0], j = op.r.rs[1] + nn[1], k = op.r.rs[2] + nn[2];
  switch (op.kind) {
    case CFX_BC_OUTFLOW: bc_outflow_point(B, op.r, i, j, k); break;
    case CFX_BC_INFLOW: bc_inflow_point(B, op.r, i, j, k, inflow + op.cond); break;
    case CFX_BC_WALL: bc_wall_point(B, op.r, i, j, k, sp, S, op.wl); break;
    case CFX_BC_SYMMETRY: bc_symmetry_point(B, op.r, i, j, k, sp, S); break;
    case kOpExchange: {
      const GView T = gview(views[op.tgt]);
      exchange_point(B, T, op.f, nn);
      break;
    }
    default: break;
  }
}

int launch_update(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, int mixed, int stage, const double *dt,
                  int *flags, unsigned long long *nf, hipStream_t st) {
  if (b.total <= 0) return 0;
  if (mixed)
    update_kernel<true><<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, sp, S, stage, dt, flags, nf);
  else
    update_kernel<false><<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, sp, S, stage, dt, flags, nf);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------- LU-SGS ---
// Matrix-free LU-SGS (Yoon & Jameson, AIAA J. 26 (1988) 1025-1026), the
// implicit treatment the reference declares without a body (LUSGS,
// TemporalScheme.cuh:19-21).  Per interior cell:
//   D = J/dt + sum_d (omega r_d + 2 v_d),  a_d = omega r_d / 2 + v_d,
//   r_d = J (|V.grad xi_d| + c |grad xi_d|),  v_d = max(gamma, 4/3) mu/(rho Pr) J |grad xi_d|^2;
//   forward  D dQ*_i = R_i + sum_d [ dG_d(i-e_d)/2 + a_d(i-e_d) dQ*_(i-e_d) ]
//   backward dQ_i = dQ*_i + sum_d [ a_d(i+e_d) dQ_(i+e_d) - dG_d(i+e_d)/2 ] / D_i
// dG_d(n) = G_d(Q_n + dQ_n) - G_d(Q_n): the inviscid flux of cell n through
// S_d = J grad xi_d, the pressure change at the cell's frozen gamma; dQ = 0
// outside the block.  dq holds R on entry and dQ on exit (in place: a cell
// reads only final neighbour values).

// lus: D over the interior, then a_d for d = 0, 1, 2
__global__ void __launch_bounds__(kTpb) lusgs_coef_kernel(const BlockView *__restrict__ views,
                                                          const int *__restrict__ off, int nb, SchemeDev S) {
  const PointWork pw = point_work(off, nb);
  const GView B = gview(views[pw.blk]);
  const long long np = B.np, ni = B.ni;
  const long long q = pw.base + threadIdx.x;
  if (q >= ni) return;
  const long long p = interior_pidx(B, q);
  const double J = B.jac[p], rho = B.bv[p];
  const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
  const double gam = B.props[p], c = B.props[np + p], mu = B.props[2 * np + p];
  const double vc = fmax(gam, 4.0 / 3.0) * mu / rho / S.pr;
  double diag = J / B.dt_local[q];
  for (int d = 0; d < S.dim; ++d) {
    const double gx = B.metric[(3 * d) * np + p], gy = B.metric[(3 * d + 1) * np + p],
                 gz = B.metric[(3 * d + 2) * np + p];
    const double g2 = gx * gx + gy * gy + gz * gz;
    const double r = J * (fabs(u * gx + v * gy + w * gz) + c * sqrt(g2));
    const double vd = vc * J * g2;
    diag += S.omega * r + 2 * vd;
    B.lus[(1 + d) * ni + q] = 0.5 * S.omega * r + vd;
  }
  B.lus[q] = diag;
}

// The sweeps in tiles of kLuTi x kLuTj x kLuTk cells.  Tiles (I, J, K) on
// one tile hyperplane I + J + K = tplane depend only on tiles of the
// previous one, so one launch runs a tile hyperplane, one workgroup per
// tile.  The workgroup first stages in LDS, for its tile and the upstream
// face layers around it (the cells at i0-1, j0-1, k0-1 for the forward
// sweep, past the far faces for the backward one, final since an earlier
// launch), each cell's dQ slot (R or dQ* on entry), conserved state,
// pressure, gamma, J grad xi_d, a_d and D.  Lane (jl, kl) of a wavefront
// then stands for the cell column (j, k), marching along i: at local step s
// the cell il = s - jl - kl is updated, whose upstream neighbours were
// updated at step s-1 or
// are staged; every read of the dependent steps is an LDS read, and a
// cell's slot is written once, after every read of its input.  The tile's
// increments go back to dq at the end.  The divisions share a
// Newton-refined reciprocal per denominator; the three neighbour terms of a
// cell are formed by three wavefronts at once and added in a fixed order.
constexpr int kLuEi = kLuTi + 1, kLuEj = kLuTj + 1, kLuEk = kLuTk + 1, kLuExt = kLuEi * kLuEj * kLuEk;
// LDS slot of staged cell (i, j, k): i + kLuEi j + kLuSk k, the k planes one
// slot apart so that the lanes of a step, which differ in j and k, spread
// over the LDS banks (measured: bank-conflict cycles 40% of the LDS-active
// cycles with kLuSk = 45, 7% with 46); kLuSlots slots per staged variable
constexpr int kLuSk = kLuEi * kLuEj + 1, kLuSlots = (kLuEk - 1) * kLuSk + kLuEi * kLuEj;
__device__ __forceinline__ int lu_slot(int i, int j, int k) { return (k * kLuSk + j * kLuEi) + i; }
// a / b with r = frcp(b) shared by the quotients over one b: one residual
// correction of a r (within an ulp of the IEEE quotient)
__device__ __forceinline__ double lu_div(double a, double b, double r) {
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}
// staged values per cell: dQ (NV), state (NV), p, gamma, S_d (3 x 3), a_d (3), D
constexpr int kLuVars = 2 * (5 + NS) + 15;
// four wavefronts stage the tile and write it back; in the dependent steps
// wavefronts 0-2 form one upstream direction's terms each
constexpr int kLuThreads = 256;

template <bool FWD>
__global__ void __launch_bounds__(kLuThreads) lusgs_tile_kernel(const BlockView *__restrict__ views,
                                                        const int *__restrict__ off, int nb, SchemeDev S,
                                                        int tplane) {
  constexpr int NV = 5 + NS, STEPS = kLuTi + kLuTj + kLuTk - 2;
  constexpr int XQ = 0, XC = NV, XP = 2 * NV, XG = 2 * NV + 1, XS = 2 * NV + 2, XA = 2 * NV + 11, XD = 2 * NV + 14;
  static_assert(kLuTj * kLuTk == 64, "one lane per (j, k) column of a tile");
  static_assert((kLuVars * kLuSlots + 3 * NV * 64) * 8 <= 160 * 1024, "the staged tile fits in LDS");
  __shared__ double ext[kLuVars][kLuSlots];
  __shared__ double term[3][NV][64];  // the step's neighbour terms, by direction
  const int w = blockIdx.x + off[0];
  const int blk = find_seg(off, nb, w);
  const GView B = gview(views[blk]);
  const int m[3] = {B.mx, B.my, B.mz};
  const int nt0 = (m[0] + kLuTi - 1) / kLuTi, nt1 = (m[1] + kLuTj - 1) / kLuTj, nt2 = (m[2] + kLuTk - 1) / kLuTk;
  const int t = w - off[blk];
  const int T1 = t % nt1, T2 = t / nt1, T0 = tplane - T1 - T2;
  if (T2 >= nt2 || T0 < 0 || T0 >= nt0) return;  // uniform over the workgroup
  const int lane = threadIdx.x;
  const int org[3] = {T0 * kLuTi - (FWD ? 1 : 0), T1 * kLuTj - (FWD ? 1 : 0), T2 * kLuTk - (FWD ? 1 : 0)};
  const long long np = B.np, ni = B.ni;
  // ---- stage the tile and its upstream faces
  for (int n = lane; n < kLuExt; n += kLuThreads) {
    const int li = n % kLuEi, lj = (n / kLuEi) % kLuEj, lk = n / (kLuEi * kLuEj);
    const int e = lu_slot(li, lj, lk);
    const int c[3] = {org[0] + li, org[1] + lj, org[2] + lk};
    if (c[0] < 0 || c[0] >= m[0] || c[1] < 0 || c[1] >= m[1] || c[2] < 0 || c[2] >= m[2]) continue;
    const long long q = ((long long)c[2] * m[1] + c[1]) * m[0] + c[0];
    const long long p = pidx(B, c[0], c[1], c[2]);
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      ext[XQ + l][e] = B.dq[l * ni + q];
      ext[XC + l][e] = B.cv[l * np + p];
    }
    ext[XP][e] = B.bv[4 * np + p];
    ext[XG][e] = B.props[p];
    const double J = B.jac[p];
#pragma unroll
    for (int v = 0; v < 9; ++v) ext[XS + v][e] = J * B.metric[v * np + p];
#pragma unroll
    for (int d = 0; d < 3; ++d) ext[XA + d][e] = B.lus[(1 + d) * ni + q];
    ext[XD][e] = B.lus[q];
  }
  __syncthreads();
  // ---- the dependent steps, in LDS: wavefront d (0..2) forms the term of
  // the upstream neighbour along d for every column of the step, wavefront 0
  // then adds them in the order d = 0, 1, 2 and updates the cells
  const int wv = lane >> 6, cl = lane & 63;
  const int jl = cl % kLuTj, kl = cl / kLuTj;
  const int j = T1 * kLuTj + jl, k = T2 * kLuTk + kl;
  const bool col = j < m[1] && k < m[2];
  const int o = FWD ? 1 : 0;  // position of the tile's first cell in the staged box
  const int es[3] = {1, kLuEi, kLuSk};
  for (int step = 0; step < STEPS; ++step) {
    const int il = (FWD ? step : STEPS - 1 - step) - jl - kl;
    const int i = T0 * kLuTi + il;
    const bool act = col && il >= 0 && il < kLuTi && i < m[0];
    const int ijk[3] = {i, j, k};
    const int e = lu_slot(il + o, jl + o, kl + o);
    if (act && wv < S.dim && (FWD ? ijk[wv] > 0 : ijk[wv] < m[wv] - 1)) {
      const int d = wv;
      const int en = FWD ? e - es[d] : e + es[d];
      // dG_d = G_d(Q + dQ) - G_d(Q) of the neighbour, pressure change at its
      // frozen gamma
      double dn[NV], q0[NV], q1[NV], dG[NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) {
        dn[l] = ext[XQ + l][en];
        q0[l] = ext[XC + l][en];
        q1[l] = q0[l] + dn[l];
      }
      const double Sx = ext[XS + 3 * d][en], Sy = ext[XS + 3 * d + 1][en], Sz = ext[XS + 3 * d + 2][en];
      const double gam = ext[XG][en], p0 = ext[XP][en];
      const double r0 = frcp<double>(q0[0]), r1 = frcp<double>(q1[0]);
      const double ke0 = lu_div(0.5 * (q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]), q0[0], r0);
      const double ke1 = lu_div(0.5 * (q1[1] * q1[1] + q1[2] * q1[2] + q1[3] * q1[3]), q1[0], r1);
      const double p1 = p0 + (gam - 1) * (dn[4] - (ke1 - ke0));
      const double U0 = lu_div(Sx * q0[1] + Sy * q0[2] + Sz * q0[3], q0[0], r0);
      const double U1 = lu_div(Sx * q1[1] + Sy * q1[2] + Sz * q1[3], q1[0], r1);
      dG[0] = q1[0] * U1 - q0[0] * U0;
      dG[1] = (q1[1] * U1 + Sx * p1) - (q0[1] * U0 + Sx * p0);
      dG[2] = (q1[2] * U1 + Sy * p1) - (q0[2] * U0 + Sy * p0);
      dG[3] = (q1[3] * U1 + Sz * p1) - (q0[3] * U0 + Sz * p0);
      dG[4] = (q1[4] + p1) * U1 - (q0[4] + p0) * U0;
#pragma unroll
      for (int l = 5; l < NV; ++l) dG[l] = q1[l] * U1 - q0[l] * U0;
      const double an = ext[XA + d][en];
#pragma unroll
      for (int l = 0; l < NV; ++l) term[d][l][cl] = FWD ? 0.5 * dG[l] + an * dn[l] : an * dn[l] - 0.5 * dG[l];
    }
    __syncthreads();
    if (act && wv == 0) {
      double acc[NV];
#pragma unroll
      for (int l = 0; l < NV; ++l) acc[l] = FWD ? ext[XQ + l][e] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!(d < S.dim && (FWD ? ijk[d] > 0 : ijk[d] < m[d] - 1))) continue;
#pragma unroll
        for (int l = 0; l < NV; ++l) acc[l] += term[d][l][cl];
      }
      const double D = ext[XD][e], rD = frcp<double>(D);
#pragma unroll
      for (int l = 0; l < NV; ++l)
        ext[XQ + l][e] = FWD ? lu_div(acc[l], D, rD) : ext[XQ + l][e] + lu_div(acc[l], D, rD);
    }
    __syncthreads();
  }
  // ---- the tile's increments back to dq
  for (int c = lane; c < kLuTi * kLuTj * kLuTk; c += kLuThreads) {
    const int il = c % kLuTi, jc = (c / kLuTi) % kLuTj, kc = c / (kLuTi * kLuTj);
    const int i = T0 * kLuTi + il, jg = T1 * kLuTj + jc, kg = T2 * kLuTk + kc;
    if (i >= m[0] || jg >= m[1] || kg >= m[2]) continue;
    const int e = lu_slot(il + o, jc + o, kc + o);
    const long long q = ((long long)kg * m[1] + jg) * m[0] + i;
#pragma unroll
    for (int l = 0; l < NV; ++l) B.dq[l * ni + q] = ext[XQ + l][e];
  }
}

int launch_lusgs_coef(const Batch &b, const SchemeDev &S, hipStream_t st) {
  if (b.total <= 0) return 0;
  lusgs_coef_kernel<<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, S);
  return (int)hipGetLastError();
}

int launch_lusgs_sweep(const Batch &b, const SchemeDev &S, int tplane, int fwd, hipStream_t st) {
  if (b.total <= 0) return 0;
  if (fwd)
    lusgs_tile_kernel<true><<<b.total, kLuThreads, 0, st>>>(b.views, b.off, b.nb, S, tplane);
  else
    lusgs_tile_kernel<false><<<b.total, kLuThreads, 0, st>>>(b.views, b.off, b.nb, S, tplane);
  return (int)hipGetLastError();
}

int launch_cv_from_bv(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st) {
  if (b.total <= 0) return 0;
  cv_from_bv_kernel<<<b.total, kTpb, 0, st>>>(b.views, b.off, b.nb, sp, S);
  return (int)hipGetLastError();
}

int launch_face_ops(const BlockView *views, const FaceOp *ops, const int *off, int nops, int total,
                    const SpeciesDev *sp, const SchemeDev &S, const InflowDev *inflow, hipStream_t st) {
  if (total <= 0) return 0;
  face_ops_kernel<<<total, kTpb, 0, st>>>(views, ops, off, nops, sp, S, inflow);
  return (int)hipGetLastError();
}

// ------------------------------------------------- surrogate training set ---
// Counter-based uniform in [0, 1): SplitMix64 finaliser of (seed, sample, k);
// each sample owns 4096 counters (a sample draws at most 2 + ns (3*32 + 1)).
__device__ __forceinline__ double rng_uniform(unsigned long long seed, long long s, unsigned k) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)s * 4096ull + k + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Gamma(a, 1) by Marsaglia & Tsang (2000), "A simple method for generating
// gamma variables", with the boost G(a) = G(a+1) U^(1/a) for a < 1; normals
// by Box-Muller.  Bounded: after 32 rejections the mode is returned.
__device__ __forceinline__ double rng_gamma(unsigned long long seed, long long s, unsigned &k, double a) {
  const double aa = a < 1.0 ? a + 1.0 : a;
  const double d = aa - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  double g = d;
  for (int it = 0; it < 32; ++it) {
    const double u1 = rng_uniform(seed, s, k++), u2 = rng_uniform(seed, s, k++);
    const double z = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
    const double v1 = 1.0 + c * z;
    if (v1 <= 0.0) continue;
    const double v = v1 * v1 * v1;
    const double u = rng_uniform(seed, s, k++);
    if (log(1.0 - u) < 0.5 * z * z + d - d * v + d * log(v)) {
      g = d * v;
      break;
    }
  }
  if (a < 1.0) g *= pow(1.0 - rng_uniform(seed, s, k++), 1.0 / a);
  return g;
}

// Sample s: T ~ U[t_lo, t_hi], log p ~ U[lp_lo, lp_hi], Y ~ Dirichlet(alpha);
// inputs x = (T, log p, Y) and targets y = log(mu, kappa, rho*D_l) of the
// FP64 transport routine.
__global__ void __launch_bounds__(256) mlp_samples_kernel(const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                          long long n, unsigned long long seed, double t_lo,
                                                          double t_hi, double lp_lo, double lp_hi, double alpha,
                                                          double *__restrict__ x, double *__restrict__ y) {
  const SpeciesPtrT<double> sp = const_table<double>(spg);
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  unsigned k = 0;
  const double t = t_lo + (t_hi - t_lo) * rng_uniform(seed, s, k++);
  const double lp = lp_lo + (lp_hi - lp_lo) * rng_uniform(seed, s, k++);
  double Y[NSA], sum = 0.0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    Y[l] = rng_gamma(seed, s, k, alpha);
    sum += Y[l];
  }
  const double inv = sum > 0.0 ? 1.0 / sum : 0.0;
#pragma unroll
  for (int l = 0; l < NS; ++l) Y[l] = sum > 0.0 ? Y[l] * inv : 1.0 / NS;
  double cp[NSA], winv, gc[2], out[2 + NSA];
  mixture_gamma_c<double>(sp, t, Y, cp, winv, gc);
  wilke_point<double>(sp, S, t, Y, cp, winv, out);
  double *xr = x + s * (NS + 2), *yr = y + s * (NS + 2);
  xr[0] = t;
  xr[1] = lp;
#pragma unroll
  for (int l = 0; l < NS; ++l) xr[2 + l] = Y[l];
#pragma unroll
  for (int o = 0; o < NS + 2; ++o) yr[o] = log(out[o]);
}

int launch_mlp_samples(const SpeciesDev *sp, const SchemeDev &S, long long n, unsigned long long seed,
                       const double lim[4], double alpha, double *x, double *y, hipStream_t st) {
  if constexpr (NS == 0) {
    return CFX_E_UNSUPPORTED;
  } else {
    if (n <= 0) return 0;
    mlp_samples_kernel<<<(int)((n + 255) / 256), 256, 0, st>>>(sp, S, n, seed, lim[0], lim[1], lim[2], lim[3],
                                                                alpha, x, y);
    return (int)hipGetLastError();
  }
}

struct Register64 {
  Register64() {
    KernelSet64 &k = g_kernels64[NS];
    k.mlp_samples = launch_mlp_samples;
    k.props = launch_props<double>;
    k.cv_from_bv = launch_cv_from_bv;
    k.update = launch_update;
    k.face_ops = launch_face_ops;
    k.lusgs_coef = launch_lusgs_coef;
    k.lusgs_sweep = launch_lusgs_sweep;
    register_rhs<double>(k);
    k.present = 1;
  }
} register64;

#else  // CFX_PREC == 32

// ----------------------------------------------- properties, MLP transport ---
// The mixed-precision properties with mu, kappa and rho*D from the surrogate.
// Each wavefront evaluates its 64 points as four 16-point MFMA tiles (inputs
// and outputs pass through a per-wavefront LDS buffer between the
// point-per-lane and the tile layouts).  Points outside the training envelope
// take the Wilke routine; one chunk-wavefront in check_every of each
// workgroup also evaluates Wilke at all its points and keeps Wilke where the
// surrogate is off by more than check_tol.  The work runs in three phases so
// that each holds only its own registers: (1) FP32 primitives, Y, gamma, c
// and the network inputs, with every output but the transport coefficients
// stored at once; (2) the network; (3) the coefficients, with the Wilke
// fallback recomputing its inputs from the FP64 state.
// counters: [0] points outside, [1] points checked, [2] checks failed.

// Orders the LDS accesses of one wavefront (its lanes exchange values through
// its own slice of LDS, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Phase 1 from the FP32 primitives b32 (rho, u, v, w, p, T) and Y of padded
// point p (byte offset o4 in the FP32 arrays): the network inputs go to
// xl[lane], and with store set gamma, c and the FP32 copies are written.
// Returns whether the point lies outside the training envelope.
__device__ __forceinline__ bool mlp_phase1(const GView &B, SpeciesPtrT<float> sp, const MlpDev *__restrict__ mlp,
                                           const float *b32, const float *Y, float (*xl)[kMlpIO + 1], bool store,
                                           unsigned o4) {
  const int lane = threadIdx.x & 63;
  const long long np = B.np;
  float cp[NSA], winv, gc[2], x[kMlpIO];
  mixture_gamma_c<float>(sp, b32[5], Y, cp, winv, gc);
  float ymax = 0.0f;  // largest Y_l / W_l: the largest mole fraction times winv
#pragma unroll
  for (int l = 0; l < NS; ++l) ymax = fmaxf(ymax, Y[l] * sp->inv_mw[l]);
  const bool outside = mlp_inputs<NS>(mlp, b32[5], flog<float>(b32[4]), Y, x, ymax * frcp<float>(winv));
#pragma unroll
  for (int f = 0; f < kMlpIO; ++f) xl[lane][f] = x[f];
  if (store) {
    *atg<float>(B.props32, o4) = gc[0];
    *atg<float>(B.props32 + np, o4) = gc[1];
#pragma unroll
    for (int v = 0; v < 6; ++v) *atg<float>(B.bv32 + v * np, o4) = b32[v];
#pragma unroll
    for (int l = 0; l < NS; ++l) *atg<float>(B.yk32 + l * np, o4) = Y[l];
  }
  return outside;
}

// Phases 2 and 3 for the wavefront: the network on xl with the operands
// staged in LDS, the coefficients, the Wilke fallback (its inputs reloaded
// from the FP64 state at byte offset o8; at every point when check is set),
// the counters over the lanes with store set, and their stores.
__device__ __forceinline__ void mlp_phase23(const GView &B, SpeciesPtrT<float> sp, const SchemeDev &S,
                                            const MlpDev *__restrict__ mlp, lds_ptr<const bf16x8> ops,
                                            lds_ptr<const f32x4> bias, unsigned long long *__restrict__ counters,
                                            float (*xl)[kMlpIO + 1], bool outside, bool store, unsigned o8,
                                            unsigned o4, bool check) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  wave_sync();
  {
    bf16x8 xo[4];
    f32x4 yv[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int j = 0; j < 4; ++j) xo[tt][j] = (__bf16)xl[16 * tt + c][4 * g + j], xo[tt][4 + j] = (__bf16)0.0f;
    wave_sync();
    mlp_infer_tiles<4>(ops, bias, xo, yv);
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) xl[16 * tt + c][4 * g + r] = yv[tt][r];
  }
  wave_sync();
  // fresh copies of the uniform bases, so that the field addresses of the
  // first phase need not stay live (in SGPRs) through the network
  gptr<const double> cvb = B.cv, bvb = B.bv;
  gptr<float> prb = B.props32;
  long long npl = B.np;
  asm volatile("" : "+s"(cvb), "+s"(bvb), "+s"(prb), "+s"(npl));
  float tr[2 + NSA];
#pragma unroll
  for (int o = 0; o < NS + 2; ++o) tr[o] = __expf(xl[lane][o] * mlp->t_std[o] + mlp->t_mean[o]);
  bool fail = false;
  if (check || outside) {
    float Y[NSA], cp[NSA], winv, gc[2], w[2 + NSA];
    const float t = float(ldg<double>(bvb + 5 * npl, o8));
    const double ir = 1.0 / ldg<double>(cvb, o8);
#pragma unroll
    for (int l = 0; l < NS; ++l) Y[l] = float(ldg<double>(cvb + (5 + l) * npl, o8) * ir);
    mixture_gamma_c<float>(sp, t, Y, cp, winv, gc);
    wilke_point<float>(sp, S, t, Y, cp, winv, w);
    if (check) {
#pragma unroll
      for (int o = 0; o < NS + 2; ++o) fail = fail || !(fabsf(tr[o] - w[o]) <= mlp->check_tol * fabsf(w[o]));
    }
    if (outside || fail) {
#pragma unroll
      for (int o = 0; o < NS + 2; ++o) tr[o] = w[o];
    }
  }
  const unsigned long long n_out = __ballot(store && outside), n_chk = __ballot(store && check),
                           n_fail = __ballot(store && fail);
  if (lane == 0) {
    if (n_out) atomicAdd(&counters[0], (unsigned long long)__popcll(n_out));
    if (n_chk) atomicAdd(&counters[1], (unsigned long long)__popcll(n_chk));
    if (n_fail) atomicAdd(&counters[2], (unsigned long long)__popcll(n_fail));
  }
  if (!store) return;
#pragma unroll
  for (int o = 0; o < NS + 2; ++o) *atg<float>(prb + (2 + o) * npl, o4) = tr[o];
}

// Surrogate properties over the padded range.  A persistent grid: workgroup
// b walks the kTpb-point chunks b, b + gridDim.x, ... of the batch, and loads
// the FP64 inputs of its next chunk before the network runs on the current
// one, so that those loads overlap the arithmetic.  The inference operands
// are staged in LDS once per workgroup: the network then issues no vector
// memory instruction whose wait would also wait for the prefetch.
constexpr int kMlpInfBlocks = kB4;  // forward A operands (the first kB4 blocks)
constexpr int kMlpIn = 7 + NS;      // rho, u, v, w, p, T, rho and rho*Y_l of one point

__device__ __forceinline__ void mlp_fetch(const BlockView *__restrict__ views, const int *__restrict__ off, int nb,
                                          int w, double *in) {
  const int gw = w + off[0];
  const int blk = find_seg(off, nb, gw);
  const GView B = gview(views[blk]);
  const long long np = B.np;
  const long long p0 = (long long)(gw - off[blk]) * kTpb + threadIdx.x;
  const unsigned o8 = unsigned(p0 < np ? p0 : np - 1) * 8u;
#pragma unroll
  for (int v = 0; v < 6; ++v) in[v] = ldg<double>(B.bv + v * np, o8);
  in[6] = ldg<double>(B.cv, o8);
#pragma unroll
  for (int l = 0; l < NS; ++l) in[7 + l] = ldg<double>(B.cv + (5 + l) * np, o8);
}

__global__ void __launch_bounds__(kTpb) props_mlp_kernel(const BlockView *__restrict__ views,
                                                         const int *__restrict__ off, int nb, int total,
                                                         const SpeciesDev *__restrict__ spg, SchemeDev S,
                                                         const MlpDev *__restrict__ mlp,
                                                         const bf16x8 *__restrict__ ops,
                                                         const float *__restrict__ bias,
                                                         unsigned long long *__restrict__ counters) {
  __shared__ float buf[kTpb / 64][64][kMlpIO + 1];
  __shared__ bf16x8 ops_l[kMlpInfBlocks * 64];
  __shared__ f32x4 bias_l[4 * kMlpH / 4];
  for (int e = threadIdx.x; e < kMlpInfBlocks * 64; e += kTpb) ops_l[e] = ops[e];
  if (threadIdx.x < kMlpH) bias_l[threadIdx.x] = reinterpret_cast<const f32x4 *>(bias)[threadIdx.x];
  __syncthreads();
  const lds_ptr<const bf16x8> lops = (lds_ptr<const bf16x8>)ops_l;
  const lds_ptr<const f32x4> lbias = (lds_ptr<const f32x4>)bias_l;
  const SpeciesPtrT<float> sp = const_table<float>(spg);
  float(*xl)[kMlpIO + 1] = buf[threadIdx.x >> 6];
  double in[kMlpIn];
  int w = blockIdx.x;
  mlp_fetch(views, off, nb, w, in);
  // the sampled check takes every check_every-th chunk-wavefront of each
  // workgroup, phase-shifted by the workgroup, so every workgroup carries
  // the same share of Wilke evaluations
  int ck = blockIdx.x + (threadIdx.x >> 6);
  for (; w < total; w += gridDim.x, ck += kTpb / 64) {
    const int gw = w + off[0];
    const int blk = find_seg(off, nb, gw);
    const GView B = gview(views[blk]);
    const long long np = B.np;
    const long long p0 = (long long)(gw - off[blk]) * kTpb + threadIdx.x;
    const bool valid = p0 < np;
    const long long p = valid ? p0 : np - 1;  // every lane takes part in the MFMA tiles
    // field arrays are addressed by a wave-uniform base and a 32-bit byte offset
    const unsigned o8 = unsigned(p) * 8u, o4 = unsigned(p) * 4u;
    float b32[6], Y[NSA];
#pragma unroll
    for (int v = 0; v < 6; ++v) b32[v] = float(in[v]);
    const double ir = 1.0 / in[6];
#pragma unroll
    for (int l = 0; l < NS; ++l) Y[l] = float(in[7 + l] * ir);
    wave_sync();  // the previous chunk's reads of xl come first
    const bool outside = mlp_phase1(B, sp, mlp, b32, Y, xl, valid, o4);
    if (w + (int)gridDim.x < total) mlp_fetch(views, off, nb, w + gridDim.x, in);
    mlp_phase23(B, sp, S, mlp, lops, lbias, counters, xl, outside, valid, o8, o4, ck % mlp->check_every == 0);
  }
}

int launch_props_mlp(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, const MlpDev *mlp, const void *ops,
                     const float *bias, unsigned long long *counters, hipStream_t st) {
  if constexpr (NS == 0 || NS > kMlpMaxSp) {
    return CFX_E_UNSUPPORTED;
  } else {
    if (b.total <= 0) return 0;
    // persistent grid: the resident workgroups of the current device (cached
    // per device; concurrent first calls compute the same value)
    constexpr int kMaxDev = 64;
    static std::atomic<int> resident[kMaxDev] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    int r = dev >= 0 && dev < kMaxDev ? resident[dev].load(std::memory_order_relaxed) : 0;
    if (!r) {
      int n_cu = 0, per_cu = 0;
      if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) n_cu = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, props_mlp_kernel, kTpb, 0) != hipSuccess)
        per_cu = 1;
      r = std::max(1, n_cu * std::max(1, per_cu));
      if (dev >= 0 && dev < kMaxDev) resident[dev].store(r, std::memory_order_relaxed);
    }
    const int grid = std::min(b.total, r);
    props_mlp_kernel<<<grid, kTpb, 0, st>>>(b.views, b.off, b.nb, b.total, sp, S, mlp, (const bf16x8 *)ops, bias,
                                           counters);
    return (int)hipGetLastError();
  }
}

struct Register32 {
  Register32() {
    KernelSet32 &k = g_kernels32[NS];
    k.props_mlp = launch_props_mlp;
    k.props = launch_props<float>;
    register_rhs<float>(k);
    k.present = 1;
  }
} register32;

#endif

}  // namespace

}  // namespace cfx
