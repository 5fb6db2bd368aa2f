// cfx_types.h — plain structs shared by the host side of the library and the
// HIP kernels: the per-block device view, the species table and the scheme
// constants.  All device arrays are structure-of-arrays over the ghost-padded
// block with i fastest (see include/cfx.h for the index formula).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cfx.h"

namespace cfx {

constexpr double kRu = 8314.462618;  // universal gas constant, J/(kmol K) (Constants.h:19)
constexpr double kMwAir = 28.9647;   // Constants.h:28
constexpr int kNgg = 2;              // ghost layers for MUSCL (Parameter.cpp:32)
constexpr int kMaxSp = CFX_MAX_SPEC;

// Species table resident in device memory.  Everything the kernels need is
// precomputed here so that the per-point loops are polynomial evaluations:
//   cp_l(T) = sum_k cpc[r][l][k] T^k,   h_l(T) = T*sum_k hc[r][l][k] T^k + hc[r][l][5]
// with r = 0 for T < t_mid (low set) and r = 1 otherwise, already scaled by
// R_u/W_l; below t_low the enthalpy is extended linearly with the frozen cp
// (h_tlow, cp_tlow), as Thermo.cu:30-37 does.  Held in the arithmetic type
// of the kernel that reads it, so that every constant is a scalar operand.
// One species' NASA-7 thermo as one contiguous, 64-byte aligned record of
// 32 values, read by the point loops as a single vector (two or four wide
// scalar loads, one wait) instead of one scalar load per array:
//   [kThCp + 5r + k]  cp polynomial coefficient k of range r (0 low, 1 high)
//   [kThH + 6r + k]   h = T * (h_0 + h_1 T + ... + h_4 T^4) + h_5, range r
//   [kThTmid], [kThTlow], [kThHtlow], [kThCptlow]
constexpr int kThCp = 0, kThH = 10, kThTmid = 22, kThTlow = 23, kThHtlow = 24, kThCptlow = 25, kThLen = 32;
template <typename T>
struct alignas(64) ThermoRec {
  T v[kThLen];
};

template <typename T>
struct SpeciesT {
  ThermoRec<T> th[kMaxSp];   // the thermo arrays below, regrouped per species
  int ns;
  T mw[kMaxSp], inv_mw[kMaxSp], r_sp[kMaxSp];  // W, 1/W, R_u/W
  T t_low[kMaxSp], t_mid[kMaxSp];
  T cpc[2][kMaxSp][5];
  T hc[2][kMaxSp][6];
  T h_tlow[kMaxSp], cp_tlow[kMaxSp];
  T lj_inv[kMaxSp], vis_coeff[kMaxSp];
  T om145[kMaxSp];            // (k/eps)^-0.145: T*^-0.145 = T^-0.145 om145
  T cond_add[kMaxSp];        // 1.25 R_u / W_l (Transport.cu:59)
  T wj4[kMaxSp][kMaxSp];     // (W_n/W_m)^(1/4)           (ChemData.cpp:257)
  T sqw8[kMaxSp][kMaxSp];    // 1/sqrt(8(1+W_m/W_n))      (ChemData.cpp:258)
  // range bounds over the species: T >= hi_from puts every species in its
  // high set (and above its t_low), lo_from <= T < lo_to every species in its
  // low set above t_low (lo_from > lo_to when no such T exists)
  T hi_from, lo_from, lo_to;
};

// The FP64 table and its FP32 rounding, side by side in one allocation.
struct SpeciesDev {
  SpeciesT<double> d;
  SpeciesT<float> f;
};

struct SchemeDev {
  int reconstruction, limiter, viscous, quirks, rk, dim;
  double cfl, pr, sc, gamma_air, r_air;
  double inv_sc, inv_pr;
  int lusgs;     // implicit LU-SGS step: the update adds dq (the increment)
  double omega;  // LU-SGS over-relaxation of the inviscid spectral radius
};

// One block as seen by the kernels.
struct BlockView {
  int mx, my, mz;
  int px, py, pz;   // padded extents
  long long np;     // padded point count
  long long ni;     // interior point count
  long long sy, sz; // padded strides in j and k
  double *cv, *bv, *yk, *props;  // props: gamma, c, mu, kappa, rhoD[ns]
  double *metric, *jac;          // metric: 9 x np (xi_x..zeta_z), jac: np
  double *dq, *qn, *dt_local, *bv_last;
  double *lus;  // LU-SGS: D, then a_d for d = 0..2, each over the interior (null otherwise)
  // mixed precision: FP32 copies of the RHS inputs (refreshed by the
  // properties kernel over the padded range) and the FP32 dq; null in FP64 mode
  int mixed;
  float *bv32, *yk32, *props32, *metric32, *jac32, *dq32;
  // fuse[2a + s]: the outflow condition on face (axis a, low s = 0 / high
  // s = 1) covers the whole face and is applied by the update kernel to the
  // ghost layers of the face's interior rows (see the ghost-fill schedule)
  int fuse[6];
  int fuse_any;
};

// Pointers in the global address space.  A view loaded from device memory
// (the block table) hands out generic pointers, for which the compiler emits
// flat loads and stores; flat accesses also share the LDS wait counter, so
// they serialise the LDS staging with the HBM traffic.  Kernels therefore use
// GView, the same view with every array pointer marked global.
template <class T>
using gptr = __attribute__((address_space(1))) T *;

// Element at a 32-bit byte offset from a wave-uniform base.  The compiler
// can then use the scalar-base + 32-bit vector-offset addressing, so several
// field arrays indexed by the same point share one offset register instead
// of each needing a 64-bit address computed per lane.  Offsets must stay
// below 4 GiB (the library rejects blocks with more than 2^29 padded points).
template <class T>
__device__ __forceinline__ T ldg(gptr<const T> base, unsigned byte_off) {
  return *(gptr<const T>)((gptr<const char>)base + byte_off);
}
template <class T>
__device__ __forceinline__ gptr<T> atg(gptr<T> base, unsigned byte_off) {
  return (gptr<T>)((gptr<char>)base + byte_off);
}

struct GView {
  int mx, my, mz;
  int px, py, pz;
  long long np, ni, sy, sz;
  gptr<double> cv, bv, yk, props, metric, jac, dq, qn, dt_local, bv_last, lus;
  int mixed;
  gptr<float> bv32, yk32, props32, metric32, jac32, dq32;
  int fuse[6];
  int fuse_any;
};

__device__ __forceinline__ GView gview(const BlockView &B) {
  GView g;
  g.mx = B.mx, g.my = B.my, g.mz = B.mz, g.px = B.px, g.py = B.py, g.pz = B.pz;
  g.np = B.np, g.ni = B.ni, g.sy = B.sy, g.sz = B.sz;
  g.cv = (gptr<double>)B.cv, g.bv = (gptr<double>)B.bv, g.yk = (gptr<double>)B.yk;
  g.props = (gptr<double>)B.props, g.metric = (gptr<double>)B.metric, g.jac = (gptr<double>)B.jac;
  g.dq = (gptr<double>)B.dq, g.qn = (gptr<double>)B.qn, g.dt_local = (gptr<double>)B.dt_local;
  g.bv_last = (gptr<double>)B.bv_last, g.lus = (gptr<double>)B.lus;
  g.mixed = B.mixed;
  g.bv32 = (gptr<float>)B.bv32, g.yk32 = (gptr<float>)B.yk32, g.props32 = (gptr<float>)B.props32;
  g.metric32 = (gptr<float>)B.metric32, g.jac32 = (gptr<float>)B.jac32, g.dq32 = (gptr<float>)B.dq32;
  for (int f = 0; f < 6; ++f) g.fuse[f] = B.fuse[f];
  g.fuse_any = B.fuse_any;
  return g;
}

// Field selection by the arithmetic type of the flux work.
template <typename R>
struct Fields;
template <>
struct Fields<double> {
  __device__ static gptr<const double> bv(const GView &B) { return B.bv; }
  __device__ static gptr<const double> yk(const GView &B) { return B.yk; }
  __device__ static gptr<const double> props(const GView &B) { return B.props; }
  __device__ static gptr<const double> metric(const GView &B) { return B.metric; }
  __device__ static gptr<const double> jac(const GView &B) { return B.jac; }
  __device__ static gptr<double> dq(const GView &B) { return B.dq; }
};
template <>
struct Fields<float> {
  __device__ static gptr<const float> bv(const GView &B) { return B.bv32; }
  __device__ static gptr<const float> yk(const GView &B) { return B.yk32; }
  __device__ static gptr<const float> props(const GView &B) { return B.props32; }
  __device__ static gptr<const float> metric(const GView &B) { return B.metric32; }
  __device__ static gptr<const float> jac(const GView &B) { return B.jac32; }
  __device__ static gptr<float> dq(const GView &B) { return B.dq32; }
};

// Inflow state resident on device: rho,u,v,w,p,T, E, then Y[ns].
struct InflowDev {
  double st[7];
  double y[kMaxSp];
};

struct WallDev {
  int adiabatic;
  double temperature;
};

// A registered boundary face (Boundary::register_boundary, Mesh.cpp:14-32).
struct FaceRange {
  int rs[3], re[3];
  int face, direction;
};

// A registered inner face (InnerFace::register_boundary, Mesh.cpp:44-143).
struct InnerDev {
  int rs[3], re[3], face, direction;
  int ts[3], target_block, src_tar[3], loop_dir[3], target_loop_dir[3];
  int n_point[3];
};

// One ghost-fill operation over a box of points: a boundary condition on a
// face of block blk (kind = CFX_BC_*), or an inner-face exchange from
// target block tgt into blk (kind = kOpExchange).  Point n of the op is
// (n % ext[0], n / ext[0] % ext[1], n / (ext[0] ext[1])) relative to the
// range start.
constexpr int kOpExchange = 100;
struct FaceOp {
  int kind, blk, tgt, cond;
  int ext[3];
  long long npts;
  FaceRange r;  // boundary range (wall: shrunk by ngg tangentially)
  InnerDev f;   // exchange mapping
  WallDev wl;
};

// Work segments: workgroup w (counted from off[0]) belongs to segment s with
// off[s] <= w < off[s+1], s in [0, n).  The search is uniform per workgroup.
__device__ __forceinline__ int find_seg(const int *__restrict__ off, int n, int w) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace cfx
