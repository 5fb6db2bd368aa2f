// device_physics.h — point-wise physics on the device: NASA-7 thermo,
// Chapman–Enskog/Wilke transport, limiters, the AUSM+ face flux with MUSCL
// reconstruction and the central viscous face flux.  Templated on the
// species count NS (compile time, so all species loops unroll and the
// species table is read with uniform scalar loads) and on the arithmetic
// type R of the flux/transport work (double for the plain path, float for
// the mixed-precision path).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cfx_types.h"

namespace cfx {

// The species table is read through the constant address space: every index
// into it is uniform across the wavefront, so the loads become scalar loads
// into SGPRs instead of vector loads into VGPRs.
template <typename R>
using SpeciesPtrT = const __attribute__((address_space(4))) SpeciesT<R> *;

template <typename R>
__device__ __forceinline__ SpeciesPtrT<R> const_table(const SpeciesDev *p);
template <>
__device__ __forceinline__ SpeciesPtrT<double> const_table<double>(const SpeciesDev *p) {
  return (SpeciesPtrT<double>)&p->d;
}
template <>
__device__ __forceinline__ SpeciesPtrT<float> const_table<float>(const SpeciesDev *p) {
  return (SpeciesPtrT<float>)&p->f;
}

template <typename R>
struct Tiny;
template <>
struct Tiny<double> {
  static constexpr double v = 1e-300;
};
template <>
struct Tiny<float> {
  static constexpr float v = 1e-30f;
};

// Division, reciprocal and square root of the flux work.  FP64: the hardware
// reciprocal (v_rcp_f64) refined by two Newton steps, r <- r + r (1 - b r),
// and the quotient corrected once, q <- q + r (a - b q): within an ulp of the
// IEEE quotient for the normal, finite operands of the flux work (positive
// denominators well inside the exponent range), without the scaling and
// fix-up steps the IEEE sequence spends on special values.  FP32 (mixed
// precision): the hardware reciprocal and square root (v_rcp_f32,
// v_sqrt_f32, 1 ulp), far below the FP32 rounding already accepted there.
template <typename R>
__device__ __forceinline__ R frcp(R a) { return R(1) / a; }
template <>
__device__ __forceinline__ double frcp<double>(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}
template <>
__device__ __forceinline__ float frcp<float>(float a) { return __builtin_amdgcn_rcpf(a); }
template <typename R>
__device__ __forceinline__ R fdiv(R a, R b) { return a / b; }
template <>
__device__ __forceinline__ double fdiv<double>(double a, double b) {
  const double r = frcp<double>(b);
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}
template <>
__device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R>
__device__ __forceinline__ R fsqrt(R a) { return sqrt(a); }
template <>
__device__ __forceinline__ float fsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }
// exp and log of the transport work: the library functions in FP64; in FP32
// e^a = 2^(a log2 e) and ln a = log2(a) ln 2 with the hardware exp2/log2
// (v_exp_f32, v_log_f32), for the moderate arguments of the transport
// (|a| < 10 in exp, temperatures in log) within a few ulp.
template <typename R>
__device__ __forceinline__ R fexp(R a) { return exp(a); }
template <>
__device__ __forceinline__ float fexp<float>(float a) { return __builtin_amdgcn_exp2f(a * 1.4426950408889634f); }
template <typename R>
__device__ __forceinline__ R flog(R a) { return log(a); }
template <>
__device__ __forceinline__ float flog<float>(float a) { return __builtin_amdgcn_logf(a) * 0.6931471805599453f; }

// ---------------------------------------------------------------- thermo ---
// h_l(T) and cp_l(T) of one species (Thermo.cu:27-47): NASA-7 polynomial in
// Horner form with the low/high set chosen at t_mid and the linear enthalpy
// extension with frozen cp below t_low.  The coefficients are wave-uniform
// (one record per species); T varies per lane.  Both ranges are evaluated
// with uniform operands and the two results selected per lane: selecting per
// coefficient instead costs a move and a select for each one, or (if the
// compiler folds the select into the loads) a vector load from a per-lane
// address.  The chosen value is bit-for-bit the same.

// Marks v as computed on every lane at this point.  Without it the compiler
// turns the selects below into branches and sinks the polynomial work (and
// its coefficient loads, each with its own wait) into them.
template <typename R>
__device__ __forceinline__ R on_all_lanes(R v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The NASA-7 range every lane of the wavefront is in, for every species:
// kRangeHi (all in the high set), kRangeLo (all in the low set, above
// t_low), or kRangeMixed.  Smooth fields put most wavefronts on one side of
// t_mid; those evaluate one polynomial set instead of both (the value is
// the same bit for bit: the general path selects the same expression).
constexpr int kRangeMixed = 0, kRangeLo = 1, kRangeHi = 2;
template <typename R>
__device__ __forceinline__ int thermo_range(SpeciesPtrT<R> sp, R t0, R t1) {
  const R hf = R(sp->hi_from), lf = R(sp->lo_from), lt = R(sp->lo_to);
  if (__all(t0 >= hf && t1 >= hf)) return kRangeHi;
  if (__all(t0 >= lf && t0 < lt && t1 >= lf && t1 < lt)) return kRangeLo;
  return kRangeMixed;
}

// h and cp of one species from its thermo record c (c[i] = value i of the
// record: constant memory read with scalar loads, or a copy in LDS).
// RG: kRangeMixed (per-lane selection), or the wavefront's single range.
template <typename R, int RG = kRangeMixed, typename C>
__device__ __forceinline__ void thermo_h_cp(const C &c, R t, R &h, R &cp) {
  auto p5 = [&](int o) { return R(c[o]) + t * (R(c[o + 1]) + t * (R(c[o + 2]) + t * (R(c[o + 3]) + t * R(c[o + 4])))); };
  if constexpr (RG != kRangeMixed) {
    const int r = RG == kRangeHi ? 1 : 0;
    cp = p5(kThCp + 5 * r);
    h = t * p5(kThH + 6 * r) + R(c[kThH + 6 * r + 5]);
    return;
  }
  const R cp_lo = p5(kThCp), cp_hi = p5(kThCp + 5);
  const R h_lo = t * p5(kThH) + R(c[kThH + 5]);
  const R h_hi = t * p5(kThH + 6) + R(c[kThH + 11]);
  const bool hi = t >= R(c[kThTmid]);
  const R cpp = on_all_lanes<R>(hi ? cp_hi : cp_lo);
  const R hp = on_all_lanes<R>(hi ? h_hi : h_lo);
  const R tl = R(c[kThTlow]);
  const bool below = t < tl;
  cp = below ? R(c[kThCptlow]) : cpp;
  h = below ? R(c[kThHtlow]) + R(c[kThCptlow]) * (t - tl) : hp;
}

template <typename R>
__device__ __forceinline__ void species_h_cp(SpeciesPtrT<R> sp, int l, R t, R &h, R &cp) {
  thermo_h_cp<R>(sp->th[l].v, t, h, cp);
}

// h_l(T) alone (the same value as thermo_h_cp's h), for loops that need no
// cp: the marker on the selected cp would otherwise keep both cp
// polynomials alive.
template <typename R, int RG = kRangeMixed>
__device__ __forceinline__ R species_h(SpeciesPtrT<R> sp, int l, R t) {
  const auto &c = sp->th[l].v;
  auto p5 = [&](int o) { return R(c[o]) + t * (R(c[o + 1]) + t * (R(c[o + 2]) + t * (R(c[o + 3]) + t * R(c[o + 4])))); };
  if constexpr (RG != kRangeMixed) {
    const int r = RG == kRangeHi ? 1 : 0;
    return t * p5(kThH + 6 * r) + R(c[kThH + 6 * r + 5]);
  }
  const R h_lo = t * p5(kThH) + R(c[kThH + 5]);
  const R h_hi = t * p5(kThH + 6) + R(c[kThH + 11]);
  const R hp = on_all_lanes<R>(t >= R(c[kThTmid]) ? h_hi : h_lo);
  const R tl = R(c[kThTlow]);
  return t < tl ? R(c[kThHtlow]) + R(c[kThCptlow]) * (t - tl) : hp;
}

// cp_l(T) only, frozen below t_low (Thermo.cu:49-63).  Loops over species
// that need only cp read the per-array layout: the cp coefficients of all
// species are contiguous there and arrive in a few wide loads.
template <typename R>
__device__ __forceinline__ R species_cp(SpeciesPtrT<R> sp, int l, R t) {
  auto p5 = [&](const auto &c) { return R(c[0]) + t * (R(c[1]) + t * (R(c[2]) + t * (R(c[3]) + t * R(c[4])))); };
  R cpp;
  if constexpr (sizeof(R) == 4) {
    // both ranges from scalar loads, then a per-lane select
    const R lo = p5(sp->cpc[0][l]), hi = p5(sp->cpc[1][l]);
    cpp = on_all_lanes<R>(t >= R(sp->t_mid[l]) ? hi : lo);
  } else {
    cpp = t >= R(sp->t_mid[l]) ? p5(sp->cpc[1][l]) : p5(sp->cpc[0][l]);
  }
  return t < R(sp->t_low[l]) ? R(sp->cp_tlow[l]) : cpp;
}

// Where a kernel reads the thermo records from: constant memory, or a copy
// of the species' records in LDS (loaded once per workgroup).
template <typename R>
struct ThermoConst {
  SpeciesPtrT<R> sp;
  __device__ auto rec(int l) const { return sp->th[l].v; }
};
template <typename R>
struct ThermoShared {
  const R *p;
  __device__ const R *rec(int l) const { return p + l * kThLen; }
};

// -------------------------------------------------------------- limiters ---
// minmod (Limiter.cuh:13-16) and Van Albada (tag 1)
template <typename R>
__device__ __forceinline__ R limiter(int tag, R a, R b) {
  if (tag == 1) {
    const R ab = a * b;
    return ab > R(0) ? fdiv<R>(ab * (a + b), a * a + b * b + Tiny<R>::v) : R(0);
  }
  const R m = fmin(fabs(a), fabs(b));
  return (a > R(0) && b > R(0)) ? m : ((a < R(0) && b < R(0)) ? -m : R(0));
}

// MUSCL kappa = 1/3 on one variable (Reconstruction.cu:26-44), or first
// order.  Both limiters are symmetric, lim(a,b) == lim(b,a) bit for bit, so
// the reference's four calls per face reduce to the two limited slopes of
// the two cells: (1-kappa)/4 + (1+kappa)/4 = 1/2.
template <typename R>
__device__ __forceinline__ void muscl(const SchemeDev &S, R q0, R q1, R q2, R q3, R &l, R &r) {
  if (S.reconstruction == 2) {
    const R sl = limiter<R>(S.limiter, q1 - q0, q2 - q1);
    const R sr = limiter<R>(S.limiter, q2 - q1, q3 - q2);
    l = q1 + R(0.5) * sl;
    r = q2 - R(0.5) * sr;
  } else {
    l = q1;
    r = q2;
  }
}

// AUSM+ flux (InviscidScheme.cu:39-106) from the reconstructed face states
// pl/pr = (rho,u,v,w,p,Y..) and the face area vector k = 0.5(J_L grad_L +
// J_R grad_R) of the face direction; the face temperature, energy and gamma
// follow Reconstruction.cu:46-89.
template <int NS, typename R, typename TH>
__device__ __forceinline__ void ausm_from_states(SpeciesPtrT<R> sp, const TH &th, const SchemeDev &S, const R *pl,
                                                 const R *pr, R k1, R k2, R k3, R *F) {
  constexpr int NV = 5 + NS;
  R el = R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  R er = R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  R gam_l, gam_r, E_l, E_r;
  if constexpr (NS > 0) {
    R wl = R(0), wr = R(0);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      wl += pl[5 + l] * R(sp->inv_mw[l]);
      wr += pr[5 + l] * R(sp->inv_mw[l]);
    }
    const R Ru = R(kRu);
    const R t_l = fdiv<R>(pl[4], pl[0] * Ru * wl);
    const R t_r = fdiv<R>(pr[4], pr[0] * Ru * wr);
    R cpl = R(0), cpr = R(0);
    auto species_loop = [&](auto rgc) {
      constexpr int RG = decltype(rgc)::value;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        R h, cp;
        thermo_h_cp<R, RG>(th.rec(l), t_l, h, cp);
        el += h * pl[5 + l];
        cpl += cp * pl[5 + l];
        thermo_h_cp<R, RG>(th.rec(l), t_r, h, cp);
        er += h * pr[5 + l];
        cpr += cp * pr[5 + l];
      }
    };
    const int rg = thermo_range<R>(sp, t_l, t_r);
    if (rg == kRangeHi) species_loop(std::integral_constant<int, kRangeHi>{});
    else if (rg == kRangeLo) species_loop(std::integral_constant<int, kRangeLo>{});
    else species_loop(std::integral_constant<int, kRangeMixed>{});
    E_l = pl[0] * el - pl[4];
    E_r = pr[0] * er - pr[4];
    gam_l = fdiv<R>(cpl, cpl - Ru * wl);
    gam_r = fdiv<R>(cpr, cpr - Ru * wr);
  } else {
    const R g1 = R(1) / R(S.gamma_air - 1);
    E_l = el * pl[0] + pl[4] * g1;
    E_r = er * pr[0] + pr[4] * g1;
    gam_l = gam_r = R(S.gamma_air);
  }
  const R gk = fsqrt<R>(k1 * k1 + k2 * k2 + k3 * k3);
  const R igk = frcp<R>(gk);
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) * igk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) * igk;
  const R c = R(0.5) * (fsqrt<R>(fdiv<R>(gam_l * pl[4], pl[0])) + fsqrt<R>(fdiv<R>(gam_r * pr[4], pr[0])));
  const R ic = frcp<R>(c);
  const R ml = ul * ic, mr = ur * ic;
  constexpr double alpha = 3 / 16.0;
  R mlp, plp, mrn, prn;
  if (fabs(ml) > R(1)) {
    mlp = R(0.5) * (ml + fabs(ml));
    plp = fdiv<R>(mlp, ml);
  } else {
    const R a = (ml + R(1)) * (ml + R(1)) * R(0.25);
    const R b = (ml * ml - R(1)) * (ml * ml - R(1));
    mlp = a + R(0.125) * b;
    plp = a * (R(2) - ml) + R(alpha) * ml * b;
  }
  if (fabs(mr) > R(1)) {
    mrn = R(0.5) * (mr - fabs(mr));
    prn = fdiv<R>(mrn, mr);
  } else {
    const R a = (mr - R(1)) * (mr - R(1)) * R(0.25);
    const R b = (mr * mr - R(1)) * (mr * mr - R(1));
    mrn = -a - R(0.125) * b;
    prn = a * (R(2) + mr) - R(alpha) * mr * b;
  }
  const R p_coeff = plp * pl[4] + prn * pr[4];
  const R m_half = mlp + mrn;
  const R mpos = R(0.5) * (m_half + fabs(m_half));
  const R mneg = R(0.5) * (m_half - fabs(m_half));
  const R mass = c * (pl[0] * mpos + pr[0] * mneg);
  const R coeff = mass * gk;
  const bool left = mass >= R(0);
  const R rho = left ? pl[0] : pr[0];
  F[0] = coeff;
  F[1] = coeff * (left ? pl[1] : pr[1]) + p_coeff * k1;
  F[2] = coeff * (left ? pl[2] : pr[2]) + p_coeff * k2;
  F[3] = coeff * (left ? pl[3] : pr[3]) + p_coeff * k3;
  F[4] = fdiv<R>(coeff * ((left ? E_l : E_r) + (left ? pl[4] : pr[4])), rho);
#pragma unroll
  for (int l = 5; l < NV; ++l) F[l] = coeff * (left ? pl[l] : pr[l]);
}

}  // namespace cfx
