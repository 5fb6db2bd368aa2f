// device_physics.h — point-wise physics on the device: NASA-7 thermo,
// Chapman–Enskog/Wilke transport, limiters, the AUSM+ face flux with MUSCL
// reconstruction and the central viscous face flux.  Templated on the
// species count NS (compile time, so all species loops unroll and the
// species table is read with uniform scalar loads) and on the arithmetic
// type R of the flux/transport work (double for the plain path, float for
// the mixed-precision path).
#pragma once

#include <hip/hip_runtime.h>

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

// ---------------------------------------------------------------- thermo ---
// h_l(T) and cp_l(T) of one species (Thermo.cu:46-66): NASA-7 polynomial in
// Horner form with the low/high set chosen at t_mid and the linear enthalpy
// extension with frozen cp below t_low.
template <typename R>
__device__ __forceinline__ void species_h_cp(SpeciesPtrT<R> sp, int l, R t, R &h, R &cp) {
  const bool hi = t >= R(sp->t_mid[l]);
  const R c0 = hi ? R(sp->cpc[1][l][0]) : R(sp->cpc[0][l][0]), c1 = hi ? R(sp->cpc[1][l][1]) : R(sp->cpc[0][l][1]),
          c2 = hi ? R(sp->cpc[1][l][2]) : R(sp->cpc[0][l][2]), c3 = hi ? R(sp->cpc[1][l][3]) : R(sp->cpc[0][l][3]),
          c4 = hi ? R(sp->cpc[1][l][4]) : R(sp->cpc[0][l][4]);
  const R e0 = hi ? R(sp->hc[1][l][0]) : R(sp->hc[0][l][0]), e1 = hi ? R(sp->hc[1][l][1]) : R(sp->hc[0][l][1]),
          e2 = hi ? R(sp->hc[1][l][2]) : R(sp->hc[0][l][2]), e3 = hi ? R(sp->hc[1][l][3]) : R(sp->hc[0][l][3]),
          e4 = hi ? R(sp->hc[1][l][4]) : R(sp->hc[0][l][4]), e5 = hi ? R(sp->hc[1][l][5]) : R(sp->hc[0][l][5]);
  const R cpp = c0 + t * (c1 + t * (c2 + t * (c3 + t * c4)));
  const R hp = t * (e0 + t * (e1 + t * (e2 + t * (e3 + t * e4)))) + e5;
  const R tl = R(sp->t_low[l]);
  const bool below = t < tl;
  cp = below ? R(sp->cp_tlow[l]) : cpp;
  h = below ? R(sp->h_tlow[l]) + R(sp->cp_tlow[l]) * (t - tl) : hp;
}

// cp_l(T) only, frozen below t_low (Thermo.cu:68-82)
template <typename R>
__device__ __forceinline__ R species_cp(SpeciesPtrT<R> sp, int l, R t) {
  const bool hi = t >= R(sp->t_mid[l]);
  const R c0 = hi ? R(sp->cpc[1][l][0]) : R(sp->cpc[0][l][0]), c1 = hi ? R(sp->cpc[1][l][1]) : R(sp->cpc[0][l][1]),
          c2 = hi ? R(sp->cpc[1][l][2]) : R(sp->cpc[0][l][2]), c3 = hi ? R(sp->cpc[1][l][3]) : R(sp->cpc[0][l][3]),
          c4 = hi ? R(sp->cpc[1][l][4]) : R(sp->cpc[0][l][4]);
  const R cpp = c0 + t * (c1 + t * (c2 + t * (c3 + t * c4)));
  return t < R(sp->t_low[l]) ? R(sp->cp_tlow[l]) : cpp;
}

// -------------------------------------------------------------- limiters ---
// minmod (Limiter.cuh:135-138) and Van Albada (tag 1)
template <typename R>
__device__ __forceinline__ R limiter(int tag, R a, R b) {
  if (tag == 1) {
    const R ab = a * b;
    return ab > R(0) ? ab * (a + b) / (a * a + b * b + Tiny<R>::v) : R(0);
  }
  const R m = fmin(fabs(a), fabs(b));
  return (a > R(0) && b > R(0)) ? m : ((a < R(0) && b < R(0)) ? -m : R(0));
}

// MUSCL kappa = 1/3 on one variable (Reconstruction.cu:59-77), or first
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

// AUSM+ flux (InviscidScheme.cu:235-304) from the reconstructed face states
// pl/pr = (rho,u,v,w,p,Y..) and the face area vector k = 0.5(J_L grad_L +
// J_R grad_R) of the face direction; the face temperature, energy and gamma
// follow Reconstruction.cu:79-122.
template <int NS, typename R>
__device__ __forceinline__ void ausm_from_states(SpeciesPtrT<R> sp, const SchemeDev &S, const R *pl,
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
    const R t_l = pl[4] / (pl[0] * Ru * wl);
    const R t_r = pr[4] / (pr[0] * Ru * wr);
    R cpl = R(0), cpr = R(0);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      R h, cp;
      species_h_cp<R>(sp, l, t_l, h, cp);
      el += h * pl[5 + l];
      cpl += cp * pl[5 + l];
      species_h_cp<R>(sp, l, t_r, h, cp);
      er += h * pr[5 + l];
      cpr += cp * pr[5 + l];
    }
    E_l = pl[0] * el - pl[4];
    E_r = pr[0] * er - pr[4];
    gam_l = cpl / (cpl - Ru * wl);
    gam_r = cpr / (cpr - Ru * wr);
  } else {
    const R g1 = R(1) / R(S.gamma_air - 1);
    E_l = el * pl[0] + pl[4] * g1;
    E_r = er * pr[0] + pr[4] * g1;
    gam_l = gam_r = R(S.gamma_air);
  }
  const R gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const R igk = R(1) / gk;
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) * igk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) * igk;
  const R c = R(0.5) * (sqrt(gam_l * pl[4] / pl[0]) + sqrt(gam_r * pr[4] / pr[0]));
  const R ic = R(1) / c;
  const R ml = ul * ic, mr = ur * ic;
  constexpr double alpha = 3 / 16.0;
  R mlp, plp, mrn, prn;
  if (fabs(ml) > R(1)) {
    mlp = R(0.5) * (ml + fabs(ml));
    plp = mlp / ml;
  } else {
    const R a = (ml + R(1)) * (ml + R(1)) * R(0.25);
    const R b = (ml * ml - R(1)) * (ml * ml - R(1));
    mlp = a + R(0.125) * b;
    plp = a * (R(2) - ml) + R(alpha) * ml * b;
  }
  if (fabs(mr) > R(1)) {
    mrn = R(0.5) * (mr - fabs(mr));
    prn = mrn / mr;
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
  F[4] = coeff * ((left ? E_l : E_r) + (left ? pl[4] : pr[4])) / rho;
#pragma unroll
  for (int l = 5; l < NV; ++l) F[l] = coeff * (left ? pl[l] : pr[l]);
}

// ------------------------------------------------------- inviscid flux ----
// AUSM+ flux of the face between padded points p and p+s in direction d
// (InviscidScheme.cu:220-304) with the face states from MUSCL
// (Reconstruction.cu:79-122).  F has 5+NS entries.
template <int NS, typename R>
__device__ __forceinline__ void inviscid_flux(const BlockView &B, SpeciesPtrT<R> sp,
                                              const SchemeDev &S, int d, long long p, long long s, R *F) {
  constexpr int NV = 5 + NS;
  const long long np = B.np;
  R pl[NV], pr[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double *q = l < 5 ? B.bv + l * np : B.yk + (l - 5) * np;
    const R q0 = R(q[p - s]), q1 = R(q[p]), q2 = R(q[p + s]), q3 = R(q[p + 2 * s]);
    muscl<R>(S, q0, q1, q2, q3, pl[l], pr[l]);
  }
  const double *M = B.metric + 3 * d * np;
  const R jl = R(B.jac[p]), jr = R(B.jac[p + s]);
  const R k1 = R(0.5) * (jl * R(M[p]) + jr * R(M[p + s]));
  const R k2 = R(0.5) * (jl * R(M[np + p]) + jr * R(M[np + p + s]));
  const R k3 = R(0.5) * (jl * R(M[2 * np + p]) + jr * R(M[2 * np + p + s]));
  ausm_from_states<NS, R>(sp, S, pl, pr, k1, k2, k3, F);
}

// -------------------------------------------------------- viscous flux ----
// Central 2nd-order viscous flux of the face between p and p+s_d
// (SecOrdViscScheme::compute_fv/gv/hv, ViscousScheme.cu:11-342), generic in
// the face direction d.  Quirk bits reproduce the reference's face-metric
// averaging (fv, hv) and the (i,j+1,k) rho*D neighbour of hv.
template <int NS, typename R>
__device__ __forceinline__ void viscous_flux(const BlockView &B, SpeciesPtrT<R> sp,
                                             const SchemeDev &S, int d, long long p, const long long st[3],
                                             R *G) {
  const long long np = B.np;
  const long long pd = p + st[d];
  const double *M = B.metric;
  R mf[9];
  const bool q5 = (S.quirks & CFX_Q_METRIC_AVG) && d != 1;
#pragma unroll
  for (int rc = 0; rc < 9; ++rc) {
    const R a = R(M[rc * np + p]);
    mf[rc] = (q5 && rc != 0) ? a : R(0.5) * (a + R(M[rc * np + pd]));
  }
  // derivatives along (xi, eta, zeta): normal one-sided, tangential 4-point
  auto grad = [&](const double *q, R g[3]) {
    const R a = R(q[p]), b = R(q[pd]);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      if (t == d) {
        g[t] = b - a;
      } else {
        const long long o = st[t];
        g[t] = R(0.25) * (R(q[p + o]) - R(q[p - o]) + R(q[pd + o]) - R(q[pd - o]));
      }
    }
  };
  auto chain = [&](const R g[3], R out[3]) {
    out[0] = g[0] * mf[0] + g[1] * mf[3] + g[2] * mf[6];
    out[1] = g[0] * mf[1] + g[1] * mf[4] + g[2] * mf[7];
    out[2] = g[0] * mf[2] + g[1] * mf[5] + g[2] * mf[8];
  };
  R g[3], du[3], dv[3], dw[3], dT[3];
  grad(B.bv + np, g);
  chain(g, du);
  grad(B.bv + 2 * np, g);
  chain(g, dv);
  grad(B.bv + 3 * np, g);
  chain(g, dw);
  grad(B.bv + 5 * np, g);
  chain(g, dT);
  const double *mu = B.props + 2 * np;
  const double *kap = B.props + 3 * np;
  const R vis = R(0.5) * (R(mu[p]) + R(mu[pd]));
  const R third = R(1.0 / 3.0);
  const R txx = vis * (R(4) * du[0] - R(2) * dv[1] - R(2) * dw[2]) * third;
  const R tyy = vis * (R(4) * dv[1] - R(2) * du[0] - R(2) * dw[2]) * third;
  const R tzz = vis * (R(4) * dw[2] - R(2) * du[0] - R(2) * dv[1]) * third;
  const R txy = vis * (du[1] + dv[0]);
  const R txz = vis * (du[2] + dw[0]);
  const R tyz = vis * (dv[2] + dw[1]);
  const double *Md = M + 3 * d * np;
  const R j0 = R(B.jac[p]), j1 = R(B.jac[pd]);
  const R S0 = R(0.5) * (R(Md[p]) * j0 + R(Md[pd]) * j1);
  const R S1 = R(0.5) * (R(Md[np + p]) * j0 + R(Md[np + pd]) * j1);
  const R S2 = R(0.5) * (R(Md[2 * np + p]) * j0 + R(Md[2 * np + pd]) * j1);
  G[0] = R(0);
  G[1] = S0 * txx + S1 * txy + S2 * txz;
  G[2] = S0 * txy + S1 * tyy + S2 * tyz;
  G[3] = S0 * txz + S1 * tyz + S2 * tzz;
  const R um = R(0.5) * (R(B.bv[np + p]) + R(B.bv[np + pd]));
  const R vm = R(0.5) * (R(B.bv[2 * np + p]) + R(B.bv[2 * np + pd]));
  const R wm = R(0.5) * (R(B.bv[3 * np + p]) + R(B.bv[3 * np + pd]));
  const R cond = R(0.5) * (R(kap[p]) + R(kap[pd]));
  G[4] = um * G[1] + vm * G[2] + wm * G[3] + cond * (S0 * dT[0] + S1 * dT[1] + S2 * dT[2]);
  if constexpr (NS > 0) {
    const long long pD = (d == 2 && (S.quirks & CFX_Q_HV_RHOD_J)) ? p + st[1] : pd;
    const R tm = R(0.5) * (R(B.bv[5 * np + p]) + R(B.bv[5 * np + pd]));
    const double *rd = B.props + 4 * np;
    R proj[NS], dif[NS];
    R ucx = R(0), ucy = R(0), ucz = R(0);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      R h, cp;
      species_h_cp<R>(sp, l, tm, h, cp);
      dif[l] = R(0.5) * (R(rd[l * np + p]) + R(rd[l * np + pD]));
      R gy[3], dy[3];
      grad(B.yk + l * np, gy);
      chain(gy, dy);
      proj[l] = S0 * dy[0] + S1 * dy[1] + S2 * dy[2];
      G[4] += dif[l] * h * proj[l];
      ucx += dif[l] * dy[0];
      ucy += dif[l] * dy[1];
      ucz += dif[l] * dy[2];
    }
    const R uc = S0 * ucx + S1 * ucy + S2 * ucz;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double *Y = B.yk + l * np;
      G[5 + l] = dif[l] * proj[l] - R(0.5) * (R(Y[p]) + R(Y[pd])) * uc;
    }
  }
}

}  // namespace cfx
