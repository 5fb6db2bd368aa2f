// oracle.cpp — TEST INFRASTRUCTURE, not product code.
//
// A host C++17/OpenMP restatement of the reference's explicit step
// (Liangerty/CombustionCodeGPU) for use as the parity checker by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg only.  Every
// routine follows the reference source text it cites (file:line); the only
// departures are the ones the reference leaves undefined or that the step
// adds, each marked "corrected" or "new":
//   - Transport.cu:88 indexes X with the grid index i (out of range); the
//     oracle uses X[l].                                          (corrected)
//   - Field.cu:216 (n[1] = blockDim.y + blockIdx.y + ...) and the host array
//     passed to the device at Driver.cu:214; the oracle loops the face range
//     with the index map of Mesh.cpp:905-917.                     (corrected)
//   - Newton stop |dT| <= 1e-10 K unless CFX_Q_NEWTON_REL.         (corrected)
//   - Van Albada limiter (tag 1), slip wall (type 3), SSP-RK3 with a global
//     CFL time step (rk_stages == 3).                                   (new)
// Parity status of each function is listed in the header of tests and in the
// repository's design notes.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kRu = 8314.462618;   // Constants.h:19
constexpr double kMwAir = 28.9647;    // Constants.h:28
constexpr int kMaxSp = CFX_MAX_SPEC;
constexpr int kMaxVar = 5 + kMaxSp + 2;

struct Species {
  int ns = 0;
  std::vector<double> mw, t_low, t_mid, t_high, lo, hi, lj_inv, vis_coeff, wj4, sqw8;
  void set(const cfx_species *s) {
    ns = s ? s->n_spec : 0;
    mw.assign(ns, 0); t_low.assign(ns, 0); t_mid.assign(ns, 0); t_high.assign(ns, 0);
    lo.assign(ns * 7, 0); hi.assign(ns * 7, 0); lj_inv.assign(ns, 0); vis_coeff.assign(ns, 0);
    wj4.assign(ns * ns, 0); sqw8.assign(ns * ns, 0);
    for (int l = 0; l < ns; ++l) {
      mw[l] = s->mw[l]; t_low[l] = s->t_low[l]; t_mid[l] = s->t_mid[l]; t_high[l] = s->t_high[l];
      for (int c = 0; c < 7; ++c) { lo[l * 7 + c] = s->coeff_lo[l * 7 + c]; hi[l * 7 + c] = s->coeff_hi[l * 7 + c]; }
      // ChemData.cpp:245-250
      lj_inv[l] = 1.0 / s->eps_k[l];
      vis_coeff[l] = 2.6693e-6 * std::sqrt(mw[l]) / (s->sigma[l] * s->sigma[l]);
    }
    // ChemData.cpp:255-260
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) {
        wj4[i * ns + j] = std::pow(mw[j] / mw[i], 0.25);
        sqw8[i * ns + j] = 1.0 / std::sqrt(8 * (1 + mw[i] / mw[j]));
      }
  }
};

struct Params {
  int reconstruction = 2, limiter = 0, viscous = 2, rk = 1, quirks = CFX_QUIRKS_REFERENCE, lusgs = 0;
  double cfl = 0.5, pr = 0.72, sc = 0.9, gamma_air = 1.4, r_air = kRu / kMwAir, omega = 1.0;
};

inline int isgn(int a) { return (a > 0) - (a < 0); }
inline double dsgn(double a) { return (a > 0) - (a < 0); }

// ---------------------------------------------------------------- thermo ---
// Thermo.cu:27-47 compute_enthalpy_and_cp
void enthalpy_and_cp(const Species &s, double t, double *h, double *cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t}, t5{t4 * t};
  for (int i = 0; i < s.ns; ++i) {
    if (t < s.t_low[i]) {
      const double tt = s.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &s.lo[i * 7];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp[i] * (t - tt);
    } else {
      const double *c = t < s.t_mid[i] ? &s.lo[i * 7] : &s.hi[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= kRu / s.mw[i];
    cp[i] *= kRu / s.mw[i];
  }
}

// Thermo.cu:7-25 compute_enthalpy (same h as above)
void enthalpy(const Species &s, double t, double *h) {
  double cp[kMaxSp];
  enthalpy_and_cp(s, t, h, cp);
}

// Thermo.cu:49-63 compute_cp (cp frozen below t_low)
void cp_only(const Species &s, double t, double *cp) {
  const double t2{t * t}, t3{t2 * t}, t4{t3 * t};
  for (int i = 0; i < s.ns; ++i) {
    if (t < s.t_low[i]) {
      const double tt = s.t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double *c = &s.lo[i * 7];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double *c = t < s.t_mid[i] ? &s.lo[i * 7] : &s.hi[i * 7];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= kRu / s.mw[i];
  }
}

// Thermo.cu:89-125 compute_temperature.  Returns the iteration count; the
// stop rule is |dT| <= 1e-10 K unless quirk CFX_Q_NEWTON_REL (|1-T1/T|<=1e-3).
int newton_t(const Species &s, int quirks, const double *Y, double e, double t, double *t_out,
             double *gas_const_out) {
  double mw{0};
  for (int l = 0; l < s.ns; ++l) mw += Y[l] / s.mw[l];
  mw = 1 / mw;
  const double gas_const = kRu / mw;
  const bool rel = (quirks & CFX_Q_NEWTON_REL) != 0;
  const double eps = rel ? 1e-3 : 1e-10;
  constexpr int max_iter{1000};
  double err{1e30};
  int iter = 0;
  double h_i[kMaxSp], cp_i[kMaxSp];
  while (err > eps && iter++ < max_iter) {
    enthalpy_and_cp(s, t, h_i, cp_i);
    double cp_tot{0}, h{0};
    for (int l = 0; l < s.ns; ++l) {
      cp_tot += cp_i[l] * Y[l];
      h += h_i[l] * Y[l];
    }
    const double e_t = h - gas_const * t;
    const double cv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cv;
    err = rel ? std::abs(1 - t1 / t) : std::abs(t1 - t);
    t = t1;
  }
  *t_out = t;
  *gas_const_out = gas_const;
  return iter;
}

// Transport.cu:7-10
double sutherland(double t) { return 1.716e-5 * std::pow(t / 273, 1.5) * (273 + 111) / (t + 111); }

// Transport.cu:43-96 compute_transport_property (X[l] in the rho*D branch)
void transport(const Species &s, double sc, double t, double mw_total, const double *cp, const double *Y,
               double *mu_out, double *kappa_out, double *rhoD) {
  const int ns = s.ns;
  double X[kMaxSp], vis[kMaxSp], lambda[kMaxSp];
  for (int l = 0; l < ns; ++l) {
    X[l] = Y[l] * mw_total / s.mw[l];
    const double t_dl{t * s.lj_inv[l]};
    const double collision_integral{1.147 * std::pow(t_dl, -0.145) + std::pow(t_dl + 0.5, -2)};
    vis[l] = s.vis_coeff[l] * std::sqrt(t) / collision_integral;
  }
  for (int l = 0; l < ns; ++l) lambda[l] = vis[l] * (cp[l] + 1.25 * kRu / s.mw[l]);
  double phi[kMaxSp * kMaxSp];
  for (int m = 0; m < ns; ++m)
    for (int n = 0; n < ns; ++n) {
      if (m == n) {
        phi[m * ns + n] = 1.0;
      } else {
        const double numerator{1 + std::sqrt(vis[m] / vis[n]) * s.wj4[m * ns + n]};
        phi[m * ns + n] = numerator * numerator * s.sqw8[m * ns + n];
      }
    }
  double viscosity = 0, conductivity = 0;
  for (int m = 0; m < ns; ++m) {
    double vis_temp{0};
    for (int n = 0; n < ns; ++n) vis_temp += phi[m * ns + n] * X[n];
    const double cond_temp = 1.065 * vis_temp - 0.065 * X[m];
    viscosity += vis[m] * X[m] / vis_temp;
    conductivity += lambda[m] * X[m] / cond_temp;
  }
  *mu_out = viscosity;
  *kappa_out = conductivity;
  for (int l = 0; l < ns; ++l) {
    if (std::abs(X[l] - 1) < 1e-3)
      rhoD[l] = viscosity / sc;
    else
      rhoD[l] = (1 - Y[l]) * viscosity / ((1 - X[l]) * sc);
  }
}

// Limiter.cuh:13-16 (minmod, tags 0 and 1 in the reference); tag 1 here is
// Van Albada: ab(a+b)/(a^2+b^2) for ab > 0, else 0 (new).
double limiter(int tag, double a, double b) {
  if (tag == 1) {
    const double ab = a * b;
    return ab > 0 ? ab * (a + b) / (a * a + b * b + 1e-300) : 0.0;
  }
  return 0.5 * (dsgn(a) + dsgn(b)) * std::min(std::abs(a), std::abs(b));
}

// Reconstruction.cu:12-44 (first order / MUSCL kappa=1/3) + 46-89 (face
// thermo) and InviscidScheme.cu:24-108 (AUSM+).  pv4 holds 4 points of
// (rho,u,v,w,p,Y..) with stride nrec; metr is the metric row of the face
// direction at the 2 face points, jac the two jacobians.
void ausm_face(const Species &s, const Params &P, const double *pv4, int nrec, const double *metr,
               const double *jac, double *fc) {
  const int ns = s.ns;
  const int n_var = 5 + ns;
  double pv_l[kMaxVar], pv_r[kMaxVar];
  const double *q0 = pv4, *q1 = pv4 + nrec, *q2 = pv4 + 2 * nrec, *q3 = pv4 + 3 * nrec;
  if (P.reconstruction == 2) {
    constexpr double kappa{1.0 / 3.0};
    for (int l = 0; l < n_var; ++l) {
      const double delta_i{q1[l] - q0[l]};
      const double delta_i1{q2[l] - q1[l]};
      const double delta_i2{q3[l] - q2[l]};
      const double delta_neg_l = limiter(P.limiter, delta_i, delta_i1);
      const double delta_pos_l = limiter(P.limiter, delta_i1, delta_i);
      const double delta_neg_r = limiter(P.limiter, delta_i1, delta_i2);
      const double delta_pos_r = limiter(P.limiter, delta_i2, delta_i1);
      pv_l[l] = q1[l] + 0.25 * ((1 - kappa) * delta_neg_l + (1 + kappa) * delta_pos_l);
      pv_r[l] = q2[l] - 0.25 * ((1 - kappa) * delta_pos_r + (1 + kappa) * delta_neg_r);
    }
  } else {
    for (int l = 0; l < n_var; ++l) { pv_l[l] = q1[l]; pv_r[l] = q2[l]; }
  }
  double el = 0.5 * (pv_l[1] * pv_l[1] + pv_l[2] * pv_l[2] + pv_l[3] * pv_l[3]);
  double er = 0.5 * (pv_r[1] * pv_r[1] + pv_r[2] * pv_r[2] + pv_r[3] * pv_r[3]);
  double gam_l, gam_r;
  if (ns > 0) {
    double mw_inv_l{0.0}, mw_inv_r{0.0};
    for (int l = 0; l < ns; ++l) {
      mw_inv_l += pv_l[5 + l] / s.mw[l];
      mw_inv_r += pv_r[5 + l] / s.mw[l];
    }
    const double t_l = pv_l[4] / (pv_l[0] * kRu * mw_inv_l);
    const double t_r = pv_r[4] / (pv_r[0] * kRu * mw_inv_r);
    double hl[kMaxSp], hr[kMaxSp], cpl_i[kMaxSp], cpr_i[kMaxSp];
    enthalpy_and_cp(s, t_l, hl, cpl_i);
    enthalpy_and_cp(s, t_r, hr, cpr_i);
    double cpl{0}, cpr{0}, cvl{0}, cvr{0};
    for (int l = 0; l < ns; ++l) {
      cpl += cpl_i[l] * pv_l[l + 5];
      cpr += cpr_i[l] * pv_r[l + 5];
      cvl += pv_l[l + 5] * (cpl_i[l] - kRu / s.mw[l]);
      cvr += pv_r[l + 5] * (cpr_i[l] - kRu / s.mw[l]);
      el += hl[l] * pv_l[l + 5];
      er += hr[l] * pv_r[l + 5];
    }
    pv_l[5 + ns] = pv_l[0] * el - pv_l[4];
    pv_r[5 + ns] = pv_r[0] * er - pv_r[4];
    gam_l = cpl / cvl;
    gam_r = cpr / cvr;
  } else {
    pv_l[5] = el * pv_l[0] + pv_l[4] / (P.gamma_air - 1);
    pv_r[5] = er * pv_r[0] + pv_r[4] / (P.gamma_air - 1);
    gam_l = gam_r = P.gamma_air;
  }
  // InviscidScheme.cu:39-106
  constexpr double alpha{3 / 16.0};
  const double *metr_l = metr, *metr_r = metr + 3;
  const double jac_l = jac[0], jac_r = jac[1];
  const double k1 = 0.5 * (jac_l * metr_l[0] + jac_r * metr_r[0]);
  const double k2 = 0.5 * (jac_l * metr_l[1] + jac_r * metr_r[1]);
  const double k3 = 0.5 * (jac_l * metr_l[2] + jac_r * metr_r[2]);
  const double grad_k_div_jac = std::sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pv_l[1] + k2 * pv_l[2] + k3 * pv_l[3]) / grad_k_div_jac;
  const double ur = (k1 * pv_r[1] + k2 * pv_r[2] + k3 * pv_r[3]) / grad_k_div_jac;
  const double pl = pv_l[4], pr = pv_r[4], rho_l = pv_l[0], rho_r = pv_r[0];
  const double c = 0.5 * (std::sqrt(gam_l * pl / rho_l) + std::sqrt(gam_r * pr / rho_r));
  const double mach_l = ul / c, mach_r = ur / c;
  double mlp{0}, mrn{0}, plp{0}, prn{0};
  if (std::abs(mach_l) > 1) {
    mlp = 0.5 * (mach_l + std::abs(mach_l));
    plp = mlp / mach_l;
  } else {
    const double ml_plus1_squared_div4 = (mach_l + 1) * (mach_l + 1) * 0.25;
    const double ml_squared_minus_1_squared = (mach_l * mach_l - 1) * (mach_l * mach_l - 1);
    mlp = ml_plus1_squared_div4 + 0.125 * ml_squared_minus_1_squared;
    plp = ml_plus1_squared_div4 * (2 - mach_l) + alpha * mach_l * ml_squared_minus_1_squared;
  }
  if (std::abs(mach_r) > 1) {
    mrn = 0.5 * (mach_r - std::abs(mach_r));
    prn = mrn / mach_r;
  } else {
    const double mr_minus1_squared_div4 = (mach_r - 1) * (mach_r - 1) * 0.25;
    const double mr_squared_minus_1_squared = (mach_r * mach_r - 1) * (mach_r * mach_r - 1);
    mrn = -mr_minus1_squared_div4 - 0.125 * mr_squared_minus_1_squared;
    prn = mr_minus1_squared_div4 * (2 + mach_r) - alpha * mach_r * mr_squared_minus_1_squared;
  }
  const double p_coeff = plp * pl + prn * pr;
  const double m_half = mlp + mrn;
  const double mach_pos = 0.5 * (m_half + std::abs(m_half));
  const double mach_neg = 0.5 * (m_half - std::abs(m_half));
  const double mass_flux_half = c * (rho_l * mach_pos + rho_r * mach_neg);
  const double coeff = mass_flux_half * grad_k_div_jac;
  const double *pv = mass_flux_half >= 0 ? pv_l : pv_r;
  fc[0] = coeff;
  fc[1] = coeff * pv[1] + p_coeff * k1;
  fc[2] = coeff * pv[2] + p_coeff * k2;
  fc[3] = coeff * pv[3] + p_coeff * k3;
  fc[4] = coeff * (pv[5 + ns] + pv[4]) / pv[0];
  for (int l = 5; l < n_var; ++l) fc[l] = coeff * pv[l];
}

// ----------------------------------------------------------------- block ---
struct Boundary {
  int rs[3], re[3], type, cond, face = 0, direction = 1;
};

struct InnerFace {
  int rs[3], re[3], face = 0, direction = 1;
  int ts[3], te[3], target_block = 0, target_face = 0, target_direction = 1;
  int src_tar[3]{0, 0, 0}, loop_dir[3]{1, 1, 1}, target_loop_dir[3]{1, 1, 1};
};

// Boundary::register_boundary (Mesh.cpp:14-32)
void register_boundary(Boundary &b, int ngg, int dim) {
  for (int i = 0; i < dim; ++i) {
    if (b.rs[i] == b.re[i]) {
      b.face = i;
      if (b.rs[i] == 0) b.direction = -1;
      break;
    }
  }
  for (int i = 0; i < dim; ++i) {
    if (i == b.face) continue;
    if (b.re[i] < b.rs[i]) std::swap(b.re[i], b.rs[i]);
    b.rs[i] -= ngg;
    b.re[i] += ngg;
  }
  if (dim == 2) {
    b.rs[2] = -ngg;
    b.re[2] = ngg;
  }
}

// InnerFace::register_boundary (Mesh.cpp:44-143)
void register_inner(InnerFace &f, int ngg, int dim) {
  for (int i = 0; i < dim; ++i) {
    if (f.rs[i] == f.re[i]) {
      f.face = i;
      if (f.rs[i] == 1) f.direction = -1;
      f.rs[i] -= 1;
      f.re[i] -= 1;
      break;
    }
  }
  for (int i = 0; i < dim; ++i) {
    if (f.ts[i] == f.te[i]) {
      f.target_face = i;
      if (f.ts[i] == 1) f.target_direction = -1;
      f.ts[i] -= 1;
      f.te[i] -= 1;
      break;
    }
  }
  f.src_tar[f.target_face] = -(f.direction * f.target_direction) * (f.face + 1);
  if (dim == 2) {
    const int next_face{1 - f.face}, target_next_face{1 - f.target_face};
    f.rs[next_face] = std::abs(f.rs[next_face]) - 1;
    f.re[next_face] = std::abs(f.re[next_face]) - 1;
    f.ts[next_face] = std::abs(f.ts[next_face]) - 1;
    f.te[next_face] = std::abs(f.te[next_face]) - 1;
    f.src_tar[target_next_face] =
        isgn((f.te[target_next_face] - f.ts[target_next_face]) * (f.re[next_face] - f.rs[next_face])) *
        (next_face + 1);
    f.src_tar[2] = 3;
    f.rs[2] = 0; f.re[2] = 0; f.ts[2] = 0; f.te[2] = 0;
  } else {
    int next_face{0}, target_next_face{0};
    for (int i = 0; i < dim; ++i) {
      if (f.rs[i] < 0) {
        next_face = i;
        f.rs[i] = -f.rs[i] - 1;
        f.re[i] = -f.re[i] - 1;
        break;
      }
    }
    for (int i = 0; i < dim; ++i) {
      if (f.ts[i] < 0) {
        target_next_face = i;
        f.ts[i] = -f.ts[i] - 1;
        f.te[i] = -f.te[i] - 1;
        break;
      }
    }
    f.src_tar[target_next_face] =
        isgn((f.te[target_next_face] - f.ts[target_next_face]) * (f.re[next_face] - f.rs[next_face])) *
        (next_face + 1);
    next_face = 3 - f.face - next_face;
    target_next_face = 3 - f.target_face - target_next_face;
    f.rs[next_face] -= 1;
    f.re[next_face] -= 1;
    f.ts[target_next_face] -= 1;
    f.te[target_next_face] -= 1;
    f.src_tar[target_next_face] =
        isgn((f.te[target_next_face] - f.ts[target_next_face]) * (f.re[next_face] - f.rs[next_face])) *
        (next_face + 1);
  }
  f.rs[f.face] = f.direction < 0 ? -ngg : f.re[f.face];
  f.re[f.face] = f.direction < 0 ? 0 : f.re[f.face] + ngg;
  const int sgn_tar = isgn(f.src_tar[f.target_face]);
  if (f.target_direction > 0) {
    f.ts[f.target_face] = sgn_tar > 0 ? f.te[f.target_face] - ngg : f.te[f.target_face];
  } else {
    f.ts[f.target_face] = (sgn_tar > 0) ? 0 : ngg;
  }
  for (int i = 0; i < dim; ++i) {
    f.loop_dir[i] = f.rs[i] <= f.re[i] ? 1 : -1;
    f.target_loop_dir[i] = isgn(f.src_tar[i]);
    f.src_tar[i] = std::abs(f.src_tar[i]) - 1;
  }
  // In 2-D the reference leaves src_tar[2] = 3 (Mesh.cpp:75,129-133), which
  // Mesh.cpp:913 then uses as an index into di[3]; k maps to k.  (corrected)
  if (dim == 2) f.src_tar[2] = 2;
  for (int i = 0; i < 3; ++i) {
    if (i != f.face) {
      f.rs[i] -= f.loop_dir[i] * ngg;
      f.re[i] += f.loop_dir[i] * ngg;
      f.ts[i] -= f.target_loop_dir[i] * f.loop_dir[i] * ngg;
      f.te[i] += f.target_loop_dir[i] * f.loop_dir[i] * ngg;
    }
  }
}

struct Block {
  int mx, my, mz, ngg, dim;
  int px, py, pz;
  long np, ni;
  // coordinates with ngg+1 ghost layers (Mesh.cpp:220-222)
  int gx, gy, gz;
  std::vector<double> x, y, z;
  std::vector<double> jac, metric;  // metric[9][np], row-major (xi, eta, zeta) x (x, y, z)
  std::vector<double> cv, bv, yk, gamma, c, mu, kappa, rhoD;
  std::vector<double> dq, dt_local, bv_last, qn;
  std::vector<Boundary> bc;
  std::vector<InnerFace> inner;

  long id(int i, int j, int k) const { return ((long)(k + ngg) * py + (j + ngg)) * px + (i + ngg); }
  long gid(int i, int j, int k) const {
    const int g = ngg + 1;
    return ((long)(k + g) * gy + (j + g)) * gx + (i + g);
  }
  long iid(int i, int j, int k) const { return ((long)k * my + j) * mx + i; }
  double &X(int i, int j, int k) { return x[gid(i, j, k)]; }
  double &Y(int i, int j, int k) { return y[gid(i, j, k)]; }
  double &Z(int i, int j, int k) { return z[gid(i, j, k)]; }
};

}  // namespace

struct orc_ctx {
  Species sp;
  Params P;
  std::vector<Block> blocks;
  std::vector<cfx_inflow> inflow;
  std::vector<std::vector<double>> inflow_y;
  std::vector<double> inflow_state;  // per inflow: rho,u,v,w,p,T
  std::vector<cfx_wall> wall;
  int dim = 3, ngg = 2, nv = 5;
  long n_grid_total = 0;
  double dt = 0;
  long long newton_fail = 0;
  std::string err;
};

namespace {

// Mesh.cpp:226-488 Block::compute_jac_metric: central differences of the
// (ngg+1)-layer coordinate arrays at every point of the padded range, in the
// reference's loop order so that the negative-Jacobian fallback (copy the
// neighbour toward the interior) sees the same values.
int compute_jac_metric(Block &B, std::string &err) {
  const int mx = B.mx, my = B.my, mz = B.mz, ngg = B.ngg;
  auto point = [&](int i, int j, int k, double &jac, double m[9]) {
    const double dxd1 = (B.X(i + 1, j, k) - B.X(i - 1, j, k)) * 0.5, dyd1 = (B.Y(i + 1, j, k) - B.Y(i - 1, j, k)) * 0.5,
                 dzd1 = (B.Z(i + 1, j, k) - B.Z(i - 1, j, k)) * 0.5;
    const double dxd2 = (B.X(i, j + 1, k) - B.X(i, j - 1, k)) * 0.5, dyd2 = (B.Y(i, j + 1, k) - B.Y(i, j - 1, k)) * 0.5,
                 dzd2 = (B.Z(i, j + 1, k) - B.Z(i, j - 1, k)) * 0.5;
    const double dxd3 = (B.X(i, j, k + 1) - B.X(i, j, k - 1)) * 0.5, dyd3 = (B.Y(i, j, k + 1) - B.Y(i, j, k - 1)) * 0.5,
                 dzd3 = (B.Z(i, j, k + 1) - B.Z(i, j, k - 1)) * 0.5;
    jac = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 - dxd2 * dyd1 * dzd3 -
          dxd3 * dyd2 * dzd1;
    if (jac <= 0) return;
    m[0] = (dyd2 * dzd3 - dyd3 * dzd2) / jac;
    m[1] = (dxd3 * dzd2 - dxd2 * dzd3) / jac;
    m[2] = (dxd2 * dyd3 - dxd3 * dyd2) / jac;
    m[3] = (dyd3 * dzd1 - dyd1 * dzd3) / jac;
    m[4] = (dxd1 * dzd3 - dxd3 * dzd1) / jac;
    m[5] = (dxd3 * dyd1 - dxd1 * dyd3) / jac;
    m[6] = (dyd1 * dzd2 - dyd2 * dzd1) / jac;
    m[7] = (dxd2 * dzd1 - dxd1 * dzd2) / jac;
    m[8] = (dxd1 * dyd2 - dxd2 * dyd1) / jac;
  };
  const long np = B.np;
  auto store = [&](int i, int j, int k, int fi, int fj, int fk) -> bool {
    double jac, m[9];
    point(i, j, k, jac, m);
    const long p = B.id(i, j, k);
    if (jac <= 0) {
      if (fi == 0 && fj == 0 && fk == 0) return false;
      const long q = B.id(i + fi, j + fj, k + fk);
      B.jac[p] = B.jac[q];
      for (int c = 0; c < 9; ++c) B.metric[c * np + p] = B.metric[c * np + q];
      return true;
    }
    B.jac[p] = jac;
    for (int c = 0; c < 9; ++c) B.metric[c * np + p] = m[c];
    return true;
  };
  int kmin = 1, kmax = mz - 1;
  if (mz == 1) { kmin = 0; kmax = 1; }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = kmin; k < kmax; ++k)
        if (!store(i, j, k, 0, 0, 0)) { err = "negative Jacobian in the interior"; return CFX_E_ARG; }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = 0; k >= -ngg; --k) store(i, j, k, 0, 0, 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + ngg; ++k) store(i, j, k, 0, 0, -1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -ngg; --j)
      for (int k = -ngg; k < mz + ngg; ++k) store(i, j, k, 0, 1, 0);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) store(i, j, k, 0, -1, 0);
  for (int i = 0; i >= -ngg; --i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) store(i, j, k, 1, 0, 0);
  for (int i = mx - 1; i < mx + ngg; ++i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) store(i, j, k, -1, 0, 0);
  return 0;
}

// Mesh.cpp:718-786 (per-direction ghost grid), 874-922 (inner ghost grid),
// 793-867 (edges/corners by the parallelogram rule).
void init_ghost_grid(orc_ctx &C) {
  const int ngg = C.ngg;
  for (auto &B : C.blocks) {
    const int mx = B.mx, my = B.my, mz = B.mz;
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        for (int i = -1; i >= -ngg - 1; --i) {
          B.X(i, j, k) = 2 * B.X(0, j, k) - B.X(-i, j, k);
          B.Y(i, j, k) = 2 * B.Y(0, j, k) - B.Y(-i, j, k);
          B.Z(i, j, k) = 2 * B.Z(0, j, k) - B.Z(-i, j, k);
        }
        for (int i = mx; i < mx + ngg + 1; ++i) {
          B.X(i, j, k) = 2 * B.X(mx - 1, j, k) - B.X(2 * (mx - 1) - i, j, k);
          B.Y(i, j, k) = 2 * B.Y(mx - 1, j, k) - B.Y(2 * (mx - 1) - i, j, k);
          B.Z(i, j, k) = 2 * B.Z(mx - 1, j, k) - B.Z(2 * (mx - 1) - i, j, k);
        }
      }
    for (int k = 0; k < mz; ++k)
      for (int i = 0; i < mx; ++i) {
        for (int j = -1; j >= -ngg - 1; --j) {
          B.X(i, j, k) = 2 * B.X(i, 0, k) - B.X(i, -j, k);
          B.Y(i, j, k) = 2 * B.Y(i, 0, k) - B.Y(i, -j, k);
          B.Z(i, j, k) = 2 * B.Z(i, 0, k) - B.Z(i, -j, k);
        }
        for (int j = my; j < my + ngg + 1; ++j) {
          B.X(i, j, k) = 2 * B.X(i, my - 1, k) - B.X(i, 2 * (my - 1) - j, k);
          B.Y(i, j, k) = 2 * B.Y(i, my - 1, k) - B.Y(i, 2 * (my - 1) - j, k);
          B.Z(i, j, k) = 2 * B.Z(i, my - 1, k) - B.Z(i, 2 * (my - 1) - j, k);
        }
      }
    if (C.dim == 3) {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) {
            B.X(i, j, k) = 2 * B.X(i, j, 0) - B.X(i, j, -k);
            B.Y(i, j, k) = 2 * B.Y(i, j, 0) - B.Y(i, j, -k);
            B.Z(i, j, k) = 2 * B.Z(i, j, 0) - B.Z(i, j, -k);
          }
          for (int k = mz; k < mz + ngg + 1; ++k) {
            B.X(i, j, k) = 2 * B.X(i, j, mz - 1) - B.X(i, j, 2 * (mz - 1) - k);
            B.Y(i, j, k) = 2 * B.Y(i, j, mz - 1) - B.Y(i, j, 2 * (mz - 1) - k);
            B.Z(i, j, k) = 2 * B.Z(i, j, mz - 1) - B.Z(i, j, 2 * (mz - 1) - k);
          }
        }
    } else {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) {
            B.X(i, j, k) = B.X(i, j, 0);
            B.Y(i, j, k) = B.Y(i, j, 0);
            B.Z(i, j, k) = B.Z(i, j, 0) + k;
          }
          for (int k = mz; k < mz + ngg + 1; ++k) {
            B.X(i, j, k) = B.X(i, j, mz - 1);
            B.Y(i, j, k) = B.Y(i, j, mz - 1);
            B.Z(i, j, k) = B.Z(i, j, mz - 1) + k - mz + 1;
          }
        }
    }
  }
  // Mesh.cpp:874-922 init_inner_ghost_grid
  for (auto &B : C.blocks) {
    for (const auto &fc : B.inner) {
      Block &nB = C.blocks[fc.target_block];
      int di[3]{0, 0, 0};
      int min_[3]{fc.rs[0], fc.rs[1], fc.rs[2]}, max_[3]{fc.re[0], fc.re[1], fc.re[2]},
          mint_[3]{fc.ts[0], fc.ts[1], fc.ts[2]};
      if (fc.direction < 0) {
        min_[fc.face] -= 1;
        max_[fc.face] -= 1;
      } else {
        max_[fc.face] += 1;
        min_[fc.face] += 1;
      }
      if (fc.target_direction > 0) mint_[fc.target_face] -= 1;
      else mint_[fc.target_face] += 1;
      const int stx = fc.target_loop_dir[0], sty = fc.target_loop_dir[1], stz = fc.target_loop_dir[2];
      const int di1 = fc.loop_dir[0], dj1 = fc.loop_dir[1], dk1 = fc.loop_dir[2];
      for (int i1 = min_[0]; di1 * (i1 - max_[0]) != 1; i1 += di1) {
        di[0] = i1 - min_[0];
        for (int j1 = min_[1]; dj1 * (j1 - max_[1]) != 1; j1 += dj1) {
          di[1] = j1 - min_[1];
          for (int k1 = min_[2]; dk1 * (k1 - max_[2]) != 1; k1 += dk1) {
            di[2] = k1 - min_[2];
            const int i2 = mint_[0] + stx * di[fc.src_tar[0]];
            const int j2 = mint_[1] + sty * di[fc.src_tar[1]];
            const int k2 = mint_[2] + stz * di[fc.src_tar[2]];
            B.X(i1, j1, k1) = nB.X(i2, j2, k2);
            B.Y(i1, j1, k1) = nB.Y(i2, j2, k2);
            B.Z(i1, j1, k1) = nB.Z(i2, j2, k2);
          }
        }
      }
    }
  }
  // Mesh.cpp:793-867 corners
  for (auto &B : C.blocks) {
    const int mx = B.mx, my = B.my, mz = B.mz;
    std::vector<double> *arr[3] = {&B.x, &B.y, &B.z};
    auto A = [&](int c, int i, int j, int k) -> double & { return (*arr[c])[B.gid(i, j, k)]; };
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 0; k < mz; ++k) {
          const int i1{-i}, j1{-j}, i2{mx - 1 + i}, j2{my - 1 + j};
          for (int c = 0; c < 3; ++c) {
            A(c, i1, j1, k) = A(c, i1, 0, k) + A(c, 0, j1, k) - A(c, 0, 0, k);
            A(c, i2, j2, k) = A(c, i2, my - 1, k) + A(c, mx - 1, j2, k) - A(c, mx - 1, my - 1, k);
            A(c, i1, j2, k) = A(c, i1, my - 1, k) + A(c, 0, j2, k) - A(c, 0, my - 1, k);
            A(c, i2, j1, k) = A(c, i2, 0, k) + A(c, mx - 1, j1, k) - A(c, mx - 1, 0, k);
          }
        }
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 0; j < my; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int i1{-i}, k1{-k}, i2{mx - 1 + i}, k2{mz - 1 + k};
          for (int c = 0; c < 3; ++c) {
            A(c, i1, j, k1) = A(c, i1, j, 0) + A(c, 0, j, k1) - A(c, 0, j, 0);
            A(c, i2, j, k2) = A(c, i2, j, mz - 1) + A(c, mx - 1, j, k2) - A(c, mx - 1, j, mz - 1);
            A(c, i1, j, k2) = A(c, i1, j, mz - 1) + A(c, 0, j, k2) - A(c, 0, j, mz - 1);
            A(c, i2, j, k1) = A(c, i2, j, 0) + A(c, mx - 1, j, k1) - A(c, mx - 1, j, 0);
          }
        }
    for (int i = -ngg - 1; i <= mx + ngg; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int j1{-j}, k1{-k}, j2{my - 1 + j}, k2{mz - 1 + k};
          for (int c = 0; c < 3; ++c) {
            A(c, i, j1, k1) = A(c, i, j1, 0) + A(c, i, 0, k1) - A(c, i, 0, 0);
            A(c, i, j2, k2) = A(c, i, j2, mz - 1) + A(c, i, my - 1, k2) - A(c, i, my - 1, mz - 1);
            A(c, i, j1, k2) = A(c, i, j1, mz - 1) + A(c, i, 0, k2) - A(c, i, 0, mz - 1);
            A(c, i, j2, k1) = A(c, i, j2, 0) + A(c, i, my - 1, k1) - A(c, i, my - 1, 0);
          }
        }
  }
}

// Thermo.cu:66-86 compute_total_energy at padded index p
void total_energy(const orc_ctx &C, Block &B, long p) {
  const long np = B.np;
  const double u = B.bv[1 * np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
  const double vel = u * u + v * v + w * w;
  double e = 0.5 * B.bv[p] * vel;
  if (C.sp.ns > 0) {
    double h[kMaxSp];
    enthalpy(C.sp, B.bv[5 * np + p], h);
    for (int l = 0; l < C.sp.ns; ++l) e += h[l] * B.cv[(5 + l) * np + p];
    e -= B.bv[4 * np + p];
  } else {
    e += B.bv[4 * np + p] / (C.P.gamma_air - 1);
  }
  B.cv[4 * np + p] = e;
}

// Field.cu:150-176 compute_cv_from_bv over the padded range
void cv_from_bv(const orc_ctx &C, Block &B) {
  const long np = B.np;
  const int ns = C.sp.ns;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < np; ++p) {
    const double rho = B.bv[p];
    B.cv[p] = rho;
    B.cv[np + p] = rho * B.bv[np + p];
    B.cv[2 * np + p] = rho * B.bv[2 * np + p];
    B.cv[3 * np + p] = rho * B.bv[3 * np + p];
    for (int l = 0; l < ns; ++l) B.cv[(5 + l) * np + p] = rho * B.yk[l * np + p];
    total_energy(C, B, p);
  }
}

// Field.cu:178-210 update_physical_properties over the padded range
void properties(const orc_ctx &C, Block &B) {
  const long np = B.np;
  const int ns = C.sp.ns;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < np; ++p) {
    const double t = B.bv[5 * np + p];
    if (ns > 0) {
      double cp[kMaxSp], Y[kMaxSp], rd[kMaxSp];
      cp_only(C.sp, t, cp);
      double mw{0}, cp_tot{0}, cvm{0};
      for (int l = 0; l < ns; ++l) {
        Y[l] = B.yk[l * np + p];
        mw += Y[l] / C.sp.mw[l];
        cp_tot += Y[l] * cp[l];
        cvm += Y[l] * (cp[l] - kRu / C.sp.mw[l]);
      }
      mw = 1 / mw;
      B.gamma[p] = cp_tot / cvm;
      B.c[p] = std::sqrt(B.gamma[p] * kRu * t / mw);
      double mu, kap;
      transport(C.sp, C.P.sc, t, mw, cp, Y, &mu, &kap, rd);
      B.mu[p] = mu;
      B.kappa[p] = kap;
      for (int l = 0; l < ns; ++l) B.rhoD[l * np + p] = rd[l];
    } else {
      const double c_temp{C.P.gamma_air * C.P.r_air};
      B.gamma[p] = C.P.gamma_air;
      B.c[p] = std::sqrt(c_temp * t);
      B.mu[p] = sutherland(t);
      B.kappa[p] = B.mu[p] * c_temp / (C.P.gamma_air - 1) / C.P.pr;
    }
  }
}

// InviscidScheme face at padded index p (left point) in direction d.
void inviscid_face_at(const orc_ctx &C, const Block &B, int d, long p, double *F) {
  const long np = B.np;
  const long s = d == 0 ? 1 : (d == 1 ? B.px : (long)B.px * B.py);
  const int ns = C.sp.ns;
  const int nrec = 5 + ns;
  double pv[4 * kMaxVar];
  for (int q = 0; q < 4; ++q) {
    const long pp = p + (q - 1) * s;
    for (int l = 0; l < 5; ++l) pv[q * nrec + l] = B.bv[l * np + pp];
    for (int l = 0; l < ns; ++l) pv[q * nrec + 5 + l] = B.yk[l * np + pp];
  }
  double metr[6], jac[2];
  for (int c = 0; c < 3; ++c) {
    metr[c] = B.metric[(d * 3 + c) * np + p];
    metr[3 + c] = B.metric[(d * 3 + c) * np + p + s];
  }
  jac[0] = B.jac[p];
  jac[1] = B.jac[p + s];
  ausm_face(C.sp, C.P, pv, nrec, metr, jac, F);
}

// SecOrdViscScheme::compute_fv/gv/hv (ViscousScheme.cu:11-342) generalised
// over the face direction d, with the reference's quirks selectable.
void viscous_face_at(const orc_ctx &C, const Block &B, int d, long p, double *G) {
  const long np = B.np;
  const long st[3] = {1, B.px, (long)B.px * B.py};
  const long pd = p + st[d];
  const int ns = C.sp.ns;
  const double *M = B.metric.data();
  // face metric (ViscousScheme.cu:16-24, 127-135, 238-246)
  double mf[9];
  const bool q5 = (C.P.quirks & CFX_Q_METRIC_AVG) && d != 1;
  for (int rc = 0; rc < 9; ++rc) {
    if (q5 && rc != 0) mf[rc] = 0.5 * (M[rc * np + p] + M[rc * np + p]);
    else mf[rc] = 0.5 * (M[rc * np + p] + M[rc * np + pd]);
  }
  auto grad = [&](const double *q, double g[3]) {
    for (int t = 0; t < 3; ++t) {
      if (t == d) {
        g[t] = q[pd] - q[p];
      } else {
        g[t] = 0.25 * (q[p + st[t]] - q[p - st[t]] + q[pd + st[t]] - q[pd - st[t]]);
      }
    }
  };
  auto chain = [&](const double g[3], double out[3]) {
    out[0] = g[0] * mf[0] + g[1] * mf[3] + g[2] * mf[6];
    out[1] = g[0] * mf[1] + g[1] * mf[4] + g[2] * mf[7];
    out[2] = g[0] * mf[2] + g[1] * mf[5] + g[2] * mf[8];
  };
  const double *U = &B.bv[np], *V = &B.bv[2 * np], *W = &B.bv[3 * np], *T = &B.bv[5 * np];
  double gu[3], gv[3], gw[3], gt[3], du[3], dv[3], dw[3], dtt[3];
  grad(U, gu); grad(V, gv); grad(W, gw); grad(T, gt);
  chain(gu, du); chain(gv, dv); chain(gw, dw); chain(gt, dtt);
  const double u_x = du[0], u_y = du[1], u_z = du[2];
  const double v_x = dv[0], v_y = dv[1], v_z = dv[2];
  const double w_x = dw[0], w_y = dw[1], w_z = dw[2];
  const double viscosity = 0.5 * (B.mu[p] + B.mu[pd]);
  const double tau_xx = viscosity * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  const double tau_yy = viscosity * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  const double tau_zz = viscosity * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double tau_xy = viscosity * (u_y + v_x);
  const double tau_xz = viscosity * (u_z + w_x);
  const double tau_yz = viscosity * (v_z + w_y);
  double S[3];
  for (int c = 0; c < 3; ++c) S[c] = 0.5 * (M[(d * 3 + c) * np + p] * B.jac[p] + M[(d * 3 + c) * np + pd] * B.jac[pd]);
  G[0] = 0;
  G[1] = S[0] * tau_xx + S[1] * tau_xy + S[2] * tau_xz;
  G[2] = S[0] * tau_xy + S[1] * tau_yy + S[2] * tau_yz;
  G[3] = S[0] * tau_xz + S[1] * tau_yz + S[2] * tau_zz;
  const double um = 0.5 * (U[p] + U[pd]);
  const double vm = 0.5 * (V[p] + V[pd]);
  const double wm = 0.5 * (W[p] + W[pd]);
  const double conductivity = 0.5 * (B.kappa[p] + B.kappa[pd]);
  G[4] = um * G[1] + vm * G[2] + wm * G[3] + conductivity * (S[0] * dtt[0] + S[1] * dtt[1] + S[2] * dtt[2]);
  if (ns > 0) {
    // ViscousScheme.cu:81-119
    const long pD = (d == 2 && (C.P.quirks & CFX_Q_HV_RHOD_J)) ? p + st[1] : pd;
    const double tm = 0.5 * (T[p] + T[pd]);
    double h[kMaxSp], diff[kMaxSp], yx[kMaxSp], yy[kMaxSp], yz[kMaxSp];
    enthalpy(C.sp, tm, h);
    double rho_uc{0}, rho_vc{0}, rho_wc{0};
    for (int l = 0; l < ns; ++l) {
      diff[l] = 0.5 * (B.rhoD[l * np + p] + B.rhoD[l * np + pD]);
      double gy[3], dy[3];
      grad(&B.yk[l * np], gy);
      chain(gy, dy);
      yx[l] = dy[0]; yy[l] = dy[1]; yz[l] = dy[2];
      const double DMulH = diff[l] * h[l];
      G[4] += S[0] * DMulH * yx[l];
      G[4] += S[1] * DMulH * yy[l];
      G[4] += S[2] * DMulH * yz[l];
      rho_uc += diff[l] * yx[l];
      rho_vc += diff[l] * yy[l];
      rho_wc += diff[l] * yz[l];
    }
    for (int l = 0; l < ns; ++l) {
      const double *Yl = &B.yk[l * np];
      G[5 + l] = diff[l] * (S[0] * yx[l] + S[1] * yy[l] + S[2] * yz[l]) -
                 0.5 * (Yl[p] + Yl[pd]) * (S[0] * rho_uc + S[1] * rho_vc + S[2] * rho_wc);
    }
  }
}

// compute_inviscid_flux + compute_viscous_flux (TimeAdvanceFunc.cu:38-243):
// dq(i) = -sum_d (F(i+1/2)-F(i-1/2)) + sum_d (G(i+1/2)-G(i-1/2)).
void rhs(orc_ctx &C, Block &B) {
  const int nv = C.nv;
  const int mx = B.mx, my = B.my, mz = B.mz;
  const long ni = B.ni;
  std::fill(B.dq.begin(), B.dq.end(), 0.0);  // set_dq_to_0 (TimeAdvanceFunc.cu:26-36)
  const int dim = C.dim;
  for (int pass = 0; pass < 2; ++pass) {  // 0 inviscid, 1 viscous
    if (pass == 1 && C.P.viscous != 2) break;
    for (int d = 0; d < dim; ++d) {
      const int fx = mx + (d == 0), fy = my + (d == 1), fz = mz + (d == 2);
      const long nf = (long)fx * fy * fz;
      std::vector<double> F((size_t)nf * nv);
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < fz; ++k)
        for (int j = 0; j < fy; ++j)
          for (int i = 0; i < fx; ++i) {
            // face index (i,j,k) is the face between point (i,j,k)-e_d and (i,j,k)
            const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
            double f[kMaxVar];
            if (pass == 0) inviscid_face_at(C, B, d, B.id(li, lj, lk), f);
            else viscous_face_at(C, B, d, B.id(li, lj, lk), f);
            const long fi = ((long)k * fy + j) * fx + i;
            for (int l = 0; l < nv; ++l) F[(size_t)fi * nv + l] = f[l];
          }
      const long sf = d == 0 ? 1 : (d == 1 ? fx : (long)fx * fy);
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < mx; ++i) {
            const long f_lo = ((long)k * fy + j) * fx + i;  // face i-1/2
            const long f_hi = f_lo + sf;                    // face i+1/2
            const long q = B.iid(i, j, k);
            for (int l = 0; l < nv; ++l) {
              const double diff = F[(size_t)f_hi * nv + l] - F[(size_t)f_lo * nv + l];
              if (pass == 0) B.dq[l * ni + q] -= diff;
              else B.dq[l * ni + q] += diff;
            }
          }
    }
  }
}

// SteadyTemporalScheme::compute_time_step (TemporalScheme.cu:10-43)
void local_dt(const orc_ctx &C, Block &B) {
  const long np = B.np;
  const int dim = C.dim;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < B.mz; ++k)
    for (int j = 0; j < B.my; ++j)
      for (int i = 0; i < B.mx; ++i) {
        const long p = B.id(i, j, k);
        const double *M = B.metric.data();
        double m[9];
        for (int c = 0; c < 9; ++c) m[c] = M[c * np + p];
        const double grad_xi = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const double grad_eta = std::sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
        const double grad_zeta = std::sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
        const double u{B.bv[np + p]}, v{B.bv[2 * np + p]}, w{B.bv[3 * np + p]};
        const double U = u * m[0] + v * m[1] + w * m[2];
        const double Vv = u * m[3] + v * m[4] + w * m[5];
        const double Wv = u * m[6] + v * m[7] + w * m[8];
        const double acoustic_speed = B.c[p];
        double sr_inv = std::abs(U) + std::abs(Vv) + acoustic_speed * (grad_xi + grad_eta);
        if (dim == 3) sr_inv += std::abs(Wv) + acoustic_speed * grad_zeta;
        const double coeff_1 = std::max(C.sp.ns > 0 ? B.gamma[p] : C.P.gamma_air, 4.0 / 3.0);
        const double coeff_2 = B.mu[p] / B.bv[p] / C.P.pr;
        double sr_visc = grad_xi * grad_xi + grad_eta * grad_eta;
        if (dim == 3) sr_visc += grad_zeta * grad_zeta;
        sr_visc *= coeff_1 * coeff_2;
        B.dt_local[B.iid(i, j, k)] = C.P.cfl / (sr_inv + sr_visc);
      }
}

// Matrix-free LU-SGS (Yoon & Jameson, AIAA J. 26 (1988) 1025-1026) for the
// implicit treatment the reference declares but leaves empty (LUSGS,
// TemporalScheme.cuh:19-21; "implicit treatment if needed", Driver.cu:156).
// Linearised implicit Euler with the local time step,
//   (J/dt - dR/dQ) dQ = R,   R = dq of the RHS,
// with the inviscid flux Jacobian through S_d = J grad xi_d split as
// A+- = (A +- omega r_d I)/2, r_d = |V.S_d| + c|S_d|, and the viscous part
// approximated by its spectral radius v_d = max(gamma, 4/3) mu/(rho Pr) J |grad xi_d|^2
// (the viscous terms of the local time step, TemporalScheme.cu:30-40):
//   D   = J/dt + sum_d (omega r_d + 2 v_d)            (at the cell)
//   fwd D dQ*_i = R_i + sum_d [ dG_d(i-e_d)/2 + a_d(i-e_d) dQ*_(i-e_d) ]
//   bwd dQ_i   = dQ*_i + sum_d [ a_d(i+e_d) dQ_(i+e_d) - dG_d(i+e_d)/2 ] / D_i
// with a_d = omega r_d / 2 + v_d, dG_d(n) = G_d(Q_n + dQ_n) - G_d(Q_n) the
// inviscid flux of cell n through S_d(n), the pressure change taken at the
// cell's frozen gamma (exact for the calorically perfect gas), and dQ = 0
// outside the block.  Every dQ*_i depends only on final lower-neighbour
// values, so any order that respects the i+j+k partial order (lexicographic
// here, hyperplanes on the GPU) gives the same result.
void lusgs_flux_change(const orc_ctx &C, const Block &B, int d, long p, const double *dQ, double *dG) {
  const long np = B.np;
  const int nv = C.nv;
  const double *M = B.metric.data();
  const double J = B.jac[p];
  const double Sx = J * M[(3 * d) * np + p], Sy = J * M[(3 * d + 1) * np + p], Sz = J * M[(3 * d + 2) * np + p];
  const double gam = C.sp.ns > 0 ? B.gamma[p] : C.P.gamma_air;
  double q0[kMaxVar], q1[kMaxVar];
  for (int l = 0; l < nv; ++l) {
    q0[l] = B.cv[l * np + p];
    q1[l] = q0[l] + dQ[l];
  }
  const double p0 = B.bv[4 * np + p];
  const double ke0 = 0.5 * (q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]) / q0[0];
  const double ke1 = 0.5 * (q1[1] * q1[1] + q1[2] * q1[2] + q1[3] * q1[3]) / q1[0];
  const double p1 = p0 + (gam - 1) * (dQ[4] - (ke1 - ke0));
  const double U0 = (Sx * q0[1] + Sy * q0[2] + Sz * q0[3]) / q0[0];
  const double U1 = (Sx * q1[1] + Sy * q1[2] + Sz * q1[3]) / q1[0];
  dG[0] = q1[0] * U1 - q0[0] * U0;
  dG[1] = (q1[1] * U1 + Sx * p1) - (q0[1] * U0 + Sx * p0);
  dG[2] = (q1[2] * U1 + Sy * p1) - (q0[2] * U0 + Sy * p0);
  dG[3] = (q1[3] * U1 + Sz * p1) - (q0[3] * U0 + Sz * p0);
  dG[4] = (q1[4] + p1) * U1 - (q0[4] + p0) * U0;
  for (int l = 5; l < nv; ++l) dG[l] = q1[l] * U1 - q0[l] * U0;
}

void lusgs(orc_ctx &C, Block &B) {
  const long np = B.np, ni = B.ni;
  const int nv = C.nv, dim = C.dim;
  const double om = C.P.omega;
  std::vector<double> Dg(ni), a(3 * ni, 0.0);
  const double *M = B.metric.data();
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < B.mz; ++k)
    for (int j = 0; j < B.my; ++j)
      for (int i = 0; i < B.mx; ++i) {
        const long p = B.id(i, j, k), q = B.iid(i, j, k);
        const double J = B.jac[p], rho = B.bv[p];
        const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
        const double gam = C.sp.ns > 0 ? B.gamma[p] : C.P.gamma_air;
        const double vc = std::max(gam, 4.0 / 3.0) * B.mu[p] / rho / C.P.pr;
        double diag = J / B.dt_local[q];
        for (int d = 0; d < dim; ++d) {
          const double gx = M[(3 * d) * np + p], gy = M[(3 * d + 1) * np + p], gz = M[(3 * d + 2) * np + p];
          const double g2 = gx * gx + gy * gy + gz * gz;
          const double r = J * (std::abs(u * gx + v * gy + w * gz) + B.c[p] * std::sqrt(g2));
          const double vd = vc * J * g2;
          diag += om * r + 2 * vd;
          a[d * ni + q] = 0.5 * om * r + vd;
        }
        Dg[q] = diag;
      }
  double *dQ = B.dq.data();
  const int m[3] = {B.mx, B.my, B.mz};
  // forward sweep: lower neighbours final
  for (int k = 0; k < B.mz; ++k)
    for (int j = 0; j < B.my; ++j)
      for (int i = 0; i < B.mx; ++i) {
        const long q = B.iid(i, j, k);
        double acc[kMaxVar];
        for (int l = 0; l < nv; ++l) acc[l] = dQ[l * ni + q];
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) {
          if (ijk[d] == 0) continue;
          int n3[3] = {i, j, k};
          n3[d] -= 1;
          const long qn = B.iid(n3[0], n3[1], n3[2]), pn = B.id(n3[0], n3[1], n3[2]);
          double dn[kMaxVar], dG[kMaxVar];
          for (int l = 0; l < nv; ++l) dn[l] = dQ[l * ni + qn];
          lusgs_flux_change(C, B, d, pn, dn, dG);
          const double an = a[d * ni + qn];
          for (int l = 0; l < nv; ++l) acc[l] += 0.5 * dG[l] + an * dn[l];
        }
        for (int l = 0; l < nv; ++l) dQ[l * ni + q] = acc[l] / Dg[q];
      }
  // backward sweep: upper neighbours final
  for (int k = B.mz - 1; k >= 0; --k)
    for (int j = B.my - 1; j >= 0; --j)
      for (int i = B.mx - 1; i >= 0; --i) {
        const long q = B.iid(i, j, k);
        double acc[kMaxVar];
        for (int l = 0; l < nv; ++l) acc[l] = 0.0;
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < dim; ++d) {
          if (ijk[d] == m[d] - 1) continue;
          int n3[3] = {i, j, k};
          n3[d] += 1;
          const long qn = B.iid(n3[0], n3[1], n3[2]), pn = B.id(n3[0], n3[1], n3[2]);
          double dn[kMaxVar], dG[kMaxVar];
          for (int l = 0; l < nv; ++l) dn[l] = dQ[l * ni + qn];
          lusgs_flux_change(C, B, d, pn, dn, dG);
          const double an = a[d * ni + qn];
          for (int l = 0; l < nv; ++l) acc[l] += an * dn[l] - 0.5 * dG[l];
        }
        for (int l = 0; l < nv; ++l) dQ[l * ni + q] += acc[l] / Dg[q];
      }
}

// update_cv_and_bv (TimeAdvanceFunc.cu:255-290) with the RK stage combination
// q = a*qn + b*(q + dt*dq/J) (a=0,b=1 for the reference's Euler step).
void update(orc_ctx &C, Block &B, int stage) {
  const long np = B.np, ni = B.ni;
  const int nv = C.nv, ns = C.sp.ns;
  const bool rk3 = C.P.rk == 3;
  static const double A[3] = {0.0, 0.75, 1.0 / 3.0};
  static const double Bc[3] = {1.0, 0.25, 2.0 / 3.0};
  long long fails = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : fails)
  for (int k = 0; k < B.mz; ++k)
    for (int j = 0; j < B.my; ++j)
      for (int i = 0; i < B.mx; ++i) {
        const long p = B.id(i, j, k);
        const long q = B.iid(i, j, k);
        double *cv = B.cv.data();
        if (rk3) {
          const double dt_div_jac = C.dt / B.jac[p];
          if (stage == 0)
            for (int l = 0; l < nv; ++l) B.qn[l * ni + q] = cv[l * np + p];
          for (int l = 0; l < nv; ++l)
            cv[l * np + p] = A[stage] * B.qn[l * ni + q] + Bc[stage] * (cv[l * np + p] + B.dq[l * ni + q] * dt_div_jac);
        } else if (C.P.lusgs) {
          for (int l = 0; l < nv; ++l) cv[l * np + p] += B.dq[l * ni + q];
        } else {
          const double dt_div_jac = B.dt_local[q] / B.jac[p];
          for (int l = 0; l < nv; ++l) cv[l * np + p] += B.dq[l * ni + q] * dt_div_jac;
        }
        if (B.mz == 1) cv[3 * np + p] = 0;
        double *bv = B.bv.data();
        bv[p] = cv[p];
        const double density_inv = 1.0 / cv[p];
        bv[np + p] = cv[np + p] * density_inv;
        bv[2 * np + p] = cv[2 * np + p] * density_inv;
        bv[3 * np + p] = cv[3 * np + p] * density_inv;
        const double velocity = bv[np + p] * bv[np + p] + bv[2 * np + p] * bv[2 * np + p] + bv[3 * np + p] * bv[3 * np + p];
        if (ns > 0) {
          double Y[kMaxSp];
          for (int l = 0; l < ns; ++l) {
            B.yk[l * np + p] = cv[(5 + l) * np + p] * density_inv;
            Y[l] = B.yk[l * np + p];
          }
          // Thermo.cu:89-125
          const double e = cv[4 * np + p] / cv[p] - 0.5 * velocity;
          double t, gas_const;
          const int it = newton_t(C.sp, C.P.quirks, Y, e, bv[5 * np + p], &t, &gas_const);
          if (it > 1000) ++fails;
          bv[5 * np + p] = t;
          bv[4 * np + p] = bv[p] * t * gas_const;
        } else {
          bv[4 * np + p] = (C.P.gamma_air - 1) * (cv[4 * np + p] - 0.5 * bv[p] * velocity);
          bv[5 * np + p] = bv[4 * np + p] * density_inv / C.P.r_air;
        }
      }
  C.newton_fail += fails;
}

// apply_outflow (BoundCond.cu:455-487)
void bc_outflow(const orc_ctx &C, Block &B, const Boundary &b) {
  const long np = B.np;
  int dir[3]{0, 0, 0};
  dir[b.face] = b.direction;
  for (int k = b.rs[2]; k <= b.re[2]; ++k)
    for (int j = b.rs[1]; j <= b.re[1]; ++j)
      for (int i = b.rs[0]; i <= b.re[0]; ++i) {
        const long p = B.id(i, j, k);
        for (int g = 1; g <= C.ngg; ++g) {
          const long q = B.id(i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          for (int l = 0; l < 6; ++l) B.bv[l * np + q] = B.bv[l * np + p];
          for (int l = 0; l < C.nv; ++l) B.cv[l * np + q] = B.cv[l * np + p];
          for (int l = 0; l < C.sp.ns; ++l) B.yk[l * np + q] = B.yk[l * np + p];
        }
      }
}

// apply_inflow (BoundCond.cu:489-527)
void bc_inflow(const orc_ctx &C, Block &B, const Boundary &b) {
  const long np = B.np;
  int dir[3]{0, 0, 0};
  dir[b.face] = b.direction;
  const double *st = &C.inflow_state[b.cond * 6];
  const double *iy = C.sp.ns > 0 ? C.inflow_y[b.cond].data() : nullptr;
  for (int k = b.rs[2]; k <= b.re[2]; ++k)
    for (int j = b.rs[1]; j <= b.re[1]; ++j)
      for (int i = b.rs[0]; i <= b.re[0]; ++i)
        for (int g = 1; g <= C.ngg; ++g) {
          const long q = B.id(i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          const double density = st[0];
          for (int l = 0; l < 6; ++l) B.bv[l * np + q] = st[l];
          for (int l = 0; l < C.sp.ns; ++l) {
            B.yk[l * np + q] = iy[l];
            B.cv[(5 + l) * np + q] = density * iy[l];
          }
          B.cv[q] = density;
          B.cv[np + q] = density * st[1];
          B.cv[2 * np + q] = density * st[2];
          B.cv[3 * np + q] = density * st[3];
          total_energy(C, B, q);
        }
}

double mixture_mw(const orc_ctx &C, const Block &B, long p) {
  if (C.sp.ns == 0) return kRu / C.P.r_air;
  double mw = 0;
  for (int l = 0; l < C.sp.ns; ++l) mw += B.yk[l * B.np + p] / C.sp.mw[l];
  return 1 / mw;
}

// apply_wall (BoundCond.cu:529-629); the range is the face's own vertex range
// (BoundCond.cu:229-239 shrinks the registered range back by ngg).
void bc_wall(const orc_ctx &C, Block &B, const Boundary &b) {
  const long np = B.np;
  const int ns = C.sp.ns;
  int dir[3]{0, 0, 0};
  dir[b.face] = b.direction;
  const cfx_wall &wl = C.wall[b.cond];
  int rs[3], re[3];
  for (int q = 0; q < 3; ++q) {
    rs[q] = b.rs[q]; re[q] = b.re[q];
    if (q == b.face) continue;
    rs[q] += C.ngg; re[q] -= C.ngg;
  }
  for (int k = rs[2]; k <= re[2]; ++k)
    for (int j = rs[1]; j <= re[1]; ++j)
      for (int i = rs[0]; i <= re[0]; ++i) {
        double t_wall{wl.temperature};
        const long p = B.id(i, j, k);
        const long pin = B.id(i - dir[0], j - dir[1], k - dir[2]);
        if (wl.adiabatic) t_wall = B.bv[5 * np + pin];
        const double pr{B.bv[4 * np + pin]};
        for (int l = 0; l < ns; ++l) B.yk[l * np + p] = B.yk[l * np + pin];
        double mw = mixture_mw(C, B, p);
        B.bv[p] = pr * mw / (t_wall * kRu);
        B.bv[np + p] = 0; B.bv[2 * np + p] = 0; B.bv[3 * np + p] = 0;
        B.bv[4 * np + p] = pr;
        B.bv[5 * np + p] = t_wall;
        B.cv[p] = B.bv[p];
        B.cv[np + p] = 0; B.cv[2 * np + p] = 0; B.cv[3 * np + p] = 0;
        for (int l = 0; l < ns; ++l) B.cv[(l + 5) * np + p] = B.yk[l * np + p] * B.bv[p];
        total_energy(C, B, p);
        for (int g = 1; g <= C.ngg; ++g) {
          const long qi = B.id(i - g * dir[0], j - g * dir[1], k - g * dir[2]);
          const long qg = B.id(i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          for (int l = 0; l < ns; ++l) B.yk[l * np + qg] = B.yk[l * np + qi];
          mw = mixture_mw(C, B, qg);
          const double u_i{B.bv[np + qi]}, v_i{B.bv[2 * np + qi]}, w_i{B.bv[3 * np + qi]};
          const double p_i{B.bv[4 * np + qi]}, t_i{B.bv[5 * np + qi]};
          double t_g{t_i};
          if (!wl.adiabatic) {
            t_g = 2 * t_wall - t_i;
            if (t_g <= 0.1 * t_wall) t_g = t_wall;
          }
          const double rho_g{p_i * mw / (t_g * kRu)};
          B.bv[qg] = rho_g;
          B.bv[np + qg] = -u_i; B.bv[2 * np + qg] = -v_i; B.bv[3 * np + qg] = -w_i;
          B.bv[4 * np + qg] = p_i;
          B.bv[5 * np + qg] = t_g;
          B.cv[qg] = rho_g;
          B.cv[np + qg] = -rho_g * u_i; B.cv[2 * np + qg] = -rho_g * v_i; B.cv[3 * np + qg] = -rho_g * w_i;
          for (int l = 0; l < ns; ++l) B.cv[(l + 5) * np + qg] = B.yk[l * np + qg] * rho_g;
          total_energy(C, B, qg);
        }
      }
}

// Slip wall (type 3, new): the boundary point keeps rho, p, T, Y and loses the
// velocity component along the face normal n = grad(xi_face)/|grad(xi_face)|;
// ghost g mirrors interior point g with the normal velocity negated.
void bc_symmetry(const orc_ctx &C, Block &B, const Boundary &b) {
  const long np = B.np;
  const int ns = C.sp.ns;
  int dir[3]{0, 0, 0};
  dir[b.face] = b.direction;
  for (int k = b.rs[2]; k <= b.re[2]; ++k)
    for (int j = b.rs[1]; j <= b.re[1]; ++j)
      for (int i = b.rs[0]; i <= b.re[0]; ++i) {
        const long p = B.id(i, j, k);
        double n[3];
        for (int c = 0; c < 3; ++c) n[c] = B.metric[(b.face * 3 + c) * np + p];
        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] /= nn; n[1] /= nn; n[2] /= nn;
        {
          const double vn = B.bv[np + p] * n[0] + B.bv[2 * np + p] * n[1] + B.bv[3 * np + p] * n[2];
          for (int c = 0; c < 3; ++c) B.bv[(1 + c) * np + p] -= vn * n[c];
          for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + p] = B.bv[p] * B.bv[(1 + c) * np + p];
          total_energy(C, B, p);
        }
        for (int g = 1; g <= C.ngg; ++g) {
          const long qi = B.id(i - g * dir[0], j - g * dir[1], k - g * dir[2]);
          const long qg = B.id(i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          const double vn = B.bv[np + qi] * n[0] + B.bv[2 * np + qi] * n[1] + B.bv[3 * np + qi] * n[2];
          B.bv[qg] = B.bv[qi];
          for (int c = 0; c < 3; ++c) B.bv[(1 + c) * np + qg] = B.bv[(1 + c) * np + qi] - 2 * vn * n[c];
          B.bv[4 * np + qg] = B.bv[4 * np + qi];
          B.bv[5 * np + qg] = B.bv[5 * np + qi];
          for (int l = 0; l < ns; ++l) {
            B.yk[l * np + qg] = B.yk[l * np + qi];
            B.cv[(5 + l) * np + qg] = B.bv[qg] * B.yk[l * np + qg];
          }
          B.cv[qg] = B.bv[qg];
          for (int c = 0; c < 3; ++c) B.cv[(1 + c) * np + qg] = B.bv[qg] * B.bv[(1 + c) * np + qg];
          total_energy(C, B, qg);
        }
      }
}

// DBoundCond::apply_boundary_conditions (BoundCond.cu:251-313): by priority,
// lowest first (outflow, inflow, symmetry, wall), then condition record, then
// block, then face order.
void apply_bc(orc_ctx &C) {
  const int order[4] = {CFX_BC_OUTFLOW, CFX_BC_INFLOW, CFX_BC_SYMMETRY, CFX_BC_WALL};
  for (int type : order) {
    int ncond = 1;
    if (type == CFX_BC_INFLOW) ncond = (int)C.inflow.size();
    if (type == CFX_BC_WALL) ncond = (int)C.wall.size();
    if (type == CFX_BC_OUTFLOW || type == CFX_BC_SYMMETRY) ncond = 1;
    for (int cond = 0; cond < ncond; ++cond)
      for (auto &B : C.blocks)
        for (const auto &b : B.bc) {
          if (b.type != type) continue;
          if ((type == CFX_BC_INFLOW || type == CFX_BC_WALL) && b.cond != cond) continue;
          switch (type) {
            case CFX_BC_OUTFLOW: bc_outflow(C, B, b); break;
            case CFX_BC_INFLOW: bc_inflow(C, B, b); break;
            case CFX_BC_SYMMETRY: bc_symmetry(C, B, b); break;
            case CFX_BC_WALL: bc_wall(C, B, b); break;
          }
        }
  }
}

// Driver::data_communication + inner_communication (Driver.cu:194-217,
// Field.cu:213-266) with the face range looped as in Mesh.cpp:905-917.
void exchange(orc_ctx &C) {
  const int ns = C.sp.ns, nv = C.nv;
  for (auto &B : C.blocks) {
    for (const auto &f : B.inner) {
      Block &T = C.blocks[f.target_block];
      const long np = B.np, tnp = T.np;
      int n_point[3];
      for (int q = 0; q < 3; ++q) n_point[q] = std::abs(f.rs[q] - f.re[q]) + 1;
      const int face_dir = f.direction > 0 ? f.rs[f.face] : f.re[f.face];
      for (int n2 = 0; n2 < n_point[2]; ++n2)
        for (int n1 = 0; n1 < n_point[1]; ++n1)
          for (int n0 = 0; n0 < n_point[0]; ++n0) {
            const int n[3] = {n0, n1, n2};
            int idx[3], d_idx[3], idx_tar[3];
            for (int q = 0; q < 3; ++q) {
              d_idx[q] = f.loop_dir[q] * n[q];
              idx[q] = f.rs[q] + d_idx[q];
            }
            for (int q = 0; q < 3; ++q) idx_tar[q] = f.ts[q] + f.target_loop_dir[q] * d_idx[f.src_tar[q]];
            const long p = B.id(idx[0], idx[1], idx[2]);
            const long t = T.id(idx_tar[0], idx_tar[1], idx_tar[2]);
            if (idx[f.face] == face_dir) {
              for (int l = 0; l < 6; ++l) {
                const double ave = 0.5 * (T.bv[l * tnp + t] + B.bv[l * np + p]);
                B.bv[l * np + p] = ave;
                T.bv[l * tnp + t] = ave;
              }
              for (int l = 0; l < nv; ++l) {
                const double ave = 0.5 * (T.cv[l * tnp + t] + B.cv[l * np + p]);
                B.cv[l * np + p] = ave;
                T.cv[l * tnp + t] = ave;
              }
              for (int l = 0; l < ns; ++l) {
                const double ave = 0.5 * (T.yk[l * tnp + t] + B.yk[l * np + p]);
                B.yk[l * np + p] = ave;
                T.yk[l * tnp + t] = ave;
              }
            } else {
              for (int l = 0; l < 5; ++l) {
                B.bv[l * np + p] = T.bv[l * tnp + t];
                B.cv[l * np + p] = T.cv[l * tnp + t];
              }
              B.bv[5 * np + p] = T.bv[5 * tnp + t];
              for (int l = 0; l < ns; ++l) {
                B.yk[l * np + p] = T.yk[l * tnp + t];
                B.cv[(l + 5) * np + p] = T.cv[(l + 5) * tnp + t];
              }
            }
          }
    }
  }
}

// eliminate_k_gradient (Field.cu:268-292)
void eliminate_k(orc_ctx &C, Block &B) {
  const long np = B.np;
  for (int j = -B.ngg; j < B.my + B.ngg; ++j)
    for (int i = -B.ngg; i < B.mx + B.ngg; ++i) {
      const long p0 = B.id(i, j, 0);
      for (int k = 1; k <= B.ngg; ++k) {
        const long pp = B.id(i, j, k), pm = B.id(i, j, -k);
        for (int l = 0; l < 6; ++l) { B.bv[l * np + pp] = B.bv[l * np + p0]; B.bv[l * np + pm] = B.bv[l * np + p0]; }
        for (int l = 0; l < C.sp.ns; ++l) { B.yk[l * np + pp] = B.yk[l * np + p0]; B.yk[l * np + pm] = B.yk[l * np + p0]; }
      }
    }
}

// Inflow ctor state (BoundCond.cu:13-105) for (mach|velocity, p, T|rho).
int inflow_state(orc_ctx &C, const cfx_inflow &in, double *st, std::vector<double> &Y) {
  const int ns = C.sp.ns;
  double mw = kRu / C.P.r_air;
  Y.assign(ns, 0.0);
  if (ns > 0) {
    if (!in.yk) return CFX_E_ARG;
    mw = 0;
    for (int i = 0; i < ns; ++i) { Y[i] = in.yk[i]; mw += Y[i] / C.sp.mw[i]; }
    mw = 1 / mw;
  }
  double temperature = in.temperature, mach = in.mach, velocity = in.velocity, density = in.density;
  const double pressure = in.pressure;
  if (temperature < 0) temperature = pressure * mw / (density * kRu);
  double gamma{C.P.gamma_air};
  if (ns > 0) {
    double cp{0}, cvv{0}, cpi[kMaxSp];
    cp_only(C.sp, temperature, cpi);
    for (int i = 0; i < ns; ++i) {
      cp += Y[i] * cpi[i];
      cvv += Y[i] * (cpi[i] - kRu / C.sp.mw[i]);
    }
    gamma = cp / cvv;
  }
  const double c{std::sqrt(gamma * kRu / mw * temperature)};
  if (mach < 0) mach = velocity / c;
  else velocity = mach * c;
  if (density < 0) density = pressure * mw / (kRu * temperature);
  st[0] = density;
  st[1] = in.u * velocity;
  st[2] = in.v * velocity;
  st[3] = in.w * velocity;
  st[4] = pressure;
  st[5] = temperature;
  return 0;
}

}  // namespace

extern "C" {

int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_create(const cfx_params *params, const cfx_species *species, const cfx_block *blocks, int n_block,
               orc_ctx **out) {
  if (!params || !blocks || n_block <= 0 || !out) return CFX_E_ARG;
  auto *C = new orc_ctx;
  C->sp.set(species);
  if (C->sp.ns > kMaxSp) { delete C; return CFX_E_UNSUPPORTED; }
  Params &P = C->P;
  P.reconstruction = params->reconstruction;
  P.limiter = params->limiter;
  P.viscous = params->viscous_order;
  P.rk = params->rk_stages == 3 ? 3 : 1;
  P.cfl = params->cfl;
  P.pr = params->prandtl;
  P.sc = params->schmidt;
  P.gamma_air = params->gamma_air > 0 ? params->gamma_air : 1.4;
  P.r_air = params->r_air > 0 ? params->r_air : kRu / kMwAir;
  P.quirks = params->quirks;
  P.lusgs = params->lusgs ? 1 : 0;
  P.omega = params->lusgs_omega > 0 ? params->lusgs_omega : 1.0;
  if (P.lusgs && P.rk != 1) { delete C; return CFX_E_ARG; }
  C->nv = 5 + C->sp.ns;
  C->dim = blocks[0].mz == 1 ? 2 : 3;  // Mesh.cpp:579-581
  C->ngg = 2;                          // Parameter.cpp:32-36
  for (int i = 0; i < params->n_wall; ++i) C->wall.push_back(params->wall[i]);
  for (int i = 0; i < params->n_inflow; ++i) C->inflow.push_back(params->inflow[i]);
  C->inflow_state.assign(6 * C->inflow.size(), 0.0);
  C->inflow_y.resize(C->inflow.size());
  for (size_t i = 0; i < C->inflow.size(); ++i)
    if (inflow_state(*C, C->inflow[i], &C->inflow_state[6 * i], C->inflow_y[i])) {
      delete C;
      return CFX_E_ARG;
    }
  const int ngg = C->ngg;
  C->blocks.resize(n_block);
  for (int b = 0; b < n_block; ++b) {
    const cfx_block &d = blocks[b];
    Block &B = C->blocks[b];
    B.mx = d.mx; B.my = d.my; B.mz = d.mz; B.ngg = ngg; B.dim = C->dim;
    B.px = d.mx + 2 * ngg; B.py = d.my + 2 * ngg; B.pz = d.mz + 2 * ngg;
    B.np = (long)B.px * B.py * B.pz;
    B.ni = (long)d.mx * d.my * d.mz;
    C->n_grid_total += B.ni;
    B.gx = d.mx + 2 * (ngg + 1); B.gy = d.my + 2 * (ngg + 1); B.gz = d.mz + 2 * (ngg + 1);
    const long ng = (long)B.gx * B.gy * B.gz;
    B.x.assign(ng, 0.0); B.y.assign(ng, 0.0); B.z.assign(ng, 0.0);
    for (int k = 0; k < d.mz; ++k)
      for (int j = 0; j < d.my; ++j)
        for (int i = 0; i < d.mx; ++i) {
          const long s = B.iid(i, j, k);
          B.X(i, j, k) = d.x[s]; B.Y(i, j, k) = d.y[s]; B.Z(i, j, k) = d.z[s];
        }
    B.jac.assign(B.np, 0.0); B.metric.assign(9 * B.np, 0.0);
    B.cv.assign(C->nv * B.np, 0.0); B.bv.assign(6 * B.np, 0.0); B.yk.assign(C->sp.ns * B.np, 0.0);
    B.gamma.assign(B.np, 0.0); B.c.assign(B.np, 0.0); B.mu.assign(B.np, 0.0); B.kappa.assign(B.np, 0.0);
    B.rhoD.assign(C->sp.ns * B.np, 0.0);
    B.dq.assign(C->nv * B.ni, 0.0); B.dt_local.assign(B.ni, 0.0); B.bv_last.assign(4 * B.ni, 0.0);
    B.qn.assign(C->nv * B.ni, 0.0);
    for (int f = 0; f < d.n_bc; ++f) {
      Boundary bb;
      for (int q = 0; q < 3; ++q) { bb.rs[q] = d.bc[f].range_start[q]; bb.re[q] = d.bc[f].range_end[q]; }
      bb.type = d.bc[f].type;
      bb.cond = d.bc[f].cond;
      register_boundary(bb, ngg, C->dim);
      B.bc.push_back(bb);
    }
    for (int f = 0; f < d.n_inner; ++f) {
      InnerFace fc;
      for (int q = 0; q < 3; ++q) {
        fc.rs[q] = d.inner[f].src[2 * q]; fc.re[q] = d.inner[f].src[2 * q + 1];
        fc.ts[q] = d.inner[f].tar[2 * q]; fc.te[q] = d.inner[f].tar[2 * q + 1];
      }
      fc.target_block = d.inner[f].target_block - 1;
      register_inner(fc, ngg, C->dim);
      B.inner.push_back(fc);
    }
  }
  init_ghost_grid(*C);
  for (auto &B : C->blocks) {
    const int rc = compute_jac_metric(B, C->err);
    if (rc) { delete C; return rc; }
  }
  *out = C;
  return 0;
}

void orc_destroy(orc_ctx *ctx) { delete ctx; }
const char *orc_last_error(const orc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int orc_upload_primitive(orc_ctx *C, int blk, const double *bv, const double *yk) {
  if (!C || blk < 0 || blk >= (int)C->blocks.size()) return CFX_E_ARG;
  Block &B = C->blocks[blk];
  std::memcpy(B.bv.data(), bv, sizeof(double) * B.bv.size());
  if (C->sp.ns > 0) std::memcpy(B.yk.data(), yk, sizeof(double) * B.yk.size());
  return 0;
}

int orc_upload_state(orc_ctx *C, int blk, const double *cv, const double *bv, const double *yk) {
  const int rc = orc_upload_primitive(C, blk, bv, yk);
  if (rc) return rc;
  Block &B = C->blocks[blk];
  std::memcpy(B.cv.data(), cv, sizeof(double) * B.cv.size());
  return 0;
}

int orc_download(orc_ctx *C, int blk, int what, double *dst) {
  if (!C || blk < 0 || blk >= (int)C->blocks.size() || !dst) return CFX_E_ARG;
  Block &B = C->blocks[blk];
  const long np = B.np;
  switch (what) {
    case CFX_CV: std::memcpy(dst, B.cv.data(), sizeof(double) * B.cv.size()); break;
    case CFX_BV: std::memcpy(dst, B.bv.data(), sizeof(double) * B.bv.size()); break;
    case CFX_YK: std::memcpy(dst, B.yk.data(), sizeof(double) * B.yk.size()); break;
    case CFX_PROPS:
      std::memcpy(dst, B.gamma.data(), sizeof(double) * np);
      std::memcpy(dst + np, B.c.data(), sizeof(double) * np);
      std::memcpy(dst + 2 * np, B.mu.data(), sizeof(double) * np);
      std::memcpy(dst + 3 * np, B.kappa.data(), sizeof(double) * np);
      if (C->sp.ns > 0) std::memcpy(dst + 4 * np, B.rhoD.data(), sizeof(double) * B.rhoD.size());
      break;
    case CFX_DQ: std::memcpy(dst, B.dq.data(), sizeof(double) * B.dq.size()); break;
    case CFX_METRIC:
      std::memcpy(dst, B.metric.data(), sizeof(double) * 9 * np);
      std::memcpy(dst + 9 * np, B.jac.data(), sizeof(double) * np);
      break;
    case CFX_DTLOCAL: std::memcpy(dst, B.dt_local.data(), sizeof(double) * B.ni); break;
    case CFX_COORD:
      for (int k = -B.ngg; k < B.mz + B.ngg; ++k)
        for (int j = -B.ngg; j < B.my + B.ngg; ++j)
          for (int i = -B.ngg; i < B.mx + B.ngg; ++i) {
            dst[B.id(i, j, k)] = B.X(i, j, k);
            dst[np + B.id(i, j, k)] = B.Y(i, j, k);
            dst[2 * np + B.id(i, j, k)] = B.Z(i, j, k);
          }
      break;
    default: return CFX_E_ARG;
  }
  return 0;
}

int orc_init_conservative(orc_ctx *C) {
  for (auto &B : C->blocks) cv_from_bv(*C, B);
  apply_bc(*C);
  exchange(*C);
  for (auto &B : C->blocks) properties(*C, B);
  return 0;
}

// store_last_step (TimeAdvanceFunc.cu:13-24)
int orc_store_last_step(orc_ctx *C) {
  for (auto &B : C->blocks) {
    const long np = B.np, ni = B.ni;
    for (int k = 0; k < B.mz; ++k)
      for (int j = 0; j < B.my; ++j)
        for (int i = 0; i < B.mx; ++i) {
          const long p = B.id(i, j, k), q = B.iid(i, j, k);
          const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
          B.bv_last[q] = B.bv[p];
          B.bv_last[ni + q] = std::sqrt(u * u + v * v + w * w);
          B.bv_last[2 * ni + q] = B.bv[4 * np + p];
          B.bv_last[3 * ni + q] = B.bv[5 * np + p];
        }
  }
  return 0;
}

int orc_rhs(orc_ctx *C, int blk) {
  if (blk < 0) {
    for (auto &B : C->blocks) rhs(*C, B);
    return 0;
  }
  if (blk >= (int)C->blocks.size()) return CFX_E_ARG;
  rhs(*C, C->blocks[blk]);
  return 0;
}

int orc_time_step(orc_ctx *C, int global, double *dt_out) {
  double dtmin = 1e300;
  for (auto &B : C->blocks) {
    local_dt(*C, B);
    if (global)
      for (long q = 0; q < B.ni; ++q) dtmin = std::min(dtmin, B.dt_local[q]);
  }
  if (global) C->dt = dtmin;
  if (dt_out) *dt_out = C->dt;
  return 0;
}

int orc_update(orc_ctx *C, int stage) {
  for (auto &B : C->blocks) update(*C, B, stage);
  return 0;
}

int orc_set_dq(orc_ctx *C, int blk, const double *dq) {
  if (!C || blk < 0 || blk >= (int)C->blocks.size() || !dq) return CFX_E_ARG;
  Block &B = C->blocks[blk];
  std::memcpy(B.dq.data(), dq, sizeof(double) * B.dq.size());
  return 0;
}

int orc_lusgs(orc_ctx *C) {
  if (!C || !C->P.lusgs) return CFX_E_ARG;
  for (auto &B : C->blocks) lusgs(*C, B);
  return 0;
}

int orc_apply_bc(orc_ctx *C) { apply_bc(*C); return 0; }
int orc_exchange(orc_ctx *C) { exchange(*C); return 0; }
int orc_eliminate_k_gradient(orc_ctx *C) {
  if (C->dim == 2)
    for (auto &B : C->blocks) eliminate_k(*C, B);
  return 0;
}
int orc_update_properties(orc_ctx *C) {
  for (auto &B : C->blocks) properties(*C, B);
  return 0;
}

// compute_square_of_dbv + reduction_of_dbv_squared + the root (Driver.cu:219-257)
int orc_residual(orc_ctx *C, double res[4]) {
  long double acc[4] = {0, 0, 0, 0};
  for (auto &B : C->blocks) {
    const long np = B.np, ni = B.ni;
    for (int k = 0; k < B.mz; ++k)
      for (int j = 0; j < B.my; ++j)
        for (int i = 0; i < B.mx; ++i) {
          const long p = B.id(i, j, k), q = B.iid(i, j, k);
          const double u = B.bv[np + p], v = B.bv[2 * np + p], w = B.bv[3 * np + p];
          const double cur[4] = {B.bv[p], std::sqrt(u * u + v * v + w * w), B.bv[4 * np + p], B.bv[5 * np + p]};
          for (int l = 0; l < 4; ++l) {
            const double dd = cur[l] - B.bv_last[l * ni + q];
            acc[l] += (long double)(dd * dd);
          }
        }
  }
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt((double)acc[l] / (double)C->n_grid_total);
  return 0;
}

// Driver::steady_simulation loop body (Driver.cu:135-190); with rk_stages == 3
// the body runs once per SSP-RK3 stage with the global dt of stage 0.
int orc_step(orc_ctx *C, int n_steps) {
  for (int s = 0; s < n_steps; ++s) {
    orc_store_last_step(C);
    const int stages = C->P.rk == 3 ? 3 : 1;
    for (int st = 0; st < stages; ++st) {
      orc_rhs(C, -1);
      if (C->P.rk == 3) {
        if (st == 0) orc_time_step(C, 1, nullptr);
      } else {
        orc_time_step(C, 0, nullptr);
      }
      if (C->P.lusgs) orc_lusgs(C);
      orc_update(C, st);
      apply_bc(*C);
      exchange(*C);
      orc_eliminate_k_gradient(C);
      orc_update_properties(C);
    }
  }
  return 0;
}

int orc_get_dt(orc_ctx *C, double *dt) { *dt = C->dt; return 0; }
int orc_set_dt(orc_ctx *C, double dt) { C->dt = dt; return 0; }
long long orc_newton_failures(orc_ctx *C) { return C->newton_fail; }

int orc_enthalpy_cp(const cfx_species *sp, double t, double *h, double *cp) {
  Species s;
  s.set(sp);
  enthalpy_and_cp(s, t, h, cp);
  return 0;
}

int orc_transport(const cfx_species *sp, double schmidt, double t, const double *Y, double *mu, double *kappa,
                  double *rhoD) {
  Species s;
  s.set(sp);
  double cp[kMaxSp];
  cp_only(s, t, cp);
  double mw = 0;
  for (int l = 0; l < s.ns; ++l) mw += Y[l] / s.mw[l];
  mw = 1 / mw;
  transport(s, schmidt, t, mw, cp, Y, mu, kappa, rhoD);
  return 0;
}

int orc_newton_temperature(const cfx_species *sp, int quirks, const double *Y, double e, double t_guess,
                           double *t_out) {
  Species s;
  s.set(sp);
  double gc;
  return newton_t(s, quirks, Y, e, t_guess, t_out, &gc);
}

double orc_limiter(int tag, double a, double b) { return limiter(tag, a, b); }
double orc_sutherland(double t) { return sutherland(t); }

int orc_inviscid_face(const cfx_species *sp, const cfx_params *prm, const double *pv, const double *metric,
                      const double *jac, double *flux) {
  Species s;
  s.set(sp);
  Params P;
  P.reconstruction = prm->reconstruction;
  P.limiter = prm->limiter;
  P.gamma_air = prm->gamma_air > 0 ? prm->gamma_air : 1.4;
  P.r_air = prm->r_air > 0 ? prm->r_air : kRu / kMwAir;
  ausm_face(s, P, pv, 5 + s.ns, metric, jac, flux);
  return 0;
}

int orc_face_flux(orc_ctx *C, int blk, int dir, int i, int j, int k, int pass, double *out) {
  if (!C || blk < 0 || blk >= (int)C->blocks.size() || dir < 0 || dir >= C->dim || !out) return CFX_E_ARG;
  const Block &B = C->blocks[blk];
  const int m[3] = {B.mx, B.my, B.mz}, p[3] = {i, j, k};
  for (int q = 0; q < 3; ++q) {
    const int lo = q == dir ? -1 : 0;
    if (q < C->dim && (p[q] < lo || p[q] >= m[q])) return CFX_E_ARG;
  }
  if (pass == 0) inviscid_face_at(*C, B, dir, B.id(i, j, k), out);
  else viscous_face_at(*C, B, dir, B.id(i, j, k), out);
  return 0;
}

}  // extern "C"
