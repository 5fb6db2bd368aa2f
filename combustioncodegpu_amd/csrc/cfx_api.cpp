// cfx_api.cpp — the C ABI of include/cfx.h: context, device memory, the
// reference's Driver step loop (Driver.cu:167-246) over the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cfx.h"
#include "cfx_types.h"
#include "kernels.h"
#include "mesh.h"

using namespace cfx;

namespace {

enum KClass { K_RHS = 0, K_UPDATE, K_BC, K_PROPS, K_DT, K_OTHER, K_NCLASS };

struct BcOp {
  int type, blk, face;
};

struct TimedLaunch {
  hipEvent_t a, b;
  int cls;
};

}  // namespace

struct cfx_ctx {
  SchemeDev S{};
  int ns = 0, nv = 5, dim = 3;
  int rk = 3;
  std::vector<HostBlock> hb;
  std::vector<BlockView> views;
  std::vector<void *> allocs;
  SpeciesDev sp_host{};
  SpeciesDev *sp_dev = nullptr;
  std::vector<InflowDev> inflow_host;
  InflowDev *inflow_dev = nullptr;
  std::vector<WallDev> walls;
  std::vector<BcOp> bc_order;
  hipStream_t stream = nullptr;
  double *d_dt = nullptr;  // device dt; its bits are reduced with atomicMin
  double *d_res = nullptr;
  int *d_flags = nullptr;
  unsigned long long *d_newton = nullptr;
  unsigned long long *d_minscratch = nullptr;  // min target of non-global dt passes
  double *d_partial = nullptr;
  std::vector<int> partial_off;
  int n_partial = 0;
  long long n_grid_total = 0;
  int rhs_tile = 0;
  bool mixed = false;
  bool profiling = false;
  std::vector<TimedLaunch> timed;
  std::vector<hipEvent_t> event_pool;
  double class_ms[K_NCLASS] = {0};
  long long class_n[K_NCLASS] = {0};
  std::string err;

  int fail(int rc, const std::string &msg) {
    err = msg;
    if (rc > 0) err += std::string(": ") + hipGetErrorString((hipError_t)rc);
    return rc;
  }
  template <class T>
  T *dalloc(size_t n, int &rc) {
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
      rc = (int)e;
      return nullptr;
    }
    allocs.push_back(p);
    return (T *)p;
  }
  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  // brackets a launch with events when profiling
  template <class F>
  int timed_launch(int cls, F &&f) {
    if (!profiling) return f();
    TimedLaunch t{get_event(), get_event(), cls};
    (void)hipEventRecord(t.a, stream);
    const int rc = f();
    (void)hipEventRecord(t.b, stream);
    timed.push_back(t);
    return rc;
  }
};

namespace {

constexpr double kRuH = 8314.462618;

// Host NASA-7 evaluation, identical to species_h_cp (device_physics.h).
void host_h_cp(const SpeciesT<double> &s, int l, double t, double &h, double &cp) {
  const int r = t >= s.t_mid[l] ? 1 : 0;
  const double *c = s.cpc[r][l], *e = s.hc[r][l];
  const double cpp = c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * c[4])));
  const double hp = t * (e[0] + t * (e[1] + t * (e[2] + t * (e[3] + t * e[4])))) + e[5];
  if (t < s.t_low[l]) {
    cp = s.cp_tlow[l];
    h = s.h_tlow[l] + s.cp_tlow[l] * (t - s.t_low[l]);
  } else {
    cp = cpp;
    h = hp;
  }
}

template <typename A, typename B>
void round_table(const A &a, B &b) {
  b.ns = a.ns;
  using T = std::remove_reference_t<decltype(b.mw[0])>;
  auto cp1 = [](const auto &src, auto &dst) {
    const double *x = reinterpret_cast<const double *>(&src);
    T *y = reinterpret_cast<T *>(&dst);
    for (size_t i = 0; i < sizeof(src) / sizeof(double); ++i) y[i] = (T)x[i];
  };
  cp1(a.mw, b.mw); cp1(a.inv_mw, b.inv_mw); cp1(a.r_sp, b.r_sp); cp1(a.t_low, b.t_low); cp1(a.t_mid, b.t_mid);
  cp1(a.cpc, b.cpc); cp1(a.hc, b.hc); cp1(a.h_tlow, b.h_tlow); cp1(a.cp_tlow, b.cp_tlow); cp1(a.lj_inv, b.lj_inv);
  cp1(a.log_lj_inv, b.log_lj_inv); cp1(a.vis_coeff, b.vis_coeff); cp1(a.cond_add, b.cond_add); cp1(a.wj4, b.wj4);
  cp1(a.sqw8, b.sqw8);
}

int build_species(const cfx_species *sp, SpeciesDev &tables, std::string &err) {
  std::memset(&tables, 0, sizeof(tables));
  SpeciesT<double> &s = tables.d;
  s.ns = sp ? sp->n_spec : 0;
  if (s.ns < 0 || s.ns > kMaxSp) {
    err = "n_spec out of range";
    return CFX_E_ARG;
  }
  for (int l = 0; l < s.ns; ++l) {
    const double W = sp->mw[l];
    if (!(W > 0) || !(sp->eps_k[l] > 0) || !(sp->sigma[l] > 0)) {
      err = "species " + std::to_string(l) + ": non-positive mw / eps_k / sigma";
      return CFX_E_ARG;
    }
    const double rs = kRuH / W;
    s.mw[l] = W;
    s.inv_mw[l] = 1.0 / W;
    s.r_sp[l] = rs;
    s.t_low[l] = sp->t_low[l];
    s.t_mid[l] = sp->t_mid[l];
    for (int r = 0; r < 2; ++r) {
      const double *a = (r == 0 ? sp->coeff_lo : sp->coeff_hi) + 7 * l;
      for (int k = 0; k < 5; ++k) {
        s.cpc[r][l][k] = a[k] * rs;
        s.hc[r][l][k] = a[k] / (k + 1) * rs;
      }
      s.hc[r][l][5] = a[5] * rs;
    }
    // enthalpy and cp at t_low with the low set (Thermo.cu:29-36)
    const double tt = sp->t_low[l];
    const double *c = s.cpc[0][l], *e = s.hc[0][l];
    s.cp_tlow[l] = c[0] + tt * (c[1] + tt * (c[2] + tt * (c[3] + tt * c[4])));
    s.h_tlow[l] = tt * (e[0] + tt * (e[1] + tt * (e[2] + tt * (e[3] + tt * e[4])))) + e[5];
    s.lj_inv[l] = 1.0 / sp->eps_k[l];
    s.log_lj_inv[l] = std::log(s.lj_inv[l]);
    s.vis_coeff[l] = 2.6693e-6 * std::sqrt(W) / (sp->sigma[l] * sp->sigma[l]);  // ChemData.cpp:249-250
    s.cond_add[l] = 1.25 * rs;
  }
  for (int m = 0; m < s.ns; ++m)
    for (int n = 0; n < s.ns; ++n) {
      s.wj4[m][n] = std::pow(s.mw[n] / s.mw[m], 0.25);
      s.sqw8[m][n] = 1.0 / std::sqrt(8 * (1 + s.mw[m] / s.mw[n]));
    }
  round_table(tables.d, tables.f);
  return 0;
}

// Inflow::Inflow state (BoundCond.cu:13-105): (mach|velocity, p, T|rho).
int build_inflow(const cfx_inflow &in, const SpeciesT<double> &s, const SchemeDev &S, InflowDev &out,
                 std::string &err) {
  std::memset(&out, 0, sizeof(out));
  const int ns = s.ns;
  double mw = kRuH / S.r_air;
  if (ns > 0) {
    if (!in.yk) {
      err = "inflow without mass fractions on a multispecies case";
      return CFX_E_ARG;
    }
    double w = 0;
    for (int l = 0; l < ns; ++l) {
      out.y[l] = in.yk[l];
      w += in.yk[l] / s.mw[l];
    }
    mw = 1 / w;
  }
  double T = in.temperature, mach = in.mach, vel = in.velocity, rho = in.density;
  const double p = in.pressure;
  if (T < 0) T = p * mw / (rho * kRuH);
  double gamma = S.gamma_air;
  if (ns > 0) {
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      double h, c;
      host_h_cp(s, l, T, h, c);
      cp += out.y[l] * c;
      cv += out.y[l] * (c - s.r_sp[l]);
    }
    gamma = cp / cv;
  }
  const double c = std::sqrt(gamma * kRuH / mw * T);
  if (mach < 0) mach = vel / c;
  else vel = mach * c;
  if (rho < 0) rho = p * mw / (kRuH * T);
  out.st[0] = rho;
  out.st[1] = in.u * vel;
  out.st[2] = in.v * vel;
  out.st[3] = in.w * vel;
  out.st[4] = p;
  out.st[5] = T;
  // total energy (Thermo.cu:85-105)
  double e = 0.5 * rho * (out.st[1] * out.st[1] + out.st[2] * out.st[2] + out.st[3] * out.st[3]);
  if (ns > 0) {
    for (int l = 0; l < ns; ++l) {
      double h, cp;
      host_h_cp(s, l, T, h, cp);
      e += h * (rho * out.y[l]);
    }
    e -= p;
  } else {
    e += p / (S.gamma_air - 1);
  }
  out.st[6] = e;
  return 0;
}

#define CFX_TRY(expr)                  \
  do {                                 \
    const int rc_ = (expr);            \
    if (rc_ != 0) return ctx->fail(rc_, #expr); \
  } while (0)

int do_rhs(cfx_ctx *ctx, int blk) {
  const int b0 = blk < 0 ? 0 : blk, b1 = blk < 0 ? (int)ctx->views.size() : blk + 1;
  for (int b = b0; b < b1; ++b)
    CFX_TRY(ctx->timed_launch(K_RHS, [&] {
      return launch_rhs(ctx->views[b], ctx->sp_dev, ctx->S, ctx->ns, ctx->rhs_tile, ctx->stream);
    }));
  return 0;
}

int do_dt(cfx_ctx *ctx, int global, int store_local) {
  auto *bits = reinterpret_cast<unsigned long long *>(ctx->d_dt);
  if (global) CFX_TRY(ctx->timed_launch(K_DT, [&] { return launch_dt_reset(bits, ctx->stream); }));
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_DT, [&] {
      return launch_dt(B, ctx->S, ctx->ns, store_local, global ? bits : ctx->d_minscratch, ctx->stream);
    }));
  return 0;
}

int do_update(cfx_ctx *ctx, int stage) {
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_UPDATE, [&] {
      return launch_update(B, ctx->sp_dev, ctx->S, ctx->ns, stage, ctx->d_dt, ctx->d_flags, ctx->d_newton,
                           ctx->stream);
    }));
  return 0;
}

int do_bc(cfx_ctx *ctx) {
  for (const BcOp &op : ctx->bc_order) {
    const HostBlock &H = ctx->hb[op.blk];
    const int cond = H.bc_cond[op.face];
    const InflowDev *in = op.type == CFX_BC_INFLOW ? ctx->inflow_dev + cond : nullptr;
    const WallDev wl = op.type == CFX_BC_WALL ? ctx->walls[cond] : WallDev{0, 300.0};
    CFX_TRY(ctx->timed_launch(K_BC, [&] {
      return launch_bc(op.type, ctx->views[op.blk], H.bc_range[op.face], ctx->sp_dev, ctx->S, ctx->ns, in, wl,
                       ctx->stream);
    }));
  }
  return 0;
}

int do_exchange(cfx_ctx *ctx) {
  for (size_t b = 0; b < ctx->hb.size(); ++b)
    for (const InnerDev &f : ctx->hb[b].inner)
      CFX_TRY(ctx->timed_launch(K_OTHER, [&] {
        return launch_exchange(ctx->views[b], ctx->views[f.target_block], f, ctx->ns, ctx->stream);
      }));
  return 0;
}

int do_elim(cfx_ctx *ctx) {
  if (ctx->dim != 2) return 0;
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_OTHER, [&] { return launch_eliminate_k(B, ctx->ns, ctx->stream); }));
  return 0;
}

int do_props(cfx_ctx *ctx) {
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_PROPS, [&] { return launch_props(B, ctx->sp_dev, ctx->S, ctx->ns, ctx->stream); }));
  return 0;
}

int do_store_last(cfx_ctx *ctx) {
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_OTHER, [&] { return launch_store_last(B, ctx->stream); }));
  return 0;
}

// Driver::steady_simulation loop body (Driver.cu:189-244): with rk_stages 1
// the reference's explicit Euler with local dt; with 3 the SSP-RK3 stages
// with the global CFL dt taken at stage 0.
int do_step(cfx_ctx *ctx) {
  int rc;
  if ((rc = do_store_last(ctx))) return rc;
  const int stages = ctx->rk == 3 ? 3 : 1;
  for (int s = 0; s < stages; ++s) {
    if ((rc = do_rhs(ctx, -1))) return rc;
    if (ctx->rk == 3) {
      if (s == 0 && (rc = do_dt(ctx, 1, 0))) return rc;
    } else {
      if ((rc = do_dt(ctx, 0, 1))) return rc;
    }
    if ((rc = do_update(ctx, s))) return rc;
    if ((rc = do_bc(ctx))) return rc;
    if ((rc = do_exchange(ctx))) return rc;
    if ((rc = do_elim(ctx))) return rc;
    if ((rc = do_props(ctx))) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

int cfx_version(void) { return CFX_ABI_VERSION; }

int cfx_create(const cfx_params *params, const cfx_species *species, const cfx_block *blocks, int n_block,
               cfx_ctx **out) {
  if (!params || !blocks || n_block <= 0 || !out) return CFX_E_ARG;
  *out = nullptr;
  auto *ctx = new cfx_ctx;
  auto bail = [&](int rc) {
    cfx_destroy(ctx);
    return rc;
  };
  int rc = build_species(species, ctx->sp_host, ctx->err);
  if (rc) return bail(rc);
  ctx->ns = ctx->sp_host.d.ns;
  ctx->nv = 5 + ctx->ns;
  if (!ns_supported(ctx->ns)) return bail(CFX_E_UNSUPPORTED);
  if (params->inviscid_scheme != 3 && params->inviscid_scheme != 0) return bail(CFX_E_UNSUPPORTED);
  if ((params->precision != CFX_FP64 && params->precision != CFX_MIXED) || params->transport != CFX_TRANSPORT_WILKE)
    return bail(CFX_E_UNSUPPORTED);
  ctx->mixed = params->precision == CFX_MIXED;
  rc = build_mesh(blocks, n_block, ctx->hb, ctx->dim, ctx->err);
  if (rc) return bail(rc);
  SchemeDev &S = ctx->S;
  S.reconstruction = params->reconstruction;
  S.limiter = params->limiter;
  S.viscous = params->viscous_order;
  S.quirks = params->quirks;
  S.rk = params->rk_stages == 3 ? 3 : 1;
  S.dim = ctx->dim;
  S.cfl = params->cfl;
  S.pr = params->prandtl;
  S.sc = params->schmidt;
  S.gamma_air = params->gamma_air > 0 ? params->gamma_air : 1.4;
  S.r_air = params->r_air > 0 ? params->r_air : kRuH / kMwAir;
  S.inv_sc = 1.0 / S.sc;
  S.inv_pr = 1.0 / S.pr;
  ctx->rk = S.rk;
  for (int i = 0; i < params->n_inflow; ++i) {
    InflowDev d;
    rc = build_inflow(params->inflow[i], ctx->sp_host.d, S, d, ctx->err);
    if (rc) return bail(rc);
    ctx->inflow_host.push_back(d);
  }
  for (int i = 0; i < params->n_wall; ++i)
    ctx->walls.push_back(WallDev{params->wall[i].adiabatic, params->wall[i].temperature});
  // boundary application order (BoundCond.cu:251-313): outflow, inflow,
  // symmetry, wall; within a type by condition record, block, face
  const int order[4] = {CFX_BC_OUTFLOW, CFX_BC_INFLOW, CFX_BC_SYMMETRY, CFX_BC_WALL};
  for (int type : order) {
    const bool by_cond = type == CFX_BC_INFLOW || type == CFX_BC_WALL;
    const int ncond = type == CFX_BC_INFLOW ? (int)ctx->inflow_host.size()
                                            : (type == CFX_BC_WALL ? (int)ctx->walls.size() : 1);
    for (int cond = 0; cond < ncond; ++cond)
      for (int b = 0; b < n_block; ++b)
        for (size_t f = 0; f < ctx->hb[b].bc_type.size(); ++f)
          if (ctx->hb[b].bc_type[f] == type && (!by_cond || ctx->hb[b].bc_cond[f] == cond))
            ctx->bc_order.push_back(BcOp{type, b, (int)f});
  }
  for (int b = 0; b < n_block; ++b)
    for (size_t f = 0; f < ctx->hb[b].bc_type.size(); ++f) {
      const int t = ctx->hb[b].bc_type[f], c = ctx->hb[b].bc_cond[f];
      if ((t == CFX_BC_INFLOW && (c < 0 || c >= (int)ctx->inflow_host.size())) ||
          (t == CFX_BC_WALL && (c < 0 || c >= (int)ctx->walls.size()))) {
        ctx->err = "boundary condition record out of range";
        return bail(CFX_E_ARG);
      }
    }
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(ctx->fail((int)e, "hipStreamCreate"));
  rc = 0;
  ctx->sp_dev = ctx->dalloc<SpeciesDev>(1, rc);
  ctx->inflow_dev = ctx->dalloc<InflowDev>(std::max<size_t>(1, ctx->inflow_host.size()), rc);
  ctx->d_dt = ctx->dalloc<double>(1, rc);
  ctx->d_res = ctx->dalloc<double>(4, rc);
  ctx->d_flags = ctx->dalloc<int>(1, rc);
  ctx->d_newton = ctx->dalloc<unsigned long long>(1, rc);
  ctx->d_minscratch = ctx->dalloc<unsigned long long>(1, rc);
  if (rc) return bail(ctx->fail(rc, "hipMalloc"));
  (void)hipMemcpy(ctx->sp_dev, &ctx->sp_host, sizeof(SpeciesDev), hipMemcpyHostToDevice);
  if (!ctx->inflow_host.empty())
    (void)hipMemcpy(ctx->inflow_dev, ctx->inflow_host.data(), sizeof(InflowDev) * ctx->inflow_host.size(),
              hipMemcpyHostToDevice);
  (void)hipMemset(ctx->d_flags, 0, sizeof(int));
  (void)hipMemset(ctx->d_newton, 0, sizeof(unsigned long long));
  const double dt0 = 0.0;
  (void)hipMemcpy(ctx->d_dt, &dt0, sizeof(double), hipMemcpyHostToDevice);
  for (int b = 0; b < n_block; ++b) {
    const HostBlock &H = ctx->hb[b];
    BlockView V{};
    V.mx = H.mx;
    V.my = H.my;
    V.mz = H.mz;
    V.px = H.px;
    V.py = H.py;
    V.pz = H.pz;
    V.np = H.np;
    V.ni = H.ni;
    V.sy = H.px;
    V.sz = (long long)H.px * H.py;
    const size_t np = (size_t)H.np, ni = (size_t)H.ni;
    V.cv = ctx->dalloc<double>(ctx->nv * np, rc);
    V.bv = ctx->dalloc<double>(6 * np, rc);
    V.yk = ctx->dalloc<double>(ctx->ns * np, rc);
    V.props = ctx->dalloc<double>((4 + ctx->ns) * np, rc);
    V.metric = ctx->dalloc<double>(9 * np, rc);
    V.jac = ctx->dalloc<double>(np, rc);
    V.dq = ctx->dalloc<double>(ctx->nv * ni, rc);
    V.qn = ctx->dalloc<double>(ctx->nv * ni, rc);
    V.dt_local = ctx->dalloc<double>(ni, rc);
    V.bv_last = ctx->dalloc<double>(4 * ni, rc);
    if (rc) return bail(ctx->fail(rc, "hipMalloc block arrays"));
    (void)hipMemcpy(V.metric, H.metric.data(), sizeof(double) * 9 * np, hipMemcpyHostToDevice);
    (void)hipMemcpy(V.jac, H.jac.data(), sizeof(double) * np, hipMemcpyHostToDevice);
    V.mixed = ctx->mixed ? 1 : 0;
    if (ctx->mixed) {
      V.bv32 = ctx->dalloc<float>(6 * np, rc);
      V.yk32 = ctx->dalloc<float>(ctx->ns * np, rc);
      V.props32 = ctx->dalloc<float>((4 + ctx->ns) * np, rc);
      V.metric32 = ctx->dalloc<float>(9 * np, rc);
      V.jac32 = ctx->dalloc<float>(np, rc);
      V.dq32 = ctx->dalloc<float>(ctx->nv * ni, rc);
      if (rc) return bail(ctx->fail(rc, "hipMalloc FP32 arrays"));
      std::vector<float> m32(9 * np), j32(np);
      for (size_t i = 0; i < 9 * np; ++i) m32[i] = (float)H.metric[i];
      for (size_t i = 0; i < np; ++i) j32[i] = (float)H.jac[i];
      (void)hipMemcpy(V.metric32, m32.data(), sizeof(float) * 9 * np, hipMemcpyHostToDevice);
      (void)hipMemcpy(V.jac32, j32.data(), sizeof(float) * np, hipMemcpyHostToDevice);
      (void)hipMemset(V.props32, 0, sizeof(float) * (4 + ctx->ns) * np);
      (void)hipMemset(V.dq32, 0, sizeof(float) * ctx->nv * ni);
    }
    (void)hipMemset(V.props, 0, sizeof(double) * (4 + ctx->ns) * np);
    (void)hipMemset(V.dq, 0, sizeof(double) * ctx->nv * ni);
    ctx->views.push_back(V);
    ctx->partial_off.push_back(ctx->n_partial);
    ctx->n_partial += residual_partials(V);
    ctx->n_grid_total += H.ni;
  }
  ctx->d_partial = ctx->dalloc<double>(4 * (size_t)ctx->n_partial, rc);
  if (rc) return bail(ctx->fail(rc, "hipMalloc partials"));
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return bail(ctx->fail((int)e, "device setup"));
  *out = ctx;
  return 0;
}

void cfx_destroy(cfx_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &t : ctx->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (void *p : ctx->allocs) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *cfx_last_error(const cfx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static int check_blk(cfx_ctx *ctx, int blk) {
  if (!ctx) return CFX_E_ARG;
  if (blk < 0 || blk >= (int)ctx->views.size()) return ctx->fail(CFX_E_ARG, "block index out of range");
  return 0;
}

int cfx_upload_primitive(cfx_ctx *ctx, int blk, const double *bv, const double *yk) {
  if (int rc = check_blk(ctx, blk)) return rc;
  if (!bv || (ctx->ns > 0 && !yk)) return ctx->fail(CFX_E_ARG, "null field");
  BlockView &V = ctx->views[blk];
  (void)hipStreamSynchronize(ctx->stream);
  CFX_TRY((int)hipMemcpy(V.bv, bv, sizeof(double) * 6 * V.np, hipMemcpyHostToDevice));
  if (ctx->ns > 0) CFX_TRY((int)hipMemcpy(V.yk, yk, sizeof(double) * ctx->ns * V.np, hipMemcpyHostToDevice));
  return 0;
}

int cfx_upload_state(cfx_ctx *ctx, int blk, const double *cv, const double *bv, const double *yk) {
  if (int rc = cfx_upload_primitive(ctx, blk, bv, yk)) return rc;
  if (!cv) return ctx->fail(CFX_E_ARG, "null cv");
  BlockView &V = ctx->views[blk];
  CFX_TRY((int)hipMemcpy(V.cv, cv, sizeof(double) * ctx->nv * V.np, hipMemcpyHostToDevice));
  return 0;
}

int cfx_download(cfx_ctx *ctx, int blk, int what, double *dst) {
  if (int rc = check_blk(ctx, blk)) return rc;
  if (!dst) return ctx->fail(CFX_E_ARG, "null destination");
  BlockView &V = ctx->views[blk];
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  const size_t np = (size_t)V.np, ni = (size_t)V.ni;
  const double *src = nullptr;
  size_t n = 0;
  if (V.mixed && (what == CFX_PROPS || what == CFX_DQ)) {  // FP32 arrays, returned as double
    const float *s32 = what == CFX_PROPS ? V.props32 : V.dq32;
    const size_t n32 = what == CFX_PROPS ? (4 + ctx->ns) * np : ctx->nv * ni;
    std::vector<float> tmp(n32);
    CFX_TRY((int)hipMemcpy(tmp.data(), s32, sizeof(float) * n32, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n32; ++i) dst[i] = tmp[i];
    return 0;
  }
  switch (what) {
    case CFX_CV: src = V.cv, n = ctx->nv * np; break;
    case CFX_BV: src = V.bv, n = 6 * np; break;
    case CFX_YK: src = V.yk, n = ctx->ns * np; break;
    case CFX_PROPS: src = V.props, n = (4 + ctx->ns) * np; break;
    case CFX_DQ: src = V.dq, n = ctx->nv * ni; break;
    case CFX_DTLOCAL: src = V.dt_local, n = ni; break;
    case CFX_METRIC:
      CFX_TRY((int)hipMemcpy(dst, V.metric, sizeof(double) * 9 * np, hipMemcpyDeviceToHost));
      CFX_TRY((int)hipMemcpy(dst + 9 * np, V.jac, sizeof(double) * np, hipMemcpyDeviceToHost));
      return 0;
    default: return ctx->fail(CFX_E_ARG, "unknown download selector");
  }
  if (n) CFX_TRY((int)hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int cfx_init_conservative(cfx_ctx *ctx) {
  if (!ctx) return CFX_E_ARG;
  for (auto &B : ctx->views) CFX_TRY(launch_cv_from_bv(B, ctx->sp_dev, ctx->S, ctx->ns, ctx->stream));
  if (int rc = do_bc(ctx)) return rc;
  if (int rc = do_exchange(ctx)) return rc;
  if (int rc = do_props(ctx)) return rc;
  return 0;
}

int cfx_store_last_step(cfx_ctx *ctx) { return ctx ? do_store_last(ctx) : CFX_E_ARG; }

int cfx_rhs(cfx_ctx *ctx, int blk) {
  if (!ctx) return CFX_E_ARG;
  if (blk >= (int)ctx->views.size()) return ctx->fail(CFX_E_ARG, "block index out of range");
  return do_rhs(ctx, blk);
}

int cfx_time_step(cfx_ctx *ctx, int global, double *dt_out) {
  if (!ctx) return CFX_E_ARG;
  if (int rc = do_dt(ctx, global ? 1 : 0, 1)) return rc;
  if (dt_out) return cfx_get_dt(ctx, dt_out);
  return 0;
}

int cfx_update(cfx_ctx *ctx, int stage) {
  if (!ctx) return CFX_E_ARG;
  if (stage < 0 || stage > 2 || (ctx->rk == 1 && stage != 0)) return ctx->fail(CFX_E_ARG, "bad stage");
  return do_update(ctx, stage);
}

int cfx_apply_bc(cfx_ctx *ctx) { return ctx ? do_bc(ctx) : CFX_E_ARG; }
int cfx_exchange(cfx_ctx *ctx) { return ctx ? do_exchange(ctx) : CFX_E_ARG; }
int cfx_eliminate_k_gradient(cfx_ctx *ctx) { return ctx ? do_elim(ctx) : CFX_E_ARG; }
int cfx_update_properties(cfx_ctx *ctx) { return ctx ? do_props(ctx) : CFX_E_ARG; }

int cfx_residual(cfx_ctx *ctx, double res[4]) {
  if (!ctx || !res) return CFX_E_ARG;
  for (size_t b = 0; b < ctx->views.size(); ++b)
    CFX_TRY(launch_residual_partial(ctx->views[b], ctx->d_partial + 4 * (size_t)ctx->partial_off[b], ctx->stream));
  CFX_TRY(launch_residual_final(ctx->d_partial, ctx->n_partial, ctx->d_res, ctx->stream));
  double h[4];
  CFX_TRY((int)hipMemcpyAsync(h, ctx->d_res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(h[l] / (double)ctx->n_grid_total);
  return 0;
}

int cfx_step(cfx_ctx *ctx, int n_steps) {
  if (!ctx || n_steps < 0) return CFX_E_ARG;
  for (int s = 0; s < n_steps; ++s)
    if (int rc = do_step(ctx)) return rc;
  return 0;
}

int cfx_get_dt(cfx_ctx *ctx, double *dt) {
  if (!ctx || !dt) return CFX_E_ARG;
  CFX_TRY((int)hipMemcpyAsync(dt, ctx->d_dt, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  return 0;
}

int cfx_set_dt(cfx_ctx *ctx, double dt) {
  if (!ctx) return CFX_E_ARG;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(ctx->d_dt, &dt, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int cfx_check(cfx_ctx *ctx, long long *n_newton_fail) {
  if (!ctx) return CFX_E_ARG;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  int flags = 0;
  unsigned long long nf = 0;
  CFX_TRY((int)hipMemcpy(&flags, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost));
  CFX_TRY((int)hipMemcpy(&nf, ctx->d_newton, sizeof(nf), hipMemcpyDeviceToHost));
  if (n_newton_fail) *n_newton_fail = (long long)nf;
  if (flags) return ctx->fail(CFX_E_NONFINITE, "non-finite or non-positive rho/T/p in the update");
  if (nf) return ctx->fail(CFX_E_NEWTON, std::to_string(nf) + " Newton temperature inversions did not converge");
  return 0;
}

int cfx_sync(cfx_ctx *ctx) {
  if (!ctx) return CFX_E_ARG;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  return 0;
}

int cfx_time_steps(cfx_ctx *ctx, int n_steps, double *ms_total) {
  if (!ctx || n_steps < 0) return CFX_E_ARG;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, ctx->stream);
  int rc = cfx_step(ctx, n_steps);
  (void)hipEventRecord(b, ctx->stream);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (ms_total) *ms_total = ms;
  if (rc) return rc;
  if (ctx->profiling) {
    for (auto &t : ctx->timed) {
      float m = 0;
      (void)hipEventElapsedTime(&m, t.a, t.b);
      ctx->class_ms[t.cls] += m;
      ctx->class_n[t.cls] += 1;
      ctx->event_pool.push_back(t.a);
      ctx->event_pool.push_back(t.b);
    }
    ctx->timed.clear();
  }
  return 0;
}

int cfx_set_rhs_variant(cfx_ctx *ctx, int variant) {
  if (!ctx) return CFX_E_ARG;
  if (variant < 0 || variant >= n_rhs_tiles(ctx->dim)) return ctx->fail(CFX_E_ARG, "rhs variant out of range");
  ctx->rhs_tile = variant;
  return 0;
}

int cfx_rhs_variants(cfx_ctx *ctx) { return ctx ? n_rhs_tiles(ctx->dim) : CFX_E_ARG; }

int cfx_kernel_times(cfx_ctx *ctx, int cls, double *ms, long long *launches) {
  if (!ctx || cls < 0 || cls >= K_NCLASS) return CFX_E_ARG;
  if (ms) *ms = ctx->class_ms[cls];
  if (launches) *launches = ctx->class_n[cls];
  return 0;
}

int cfx_set_profiling(cfx_ctx *ctx, int on) {
  if (!ctx) return CFX_E_ARG;
  ctx->profiling = on != 0;
  for (int c = 0; c < K_NCLASS; ++c) ctx->class_ms[c] = 0, ctx->class_n[c] = 0;
  return 0;
}

}  // extern "C"
