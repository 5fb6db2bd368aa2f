// cfx_api.cpp — the C ABI of include/cfx.h: context, device memory, the
// reference's Driver step loop (Driver.cu:113-192) over the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cfx.h"
#include "context.h"
#include "kernels.h"
#include "mesh.h"

using namespace cfx;

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
  cp1(a.om145, b.om145); cp1(a.vis_coeff, b.vis_coeff); cp1(a.cond_add, b.cond_add); cp1(a.wj4, b.wj4);
  cp1(a.sqw8, b.sqw8);
}

// the per-species thermo records from the arrays (same values, regrouped)
template <typename T>
void fill_thermo(SpeciesT<T> &s) {
  for (int l = 0; l < s.ns; ++l) {
    T *v = s.th[l].v;
    for (int g = 0; g < 2; ++g) {
      for (int k = 0; k < 5; ++k) v[kThCp + 5 * g + k] = s.cpc[g][l][k];
      for (int k = 0; k < 6; ++k) v[kThH + 6 * g + k] = s.hc[g][l][k];
    }
    v[kThTmid] = s.t_mid[l];
    v[kThTlow] = s.t_low[l];
    v[kThHtlow] = s.h_tlow[l];
    v[kThCptlow] = s.cp_tlow[l];
  }
  T tmid_min = std::numeric_limits<T>::max(), tmid_max = std::numeric_limits<T>::lowest();
  T tlow_max = std::numeric_limits<T>::lowest();
  for (int l = 0; l < s.ns; ++l) {
    tmid_min = std::min(tmid_min, s.t_mid[l]);
    tmid_max = std::max(tmid_max, s.t_mid[l]);
    tlow_max = std::max(tlow_max, s.t_low[l]);
  }
  s.hi_from = std::max(tmid_max, tlow_max);
  s.lo_from = tlow_max;
  s.lo_to = tmid_min;
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
    // enthalpy and cp at t_low with the low set (Thermo.cu:30-37)
    const double tt = sp->t_low[l];
    const double *c = s.cpc[0][l], *e = s.hc[0][l];
    s.cp_tlow[l] = c[0] + tt * (c[1] + tt * (c[2] + tt * (c[3] + tt * c[4])));
    s.h_tlow[l] = tt * (e[0] + tt * (e[1] + tt * (e[2] + tt * (e[3] + tt * e[4])))) + e[5];
    s.lj_inv[l] = 1.0 / sp->eps_k[l];
    s.om145[l] = std::pow(s.lj_inv[l], -0.145);
    s.vis_coeff[l] = 2.6693e-6 * std::sqrt(W) / (sp->sigma[l] * sp->sigma[l]);  // ChemData.cpp:249-250
    s.cond_add[l] = 1.25 * rs;
  }
  for (int m = 0; m < s.ns; ++m)
    for (int n = 0; n < s.ns; ++n) {
      s.wj4[m][n] = std::pow(s.mw[n] / s.mw[m], 0.25);
      s.sqw8[m][n] = 1.0 / std::sqrt(8 * (1 + s.mw[m] / s.mw[n]));
    }
  round_table(tables.d, tables.f);
  fill_thermo(tables.d);
  fill_thermo(tables.f);
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
  // total energy (Thermo.cu:66-86)
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


// ---------------------------------------------------------------- tables ---
int upload_table(cfx_ctx *ctx, Table &t, const std::vector<int> &counts) {
  t.host.assign(counts.size() + 1, 0);
  for (size_t b = 0; b < counts.size(); ++b) t.host[b + 1] = t.host[b] + counts[b];
  int rc = 0;
  t.dev = ctx->dalloc<int>(t.host.size(), rc);
  if (rc) return rc;
  return (int)hipMemcpy(t.dev, t.host.data(), sizeof(int) * t.host.size(), hipMemcpyHostToDevice);
}

int chunks(long long n, int per) { return (int)((n + per - 1) / per); }

int build_tables(cfx_ctx *ctx) {
  const int nb = (int)ctx->hb.size();
  std::vector<int> pad(nb), in1(nb), dtc(nb), last(nb), res(nb), til(nb);
  ctx->n_tplanes = 0;
  for (int b = 0; b < nb; ++b) {
    const HostBlock &H = ctx->hb[b];
    const int nt[3] = {chunks(H.mx, kLuTi), chunks(H.my, kLuTj), chunks(H.mz, kLuTk)};
    til[b] = nt[1] * nt[2];
    ctx->n_tplanes = std::max(ctx->n_tplanes, nt[0] + nt[1] + nt[2] - 2);
    pad[b] = chunks(H.np, kTpb);
    in1[b] = chunks(H.ni, kTpb);
    dtc[b] = chunks(H.ni, kTpb * kDtIter);
    last[b] = chunks(H.ni, kTpb * kLastIter);
    res[b] = chunks(H.ni, kTpb * kResIter);
  }
  int rc;
  if ((rc = upload_table(ctx, ctx->t_pad, pad)) || (rc = upload_table(ctx, ctx->t_int, in1)) ||
      (rc = upload_table(ctx, ctx->t_dt, dtc)) || (rc = upload_table(ctx, ctx->t_last, last)) ||
      (rc = upload_table(ctx, ctx->t_res, res)) || (rc = upload_table(ctx, ctx->t_tile, til)))
    return rc;
  for (int d = 0; d < ctx->dim; ++d)
    for (int v = 0; v < rhs_variants(ctx->dim); ++v) {
      std::vector<int> tiles(nb);
      for (int b = 0; b < nb; ++b) {
        const int m[3] = {ctx->hb[b].mx, ctx->hb[b].my, ctx->hb[b].mz};
        tiles[b] = pencil_tiles(m, d, rhs_tile_shape(ctx->dim, d, v));
      }
      if ((rc = upload_table(ctx, ctx->t_rhs[d][v], tiles))) return rc;
    }
  return 0;
}

// ------------------------------------------------------- ghost-fill ops ---
// Boxes of points an op reads or writes, per block (unpadded indices, ghosts
// negative or beyond the extent).
struct Touch {
  int blk, lo[3], hi[3];
  bool write;
};

Touch box_of(int blk, const int a[3], const int b[3], bool write) {
  Touch t{blk, {0, 0, 0}, {0, 0, 0}, write};
  for (int q = 0; q < 3; ++q) {
    t.lo[q] = std::min(a[q], b[q]);
    t.hi[q] = std::max(a[q], b[q]);
  }
  return t;
}

std::vector<Touch> touches(const FaceOp &op) {
  std::vector<Touch> out;
  if (op.kind == kOpExchange) {
    const InnerDev &f = op.f;
    out.push_back(box_of(op.blk, f.rs, f.re, true));
    int lo[3], hi[3], flo[3], fhi[3];
    const int face_layer = f.direction > 0 ? f.rs[f.face] : f.re[f.face];
    const int nface = (face_layer - f.rs[f.face]) * f.loop_dir[f.face];
    for (int q = 0; q < 3; ++q) {
      const int s = f.src_tar[q];
      const int step = f.target_loop_dir[q] * f.loop_dir[s];
      lo[q] = f.ts[q];
      hi[q] = f.ts[q] + step * (f.n_point[s] - 1);
      flo[q] = lo[q];
      fhi[q] = hi[q];
      if (s == f.face) flo[q] = fhi[q] = f.ts[q] + step * nface;
    }
    out.push_back(box_of(op.tgt, lo, hi, false));
    out.push_back(box_of(op.tgt, flo, fhi, true));
    return out;
  }
  const FaceRange &r = op.r;
  const int a = r.face, s = r.direction;
  int lo[3], hi[3];
  for (int q = 0; q < 3; ++q) lo[q] = r.rs[q], hi[q] = r.re[q];
  auto layer = [&](int from, int to, bool write) {
    int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
    l2[a] = r.rs[a] + s * from;
    h2[a] = r.rs[a] + s * to;
    out.push_back(box_of(op.blk, l2, h2, write));
  };
  switch (op.kind) {
    case CFX_BC_OUTFLOW: layer(0, 0, false); layer(1, kNgg, true); break;
    case CFX_BC_INFLOW: layer(1, kNgg, true); break;
    default: layer(-kNgg, -1, false); layer(0, 0, true); layer(1, kNgg, true); break;  // wall, symmetry
  }
  return out;
}

bool conflict(const std::vector<Touch> &x, const std::vector<Touch> &y) {
  for (const Touch &a : x)
    for (const Touch &b : y) {
      if (a.blk != b.blk || !(a.write || b.write)) continue;
      bool meet = true;
      for (int q = 0; q < 3; ++q) meet = meet && a.lo[q] <= b.hi[q] && b.lo[q] <= a.hi[q];
      if (meet) return true;
    }
  return false;
}

// Levels for ops[first..last): an op goes one level after the latest earlier
// op it conflicts with, so every conflicting pair keeps the application
// order and the result equals the one-op-at-a-time loop.  levelize=false
// gives each op its own level (that loop itself).
int build_schedule(cfx_ctx *ctx, const std::vector<FaceOp> &all, Schedule &sc, int first, int last, bool levelize,
                   const std::vector<int> *group = nullptr) {
  const int n = last - first;
  std::vector<int> level(n, 0);
  if (!levelize && group) {  // one level per group of pieces of one op (disjoint boxes)
    for (int i = 0; i < n; ++i) level[i] = (*group)[first + i] - (*group)[first];
  } else if (levelize) {
    std::vector<std::vector<Touch>> tc(n);
    for (int i = 0; i < n; ++i) tc[i] = touches(all[first + i]);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < j; ++i)
        if (level[i] + 1 > level[j] && conflict(tc[i], tc[j])) level[j] = level[i] + 1;
  } else {
    for (int i = 0; i < n; ++i) level[i] = i;
  }
  const int nlev = n ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  std::vector<FaceOp> ordered;
  sc.lvl.assign(1, 0);
  for (int L = 0; L < nlev; ++L) {
    for (int i = 0; i < n; ++i)
      if (level[i] == L) ordered.push_back(all[first + i]);
    sc.lvl.push_back((int)ordered.size());
  }
  sc.off_host.assign(ordered.size() + 1, 0);
  for (size_t i = 0; i < ordered.size(); ++i) sc.off_host[i + 1] = sc.off_host[i] + chunks(ordered[i].npts, kTpb);
  int rc = 0;
  sc.ops = ctx->dalloc<FaceOp>(ordered.size(), rc);
  sc.off = ctx->dalloc<int>(sc.off_host.size(), rc);
  if (rc) return rc;
  if (!ordered.empty())
    if ((rc = (int)hipMemcpy(sc.ops, ordered.data(), sizeof(FaceOp) * ordered.size(), hipMemcpyHostToDevice)))
      return rc;
  return (int)hipMemcpy(sc.off, sc.off_host.data(), sizeof(int) * sc.off_host.size(), hipMemcpyHostToDevice);
}

// An outflow op whose range is the whole face (the face's points and the
// ngg layers around them tangentially) copies the face layer into the ghost
// layers behind it.  Its interior rows (tangential indices inside the block)
// read only points the update kernel has just written; when no op before it
// in the application order touches those rows (box conflict as for the
// levels), applying them inside the update kernel, right after the point's
// own stores, gives the same values in the same order.  The op is then
// replaced in the step's list by its ring of edge rows.  Sets the blocks'
// fuse flags (uploaded with the views).
void fold_outflow(cfx_ctx *ctx) {
  ctx->ops_step.clear();
  ctx->ops_step_src.clear();
  ctx->n_fused = 0;
  for (auto &V : ctx->views) {
    for (int f = 0; f < 6; ++f) V.fuse[f] = 0;
    V.fuse_any = 0;
  }
  const int n = (int)ctx->ops.size();
  std::vector<std::vector<Touch>> before;
  for (int i = 0; i < n; ++i) {
    const FaceOp &op = ctx->ops[i];
    bool fold = false;
    FaceOp inner = op;
    if (op.kind == CFX_BC_OUTFLOW) {
      const HostBlock &H = ctx->hb[op.blk];
      const int m[3] = {H.mx, H.my, H.mz};
      const int a = op.r.face;
      fold = op.r.rs[a] == op.r.re[a] && op.r.rs[a] == (op.r.direction < 0 ? 0 : m[a] - 1);
      for (int t = 0; t < 3 && fold; ++t) {
        if (t == a) continue;
        fold = op.r.rs[t] == -kNgg && op.r.re[t] == m[t] - 1 + kNgg;
        inner.r.rs[t] = 0;
        inner.r.re[t] = m[t] - 1;
      }
      if (fold) {
        const std::vector<Touch> ti = touches(inner);
        for (const auto &tb : before) fold = fold && !conflict(tb, ti);
      }
      if (fold) {
        // the ring: rows with a tangential index outside the block
        const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
        auto add = [&](int lo1, int hi1, int lo2, int hi2) {
          if (lo1 > hi1 || lo2 > hi2) return;
          FaceOp r = op;
          r.r.rs[t1] = lo1, r.r.re[t1] = hi1, r.r.rs[t2] = lo2, r.r.re[t2] = hi2;
          for (int q = 0; q < 3; ++q) r.ext[q] = r.r.re[q] - r.r.rs[q] + 1;
          r.npts = (long long)r.ext[0] * r.ext[1] * r.ext[2];
          ctx->ops_step.push_back(r);
          ctx->ops_step_src.push_back(i);
        };
        add(-kNgg, -1, -kNgg, m[t2] - 1 + kNgg);
        add(m[t1], m[t1] - 1 + kNgg, -kNgg, m[t2] - 1 + kNgg);
        add(0, m[t1] - 1, -kNgg, -1);
        add(0, m[t1] - 1, m[t2], m[t2] - 1 + kNgg);
        BlockView &V = ctx->views[op.blk];
        V.fuse[2 * a + (op.r.direction < 0 ? 0 : 1)] = 1;
        V.fuse_any = 1;
        ++ctx->n_fused;
      }
    }
    if (!fold) {
      ctx->ops_step.push_back(op);
      ctx->ops_step_src.push_back(i);
    }
    before.push_back(touches(op));
  }
}

// BC ops in the application order (BoundCond.cu:251-313), then the inner-face
// exchanges block by block, face by face (Driver.cu:194-217).
int build_ops(cfx_ctx *ctx) {
  ctx->ops.clear();
  for (const BcOp &b : ctx->bc_order) {
    const HostBlock &H = ctx->hb[b.blk];
    FaceOp op{};
    op.kind = b.type;
    op.blk = b.blk;
    op.tgt = b.blk;
    op.cond = H.bc_cond[b.face];
    op.r = H.bc_range[b.face];
    if (b.type == CFX_BC_WALL) {
      op.wl = ctx->walls[op.cond];
      for (int q = 0; q < 3; ++q) {
        if (q == op.r.face) continue;
        op.r.rs[q] += kNgg;
        op.r.re[q] -= kNgg;
      }
      if (op.r.rs[0] > op.r.re[0] || op.r.rs[1] > op.r.re[1] || op.r.rs[2] > op.r.re[2]) continue;
    }
    for (int q = 0; q < 3; ++q) op.ext[q] = op.r.re[q] - op.r.rs[q] + 1;
    op.npts = (long long)op.ext[0] * op.ext[1] * op.ext[2];
    ctx->ops.push_back(op);
  }
  ctx->n_bc_ops = (int)ctx->ops.size();
  for (size_t b = 0; b < ctx->hb.size(); ++b)
    for (const InnerDev &f : ctx->hb[b].inner) {
      FaceOp op{};
      op.kind = kOpExchange;
      op.blk = (int)b;
      op.tgt = f.target_block;
      op.f = f;
      for (int q = 0; q < 3; ++q) op.ext[q] = f.n_point[q];
      op.npts = (long long)op.ext[0] * op.ext[1] * op.ext[2];
      ctx->ops.push_back(op);
    }
  const int n = (int)ctx->ops.size(), nbc = ctx->n_bc_ops;
  int rc;
  if ((rc = build_schedule(ctx, ctx->ops, ctx->sched_seq, 0, n, false)) ||
      (rc = build_schedule(ctx, ctx->ops, ctx->sched_bc, 0, nbc, true)) ||
      (rc = build_schedule(ctx, ctx->ops, ctx->sched_ex, nbc, n, true)) ||
      (rc = build_schedule(ctx, ctx->ops, ctx->sched_all, 0, n, true)))
    return rc;
  // the step's ghost fill, with the outflow faces folded into the update
  fold_outflow(ctx);
  const int ns = (int)ctx->ops_step.size();
  if ((rc = build_schedule(ctx, ctx->ops_step, ctx->sched_seq_step, 0, ns, false, &ctx->ops_step_src)) ||
      (rc = build_schedule(ctx, ctx->ops_step, ctx->sched_all_step, 0, ns, true)))
    return rc;
  return (int)hipMemcpy(ctx->views_dev, ctx->views.data(), sizeof(BlockView) * ctx->views.size(),
                        hipMemcpyHostToDevice);
}

// levels [l0, l1) of a schedule, one launch each
int run_levels(cfx_ctx *ctx, const Schedule &sc, int l0, int l1) {
  const KernelSet64 &k = g_kernels64[ctx->ns];
  for (int L = l0; L < l1; ++L) {
    const int a = sc.lvl[L], b = sc.lvl[L + 1];
    const int total = sc.off_host[b] - sc.off_host[a];
    CFX_TRY(ctx->timed_launch(K_BC, [&] {
      return k.face_ops(ctx->views_dev, sc.ops + a, sc.off + a, b - a, total, ctx->sp_dev, ctx->S, ctx->inflow_dev,
                        ctx->stream);
    }));
  }
  return 0;
}

// ------------------------------------------------------------ step parts ---
int rhs_dir(cfx_ctx *ctx, int d, const Batch &b, int accumulate) {
  const int v = ctx->rhs_var[d];
  const RhsEntry &e = ctx->mixed ? (ctx->dim == 3 ? g_kernels32[ctx->ns].rhs3[d][v] : g_kernels32[ctx->ns].rhs2[d][v])
                                 : (ctx->dim == 3 ? g_kernels64[ctx->ns].rhs3[d][v] : g_kernels64[ctx->ns].rhs2[d][v]);
  const int padded = (b.total + 7) / 8 * 8;
  int grid = padded;
  const int occ = ctx->occ[ctx->mixed ? 1 : 0][d][v];
  if (ctx->rhs_persist[d] > 0 && occ > 0) grid = std::min(grid, (ctx->rhs_persist[d] * occ * ctx->n_cu + 7) / 8 * 8);
  return ctx->timed_launch(K_RHS, [&] { return e.launch(b, grid, ctx->sp_dev, ctx->S, accumulate, ctx->stream); });
}

int do_rhs(cfx_ctx *ctx, int blk) {
  if (blk < 0 && ctx->batched) {
    for (int d = 0; d < ctx->dim; ++d) CFX_TRY(rhs_dir(ctx, d, ctx->all(ctx->t_rhs[d][ctx->rhs_var[d]]), d > 0));
    return 0;
  }
  const int b0 = blk < 0 ? 0 : blk, b1 = blk < 0 ? (int)ctx->views.size() : blk + 1;
  for (int b = b0; b < b1; ++b)
    for (int d = 0; d < ctx->dim; ++d) CFX_TRY(rhs_dir(ctx, d, ctx->one(ctx->t_rhs[d][ctx->rhs_var[d]], b), d > 0));
  return 0;
}

// runs f(batch) over all blocks at once, or block by block
template <class F>
int per_batch(cfx_ctx *ctx, const Table &t, F &&f) {
  if (ctx->batched) return f(ctx->all(t));
  for (size_t b = 0; b < ctx->views.size(); ++b)
    if (int rc = f(ctx->one(t, (int)b))) return rc;
  return 0;
}

int do_dt(cfx_ctx *ctx, int global, int store_local) {
  auto *bits = reinterpret_cast<unsigned long long *>(ctx->d_dt);
  if (global) CFX_TRY(ctx->timed_launch(K_DT, [&] { return launch_dt_reset(bits, ctx->stream); }));
  CFX_TRY(per_batch(ctx, ctx->t_dt, [&](const Batch &b) {
    return ctx->timed_launch(K_DT, [&] {
      return launch_dt(b, ctx->S, ctx->ns, store_local, global ? bits : ctx->d_minscratch, ctx->stream);
    });
  }));
  return 0;
}

int do_update(cfx_ctx *ctx, int stage) {
  const KernelSet64 &k = g_kernels64[ctx->ns];
  CFX_TRY(per_batch(ctx, ctx->t_int, [&](const Batch &b) {
    return ctx->timed_launch(K_UPDATE, [&] {
      return k.update(b, ctx->sp_dev, ctx->S, ctx->mixed ? 1 : 0, stage, ctx->d_dt, ctx->d_flags, ctx->d_newton,
                      ctx->stream);
    });
  }));
  return 0;
}

// LU-SGS solve (the implicit treatment of Driver.cu:156): coefficients, then
// the forward and backward sweeps tile hyperplane by tile hyperplane; dq <- dQ
int do_lusgs(cfx_ctx *ctx) {
  const KernelSet64 &k = g_kernels64[ctx->ns];
  CFX_TRY(per_batch(ctx, ctx->t_int, [&](const Batch &b) {
    return ctx->timed_launch(K_UPDATE, [&] { return k.lusgs_coef(b, ctx->S, ctx->stream); });
  }));
  for (int pass = 0; pass < 2; ++pass)
    for (int n = 0; n < ctx->n_tplanes; ++n) {
      const int tplane = pass == 0 ? n : ctx->n_tplanes - 1 - n;
      CFX_TRY(per_batch(ctx, ctx->t_tile, [&](const Batch &b) {
        return ctx->timed_launch(K_UPDATE, [&] { return k.lusgs_sweep(b, ctx->S, tplane, pass == 0, ctx->stream); });
      }));
    }
  return 0;
}

int do_bc(cfx_ctx *ctx) {
  if (ctx->batched) return run_levels(ctx, ctx->sched_bc, 0, ctx->sched_bc.levels());
  return run_levels(ctx, ctx->sched_seq, 0, ctx->n_bc_ops);
}

int do_exchange(cfx_ctx *ctx) {
  if (ctx->batched) return run_levels(ctx, ctx->sched_ex, 0, ctx->sched_ex.levels());
  return run_levels(ctx, ctx->sched_seq, ctx->n_bc_ops, ctx->sched_seq.levels());
}

// boundary conditions followed by the inner-face exchange; after_update:
// the step's schedule, whose folded outflow rows the update kernel has filled
int do_ghosts(cfx_ctx *ctx, bool after_update) {
  const Schedule &all = after_update ? ctx->sched_all_step : ctx->sched_all;
  const Schedule &seq = after_update ? ctx->sched_seq_step : ctx->sched_seq;
  if (ctx->batched) return run_levels(ctx, all, 0, all.levels());
  return run_levels(ctx, seq, 0, seq.levels());
}

int do_elim(cfx_ctx *ctx) {
  if (ctx->dim != 2) return 0;
  for (auto &B : ctx->views)
    CFX_TRY(ctx->timed_launch(K_OTHER, [&] { return launch_eliminate_k(B, ctx->ns, ctx->stream); }));
  return 0;
}

int do_props(cfx_ctx *ctx) {
  if (ctx->transport == CFX_TRANSPORT_MLP) {
    if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate transport selected but not trained or loaded");
    auto fn = g_kernels32[ctx->ns].props_mlp;
    CFX_TRY(per_batch(ctx, ctx->t_pad, [&](const Batch &b) {
      return ctx->timed_launch(K_PROPS, [&] {
        return fn(b, ctx->sp_dev, ctx->S, ctx->mlp_dev, ctx->mlp_ops_inf, ctx->mlp_bias_inf, ctx->mlp_counters,
                  ctx->stream);
      });
    }));
    return 0;
  }
  auto fn = ctx->mixed ? g_kernels32[ctx->ns].props : g_kernels64[ctx->ns].props;
  CFX_TRY(per_batch(ctx, ctx->t_pad, [&](const Batch &b) {
    return ctx->timed_launch(K_PROPS, [&] { return fn(b, ctx->sp_dev, ctx->S, ctx->stream); });
  }));
  return 0;
}

int do_store_last(cfx_ctx *ctx) {
  CFX_TRY(per_batch(ctx, ctx->t_last, [&](const Batch &b) {
    return ctx->timed_launch(K_OTHER, [&] { return launch_store_last(b, ctx->stream); });
  }));
  return 0;
}

// Driver::steady_simulation loop body (Driver.cu:135-190): with rk_stages 1
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
    if (ctx->lusgs && (rc = do_lusgs(ctx))) return rc;
    if ((rc = do_update(ctx, s))) return rc;
    if ((rc = do_ghosts(ctx, true))) return rc;
    if ((rc = do_elim(ctx))) return rc;
    if ((rc = do_props(ctx))) return rc;
  }
  return 0;
}

// One step captured as a graph (kernel nodes only: the dt reset is a kernel,
// nothing is copied or set by the runtime inside the step).
int ensure_graph(cfx_ctx *ctx) {
  if (ctx->graph_exec) return 0;
  CFX_TRY((int)hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = do_step(ctx);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return ctx->fail((int)e, "hipStreamEndCapture");
  ctx->graph = g;
  CFX_TRY((int)hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
  return 0;
}

int run_steps(cfx_ctx *ctx, int n_steps) {
  if (ctx->use_graph && !ctx->profiling) {
    if (n_steps > 0) CFX_TRY(ensure_graph(ctx));
    for (int s = 0; s < n_steps; ++s) CFX_TRY((int)hipGraphLaunch(ctx->graph_exec, ctx->stream));
    return 0;
  }
  for (int s = 0; s < n_steps; ++s)
    if (int rc = do_step(ctx)) return rc;
  return 0;
}

}  // namespace

extern "C" {

namespace cfx {
namespace {

// The (ngg+1)-layer coordinate arrays of every block built on the device
// (Mesh::init_ghost_grid, Mesh.cpp:718-922): interior vertices uploaded and
// scattered, mirror extrapolation per block, the inter-block copies in the
// reference's block and face order, then edges and corners per block; the
// arrays stay resident (CFX_COORD, the metric fallback).  A block whose inner
// face names the block itself would read points the same copy writes, so
// such meshes keep the host path (ctx->coords empty, hb[].xyz filled).
int device_ghost_grid(cfx_ctx *ctx, const cfx_block *blocks, int n_block) {
  const int nb = n_block;
  ctx->coords.assign(nb, nullptr);
  int rc = 0;
  long long max_ni = 0;
  for (int b = 0; b < nb; ++b) {
    const HostBlock &H = ctx->hb[b];
    ctx->coords[b] = ctx->dalloc<double>(3 * (size_t)H.gx * H.gy * H.gz, rc);
    max_ni = std::max(max_ni, H.ni);
  }
  if (rc) return ctx->fail(rc, "hipMalloc coordinates");
  double *raw = nullptr;
  rc = (int)hipMalloc(&raw, sizeof(double) * 3 * (size_t)max_ni);
  for (int b = 0; b < nb && !rc; ++b) {
    const HostBlock &H = ctx->hb[b];
    const size_t ni = (size_t)H.ni;
    const double *src[3] = {blocks[b].x, blocks[b].y, blocks[b].z};
    for (int c = 0; c < 3 && !rc; ++c) rc = (int)hipMemcpy(raw + c * ni, src[c], sizeof(double) * ni, hipMemcpyHostToDevice);
    if (!rc) rc = launch_coord_scatter(raw, H.mx, H.my, H.mz, ctx->coords[b], nullptr);
    if (!rc) rc = launch_ghost_extrapolate(ctx->coords[b], H.mx, H.my, H.mz, ctx->dim, nullptr);
    if (!rc) rc = (int)hipDeviceSynchronize();  // raw is reused by the next block
  }
  (void)hipFree(raw);
  if (rc) return ctx->fail(rc, "device ghost grid");
  for (const GhostCopy &g : inner_ghost_copies(ctx->hb, blocks, ctx->dim)) {
    const HostBlock &D = ctx->hb[g.blk], &S = ctx->hb[g.src];
    if ((rc = launch_ghost_inner(g, ctx->coords[g.blk], D.mx, D.my, D.mz, ctx->coords[g.src], S.mx, S.my, S.mz,
                                 nullptr)))
      return ctx->fail(rc, "device inner ghost grid");
  }
  for (int b = 0; b < nb; ++b) {
    const HostBlock &H = ctx->hb[b];
    if ((rc = launch_ghost_corners(ctx->coords[b], H.mx, H.my, H.mz, nullptr))) return ctx->fail(rc, "device corners");
  }
  rc = (int)hipDeviceSynchronize();
  return rc ? ctx->fail(rc, "device ghost grid") : 0;
}

// The host copy of block b's coordinates (downloaded from the device arrays
// when the ghost grid was built there).
int host_coords(cfx_ctx *ctx, int b) {
  HostBlock &H = ctx->hb[b];
  if (!H.xyz[0].empty() || ctx->coords.empty()) return 0;
  const size_t ng = (size_t)H.gx * H.gy * H.gz;
  for (int c = 0; c < 3; ++c) {
    H.xyz[c].resize(ng);
    if (int rc = (int)hipMemcpy(H.xyz[c].data(), ctx->coords[b] + c * ng, sizeof(double) * ng, hipMemcpyDeviceToHost))
      return ctx->fail(rc, "coordinate download");
  }
  return 0;
}

// Metrics of one block computed on the device from its ghost-grid
// coordinates; a block with a non-positive Jacobian anywhere is recomputed on
// the host with the reference's fallback rules and uploaded.
int device_metrics(cfx_ctx *ctx, int b, const BlockView &V) {
  HostBlock &H = ctx->hb[b];
  const size_t ng = (size_t)H.gx * H.gy * H.gz, np = (size_t)H.np;
  const bool resident = !ctx->coords.empty();
  double *xyz = resident ? ctx->coords[b] : nullptr;
  int *bad = nullptr;
  int rc = resident ? 0 : (int)hipMalloc(&xyz, sizeof(double) * 3 * ng);
  if (!rc) rc = (int)hipMalloc(&bad, sizeof(int));
  for (int c = 0; c < 3 && !rc && !resident; ++c)
    rc = (int)hipMemcpy(xyz + c * ng, H.xyz[c].data(), sizeof(double) * ng, hipMemcpyHostToDevice);
  if (!rc) rc = (int)hipMemset(bad, 0, sizeof(int));
  if (!rc) rc = launch_metrics(xyz, H.mx, H.my, H.mz, V.metric, V.jac, V.mixed ? V.metric32 : nullptr,
                               V.mixed ? V.jac32 : nullptr, bad, nullptr);
  int n_bad = 0;
  if (!rc) rc = (int)hipMemcpy(&n_bad, bad, sizeof(int), hipMemcpyDeviceToHost);
  if (!resident) (void)hipFree(xyz);
  (void)hipFree(bad);
  if (rc) return ctx->fail(rc, "device metrics");
  if (n_bad == 0) return 0;
  if ((rc = host_coords(ctx, b))) return rc;
  if ((rc = compute_metrics(H, ctx->err))) return rc;
  rc = (int)hipMemcpy(V.metric, H.metric.data(), sizeof(double) * 9 * np, hipMemcpyHostToDevice);
  if (!rc) rc = (int)hipMemcpy(V.jac, H.jac.data(), sizeof(double) * np, hipMemcpyHostToDevice);
  if (!rc && V.mixed) {
    std::vector<float> m32(9 * np), j32(np);
    for (size_t i = 0; i < 9 * np; ++i) m32[i] = (float)H.metric[i];
    for (size_t i = 0; i < np; ++i) j32[i] = (float)H.jac[i];
    rc = (int)hipMemcpy(V.metric32, m32.data(), sizeof(float) * 9 * np, hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(V.jac32, j32.data(), sizeof(float) * np, hipMemcpyHostToDevice);
  }
  std::vector<double>().swap(H.metric);
  std::vector<double>().swap(H.jac);
  if (resident)
    for (int c = 0; c < 3; ++c) std::vector<double>().swap(H.xyz[c]);
  return rc ? ctx->fail(rc, "metric upload") : 0;
}

}  // namespace
}  // namespace cfx

int cfx_version(void) { return CFX_ABI_VERSION; }

int cfx_ghost_grid(const cfx_block *blocks, int n_block, double *const *xyz) {
  if (!blocks || n_block <= 0 || !xyz) return CFX_E_ARG;
  std::vector<HostBlock> hb;
  int dim = 0;
  std::string err;
  if (int rc = build_mesh(blocks, n_block, hb, dim, err, /*metrics=*/false)) return rc;
  for (int b = 0; b < n_block; ++b) {
    const HostBlock &B = hb[b];
    if (!xyz[b]) return CFX_E_ARG;
    for (int c = 0; c < 3; ++c)
      for (int k = -kNgg; k < B.mz + kNgg; ++k)
        for (int j = -kNgg; j < B.my + kNgg; ++j)
          for (int i = -kNgg; i < B.mx + kNgg; ++i)
            xyz[b][c * B.np + B.pidx(i, j, k)] = B.xyz[c][B.gidx(i, j, k)];
  }
  return 0;
}

int cfx_inflow_state(const cfx_params *params, const cfx_species *species, int index, double bv[6], double *yk) {
  if (!params || !bv || index < 0 || index >= params->n_inflow || !params->inflow) return CFX_E_ARG;
  SpeciesDev tables;
  std::string err;
  if (int rc = build_species(species, tables, err)) return rc;
  SchemeDev S{};
  S.gamma_air = params->gamma_air > 0 ? params->gamma_air : 1.4;
  S.r_air = params->r_air > 0 ? params->r_air : kRuH / kMwAir;
  InflowDev d;
  if (int rc = build_inflow(params->inflow[index], tables.d, S, d, err)) return rc;
  for (int v = 0; v < 6; ++v) bv[v] = d.st[v];
  if (yk)
    for (int l = 0; l < tables.d.ns; ++l) yk[l] = d.y[l];
  return 0;
}

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
  if (!g_kernels64[ctx->ns].present) {
    ctx->err = "no kernels built for " + std::to_string(ctx->ns) + " species";
    return bail(CFX_E_UNSUPPORTED);
  }
  if (params->inviscid_scheme != 3 && params->inviscid_scheme != 0) return bail(CFX_E_UNSUPPORTED);
  if ((params->precision != CFX_FP64 && params->precision != CFX_MIXED) ||
      (params->transport != CFX_TRANSPORT_WILKE && params->transport != CFX_TRANSPORT_MLP))
    return bail(CFX_E_UNSUPPORTED);
  ctx->mixed = params->precision == CFX_MIXED;
  if (params->transport == CFX_TRANSPORT_MLP &&
      (!ctx->mixed || ctx->ns < 1 || ctx->ns > kMlpMaxSp || !g_kernels32[ctx->ns].props_mlp)) {
    ctx->err = "surrogate transport needs the mixed-precision path and 1..14 species";
    return bail(CFX_E_UNSUPPORTED);
  }
  ctx->transport = params->transport;
  if (ctx->mixed && !g_kernels32[ctx->ns].present) {
    ctx->err = "no FP32 kernels built for " + std::to_string(ctx->ns) + " species";
    return bail(CFX_E_UNSUPPORTED);
  }
  bool self_face = false;  // an inner face that names its own block keeps the host ghost grid
  for (int b = 0; b < n_block; ++b)
    for (int f = 0; f < blocks[b].n_inner; ++f) self_face |= blocks[b].inner[f].target_block == b + 1;
  rc = build_mesh(blocks, n_block, ctx->hb, ctx->dim, ctx->err, /*metrics=*/false, /*ghosts=*/self_face);
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
  S.lusgs = params->lusgs ? 1 : 0;
  S.omega = params->lusgs_omega > 0 ? params->lusgs_omega : 1.0;
  ctx->rk = S.rk;
  // default RHS tiles of a 3-D context: shapes among the fastest the
  // autotuner measures on MI355X for the 256x128x128 9-species block (FP64:
  // x 32x4x2, y 8x32x1 and z 8x32x1 with c fastest; FP32: x 32x4x2 at 4
  // waves, y 8x32x1, z 8x64x1 (512 threads) with c fastest); cfx_autotune
  // re-picks per context
  if (ctx->dim == 3) {
    const int d64[3] = {5, 1, 1}, d32[3] = {5, 2, 3};
    for (int d = 0; d < 3; ++d) ctx->rhs_var[d] = ctx->mixed ? d32[d] : d64[d];
  }
  ctx->lusgs = S.lusgs != 0;
  if (ctx->lusgs && (S.rk != 1 || ctx->mixed)) {
    ctx->err = "the LU-SGS step needs rk_stages 1 (local time step) and FP64";
    return bail(CFX_E_UNSUPPORTED);
  }
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
  rc = (int)hipMemcpy(ctx->sp_dev, &ctx->sp_host, sizeof(SpeciesDev), hipMemcpyHostToDevice);
  if (!rc && !ctx->inflow_host.empty())
    rc = (int)hipMemcpy(ctx->inflow_dev, ctx->inflow_host.data(), sizeof(InflowDev) * ctx->inflow_host.size(),
                        hipMemcpyHostToDevice);
  if (!rc) rc = (int)hipMemset(ctx->d_flags, 0, sizeof(int));
  if (!rc) rc = (int)hipMemset(ctx->d_newton, 0, sizeof(unsigned long long));
  const double dt0 = 0.0;
  if (!rc) rc = (int)hipMemcpy(ctx->d_dt, &dt0, sizeof(double), hipMemcpyHostToDevice);
  if (rc) return bail(ctx->fail(rc, "context upload"));
  for (int b = 0; b < n_block; ++b)
    if (ctx->hb[b].np >= (1LL << 29)) {  // the RHS kernels address a block with 32-bit byte offsets
      ctx->err = "block " + std::to_string(b) + " exceeds 2^29 padded points";
      return bail(CFX_E_UNSUPPORTED);
    }
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
    if (ctx->lusgs) V.lus = ctx->dalloc<double>(4 * ni, rc);
    if (rc) return bail(ctx->fail(rc, "hipMalloc block arrays"));
    V.mixed = ctx->mixed ? 1 : 0;
    if (ctx->mixed) {
      V.bv32 = ctx->dalloc<float>(6 * np, rc);
      V.yk32 = ctx->dalloc<float>(ctx->ns * np, rc);
      V.props32 = ctx->dalloc<float>((4 + ctx->ns) * np, rc);
      V.metric32 = ctx->dalloc<float>(9 * np, rc);
      V.jac32 = ctx->dalloc<float>(np, rc);
      V.dq32 = ctx->dalloc<float>(ctx->nv * ni, rc);
      if (rc) return bail(ctx->fail(rc, "hipMalloc FP32 arrays"));
      rc = (int)hipMemset(V.props32, 0, sizeof(float) * (4 + ctx->ns) * np);
      if (!rc) rc = (int)hipMemset(V.dq32, 0, sizeof(float) * ctx->nv * ni);
    }
    if (!rc) rc = (int)hipMemset(V.props, 0, sizeof(double) * (4 + ctx->ns) * np);
    if (!rc) rc = (int)hipMemset(V.dq, 0, sizeof(double) * ctx->nv * ni);
    if (rc) return bail(ctx->fail(rc, "block array clear"));
    ctx->views.push_back(V);
    ctx->n_grid_total += H.ni;
  }
  if (!self_face && (rc = device_ghost_grid(ctx, blocks, n_block))) return bail(rc);
  for (int b = 0; b < n_block; ++b)
    if ((rc = device_metrics(ctx, b, ctx->views[b]))) return bail(rc);
  ctx->views_dev = ctx->dalloc<BlockView>(ctx->views.size(), rc);
  if (rc) return bail(ctx->fail(rc, "hipMalloc block table"));
  if ((rc = (int)hipMemcpy(ctx->views_dev, ctx->views.data(), sizeof(BlockView) * ctx->views.size(),
                           hipMemcpyHostToDevice)))
    return bail(ctx->fail(rc, "block table upload"));
  if ((rc = build_tables(ctx))) return bail(ctx->fail(rc, "work tables"));
  if ((rc = build_ops(ctx))) return bail(ctx->fail(rc, "ghost-fill schedule"));
  ctx->n_partial = ctx->t_res.host.back();
  ctx->d_partial = ctx->dalloc<double>(4 * (size_t)ctx->n_partial, rc);
  if (rc) return bail(ctx->fail(rc, "hipMalloc partials"));
  {
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      ctx->n_cu = ncu;
    for (int f32 = 0; f32 < 2; ++f32) {
      if (f32 && !ctx->mixed) continue;
      for (int d = 0; d < ctx->dim; ++d)
        for (int v = 0; v < rhs_variants(ctx->dim); ++v) {
          const RhsEntry &e = f32 ? (ctx->dim == 3 ? g_kernels32[ctx->ns].rhs3[d][v] : g_kernels32[ctx->ns].rhs2[d][v])
                                  : (ctx->dim == 3 ? g_kernels64[ctx->ns].rhs3[d][v] : g_kernels64[ctx->ns].rhs2[d][v]);
          ctx->occ[f32][d][v] = e.occupancy();
        }
    }
  }
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return bail(ctx->fail((int)e, "device setup"));
  *out = ctx;
  return 0;
}

void cfx_destroy(cfx_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->drop_graph();
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
  if (V.mixed && what == CFX_YK) {  // the mixed path keeps Y only as rho*Y / rho
    std::vector<double> cv(ctx->nv * np);
    CFX_TRY((int)hipMemcpy(cv.data(), V.cv, sizeof(double) * cv.size(), hipMemcpyDeviceToHost));
    for (size_t p = 0; p < np; ++p) {
      const double ir = 1.0 / cv[p];
      for (int l = 0; l < ctx->ns; ++l) dst[l * np + p] = cv[(5 + l) * np + p] * ir;
    }
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
    case CFX_COORD: {
      CFX_TRY(host_coords(ctx, blk));
      const HostBlock &B = ctx->hb[blk];
      for (int c = 0; c < 3; ++c)
        for (int k = -kNgg; k < B.mz + kNgg; ++k)
          for (int j = -kNgg; j < B.my + kNgg; ++j)
            for (int i = -kNgg; i < B.mx + kNgg; ++i) dst[c * np + B.pidx(i, j, k)] = B.xyz[c][B.gidx(i, j, k)];
      return 0;
    }
    default: return ctx->fail(CFX_E_ARG, "unknown download selector");
  }
  if (n) CFX_TRY((int)hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int cfx_init_conservative(cfx_ctx *ctx) {
  if (!ctx) return CFX_E_ARG;
  const KernelSet64 &k = g_kernels64[ctx->ns];
  CFX_TRY(per_batch(ctx, ctx->t_pad, [&](const Batch &b) { return k.cv_from_bv(b, ctx->sp_dev, ctx->S, ctx->stream); }));
  if (int rc = do_ghosts(ctx, false)) return rc;
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

int cfx_lusgs(cfx_ctx *ctx) {
  if (!ctx) return CFX_E_ARG;
  if (!ctx->lusgs) return ctx->fail(CFX_E_ARG, "cfx_lusgs needs cfx_params.lusgs = 1");
  return do_lusgs(ctx);
}

int cfx_apply_bc(cfx_ctx *ctx) { return ctx ? do_bc(ctx) : CFX_E_ARG; }
int cfx_exchange(cfx_ctx *ctx) { return ctx ? do_exchange(ctx) : CFX_E_ARG; }
int cfx_eliminate_k_gradient(cfx_ctx *ctx) { return ctx ? do_elim(ctx) : CFX_E_ARG; }
int cfx_update_properties(cfx_ctx *ctx) { return ctx ? do_props(ctx) : CFX_E_ARG; }

int cfx_residual(cfx_ctx *ctx, double res[4]) {
  if (!ctx || !res) return CFX_E_ARG;
  CFX_TRY(per_batch(ctx, ctx->t_res, [&](const Batch &b) { return launch_residual_partial(b, ctx->d_partial, ctx->stream); }));
  CFX_TRY(launch_residual_final(ctx->d_partial, ctx->n_partial, ctx->d_res, ctx->stream));
  double h[4];
  CFX_TRY((int)hipMemcpyAsync(h, ctx->d_res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  for (int l = 0; l < 4; ++l) res[l] = std::sqrt(h[l] / (double)ctx->n_grid_total);
  return 0;
}

int cfx_step(cfx_ctx *ctx, int n_steps) {
  if (!ctx || n_steps < 0) return CFX_E_ARG;
  return run_steps(ctx, n_steps);
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
  if (variant < 0 || variant >= rhs_variants(ctx->dim)) return ctx->fail(CFX_E_ARG, "rhs variant out of range");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  for (int d = 0; d < 3; ++d) ctx->rhs_var[d] = variant, ctx->rhs_persist[d] = 0;
  ctx->drop_graph();
  return 0;
}

int cfx_rhs_variants(cfx_ctx *ctx) { return ctx ? rhs_variants(ctx->dim) : CFX_E_ARG; }

int cfx_set_rhs_tiles(cfx_ctx *ctx, const int variant[3], const int persist[3]) {
  if (!ctx || !variant || !persist) return CFX_E_ARG;
  for (int d = 0; d < ctx->dim; ++d)
    if (variant[d] < 0 || variant[d] >= rhs_variants(ctx->dim) || persist[d] < 0 || persist[d] > 64)
      return ctx->fail(CFX_E_ARG, "rhs tile choice out of range");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  for (int d = 0; d < ctx->dim; ++d) ctx->rhs_var[d] = variant[d], ctx->rhs_persist[d] = persist[d];
  ctx->drop_graph();
  return 0;
}

int cfx_get_rhs_tiles(cfx_ctx *ctx, int variant[3], int persist[3]) {
  if (!ctx || !variant || !persist) return CFX_E_ARG;
  for (int d = 0; d < 3; ++d) variant[d] = ctx->rhs_var[d], persist[d] = ctx->rhs_persist[d];
  return 0;
}

// Times every tile shape and persistence factor of each face direction on
// this context's blocks (reps launches each, after one warm-up launch) and
// keeps the fastest.  dq holds no result afterwards; the next RHS rewrites it.
int cfx_autotune(cfx_ctx *ctx, int reps, double *best_ms) {
  if (!ctx || reps <= 0) return CFX_E_ARG;
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  static const int kPersist[] = {0, 1, 2, 4};
  int rc = 0;
  for (int d = 0; d < ctx->dim && !rc; ++d) {
    double best = 1e30;
    int bv = 0, bp = 0;
    for (int v = 0; v < rhs_variants(ctx->dim) && !rc; ++v)
      for (int p : kPersist) {
        ctx->rhs_var[d] = v;
        ctx->rhs_persist[d] = p;
        auto run = [&] {
          if (ctx->batched) return rhs_dir(ctx, d, ctx->all(ctx->t_rhs[d][v]), d > 0);
          for (size_t k = 0; k < ctx->views.size(); ++k)
            if (int r = rhs_dir(ctx, d, ctx->one(ctx->t_rhs[d][v], (int)k), d > 0)) return r;
          return 0;
        };
        if ((rc = run())) break;
        (void)hipEventRecord(a, ctx->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = run();
        (void)hipEventRecord(b, ctx->stream);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms / reps < best) best = ms / reps, bv = v, bp = p;
      }
    ctx->rhs_var[d] = bv;
    ctx->rhs_persist[d] = bp;
    if (best_ms) best_ms[d] = best;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->profiling = prof;
  ctx->drop_graph();
  if (rc) return ctx->fail(rc, "autotune launch");
  return (int)hipStreamSynchronize(ctx->stream);
}

int cfx_set_execution(cfx_ctx *ctx, int batched, int graph) {
  if (!ctx) return CFX_E_ARG;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  ctx->batched = batched != 0;
  ctx->use_graph = graph != 0;
  ctx->drop_graph();
  return 0;
}

int cfx_schedule_info(cfx_ctx *ctx, int *n_ops, int *n_levels) {
  if (!ctx) return CFX_E_ARG;
  if (n_ops) *n_ops = (int)ctx->ops_step.size();
  if (n_levels) *n_levels = ctx->sched_all_step.levels();
  return 0;
}

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
