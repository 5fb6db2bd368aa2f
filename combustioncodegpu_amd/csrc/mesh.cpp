// mesh.cpp — host geometry preprocessing of the step library.
//
// Face registration follows Boundary::register_boundary (Mesh.cpp:14-32) and
// InnerFace::register_boundary (Mesh.cpp:44-143); the ghost grid follows
// Mesh::init_ghost_grid (Mesh.cpp:718-872) including the inter-block ghost
// grid (Mesh.cpp:874-922); the Jacobian and metrics follow
// Block::compute_jac_metric (Mesh.cpp:226-488).  The reference stores the
// metric as a 1-based 3x3 matrix per point; here it is 9 SoA planes.
#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace cfx {
namespace {

inline int sgn(int a) { return (a > 0) - (a < 0); }

struct Span {
  int lo, hi, step;  // inclusive lo..hi walked with step (+1 / -1)
};

template <class F>
void walk(const Span &a, const Span &b, const Span &c, F &&f) {
  for (int i = a.lo; a.step > 0 ? i <= a.hi : i >= a.hi; i += a.step)
    for (int j = b.lo; b.step > 0 ? j <= b.hi : j >= b.hi; j += b.step)
      for (int k = c.lo; c.step > 0 ? k <= c.hi : k >= c.hi; k += c.step) f(i, j, k);
}

FaceRange register_boundary_face(const cfx_bc &d, int dim) {
  FaceRange r{};
  r.direction = 1;
  r.face = 0;
  for (int q = 0; q < 3; ++q) {
    r.rs[q] = d.range_start[q];
    r.re[q] = d.range_end[q];
  }
  for (int q = 0; q < dim; ++q)
    if (r.rs[q] == r.re[q]) {
      r.face = q;
      if (r.rs[q] == 0) r.direction = -1;
      break;
    }
  for (int q = 0; q < dim; ++q) {
    if (q == r.face) continue;
    if (r.re[q] < r.rs[q]) std::swap(r.rs[q], r.re[q]);
    r.rs[q] -= kNgg;
    r.re[q] += kNgg;
  }
  if (dim == 2) {
    r.rs[2] = -kNgg;
    r.re[2] = kNgg;
  }
  return r;
}

// Turns the signed 1-based Plot3D-style ranges of an inner face into the
// loop description used by the exchange (Mesh.cpp:44-143).
InnerDev register_inner_face(const cfx_inner &d, int dim) {
  InnerDev f{};
  int rs[3], re[3], ts[3], te[3];
  for (int q = 0; q < 3; ++q) {
    rs[q] = d.src[2 * q];
    re[q] = d.src[2 * q + 1];
    ts[q] = d.tar[2 * q];
    te[q] = d.tar[2 * q + 1];
  }
  int face = 0, direction = 1, tface = 0, tdirection = 1;
  int src_tar[3] = {0, 0, 0};
  for (int q = 0; q < dim; ++q)
    if (rs[q] == re[q]) {
      face = q;
      if (rs[q] == 1) direction = -1;
      --rs[q];
      --re[q];
      break;
    }
  for (int q = 0; q < dim; ++q)
    if (ts[q] == te[q]) {
      tface = q;
      if (ts[q] == 1) tdirection = -1;
      --ts[q];
      --te[q];
      break;
    }
  src_tar[tface] = -(direction * tdirection) * (face + 1);
  auto orient = [&](int src_axis, int tar_axis) {
    src_tar[tar_axis] = sgn((te[tar_axis] - ts[tar_axis]) * (re[src_axis] - rs[src_axis])) * (src_axis + 1);
  };
  if (dim == 2) {
    const int nf = 1 - face, tnf = 1 - tface;
    rs[nf] = std::abs(rs[nf]) - 1;
    re[nf] = std::abs(re[nf]) - 1;
    ts[nf] = std::abs(ts[nf]) - 1;
    te[nf] = std::abs(te[nf]) - 1;
    orient(nf, tnf);
    src_tar[2] = 3;
    rs[2] = re[2] = ts[2] = te[2] = 0;
  } else {
    // the axis flagged by negative values is the first in-plane axis
    int nf = 0, tnf = 0;
    for (int q = 0; q < dim; ++q)
      if (rs[q] < 0) {
        nf = q;
        rs[q] = -rs[q] - 1;
        re[q] = -re[q] - 1;
        break;
      }
    for (int q = 0; q < dim; ++q)
      if (ts[q] < 0) {
        tnf = q;
        ts[q] = -ts[q] - 1;
        te[q] = -te[q] - 1;
        break;
      }
    orient(nf, tnf);
    nf = 3 - face - nf;
    tnf = 3 - tface - tnf;
    --rs[nf];
    --re[nf];
    --ts[tnf];
    --te[tnf];
    orient(nf, tnf);
  }
  // normal range: the shared layer plus the ngg ghost layers
  const int re_face = re[face];
  rs[face] = direction < 0 ? -kNgg : re_face;
  re[face] = direction < 0 ? 0 : re_face + kNgg;
  const int sgn_tar = sgn(src_tar[tface]);
  if (tdirection > 0) ts[tface] = sgn_tar > 0 ? te[tface] - kNgg : te[tface];
  else ts[tface] = sgn_tar > 0 ? 0 : kNgg;
  for (int q = 0; q < dim; ++q) {
    f.loop_dir[q] = rs[q] <= re[q] ? 1 : -1;
    f.target_loop_dir[q] = sgn(src_tar[q]);
    src_tar[q] = std::abs(src_tar[q]) - 1;
  }
  if (dim == 2) {
    f.loop_dir[2] = 1;
    f.target_loop_dir[2] = 1;
    src_tar[2] = 2;
  }
  for (int q = 0; q < 3; ++q)
    if (q != face) {  // corners travel with the face
      rs[q] -= f.loop_dir[q] * kNgg;
      re[q] += f.loop_dir[q] * kNgg;
      ts[q] -= f.target_loop_dir[q] * f.loop_dir[q] * kNgg;
    }
  for (int q = 0; q < 3; ++q) {
    f.rs[q] = rs[q];
    f.re[q] = re[q];
    f.ts[q] = ts[q];
    f.src_tar[q] = src_tar[q];
    f.n_point[q] = std::abs(rs[q] - re[q]) + 1;
  }
  f.face = face;
  f.direction = direction;
  f.target_block = d.target_block - 1;
  // the target face and direction are folded into ts / target_loop_dir; the
  // ghost-grid copy recomputes them from the descriptor (inner_target_geom)
  return f;
}

struct InnerGeom {  // what the ghost-grid copy needs beyond InnerDev
  int tface, tdirection;
};

InnerGeom inner_target_geom(const cfx_inner &d, int dim) {
  InnerGeom g{0, 1};
  for (int q = 0; q < dim; ++q)
    if (d.tar[2 * q] == d.tar[2 * q + 1]) {
      g.tface = q;
      if (d.tar[2 * q] == 1) g.tdirection = -1;
      break;
    }
  return g;
}

void extrapolate_ghost_grid(HostBlock &B, int dim) {
  const int G = kNgg + 1;
  const int m[3] = {B.mx, B.my, B.mz};
  // per direction, over the interior of the other two directions
  for (int a = 0; a < 3; ++a) {
    if (a == 2 && dim == 2) break;
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    for (int u = 0; u < m[b]; ++u)
      for (int v = 0; v < m[c]; ++v) {
        int p[3];
        auto at = [&](int cc, int s) -> double & {
          p[a] = s;
          p[b] = u;
          p[c] = v;
          return B.coord(cc, p[0], p[1], p[2]);
        };
        for (int g = 1; g <= G; ++g)
          for (int cc = 0; cc < 3; ++cc) {
            at(cc, -g) = 2 * at(cc, 0) - at(cc, g);
            at(cc, m[a] - 1 + g) = 2 * at(cc, m[a] - 1) - at(cc, m[a] - 1 - g);
          }
      }
  }
  if (dim == 2) {  // unit spacing in z (Mesh.cpp:769-785)
    for (int j = 0; j < B.my; ++j)
      for (int i = 0; i < B.mx; ++i)
        for (int g = 1; g <= G; ++g) {
          B.coord(0, i, j, -g) = B.coord(0, i, j, 0);
          B.coord(1, i, j, -g) = B.coord(1, i, j, 0);
          B.coord(2, i, j, -g) = B.coord(2, i, j, 0) - g;
          B.coord(0, i, j, B.mz - 1 + g) = B.coord(0, i, j, B.mz - 1);
          B.coord(1, i, j, B.mz - 1 + g) = B.coord(1, i, j, B.mz - 1);
          B.coord(2, i, j, B.mz - 1 + g) = B.coord(2, i, j, B.mz - 1) + g;
        }
  }
}

void corner_ghost_grid(HostBlock &B) {
  const int G = kNgg + 1;
  const int mx = B.mx, my = B.my, mz = B.mz;
  // parallelogram rule: P(a,b) = P(a,edge_b) + P(edge_a,b) - P(edge_a,edge_b)
  for (int cc = 0; cc < 3; ++cc) {
    auto X = [&](int i, int j, int k) -> double & { return B.coord(cc, i, j, k); };
    for (int s = 1; s <= G; ++s)
      for (int t = 1; t <= G; ++t)
        for (int k = 0; k < mz; ++k) {
          const int i1 = -s, j1 = -t, i2 = mx - 1 + s, j2 = my - 1 + t;
          X(i1, j1, k) = X(i1, 0, k) + X(0, j1, k) - X(0, 0, k);
          X(i2, j2, k) = X(i2, my - 1, k) + X(mx - 1, j2, k) - X(mx - 1, my - 1, k);
          X(i1, j2, k) = X(i1, my - 1, k) + X(0, j2, k) - X(0, my - 1, k);
          X(i2, j1, k) = X(i2, 0, k) + X(mx - 1, j1, k) - X(mx - 1, 0, k);
        }
    for (int s = 1; s <= G; ++s)
      for (int j = 0; j < my; ++j)
        for (int t = 1; t <= G; ++t) {
          const int i1 = -s, k1 = -t, i2 = mx - 1 + s, k2 = mz - 1 + t;
          X(i1, j, k1) = X(i1, j, 0) + X(0, j, k1) - X(0, j, 0);
          X(i2, j, k2) = X(i2, j, mz - 1) + X(mx - 1, j, k2) - X(mx - 1, j, mz - 1);
          X(i1, j, k2) = X(i1, j, mz - 1) + X(0, j, k2) - X(0, j, mz - 1);
          X(i2, j, k1) = X(i2, j, 0) + X(mx - 1, j, k1) - X(mx - 1, j, 0);
        }
    for (int i = -G; i <= mx + kNgg; ++i)
      for (int s = 1; s <= G; ++s)
        for (int t = 1; t <= G; ++t) {
          const int j1 = -s, k1 = -t, j2 = my - 1 + s, k2 = mz - 1 + t;
          X(i, j1, k1) = X(i, j1, 0) + X(i, 0, k1) - X(i, 0, 0);
          X(i, j2, k2) = X(i, j2, mz - 1) + X(i, my - 1, k2) - X(i, my - 1, mz - 1);
          X(i, j1, k2) = X(i, j1, mz - 1) + X(i, 0, k2) - X(i, 0, mz - 1);
          X(i, j2, k1) = X(i, j2, 0) + X(i, my - 1, k1) - X(i, my - 1, 0);
        }
  }
}

// Inter-block ghost grid (Mesh.cpp:874-922): the ngg+1 layers beyond the
// shared face come from the neighbour's points next to its face.
void inner_ghost_grid(std::vector<HostBlock> &blocks, const cfx_block *desc, int dim) {
  for (const GhostCopy &g : inner_ghost_copies(blocks, desc, dim)) {
    HostBlock &B = blocks[g.blk];
    const HostBlock &N = blocks[g.src];
    for (int a = 0; a < g.n[0]; ++a)
      for (int b = 0; b < g.n[1]; ++b)
        for (int c = 0; c < g.n[2]; ++c) {
          const int di[3] = {g.step[0] * a, g.step[1] * b, g.step[2] * c};
          const int t[3] = {g.t0[0] + g.tstep[0] * di[g.src_tar[0]], g.t0[1] + g.tstep[1] * di[g.src_tar[1]],
                            g.t0[2] + g.tstep[2] * di[g.src_tar[2]]};
          for (int cc = 0; cc < 3; ++cc)
            B.coord(cc, g.lo[0] + di[0], g.lo[1] + di[1], g.lo[2] + di[2]) = N.xyz[cc][N.gidx(t[0], t[1], t[2])];
        }
  }
}

}  // namespace

std::vector<GhostCopy> inner_ghost_copies(const std::vector<HostBlock> &blocks, const cfx_block *desc, int dim) {
  std::vector<GhostCopy> out;
  for (size_t b = 0; b < blocks.size(); ++b) {
    const HostBlock &B = blocks[b];
    for (size_t f = 0; f < B.inner.size(); ++f) {
      const InnerDev &fc = B.inner[f];
      const InnerGeom tg = inner_target_geom(desc[b].inner[f], dim);
      GhostCopy g{};
      g.blk = (int)b;
      g.src = fc.target_block;
      int hi[3];
      for (int q = 0; q < 3; ++q) {
        g.lo[q] = fc.rs[q];
        hi[q] = fc.re[q];
        g.t0[q] = fc.ts[q];
        g.step[q] = fc.loop_dir[q];
        g.tstep[q] = fc.target_loop_dir[q];
        g.src_tar[q] = fc.src_tar[q];
      }
      const int shift = fc.direction < 0 ? -1 : 1;
      g.lo[fc.face] += shift;
      hi[fc.face] += shift;
      g.t0[tg.tface] += tg.tdirection > 0 ? -1 : 1;
      // the loop runs lo, lo + step, ... until it has passed hi
      for (int q = 0; q < 3; ++q) g.n[q] = g.step[q] * (hi[q] - g.lo[q]) + 1;
      out.push_back(g);
    }
  }
  return out;
}

namespace {

bool metric_at(HostBlock &B, int i, int j, int k, double &J, double m[9]) {
  const auto X = [&](int cc, int di, int dj, int dk) { return B.coord(cc, i + di, j + dj, k + dk); };
  return metric_point(X, J, m);
}

}  // namespace

int compute_metrics(HostBlock &B, std::string &err) {
  const int mx = B.mx, my = B.my, mz = B.mz, g = kNgg;
  B.metric.assign(9 * B.np, 0.0);
  B.jac.assign(B.np, 0.0);
  const int kmin = mz == 1 ? 0 : 1, kmax = mz == 1 ? 0 : mz - 2;
  struct Region {
    Span a, b, c;
    int fb[3];  // neighbour to copy from on a non-positive Jacobian (0: fatal)
  };
  const Region regions[7] = {
      {{1, mx - 2, 1}, {1, my - 2, 1}, {kmin, kmax, 1}, {0, 0, 0}},
      {{1, mx - 2, 1}, {1, my - 2, 1}, {0, -g, -1}, {0, 0, 1}},
      {{1, mx - 2, 1}, {1, my - 2, 1}, {mz - 1, mz + g - 1, 1}, {0, 0, -1}},
      {{1, mx - 2, 1}, {0, -g, -1}, {-g, mz + g - 1, 1}, {0, 1, 0}},
      {{1, mx - 2, 1}, {my - 1, my + g - 1, 1}, {-g, mz + g - 1, 1}, {0, -1, 0}},
      {{0, -g, -1}, {-g, my + g - 1, 1}, {-g, mz + g - 1, 1}, {1, 0, 0}},
      {{mx - 1, mx + g - 1, 1}, {-g, my + g - 1, 1}, {-g, mz + g - 1, 1}, {-1, 0, 0}},
  };
  bool ok = true;
  for (const Region &r : regions) {
    walk(r.a, r.b, r.c, [&](int i, int j, int k) {
      double J, m[9];
      const long long p = B.pidx(i, j, k);
      if (metric_at(B, i, j, k, J, m)) {
        B.jac[p] = J;
        for (int c = 0; c < 9; ++c) B.metric[c * B.np + p] = m[c];
      } else if (r.fb[0] == 0 && r.fb[1] == 0 && r.fb[2] == 0) {
        ok = false;
      } else {
        const long long q = B.pidx(i + r.fb[0], j + r.fb[1], k + r.fb[2]);
        B.jac[p] = B.jac[q];
        for (int c = 0; c < 9; ++c) B.metric[c * B.np + p] = B.metric[c * B.np + q];
      }
    });
  }
  if (!ok) {
    err = "non-positive Jacobian inside a block";
    return CFX_E_ARG;
  }
  return 0;
}

int build_mesh(const cfx_block *blocks, int n_block, std::vector<HostBlock> &out, int &dim, std::string &err,
               bool metrics, bool ghosts) {
  if (!blocks || n_block <= 0) {
    err = "no blocks";
    return CFX_E_ARG;
  }
  dim = blocks[0].mz == 1 ? 2 : 3;
  out.assign(n_block, HostBlock{});
  for (int b = 0; b < n_block; ++b) {
    const cfx_block &d = blocks[b];
    if (d.mx < 3 || d.my < 3 || d.mz < 1 || (dim == 3 && d.mz < 3) || (dim == 2 && d.mz != 1) || !d.x || !d.y ||
        !d.z) {
      err = "block " + std::to_string(b) + ": bad extents or coordinates";
      return CFX_E_ARG;
    }
    HostBlock &B = out[b];
    B.mx = d.mx;
    B.my = d.my;
    B.mz = d.mz;
    B.px = d.mx + 2 * kNgg;
    B.py = d.my + 2 * kNgg;
    B.pz = d.mz + 2 * kNgg;
    B.np = (long long)B.px * B.py * B.pz;
    B.ni = (long long)d.mx * d.my * d.mz;
    B.gx = d.mx + 2 * (kNgg + 1);
    B.gy = d.my + 2 * (kNgg + 1);
    B.gz = d.mz + 2 * (kNgg + 1);
    const long long ng = (long long)B.gx * B.gy * B.gz;
    if (ghosts) {
      for (int c = 0; c < 3; ++c) B.xyz[c].assign(ng, 0.0);
      for (int k = 0; k < d.mz; ++k)
        for (int j = 0; j < d.my; ++j)
          for (int i = 0; i < d.mx; ++i) {
            const long long s = ((long long)k * d.my + j) * d.mx + i;
            B.coord(0, i, j, k) = d.x[s];
            B.coord(1, i, j, k) = d.y[s];
            B.coord(2, i, j, k) = d.z[s];
          }
    }
    for (int f = 0; f < d.n_bc; ++f) {
      const cfx_bc &bc = d.bc[f];
      if (bc.type != CFX_BC_WALL && bc.type != CFX_BC_SYMMETRY && bc.type != CFX_BC_INFLOW &&
          bc.type != CFX_BC_OUTFLOW) {
        err = "block " + std::to_string(b) + ": unknown boundary type " + std::to_string(bc.type);
        return CFX_E_ARG;
      }
      B.bc_range.push_back(register_boundary_face(bc, dim));
      B.bc_type.push_back(bc.type);
      B.bc_cond.push_back(bc.cond);
    }
    for (int f = 0; f < d.n_inner; ++f) {
      if (d.inner[f].target_block < 1 || d.inner[f].target_block > n_block) {
        err = "block " + std::to_string(b) + ": inner face target out of range";
        return CFX_E_ARG;
      }
      B.inner.push_back(register_inner_face(d.inner[f], dim));
    }
  }
  if (!ghosts) return 0;
  for (auto &B : out) extrapolate_ghost_grid(B, dim);
  inner_ghost_grid(out, blocks, dim);
  for (auto &B : out) corner_ghost_grid(B);
  if (metrics)
    for (auto &B : out) {
      const int rc = compute_metrics(B, err);
      if (rc) return rc;
    }
  return 0;
}

}  // namespace cfx
