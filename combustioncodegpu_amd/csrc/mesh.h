// mesh.h — host-side block geometry: ghost grid, Jacobian and metrics, and
// registration of boundary and inner faces.  Restates the reference's Mesh
// (Mesh.h:9-232, Mesh.cpp:14-143, 212-488, 718-922) as plain arrays.
#pragma once

#include <string>
#include <vector>

#include "cfx_types.h"

namespace cfx {

struct HostBlock {
  int mx = 0, my = 0, mz = 0;
  int px = 0, py = 0, pz = 0;
  long long np = 0, ni = 0;
  // coordinates over ngg+1 ghost layers (Mesh.cpp:220-222)
  int gx = 0, gy = 0, gz = 0;
  std::vector<double> xyz[3];
  // metric (9 x np) and Jacobian (np) over the ngg-padded range
  std::vector<double> metric, jac;
  std::vector<FaceRange> bc_range;
  std::vector<int> bc_type, bc_cond;
  std::vector<InnerDev> inner;

  long long pidx(int i, int j, int k) const {
    return ((long long)(k + kNgg) * py + (j + kNgg)) * px + (i + kNgg);
  }
  long long gidx(int i, int j, int k) const {
    const int g = kNgg + 1;
    return ((long long)(k + g) * gy + (j + g)) * gx + (i + g);
  }
  double &coord(int c, int i, int j, int k) { return xyz[c][gidx(i, j, k)]; }
};

// Jacobian J = det(d(x,y,z)/d(xi,eta,zeta)) and metric grad(xi), grad(eta),
// grad(zeta) (row-major 3x3) at one point from central differences of the
// vertex coordinates (Block::compute_jac_metric, Mesh.cpp:226-488).
// X(c, di, dj, dk) returns coordinate c at the offset point.  Shared by the
// device kernel and the host path; contraction is off so that both round
// every product and sum the same way.  Returns false for J <= 0.
template <class C>
__host__ __device__ inline bool metric_point(const C &X, double &J, double m[9]) {
#pragma clang fp contract(off)
  double d[3][3];  // d[axis][coord]: d(coord)/d(axis)
  for (int cc = 0; cc < 3; ++cc) {
    d[0][cc] = (X(cc, 1, 0, 0) - X(cc, -1, 0, 0)) * 0.5;
    d[1][cc] = (X(cc, 0, 1, 0) - X(cc, 0, -1, 0)) * 0.5;
    d[2][cc] = (X(cc, 0, 0, 1) - X(cc, 0, 0, -1)) * 0.5;
  }
  const double xa = d[0][0], ya = d[0][1], za = d[0][2];
  const double xb = d[1][0], yb = d[1][1], zb = d[1][2];
  const double xc = d[2][0], yc = d[2][1], zc = d[2][2];
  J = xa * yb * zc + xb * yc * za + xc * ya * zb - xa * yc * zb - xb * ya * zc - xc * yb * za;
  if (!(J > 0)) return false;
  // grad(xi) = (dr/deta x dr/dzeta)/J and cyclic
  m[0] = (yb * zc - yc * zb) / J;
  m[1] = (xc * zb - xb * zc) / J;
  m[2] = (xb * yc - xc * yb) / J;
  m[3] = (yc * za - ya * zc) / J;
  m[4] = (xa * zc - xc * za) / J;
  m[5] = (xc * ya - xa * yc) / J;
  m[6] = (ya * zb - yb * za) / J;
  m[7] = (xb * za - xa * zb) / J;
  m[8] = (xa * yb - xb * ya) / J;
  return true;
}

// One inter-block ghost-grid copy (Mesh.cpp:874-922): block blk's points
// lo + step * d (d in [0, n)) take the coordinates of block src's points
// t0 + tstep * d[src_tar], in the reference's block and face order.
struct GhostCopy {
  int blk, src;
  int lo[3], n[3], step[3], t0[3], tstep[3], src_tar[3];
};

// Builds every block from the ABI descriptors: registers faces and, if
// ghosts is set, fills the (ngg+1)-layer coordinate arrays: interior
// vertices, the extrapolated ghost grid, the inter-block ghost grid and the
// corners; if metrics is set too, computes the metrics on the host.  With
// ghosts off the coordinate arrays stay empty (the device builds them).
// Returns 0 or a CFX_E_* code with a message in err.
int build_mesh(const cfx_block *blocks, int n_block, std::vector<HostBlock> &out, int &dim, std::string &err,
               bool metrics = true, bool ghosts = true);
// The inter-block ghost-grid copies of registered blocks, in order.
std::vector<GhostCopy> inner_ghost_copies(const std::vector<HostBlock> &blocks, const cfx_block *desc, int dim);
// Host metrics of one block, with the reference's fallback for ghost points
// whose Jacobian is not positive (copy from the neighbour towards the block).
int compute_metrics(HostBlock &B, std::string &err);

}  // namespace cfx
