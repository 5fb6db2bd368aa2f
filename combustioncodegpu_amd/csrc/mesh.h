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

// Builds every block from the ABI descriptors: registers faces, extrapolates
// the ghost grid (including inter-block ghost grid and corners) and computes
// the metrics.  Returns 0 or a CFX_E_* code with a message in err.
int build_mesh(const cfx_block *blocks, int n_block, std::vector<HostBlock> &out, int &dim, std::string &err);

}  // namespace cfx
