// kernels.h — host launchers of the step kernels.  Every kernel works on a
// batch of blocks: a device table of BlockViews and a prefix table off[] of
// workgroups per block, so one launch covers any subset of the blocks (a
// single block is a batch of one).  The species-count templated launchers
// are compiled once per supported species count and precision
// (kernels_ns.hip) and register themselves in g_kernels64 / g_kernels32.
// All launchers return 0 or a hipError_t / CFX_E_* code and are asynchronous
// on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include "cfx_types.h"

namespace cfx {

// ------------------------------------------------------------ work units ---
constexpr int kTpb = 256;        // threads per workgroup of the point kernels
constexpr int kDtIter = 8;       // points per thread of the time-step reduction
constexpr int kLastIter = 4;     // points per thread of the last-step snapshot
constexpr int kResIter = 8;      // points per thread of the residual partials
// LU-SGS sweep tiles: kLuTi x kLuTj x kLuTk cells (i, j, k), one workgroup
// each, whose first wavefront's lanes own the (j, k) columns
constexpr int kLuTi = 4, kLuTj = 8, kLuTk = 8;

// A batch of blocks: views[0..nb) with workgroup prefix off[0..nb] (device
// pointers, off[0] need not be 0) and total = off[nb] - off[0] workgroups.
struct Batch {
  const BlockView *views;
  const int *off;
  int nb;
  int total;
};

// --------------------------------------------------------- RHS tile shapes ---
// Per face direction, the pencil tiles (faces along the direction x two
// tangential extents) and the tile-group width G that the autotuner chooses
// among.  Index v of each list is "variant v" of that direction.
// w32: target waves per SIMD of the FP32 (mixed-precision) kernel; FP64
// kernels target 2.
struct PencilTile {
  int ta, tb, tc, g, w32;
};
constexpr int kNVar3 = 6, kNVar2 = 2;
constexpr PencilTile kTiles3[3][kNVar3] = {
    {{32, 4, 2, 1, 3}, {32, 4, 4, 1, 3}, {64, 2, 2, 1, 3}, {32, 8, 1, 1, 3}, {32, 8, 1, 1, 4}, {32, 4, 2, 1, 4}},
    {{8, 64, 1, 1, 4}, {8, 32, 1, 0, 4}, {8, 32, 1, 1, 4}, {8, 64, 1, 0, 4}, {16, 16, 1, 1, 4}, {16, 8, 2, 1, 4}},
    {{8, 64, 1, 1, 4}, {8, 32, 1, 0, 4}, {8, 32, 1, 1, 4}, {8, 64, 1, 0, 4}, {16, 16, 1, 1, 4}, {16, 8, 2, 1, 4}}};
constexpr PencilTile kTiles2[2][kNVar2] = {{{32, 8, 1, 1, 3}, {64, 4, 1, 1, 3}}, {{16, 16, 1, 1, 3}, {32, 8, 1, 1, 3}}};

inline int rhs_variants(int dim) { return dim == 3 ? kNVar3 : kNVar2; }
inline const PencilTile &rhs_tile_shape(int dim, int dir, int v) { return dim == 3 ? kTiles3[dir][v] : kTiles2[dir][v]; }

// Tiles one pencil launch of shape s covers on a block of m[3] points
// (including the padding tiles of partial G x G groups).
inline int pencil_tiles(const int m[3], int dir, const PencilTile &s) {
  const int a = dir, b = dir == 0 ? 1 : 0, c = dir == 2 ? 1 : 2;
  const int nta = (m[a] + s.ta - 2) / (s.ta - 1);
  const int ntb = (m[b] + s.tb - 1) / s.tb, ntc = (m[c] + s.tc - 1) / s.tc;
  if (s.g == 0) return nta * ntb * ntc;
  return nta * s.g * s.g * ((ntb + s.g - 1) / s.g) * ((ntc + s.g - 1) / s.g);
}

// grid: workgroups launched (a multiple of 8, at most the padded tile count)
using RhsFn = int (*)(const Batch &b, int grid, const SpeciesDev *sp, const SchemeDev &S, int accumulate,
                      hipStream_t st);
using OccFn = int (*)();  // resident workgroups per CU of the kernel
struct RhsEntry {
  RhsFn launch;
  OccFn occupancy;
};

struct KernelSet64 {
  int present;
  int (*props)(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st);
  int (*cv_from_bv)(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st);
  int (*update)(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, int mixed, int stage, const double *dt,
                int *flags, unsigned long long *newton_fail, hipStream_t st);
  // ops[0..nops) with workgroup prefix off[0..nops], total workgroups
  int (*face_ops)(const BlockView *views, const FaceOp *ops, const int *off, int nops, int total,
                  const SpeciesDev *sp, const SchemeDev &S, const InflowDev *inflow, hipStream_t st);
  RhsEntry rhs3[3][kNVar3];
  RhsEntry rhs2[2][kNVar2];
  // surrogate training set: n samples of (T, log p, Y) and log(mu, kappa,
  // rho*D) from the FP64 transport; lim = t_lo, t_hi, log p_lo, log p_hi
  int (*mlp_samples)(const SpeciesDev *sp, const SchemeDev &S, long long n, unsigned long long seed,
                     const double lim[4], double alpha, double *x, double *y, hipStream_t st);
  // LU-SGS: the diagonal and neighbour coefficients over the interior (point
  // batch), and the tiles of one tile hyperplane I + J + K = tplane of the
  // forward (fwd = 1) or backward sweep (batch of one workgroup per (J, K)
  // tile column of each block)
  int (*lusgs_coef)(const Batch &b, const SchemeDev &S, hipStream_t st);
  int (*lusgs_sweep)(const Batch &b, const SchemeDev &S, int tplane, int fwd, hipStream_t st);
};

struct MlpDev;
struct KernelSet32 {
  int present;
  int (*props)(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, hipStream_t st);
  // properties with the surrogate transport (mixed precision)
  int (*props_mlp)(const Batch &b, const SpeciesDev *sp, const SchemeDev &S, const MlpDev *mlp, const void *ops,
                   const float *bias, unsigned long long *counters, hipStream_t st);
  RhsEntry rhs3[3][kNVar3];
  RhsEntry rhs2[2][kNVar2];
};

extern KernelSet64 g_kernels64[kMaxSp + 1];
extern KernelSet32 g_kernels32[kMaxSp + 1];

// ------------------------------------------------ species-independent ones ---
int launch_dt_reset(unsigned long long *dt_bits, hipStream_t st);
int launch_dt(const Batch &b, const SchemeDev &S, int ns, int store_local, unsigned long long *dt_bits,
              hipStream_t st);
int launch_store_last(const Batch &b, hipStream_t st);
// metrics and Jacobian of one block from its coordinates over ngg+1 ghost
// layers (3 planes of (mx+2ngg+2)(my+2ngg+2)(mz+2ngg+2)); FP32 copies when
// metric32 is non-null; *bad counts points with a non-positive Jacobian
int launch_metrics(const double *xyz, int mx, int my, int mz, double *metric, double *jac, float *metric32,
                   float *jac32, int *bad, hipStream_t st);
// The ghost grid of Mesh::init_ghost_grid on the device, over the (ngg+1)-
// layer coordinate array of a block (3 planes of (mx+2ngg+2)(my+2ngg+2)(mz+2ngg+2)):
// raw (3 x mx*my*mz, i fastest) scattered into it; the mirror extrapolation
// (and the 2-D z layers); one inter-block copy; the edges and corners.
struct GhostCopy;
int launch_coord_scatter(const double *raw, int mx, int my, int mz, double *xyz, hipStream_t st);
int launch_ghost_extrapolate(double *xyz, int mx, int my, int mz, int dim, hipStream_t st);
int launch_ghost_inner(const GhostCopy &g, double *dst, int dmx, int dmy, int dmz, const double *src, int smx,
                       int smy, int smz, hipStream_t st);
int launch_ghost_corners(double *xyz, int mx, int my, int mz, hipStream_t st);
// partial sums of workgroup w go to partial[4 * (w)] with w counted from 0 of
// the full prefix, so a per-block launch and a batched one fill the same slots
int launch_residual_partial(const Batch &b, double *partial, hipStream_t st);
int launch_residual_final(const double *partial, int n, double *out, hipStream_t st);
int launch_eliminate_k(const BlockView &B, int ns, hipStream_t st);

}  // namespace cfx
