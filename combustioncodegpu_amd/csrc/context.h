// context.h — the opaque cfx_ctx of include/cfx.h: device memory, block
// tables, ghost-fill schedules, execution mode, step graph, surrogate state.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cfx.h"
#include "cfx_types.h"
#include "kernels.h"
#include "mesh.h"
#include "mlp.h"

namespace cfx {

enum KClass { K_RHS = 0, K_UPDATE, K_BC, K_PROPS, K_DT, K_OTHER, K_NCLASS };

struct BcOp {
  int type, blk, face;
};

struct TimedLaunch {
  hipEvent_t a, b;
  int cls;
};

// A workgroup prefix over the blocks (nb+1 entries), device and host copies.
struct Table {
  int *dev = nullptr;
  std::vector<int> host;
};

// Ghost-fill ops grouped into levels: the ops of one level touch disjoint
// data and run as one launch; lvl[L]..lvl[L+1] are the ops of level L.
struct Schedule {
  FaceOp *ops = nullptr;
  int *off = nullptr;
  std::vector<int> off_host;
  std::vector<int> lvl;
  int levels() const { return lvl.empty() ? 0 : (int)lvl.size() - 1; }
};

}  // namespace cfx

using namespace cfx;

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
  int n_partial = 0;
  long long n_grid_total = 0;
  BlockView *views_dev = nullptr;
  std::vector<double *> coords;  // per block: the (ngg+1)-layer coordinates built on the device
  Table t_pad, t_int, t_dt, t_last, t_res;     // point-kernel workgroup prefixes
  Table t_tile;                                // LU-SGS sweeps: one workgroup per (j, k) tile column
  int n_tplanes = 0;                           // LU-SGS tile hyperplanes of the largest block
  bool lusgs = false;
  Table t_rhs[3][kNVar3];                      // pencil tiles per direction and variant
  int occ[2][3][kNVar3] = {};                  // resident workgroups per CU [fp32][dir][variant]
  int n_cu = 256;
  int rhs_var[3] = {0, 0, 0}, rhs_persist[3] = {0, 0, 0};
  std::vector<FaceOp> ops;                     // BC ops in application order, then exchange ops
  int n_bc_ops = 0;
  Schedule sched_seq, sched_bc, sched_ex, sched_all;
  // the step's ghost fill: ops with the whole-face outflow conditions that
  // the update kernel applies (their interior rows) replaced by the rest
  std::vector<FaceOp> ops_step;
  std::vector<int> ops_step_src;  // index of the op in `ops` each step op comes from
  Schedule sched_seq_step, sched_all_step;
  int n_fused = 0;
  bool batched = true, use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool mixed = false;
  bool profiling = false;
  // transport surrogate
  int transport = CFX_TRANSPORT_WILKE;
  MlpShape mlp_shape{};
  float *mlp_params = nullptr, *mlp_m = nullptr, *mlp_v = nullptr;
  void *mlp_ops = nullptr;      // training operands
  float *mlp_bias = nullptr;
  void *mlp_ops_inf = nullptr;  // inference operands (folded sigmoid factor)
  float *mlp_bias_inf = nullptr;
  MlpDev mlp_host{};
  MlpDev *mlp_dev = nullptr;
  unsigned long long *mlp_counters = nullptr;
  bool mlp_ready = false;
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
  Batch all(const Table &t) const { return Batch{views_dev, t.dev, (int)views.size(), t.host.back() - t.host[0]}; }
  Batch one(const Table &t, int b) const { return Batch{views_dev + b, t.dev + b, 1, t.host[b + 1] - t.host[b]}; }
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
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

#define CFX_TRY(expr)                  \
  do {                                 \
    const int rc_ = (expr);            \
    if (rc_ != 0) return ctx->fail(rc_, #expr); \
  } while (0)
