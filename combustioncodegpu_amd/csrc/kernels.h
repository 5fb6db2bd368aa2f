// kernels.h — host launchers of the step kernels (kernels.hip).  All return 0
// or a hipError_t / CFX_E_* code; all are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include "cfx_types.h"

namespace cfx {

bool ns_supported(int ns);
int rhs_tile(int dim, int tile, int &tx, int &ty, int &tz);
int n_rhs_tiles(int dim);

int launch_props(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, hipStream_t st);
int launch_rhs(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, int tile, hipStream_t st);
int launch_dt_reset(unsigned long long *dt_bits, hipStream_t st);
int launch_dt(const BlockView &B, const SchemeDev &S, int ns, int store_local, unsigned long long *dt_bits,
              hipStream_t st);
int launch_update(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, int stage, const double *dt,
                  int *flags, unsigned long long *newton_fail, hipStream_t st);
int launch_cv_from_bv(const BlockView &B, const SpeciesDev *sp, const SchemeDev &S, int ns, hipStream_t st);
int launch_store_last(const BlockView &B, hipStream_t st);
int residual_partials(const BlockView &B);
int launch_residual_partial(const BlockView &B, double *partial, hipStream_t st);
int launch_residual_final(const double *partial, int n, double *out, hipStream_t st);
int launch_bc(int type, const BlockView &B, const FaceRange &f, const SpeciesDev *sp, const SchemeDev &S, int ns,
              const InflowDev *inflow, const WallDev &wall, hipStream_t st);
int launch_exchange(const BlockView &B, const BlockView &T, const InnerDev &f, int ns, hipStream_t st);
int launch_eliminate_k(const BlockView &B, int ns, hipStream_t st);

}  // namespace cfx
