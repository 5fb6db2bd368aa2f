// mlp_kernels.h — host launchers of the transport-surrogate kernels that do
// not depend on the species count (mlp_kernels.hip).  All are asynchronous on
// the given stream and return 0 or a hipError_t.
#pragma once

#include <hip/hip_runtime.h>

#include "mlp.h"

namespace cfx {

// waves x 32 samples (indices (first + k + offset) * mult mod n_train) through
// forward and backward; per-wave gradients to partial[w * P], loss to loss[w]
int launch_mlp_train(const float *xs, const float *ts, long long n_train, long long first, long long offset,
                     long long mult, int waves, const void *ops, const float *bias, const MlpShape &shp,
                     float scale, float *partial, float *loss, hipStream_t st);
// sum of n_partial gradients (into grad_out if non-null) and, when lr > 0, an
// Adam step with bias corrections c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
int launch_mlp_adam(float *params, float *m, float *v, const float *partial, int n_partial, int P, float lr,
                    float b1, float b2, float eps, float c1, float c2, float *grad_out, hipStream_t st);
// fold = 0: the training operands; fold = 1: the inference operands (the
// sigmoid's exponent factor folded in, see mlp.h)
int launch_mlp_pack(const float *params, const MlpShape &shp, void *ops, float *bias, int fold, hipStream_t st);
// ops/bias: the inference operands
int launch_mlp_predict(const float *xs, long long n, const void *ops, const float *bias, int nout, float *ys,
                       hipStream_t st);
int launch_mlp_moments(const double *x, long long n, int nf, int cross, int groups, double *out, hipStream_t st);
int launch_mlp_normalize(const double *x, long long n, int nf, const float *mean, const float *inv_std, float *xs,
                         hipStream_t st);

}  // namespace cfx
