// noalloc.hip — surrogate inference that allocates nothing.
//
// cfx_mlp_predict stages its host inputs and outputs through two device
// buffers it allocates on every call.  cfx_mlp_predict_noalloc does the same
// work through `scratch`, a fixed device array of this file, so a caller that
// may not allocate device memory (a serving loop, code that runs while a
// graph is captured elsewhere on the device) can still evaluate the network.
//
// `scratch` holds kNoAllocMaxPoints inputs followed by as many outputs, 16
// floats each (8 MiB).  Each device the library runs on keeps its own copy
// until the process ends; its address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels and copies.
// Every call on a device shares that device's array and it keeps the last
// call's contents: the output region is cleared by a fill kernel on each
// call.  Calls that overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.h"
#include "mlp_kernels.h"

constexpr long long kNoAllocMaxPoints = 65536;
constexpr long long kNoAllocFloats = 2 * kNoAllocMaxPoints * cfx::kMlpIO;

static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mutex;
float *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
int scratch_address(float **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return CFX_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return CFX_E_UNSUPPORTED;
    g_scratch_ptr[dev] = static_cast<float *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ void fill_zero_kernel(float *__restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

}  // namespace

extern "C" int cfx_mlp_predict_noalloc(cfx_ctx *ctx, const float *x, int n, float *y) {
  using namespace cfx;
  if (!ctx || !x || !y || n <= 0) return CFX_E_ARG;
  if ((long long)n > kNoAllocMaxPoints)
    return ctx->fail(CFX_E_ARG, "cfx_mlp_predict_noalloc: more than 65536 points");
  if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  float *base = nullptr;
  if (int rc = scratch_address(&base)) return ctx->fail(rc, "scratch address");
  float *dx = base, *dy = base + kNoAllocMaxPoints * kMlpIO;
  const long long e = (long long)n * kMlpIO;
  hipStream_t st = ctx->stream;
  int rc = (int)hipStreamSynchronize(st);
  if (!rc) rc = (int)hipMemcpyAsync(dx, x, sizeof(float) * e, hipMemcpyHostToDevice, st);
  if (!rc) {
    fill_zero_kernel<<<(int)((e + 255) / 256), 256, 0, st>>>(dy, e);
    rc = (int)hipGetLastError();
  }
  if (!rc) rc = launch_mlp_predict(dx, n, ctx->mlp_ops_inf, ctx->mlp_bias_inf, ctx->mlp_shape.nout, dy, st);
  if (!rc) rc = (int)hipMemcpyAsync(y, dy, sizeof(float) * e, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  return rc ? ctx->fail(rc, "cfx_mlp_predict_noalloc") : 0;
}
