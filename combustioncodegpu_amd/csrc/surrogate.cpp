// surrogate.cpp — the learned transport surrogate on the host side: training
// set generation, normalisation, training envelope, training on the GPU,
// validation, loading and inspection.  No counterpart in the reference; its
// transport routine (Transport.cu:43-96) produces the samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "context.h"
#include "mlp_kernels.h"

namespace {

// Cyclic Jacobi eigen-decomposition of the symmetric n x n matrix a (n <= 16):
// a = V diag(w) V^T, eigenvectors in the rows of v.
void jacobi_eigen(int n, double a[kMlpIO][kMlpIO], double v[kMlpIO][kMlpIO], double w[kMlpIO]) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[i][j] * a[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        const double th = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  // eigenvectors as rows
          const double vpk = v[p][k], vqk = v[q][k];
          v[p][k] = c * vpk - s * vqk;
          v[q][k] = s * vpk + c * vqk;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = a[i][i];
}

unsigned long long splitmix(unsigned long long &s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// device buffers that live only for one call
struct Scratch {
  std::vector<void *> p;
  template <class T>
  T *get(size_t n, int &rc) {
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
      rc = (int)e;
      return nullptr;
    }
    p.push_back(q);
    return (T *)q;
  }
  ~Scratch() {
    for (void *q : p) (void)hipFree(q);
  }
};

int mlp_alloc(cfx_ctx *ctx) {
  if (ctx->mlp_params) return 0;
  if (ctx->ns < 1 || ctx->ns > kMlpMaxSp) return ctx->fail(CFX_E_UNSUPPORTED, "surrogate needs 1..14 species");
  ctx->mlp_shape = MlpShape{ctx->ns + 2, ctx->ns + 2};
  const int P = ctx->mlp_shape.count();
  int rc = 0;
  ctx->mlp_params = ctx->dalloc<float>(P, rc);
  ctx->mlp_m = ctx->dalloc<float>(P, rc);
  ctx->mlp_v = ctx->dalloc<float>(P, rc);
  ctx->mlp_ops = ctx->dalloc<__bf16>((size_t)kNBlocks * kOpBlock, rc);
  ctx->mlp_bias = ctx->dalloc<float>(4 * kMlpH, rc);
  ctx->mlp_ops_inf = ctx->dalloc<__bf16>((size_t)kNBlocks * kOpBlock, rc);
  ctx->mlp_bias_inf = ctx->dalloc<float>(4 * kMlpH, rc);
  ctx->mlp_dev = ctx->dalloc<MlpDev>(1, rc);
  ctx->mlp_counters = ctx->dalloc<unsigned long long>(3, rc);
  if (rc) return ctx->fail(rc, "hipMalloc surrogate");
  CFX_TRY((int)hipMemsetAsync(ctx->mlp_counters, 0, 3 * sizeof(unsigned long long), ctx->stream));
  std::memset(&ctx->mlp_host, 0, sizeof(MlpDev));
  ctx->mlp_host.nin = ctx->mlp_shape.nin;
  ctx->mlp_host.nout = ctx->mlp_shape.nout;
  return 0;
}

// an envelope that admits every input (parameters loaded without training data)
void open_envelope(MlpDev &m) {
  m.t_lo = -1e30f;
  m.t_hi = 1e30f;
  m.lp_lo = -1e30f;
  m.lp_hi = 1e30f;
  for (int k = 0; k < kMlpIO; ++k) {
    m.pca_inv_var[k] = 0.0f;
    for (int i = 0; i < kMlpIO; ++i) m.pca[k][i] = k == i ? 1.0f : 0.0f;
  }
  m.d2_max = 1e30f;
  m.x_pure = 2.0f;
}

// the envelope distance in the principal axes as one quadratic form
void envelope_quadratic(MlpDev &m) {
  for (int i = 0; i < kMlpIO; ++i)
    for (int j = 0; j < kMlpIO; ++j) {
      double a = 0.0;
      for (int k = 0; k < kMlpIO; ++k) a += (double)m.pca_inv_var[k] * m.pca[k][i] * m.pca[k][j];
      m.quad[i][j] = j < i ? 0.0f : (float)(j == i ? a : 2.0 * a);
    }
}

int upload_mlp(cfx_ctx *ctx) {
  envelope_quadratic(ctx->mlp_host);
  CFX_TRY((int)hipMemcpy(ctx->mlp_dev, &ctx->mlp_host, sizeof(MlpDev), hipMemcpyHostToDevice));
  CFX_TRY(launch_mlp_pack(ctx->mlp_params, ctx->mlp_shape, ctx->mlp_ops, ctx->mlp_bias, 0, ctx->stream));
  CFX_TRY(launch_mlp_pack(ctx->mlp_params, ctx->mlp_shape, ctx->mlp_ops_inf, ctx->mlp_bias_inf, 1, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  return 0;
}

long long gcd(long long a, long long b) { return b ? gcd(b, a % b) : a; }

}  // namespace

extern "C" {

void cfx_mlp_default_config(cfx_mlp_config *c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->n_samples = 2000000;
  c->holdout = 0.02;
  c->epochs = 60;
  c->batch = 4096;
  c->lr = 4e-3;
  c->lr_final = 1e-5;
  c->t_min = 250.0;
  c->t_max = 3000.0;
  c->p_min = 0.5 * 101325.0;
  c->p_max = 2.0 * 101325.0;
  c->dirichlet_alpha = 0.5;
  c->check_tol = 0.05;
  c->check_every = 64;
  c->seed = 20240531ull;
  c->envelope_quantile = 0.99;
}

int cfx_mlp_train(cfx_ctx *ctx, const cfx_mlp_config *cfg_in, cfx_mlp_report *rep) {
  if (!ctx) return CFX_E_ARG;
  cfx_mlp_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else cfx_mlp_default_config(&cfg);
  if (cfg.n_samples < 1024 || cfg.batch < 32 || cfg.batch % 32 || cfg.epochs < 0 || !(cfg.holdout >= 0) ||
      cfg.holdout >= 0.5 || !(cfg.t_min > 0) || !(cfg.t_max > cfg.t_min) || !(cfg.p_min > 0) ||
      !(cfg.p_max >= cfg.p_min) || !(cfg.dirichlet_alpha > 0) || cfg.check_every < 1 ||
      !(cfg.envelope_quantile > 0) || cfg.envelope_quantile > 1)
    return ctx->fail(CFX_E_ARG, "surrogate training configuration out of range");
  if (int rc = mlp_alloc(ctx)) return rc;
  const KernelSet64 &k64 = g_kernels64[ctx->ns];
  if (!k64.mlp_samples) return ctx->fail(CFX_E_UNSUPPORTED, "no sample kernel for this species count");
  const MlpShape shp = ctx->mlp_shape;
  const int nin = shp.nin, nout = shp.nout, P = shp.count();
  const long long N = cfg.n_samples;
  const long long n_val = (long long)(cfg.holdout * (double)N);
  const long long n_train = N - n_val;
  if (n_train < cfg.batch) return ctx->fail(CFX_E_ARG, "fewer training samples than one batch");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  hipEvent_t e0, e1, e2;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventCreate(&e2);
  Scratch sc;
  int rc = 0;
  double *xr = sc.get<double>((size_t)N * nin, rc), *tr = sc.get<double>((size_t)N * nout, rc);
  float *xs = sc.get<float>((size_t)N * kMlpIO, rc), *ts = sc.get<float>((size_t)N * kMlpIO, rc);
  const int groups = 512;
  const int nmx = 2 * nin + nin * nin, nmt = 2 * nout;
  double *mom_x = sc.get<double>((size_t)groups * nmx, rc), *mom_t = sc.get<double>((size_t)groups * nmt, rc);
  float *stats = sc.get<float>(4 * kMlpIO, rc);
  if (rc) return ctx->fail(rc, "hipMalloc training set");
  // ---- samples from the FP64 transport routine
  (void)hipEventRecord(e0, ctx->stream);
  const double lim[4] = {cfg.t_min, cfg.t_max, std::log(cfg.p_min), std::log(cfg.p_max)};
  CFX_TRY(k64.mlp_samples(ctx->sp_dev, ctx->S, N, cfg.seed, lim, cfg.dirichlet_alpha, xr, tr, ctx->stream));
  (void)hipEventRecord(e1, ctx->stream);
  // ---- normalisation from the training part
  CFX_TRY(launch_mlp_moments(xr, n_train, nin, 1, groups, mom_x, ctx->stream));
  CFX_TRY(launch_mlp_moments(tr, n_train, nout, 0, groups, mom_t, ctx->stream));
  std::vector<double> hx((size_t)groups * nmx), ht((size_t)groups * nmt);
  CFX_TRY((int)hipMemcpyAsync(hx.data(), mom_x, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipMemcpyAsync(ht.data(), mom_t, sizeof(double) * ht.size(), hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  std::vector<double> sx(nmx, 0.0), st(nmt, 0.0);
  for (int g = 0; g < groups; ++g) {
    for (int q = 0; q < nmx; ++q) sx[q] += hx[(size_t)g * nmx + q];
    for (int q = 0; q < nmt; ++q) st[q] += ht[(size_t)g * nmt + q];
  }
  MlpDev &m = ctx->mlp_host;
  std::memset(&m, 0, sizeof(m));
  m.nin = nin;
  m.nout = nout;
  double mean_x[kMlpIO] = {0}, isd_x[kMlpIO] = {0};
  const double nt = (double)n_train;
  for (int f = 0; f < nin; ++f) {
    mean_x[f] = sx[f] / nt;
    const double var = std::max(sx[nin + f] / nt - mean_x[f] * mean_x[f], 0.0);
    isd_x[f] = var > 1e-30 ? 1.0 / std::sqrt(var) : 1.0;
    m.x_mean[f] = (float)mean_x[f];
    m.x_inv_std[f] = (float)isd_x[f];
  }
  for (int o = 0; o < nout; ++o) {
    const double mu = st[o] / nt;
    const double var = std::max(st[nout + o] / nt - mu * mu, 0.0);
    m.t_mean[o] = (float)mu;
    m.t_std[o] = (float)(var > 1e-30 ? std::sqrt(var) : 1.0);
  }
  float hstats[4 * kMlpIO];
  std::memcpy(hstats, m.x_mean, sizeof(m.x_mean));
  std::memcpy(hstats + kMlpIO, m.x_inv_std, sizeof(m.x_inv_std));
  std::memcpy(hstats + 2 * kMlpIO, m.t_mean, sizeof(m.t_mean));
  for (int o = 0; o < kMlpIO; ++o) hstats[3 * kMlpIO + o] = o < nout ? 1.0f / m.t_std[o] : 1.0f;
  CFX_TRY((int)hipMemcpyAsync(stats, hstats, sizeof(hstats), hipMemcpyHostToDevice, ctx->stream));
  CFX_TRY(launch_mlp_normalize(xr, N, nin, stats, stats + kMlpIO, xs, ctx->stream));
  CFX_TRY(launch_mlp_normalize(tr, N, nout, stats + 2 * kMlpIO, stats + 3 * kMlpIO, ts, ctx->stream));
  // ---- envelope: box, and principal axes of the normalised inputs
  m.t_lo = (float)cfg.t_min;
  m.t_hi = (float)cfg.t_max;
  m.lp_lo = (float)lim[2];
  m.lp_hi = (float)lim[3];
  double cov[kMlpIO][kMlpIO] = {{0}}, vec[kMlpIO][kMlpIO], lam[kMlpIO];
  for (int i = 0; i < nin; ++i)
    for (int j = 0; j < nin; ++j)
      cov[i][j] = (sx[2 * nin + i * nin + j] / nt - mean_x[i] * mean_x[j]) * isd_x[i] * isd_x[j];
  jacobi_eigen(nin, cov, vec, lam);
  double lmax = 0.0;
  for (int k = 0; k < nin; ++k) lmax = std::max(lmax, lam[k]);
  for (int k = 0; k < nin; ++k) {
    m.pca_inv_var[k] = lam[k] > 1e-9 * lmax ? (float)(1.0 / lam[k]) : 0.0f;
    for (int i = 0; i < nin; ++i) m.pca[k][i] = (float)vec[k][i];
  }
  // squared Mahalanobis distance of a normalised input in the principal axes
  auto dist2 = [&](const float *x) {
    double d2 = 0.0;
    for (int k = 0; k < nin; ++k) {
      double a = 0.0;
      for (int i = 0; i < nin; ++i) a += (double)m.pca[k][i] * x[i];
      d2 += (double)m.pca_inv_var[k] * a * a;
    }
    return d2;
  };
  {
    // the envelope's radius: the envelope_quantile of the training inputs'
    // distances (the network's largest held-out errors sit in the sparse
    // tail of the input distribution, at the dominant species' rho*D)
    std::vector<float> hxs((size_t)n_train * kMlpIO);
    CFX_TRY((int)hipMemcpyAsync(hxs.data(), xs, sizeof(float) * hxs.size(), hipMemcpyDeviceToHost, ctx->stream));
    CFX_TRY((int)hipStreamSynchronize(ctx->stream));
    std::vector<double> d2s((size_t)n_train);
    for (long long s = 0; s < n_train; ++s) d2s[(size_t)s] = dist2(hxs.data() + s * kMlpIO);
    if (cfg.envelope_quantile >= 1.0) {
      m.d2_max = (float)(*std::max_element(d2s.begin(), d2s.end()) * 1.0001);
    } else {
      const size_t r = (size_t)(cfg.envelope_quantile * (double)(n_train - 1));
      std::nth_element(d2s.begin(), d2s.begin() + (long long)r, d2s.end());
      m.d2_max = (float)d2s[r];
    }
  }
  m.x_pure = kMlpPureX;
  m.check_tol = (float)cfg.check_tol;
  m.check_every = cfg.check_every;
  // ---- initial weights: uniform with variance 1/fan_in, zero biases
  {
    std::vector<float> w(P, 0.0f);
    unsigned long long s = cfg.seed ^ 0x5DEECE66Dull;
    auto fill = [&](int off, int rows, int fan_in) {
      const double a = std::sqrt(3.0 / fan_in);
      for (int i = 0; i < rows * fan_in; ++i) w[off + i] = (float)(a * (2.0 * ((splitmix(s) >> 11) * 0x1.0p-53) - 1.0));
    };
    fill(shp.w1(), kMlpH, nin);
    fill(shp.w2(), kMlpH, kMlpH);
    fill(shp.w3(), kMlpH, kMlpH);
    fill(shp.w4(), nout, kMlpH);
    CFX_TRY((int)hipMemcpyAsync(ctx->mlp_params, w.data(), sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream));
    CFX_TRY((int)hipMemsetAsync(ctx->mlp_m, 0, sizeof(float) * P, ctx->stream));
    CFX_TRY((int)hipMemsetAsync(ctx->mlp_v, 0, sizeof(float) * P, ctx->stream));
  }
  // ---- training: Adam, cosine-decayed step size, reshuffled every epoch
  const int waves = cfg.batch / 32;
  float *partial = sc.get<float>((size_t)waves * P, rc), *loss = sc.get<float>(waves, rc);
  if (rc) return ctx->fail(rc, "hipMalloc gradients");
  const long long steps_per_epoch = n_train / cfg.batch;
  const long long total_steps = steps_per_epoch * cfg.epochs;
  const float scale = 1.0f / ((float)cfg.batch * (float)nout);
  const double b1 = 0.9, b2 = 0.999;
  static const long long kPrimes[] = {1000003, 1000033, 1000037, 1000039, 1000081, 1000099, 1000117, 1000121,
                                      1000133, 1000151, 1000159, 1000171, 1000183, 1000187, 1000193, 1000199};
  std::vector<float> hloss(waves);
  double loss_first = -1.0, loss_last = -1.0;
  long long step = 0;
  for (int ep = 0; ep < cfg.epochs; ++ep) {
    long long mult = 1;
    for (int q = 0; q < 16; ++q) {
      const long long cand = kPrimes[(ep + q) % 16] % n_train;
      if (cand > 1 && gcd(cand, n_train) == 1) {
        mult = cand;
        break;
      }
    }
    const long long offset = ((long long)ep * 7919LL) % n_train;
    for (long long b = 0; b < steps_per_epoch; ++b, ++step) {
      const double frac = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 1.0;
      const double lr = cfg.lr_final + 0.5 * (cfg.lr - cfg.lr_final) * (1.0 + std::cos(M_PI * frac));
      const double c1 = 1.0 / (1.0 - std::pow(b1, (double)(step + 1)));
      const double c2 = 1.0 / (1.0 - std::pow(b2, (double)(step + 1)));
      CFX_TRY(launch_mlp_pack(ctx->mlp_params, shp, ctx->mlp_ops, ctx->mlp_bias, 0, ctx->stream));
      CFX_TRY(launch_mlp_train(xs, ts, n_train, b * cfg.batch, offset, mult, waves, ctx->mlp_ops, ctx->mlp_bias,
                               shp, scale, partial, loss, ctx->stream));
      CFX_TRY(launch_mlp_adam(ctx->mlp_params, ctx->mlp_m, ctx->mlp_v, partial, waves, P, (float)lr, (float)b1,
                              (float)b2, 1e-8f, (float)c1, (float)c2, nullptr, ctx->stream));
      if (b == 0 || b == steps_per_epoch - 1) {
        CFX_TRY((int)hipMemcpyAsync(hloss.data(), loss, sizeof(float) * waves, hipMemcpyDeviceToHost, ctx->stream));
        CFX_TRY((int)hipStreamSynchronize(ctx->stream));
        double l = 0.0;
        for (float v : hloss) l += v;
        if (loss_first < 0) loss_first = l;
        loss_last = l;
      }
    }
  }
  CFX_TRY(launch_mlp_pack(ctx->mlp_params, shp, ctx->mlp_ops, ctx->mlp_bias, 0, ctx->stream));
  CFX_TRY(launch_mlp_pack(ctx->mlp_params, shp, ctx->mlp_ops_inf, ctx->mlp_bias_inf, 1, ctx->stream));
  (void)hipEventRecord(e2, ctx->stream);
  // ---- validation on the held-out samples
  double rmse = 0.0, vmax[3] = {0, 0, 0}, vp99[3] = {0, 0, 0}, val_answered = 0.0;
  if (n_val > 0) {
    float *ys = sc.get<float>((size_t)n_val * kMlpIO, rc);
    if (rc) return ctx->fail(rc, "hipMalloc validation");
    CFX_TRY((int)hipMemsetAsync(ys, 0, sizeof(float) * n_val * kMlpIO, ctx->stream));
    CFX_TRY(launch_mlp_predict(xs + n_train * kMlpIO, n_val, ctx->mlp_ops_inf, ctx->mlp_bias_inf, nout, ys,
                               ctx->stream));
    std::vector<float> hy((size_t)n_val * kMlpIO), ht2((size_t)n_val * kMlpIO);
    std::vector<double> hxr((size_t)n_val * nin);  // raw inputs: T, log p, Y_l
    std::vector<float> hxv((size_t)n_val * kMlpIO);  // normalised inputs
    CFX_TRY((int)hipMemcpyAsync(hxr.data(), xr + n_train * nin, sizeof(double) * hxr.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    CFX_TRY((int)hipMemcpyAsync(hxv.data(), xs + n_train * kMlpIO, sizeof(float) * hxv.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
    CFX_TRY((int)hipMemcpyAsync(hy.data(), ys, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, ctx->stream));
    CFX_TRY((int)hipMemcpyAsync(ht2.data(), ts + n_train * kMlpIO, sizeof(float) * ht2.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
    CFX_TRY((int)hipStreamSynchronize(ctx->stream));
    std::vector<double> rel[3];
    long long answered = 0;
    for (long long s = 0; s < n_val; ++s) {
      // statistics over the samples the surrogate answers (inside the
      // envelope, largest mole fraction below x_pure)
      if (!(dist2(hxv.data() + s * kMlpIO) <= (double)m.d2_max)) continue;
      double winv = 0.0, ymax = 0.0;
      for (int l = 0; l < nin - 2; ++l) {
        const double yl = hxr[s * nin + 2 + l] * ctx->sp_host.d.inv_mw[l];
        winv += yl;
        ymax = std::max(ymax, yl);
      }
      if (!(ymax < (double)m.x_pure * winv)) continue;
      ++answered;
      double worst_d = 0.0;
      for (int o = 0; o < nout; ++o) {
        const double d = (double)hy[s * kMlpIO + o] - (double)ht2[s * kMlpIO + o];
        rmse += d * d;
        const double r = std::fabs(std::expm1(d * (double)m.t_std[o]));
        if (o < 2) rel[o].push_back(r);
        else worst_d = std::max(worst_d, r);
      }
      rel[2].push_back(worst_d);
    }
    val_answered = (double)answered / (double)n_val;
    rmse = std::sqrt(rmse / ((double)std::max(answered, 1LL) * nout));
    for (int q = 0; q < 3 && answered > 0; ++q) {
      std::sort(rel[q].begin(), rel[q].end());
      vmax[q] = rel[q].back();
      vp99[q] = rel[q][(size_t)(0.99 * (double)(rel[q].size() - 1))];
    }
  }
  m.ready = 1;
  if (int r = upload_mlp(ctx)) return r;
  CFX_TRY((int)hipMemsetAsync(ctx->mlp_counters, 0, 3 * sizeof(unsigned long long), ctx->stream));
  ctx->mlp_ready = true;
  ctx->drop_graph();
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    float a = 0, b = 0;
    (void)hipEventElapsedTime(&a, e0, e1);
    (void)hipEventElapsedTime(&b, e1, e2);
    rep->samples_ms = a;
    rep->train_ms = b;
    rep->epochs = cfg.epochs;
    rep->steps = (int)total_steps;
    rep->loss_first = loss_first;
    rep->loss_last = loss_last;
    rep->val_rmse = rmse;
    for (int q = 0; q < 3; ++q) rep->val_max_rel[q] = vmax[q], rep->val_p99_rel[q] = vp99[q];
    rep->n_train = n_train;
    rep->n_val = n_val;
    rep->n_params = P;
    rep->val_answered = val_answered;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  return 0;
}

int cfx_mlp_n_params(cfx_ctx *ctx) {
  if (!ctx) return CFX_E_ARG;
  if (ctx->ns < 1 || ctx->ns > kMlpMaxSp) return CFX_E_UNSUPPORTED;
  return MlpShape{ctx->ns + 2, ctx->ns + 2}.count();
}

int cfx_mlp_set_params(cfx_ctx *ctx, const float *params, const float *norm) {
  if (!ctx || !params || !norm) return CFX_E_ARG;
  if (int rc = mlp_alloc(ctx)) return rc;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(ctx->mlp_params, params, sizeof(float) * ctx->mlp_shape.count(), hipMemcpyHostToDevice));
  MlpDev &m = ctx->mlp_host;
  for (int f = 0; f < kMlpIO; ++f) {
    m.x_mean[f] = norm[f];
    m.x_inv_std[f] = norm[kMlpIO + f];
    m.t_mean[f] = norm[2 * kMlpIO + f];
    m.t_std[f] = norm[3 * kMlpIO + f];
  }
  open_envelope(m);
  if (m.check_every < 1) m.check_every = 64, m.check_tol = 0.05f;
  m.ready = 1;
  if (int rc = upload_mlp(ctx)) return rc;
  ctx->mlp_ready = true;
  ctx->drop_graph();
  return 0;
}

int cfx_mlp_get_params(cfx_ctx *ctx, float *params, float *norm) {
  if (!ctx || !params) return CFX_E_ARG;
  if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(params, ctx->mlp_params, sizeof(float) * ctx->mlp_shape.count(), hipMemcpyDeviceToHost));
  if (norm) {
    const MlpDev &m = ctx->mlp_host;
    for (int f = 0; f < kMlpIO; ++f) {
      norm[f] = m.x_mean[f];
      norm[kMlpIO + f] = m.x_inv_std[f];
      norm[2 * kMlpIO + f] = m.t_mean[f];
      norm[3 * kMlpIO + f] = m.t_std[f];
    }
  }
  return 0;
}

int cfx_mlp_set_check(cfx_ctx *ctx, double tol, int every) {
  if (!ctx || !(tol > 0) || every < 1) return CFX_E_ARG;
  if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  ctx->mlp_host.check_tol = (float)tol;
  ctx->mlp_host.check_every = every;
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(ctx->mlp_dev, &ctx->mlp_host, sizeof(MlpDev), hipMemcpyHostToDevice));
  return 0;
}

int cfx_mlp_gradient(cfx_ctx *ctx, const float *x, const float *t, int n, float *grad, double *loss_out) {
  if (!ctx || !x || !t || !grad || n <= 0 || n % 32) return CFX_E_ARG;
  if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  const MlpShape shp = ctx->mlp_shape;
  const int P = shp.count(), waves = n / 32;
  Scratch sc;
  int rc = 0;
  float *dx = sc.get<float>((size_t)n * kMlpIO, rc), *dt = sc.get<float>((size_t)n * kMlpIO, rc);
  float *partial = sc.get<float>((size_t)waves * P, rc), *loss = sc.get<float>(waves, rc);
  float *g = sc.get<float>(P, rc);
  if (rc) return ctx->fail(rc, "hipMalloc gradient check");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(dx, x, sizeof(float) * n * kMlpIO, hipMemcpyHostToDevice));
  CFX_TRY((int)hipMemcpy(dt, t, sizeof(float) * n * kMlpIO, hipMemcpyHostToDevice));
  CFX_TRY(launch_mlp_pack(ctx->mlp_params, shp, ctx->mlp_ops, ctx->mlp_bias, 0, ctx->stream));
  const float scale = 1.0f / ((float)n * (float)shp.nout);
  CFX_TRY(launch_mlp_train(dx, dt, n, 0, 0, 1, waves, ctx->mlp_ops, ctx->mlp_bias, shp, scale, partial, loss,
                           ctx->stream));
  CFX_TRY(launch_mlp_adam(ctx->mlp_params, ctx->mlp_m, ctx->mlp_v, partial, waves, P, 0.0f, 0.9f, 0.999f, 1e-8f, 1.0f,
                          1.0f, g, ctx->stream));
  std::vector<float> hl(waves);
  CFX_TRY((int)hipMemcpyAsync(grad, g, sizeof(float) * P, hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipMemcpyAsync(hl.data(), loss, sizeof(float) * waves, hipMemcpyDeviceToHost, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  if (loss_out) {
    double l = 0.0;
    for (float v : hl) l += v;
    *loss_out = l;
  }
  return 0;
}

int cfx_mlp_predict(cfx_ctx *ctx, const float *x, int n, float *y) {
  if (!ctx || !x || !y || n <= 0) return CFX_E_ARG;
  if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  Scratch sc;
  int rc = 0;
  float *dx = sc.get<float>((size_t)n * kMlpIO, rc), *dy = sc.get<float>((size_t)n * kMlpIO, rc);
  if (rc) return ctx->fail(rc, "hipMalloc predict");
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(dx, x, sizeof(float) * n * kMlpIO, hipMemcpyHostToDevice));
  // on the context's stream: the stream does not synchronise with the null
  // stream, so a plain hipMemset could land after the kernel below
  CFX_TRY((int)hipMemsetAsync(dy, 0, sizeof(float) * n * kMlpIO, ctx->stream));
  CFX_TRY(launch_mlp_predict(dx, n, ctx->mlp_ops_inf, ctx->mlp_bias_inf, ctx->mlp_shape.nout, dy, ctx->stream));
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(y, dy, sizeof(float) * n * kMlpIO, hipMemcpyDeviceToHost));
  return 0;
}

int cfx_set_transport(cfx_ctx *ctx, int transport) {
  if (!ctx) return CFX_E_ARG;
  if (transport == CFX_TRANSPORT_MLP) {
    if (!ctx->mixed) return ctx->fail(CFX_E_UNSUPPORTED, "surrogate transport runs on the mixed-precision path");
    if (!g_kernels32[ctx->ns].props_mlp || ctx->ns < 1 || ctx->ns > kMlpMaxSp)
      return ctx->fail(CFX_E_UNSUPPORTED, "surrogate transport not built for this species count");
    if (!ctx->mlp_ready) return ctx->fail(CFX_E_SURROGATE, "surrogate not trained or loaded");
  } else if (transport != CFX_TRANSPORT_WILKE) {
    return ctx->fail(CFX_E_ARG, "unknown transport model");
  }
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  ctx->transport = transport;
  ctx->drop_graph();
  return 0;
}

int cfx_mlp_stats(cfx_ctx *ctx, long long counts[3]) {
  if (!ctx || !counts) return CFX_E_ARG;
  counts[0] = counts[1] = counts[2] = 0;
  if (!ctx->mlp_counters) return 0;
  unsigned long long h[3];
  CFX_TRY((int)hipStreamSynchronize(ctx->stream));
  CFX_TRY((int)hipMemcpy(h, ctx->mlp_counters, sizeof(h), hipMemcpyDeviceToHost));
  for (int q = 0; q < 3; ++q) counts[q] = (long long)h[q];
  return 0;
}

}  // extern "C"
