// Colour GMMs on the device: per-pixel scores (negative log-likelihood),
// EM fitting, and cut unaries from the scores of a background and a
// foreground model.
//
// The reference (GMM.pyx, gmm.cl) fits and scores in float32 with its
// mixture packed into a matrix A (per component: log weight + constant,
// mean / var, -1 / (2 var); initA / em2 gmm.cl:33-310) and reduces EM's
// sums in pairwise tree launches, one blocking wait per level
// (GMM.pyx:245-330).  Here everything is float64; a score evaluates the
// mixture in the order the C ABI documents (products and sums rounded one
// at a time: contraction is off), so a host statement in the same order
// gives the same values; an EM iteration is two launches (per-block sums of
// the responsibilities, then one block adds them in block order and updates
// the model on the device), and only a tolerance check reads back.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "gcmi_pool.h"

#pragma clang fp contract(off)

namespace gcmi {

constexpr int kGK = GCMI_GMM_MAX_COMPONENTS;
constexpr double kTwoPi = 6.283185307179586;

// A mixture in the form a score evaluates: e_i(x) = c[i] - sum_j (x_j -
// mu[i][j])^2 * h[i][j].
struct GmmConst {
  int k;
  double c[kGK];
  double mu[kGK][3];
  double h[kGK][3];
};

// The model the EM kernels read and update in device memory.
struct DevGmm {
  int k;
  int pad;
  double w[kGK], mu[kGK][3], v[kGK][3];
  GmmConst c;
  double ll;  // mean log-likelihood of the last E step
};

namespace {

int check_gmm(const gcmi_gmm* m) {
  if (!m) {
    set_error("null GMM");
    return GCMI_EINVAL;
  }
  bool ok = m->k >= 1 && m->k <= kGK;
  double wsum = 0;
  for (int i = 0; ok && i < m->k; ++i) {
    ok = std::isfinite(m->weight[i]) && m->weight[i] >= 0;
    wsum += m->weight[i];
    for (int j = 0; ok && j < 3; ++j)
      ok = std::isfinite(m->mean[i][j]) && std::isfinite(m->var[i][j]) && m->var[i][j] > 0;
  }
  if (!ok || !(wsum > 0)) {
    set_error("GMM out of range (k in 1..%d, weights >= 0 with a positive sum, "
              "finite means, variances > 0)", kGK);
    return GCMI_EINVAL;
  }
  return GCMI_OK;
}

__host__ __device__ inline void gmm_const(int k, const double* w, const double (*mu)[3],
                                          const double (*v)[3], GmmConst* c) {
  c->k = k;
  for (int i = 0; i < kGK; ++i) {
    const bool on = i < k;
    c->c[i] = on ? log(w[i]) - 0.5 * ((log(kTwoPi * v[i][0]) + log(kTwoPi * v[i][1])) +
                                      log(kTwoPi * v[i][2]))
                 : 0.0;
    for (int j = 0; j < 3; ++j) {
      c->mu[i][j] = on ? mu[i][j] : 0.0;
      c->h[i][j] = on ? 0.5 / v[i][j] : 0.0;
    }
  }
}

void host_const(const gcmi_gmm* m, GmmConst* c) { gmm_const(m->k, m->weight, m->mean, m->var, c); }

// The component terms e_i of colour (x0, x1, x2) and their maximum.
__device__ __forceinline__ double gmm_terms(const GmmConst& g, double x0, double x1, double x2,
                                            double (&e)[kGK]) {
  double m = -INFINITY;
#pragma unroll
  for (int i = 0; i < kGK; ++i) {
    if (i < g.k) {
      const double d0 = x0 - g.mu[i][0], d1 = x1 - g.mu[i][1], d2 = x2 - g.mu[i][2];
      const double t0 = (d0 * d0) * g.h[i][0];
      const double t1 = (d1 * d1) * g.h[i][1];
      const double t2 = (d2 * d2) * g.h[i][2];
      e[i] = g.c[i] - ((t0 + t1) + t2);
      m = fmax(m, e[i]);
    }
  }
  return m;
}

// log sum_i exp(e_i) = m + log(sum_i exp(e_i - m)).
__device__ __forceinline__ double gmm_lse(const GmmConst& g, const double (&e)[kGK], double m) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < kGK; ++i)
    if (i < g.k) s += exp(e[i] - m);
  return m + log(s);
}

__device__ __forceinline__ void load_colour(const uint8_t* __restrict__ rgb, int ch, long long i,
                                            double& x0, double& x1, double& x2) {
  const uint8_t* p = rgb + i * ch;
  x0 = p[0];
  x1 = p[1];
  x2 = p[2];
}

__global__ void __launch_bounds__(256)
k_gmm_score(GmmConst g, const uint8_t* __restrict__ rgb, int ch, long long n, float* out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double x0, x1, x2, e[kGK];
    load_colour(rgb, ch, i, x0, x1, x2);
    const double m = gmm_terms(g, x0, x1, x2, e);
    out[i] = (float)(-gmm_lse(g, e, m));
  }
}

__device__ __forceinline__ int32_t quantise(double s, double scale, double cap) {
  return (int32_t)rint(scale * fmin(fmax(s, 0.0), cap));
}

// Unaries of n pixels of one image: cs[i] = q(score under bg), ct[i] =
// q(score under fg).
__global__ void __launch_bounds__(256)
k_gmm_unaries(GmmConst fg, GmmConst bg, const uint8_t* __restrict__ rgb, int ch, long long n,
              double scale, double cap, int32_t* cs, int32_t* ct) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double x0, x1, x2, e[kGK];
    load_colour(rgb, ch, i, x0, x1, x2);
    const double mb = gmm_terms(bg, x0, x1, x2, e);
    cs[i] = quantise(-gmm_lse(bg, e, mb), scale, cap);
    const double mf = gmm_terms(fg, x0, x1, x2, e);
    ct[i] = quantise(-gmm_lse(fg, e, mf), scale, cap);
  }
}

// ------------------------------------------------------------------ EM
// Sums of one E step per block: [0, 8) sum r_i, [8, 32) sum r_i x_j,
// [32, 56) sum r_i x_j^2 (component-major), [56] sum of the samples'
// log-likelihoods.
constexpr int kEmSums = kGK * 7 + 1;
constexpr int kEmNT = 256;
constexpr int kEmBlocksMax = 1024;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__global__ void __launch_bounds__(kEmNT)
k_gmm_estep(const DevGmm* __restrict__ model, const uint8_t* __restrict__ rgb, int ch,
            long long n, double* partial) {
  __shared__ double red[kEmNT / kWave][kEmSums];
  const GmmConst g = model->c;
  double a[kEmSums];
#pragma unroll
  for (int j = 0; j < kEmSums; ++j) a[j] = 0;
  for (long long i = (long long)blockIdx.x * kEmNT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kEmNT) {
    double x0, x1, x2, e[kGK];
    load_colour(rgb, ch, i, x0, x1, x2);
    const double m = gmm_terms(g, x0, x1, x2, e);
    const double lse = gmm_lse(g, e, m);
    a[kEmSums - 1] += lse;
#pragma unroll
    for (int c = 0; c < kGK; ++c) {
      if (c < g.k) {
        const double r = exp(e[c] - lse);
        a[c] += r;
        a[kGK + 3 * c + 0] += r * x0;
        a[kGK + 3 * c + 1] += r * x1;
        a[kGK + 3 * c + 2] += r * x2;
        a[4 * kGK + 3 * c + 0] += r * (x0 * x0);
        a[4 * kGK + 3 * c + 1] += r * (x1 * x1);
        a[4 * kGK + 3 * c + 2] += r * (x2 * x2);
      }
    }
  }
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < kEmSums; ++j) {
    const double v = wave_sum_f64(a[j]);
    if (lane_id() == 0) red[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < kEmSums) {
    double t = 0;
    for (int w = 0; w < kEmNT / kWave; ++w) t += red[w][threadIdx.x];
    partial[(size_t)blockIdx.x * kEmSums + threadIdx.x] = t;
  }
}

// Adds the blocks' sums in block order and updates the model (M step).
__global__ void __launch_bounds__(64)
k_gmm_mstep(const double* __restrict__ partial, int nblocks, long long n, double min_covar,
            DevGmm* model) {
  __shared__ double tot[kEmSums];
  if (threadIdx.x < kEmSums) {
    double t = 0;
    for (int b = 0; b < nblocks; ++b) t += partial[(size_t)b * kEmSums + threadIdx.x];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  DevGmm& M = *model;
  const double nd = (double)n;
  M.ll = tot[kEmSums - 1] / nd;
  for (int c = 0; c < M.k; ++c) {
    const double r = tot[c];
    M.w[c] = r / nd;
    if (r > 0) {
      for (int j = 0; j < 3; ++j) {
        const double mean = tot[kGK + 3 * c + j] / r;
        double var = tot[4 * kGK + 3 * c + j] / r - mean * mean + min_covar;
        if (!(var > 0)) var = min_covar;
        M.mu[c][j] = mean;
        M.v[c][j] = var;
      }
    }
  }
  gmm_const(M.k, M.w, M.mu, M.v, &M.c);
}

int launch_grid(long long n) {
  return (int)std::max(1LL, std::min<long long>((n + 255) / 256, 4096));
}

}  // namespace

int pool_upload_rgb_gmm_host(Pool* p, int img, const uint8_t* h_rgb, int channels,
                             const gcmi_rgb_model* pair, const gcmi_gmm* fg, const gcmi_gmm* bg,
                             double scale, double cap) {
  GCMI_TRY(check_gmm(fg));
  GCMI_TRY(check_gmm(bg));
  if (!(scale >= 0) || !(cap >= 0) || !std::isfinite(scale) || !std::isfinite(cap) ||
      scale * cap > double(1 << 26)) {
    set_error("GMM unary scale and cap must be >= 0 with scale * cap <= 2^26");
    return GCMI_EINVAL;
  }
  if (!pair) {
    set_error("null colour model");
    return GCMI_EINVAL;
  }
  // the pairwise model's own unaries are built and then replaced; they only
  // need to be in range
  gcmi_rgb_model pm = *pair;
  pm.unary_scale = 0;
  pm.unary_div = 1;
  pm.unary_cap = 0;
  GCMI_TRY(check_rgb_model(&pm, nullptr));
  const uint8_t* d = nullptr;
  GCMI_TRY(pool_stage_rgb(p, img, h_rgb, channels, &d));
  GCMI_TRY(pool_upload_rgb(p, img, d, channels, &pm));
  GmmConst cf, cb;
  host_const(fg, &cf);
  host_const(bg, &cb);
  const PoolImage& im = p->imgs[img];
  const long long n = (long long)im.W * im.H;
  hipLaunchKernelGGL(k_gmm_unaries, dim3(launch_grid(n)), dim3(256), 0, p->stream, cf, cb, d,
                     channels, n, scale, cap, p->dev.cs + im.pix0, p->dev.ct + im.pix0);
  GCMI_HIP(hipGetLastError());
  GCMI_HIP(hipStreamSynchronize(p->stream));
  return GCMI_OK;
}

}  // namespace gcmi

using namespace gcmi;

extern "C" {

int gcmi_gmm_score(const gcmi_gmm* model, const uint8_t* d_rgb, int channels, long long n,
                   float* d_out, void* stream) {
  GCMI_TRY(require_device());
  GCMI_TRY(check_gmm(model));
  if (channels != 3 && channels != 4) {
    set_error("a colour image has 3 or 4 channels (got %d)", channels);
    return GCMI_EINVAL;
  }
  if (n < 0 || (n > 0 && (!d_rgb || !d_out))) {
    set_error("bad score arguments");
    return GCMI_EINVAL;
  }
  if (n == 0) return GCMI_OK;
  GmmConst c;
  host_const(model, &c);
  hipLaunchKernelGGL(k_gmm_score, dim3(launch_grid(n)), dim3(256), 0, pick_stream(stream), c,
                     d_rgb, channels, n, d_out);
  GCMI_HIP(hipGetLastError());
  return GCMI_OK;
}

int gcmi_gmm_fit(const uint8_t* d_rgb, int channels, long long n, int k, int iters, double tol,
                 double init_covar, double min_covar, int init_from_model, gcmi_gmm* model_io,
                 double* loglik_out, int* iters_out) {
  GCMI_TRY(require_device());
  if (!model_io || !d_rgb || (channels != 3 && channels != 4) || k < 1 || k > kGK ||
      n < k || iters < 0 || !(min_covar >= 0) || !std::isfinite(min_covar)) {
    set_error("bad GMM fit arguments (k in 1..%d, n >= k samples, 3 or 4 channels)", kGK);
    return GCMI_EINVAL;
  }
  gcmi_gmm m0{};
  if (init_from_model) {
    GCMI_TRY(check_gmm(model_io));
    if (model_io->k != k) {
      set_error("initial model has %d components, not %d", model_io->k, k);
      return GCMI_EINVAL;
    }
    m0 = *model_io;
  } else {
    if (!(init_covar > 0) || !std::isfinite(init_covar)) {
      set_error("init_covar must be > 0");
      return GCMI_EINVAL;
    }
    // initA (gmm.cl:33-72): means at evenly spaced samples
    m0.k = k;
    for (int i = 0; i < k; ++i) {
      const long long at = k > 1 ? (long long)i * (n - 1) / (k - 1) : 0;
      uint8_t px[4] = {0, 0, 0, 0};
      GCMI_HIP(hipMemcpy(px, d_rgb + at * channels, (size_t)channels, hipMemcpyDeviceToHost));
      m0.weight[i] = 1.0 / k;
      for (int j = 0; j < 3; ++j) {
        m0.mean[i][j] = px[j];
        m0.var[i][j] = init_covar;
      }
    }
  }
  DevGmm h{};
  h.k = k;
  memcpy(h.w, m0.weight, sizeof h.w);
  memcpy(h.mu, m0.mean, sizeof h.mu);
  memcpy(h.v, m0.var, sizeof h.v);
  host_const(&m0, &h.c);
  const int nb = (int)std::max(1LL, std::min<long long>((n + kEmNT * 4 - 1) / (kEmNT * 4),
                                                        kEmBlocksMax));
  hipStream_t s = default_stream();
  DevGmm* d_model = nullptr;
  double* d_partial = nullptr;
  int rc = GCMI_OK;
  if (hipMalloc((void**)&d_model, sizeof(DevGmm)) != hipSuccess ||
      hipMalloc((void**)&d_partial, sizeof(double) * kEmSums * nb) != hipSuccess) {
    (void)hipGetLastError();
    set_error("GMM fit allocation failed");
    rc = GCMI_ENOMEM;
  }
  int it = 0;
  double ll = NAN, prev = NAN;
  if (rc == GCMI_OK && hipMemcpyAsync(d_model, &h, sizeof h, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = GCMI_EHIP;
  for (; rc == GCMI_OK && it < iters; ++it) {
    hipLaunchKernelGGL(k_gmm_estep, dim3(nb), dim3(kEmNT), 0, s, (const DevGmm*)d_model, d_rgb,
                       channels, n, d_partial);
    hipLaunchKernelGGL(k_gmm_mstep, dim3(1), dim3(64), 0, s, (const double*)d_partial, nb, n,
                       min_covar, d_model);
    if (hipGetLastError() != hipSuccess) {
      rc = GCMI_EHIP;
      break;
    }
    if (tol > 0) {
      if (hipMemcpyAsync(&ll, &d_model->ll, sizeof ll, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = GCMI_EHIP;
        break;
      }
      if (it > 0 && fabs(ll - prev) < tol) {
        ++it;
        break;
      }
      prev = ll;
    }
  }
  if (rc == GCMI_OK && (hipMemcpyAsync(&h, d_model, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
    rc = GCMI_EHIP;
  if (rc == GCMI_EHIP) {
    (void)hipGetLastError();
    set_error("GMM fit failed on the device");
  }
  if (d_model) (void)hipFree(d_model);
  if (d_partial) (void)hipFree(d_partial);
  if (rc != GCMI_OK) return rc;
  gcmi_gmm out{};
  out.k = k;
  memcpy(out.weight, h.w, sizeof h.w);
  memcpy(out.mean, h.mu, sizeof h.mu);
  memcpy(out.var, h.v, sizeof h.v);
  *model_io = out;
  if (loglik_out) *loglik_out = iters > 0 ? h.ll : NAN;
  if (iters_out) *iters_out = it;
  return GCMI_OK;
}

}  // extern "C"
