/*
 * Context, buffers, grids and the t-tail tables of libeqbf_hip.
 */
#include <cstdarg>
#include <cstring>

#include "eqbf_internal.h"

int ctx_fail(eqbf_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

void* ctx_buf(eqbf_ctx* ctx, const std::string& name, size_t bytes) {
  DevBuf& b = ctx->bufs[name];
  if (bytes == 0) bytes = 256;
  if (b.bytes < bytes) {
    if (b.ptr) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(b.ptr);
    }
    b.ptr = nullptr;
    b.bytes = 0;
    size_t want = bytes + bytes / 8;
    ++ctx->buf_epoch;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
      b.ptr = nullptr;
      ctx_fail(ctx, EQBF_ERR_HIP, "hipMalloc of %zu bytes for %s failed", want, name.c_str());
      return nullptr;
    }
    b.bytes = want;
  }
  return b.ptr;
}

extern "C" int eqbf_create(eqbf_ctx** out, int device) {
  if (!out) return EQBF_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return EQBF_ERR_HIP;
  eqbf_ctx* ctx = new eqbf_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return EQBF_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return EQBF_ERR_HIP;
  }
  ctx->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return EQBF_ERR_HIP;
  }
  if (hipMalloc(&ctx->d_flag.ptr, 256) != hipSuccess) {
    delete ctx;
    return EQBF_ERR_HIP;
  }
  ctx->d_flag.bytes = 256;
  if (assoc_noalloc_setup(ctx) != EQBF_OK) {
    (void)hipFree(ctx->d_flag.ptr);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EQBF_ERR_HIP;
  }
  *out = ctx;
  return EQBF_OK;
}

extern "C" void eqbf_destroy(eqbf_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto& kv : ctx->bufs)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  if (ctx->d_grid.ptr) (void)hipFree(ctx->d_grid.ptr);
  if (ctx->d_flag.ptr) (void)hipFree(ctx->d_flag.ptr);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* eqbf_last_error(const eqbf_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int eqbf_sync(eqbf_ctx* ctx) {
  if (!ctx) return EQBF_ERR_ARG;
  EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EQBF_OK;
}

extern "C" int eqbf_set_grids(eqbf_ctx* ctx, const double* phiL, const double* omaL, int32_t nL,
                              const double* phiS, const double* omaS, int32_t nS) {
  if (!ctx || nL < 0 || nS < 0 || (nL > 0 && (!phiL || !omaL)) || (nS > 0 && (!phiS || !omaS)))
    return ctx_fail(ctx, EQBF_ERR_ARG, "eqbf_set_grids: invalid grid");
  ctx->phiL.assign(phiL, phiL + nL);
  ctx->omaL.assign(omaL, omaL + nL);
  ctx->phiS.assign(phiS, phiS + nS);
  ctx->omaS.assign(omaS, omaS + nS);
  std::vector<double> packed;
  packed.insert(packed.end(), ctx->phiL.begin(), ctx->phiL.end());
  packed.insert(packed.end(), ctx->omaL.begin(), ctx->omaL.end());
  packed.insert(packed.end(), ctx->phiS.begin(), ctx->phiS.end());
  packed.insert(packed.end(), ctx->omaS.begin(), ctx->omaS.end());
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  if (ctx->d_grid.ptr) {
    EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EQBF_HIP(ctx, hipFree(ctx->d_grid.ptr));
    ctx->d_grid.ptr = nullptr;
  }
  size_t bytes = std::max<size_t>(packed.size(), 1) * sizeof(double);
  ++ctx->buf_epoch; /* graphs and cached Bayes-factor jobs refer to the old grids */
  EQBF_HIP(ctx, hipMalloc(&ctx->d_grid.ptr, bytes));
  ctx->d_grid.bytes = bytes;
  if (!packed.empty())
    EQBF_HIP(ctx, hipMemcpy(ctx->d_grid.ptr, packed.data(), packed.size() * sizeof(double),
                            hipMemcpyHostToDevice));
  return EQBF_OK;
}

/* ---- configurations -------------------------------------------------------- */
extern "C" int64_t eqbf_n_configs(int32_t bfs, int32_t S) {
  if (bfs == EQBF_BFS_GEN) return 0;
  if (bfs == EQBF_BFS_SIN) return S;
  if (S >= 63) return -1;
  return (int64_t)((1ull << S) - 1);
}

extern "C" int eqbf_config_masks(int32_t bfs, int32_t S, uint64_t* masks) {
  if (!masks || S < 1 || S > 64) return EQBF_ERR_ARG;
  int64_t k = 0;
  if (bfs == EQBF_BFS_SIN) {
    for (int s = 0; s < S; ++s) masks[k++] = 1ull << s;
    return EQBF_OK;
  }
  if (bfs != EQBF_BFS_ALL) return EQBF_OK;
  if (S >= 63) return EQBF_ERR_ARG;
  /* subsets of size 1..S, lexicographic within a size (gsl_combination) */
  for (int sz = 1; sz <= S; ++sz) {
    std::vector<int> c(sz);
    for (int i = 0; i < sz; ++i) c[i] = i;
    while (true) {
      uint64_t m = 0;
      for (int i = 0; i < sz; ++i) m |= 1ull << c[i];
      masks[k++] = m;
      int i = sz - 1;
      while (i >= 0 && c[i] == S - sz + i) --i;
      if (i < 0) break;
      ++c[i];
      for (int j = i + 1; j < sz; ++j) c[j] = c[j - 1] + 1;
    }
  }
  return EQBF_OK;
}

extern "C" int64_t eqbf_n_raw(int32_t bfs, int32_t S, int32_t nL, int32_t nS) {
  return 3 * (int64_t)nL + eqbf_n_configs(bfs, S) * nS;
}
extern "C" int64_t eqbf_n_weighted(int32_t bfs, int32_t S) { return 3 + eqbf_n_configs(bfs, S) + 2; }

/* ---- t-tail tables -------------------------------------------------------- */
__device__ static double tz_w_at(double s, double nu) {
  double t = sqrt(nu) * sinh(s);
  return -dm::log_t_lower_neg(t, nu);
}

/* ln(cosh s) for s >= 0 without overflow. */
__device__ static double lncosh(double s) { return s + log1p(exp(-2.0 * s)) - 0.6931471805599453; }

__global__ void __launch_bounds__(256) tz_build_kernel(double* __restrict__ tz, int nu0) {
  const int nu_i = nu0 + blockIdx.x;
  const double nu = (double)nu_i;
  double* row = tz + (size_t)nu_i * TZ_ROW;
  __shared__ double s_hdr[4];
  if (threadIdx.x == 0) {
    /* smallest s with w(s) >= 760 (P below the smallest subnormal), found by
     * doubling then bisection, so the nodes are not spent past the end */
    double hi = 0.5;
    while (hi < 4096.0 && tz_w_at(hi, nu) < 760.0) hi *= 2.0;
    double lo = hi * 0.5;
    for (int it = 0; it < 40; ++it) {
      double mid = 0.5 * (lo + hi);
      if (tz_w_at(mid, nu) < 760.0)
        lo = mid;
      else
        hi = mid;
    }
    double smax = hi;
    double sA = smax < 8.0 ? smax : 8.0;
    s_hdr[0] = sA / (TZ_NODES - 1);
    s_hdr[1] = sA;
    s_hdr[2] = (smax > sA) ? (smax - sA) / (TZ_NODES - 1) : 1.0;
    s_hdr[3] = smax;
    for (int k = 0; k < 4; ++k) row[k] = s_hdr[k];
  }
  __syncthreads();
  const double lc = lgamma(0.5 * (nu + 1.0)) - lgamma(0.5 * nu) - 0.5 * log(3.141592653589793);
  for (int k = threadIdx.x; k < 2 * TZ_NODES; k += blockDim.x) {
    int seg = k / TZ_NODES, j = k % TZ_NODES;
    double s = (seg == 0) ? j * s_hdr[0] : s_hdr[1] + j * s_hdr[2];
    if (seg == 1 && !(s_hdr[3] > s_hdr[1])) s = s_hdr[1];
    double w = tz_w_at(s, nu);
    /* dw/ds = f_nu(t) / P * dt/ds = C_nu sqrt(nu) cosh(s)^(-nu) / P */
    double dw = exp(lc - nu * lncosh(s) + w);
    row[4 + 2 * k] = w;
    row[4 + 2 * k + 1] = dw;
  }
}

int ctx_ensure_tz(eqbf_ctx* ctx, int nu_max, hipStream_t st) {
  if (nu_max <= ctx->tz_nu_max) return EQBF_OK;
  int want = nu_max < 64 ? 64 : nu_max;
  size_t bytes = (size_t)(want + 1) * TZ_ROW * sizeof(double);
  DevBuf& b = ctx->bufs["tz"];
  if (b.bytes < bytes) {
    void* np = nullptr;
    EQBF_HIP(ctx, hipMalloc(&np, bytes));
    if (b.ptr && ctx->tz_nu_max > 0) {
      EQBF_HIP(ctx, hipMemcpyAsync(np, b.ptr, (size_t)(ctx->tz_nu_max + 1) * TZ_ROW * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
      EQBF_HIP(ctx, hipStreamSynchronize(st));
    }
    if (b.ptr) EQBF_HIP(ctx, hipFree(b.ptr));
    b.ptr = np;
    b.bytes = bytes;
  }
  int nu0 = ctx->tz_nu_max + 1;
  hipLaunchKernelGGL(tz_build_kernel, dim3(want - nu0 + 1), dim3(256), 0, st, (double*)b.ptr, nu0);
  EQBF_HIP(ctx, hipGetLastError());
  ctx->tz_nu_max = want;
  return EQBF_OK;
}

double* ctx_tz_table(eqbf_ctx* ctx) { return (double*)ctx->bufs["tz"].ptr; }

/* ---- test hooks: special functions on the device ---------------------------- */
__global__ void tz_probe_kernel(const double* __restrict__ tz, const double* t, const int* nu, int n,
                                double* w_tab, double* w_cf, double* z_tab, double* z_cf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double at = fabs(t[i]);
  double w1 = dm::tz_eval(tz, nu[i], at);
  double w2 = -dm::log_t_lower_neg(at, (double)nu[i]);
  w_tab[i] = w1;
  w_cf[i] = w2;
  z_tab[i] = dm::gauss_pinv_from_w(w1);
  z_cf[i] = dm::gauss_pinv(dm::t_lower_neg(at, (double)nu[i]));
}

/* Evaluates the t-tail (table and direct continued fraction) and the normal
 * quantile on the device for n points (host arrays); used by the tests. */
extern "C" int eqbf_debug_ttail(eqbf_ctx* ctx, const double* t, const int32_t* nu, int32_t n, double* w_tab,
                                double* w_cf, double* z_tab, double* z_cf) {
  if (!ctx || n <= 0) return EQBF_ERR_ARG;
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  int numax = 1;
  for (int i = 0; i < n; ++i) numax = std::max(numax, (int)nu[i]);
  int rc = ctx_ensure_tz(ctx, numax, ctx->stream);
  if (rc) return rc;
  double* d = (double*)ctx_buf(ctx, "dbg", (size_t)n * (5 * sizeof(double) + sizeof(int)));
  if (!d) return EQBF_ERR_HIP;
  int* dnu = (int*)(d + 5 * (size_t)n);
  EQBF_HIP(ctx, hipMemcpyAsync(d, t, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  EQBF_HIP(ctx, hipMemcpyAsync(dnu, nu, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(tz_probe_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx_tz_table(ctx), d,
                     dnu, n, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n);
  EQBF_HIP(ctx, hipGetLastError());
  EQBF_HIP(ctx, hipMemcpyAsync(w_tab, d + n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EQBF_HIP(ctx, hipMemcpyAsync(w_cf, d + 2 * (size_t)n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EQBF_HIP(ctx, hipMemcpyAsync(z_tab, d + 3 * (size_t)n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EQBF_HIP(ctx, hipMemcpyAsync(z_cf, d + 4 * (size_t)n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EQBF_OK;
}

/* ---- live kernel timing ------------------------------------------------------ */
extern "C" int eqbf_profile(eqbf_ctx* ctx, int enable) {
  if (!ctx) return EQBF_ERR_ARG;
  ctx->prof_on = enable != 0;
  return EQBF_OK;
}

extern "C" int eqbf_profile_read(eqbf_ctx* ctx, char* names, double* total_ms, int64_t* launches, int32_t n_max,
                                 int32_t* n_out) {
  if (!ctx) return EQBF_ERR_ARG;
  int32_t k = 0;
  for (auto& kv : ctx->prof) {
    auto& e = kv.second;
    for (auto& pr : e.pending) {
      EQBF_HIP(ctx, hipEventSynchronize(pr.second));
      float ms = 0.f;
      EQBF_HIP(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
      e.total_ms += ms;
      e.n += 1;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    e.pending.clear();
    if (k < n_max) {
      if (names) {
        strncpy(names + 64 * k, kv.first.c_str(), 63);
        names[64 * k + 63] = 0;
      }
      if (total_ms) total_ms[k] = e.total_ms;
      if (launches) launches[k] = e.n;
    }
    ++k;
  }
  if (n_out) *n_out = k;
  return EQBF_OK;
}
