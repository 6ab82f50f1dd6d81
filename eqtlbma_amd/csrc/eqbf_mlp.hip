/*
 * The learned permutation allocator's network: a multilayer perceptron
 * n_in -> n_hidden (n_layers times, GELU) -> n_out, trained and evaluated on
 * the GPU by the kernels below.  No counterpart in the reference.
 *
 * Numerics: the weights used by the forward and backward products are BF16
 * (round to nearest even from an FP32 master copy that the optimiser
 * updates); activations, products and sums are FP32.  GELU is the exact
 * x Phi(x).  Gradients with respect to the BF16 weights are applied to the
 * master copy (straight-through).
 *
 * Kernels: one tiled product C = A . B (64 x 64 tile per 256-thread
 * workgroup, 4 x 4 outputs per thread, 16-deep LDS stages) with three
 * epilogues -- forward (bias, pre-activation and GELU stored), backward
 * through the weights (times GELU'(z) of the layer below), and the weight
 * gradient -- plus bias-gradient, loss and Adam kernels.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

struct eqbf_mlp {
  int n_in = 0, n_hidden = 0, n_layers = 0, n_out = 0;
  std::vector<int> din, dout;          /* per linear layer            */
  std::vector<int64_t> woff, boff;     /* offsets in the flat vector  */
  int64_t n_params = 0;
  float* p = nullptr;     /* FP32 master parameters */
  uint16_t* w16 = nullptr;/* BF16 weights (same offsets as p) */
  float* g = nullptr;     /* gradients */
  float* m1 = nullptr;    /* Adam moments */
  float* m2 = nullptr;
  int64_t step = 0;
  /* activations for a batch of at most cap rows */
  int64_t cap = 0;
  std::vector<float*> z, h; /* per layer: pre-activation, output          */
  float* dz0 = nullptr;     /* two gradient buffers of the widest layer    */
  float* dz1 = nullptr;
  float* xin = nullptr;     /* batch inputs, targets, loss scalar          */
  float* tgt = nullptr;
  float* wout = nullptr;
  double* loss = nullptr;
  int64_t* idx = nullptr;
  float* xall = nullptr;    /* training set resident during eqbf_mlp_train */
  float* tall = nullptr;
  int64_t all_cap = 0;
};

namespace {

__device__ inline float bf16_to_f32(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ inline uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40); /* NaN stays NaN */
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);                        /* nearest even   */
  return (uint16_t)(r >> 16);
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

template <typename T>
__device__ inline float ld(const T* p, int64_t i);
template <>
__device__ inline float ld<float>(const float* p, int64_t i) { return p[i]; }
template <>
__device__ inline float ld<uint16_t>(const uint16_t* p, int64_t i) { return bf16_to_f32(p[i]); }

enum { EPI_FWD = 0, EPI_FWD_LINEAR = 1, EPI_BWD_DATA = 2, EPI_STORE = 3 };

/* C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]. */
template <typename TA, typename TB, int EPI>
__global__ void __launch_bounds__(256) mlp_gemm_kernel(const TA* __restrict__ A, int64_t sam, int64_t sak,
                                                       const TB* __restrict__ B, int64_t sbk, int64_t sbn, int M,
                                                       int N, int K, const float* __restrict__ bias,
                                                       float* __restrict__ C, float* __restrict__ C2,
                                                       const float* __restrict__ zprev) {
  __shared__ float As[16][64 + 1];
  __shared__ float Bs[16][64 + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e >> 6, r = e & 63;
      const int m = m0 + r, k = k0 + kk, n = n0 + r;
      As[kk][r] = (m < M && k < K) ? ld<TA>(A, (int64_t)m * sam + (int64_t)k * sak) : 0.f;
      Bs[kk][r] = (n < N && k < K) ? ld<TB>(B, (int64_t)k * sbk + (int64_t)n * sbn) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty + 16 * i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx + 16 * j;
      if (n >= N) continue;
      const int64_t o = (int64_t)m * N + n;
      float v = acc[i][j];
      if (EPI == EPI_FWD || EPI == EPI_FWD_LINEAR) {
        v += bias[n];
        C[o] = v;                                    /* pre-activation */
        if (C2) C2[o] = EPI == EPI_FWD ? gelu(v) : v; /* output         */
      } else if (EPI == EPI_BWD_DATA) {
        C[o] = v * gelu_d(zprev[o]);
      } else {
        C[o] = v;
      }
    }
  }
}

/* db[n] = sum_b dz[b][n] */
__global__ void mlp_bias_grad_kernel(const float* __restrict__ dz, int64_t B, int N, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int64_t b = 0; b < B; ++b) s += dz[b * N + n];
  db[n] = s;
}

/* Weighted squared error over the outputs and its gradient:
 * loss = sum_b sum_j w_j (y - t)^2 / B,  dy = 2 w_j (y - t) / B. */
__global__ void __launch_bounds__(256) mlp_loss_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                                       const float* __restrict__ w, int64_t B, int J,
                                                       float* __restrict__ dy, double* __restrict__ loss) {
  __shared__ double red[256];
  double s = 0.0;
  const float invB = 1.f / (float)B;
  for (int64_t e = threadIdx.x; e < B * J; e += 256) {
    const int j = (int)(e % J);
    const float d = y[e] - t[e];
    s += (double)w[j] * d * d;
    dy[e] = 2.f * w[j] * d * invB;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / (double)B;
}

__global__ void mlp_adam_kernel(float* __restrict__ p, uint16_t* __restrict__ w16, const float* __restrict__ g,
                                float* __restrict__ m1, float* __restrict__ m2, int64_t n, float lr, float b1,
                                float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float v = p[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
  p[i] = v;
  w16[i] = f32_to_bf16(v);
}

__global__ void mlp_to_bf16_kernel(const float* __restrict__ p, uint16_t* __restrict__ w16, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w16[i] = f32_to_bf16(p[i]);
}

/* rows of the minibatch from the resident training set */
__global__ void mlp_gather_kernel(const float* __restrict__ xall, const float* __restrict__ tall, int nin, int nout,
                                  const int64_t* __restrict__ idx, int64_t B, float* __restrict__ x,
                                  float* __restrict__ t) {
  const int64_t b = blockIdx.x;
  if (b >= B) return;
  const int64_t r = idx[b];
  for (int k = threadIdx.x; k < nin; k += blockDim.x) x[b * nin + k] = xall[r * nin + k];
  for (int k = threadIdx.x; k < nout; k += blockDim.x) t[b * nout + k] = tall[r * nout + k];
}

dim3 gemm_grid(int M, int N) { return dim3((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64)); }

void free_mlp_buffers(eqbf_mlp* m) {
  for (float* q : m->z) (void)hipFree(q);
  for (float* q : m->h) (void)hipFree(q);
  m->z.clear();
  m->h.clear();
  void* ps[] = {m->dz0, m->dz1, m->xin, m->tgt, m->wout, m->loss, m->idx};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  m->dz0 = m->dz1 = m->xin = m->tgt = m->wout = nullptr;
  m->loss = nullptr;
  m->idx = nullptr;
  m->cap = 0;
}

int ensure_batch(eqbf_ctx* ctx, eqbf_mlp* m, int64_t B) {
  if (B <= m->cap) return EQBF_OK;
  (void)hipStreamSynchronize(ctx->stream);
  free_mlp_buffers(m);
  const int L = (int)m->din.size();
  int wmax = m->n_in;
  for (int l = 0; l < L; ++l) {
    float *zz = nullptr, *hh = nullptr;
    EQBF_HIP(ctx, hipMalloc(&zz, (size_t)B * m->dout[l] * sizeof(float)));
    EQBF_HIP(ctx, hipMalloc(&hh, (size_t)B * m->dout[l] * sizeof(float)));
    m->z.push_back(zz);
    m->h.push_back(hh);
    wmax = std::max(wmax, m->dout[l]);
  }
  EQBF_HIP(ctx, hipMalloc(&m->dz0, (size_t)B * wmax * sizeof(float)));
  EQBF_HIP(ctx, hipMalloc(&m->dz1, (size_t)B * wmax * sizeof(float)));
  EQBF_HIP(ctx, hipMalloc(&m->xin, (size_t)B * m->n_in * sizeof(float)));
  EQBF_HIP(ctx, hipMalloc(&m->tgt, (size_t)B * m->n_out * sizeof(float)));
  EQBF_HIP(ctx, hipMalloc(&m->wout, 64 * sizeof(float)));
  EQBF_HIP(ctx, hipMalloc(&m->loss, sizeof(double)));
  EQBF_HIP(ctx, hipMalloc(&m->idx, (size_t)B * sizeof(int64_t)));
  m->cap = B;
  return EQBF_OK;
}

/* forward pass of the rows in m->xin; output in m->h[L-1] */
void forward(eqbf_ctx* ctx, eqbf_mlp* m, int64_t B) {
  const hipStream_t st = ctx->stream;
  const int L = (int)m->din.size();
  const float* x = m->xin;
  for (int l = 0; l < L; ++l) {
    const int K = m->din[l], N = m->dout[l];
    const uint16_t* W = m->w16 + m->woff[l]; /* [N][K] */
    const float* b = m->p + m->boff[l];
    ProfScope ps(ctx, "mlp_gemm_kernel", st);
    if (l + 1 < L)
      hipLaunchKernelGGL((mlp_gemm_kernel<float, uint16_t, EPI_FWD>), gemm_grid((int)B, N), dim3(256), 0, st, x,
                         (int64_t)K, (int64_t)1, W, (int64_t)1, (int64_t)K, (int)B, N, K, b, m->z[l], m->h[l],
                         (const float*)nullptr);
    else
      hipLaunchKernelGGL((mlp_gemm_kernel<float, uint16_t, EPI_FWD_LINEAR>), gemm_grid((int)B, N), dim3(256), 0, st,
                         x, (int64_t)K, (int64_t)1, W, (int64_t)1, (int64_t)K, (int)B, N, K, b, m->z[l], m->h[l],
                         (const float*)nullptr);
    x = m->h[l];
  }
}

/* loss of the batch (targets in m->tgt, output weights in m->wout) and the
 * gradients of every parameter into m->g */
void backward(eqbf_ctx* ctx, eqbf_mlp* m, int64_t B) {
  const hipStream_t st = ctx->stream;
  const int L = (int)m->din.size();
  float* dz = m->dz0;
  float* dnext = m->dz1;
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, st, m->h[L - 1], m->tgt, m->wout, B, m->n_out, dz,
                     m->loss);
  for (int l = L - 1; l >= 0; --l) {
    const int K = m->din[l], N = m->dout[l];
    const float* xin = l > 0 ? m->h[l - 1] : m->xin; /* [B][K] */
    {
      ProfScope ps(ctx, "mlp_gemm_kernel", st);
      /* dW[n][k] = sum_b dz[b][n] x[b][k] */
      hipLaunchKernelGGL((mlp_gemm_kernel<float, float, EPI_STORE>), gemm_grid(N, K), dim3(256), 0, st, dz,
                         (int64_t)1, (int64_t)N, xin, (int64_t)K, (int64_t)1, N, K, (int)B, (const float*)nullptr,
                         m->g + m->woff[l], (float*)nullptr, (const float*)nullptr);
      hipLaunchKernelGGL(mlp_bias_grad_kernel, dim3((N + 255) / 256), dim3(256), 0, st, dz, B, N,
                         m->g + m->boff[l]);
      if (l > 0) /* dz_below[b][k] = sum_n dz[b][n] W[n][k] * gelu'(z_below[b][k]) */
        hipLaunchKernelGGL((mlp_gemm_kernel<float, uint16_t, EPI_BWD_DATA>), gemm_grid((int)B, K), dim3(256), 0, st,
                           dz, (int64_t)N, (int64_t)1, m->w16 + m->woff[l], (int64_t)K, (int64_t)1, (int)B, K, N,
                           (const float*)nullptr, dnext, (float*)nullptr, m->z[l - 1]);
    }
    std::swap(dz, dnext);
  }
}

}  // namespace

extern "C" int eqbf_mlp_create(eqbf_ctx* ctx, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                               uint64_t seed, eqbf_mlp** out) {
  if (!ctx || !out || n_in < 1 || n_hidden < 1 || n_layers < 0 || n_out < 1 || n_out > 64)
    return ctx ? ctx_fail(ctx, EQBF_ERR_ARG, "invalid network shape") : EQBF_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  eqbf_mlp* m = new eqbf_mlp();
  m->n_in = n_in;
  m->n_hidden = n_hidden;
  m->n_layers = n_layers;
  m->n_out = n_out;
  int prev = n_in;
  for (int l = 0; l < n_layers; ++l) {
    m->din.push_back(prev);
    m->dout.push_back(n_hidden);
    prev = n_hidden;
  }
  m->din.push_back(prev);
  m->dout.push_back(n_out);
  int64_t off = 0;
  for (size_t l = 0; l < m->din.size(); ++l) {
    m->woff.push_back(off);
    off += (int64_t)m->din[l] * m->dout[l];
    m->boff.push_back(off);
    off += m->dout[l];
  }
  m->n_params = off;
  /* He-scaled normal weights (mt19937 + Box-Muller), zero biases */
  std::vector<float> p(off, 0.f);
  eqbf_rng r;
  eqbf_rng_seed(&r, seed);
  for (size_t l = 0; l < m->din.size(); ++l) {
    const double sd = std::sqrt(2.0 / m->din[l]) * (l + 1 == m->din.size() ? 0.5 : 1.0);
    for (int64_t i = 0; i < (int64_t)m->din[l] * m->dout[l]; ++i) {
      const double u1 = (eqbf_rng_get(&r) + 1.0) / 4294967297.0, u2 = eqbf_rng_get(&r) / 4294967296.0;
      p[m->woff[l] + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  }
  const size_t fb = (size_t)off * sizeof(float);
  if (hipMalloc(&m->p, fb) != hipSuccess || hipMalloc(&m->g, fb) != hipSuccess || hipMalloc(&m->m1, fb) != hipSuccess ||
      hipMalloc(&m->m2, fb) != hipSuccess || hipMalloc(&m->w16, (size_t)off * 2) != hipSuccess) {
    eqbf_mlp_destroy(m);
    return ctx_fail(ctx, EQBF_ERR_HIP, "allocating the network failed");
  }
  *out = m;
  return eqbf_mlp_set_params(ctx, m, p.data());
}

extern "C" void eqbf_mlp_destroy(eqbf_mlp* m) {
  if (!m) return;
  free_mlp_buffers(m);
  void* ps[] = {m->p, m->w16, m->g, m->m1, m->m2, m->xall, m->tall};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  delete m;
}

extern "C" int64_t eqbf_mlp_n_params(const eqbf_mlp* m) { return m ? m->n_params : -1; }

extern "C" int eqbf_mlp_set_params(eqbf_ctx* ctx, eqbf_mlp* m, const float* params) {
  if (!ctx || !m || !params) return EQBF_ERR_ARG;
  const hipStream_t st = ctx->stream;
  const size_t fb = (size_t)m->n_params * sizeof(float);
  EQBF_HIP(ctx, hipMemcpyAsync(m->p, params, fb, hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemsetAsync(m->m1, 0, fb, st));
  EQBF_HIP(ctx, hipMemsetAsync(m->m2, 0, fb, st));
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, st, m->p, m->w16,
                     m->n_params);
  EQBF_HIP(ctx, hipGetLastError());
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  m->step = 0;
  return EQBF_OK;
}

extern "C" int eqbf_mlp_get_params(eqbf_ctx* ctx, eqbf_mlp* m, float* params) {
  if (!ctx || !m || !params) return EQBF_ERR_ARG;
  EQBF_HIP(ctx, hipMemcpyAsync(params, m->p, (size_t)m->n_params * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EQBF_OK;
}

extern "C" int eqbf_mlp_forward(eqbf_ctx* ctx, eqbf_mlp* m, const float* x, int64_t B, float* y) {
  if (!ctx || !m || !x || !y || B < 0) return EQBF_ERR_ARG;
  if (B == 0) return EQBF_OK;
  const int64_t chunk = std::min<int64_t>(B, 1 << 16);
  int rc = ensure_batch(ctx, m, chunk);
  if (rc) return rc;
  const int L = (int)m->din.size();
  for (int64_t b0 = 0; b0 < B; b0 += chunk) {
    const int64_t nb = std::min(chunk, B - b0);
    EQBF_HIP(ctx, hipMemcpyAsync(m->xin, x + b0 * m->n_in, (size_t)nb * m->n_in * sizeof(float),
                                 hipMemcpyHostToDevice, ctx->stream));
    forward(ctx, m, nb);
    EQBF_HIP(ctx, hipGetLastError());
    EQBF_HIP(ctx, hipMemcpyAsync(y + b0 * m->n_out, m->h[L - 1], (size_t)nb * m->n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, ctx->stream));
  }
  EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EQBF_OK;
}

extern "C" int eqbf_mlp_loss_grad(eqbf_ctx* ctx, eqbf_mlp* m, const float* x, const float* target,
                                  const float* out_weight, int64_t B, double* loss, float* grad) {
  if (!ctx || !m || !x || !target || !out_weight || B < 1) return EQBF_ERR_ARG;
  int rc = ensure_batch(ctx, m, B);
  if (rc) return rc;
  const hipStream_t st = ctx->stream;
  EQBF_HIP(ctx, hipMemcpyAsync(m->xin, x, (size_t)B * m->n_in * sizeof(float), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(m->tgt, target, (size_t)B * m->n_out * sizeof(float), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(m->wout, out_weight, (size_t)m->n_out * sizeof(float), hipMemcpyHostToDevice, st));
  forward(ctx, m, B);
  backward(ctx, m, B);
  EQBF_HIP(ctx, hipGetLastError());
  double l = 0;
  EQBF_HIP(ctx, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, st));
  if (grad) EQBF_HIP(ctx, hipMemcpyAsync(grad, m->g, (size_t)m->n_params * sizeof(float), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  if (loss) *loss = l;
  return EQBF_OK;
}

extern "C" int eqbf_mlp_train(eqbf_ctx* ctx, eqbf_mlp* m, const float* x, const float* target,
                              const float* out_weight, int64_t n_rows, int32_t epochs, int32_t batch, float lr,
                              uint64_t seed, float* epoch_loss) {
  if (!ctx || !m || !x || !target || !out_weight || n_rows < 1 || epochs < 0 || batch < 1) return EQBF_ERR_ARG;
  const hipStream_t st = ctx->stream;
  const int64_t B = std::min<int64_t>(batch, n_rows);
  int rc = ensure_batch(ctx, m, B);
  if (rc) return rc;
  if (m->all_cap < n_rows) {
    if (m->xall) (void)hipFree(m->xall);
    if (m->tall) (void)hipFree(m->tall);
    m->xall = m->tall = nullptr;
    m->all_cap = 0;
    EQBF_HIP(ctx, hipMalloc(&m->xall, (size_t)n_rows * m->n_in * sizeof(float)));
    EQBF_HIP(ctx, hipMalloc(&m->tall, (size_t)n_rows * m->n_out * sizeof(float)));
    m->all_cap = n_rows;
  }
  EQBF_HIP(ctx, hipMemcpyAsync(m->xall, x, (size_t)n_rows * m->n_in * sizeof(float), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(m->tall, target, (size_t)n_rows * m->n_out * sizeof(float), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(m->wout, out_weight, (size_t)m->n_out * sizeof(float), hipMemcpyHostToDevice, st));
  eqbf_rng r;
  eqbf_rng_seed(&r, seed);
  std::vector<int64_t> order(n_rows);
  for (int64_t i = 0; i < n_rows; ++i) order[i] = i;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    /* one shuffle per epoch (Fisher-Yates with the library's generator) */
    for (int64_t i = n_rows - 1; i > 0; --i) std::swap(order[i], order[eqbf_rng_uniform_int(&r, (uint32_t)(i + 1))]);
    double sum = 0;
    int64_t nb = 0;
    for (int64_t b0 = 0; b0 + B <= n_rows; b0 += B) {
      EQBF_HIP(ctx, hipMemcpyAsync(m->idx, order.data() + b0, (size_t)B * sizeof(int64_t), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(mlp_gather_kernel, dim3((unsigned)B), dim3(64), 0, st, m->xall, m->tall, m->n_in, m->n_out,
                         m->idx, B, m->xin, m->tgt);
      forward(ctx, m, B);
      backward(ctx, m, B);
      ++m->step;
      const float c1 = 1.f / (1.f - std::pow(b1, (float)m->step)), c2 = 1.f / (1.f - std::pow(b2, (float)m->step));
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, st, m->p, m->w16,
                         m->g, m->m1, m->m2, m->n_params, lr, b1, b2, eps, c1, c2);
      EQBF_HIP(ctx, hipGetLastError());
      double l = 0;
      EQBF_HIP(ctx, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, st));
      /* the index upload of the next step reuses m->idx: wait for this one */
      EQBF_HIP(ctx, hipStreamSynchronize(st));
      sum += l;
      ++nb;
    }
    if (epoch_loss) epoch_loss[ep] = nb ? (float)(sum / nb) : 0.f;
  }
  return EQBF_OK;
}
