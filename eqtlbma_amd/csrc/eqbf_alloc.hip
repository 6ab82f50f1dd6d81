/*
 * Learned permutation allocator: the per-gene inputs of its network and the
 * targets it is trained on.  No counterpart in the reference.
 *
 *  - LD eigenvalues: the k largest eigenvalues of the correlation matrix of
 *    a gene's cis genotypes (P x P).  They are the k largest eigenvalues of
 *    Z Z^T / N (N x N), Z the N x P standardised genotypes, so the kernels
 *    form that Gram matrix (64 x 64 tiles, FP32) and run a blocked subspace
 *    iteration on it (32 vectors, FP32 products, FP64 Cholesky-QR, then the
 *    Rayleigh-Ritz matrix diagonalised by Jacobi rotations in FP64, p/2
 *    disjoint pairs per round).
 *  - GEV fit: location, scale and shape of each gene's permutation
 *    statistics by probability-weighted moments (Hosking, Wallis & Wood
 *    1985, Technometrics 27:251-261): the draws are sorted in LDS (bitonic,
 *    NaN last and not counted) and the moments b0, b1, b2 summed in FP64.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

constexpr int EIG_NMAX = 512;  /* samples for the eigenvalue kernel (LDS) */
constexpr int EIG_P = 32;      /* subspace vectors                        */
constexpr int EIG_ITERS = 20;
constexpr int EIG_NT = 512;    /* threads of the eigenvalue kernel          */
constexpr int GEV_NMAX = 16384; /* permutations per gene for the GEV fit    */

__device__ inline double al_geno(const eqbf_chunk& c, int64_t i, int64_t j) {
  const int64_t off = i * c.geno_ld + j;
  if (c.geno_dtype == EQBF_F32) return (double)((const float*)c.genotypes)[off];
  return ((const double*)c.genotypes)[off];
}

/* mean and 1/sd of each pair's SNP over the chunk's samples (missing values
 * excluded; a constant SNP gets 1/sd = 0, so it adds nothing) */
__global__ void al_snpstat_kernel(const eqbf_chunk c, const int64_t* __restrict__ bsrc, int64_t nbp,
                                  float2* __restrict__ ms) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbp) return;
  const int64_t snp = c.pair_snp[bsrc[b]];
  double s = 0, s2 = 0;
  int n = 0;
  for (int i = 0; i < c.n_samples; ++i) {
    const double x = al_geno(c, i, snp);
    if (x == x) {
      s += x;
      s2 += x * x;
      ++n;
    }
  }
  const double m = n ? s / n : 0.0;
  const double v = n ? s2 / n - m * m : 0.0;
  ms[b] = make_float2((float)m, v > 1e-12 * (s2 / (n ? n : 1)) ? (float)(1.0 / sqrt(v)) : 0.f);
}

/* K[gene][a][b] = sum_j z[a][j] z[b][j] / N over the gene's pairs j. */
__global__ void __launch_bounds__(256) al_gram_kernel(const eqbf_chunk c, const int64_t* __restrict__ bsrc,
                                                      const int64_t* __restrict__ bpoff, const float2* __restrict__ ms,
                                                      int nt, float* __restrict__ K) {
  __shared__ float Za[16][65], Zb[16][65];
  const int N = c.n_samples;
  const int t = blockIdx.x % (nt * nt), bi = blockIdx.x / (nt * nt);
  const int a0 = (t / nt) * 64, b0 = (t % nt) * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t j0 = bpoff[bi], j1 = bpoff[bi + 1];
  float acc[4][4] = {};
  for (int64_t jj = j0; jj < j1; jj += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int r = e >> 4, kk = e & 15; /* consecutive threads: consecutive SNPs of one sample */
      const int64_t j = jj + kk;
      float za = 0.f, zb = 0.f;
      if (j < j1) {
        const float2 s = ms[j];
        const int64_t snp = c.pair_snp[bsrc[j]];
        if (a0 + r < N) {
          const double x = al_geno(c, a0 + r, snp);
          za = x == x ? ((float)x - s.x) * s.y : 0.f;
        }
        if (b0 + r < N) {
          const double x = al_geno(c, b0 + r, snp);
          zb = x == x ? ((float)x - s.x) * s.y : 0.f;
        }
      }
      Za[kk][r] = za;
      Zb[kk][r] = zb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = Za[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Zb[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  const float invN = 1.f / (float)N;
  float* Kg = K + (int64_t)bi * N * N;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = a0 + ty + 16 * i, b = b0 + tx + 16 * j;
      if (a < N && b < N) Kg[(int64_t)a * N + b] = acc[i][j] * invN;
    }
}

/* k largest eigenvalues of one gene's K (one workgroup per gene). */
__global__ void __launch_bounds__(EIG_NT) al_eig_kernel(const float* __restrict__ Kall, int N, int k,
                                                     float* __restrict__ eig) {
  constexpr int p = EIG_P;
  extern __shared__ __align__(16) unsigned char smem[];
  float* V = (float*)smem;                   /* [N][p] */
  float* W = V + (size_t)N * p;              /* [N][p] */
  double* G = (double*)(W + (size_t)N * p);  /* [p][p] */
  __shared__ int ok;
  const float* K = Kall + (int64_t)blockIdx.x * N * N;
  const int tid = threadIdx.x;
  /* start: a fixed pseudo-random block */
  for (int e = tid; e < N * p; e += EIG_NT) {
    uint32_t h = (uint32_t)e * 2654435761u ^ 0x9e3779b9u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    V[e] = (float)(h & 0xffff) / 65536.f - 0.5f;
  }
  __syncthreads();
  for (int it = 0; it <= EIG_ITERS; ++it) {
    /* W = K V (K symmetric: row i read as column i, coalesced) */
    for (int i = tid; i < N; i += EIG_NT) {
      float acc[p];
#pragma unroll
      for (int c2 = 0; c2 < p; ++c2) acc[c2] = 0.f;
      for (int j0 = 0; j0 < N; j0 += 16) {
        float kv[16]; /* sixteen loads in flight before the products */
#pragma unroll
        for (int u = 0; u < 16; ++u) kv[u] = j0 + u < N ? K[(int64_t)(j0 + u) * N + i] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          if (j0 + u < N) {
            const float4* v4 = (const float4*)(V + (j0 + u) * p);
#pragma unroll
            for (int c4 = 0; c4 < p / 4; ++c4) {
              const float4 v = v4[c4];
              acc[4 * c4] = fmaf(kv[u], v.x, acc[4 * c4]);
              acc[4 * c4 + 1] = fmaf(kv[u], v.y, acc[4 * c4 + 1]);
              acc[4 * c4 + 2] = fmaf(kv[u], v.z, acc[4 * c4 + 2]);
              acc[4 * c4 + 3] = fmaf(kv[u], v.w, acc[4 * c4 + 3]);
            }
          }
        }
      }
#pragma unroll
      for (int c2 = 0; c2 < p; ++c2) W[i * p + c2] = acc[c2];
    }
    __syncthreads();
    /* G = V^T W (Rayleigh-Ritz on the last pass) or W^T W (orthonormalise) */
    const float* X = it == EIG_ITERS ? V : W;
    for (int e = tid; e < p * p; e += EIG_NT) {
      const int a = e / p, b = e - a * p;
      if (b > a && it < EIG_ITERS) continue;
      double s = 0;
      for (int i = 0; i < N; ++i) s += (double)X[i * p + a] * (double)W[i * p + b];
      G[a * p + b] = s;
    }
    __syncthreads();
    if (it == EIG_ITERS) break;
    /* Cholesky G = L L^T, right-looking, the trailing updates in parallel;
     * a pivot below 1e-14 of the first (rank-deficient block) is raised */
    if (tid == 0) ok = G[0] > 0;
    __syncthreads();
    if (!ok) break;
    const double floor_ = 1e-14 * G[0];
    for (int j = 0; j < p; ++j) {
      if (tid == 0) G[j * p + j] = sqrt(fmax(G[j * p + j], floor_));
      __syncthreads();
      const double d = G[j * p + j];
      for (int i = j + 1 + tid; i < p; i += EIG_NT) G[i * p + j] /= d;
      __syncthreads();
      /* G[i][c] -= L[i][j] L[c][j] for j < c <= i */
      const int m = p - 1 - j;
      for (int e = tid; e < m * m; e += EIG_NT) {
        const int i = j + 1 + e / m, c2 = j + 1 + e % m;
        if (c2 <= i) G[i * p + c2] -= G[i * p + j] * G[c2 * p + j];
      }
      __syncthreads();
    }
    /* V = W L^-T (rows are independent: forward substitution per row) */
    for (int i = tid; i < N; i += EIG_NT) {
      float* w = W + i * p;
      float* v = V + i * p;
      for (int c2 = 0; c2 < p; ++c2) {
        double s = w[c2];
        for (int d = 0; d < c2; ++d) s -= (double)v[d] * G[c2 * p + d];
        v[c2] = (float)(s / G[c2 * p + c2]);
      }
    }
    __syncthreads();
  }
  /* Jacobi on the symmetrised p x p Rayleigh-Ritz matrix, parallel
   * (tournament) ordering: each round rotates p/2 disjoint pairs, so their
   * row updates and then their column updates are independent. */
  __shared__ double cs_[EIG_P / 2], sn_[EIG_P / 2];
  __shared__ int pa_[EIG_P / 2], pb_[EIG_P / 2];
  for (int e = tid; e < p * p; e += EIG_NT) {
    const int a = e / p, b = e % p;
    if (b > a) {
      const double m = 0.5 * (G[a * p + b] + G[b * p + a]);
      G[a * p + b] = G[b * p + a] = m;
    }
  }
  __syncthreads();
  for (int sweep = 0; sweep < 12; ++sweep) {
    for (int r = 0; r < p - 1; ++r) {
      if (tid < p / 2) {
        int a, b;
        if (tid == 0) {
          a = r;
          b = p - 1;
        } else {
          a = (r + tid) % (p - 1);
          b = (r - tid + (p - 1)) % (p - 1);
        }
        if (a > b) {
          const int t2 = a;
          a = b;
          b = t2;
        }
        const double apq = G[a * p + b];
        double c2 = 1.0, s2 = 0.0;
        if (fabs(apq) > 1e-300) {
          const double th = (G[b * p + b] - G[a * p + a]) / (2 * apq);
          const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
          c2 = 1 / sqrt(tt * tt + 1);
          s2 = tt * c2;
        }
        pa_[tid] = a;
        pb_[tid] = b;
        cs_[tid] = c2;
        sn_[tid] = s2;
      }
      __syncthreads();
      for (int e = tid; e < (p / 2) * p; e += EIG_NT) { /* columns a, b of every row */
        const int q = e / p, rr = e % p, a = pa_[q], b = pb_[q];
        const double gra = G[rr * p + a], grb = G[rr * p + b];
        G[rr * p + a] = cs_[q] * gra - sn_[q] * grb;
        G[rr * p + b] = sn_[q] * gra + cs_[q] * grb;
      }
      __syncthreads();
      for (int e = tid; e < (p / 2) * p; e += EIG_NT) { /* rows a, b of every column */
        const int q = e / p, rr = e % p, a = pa_[q], b = pb_[q];
        const double gar = G[a * p + rr], gbr = G[b * p + rr];
        G[a * p + rr] = cs_[q] * gar - sn_[q] * gbr;
        G[b * p + rr] = sn_[q] * gar + cs_[q] * gbr;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    /* the k largest diagonal entries, descending (NaN when the iteration failed) */
    float* out = eig + (int64_t)blockIdx.x * k;
    for (int r = 0; r < k; ++r) {
      int best = -1;
      double bv = -INFINITY;
      for (int a = 0; a < p; ++a)
        if (G[a * p + a] > bv) {
          bv = G[a * p + a];
          best = a;
        }
      out[r] = ok ? (float)bv : NAN;
      if (best >= 0) G[best * p + best] = -INFINITY;
    }
  }
}

/* GEV by probability-weighted moments of one gene's statistics. */
__global__ void __launch_bounds__(1024) al_gev_kernel(const double* __restrict__ stats, int64_t nperm, int npow,
                                                      double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* x = (double*)smem;
  __shared__ double red[3][1024];
  __shared__ int nvalid;
  const int tid = threadIdx.x;
  const double* s = stats + (int64_t)blockIdx.x * nperm;
  if (tid == 0) nvalid = 0;
  __syncthreads();
  int cnt = 0;
  for (int i = tid; i < npow; i += 1024) {
    const double v = i < nperm ? s[i] : NAN;
    const bool valid = v == v;
    cnt += valid;
    x[i] = valid ? v : INFINITY;
  }
  atomicAdd(&nvalid, cnt);
  __syncthreads();
  for (int kk = 2; kk <= npow; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < npow; i += 1024) {
        const int l = i ^ j;
        if (l > i) {
          const double a = x[i], b = x[l];
          const bool up = (i & kk) == 0;
          if ((a > b) == up) {
            x[i] = b;
            x[l] = a;
          }
        }
      }
      __syncthreads();
    }
  const int n = nvalid;
  double b0 = 0, b1 = 0, b2 = 0;
  for (int i = tid; i < n; i += 1024) {
    const double v = x[i], j = (double)i;
    b0 += v;
    b1 += j / (n - 1) * v;
    b2 += j * (j - 1) / ((double)(n - 1) * (n - 2)) * v;
  }
  red[0][tid] = b0;
  red[1][tid] = b1;
  red[2][tid] = b2;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o)
      for (int q = 0; q < 3; ++q) red[q][tid] += red[q][tid + o];
    __syncthreads();
  }
  if (tid == 0) {
    double* o3 = out + (int64_t)blockIdx.x * 3;
    if (n < 3) {
      o3[0] = o3[1] = o3[2] = NAN;
      return;
    }
    b0 = red[0][0] / n;
    b1 = red[1][0] / n;
    b2 = red[2][0] / n;
    const double cc = (2 * b1 - b0) / (3 * b2 - b0) - 0.63092975357145743710; /* ln 2 / ln 3 */
    const double k = 7.8590 * cc + 2.9554 * cc * cc;
    double mu, sigma;
    if (fabs(k) < 1e-6) {
      sigma = (2 * b1 - b0) / 0.69314718055994531;
      mu = b0 - 0.57721566490153286 * sigma;
    } else {
      const double g = tgamma(1 + k);
      sigma = (2 * b1 - b0) * k / (g * (1 - exp2(-k)));
      mu = b0 + sigma * (g - 1) / k;
    }
    o3[0] = mu;
    o3[1] = sigma;
    o3[2] = fabs(k) < 1e-6 ? 0.0 : -k;
  }
}

}  // namespace

/* k largest LD eigenvalues of every gene of a chunk whose pointers are
 * device addresses (off: host copy of pair_offset). */
static int ld_eig_run(eqbf_ctx* ctx, const eqbf_chunk& d, const std::vector<int64_t>& off, int32_t n_eig,
                      float* eig) {
  const int N = d.n_samples, G = d.n_genes;
  const hipStream_t st = ctx->stream;
  const int nt = (N + 63) / 64;
  const size_t lds = (size_t)2 * N * EIG_P * sizeof(float) + (size_t)EIG_P * EIG_P * sizeof(double);
  /* batches of genes: Gram matrices of at most ~1 GB */
  const int gb = std::max(1, (int)std::min<int64_t>(G, ((int64_t)1 << 30) / ((int64_t)N * N * 4)));
  float* dK = (float*)ctx_buf(ctx, "al_gram", (size_t)gb * N * N * sizeof(float));
  float* deig = (float*)ctx_buf(ctx, "al_eig", (size_t)gb * n_eig * sizeof(float));
  if (!dK || !deig) return EQBF_ERR_HIP; /* ctx_buf has set the message */
  std::vector<int64_t> src, boff;
  for (int g0 = 0; g0 < G; g0 += gb) {
    const int g1 = std::min(G, g0 + gb), nb = g1 - g0;
    src.clear();
    boff.assign(1, 0);
    for (int g = g0; g < g1; ++g) {
      for (int64_t q = off[g]; q < off[g + 1]; ++q) src.push_back(q);
      boff.push_back((int64_t)src.size());
    }
    const size_t nbp = std::max<size_t>(src.size(), 1);
    float2* dms = (float2*)ctx_buf(ctx, "al_ms", nbp * sizeof(float2));
    int64_t* dsrc = (int64_t*)ctx_buf(ctx, "al_src", nbp * sizeof(int64_t));
    int64_t* doff = (int64_t*)ctx_buf(ctx, "al_off", (nb + 1) * sizeof(int64_t));
    if (!dms || !dsrc || !doff) return EQBF_ERR_HIP;
    if (!src.empty())
      EQBF_HIP(ctx, hipMemcpyAsync(dsrc, src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(doff, boff.data(), boff.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    const int64_t n_src = (int64_t)src.size();
    if (n_src > 0)
      hipLaunchKernelGGL(al_snpstat_kernel, dim3((unsigned)((n_src + 255) / 256)), dim3(256), 0, st, d, dsrc, n_src,
                         dms);
    {
      ProfScope ps(ctx, "al_gram_kernel", st);
      hipLaunchKernelGGL(al_gram_kernel, dim3((unsigned)(nb * nt * nt)), dim3(256), 0, st, d, dsrc, doff, dms, nt, dK);
    }
    {
      ProfScope ps(ctx, "al_eig_kernel", st);
      hipLaunchKernelGGL(al_eig_kernel, dim3((unsigned)nb), dim3(EIG_NT), lds, st, dK, N, n_eig, deig);
    }
    EQBF_HIP(ctx, hipGetLastError());
    EQBF_HIP(ctx, hipMemcpyAsync(eig + (int64_t)g0 * n_eig, deig, (size_t)nb * n_eig * sizeof(float),
                                 hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st)); /* src / boff are refilled for the next batch */
  }
  return EQBF_OK;
}

static int ld_eig_check(eqbf_ctx* ctx, const eqbf_chunk* c, int32_t n_eig, float* eig) {
  if (!ctx) return EQBF_ERR_ARG;
  if (!c || !eig || n_eig < 1 || n_eig > EIG_P - 4)
    return ctx_fail(ctx, EQBF_ERR_ARG, "n_eig must be in 1..%d", EIG_P - 4);
  if (c->n_genes > 0 && (c->n_samples < 2 || c->n_samples > EIG_NMAX))
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "LD eigenvalues support 2..%d samples", EIG_NMAX);
  return EQBF_OK;
}

extern "C" int eqbf_ld_eigenvalues_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, int32_t n_eig, float* eig) {
  int rc = ld_eig_check(ctx, chunk_dev, n_eig, eig);
  if (rc || chunk_dev->n_genes == 0) return rc;
  std::vector<int64_t> off(chunk_dev->n_genes + 1);
  EQBF_HIP(ctx, hipMemcpyAsync(off.data(), chunk_dev->pair_offset, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
  EQBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ld_eig_run(ctx, *chunk_dev, off, n_eig, eig);
}

extern "C" int eqbf_ld_eigenvalues(eqbf_ctx* ctx, const eqbf_chunk* chunk, int32_t n_eig, float* eig) {
  int rc = ld_eig_check(ctx, chunk, n_eig, eig);
  if (rc || chunk->n_genes == 0) return rc;
  const eqbf_chunk& c = *chunk;
  const hipStream_t st = ctx->stream;
  eqbf_chunk d = c;
  const size_t gbytes = (size_t)c.n_samples * c.geno_ld * (c.geno_dtype == EQBF_F32 ? 4 : 8);
  void* g = ctx_buf(ctx, "al_geno", gbytes);
  void* ps = ctx_buf(ctx, "al_psnp", (size_t)std::max<int64_t>(c.n_pairs, 1) * sizeof(int32_t));
  if (!g || !ps) return EQBF_ERR_HIP;
  EQBF_HIP(ctx, hipMemcpyAsync(g, c.genotypes, gbytes, hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(ps, c.pair_snp, (size_t)c.n_pairs * sizeof(int32_t), hipMemcpyHostToDevice, st));
  d.genotypes = g;
  d.pair_snp = (const int32_t*)ps;
  std::vector<int64_t> off(c.pair_offset, c.pair_offset + c.n_genes + 1);
  return ld_eig_run(ctx, d, off, n_eig, eig);
}

extern "C" int eqbf_gev_fit(eqbf_ctx* ctx, const double* stats, int64_t n_genes, int64_t nperm, double* params) {
  if (!ctx || !stats || !params || n_genes < 0 || nperm < 3) return ctx ? ctx_fail(ctx, EQBF_ERR_ARG, "gev") : 1;
  if (nperm > GEV_NMAX) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "GEV fit supports at most %d draws", GEV_NMAX);
  if (n_genes == 0) return EQBF_OK;
  const hipStream_t st = ctx->stream;
  int npow = 1;
  while (npow < nperm) npow <<= 1;
  const int64_t gb = std::min<int64_t>(n_genes, 65536);
  double* ds = (double*)ctx_buf(ctx, "gev_in", (size_t)gb * nperm * sizeof(double));
  double* dp = (double*)ctx_buf(ctx, "gev_out", (size_t)gb * 3 * sizeof(double));
  if (!ds || !dp) return EQBF_ERR_HIP;
  for (int64_t g0 = 0; g0 < n_genes; g0 += gb) {
    const int64_t nb = std::min(gb, n_genes - g0);
    EQBF_HIP(ctx, hipMemcpyAsync(ds, stats + g0 * nperm, (size_t)nb * nperm * sizeof(double), hipMemcpyHostToDevice,
                                 st));
    {
      ProfScope ps(ctx, "al_gev_kernel", st);
      hipLaunchKernelGGL(al_gev_kernel, dim3((unsigned)nb), dim3(1024), (size_t)npow * sizeof(double), st, ds, nperm,
                         npow, dp);
    }
    EQBF_HIP(ctx, hipGetLastError());
    EQBF_HIP(ctx, hipMemcpyAsync(params + g0 * 3, dp, (size_t)nb * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st));
  }
  return EQBF_OK;
}
