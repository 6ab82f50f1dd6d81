/*
 * eqtlbma_avg_bfs on the device: raw log10 ABFs averaged over the grid and
 * over the configurations or types, gene and SNP Bayes factors and the
 * posteriors (src/eqtlbma_avg_bfs.cpp:371-796).
 *
 * Layout: raw [n_pairs][R][L] in FP64, R = dim rows (configs model: the
 * j-th configuration row of the SNP) or n_subgroups rows (types model:
 * subgroup-specific rows); a gene's pairs are contiguous (gene_offset).
 * Every sum is utils::log10_weighted_sum (utils_math.cpp:100-159): the
 * maximum starts from the first value (a NaN first value makes the result
 * NaN), NaN terms are skipped, |result| <= DBL_EPSILON gives 0.
 *
 * Kernels:
 *   avg_grid_kernel  one thread per raw row: the row averaged over the kept
 *                    grid points (BayesFactor::avg_raw_bfs, :371-382)
 *   avg_pair_kernel  one wave per pair (four waves a block striding over
 *                    the pairs), lanes over the dimensions: the
 *                    per-dimension BFs (Snp::avg_raw_bfs, :418-475), the
 *                    SNP BF over them and the first strict maximum; then,
 *                    when asked, post_dim and its argmax (:683-718) and the
 *                    configs model's subgroup posteriors (:722-745)
 *   avg_gene_kernel  one wave per gene, lanes over its SNPs: the gene BF
 *                    (Gene::avg_raw_bfs, :598-622), gene_post (:634-643),
 *                    post_the (:647-662), post_an (:664-680)
 *   avg_post_kernel  the types model's subgroup posteriors (:749-796)
 * The roofline is HBM: each raw value is read once by avg_pair_kernel (and
 * again by avg_post_kernel for the types model's subgroup posteriors).
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

constexpr double kEps = 2.220446049250313e-16;

__device__ inline double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double l10_round(double r) { return fabs(r) <= kEps ? 0.0 : r; }

struct AvgArgs {
  const double* raw;
  const int64_t* off;
  const int32_t* keep;
  const double* gw;   /* [L] grid weights                          */
  const double* dw;   /* [D] configuration or type weights         */
  const uint64_t* cs; /* [D] subgroups of each configuration        */
  const double* sw;   /* [D][S] subgroup weights per type           */
  double* gene_bf;
  double* gene_post;
  double* snp_bf;
  double* dim_bf;
  double* post_the;
  double* post_an;
  double* post_dim;
  double* post_sub;
  int32_t* best_dim;
  double pi0;
  int64_t NP, G;
  int D, S, L, nk, types, want;
};

/* sum over the subgroups of log10(q 10^x + 1 - q) at kept grid point i,
 * skipping subgroup `skip` (types model, :464-466, :778-782) */
__device__ inline double types_point(const AvgArgs& a, int64_t p, int k, int i, int skip) {
  const double* r = a.raw + (size_t)p * a.S * a.L + a.keep[i];
  double v = 0.0;
  for (int s = 0; s < a.S; ++s) {
    if (s == skip) continue;
    const double q = a.sw[(size_t)k * a.S + s];
    v += log10(q * exp10(r[(size_t)s * a.L]) + 1 - q);
  }
  return v;
}

/* log10 sum_i w_i 10^v_i over the kept grid points, values from f(i) */
template <class F>
__device__ inline double grid_lws(const AvgArgs& a, F f) {
  double mx = f(0);
  for (int i = 1; i < a.nk; ++i) {
    const double v = f(i);
    if (v > mx) mx = v;
  }
  double sum = 0.0;
  for (int i = 0; i < a.nk; ++i) {
    const double v = f(i);
    if (!isnan(v)) sum += a.gw[a.keep[i]] * exp10(v - mx);
  }
  return l10_round(mx + log10(sum));
}

__global__ void __launch_bounds__(256) avg_grid_kernel(const double* __restrict__ raw, int64_t n_rows, int L,
                                                       const double* __restrict__ gw, const int32_t* __restrict__ keep,
                                                       int nk, double* __restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const double* v = raw + r * L;
  double mx = v[keep[0]];
  for (int i = 1; i < nk; ++i)
    if (v[keep[i]] > mx) mx = v[keep[i]];
  double sum = 0.0;
  for (int i = 0; i < nk; ++i) {
    const double x = v[keep[i]];
    if (!isnan(x)) sum += gw[keep[i]] * exp10(x - mx);
  }
  out[r] = l10_round(mx + log10(sum));
}

/* reductions over the aligned group of w lanes (w a power of two <= 64) */
__device__ inline double gmax(double v, int w) {
  for (int o = w >> 1; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double gsum(double v, int w) {
  for (int o = w >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* first index of the strict maximum scanning from index 0 (the reference's
 * "j == 0 || v[j] > v[best]"): v0 is the first value, (lmax, larg) the
 * lane's largest value and its first index (larg = n when it has none) */
__device__ inline int first_max(double v0, double lmax, int larg, int n, int w) {
  const double mx = gmax(lmax, w);
  int cand = (larg < n && lmax == mx) ? larg : n;
  for (int o = w >> 1; o > 0; o >>= 1) cand = min(cand, __shfl_xor(cand, o, 64));
  if (isnan(v0) || !(mx > v0)) return 0;
  return cand >= n ? 0 : cand;
}

constexpr int kAvgWaves = 4;

/* A group of w lanes per pair (w = the power of two >= dim, at most 64:
 * 64 / w pairs a wave), four waves a block striding over the pairs; lanes
 * over the dimensions.  Also writes post_dim, its best dimension and the
 * configs model's subgroup posteriors when asked: they need only the pair's
 * own values. */
__global__ void __launch_bounds__(64 * kAvgWaves) avg_pair_kernel(const AvgArgs a, int w) {
  const int lane = threadIdx.x & 63;
  const int gl = lane & (w - 1), g0 = lane - gl; /* lane in the group, group's first lane */
  const int per_wave = 64 / w;
  const int D = a.D;
  const bool dims_post = (a.want & 16) || (!a.types && (a.want & 8));
  const bool sub_post = !a.types && (a.want & 8);
  const int64_t step = (int64_t)gridDim.x * kAvgWaves * per_wave;
  for (int64_t pb = ((int64_t)blockIdx.x * kAvgWaves + (threadIdx.x >> 6)) * per_wave; pb < a.NP; pb += step) {
    const int64_t p = pb + lane / w;
    const bool on = p < a.NP; /* every lane still takes part in the shuffles */
    double* dims = a.dim_bf + (size_t)p * D;
    /* per-dimension BFs; the lane's largest value and its first index */
    double lmax = -INFINITY, dk = NAN;
    int larg = D;
    for (int k = gl; on && k < D; k += w) {
      double d;
      if (!a.types) {
        const double* r = a.raw + ((size_t)p * D + k) * a.L;
        d = grid_lws(a, [&](int i) { return r[a.keep[i]]; });
      } else {
        d = grid_lws(a, [&](int i) { return types_point(a, p, k, i, -1); });
      }
      dims[k] = d;
      if (k == gl) dk = d;
      if (d > lmax) {
        lmax = d;
        larg = k;
      }
    }
    const double d0 = __shfl(dk, g0, 64); /* the group's first lane holds dims[0] */
    /* maximum as the reference scans it: from the first value, strict > */
    double mx = gmax(lmax, w);
    if (isnan(d0)) mx = NAN;
    if (!(mx > d0)) mx = d0;
    const int best0 = first_max(d0, lmax, larg, D, w);
    double sum = 0.0;
    for (int k = gl; on && k < D; k += w) {
      const double d = (k == gl) ? dk : dims[k];
      if (!isnan(d)) sum += a.dw[k] * exp10(d - mx);
    }
    sum = gsum(sum, w);
    const double sbp = l10_round(mx + log10(sum));
    int best = a.types ? -1 : best0;
    if (dims_post) {
      double* pd = a.post_dim + (size_t)p * D;
      double pmax = -INFINITY, pk = NAN;
      int parg = D;
      for (int k = gl; on && k < D; k += w) {
        const double d = (k == gl) ? dk : dims[k];
        double v = exp10(log10(a.dw[k]) + d - sbp);
        if (v > 1.0) v = 1.0;
        pd[k] = v;
        if (k == gl) pk = v;
        if (v > pmax) {
          pmax = v;
          parg = k;
        }
      }
      best = first_max(__shfl(pk, g0, 64), pmax, parg, D, w);
      if (sub_post) {
        for (int s = 0; s < a.S; ++s) {
          double acc = 0.0;
          for (int k = gl; on && k < D; k += w)
            if ((a.cs[k] >> s) & 1ull) acc += (k == gl) ? pk : pd[k];
          acc = gsum(acc, w);
          if (on && gl == 0) a.post_sub[(size_t)p * a.S + s] = acc > 1.0 ? 1.0 : acc;
        }
      }
    }
    if (on && gl == 0) {
      a.snp_bf[p] = sbp;
      a.best_dim[p] = best;
    }
  }
}

__global__ void __launch_bounds__(64) avg_gene_kernel(const AvgArgs a) {
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t p0 = a.off[g], m = a.off[g + 1] - p0;
  if (m <= 0) {
    if (lane == 0) a.gene_bf[g] = NAN;
    return;
  }
  const double* sb = a.snp_bf + p0;
  const double x0 = sb[0];
  double lm = -INFINITY;
  for (int64_t q = lane; q < m; q += 64) lm = fmax(lm, sb[q]);
  double mx = wmax(lm);
  if (isnan(x0)) mx = NAN;
  if (!(mx > x0)) mx = x0;
  const double w = 1 / ((double)m);
  double s1 = 0.0, s2 = 0.0;
  for (int64_t q = lane; q < m; q += 64) {
    const double x = sb[q];
    if (!isnan(x)) {
      const double e = exp10(x - mx);
      s1 += w * e;
      s2 += e;
    }
  }
  s1 = wsum(s1);
  s2 = wsum(s2);
  const double gbf = l10_round(mx + log10(s1));
  if (lane == 0) {
    a.gene_bf[g] = gbf;
    if (a.want & 1) {
      /* lws of (0, gbf) with weights (pi0, 1 - pi0) */
      double om = 0.0;
      if (gbf > om) om = gbf;
      double s = a.pi0 * exp10(0.0 - om);
      if (!isnan(gbf)) s += (1 - a.pi0) * exp10(gbf - om);
      const double ol = l10_round(om + log10(s));
      const double pp = exp10(log10(1.0 - a.pi0) + gbf - ol);
      a.gene_post[g] = pp > 1.0 ? 1.0 : pp;
    }
  }
  if (a.want & 2) {
    const double tot = l10_round(mx + log10(s2));
    for (int64_t q = lane; q < m; q += 64) {
      const double pt = exp10(sb[q] - tot);
      a.post_the[p0 + q] = pt > 1.0 ? 1.0 : pt;
    }
  }
  if (a.want & 4) {
    const double w0 = 1.0 - 1 / (double)m, w1 = 1 / (double)m;
    for (int64_t q = lane; q < m; q += 64) {
      const double x = sb[q];
      double om = 0.0;
      if (x > om) om = x;
      double s = w0 * exp10(0.0 - om);
      if (!isnan(x)) s += w1 * exp10(x - om);
      a.post_an[p0 + q] = exp10(-log10((double)m) + x - l10_round(om + log10(s)));
    }
  }
}

/* types model's subgroup posteriors (:749-796): one wave per pair
 * (grid-stride), lanes over the subgroups */
__global__ void __launch_bounds__(64 * kAvgWaves) avg_post_kernel(const AvgArgs a) {
  const int lane = threadIdx.x & 63;
  const int D = a.D, S = a.S;
  for (int64_t p = (int64_t)blockIdx.x * kAvgWaves + (threadIdx.x >> 6); p < a.NP;
       p += (int64_t)gridDim.x * kAvgWaves) {
    const double sbp = a.snp_bf[p];
    for (int s = lane; s < S; s += 64) {
      double acc = 0.0;
      for (int k = 0; k < D; ++k) {
        double t = log10(a.dw[k]) + log10(a.sw[(size_t)k * S + s]) - sbp;
        const double* rs = a.raw + ((size_t)p * S + s) * a.L;
        t += grid_lws(a, [&](int i) { return rs[a.keep[i]] + types_point(a, p, k, i, s); });
        acc += exp10(t);
      }
      a.post_sub[(size_t)p * S + s] = acc > 1.0 ? 1.0 : acc;
    }
  }
}

}  // namespace

int avg_bfs_run(eqbf_ctx* ctx, const double* raw, int64_t NP, const int64_t* off, int64_t G, const eqbf_avg_opts* o,
                const double* d_gw, const int32_t* d_keep, const double* d_dw, const uint64_t* d_cs,
                const double* d_sw, eqbf_avg_out* out, hipStream_t st) {
  AvgArgs a;
  memset(&a, 0, sizeof a);
  a.raw = raw;
  a.off = off;
  a.keep = d_keep;
  a.gw = d_gw;
  a.dw = d_dw;
  a.cs = d_cs;
  a.sw = d_sw;
  a.gene_bf = out->gene_bf;
  a.gene_post = out->gene_post;
  a.snp_bf = out->snp_bf;
  a.dim_bf = out->dim_bf;
  a.post_the = out->post_the;
  a.post_an = out->post_an;
  a.post_dim = out->post_dim;
  a.post_sub = out->post_sub;
  a.best_dim = out->best_dim;
  a.pi0 = o->pi0;
  a.NP = NP;
  a.G = G;
  a.D = o->dim;
  a.S = o->n_subgroups;
  a.L = o->n_grid;
  a.nk = o->n_keep;
  a.types = o->model == EQBF_AVG_TYPES;
  a.want = o->want;
  /* lanes per pair: the power of two >= dim, at most 64; enough blocks to
   * fill every CU several times over, each wave then strides over the pairs */
  int w = 1;
  while (w < a.D && w < 64) w <<= 1;
  const int64_t per_block = (int64_t)kAvgWaves * (64 / w);
  const unsigned pblocks = (unsigned)std::min<int64_t>((NP + per_block - 1) / per_block, (int64_t)256 * 64);
  if (NP > 0) {
    ProfScope ps(ctx, "avg_pair_kernel", st);
    hipLaunchKernelGGL(avg_pair_kernel, dim3(pblocks), dim3(64 * kAvgWaves), 0, st, a, w);
  }
  if (G > 0) {
    ProfScope ps(ctx, "avg_gene_kernel", st);
    hipLaunchKernelGGL(avg_gene_kernel, dim3((unsigned)G), dim3(64), 0, st, a);
  }
  if (NP > 0 && a.types && (a.want & 8)) {
    ProfScope ps(ctx, "avg_post_kernel", st);
    const unsigned qblocks = (unsigned)std::min<int64_t>((NP + kAvgWaves - 1) / kAvgWaves, (int64_t)256 * 64);
    hipLaunchKernelGGL(avg_post_kernel, dim3(qblocks), dim3(64 * kAvgWaves), 0, st, a);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}

int avg_grid_run(eqbf_ctx* ctx, const double* raw, int64_t n_rows, int L, const double* gw, const int32_t* keep,
                 int nk, double* out, hipStream_t st) {
  if (n_rows <= 0) return EQBF_OK;
  {
    ProfScope ps(ctx, "avg_grid_kernel", st);
    hipLaunchKernelGGL(avg_grid_kernel, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, st, raw, n_rows, L,
                       gw, keep, nk, out);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}

namespace {

int avg_check(eqbf_ctx* ctx, int64_t NP, int64_t G, const eqbf_avg_opts* o) {
  if (!o) return ctx_fail(ctx, EQBF_ERR_ARG, "null options");
  if (o->model != EQBF_AVG_CONFIGS && o->model != EQBF_AVG_TYPES)
    return ctx_fail(ctx, EQBF_ERR_ARG, "unknown model %d", o->model);
  if (o->dim < 1 || o->n_grid < 1 || o->n_keep < 1 || o->n_keep > o->n_grid || NP < 0 || G < 0 ||
      (o->model == EQBF_AVG_TYPES && (o->n_subgroups < 1 || o->n_subgroups > 64)) || o->n_subgroups < 0 ||
      o->n_subgroups > 64)
    return ctx_fail(ctx, EQBF_ERR_ARG, "sizes out of range (dim=%d grid=%d keep=%d S=%d)", o->dim, o->n_grid,
                    o->n_keep, o->n_subgroups);
  if (!o->keep || !o->grid_weights || !o->dim_weights ||
      (o->model == EQBF_AVG_TYPES && !o->sub_weights) ||
      (o->model == EQBF_AVG_CONFIGS && (o->want & EQBF_AVG_POST_D) && !o->cfg_subgroups))
    return ctx_fail(ctx, EQBF_ERR_ARG, "missing weights");
  for (int i = 0; i < o->n_keep; ++i)
    if (o->keep[i] < 0 || o->keep[i] >= o->n_grid) return ctx_fail(ctx, EQBF_ERR_ARG, "grid index %d out of range", o->keep[i]);
  return EQBF_OK;
}

/* the option tables on the device */
struct AvgTables {
  double *gw = nullptr, *dw = nullptr, *sw = nullptr;
  int32_t* keep = nullptr;
  uint64_t* cs = nullptr;
};

int avg_tables(eqbf_ctx* ctx, const eqbf_avg_opts* o, AvgTables& t, hipStream_t st) {
  const int D = o->dim, S = o->n_subgroups;
  t.gw = (double*)ctx_buf(ctx, "avg_gw", (size_t)o->n_grid * sizeof(double));
  t.dw = (double*)ctx_buf(ctx, "avg_dw", (size_t)D * sizeof(double));
  t.keep = (int32_t*)ctx_buf(ctx, "avg_keep", (size_t)o->n_keep * sizeof(int32_t));
  t.cs = (uint64_t*)ctx_buf(ctx, "avg_cs", (size_t)D * sizeof(uint64_t));
  t.sw = (double*)ctx_buf(ctx, "avg_sw", (size_t)D * std::max(S, 1) * sizeof(double));
  if (!t.gw || !t.dw || !t.keep || !t.cs || !t.sw) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating tables failed");
  EQBF_HIP(ctx, hipMemcpyAsync(t.gw, o->grid_weights, (size_t)o->n_grid * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(t.dw, o->dim_weights, (size_t)D * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(t.keep, o->keep, (size_t)o->n_keep * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (o->cfg_subgroups)
    EQBF_HIP(ctx, hipMemcpyAsync(t.cs, o->cfg_subgroups, (size_t)D * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  if (o->sub_weights && S > 0)
    EQBF_HIP(ctx, hipMemcpyAsync(t.sw, o->sub_weights, (size_t)D * S * sizeof(double), hipMemcpyHostToDevice, st));
  return EQBF_OK;
}

}  // namespace

extern "C" int eqbf_avg_bfs_device(eqbf_ctx* ctx, const double* raw, int64_t NP, const int64_t* off, int64_t G,
                                   const eqbf_avg_opts* o, eqbf_avg_out* out, void* stream) {
  if (!ctx || !out) return EQBF_ERR_ARG;
  int rc = avg_check(ctx, NP, G, o);
  if (rc) return rc;
  if ((NP > 0 && (!raw || !out->snp_bf || !out->dim_bf || !out->best_dim)) || (G > 0 && (!off || !out->gene_bf)))
    return ctx_fail(ctx, EQBF_ERR_ARG, "missing input or output");
  const bool types = o->model == EQBF_AVG_TYPES;
  const bool dims_post = (o->want & EQBF_AVG_POST_DIMS) || (!types && (o->want & EQBF_AVG_POST_D));
  if (((o->want & EQBF_AVG_POST_A) && !out->gene_post) || ((o->want & EQBF_AVG_POST_B) && !out->post_the) ||
      ((o->want & EQBF_AVG_POST_C) && !out->post_an) || ((o->want & EQBF_AVG_POST_D) && !out->post_sub) ||
      (dims_post && !out->post_dim))
    return ctx_fail(ctx, EQBF_ERR_ARG, "missing output for a requested posterior");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  AvgTables t;
  if ((rc = avg_tables(ctx, o, t, st))) return rc;
  rc = avg_bfs_run(ctx, raw, NP, off, G, o, t.gw, t.keep, t.dw, t.cs, t.sw, out, st);
  if (rc) return rc;
  /* the host tables must outlive the copies */
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_avg_bfs(eqbf_ctx* ctx, const double* raw, int64_t NP, const int64_t* off, int64_t G,
                            const eqbf_avg_opts* o, eqbf_avg_out* out) {
  if (!ctx || !out) return EQBF_ERR_ARG;
  int rc = avg_check(ctx, NP, G, o);
  if (rc) return rc;
  if ((NP > 0 && !raw) || !off) return ctx_fail(ctx, EQBF_ERR_ARG, "missing input");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int D = o->dim, S = o->n_subgroups, R = o->model == EQBF_AVG_TYPES ? S : D;
  const size_t np = (size_t)std::max<int64_t>(NP, 1), ng = (size_t)std::max<int64_t>(G, 1);
  auto buf = [&](const char* name, size_t bytes) { return ctx_buf(ctx, name, bytes); };
  double* d_raw = (double*)buf("avg_raw", np * R * o->n_grid * sizeof(double));
  int64_t* d_off = (int64_t*)buf("avg_off", (ng + 1) * sizeof(int64_t));
  eqbf_avg_out d;
  d.gene_bf = (double*)buf("avg_gbf", ng * sizeof(double));
  d.gene_post = (double*)buf("avg_gpost", ng * sizeof(double));
  d.snp_bf = (double*)buf("avg_sbf", np * sizeof(double));
  d.dim_bf = (double*)buf("avg_dbf", np * D * sizeof(double));
  d.post_the = (double*)buf("avg_pthe", np * sizeof(double));
  d.post_an = (double*)buf("avg_pan", np * sizeof(double));
  d.post_dim = (double*)buf("avg_pdim", np * D * sizeof(double));
  d.post_sub = (double*)buf("avg_psub", np * std::max(S, 1) * sizeof(double));
  d.best_dim = (int32_t*)buf("avg_best", np * sizeof(int32_t));
  if (!d_raw || !d_off || !d.gene_bf || !d.gene_post || !d.snp_bf || !d.dim_bf || !d.post_the || !d.post_an ||
      !d.post_dim || !d.post_sub || !d.best_dim)
    return ctx_fail(ctx, EQBF_ERR_HIP, "allocating eqtlbma_avg_bfs buffers failed");
  if (NP > 0)
    EQBF_HIP(ctx, hipMemcpyAsync(d_raw, raw, (size_t)NP * R * o->n_grid * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(d_off, off, (size_t)(G + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  AvgTables t;
  if ((rc = avg_tables(ctx, o, t, st))) return rc;
  rc = avg_bfs_run(ctx, d_raw, NP, d_off, G, o, t.gw, t.keep, t.dw, t.cs, t.sw, &d, st);
  if (rc) return rc;
  auto back = [&](void* h, const void* dv, size_t bytes) -> int {
    if (h && bytes) EQBF_HIP(ctx, hipMemcpyAsync(h, dv, bytes, hipMemcpyDeviceToHost, st));
    return EQBF_OK;
  };
  const bool types = o->model == EQBF_AVG_TYPES;
  const bool dims_post = (o->want & EQBF_AVG_POST_DIMS) || (!types && (o->want & EQBF_AVG_POST_D));
  if ((rc = back(out->gene_bf, d.gene_bf, (size_t)G * sizeof(double))) ||
      (rc = back(out->snp_bf, d.snp_bf, (size_t)NP * sizeof(double))) ||
      (rc = back(out->dim_bf, d.dim_bf, (size_t)NP * D * sizeof(double))) ||
      (rc = back(out->best_dim, d.best_dim, (size_t)NP * sizeof(int32_t))) ||
      ((o->want & EQBF_AVG_POST_A) && (rc = back(out->gene_post, d.gene_post, (size_t)G * sizeof(double)))) ||
      ((o->want & EQBF_AVG_POST_B) && (rc = back(out->post_the, d.post_the, (size_t)NP * sizeof(double)))) ||
      ((o->want & EQBF_AVG_POST_C) && (rc = back(out->post_an, d.post_an, (size_t)NP * sizeof(double)))) ||
      (dims_post && (rc = back(out->post_dim, d.post_dim, (size_t)NP * D * sizeof(double)))) ||
      ((o->want & EQBF_AVG_POST_D) && (rc = back(out->post_sub, d.post_sub, (size_t)NP * S * sizeof(double)))))
    return rc;
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_avg_grid(eqbf_ctx* ctx, const double* raw, int64_t n_rows, int32_t L, const double* gw,
                             const int32_t* keep, int32_t nk, double* out) {
  if (!ctx) return EQBF_ERR_ARG;
  if (n_rows < 0 || L < 1 || nk < 1 || nk > L || !gw || !keep || (n_rows > 0 && (!raw || !out)))
    return ctx_fail(ctx, EQBF_ERR_ARG, "invalid arguments");
  for (int i = 0; i < nk; ++i)
    if (keep[i] < 0 || keep[i] >= L) return ctx_fail(ctx, EQBF_ERR_ARG, "grid index %d out of range", keep[i]);
  if (n_rows == 0) return EQBF_OK;
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  double* d_raw = (double*)ctx_buf(ctx, "avg_raw", (size_t)n_rows * L * sizeof(double));
  double* d_out = (double*)ctx_buf(ctx, "avg_gout", (size_t)n_rows * sizeof(double));
  double* d_gw = (double*)ctx_buf(ctx, "avg_gw", (size_t)L * sizeof(double));
  int32_t* d_keep = (int32_t*)ctx_buf(ctx, "avg_keep", (size_t)nk * sizeof(int32_t));
  if (!d_raw || !d_out || !d_gw || !d_keep) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating buffers failed");
  EQBF_HIP(ctx, hipMemcpyAsync(d_raw, raw, (size_t)n_rows * L * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(d_gw, gw, (size_t)L * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(d_keep, keep, (size_t)nk * sizeof(int32_t), hipMemcpyHostToDevice, st));
  int rc = avg_grid_run(ctx, d_raw, n_rows, L, d_gw, d_keep, nk, d_out, st);
  if (rc) return rc;
  EQBF_HIP(ctx, hipMemcpyAsync(out, d_out, (size_t)n_rows * sizeof(double), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}
