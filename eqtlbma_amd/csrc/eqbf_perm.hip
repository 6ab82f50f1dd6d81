/*
 * Permutation statistics, general path.
 *
 * Reference loops (eqtlbma 1.3.3): Gene::MakePermutationsJoin
 * (gene.cpp:598-717), MakePermutationsSepAllSubgroups (gene.cpp:493-570),
 * MakePermutationsSepPerSubgroup (gene.cpp:380-450).  Under a permutation
 * the expression level used for sample i is the one of sample perm(i), while
 * genotypes and covariates stay with i (gene_snp_pair.cpp:104-149); which
 * samples are used can therefore change from one permutation to the next.
 *
 * Each (gene, permutation) is laid out as a virtual gene whose expression
 * levels are gathered through the permutation, and the virtual genes of a
 * batch go through the association engine (prep, sums, Bayes factors) in
 * one call.  A reduction per virtual gene then forms the statistic:
 *   join: max over SNPs of the grid-averaged ABF[pbf] (std::max_element
 *         semantics: NaN only when the first SNP's value is NaN,
 *         gene.cpp:677-678) or its log10 average over SNPs
 *         (utils::log10_weighted_sum, gene.cpp:689-690);
 *   sep:  minimum p-value over SNPs (and subgroups for --permsep 1).
 * This path is exact for any input (missing values, uneven membership,
 * covariates).  With the mixed precision, genes whose every sample is in
 * every subgroup go through the contraction path (eqbf_perm_fast.hip) first
 * and only the others come here.
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

__global__ void perm_gather_kernel(const double* __restrict__ y, const uint8_t* __restrict__ gene_in,
                                   const int32_t* __restrict__ tables, const int32_t* __restrict__ vg_gene,
                                   const int32_t* __restrict__ vg_table, const int32_t* __restrict__ vg_perm,
                                   int64_t nperm, int S, int N, double* __restrict__ y_v,
                                   uint8_t* __restrict__ gene_in_v) {
  const int v = blockIdx.x / S, s = blockIdx.x % S;
  const int g = vg_gene[v];
  const double* src = y + ((int64_t)g * S + s) * N;
  double* dst = y_v + ((int64_t)v * S + s) * N;
  if (vg_perm[v] < 0) { /* the unpermuted data (true statistic) */
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
  } else {
    const int32_t* perm = tables + ((int64_t)vg_table[v] * nperm + vg_perm[v]) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[perm[i]];
  }
  if (threadIdx.x == 0) gene_in_v[(int64_t)v * S + s] = gene_in[(int64_t)g * S + s];
}

/* the levels as loaded, per virtual gene (the hybrid model's pairwise step) */
__global__ void perm_levels_kernel(const double* __restrict__ y, const int32_t* __restrict__ vg_gene, int S, int N,
                                   double* __restrict__ y_v) {
  const int v = blockIdx.x / S, s = blockIdx.x - (blockIdx.x / S) * S;
  const int g = vg_gene[v];
  const double* src = y + ((int64_t)g * S + s) * N;
  double* dst = y_v + ((int64_t)v * S + s) * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

__global__ void perm_pairs_kernel(const int64_t* __restrict__ pair_offset, const int32_t* __restrict__ pair_snp,
                                  const int32_t* __restrict__ vg_gene, const int64_t* __restrict__ off_v,
                                  int32_t* __restrict__ pair_snp_v) {
  const int v = blockIdx.x;
  const int g = vg_gene[v];
  const int64_t a = pair_offset[g], n = pair_offset[g + 1] - a, d = off_v[v];
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) pair_snp_v[d + k] = pair_snp[a + k];
}

__global__ void zero_u64_kernel_p(uint64_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void fill_nan_kernel(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = NAN;
}

__device__ inline double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wmin(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* One wave per virtual gene.  A virtual gene with vg_perm < 0 is the
 * unpermuted gene and gets the true statistic's semantics, which differ from
 * a permutation's: Gene::FindMaxTrueL10Abf keeps the largest value a `>`
 * comparison sees (NaN never wins, -inf when every SNP is NaN) and
 * AvgTrueL10Abfs averages the non-NaN values only (gene.cpp:575-596); the
 * sep statistics (FindMinTruePvalue*, gene.cpp:369-378, 475-488) start from
 * 1 like the permutations'.  The statistic goes to stats[g * out_stride +
 * max(vg_perm, 0)]. */
__global__ void __launch_bounds__(64) perm_reduce_kernel(
    const int64_t* __restrict__ off_v, const int32_t* __restrict__ vg_gene, const int32_t* __restrict__ vg_perm,
    const int32_t* __restrict__ pair_snp_v, const int32_t* __restrict__ n_out, const double* __restrict__ sstats,
    const double* __restrict__ w, const uint8_t* __restrict__ gene_in, const uint8_t* __restrict__ snp_in,
    int64_t NP, int S, int join, int w_row, int use_max, int sep_sbgrp, int64_t out_stride, int mvlr,
    double* __restrict__ stats) {
  const bool truth = vg_perm[blockIdx.x] < 0;
  const int v = blockIdx.x, lane = threadIdx.x;
  const int g = vg_gene[v];
  const int64_t a = off_v[v], b = off_v[v + 1];
  auto val = [&](int64_t q) -> double {
    const int64_t snp = pair_snp_v[q];
    if (join && mvlr && !truth) { /* a SNP skipped by --error mvlr keeps 0.0 */
      bool all = true;
      for (int s = 0; s < S; ++s) all &= snp_in[snp * S + s] != 0;
      if (!all) return 0.0;
    }
    if (join) return w[(int64_t)w_row * NP + q];
    if (sep_sbgrp >= 0) {
      const int s = sep_sbgrp;
      if (!(gene_in[(int64_t)g * S + s] && snp_in[snp * S + s])) return 1.0;
      return n_out[(int64_t)s * NP + q] > 0 ? sstats[((int64_t)4 * S + s) * NP + q] : 1.0;
    }
    double mn = 1.0;
    for (int s = 0; s < S; ++s) {
      if (!(gene_in[(int64_t)g * S + s] && snp_in[snp * S + s]) || n_out[(int64_t)s * NP + q] <= 0) continue;
      const double pv = sstats[((int64_t)4 * S + s) * NP + q];
      if (pv < mn) mn = pv;
    }
    return mn;
  };
  double res;
  if (b <= a) {
    res = NAN;
  } else {
    const double v0 = val(a);
    if (join && !use_max) {
      /* utils::log10_weighted_sum over the SNPs, weights 1 / (number of
       * SNPs); for the true statistic over the non-NaN values only */
      double mx = -INFINITY, cnt = 0;
      for (int64_t q = a + lane; q < b; q += 64) {
        const double x = val(q);
        if (!isnan(x)) {
          mx = fmax(mx, x);
          cnt += 1;
        }
      }
      mx = wmax(mx);
      cnt = wsum(cnt);
      if ((!truth && isnan(v0)) || (truth && cnt == 0)) {
        res = NAN;
      } else {
        double sum = 0;
        for (int64_t q = a + lane; q < b; q += 64) {
          const double x = val(q);
          if (!isnan(x)) sum += exp10(x - mx);
        }
        sum = wsum(sum) / (truth ? cnt : (double)(b - a));
        res = mx + log10(sum);
        if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
      }
    } else {
      /* std::max_element / std::min_element: NaN only if the first is NaN;
       * the true statistic: NaN never wins, sep starts from 1 */
      double m = join ? -INFINITY : (truth ? 1.0 : INFINITY);
      for (int64_t q = a + lane; q < b; q += 64) {
        const double x = val(q);
        if (!isnan(x)) m = join ? fmax(m, x) : fmin(m, x);
      }
      m = join ? wmax(m) : wmin(m);
      res = (!truth && isnan(v0)) ? NAN : m;
    }
  }
  if (lane == 0) stats[(int64_t)g * out_stride + (truth ? 0 : vg_perm[v])] = res;
}

/* Permutations whose statistic lies within the bound of the gene's true
 * statistic (absolute for join, relative for sep): their indices g * nperm +
 * p, appended through a counter. */
__global__ void tie_flag_kernel(const double* __restrict__ stats, const double* __restrict__ truth, int G,
                                int64_t nperm, double bound, int rel, unsigned long long* __restrict__ count,
                                int64_t* __restrict__ list) {
  const int64_t n = (int64_t)G * nperm;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double s = stats[i], t = truth[i / nperm];
    if (!isfinite(s) || !isfinite(t)) continue;
    if (fabs(s - t) <= bound * (rel ? fabs(t) : 1.0)) list[atomicAdd(count, 1ull)] = i;
  }
}

}  // namespace

static int pbf_bfs(int pbf) { return pbf == EQBF_PBF_GEN ? EQBF_BFS_GEN : (pbf == EQBF_PBF_GENSIN ? EQBF_BFS_SIN : EQBF_BFS_ALL); }

/* Statistics of a list of (gene, permutation) items through the association
 * engine with the given precision: items[k] = {g, p}, p < 0 for the true
 * statistic; out[g * out_stride + max(p, 0)]. */
static int general_items_run(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_perm_opts* po, const int32_t* tables_dev,
                             const int32_t* gene_table, int64_t nperm, const std::vector<int64_t>& off,
                             const std::vector<std::pair<int32_t, int32_t>>& items, int precision, double* out,
                             int64_t out_stride, hipStream_t st, bool fp64_check = false) {
  const int S = c.n_subgroups, N = c.n_samples;
  const bool join = po->analysis == EQBF_ANALYSIS_JOIN;
  const int bfs = join ? pbf_bfs(po->pbf) : EQBF_BFS_GEN;
  const int64_t n_cfg = join ? eqbf_n_configs(bfs, S) : 0;
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--pbf all with %d subgroups", S);
  const int64_t n_w = 3 + n_cfg + 2;
  const int w_row = !join ? 0 : (po->pbf == EQBF_PBF_GEN ? 0 : (po->pbf == EQBF_PBF_GENSIN ? (int)(3 + n_cfg) : (int)(3 + n_cfg + 1)));
  const bool mvlr = join && po->error_model == EQBF_ERROR_MVLR;
  const bool hybrid = join && po->error_model == EQBF_ERROR_HYBRID;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const int64_t n_raw = (mvlr || hybrid) ? 3 * (int64_t)GL + n_cfg * GS : 0;
  /* batch: at most 64k virtual genes and a bounded number of virtual pairs */
  const int64_t per_pair_bytes = (int64_t)S * 4 + (join ? (n_w + n_raw) * 8 : (int64_t)5 * S * 8) + 4;
  const int64_t pmax = std::max<int64_t>(1 << 16, std::min<int64_t>(8 << 20, ((int64_t)3 << 30) / per_pair_bytes));
  std::vector<int32_t> vg_gene, vg_table, vg_perm;
  std::vector<int64_t> off_v;
  eqbf_assoc_opts ao;
  ao.analysis = join ? EQBF_ANALYSIS_JOIN : EQBF_ANALYSIS_SEP;
  ao.bfs = bfs;
  ao.precision = precision;
  ao.raw_dtype = EQBF_F32;
  auto flush = [&]() -> int {
    const int VB = (int)vg_gene.size();
    if (VB == 0) return EQBF_OK;
    const int64_t NPv = off_v.back();
    int32_t* d_vg = (int32_t*)ctx_buf(ctx, "p_vg", (size_t)VB * 3 * sizeof(int32_t));
    int64_t* d_off = (int64_t*)ctx_buf(ctx, "p_off", (size_t)(VB + 1) * sizeof(int64_t));
    double* d_y = (double*)ctx_buf(ctx, "p_y", (size_t)VB * S * N * sizeof(double));
    uint8_t* d_gin = (uint8_t*)ctx_buf(ctx, "p_gin", (size_t)VB * S);
    int32_t* d_psnp = (int32_t*)ctx_buf(ctx, "p_psnp", (size_t)std::max<int64_t>(NPv, 1) * sizeof(int32_t));
    int32_t* d_n = (int32_t*)ctx_buf(ctx, "p_n", (size_t)S * std::max<int64_t>(NPv, 1) * sizeof(int32_t));
    double* d_w = join ? (double*)ctx_buf(ctx, "p_w", (size_t)n_w * std::max<int64_t>(NPv, 1) * sizeof(double)) : nullptr;
    double* d_ss = !join ? (double*)ctx_buf(ctx, "p_ss", (size_t)5 * S * std::max<int64_t>(NPv, 1) * sizeof(double)) : nullptr;
    double* d_raw = (mvlr || hybrid) ? (double*)ctx_buf(ctx, "p_raw", (size_t)n_raw * std::max<int64_t>(NPv, 1) * sizeof(double)) : nullptr;
    double* d_yp = hybrid ? (double*)ctx_buf(ctx, "p_yp", (size_t)VB * S * N * sizeof(double)) : nullptr;
    if (!d_vg || !d_off || !d_y || !d_gin || !d_psnp || !d_n || (join && !d_w) || (!join && !d_ss) ||
        ((mvlr || hybrid) && !d_raw) || (hybrid && !d_yp))
      return EQBF_ERR_HIP;
    std::vector<int32_t> packed;
    packed.insert(packed.end(), vg_gene.begin(), vg_gene.end());
    packed.insert(packed.end(), vg_table.begin(), vg_table.end());
    packed.insert(packed.end(), vg_perm.begin(), vg_perm.end());
    EQBF_HIP(ctx, hipMemcpyAsync(d_vg, packed.data(), packed.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_off, off_v.data(), off_v.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    {
      /* the FP64 true statistics and re-checks are timed under names of
       * their own, so that a profile shows which genes took this path */
      ProfScope ps(ctx, fp64_check ? "fp64_check_gather_kernel" : "perm_gather_kernel", st);
      hipLaunchKernelGGL(perm_gather_kernel, dim3(VB * S), dim3(256), 0, st, c.explevels, c.gene_in_sbgrp,
                         tables_dev, d_vg, d_vg + VB, d_vg + 2 * VB, nperm, S, N, d_y, d_gin);
      hipLaunchKernelGGL(perm_pairs_kernel, dim3(VB), dim3(256), 0, st, c.pair_offset, c.pair_snp, d_vg, d_off,
                         d_psnp);
      if (hybrid)
        hipLaunchKernelGGL(perm_levels_kernel, dim3(VB * S), dim3(256), 0, st,
                           po->explevels_pairs ? po->explevels_pairs : c.explevels, d_vg, S, N, d_yp);
    }
    EQBF_HIP(ctx, hipGetLastError());
    eqbf_chunk vc = c;
    vc.n_genes = VB;
    vc.n_pairs = NPv;
    vc.explevels = d_y;
    vc.gene_in_sbgrp = d_gin;
    vc.pair_offset = d_off;
    vc.pair_snp = d_psnp;
    eqbf_assoc_out vo;
    memset(&vo, 0, sizeof vo);
    vo.n = d_n;
    vo.sstats = d_ss;
    vo.w_abf = d_w;
    int rc;
    if (mvlr) {
      eqbf_mvlr_opts mo;
      mo.bfs = bfs;
      mo.fiterr = po->fiterr;
      vo.raw_abf = d_raw;
      rc = mvlr_run(ctx, &vc, &mo, &vo, st);
    } else if (hybrid) {
      eqbf_hybrid_opts ho;
      ho.bfs = bfs;
      ho.fiterr = po->fiterr;
      ho.covars_by_position = po->covars_by_position;
      ho.explevels_pairs = d_yp;
      vo.raw_abf = d_raw;
      rc = hybrid_run(ctx, &vc, &ho, &vo, st);
    } else {
      rc = assoc_uvlr_run(ctx, &vc, &ao, &vo, st);
    }
    if (rc) return rc;
    {
      ProfScope ps(ctx, fp64_check ? "fp64_check_reduce_kernel" : "perm_reduce_kernel", st);
      hipLaunchKernelGGL(perm_reduce_kernel, dim3(VB), dim3(64), 0, st, d_off, d_vg, d_vg + 2 * VB, d_psnp, d_n,
                         d_ss, d_w, c.gene_in_sbgrp, c.snp_in_sbgrp, NPv, S, (int)join, w_row, po->use_max_bf,
                         po->sep_sbgrp, out_stride, (int)(mvlr || hybrid), out);
    }
    EQBF_HIP(ctx, hipGetLastError());
    /* the host vectors are reused for the next batch */
    EQBF_HIP(ctx, hipStreamSynchronize(st));
    vg_gene.clear();
    vg_table.clear();
    vg_perm.clear();
    off_v.assign(1, 0);
    return EQBF_OK;
  };
  off_v.assign(1, 0);
  for (const auto& it : items) {
    const int g = it.first;
    const int64_t np = off[g + 1] - off[g];
    if (np <= 0) continue;
    if ((int64_t)vg_gene.size() >= 65536 || off_v.back() + np > pmax) {
      int rc = flush();
      if (rc) return rc;
    }
    vg_gene.push_back(g);
    vg_table.push_back(std::max(gene_table[g], 0));
    vg_perm.push_back(it.second);
    off_v.push_back(off_v.back() + np);
  }
  return flush();
}

static int gene_offsets(eqbf_ctx* ctx, const eqbf_chunk& c, std::vector<int64_t>& off, hipStream_t st) {
  off.resize(c.n_genes + 1);
  EQBF_HIP(ctx, hipMemcpyAsync(off.data(), c.pair_offset, (c.n_genes + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

static void nan_fill(double* p, int64_t n, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(fill_nan_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0, st,
                       p, n);
}

int perm_stats_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_perm_opts* po, const int32_t* tables_dev,
                   int32_t n_tables, const int32_t* gene_table, int64_t nperm, double* stats_dev, hipStream_t st) {
  const eqbf_chunk& c = *cp;
  const int G = c.n_genes;
  if (nperm <= 0 || G <= 0) return EQBF_OK;
  if (nperm > INT32_MAX) return ctx_fail(ctx, EQBF_ERR_ARG, "nperm %lld out of range", (long long)nperm);
  for (int g = 0; g < G; ++g)
    if (gene_table[g] >= n_tables) return ctx_fail(ctx, EQBF_ERR_ARG, "gene_table[%d] = %d out of range", g, gene_table[g]);
  std::vector<int64_t> off;
  int rc = gene_offsets(ctx, c, off, st);
  if (rc) return rc;
  /* every statistic starts as NaN (skipped genes keep it) */
  nan_fill(stats_dev, (int64_t)G * nperm, st);
  EQBF_HIP(ctx, hipGetLastError());
  /* genes that qualify go through the contraction path first */
  const bool join = po->analysis == EQBF_ANALYSIS_JOIN;
  const bool multi = join && (po->error_model == EQBF_ERROR_MVLR || po->error_model == EQBF_ERROR_HYBRID);
  if (po->error_model != EQBF_ERROR_UVLR && !multi)
    return ctx_fail(ctx, EQBF_ERR_ARG, "error_model %d is not valid for this analysis", po->error_model);
  std::vector<uint8_t> done(G, 0);
  if (!multi) { /* the contraction path is the uvlr model's */
    rc = perm_fast_run(ctx, cp, po, tables_dev, n_tables, gene_table, nperm, stats_dev, st, done);
    if (rc) return rc;
  }
  std::vector<std::pair<int32_t, int32_t>> items;
  for (int g = 0; g < G; ++g) {
    if (gene_table[g] < 0 || done[g] || off[g + 1] <= off[g]) continue;
    for (int64_t p = 0; p < nperm; ++p) items.push_back({g, (int32_t)p});
  }
  return general_items_run(ctx, c, po, tables_dev, gene_table, nperm, off, items, po->precision, stats_dev, nperm, st);
}

int perm_ties_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_perm_opts* po, const int32_t* tables_dev,
                  int32_t n_tables, const int32_t* gene_table, int64_t nperm, double tie_bound, double* stats_dev,
                  double* true_dev, int64_t* n_rechecked, hipStream_t st) {
  const eqbf_chunk& c = *cp;
  const int G = c.n_genes;
  if (n_rechecked) *n_rechecked = 0;
  if (G <= 0) return EQBF_OK;
  if (!(tie_bound >= 0)) return ctx_fail(ctx, EQBF_ERR_ARG, "tie_bound must be >= 0");
  int rc = perm_stats_run(ctx, cp, po, tables_dev, n_tables, gene_table, nperm, stats_dev, st);
  if (rc) return rc;
  /* FP64 already: the FP64 precisions and the MVLR and hybrid engines */
  const bool fp64 = po->precision == EQBF_PREC_FP64_PLAIN || po->precision == EQBF_PREC_FP64_BATCHED ||
                    (po->analysis == EQBF_ANALYSIS_JOIN &&
                     (po->error_model == EQBF_ERROR_MVLR || po->error_model == EQBF_ERROR_HYBRID));
  if (fp64 || nperm <= 0) return EQBF_OK;
  /* permutations within the bound of the caller's true statistic */
  const int64_t n = (int64_t)G * nperm;
  unsigned long long* d_cnt = (unsigned long long*)ctx_buf(ctx, "tie_cnt", 64);
  int64_t* d_list = (int64_t*)ctx_buf(ctx, "tie_list", (size_t)n * sizeof(int64_t));
  if (!d_cnt || !d_list) return EQBF_ERR_HIP;
  hipLaunchKernelGGL(zero_u64_kernel_p, dim3(1), dim3(64), 0, st, (uint64_t*)d_cnt, (int64_t)1);
  hipLaunchKernelGGL(tie_flag_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16384)), dim3(256), 0, st,
                     stats_dev, true_dev, G, nperm, tie_bound, po->analysis == EQBF_ANALYSIS_SEP ? 1 : 0, d_cnt, d_list);
  EQBF_HIP(ctx, hipGetLastError());
  unsigned long long cnt = 0;
  EQBF_HIP(ctx, hipMemcpyAsync(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  if (cnt == 0) return EQBF_OK;
  std::vector<int64_t> list(cnt);
  EQBF_HIP(ctx, hipMemcpyAsync(list.data(), d_list, cnt * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  std::sort(list.begin(), list.end());
  std::vector<int64_t> off;
  rc = gene_offsets(ctx, c, off, st);
  if (rc) return rc;
  /* the genes concerned: true statistic in FP64 (overwriting the caller's),
   * then their near-tie permutations in FP64 */
  std::vector<std::pair<int32_t, int32_t>> items;
  for (int64_t i : list) {
    const int32_t g = (int32_t)(i / nperm);
    if (items.empty() || items.back().first != g) items.push_back({g, -1});
  }
  rc = general_items_run(ctx, c, po, tables_dev, gene_table, nperm, off, items, EQBF_PREC_FP64_BATCHED, true_dev, 1,
                         st, true);
  if (rc) return rc;
  items.clear();
  for (int64_t i : list) items.push_back({(int32_t)(i / nperm), (int32_t)(i % nperm)});
  if (n_rechecked) *n_rechecked = (int64_t)cnt;
  return general_items_run(ctx, c, po, tables_dev, gene_table, nperm, off, items, EQBF_PREC_FP64_BATCHED, stats_dev,
                           nperm, st, true);
}
