/*
 * eqbf_oracle_perm.cpp -- CPU oracle for eqtlbma_bf's permutation p-values.
 *
 * TEST INFRASTRUCTURE ONLY (see eqbf_oracle.cpp).
 *
 * Follows Gene::MakePermutationsJoin (gene.cpp:598-717),
 * Gene::MakePermutationsSepAllSubgroups (gene.cpp:493-570),
 * Gene::MakePermutationsSepPerSubgroup (gene.cpp:380-450),
 * Gene::CalcPermutationPvalue (gene.cpp:348-364) and the chunk drivers
 * makePermutationsJoin / makePermutationsSep (eqtlbma_bf.cpp:773-864).
 *
 * The reference draws permutations with GSL inside these loops; here the
 * index tables are inputs (made by orc_perm_tables), so the statistic and
 * counting logic is checked independently of the generator.  Pinned by
 * brute force: on tiny inputs the exact permutation distribution is
 * enumerated in tests/test_oracle_perm.py.
 *
 * The reference takes the median over one element past the valid ones
 * (gene.cpp:713-714, undefined behaviour); this oracle takes the median of
 * the valid values and returns it.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "eqbf_oracle_internal.h"

extern "C" int orc_assoc_mvlr(const eqbf_chunk* c, int32_t bfs, double fiterr, const double* phiL,
                              const double* omaL, int32_t GL, const double* phiS, const double* omaS,
                              int32_t GS, int32_t* n_out, double* raw, double* w_abf);

using namespace orc;
static const double kNaN = std::numeric_limits<double>::quiet_NaN();

static int32_t pbf_to_bfs(int32_t pbf) {
  return pbf == EQBF_PBF_GEN ? EQBF_BFS_GEN : (pbf == EQBF_PBF_GENSIN ? EQBF_BFS_SIN : EQBF_BFS_ALL);
}
static int64_t pbf_index(int32_t pbf, int32_t S) {
  int32_t bfs = pbf_to_bfs(pbf);
  int64_t nc = n_configs(bfs, S);
  if (pbf == EQBF_PBF_GEN) return 0;
  if (pbf == EQBF_PBF_GENSIN) return 3 + nc;
  return 3 + nc + 1;
}

/* Weighted ABF[pbf] of one pair under a permutation (the body of the OpenMP
 * loop, gene.cpp:650-661). */
static double pair_stat(const eqbf_chunk* c, int64_t g, int64_t snp, const int32_t* perm, int32_t pbf,
                        const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                        const double* omaS, int32_t GS, int* err) {
  const int32_t S = c->n_subgroups;
  std::vector<int> has(S, 0);
  std::vector<double> std3(3 * S, kNaN);
  for (int32_t s = 0; s < S; ++s) {
    if (!(c->gene_in_sbgrp[g * S + s] && c->snp_in_sbgrp[snp * S + s])) continue;
    double o5[5];
    int64_t n = sstats_one(c, g, s, snp, perm, o5);
    if (n < 0) {
      *err = 1;
      return kNaN;
    }
    has[s] = n > 0;
    if (n > 0) orc_standardize((double)n, (double)c->covar_count[s], o5[1], o5[2], o5[3], &std3[3 * s]);
  }
  int32_t bfs = pbf_to_bfs(pbf);
  int64_t nc = n_configs(bfs, S);
  std::vector<double> raw(3 * (int64_t)GL + nc * GS), w(3 + nc + 2);
  pair_abfs(S, bfs, has, std3, phiL, omaL, GL, phiS, omaS, GS, raw.data(), w.data());
  return w[pbf_index(pbf, S)];
}

/* std::max_element semantics: NaN only when the first element is NaN. */
static double max_elem(const std::vector<double>& v) {
  double m = v[0];
  for (double x : v)
    if (m < x) m = x;
  return m;
}
static double min_elem(const std::vector<double>& v) {
  double m = v[0];
  for (double x : v)
    if (x < m) m = x;
  return m;
}

static double median_of(std::vector<double> v) {
  if (v.empty()) return kNaN;
  size_t n = v.size(), mid = n / 2;
  std::nth_element(v.begin(), v.begin() + mid, v.end());
  double a = v[mid];
  if (n % 2) return a;
  std::nth_element(v.begin(), v.begin() + mid - 1, v.end());
  return (a + v[mid - 1]) / 2.0;
}

/* Per-permutation statistic of gene g for every permutation of its table:
 * max (use_max_bf) or log10 average over its pairs of weighted ABF[pbf].
 * stats[p] is NaN when the statistic is NaN.  This is what the GPU engine
 * computes; the counting below is separate. */
extern "C" int orc_perm_join_stats(const eqbf_chunk* c, int64_t g, int32_t pbf, int32_t use_max_bf,
                                   const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                                   const double* omaS, int32_t GS, const int32_t* table, int64_t nperm,
                                   double* stats) {
  const int64_t N = c->n_samples;
  int64_t p0 = c->pair_offset[g], p1 = c->pair_offset[g + 1];
  std::vector<double> v(p1 - p0);
  for (int64_t p = 0; p < nperm; ++p) {
    const int32_t* perm = table + p * N;
    for (int64_t q = p0; q < p1; ++q) {
      int err = 0;
      v[q - p0] = pair_stat(c, g, c->pair_snp[q], perm, pbf, phiL, omaL, GL, phiS, omaS, GS, &err);
      if (err) return EQBF_ERR_DATA;
    }
    stats[p] = use_max_bf ? max_elem(v) : orc_log10_weighted_sum(v.data(), (int64_t)v.size());
  }
  return EQBF_OK;
}

/* --error mvlr (Gene::MakePermutationsJoin, gene.cpp:662-672): under a
 * permutation sample i takes the expression level of sample perm(i) while
 * genotypes and covariates stay with i (FillStlContainers,
 * gene_snp_pair.cpp:104-149), and CalcAbfsMvlr runs per SNP; a SNP without
 * genotypes in every subgroup is skipped and keeps its initial 0.0
 * (gene.cpp:650, 663-664).  Each permutation is the association test of a
 * one-gene chunk whose levels are read through the permutation. */
extern "C" int orc_perm_join_stats_mvlr(const eqbf_chunk* c, int64_t g, int32_t pbf, int32_t use_max_bf,
                                        double fiterr, const double* phiL, const double* omaL, int32_t GL,
                                        const double* phiS, const double* omaS, int32_t GS, const int32_t* table,
                                        int64_t nperm, double* stats) {
  const int32_t S = c->n_subgroups;
  const int64_t N = c->n_samples;
  const int64_t p0 = c->pair_offset[g], np = c->pair_offset[g + 1] - p0;
  const int32_t bfs = pbf_to_bfs(pbf);
  const int64_t nc = n_configs(bfs, S), n_raw = 3 * (int64_t)GL + nc * GS, n_w = 3 + nc + 2;
  const int64_t row = pbf_index(pbf, S);
  std::vector<double> y((size_t)S * N);
  std::vector<int64_t> poff = {0, np};
  eqbf_chunk c1 = *c;
  c1.n_genes = 1;
  c1.n_pairs = np;
  c1.explevels = y.data();
  c1.gene_in_sbgrp = c->gene_in_sbgrp + g * S;
  c1.pair_offset = poff.data();
  c1.pair_snp = c->pair_snp + p0;
  std::vector<int32_t> n_out((size_t)S * std::max<int64_t>(np, 1));
  std::vector<double> raw((size_t)n_raw * std::max<int64_t>(np, 1)), w((size_t)n_w * std::max<int64_t>(np, 1));
  std::vector<double> v(np);
  for (int64_t p = 0; p < nperm; ++p) {
    const int32_t* perm = table + p * N;
    for (int32_t s = 0; s < S; ++s)
      for (int64_t i = 0; i < N; ++i) y[(size_t)s * N + i] = expl_at(c, g, s, perm[i]);
    int rc = orc_assoc_mvlr(&c1, bfs, fiterr, phiL, omaL, GL, phiS, omaS, GS, n_out.data(), raw.data(), w.data());
    if (rc) return rc;
    for (int64_t q = 0; q < np; ++q) {
      const int64_t snp = c1.pair_snp[q];
      bool all_geno = true;
      for (int32_t s = 0; s < S; ++s) all_geno &= c->snp_in_sbgrp[snp * S + s] != 0;
      v[q] = all_geno ? w[(size_t)row * np + q] : 0.0;
    }
    stats[p] = use_max_bf ? max_elem(v) : orc_log10_weighted_sum(v.data(), (int64_t)v.size());
  }
  return EQBF_OK;
}

/* --error hybrid permutation statistics of gene g (gene.cpp:669-672): the
 * per-subgroup fits on the permuted levels, the pairwise step on the levels
 * as loaded (expl_pair, NULL: the chunk's), a SNP without genotypes in
 * every subgroup contributing 0.  Pinned by construction: the hybrid oracle
 * (pinned in tests/test_oracle_hybrid.py) on explicitly permuted levels. */
extern "C" int orc_assoc_hybrid(const eqbf_chunk* c, int32_t bfs, double fiterr, const double* covars_pos,
                                const double* expl_pair, const double* phiL, const double* omaL, int32_t GL,
                                const double* phiS, const double* omaS, int32_t GS, int32_t* n_out, double* sstats,
                                double* raw, double* w_abf);
extern "C" int orc_perm_join_stats_hybrid(const eqbf_chunk* c, int64_t g, int32_t pbf, int32_t use_max_bf,
                                          double fiterr, const double* covars_pos, const double* expl_pair,
                                          const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                                          const double* omaS, int32_t GS, const int32_t* table, int64_t nperm,
                                          double* stats) {
  const int32_t S = c->n_subgroups;
  const int64_t N = c->n_samples;
  const int64_t p0 = c->pair_offset[g], np = c->pair_offset[g + 1] - p0;
  const int32_t bfs = pbf_to_bfs(pbf);
  const int64_t nc = n_configs(bfs, S), n_raw = 3 * (int64_t)GL + nc * GS, n_w = 3 + nc + 2;
  const int64_t row = pbf_index(pbf, S);
  std::vector<double> y((size_t)S * N);
  const double* yp = (expl_pair ? expl_pair : c->explevels) + (size_t)g * S * N;
  std::vector<int64_t> poff = {0, np};
  eqbf_chunk c1 = *c;
  c1.n_genes = 1;
  c1.n_pairs = np;
  c1.explevels = y.data();
  c1.gene_in_sbgrp = c->gene_in_sbgrp + g * S;
  c1.pair_offset = poff.data();
  c1.pair_snp = c->pair_snp + p0;
  std::vector<int32_t> n_out((size_t)S * std::max<int64_t>(np, 1));
  std::vector<double> ss((size_t)5 * S * std::max<int64_t>(np, 1));
  std::vector<double> raw((size_t)n_raw * std::max<int64_t>(np, 1)), w((size_t)n_w * std::max<int64_t>(np, 1));
  std::vector<double> v(np);
  for (int64_t p = 0; p < nperm; ++p) {
    const int32_t* perm = table + p * N;
    for (int32_t s = 0; s < S; ++s)
      for (int64_t i = 0; i < N; ++i) y[(size_t)s * N + i] = expl_at(c, g, s, perm[i]);
    int rc = orc_assoc_hybrid(&c1, bfs, fiterr, covars_pos, yp, phiL, omaL, GL, phiS, omaS, GS, n_out.data(),
                              ss.data(), raw.data(), w.data());
    if (rc) return rc;
    for (int64_t q = 0; q < np; ++q) {
      const int64_t snp = c1.pair_snp[q];
      bool all_geno = true;
      for (int32_t s = 0; s < S; ++s) all_geno &= c->snp_in_sbgrp[snp * S + s] != 0;
      v[q] = all_geno ? w[(size_t)row * np + q] : 0.0;
    }
    stats[p] = use_max_bf ? max_elem(v) : orc_log10_weighted_sum(v.data(), (int64_t)v.size());
  }
  return EQBF_OK;
}

/* True statistic of a gene from its pairs' weighted ABF[pbf]
 * (FindMaxTrueL10Abf gene.cpp:575-582 / AvgTrueL10Abfs :587-596). */
extern "C" double orc_true_stat(const double* wpbf, int64_t n, int32_t use_max_bf) {
  if (use_max_bf) {
    double m = -std::numeric_limits<double>::infinity();
    for (int64_t i = 0; i < n; ++i)
      if (wpbf[i] > m) m = wpbf[i];
    return m;
  }
  std::vector<double> v;
  for (int64_t i = 0; i < n; ++i)
    if (!std::isnan(wpbf[i])) v.push_back(wpbf[i]);
  if (v.empty()) return kNaN; /* reference reads an empty vector here */
  return orc_log10_weighted_sum(v.data(), (int64_t)v.size());
}

/* Counting loop of Gene::MakePermutationsJoin (gene.cpp:637-716) replayed over
 * precomputed per-permutation statistics.  nb_permutations is the caller's
 * shared counter (decremented on NaN, gene.cpp:680-695).  rng_trick may be
 * NULL when trick == 0.  Outputs: pval, number of valid permutations, median.
 * Returns the number of permutations actually drawn (for the stream). */
extern "C" int64_t orc_perm_count(const double* stats, double true_stat, int64_t* nb_permutations, int32_t trick,
                                  int64_t trick_cutoff, orc_mt* rng_trick, double* pval, int64_t* nperm_used,
                                  double* median);
extern "C" int64_t orc_perm_count(const double* stats, double true_stat, int64_t* nb_permutations, int32_t trick,
                                  int64_t trick_cutoff, orc_mt* rng_trick, double* pval, int64_t* nperm_used,
                                  double* median) {
  int64_t sofar = 0;
  double count = 1;
  bool shuffle_only = false;
  std::vector<double> kept;
  int64_t n_all = *nb_permutations, drawn = 0;
  for (int64_t p = 0; p < n_all; ++p) {
    ++drawn;
    if (shuffle_only) continue;
    double st = stats[p];
    if (std::isnan(st)) {
      (*nb_permutations)--;
      continue;
    }
    ++sofar;
    if (st >= true_stat) ++count;
    kept.push_back(st);
    if (trick != 0 && count == 1 + trick_cutoff) {
      if (trick == 1) break;
      shuffle_only = true;
    }
  }
  /* CalcPermutationPvalue (gene.cpp:348-364) */
  if (sofar == *nb_permutations)
    *pval = count / (double)(*nb_permutations + 1);
  else if (rng_trick)
    *pval = orc_flat(rng_trick, (1 + trick_cutoff) / ((double)(sofar + 2)),
                     (1 + trick_cutoff) / ((double)(sofar + 1)));
  else
    *pval = kNaN; /* the reference dereferences a NULL generator here */
  *nperm_used = sofar;
  *median = median_of(kept);
  return drawn;
}

/* Per-permutation separate-analysis statistic (gene.cpp:409-428 for one
 * subgroup, gene.cpp:522-547 over all subgroups): minimum p-value over
 * SNPs (and subgroups).  sbgrp < 0 means all subgroups (--permsep 1). */
extern "C" int orc_perm_sep_stats(const eqbf_chunk* c, int64_t g, int32_t sbgrp, const int32_t* table,
                                  int64_t nperm, double* stats) {
  const int32_t S = c->n_subgroups;
  const int64_t N = c->n_samples;
  int64_t p0 = c->pair_offset[g], p1 = c->pair_offset[g + 1];
  std::vector<double> v(p1 - p0);
  for (int64_t p = 0; p < nperm; ++p) {
    const int32_t* perm = table + p * N;
    for (int64_t q = p0; q < p1; ++q) {
      int64_t snp = c->pair_snp[q];
      if (sbgrp >= 0) {
        double pv = 1.0;
        if (c->gene_in_sbgrp[g * S + sbgrp] && c->snp_in_sbgrp[snp * S + sbgrp]) {
          double o5[5];
          int64_t n = sstats_one(c, g, sbgrp, snp, perm, o5);
          if (n < 0) return EQBF_ERR_DATA;
          pv = n > 0 ? o5[4] : 1.0;
        }
        v[q - p0] = pv;
      } else {
        double mn = 1;
        for (int32_t s = 0; s < S; ++s) {
          if (!(c->gene_in_sbgrp[g * S + s] && c->snp_in_sbgrp[snp * S + s])) continue;
          double o5[5];
          int64_t n = sstats_one(c, g, s, snp, perm, o5);
          if (n < 0) return EQBF_ERR_DATA;
          if (n > 0 && o5[4] < mn) mn = o5[4];
        }
        v[q - p0] = mn;
      }
    }
    stats[p] = min_elem(v);
  }
  return EQBF_OK;
}

/* Counting loop of the separate analysis (gene.cpp:429-449, :549-569):
 * a permutation counts when its minimum p-value is <= the true one. */
extern "C" int64_t orc_perm_sep_count(const double* stats, double true_min, int64_t* nb_permutations,
                                      int32_t trick, int64_t trick_cutoff, orc_mt* rng_trick, double* pval,
                                      int64_t* nperm_used) {
  int64_t sofar = 0, n_all = *nb_permutations, drawn = 0;
  double count = 1;
  bool shuffle_only = false;
  for (int64_t p = 0; p < n_all; ++p) {
    ++drawn;
    if (shuffle_only) continue;
    double st = stats[p];
    if (std::isnan(st)) {
      (*nb_permutations)--;
      continue;
    }
    ++sofar;
    if (st <= true_min) ++count;
    if (trick != 0 && count == 1 + trick_cutoff) {
      if (trick == 1) break;
      shuffle_only = true;
    }
  }
  if (sofar == *nb_permutations)
    *pval = count / (double)(*nb_permutations + 1);
  else if (rng_trick)
    *pval = orc_flat(rng_trick, (1 + trick_cutoff) / ((double)(sofar + 2)),
                     (1 + trick_cutoff) / ((double)(sofar + 1)));
  else
    *pval = kNaN;
  *nperm_used = sofar;
  return drawn;
}
