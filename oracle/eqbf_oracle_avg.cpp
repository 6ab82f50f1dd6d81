/*
 * CPU oracle of eqtlbma_avg_bfs (test infrastructure only: the tests, the
 * smoke check and the benchmark's CPU baseline call it; the product never
 * does).  Restates src/eqtlbma_avg_bfs.cpp in its order of operations:
 *
 *   orc_avg_grid  BayesFactor::avg_raw_bfs (:371-382): one line of the raw
 *                 file averaged over the kept grid points with the grid
 *                 weights (the output read by eqtlbma_hm).
 *   orc_avg_bfs   per gene: Snp::avg_raw_bfs for the configs (:418-445) or
 *                 types (:447-475) model, Gene::avg_raw_bfs (:598-622), and
 *                 the posteriors calc_posterior (:634-643),
 *                 calc_cond_snp_posteriors_the_eQTL (:647-662), _an_eQTL
 *                 (:664-680), _config (:683-699), _type (:702-718),
 *                 _subgroup for configs (:722-745) and types (:749-796).
 *
 * Raw values: [n_pairs][R][L], R = dim (configs: the j-th configuration
 * row read for the SNP) or n_subgroups (types: subgroup-specific rows).
 * Quirks kept: the j-th raw row of a SNP goes with the j-th configuration
 * weight whatever the names (the reference pairs them by position); the
 * best dimension of the averaging step is the first strict maximum of the
 * per-dimension BFs and is replaced by that of the posteriors when those
 * are computed; post_an is never capped (the reference tests post_the
 * instead, :677-678).
 * Pinned in tests/test_oracle_avg.py: the log10 sums against numpy with
 * explicit weights, the nested grid-then-configuration average against the
 * flat double sum, the posteriors' normalisations (post_the sums to 1 over
 * a gene's SNPs, post_dim to 1 over the dimensions, gene_post from Bayes'
 * rule), the types model with subgroup weights 1 against the product of the
 * subgroup BFs, and the subgroup posteriors against sums over configurations.
 */
#include <cmath>
#include <cstdint>
#include <vector>

#include "eqbf_oracle_internal.h"

namespace {

double lws_w(const std::vector<double>& v, const std::vector<double>& w) {
  return orc_log10_weighted_sum_w(v.data(), w.data(), (int64_t)v.size());
}

}  // namespace

extern "C" void orc_avg_grid(const double* raw, int64_t n_rows, int32_t L, const double* gw, const int32_t* keep,
                             int32_t nk, double* out) {
  std::vector<double> v(nk), w(nk);
  for (int64_t r = 0; r < n_rows; ++r) {
    for (int32_t i = 0; i < nk; ++i) {
      v[i] = raw[r * L + keep[i]];
      w[i] = gw[keep[i]];
    }
    out[r] = lws_w(v, w);
  }
}

/* want: bit 0 gene posterior (a), 1 post_the (b), 2 post_an (c), 3 subgroup
 * posteriors (d), 4 dimension posteriors (--bestdim / --alldim) */
extern "C" void orc_avg_bfs(const double* raw, const int64_t* off, int64_t G, int32_t types, int32_t D, int32_t S,
                            int32_t L, const int32_t* keep, int32_t nk, const double* gw, const double* cw,
                            const uint64_t* csub, const double* tw, const double* sw, double pi0, int32_t want,
                            double* gene_bf, double* gene_post, double* snp_bf, double* dim_bf, double* post_the,
                            double* post_an, double* post_dim, double* post_sub, int32_t* best_dim) {
  const int32_t R = types ? S : D;
  std::vector<double> gwk(nk), v(nk), dims(D), dw(types ? tw : cw, (types ? tw : cw) + D);
  for (int32_t i = 0; i < nk; ++i) gwk[i] = gw[keep[i]];
  auto rawv = [&](int64_t p, int32_t r, int32_t i) { return raw[((size_t)p * R + r) * L + keep[i]]; };
  for (int64_t g = 0; g < G; ++g) {
    const int64_t p0 = off[g], m = off[g + 1] - off[g];
    std::vector<double> sb(m);
    for (int64_t q = 0; q < m; ++q) {
      const int64_t p = p0 + q;
      int32_t best = -1;
      for (int32_t k = 0; k < D; ++k) {
        if (!types) {
          for (int32_t i = 0; i < nk; ++i) v[i] = rawv(p, k, i);
        } else {
          for (int32_t i = 0; i < nk; ++i) {
            v[i] = 0.0;
            for (int32_t s = 0; s < S; ++s) {
              const double w = sw[(size_t)k * S + s];
              v[i] += std::log10(w * std::pow(10, rawv(p, s, i)) + 1 - w);
            }
          }
        }
        dims[k] = lws_w(v, gwk);
        if (!types && (best < 0 || dims[k] > dims[best])) best = k;
        dim_bf[(size_t)p * D + k] = dims[k];
      }
      sb[q] = lws_w(dims, dw);
      snp_bf[p] = sb[q];
      best_dim[p] = best;
    }
    gene_bf[g] = m > 0 ? orc_log10_weighted_sum(sb.data(), m) : NAN;
    if (want & 1) {
      const std::vector<double> vec{0.0, gene_bf[g]}, wts{pi0, 1 - pi0};
      const double ol = lws_w(vec, wts);
      double pp = std::pow(10, std::log10(1.0 - pi0) + gene_bf[g] - ol);
      gene_post[g] = pp > 1.0 ? 1.0 : pp;
    }
    if (want & 2) {
      const std::vector<double> ones(m, 1.0);
      const double tot = lws_w(sb, ones);
      for (int64_t q = 0; q < m; ++q) {
        double pt = std::pow(10, sb[q] - tot);
        post_the[p0 + q] = pt > 1.0 ? 1.0 : pt;
      }
    }
    if (want & 4) {
      const std::vector<double> wts{1.0 - 1 / (double)m, 1 / (double)m};
      for (int64_t q = 0; q < m; ++q) {
        const std::vector<double> vec{0.0, sb[q]};
        post_an[p0 + q] = std::pow(10, -std::log10((double)m) + sb[q] - lws_w(vec, wts));
      }
    }
    const bool dims_post = (want & 16) || (!types && (want & 8));
    if (dims_post)
      for (int64_t q = 0; q < m; ++q) {
        const int64_t p = p0 + q;
        int32_t best = 0;
        for (int32_t k = 0; k < D; ++k) {
          double pd = std::pow(10, std::log10(dw[k]) + dim_bf[(size_t)p * D + k] - sb[q]);
          if (pd > 1.0) pd = 1.0;
          post_dim[(size_t)p * D + k] = pd;
          if (k == 0 || pd > post_dim[(size_t)p * D + best]) best = k;
        }
        best_dim[p] = best;
      }
    if (want & 8)
      for (int64_t q = 0; q < m; ++q) {
        const int64_t p = p0 + q;
        for (int32_t s = 0; s < S; ++s) {
          double acc = 0.0;
          if (!types) {
            for (int32_t j = 0; j < D; ++j)
              if ((csub[j] >> s) & 1ull) acc += post_dim[(size_t)p * D + j];
          } else {
            for (int32_t k = 0; k < D; ++k) {
              double t = std::log10(tw[k]) + std::log10(sw[(size_t)k * S + s]) - sb[q];
              for (int32_t i = 0; i < nk; ++i) {
                v[i] = rawv(p, s, i);
                for (int32_t s2 = 0; s2 < S; ++s2)
                  if (s2 != s) {
                    const double w = sw[(size_t)k * S + s2];
                    v[i] += std::log10(w * std::pow(10, rawv(p, s2, i)) + 1 - w);
                  }
              }
              t += lws_w(v, gwk);
              acc += std::pow(10, t);
            }
          }
          post_sub[(size_t)p * S + s] = acc > 1.0 ? 1.0 : acc;
        }
      }
  }
}
