/*
 * eqbf_oracle_hm.cpp -- CPU oracle for eqtlbma_hm's EM on raw log10 ABFs.
 *
 * TEST INFRASTRUCTURE ONLY (see eqbf_oracle.cpp).
 *
 * Restates, in plain double arithmetic and the reference's loop order:
 *  - snp_eQTL::compute_log10_BF (configs and types), em_update_config,
 *    em_update_type, em_update_subgroup, em_update_grid and
 *    compute_log10_config_BF / compute_log10_type_BF (hm_methods.cpp:74-375);
 *  - gene_eQTL::compute_log10_BF, compute_log10_obs_lik, em_update_pi0,
 *    em_update_config, em_update_type, em_update_subgroup, em_update_grid
 *    and compute_posterior (hm_methods.cpp:392-781), with the SNP prior held
 *    at 1/P (set_snp_prior; the EM computes new SNP priors but never applies
 *    them, update_snp_prior is not called);
 *  - Controller::compute_log10_obs_lik, em_update_pi0/config/type/subgroup/
 *    grid, run_EM_fixedpoint and run_EM_classic (eqtlbma_hm.cpp:617-1107):
 *    every update of an iteration uses the previous iteration's parameters
 *    and gene likelihoods, the loop stops when the log10 likelihood moves by
 *    less than thresh (or at iteration maxit - 1), and one more iteration
 *    follows the stop.
 * log10_weighted_sum keeps the reference's rules (utils_math.cpp:133-159):
 * the maximum starts from the first element, NaN terms are skipped, and a
 * result within DBL_EPSILON of zero is zero.
 *
 * Pinned in tests/test_oracle_hm.py: the EM's fixed point against a direct
 * numerical maximisation of the same likelihood (scipy), the closed-form
 * pi0 of a single-component model, and the likelihood's monotone increase
 * over iterations.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "eqbf_oracle_internal.h"

namespace {

double lws(const double* v, const double* w, int64_t n) { return orc_log10_weighted_sum_w(v, w, n); }

struct HmData {
  const double* raw; /* [pairs][D][L]: D = configs (configs model) or subgroups (types model) */
  const int64_t* off;
  int64_t G;
  int32_t D, L;
  const double* r(int64_t p, int32_t d) const { return raw + ((size_t)p * D + d) * L; }
};

/* log10(q 10^b + 1 - q), the subgroup term of the types model
 * (hm_methods.cpp:143-146) */
inline double sub_term(double q, double b) { return std::log10(std::pow(10.0, std::log10(q) + b) + 1 - q); }

/* configs model ------------------------------------------------------------ */
double snp_bf_cfg(const HmData& d, int64_t p, const double* grid, const double* cfg) {
  std::vector<double> t(d.D);
  for (int32_t k = 0; k < d.D; ++k) t[k] = lws(d.r(p, k), grid, d.L);
  return lws(t.data(), cfg, d.D);
}

/* types model: per type k and grid point l, the sum over subgroups */
double type_grid_sum(const HmData& d, int64_t p, int32_t k, int32_t l, const double* sub, int32_t S, int32_t skip) {
  double s = 0.0;
  for (int32_t t = 0; t < S; ++t)
    if (t != skip) s += sub_term(sub[(size_t)k * S + t], d.r(p, t)[l]);
  return s;
}
double snp_bf_type(const HmData& d, int64_t p, const double* grid, const double* type, const double* sub, int32_t K) {
  const int32_t S = d.D;
  std::vector<double> gt(d.L), dt(K);
  for (int32_t k = 0; k < K; ++k) {
    for (int32_t l = 0; l < d.L; ++l) gt[l] = type_grid_sum(d, p, k, l, sub, S, -1);
    dt[k] = lws(gt.data(), grid, d.L);
  }
  return lws(dt.data(), type, K);
}

struct Params {
  double pi0;
  std::vector<double> grid, cfg, type, sub; /* sub: [K][S] */
};

/* gene log10 BF and log10 observed likelihood (hm_methods.cpp:392-520) */
void gene_lik(const HmData& d, bool types, int32_t K, int64_t g, const Params& P, double* gbf, double* olik,
              std::vector<double>* snpbf) {
  const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
  std::vector<double> sb(np), sw(np, 1.0 / (double)np);
  for (int64_t i = 0; i < np; ++i)
    sb[i] = types ? snp_bf_type(d, p0 + i, P.grid.data(), P.type.data(), P.sub.data(), K)
                  : snp_bf_cfg(d, p0 + i, P.grid.data(), P.cfg.data());
  *gbf = lws(sb.data(), sw.data(), np);
  const double vec[2] = {0.0, *gbf}, wts[2] = {P.pi0, 1 - P.pi0};
  *olik = lws(vec, wts, 2);
  if (snpbf) *snpbf = sb;
}

double total_lik(const HmData& d, bool types, int32_t K, const Params& P, std::vector<double>& olik) {
  double s = 0.0;
  for (int64_t g = 0; g < d.G; ++g) {
    double gbf;
    gene_lik(d, types, K, g, P, &gbf, &olik[g], nullptr);
    s += olik[g];
  }
  return s;
}

/* normalised weights from per-gene log10 contributions (eqtlbma_hm.cpp:
 * 678-722, 816-868): new_k = lws_g(c[k][g], 1) + log10(old_k), divided by
 * their lws */
bool normalise(const std::vector<std::vector<double>>& c, const std::vector<double>& old, std::vector<double>& out) {
  const size_t D = old.size();
  const int64_t G = c.empty() ? 0 : (int64_t)c[0].size();
  std::vector<double> ones(std::max<int64_t>(G, (int64_t)D), 1.0), t(D);
  bool ok = true;
  for (size_t k = 0; k < D; ++k) {
    t[k] = lws(c[k].data(), ones.data(), G) + std::log10(old[k]);
    if (std::isnan(t[k])) ok = false;
  }
  const double den = lws(t.data(), ones.data(), (int64_t)D);
  out.resize(D);
  for (size_t k = 0; k < D; ++k) out[k] = std::pow(10.0, t[k] - den);
  return ok;
}

/* One EM iteration from P with the genes' stored likelihoods olik (those of
 * the last parameters whose likelihood was computed, normally P).  False
 * where the reference throws: a NaN unnormalised configuration weight
 * (eqtlbma_hm.cpp:697-700) or grid weight (eqtlbma_hm.cpp:844-847); the types
 * and subgroup updates have no such check. */
bool em_step(const HmData& d, bool types, int32_t K, const Params& P, const std::vector<double>& olik,
             const int* fixed, Params& Q) {
  const int32_t L = d.L, S = d.D;
  const int64_t G = d.G;
  Q = P;
  bool ok = true;
  /* pi0 (eqtlbma_hm.cpp:659-676, hm_methods.cpp:542-545) */
  if (!fixed[0]) {
    double s = 0.0;
    for (int64_t g = 0; g < G; ++g) s += std::pow(10.0, std::log10(P.pi0) - olik[g]);
    Q.pi0 = s / (double)G;
  }
  std::vector<std::vector<double>> cg(K, std::vector<double>(G)), gg(L, std::vector<double>(G));
  if (!types) {
    /* configs (hm_methods.cpp:200-208, 570-597) */
    if (!fixed[2] && K > 1) {
      for (int64_t g = 0; g < G; ++g) {
        const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
        std::vector<double> sw(np, 1.0 / (double)np), v(np);
        for (int32_t k = 0; k < K; ++k) {
          for (int64_t i = 0; i < np; ++i) v[i] = lws(d.r(p0 + i, k), P.grid.data(), L);
          cg[k][g] = lws(v.data(), sw.data(), np) - olik[g];
        }
      }
      ok &= normalise(cg, P.cfg, Q.cfg);
    }
    /* grid (hm_methods.cpp:261-275, 679-704) */
    if (!fixed[1]) {
      std::vector<double> kv(K);
      for (int64_t g = 0; g < G; ++g) {
        const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
        std::vector<double> sw(np, 1.0 / (double)np), v(np);
        for (int32_t l = 0; l < L; ++l) {
          for (int64_t i = 0; i < np; ++i) {
            for (int32_t k = 0; k < K; ++k) kv[k] = d.r(p0 + i, k)[l];
            v[i] = lws(kv.data(), P.cfg.data(), K);
          }
          gg[l][g] = lws(v.data(), sw.data(), np) - olik[g];
        }
      }
      ok &= normalise(gg, P.grid, Q.grid);
    }
    return ok;
  }
  /* types (hm_methods.cpp:210-259, 277-300, 598-650 and eqtlbma_hm.cpp:724-814) */
  std::vector<double> gt(L), kv(K);
  if (!fixed[3]) {
    for (int64_t g = 0; g < G; ++g) {
      const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
      std::vector<double> sw(np, 1.0 / (double)np), v(np);
      for (int32_t k = 0; k < K; ++k) {
        for (int64_t i = 0; i < np; ++i) {
          for (int32_t l = 0; l < L; ++l) gt[l] = type_grid_sum(d, p0 + i, k, l, P.sub.data(), S, -1);
          v[i] = lws(gt.data(), P.grid.data(), L);
        }
        cg[k][g] = lws(v.data(), sw.data(), np) - olik[g];
      }
    }
    normalise(cg, P.type, Q.type);
  }
  if (!fixed[4]) {
    std::vector<double> den(G), num(G);
    for (int32_t k = 0; k < K; ++k) {
      std::vector<std::vector<double>> ng(S, std::vector<double>(G));
      std::vector<double> dg(G);
      for (int64_t g = 0; g < G; ++g) {
        const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
        std::vector<double> sw(np, 1.0 / (double)np), v(np);
        for (int32_t s = 0; s < S; ++s) {
          for (int64_t i = 0; i < np; ++i) {
            for (int32_t l = 0; l < L; ++l)
              gt[l] = d.r(p0 + i, s)[l] + type_grid_sum(d, p0 + i, k, l, P.sub.data(), S, s);
            v[i] = lws(gt.data(), P.grid.data(), L);
          }
          ng[s][g] = lws(v.data(), sw.data(), np) - olik[g];
        }
        for (int64_t i = 0; i < np; ++i) {
          for (int32_t l = 0; l < L; ++l) gt[l] = type_grid_sum(d, p0 + i, k, l, P.sub.data(), S, -1);
          v[i] = lws(gt.data(), P.grid.data(), L);
        }
        dg[g] = lws(v.data(), sw.data(), np) - olik[g];
      }
      std::vector<double> ones(G, 1.0);
      const double l10den = lws(dg.data(), ones.data(), G);
      for (int32_t s = 0; s < S; ++s)
        Q.sub[(size_t)k * S + s] =
            std::pow(10.0, lws(ng[s].data(), ones.data(), G) + std::log10(P.sub[(size_t)k * S + s]) - l10den);
    }
  }
  if (!fixed[1]) {
    for (int64_t g = 0; g < G; ++g) {
      const int64_t p0 = d.off[g], np = d.off[g + 1] - p0;
      std::vector<double> sw(np, 1.0 / (double)np), v(np);
      for (int32_t l = 0; l < L; ++l) {
        for (int64_t i = 0; i < np; ++i) {
          for (int32_t k = 0; k < K; ++k) kv[k] = type_grid_sum(d, p0 + i, k, l, P.sub.data(), S, -1);
          v[i] = lws(kv.data(), P.type.data(), K);
        }
        gg[l][g] = lws(v.data(), sw.data(), np) - olik[g];
      }
    }
    ok &= normalise(gg, P.grid, Q.grid);
  }
  return ok;
}

Params load_params(bool types, int32_t K, int32_t S, int32_t L, const double* pi0, const double* grid,
                   const double* cfg, const double* type, const double* sub) {
  Params P;
  P.pi0 = *pi0;
  P.grid.assign(grid, grid + L);
  if (!types) {
    P.cfg.assign(cfg, cfg + K);
  } else {
    P.type.assign(type, type + K);
    P.sub.assign(sub, sub + (size_t)K * S);
  }
  return P;
}

void store_params(bool types, const Params& P, double* pi0, double* grid, double* cfg, double* type, double* sub) {
  *pi0 = P.pi0;
  std::copy(P.grid.begin(), P.grid.end(), grid);
  if (!types) {
    std::copy(P.cfg.begin(), P.cfg.end(), cfg);
  } else {
    std::copy(P.type.begin(), P.type.end(), type);
    std::copy(P.sub.begin(), P.sub.end(), sub);
  }
}

}  // namespace

namespace {

/* The EM state of the reference's Controller: parameters, their log10
 * likelihood and the genes' stored log10 likelihoods (compute_log10_obs_lik
 * with keep = true), which every update reads. */
struct EmRun {
  HmData d;
  bool types;
  int32_t K;
  int fx[5];
  double thresh;
  int64_t maxit;
  std::vector<double> olik;
  double* trace;
  int64_t max_trace;

  /* compute_log10_obs_lik(P, keep = true); false where it throws (NaN or
   * infinite sum, eqtlbma_hm.cpp:640-647) */
  bool lik(const Params& P, double* out) {
    *out = total_lik(d, types, K, P, olik);
    return std::isfinite(*out);
  }
  /* run_EM_fixedpoint (eqtlbma_hm.cpp:925-1011): P -> Q and lik(Q) */
  bool fixedpoint(int64_t iter, const Params& P, Params& Q, double* nl) {
    if (!em_step(d, types, K, P, olik, fx, Q)) return false;
    const bool ok = lik(Q, nl);
    if (iter < max_trace) trace[iter] = *nl;
    return ok;
  }
  bool stop(double a, double b, int64_t iter) const {
    return std::fabs(a - b) < thresh || (maxit > 0 && iter == maxit - 1);
  }
};

double sq(double x) { return x * x; }

/* accumulate |r|^2 and |v|^2 of one parameter vector (compute_steplength) */
void sq_acc(const std::vector<double>& a0, const std::vector<double>& a1, const std::vector<double>& a2, double* sr2,
            double* sv2) {
  for (size_t k = 0; k < a0.size(); ++k) {
    *sr2 += sq(a1[k] - a0[k]);
    *sv2 += sq(a2[k] - 2 * a1[k] + a0[k]);
  }
}
void sq_prop(const std::vector<double>& a0, const std::vector<double>& a1, const std::vector<double>& a2, double alpha,
             std::vector<double>& o) {
  for (size_t k = 0; k < a0.size(); ++k)
    o[k] = a0[k] + 2.0 * alpha * (a1[k] - a0[k]) + std::pow(alpha, 2) * (a2[k] - 2.0 * a1[k] + a0[k]);
}

/* run_EM_square (eqtlbma_hm.cpp:1215-1324) with the reference's constants:
 * step length rule 3 (alpha = sqrt(|r|^2 / |v|^2), Varadhan and Roland
 * 2008), stepmin0 = stepmax0 = 1, mstep = 4, maxdist = 1.  When the
 * extrapolation is rejected for being too far below the start, the
 * parameters go back to the second EM iterate but the genes keep the stored
 * likelihoods of the last evaluation, which the next update reads. */
int64_t run_square(EmRun& R, Params& P, double stepmax_user) {
  const double stepmin0 = 1, stepmax0 = 1, mstep = 4, maxdist = 1;
  const bool dimv = R.types || R.K > 1;
  double steplen = 1.0, stepmin = stepmin0, stepmax = stepmax0, lik;
  int64_t iter = 0;
  if (!R.lik(P, &lik)) return -2;
  if (R.max_trace > 0) R.trace[0] = lik;
  Params Q;
  while (true) {
    const double lik0 = lik;
    const Params P0 = P;
    double nl;
    ++iter;
    if (!R.fixedpoint(iter, P, Q, &nl)) return -2;
    P = Q;
    lik = nl;
    const Params P1 = P;
    const double lik1 = lik;
    if (R.stop(lik1, lik0, iter)) break;
    ++iter;
    if (!R.fixedpoint(iter, P, Q, &nl)) return -2;
    P = Q;
    lik = nl;
    const Params P2 = P;
    const double lik2 = lik;
    if (R.stop(lik2, lik1, iter)) break;
    double sr2 = sq(P1.pi0 - P0.pi0), sv2 = sq(P2.pi0 - 2 * P1.pi0 + P0.pi0);
    if (dimv) {
      if (R.types) {
        sq_acc(P0.type, P1.type, P2.type, &sr2, &sv2);
        sq_acc(P0.sub, P1.sub, P2.sub, &sr2, &sv2);
      } else {
        sq_acc(P0.cfg, P1.cfg, P2.cfg, &sr2, &sv2);
      }
    }
    sq_acc(P0.grid, P1.grid, P2.grid, &sr2, &sv2);
    double alpha = std::sqrt(sr2 / sv2);
    alpha = std::max(stepmin, std::min(stepmax, alpha));
    alpha = std::min(stepmax_user, alpha);
    steplen = alpha;
    Params Pn = P2;
    Pn.pi0 = P0.pi0 + 2.0 * alpha * (P1.pi0 - P0.pi0) + std::pow(alpha, 2) * (P2.pi0 - 2.0 * P1.pi0 + P0.pi0);
    if (dimv) {
      if (R.types) {
        sq_prop(P0.type, P1.type, P2.type, alpha, Pn.type);
        sq_prop(P0.sub, P1.sub, P2.sub, alpha, Pn.sub);
      } else {
        sq_prop(P0.cfg, P1.cfg, P2.cfg, alpha, Pn.cfg);
      }
    }
    sq_prop(P0.grid, P1.grid, P2.grid, alpha, Pn.grid);
    if (!R.lik(Pn, &nl)) return -2;
    P = Pn;
    lik = nl;
    bool extrap = true;
    if (std::fabs(steplen - 1) > 0.01) {
      ++iter;
      if (!R.fixedpoint(iter, P, Q, &nl)) {
        P = P2;
        if (!R.lik(P, &lik)) return -2;
        extrap = false;
        if (steplen == stepmax) stepmax = std::max(stepmax0, stepmax / mstep);
        steplen = 1;
        if (steplen == stepmax) stepmax = mstep * stepmax;
        if (stepmin < 0 && steplen == stepmin) stepmin = mstep * stepmin;
        continue;
      }
      P = Q;
      lik = nl;
    }
    if (extrap && lik < lik0 - maxdist) {
      lik = lik2;
      P = P2;
      if (steplen == stepmax) stepmax = std::max(stepmax0, stepmax / mstep);
      steplen = 1;
    }
    if (steplen == stepmax) stepmax = mstep * stepmax;
    if (stepmin < 0 && steplen == stepmin) stepmin = mstep * stepmin;
  }
  ++iter;
  double nl;
  if (!R.fixedpoint(iter, P, Q, &nl)) return -2;
  P = Q;
  return iter;
}

/* run_EM_classic (eqtlbma_hm.cpp:1076-1107) */
int64_t run_classic(EmRun& R, Params& P) {
  double lik;
  if (!R.lik(P, &lik)) return -2;
  if (R.max_trace > 0) R.trace[0] = lik;
  int64_t iter = 0;
  Params Q;
  while (true) {
    ++iter;
    double nl;
    if (!R.fixedpoint(iter, P, Q, &nl)) return -2;
    if (nl < lik) return -1;
    const bool stop = R.stop(nl, lik, iter);
    P = Q;
    lik = nl;
    if (stop) break;
  }
  ++iter;
  double nl;
  if (!R.fixedpoint(iter, P, Q, &nl)) return -2;
  P = Q;
  return iter;
}

}  // namespace

/* EM of eqtlbma_hm (Controller::run_EM, eqtlbma_hm.cpp:1326-1346): classic
 * when stepmax == 1, SQUAREM otherwise.
 * raw: [n_pairs][D][L] with D = K configurations (types = 0) or D = S
 * subgroups (types = 1, K types).  fixed[5]: pi0, grid, configs, types,
 * subgroups-per-type.  Parameters are in/out; lik_trace[i] is the log10
 * likelihood after iteration i (i = 0: the initial value; SQUAREM's
 * extrapolations are not iterations).  maxit <= 0: no limit.  Returns the
 * number of iterations, -1 if the likelihood decreased under classic EM (the
 * reference exits with an error), -2 where the reference throws (a NaN or
 * infinite log10 likelihood, a NaN unnormalised weight). */
extern "C" int64_t orc_hm_em(const double* raw, const int64_t* off, int64_t G, int32_t types, int32_t K, int32_t S,
                             int32_t L, double thresh, int64_t maxit, double stepmax, const int32_t* fixed,
                             double* pi0, double* grid, double* cfg, double* type, double* sub, double* lik_trace,
                             int64_t max_trace) {
  EmRun R{HmData{raw, off, G, types ? S : K, L}, types != 0, K, {fixed[0], fixed[1], fixed[2], fixed[3], fixed[4]},
          thresh, maxit, std::vector<double>(G), lik_trace, max_trace};
  Params P = load_params(types, K, S, L, pi0, grid, cfg, type, sub);
  const int64_t it = stepmax == 1.0 ? run_classic(R, P) : run_square(R, P, stepmax);
  if (it >= 0) store_params(types, P, pi0, grid, cfg, type, sub);
  return it;
}

/* log10 likelihood of a parameter set (Controller::compute_log10_obs_lik). */
extern "C" double orc_hm_loglik(const double* raw, const int64_t* off, int64_t G, int32_t types, int32_t K, int32_t S,
                                int32_t L, const double* pi0, const double* grid, const double* cfg, const double* type,
                                const double* sub) {
  HmData d{raw, off, G, types ? S : K, L};
  Params P = load_params(types, K, S, L, pi0, grid, cfg, type, sub);
  std::vector<double> olik(G);
  return total_lik(d, types, K, P, olik);
}

/* Outputs of --getbf (gene_eQTL::compute_posterior, hm_methods.cpp:752-781,
 * and Controller::save_result, eqtlbma_hm.cpp:1683-1740): per gene the
 * posterior probability min(1, 10^(log10(1 - pi0) + BF - obslik)) and the
 * log10 BF, per SNP the log10 BF and per SNP and configuration (or type)
 * the log10 BF averaged over the grid. */
extern "C" void orc_hm_bfs(const double* raw, const int64_t* off, int64_t G, int32_t types, int32_t K, int32_t S,
                           int32_t L, const double* pi0, const double* grid, const double* cfg, const double* type,
                           const double* sub, double* gene_post, double* gene_bf, double* snp_bf, double* dim_bf) {
  HmData d{raw, off, G, types ? S : K, L};
  Params P = load_params(types, K, S, L, pi0, grid, cfg, type, sub);
  std::vector<double> sb, gt(L);
  for (int64_t g = 0; g < G; ++g) {
    double gbf, ol;
    gene_lik(d, types, K, g, P, &gbf, &ol, &sb);
    double pp = std::pow(10.0, std::log10(1.0 - P.pi0) + gbf - ol);
    gene_post[g] = pp > 1.0 ? 1.0 : pp;
    gene_bf[g] = gbf;
    for (int64_t i = 0; i < d.off[g + 1] - d.off[g]; ++i) {
      const int64_t p = d.off[g] + i;
      snp_bf[p] = sb[i];
      for (int32_t k = 0; k < K; ++k) {
        if (!types) {
          dim_bf[(size_t)p * K + k] = lws(d.r(p, k), P.grid.data(), L);
        } else {
          for (int32_t l = 0; l < L; ++l) gt[l] = type_grid_sum(d, p, k, l, P.sub.data(), S, -1);
          dim_bf[(size_t)p * K + k] = lws(gt.data(), P.grid.data(), L);
        }
      }
    }
  }
}

/* Profile-likelihood bounds of --getci (Controller::estimate_profile_ci and
 * estimate_profile_ci_pi0/configs/types/subgroups/grids,
 * eqtlbma_hm.cpp:1348-1546), at the given parameters: steps of 0.001 while
 * log10 lik / log10(e) >= max / log10(e) - 2.  left/right: pi0, configs or
 * types, subgroups per type (types model), grid points.  Returns 0, or -2
 * where the reference throws (a NaN or infinite log10 likelihood). */
extern "C" int32_t orc_hm_ci(const double* raw, const int64_t* off, int64_t G, int32_t types, int32_t K, int32_t S,
                             int32_t L, const double* pi0, const double* grid, const double* cfg, const double* type,
                             const double* sub, double* left, double* right) {
  HmData d{raw, off, G, types ? S : K, L};
  const Params P = load_params(types, K, S, L, pi0, grid, cfg, type, sub);
  std::vector<double> olik(G);
  const double tick = 0.001, le = std::log10(std::exp(1.0));
  const double mx = total_lik(d, types, K, P, olik);
  if (!std::isfinite(mx)) return -2;
  Params Q = P;
  bool bad = false;
  auto below = [&]() {
    const double v = total_lik(d, types, K, Q, olik);
    if (!std::isfinite(v)) bad = true;
    return v / le < mx / le - 2.0;
  };
  /* pi0 (estimate_profile_ci_pi0) */
  double l = P.pi0, r = P.pi0;
  while (l >= 0 && !bad) {
    l -= tick;
    if (l < 0) {
      l = 0;
      break;
    }
    Q.pi0 = l;
    if (below()) {
      l += tick;
      break;
    }
  }
  while (r <= 1 && !bad) {
    r += tick;
    if (r > 1) {
      r = 1;
      break;
    }
    Q.pi0 = r;
    if (below()) {
      r -= tick;
      break;
    }
  }
  Q.pi0 = P.pi0;
  left[0] = l;
  right[0] = r;
  /* configs and grids: one weight moves, the others share the difference
   * in proportion to their estimate */
  auto walk = [&](std::vector<double>& w, const std::vector<double>& mle, double* lo, double* hi) {
    for (size_t i = 0; i < mle.size() && !bad; ++i) {
      const double cp = mle[i], st = 1 - cp;
      double a = cp, b = cp;
      auto set = [&](double x) {
        for (size_t j = 0; j < mle.size(); ++j)
          if (j != i) w[j] = mle[j] + (cp - x) * mle[j] / st;
        w[i] = x;
      };
      while (a >= 0 && !bad) {
        a -= tick;
        if (a < 0) {
          a = 0;
          break;
        }
        set(a);
        if (below()) {
          a += tick;
          break;
        }
      }
      while (b <= 1 && !bad) {
        b += tick;
        if (b > 1) {
          b = 1;
          break;
        }
        set(b);
        if (below()) {
          b -= tick;
          break;
        }
      }
      lo[i] = a;
      hi[i] = b;
    }
    w = mle;
  };
  if (!types) {
    walk(Q.cfg, P.cfg, left + 1, right + 1);
  } else {
    for (int32_t k = 0; k < K; ++k) left[1 + k] = right[1 + k] = P.type[k];
    for (int32_t i = 0; i < K * S; ++i) left[1 + K + i] = right[1 + K + i] = P.sub[i];
  }
  const int32_t g0 = 1 + K + (types ? K * S : 0);
  walk(Q.grid, P.grid, left + g0, right + g0);
  return bad ? -2 : 0;
}
