/*
 * eqbf_oracle.cpp -- CPU oracle for the eqtlbma_bf Bayes-factor engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark use it, as the checker.
 *
 * It restates, in plain C++ with FP64 outputs (long double where a
 * factorisation or a special function benefits from it), the reference
 * algorithm of eqtlbma_bf 1.3.3 for the hot path, following the reference's
 * loop structure and operation order.  Each function names the reference lines
 * it follows.  How each part is pinned (known-answer values, closed forms,
 * scipy/mpmath, brute force) is stated in tests/test_oracle_*.py; parts with no
 * such check are marked "parity unpinned" below.
 *
 * Outputs use the structure-of-arrays layout of include/eqbf.h so that tests
 * compare the GPU library and the oracle element by element.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/eqbf.h"
#include "eqbf_oracle_internal.h"

using namespace orc;
static const double kNaN = std::numeric_limits<double>::quiet_NaN();
static const double kInf = std::numeric_limits<double>::infinity();

char orc::g_err[512];
extern "C" const char* orc_last_error(void) { return orc::g_err; }

/* ======================================================================== */
/* 1. Random numbers: GSL mt19937 + gsl_rng_uniform_int + gsl_ran_shuffle    */
/*    (GSL is not vendored; call sites eqtlbma_bf.cpp:889-916,               */
/*    gene.cpp:410,523,639,358-362).  The raw MT19937 stream is pinned by the */
/*    C++ standard's known answer for std::mt19937; uniform_int and shuffle   */
/*    follow GSL's published semantics -- parity unpinned (no GSL here).      */
/* ======================================================================== */
struct orc_mt {
  uint32_t s[624];
  int32_t idx;
};

/* gsl_rng_set(r, seed) for mt19937: seed 0 maps to 4357, then Knuth's
 * initialisation with multiplier 1812433253, state index reset. */
extern "C" void orc_mt_seed(orc_mt* st, uint64_t seed) {
  if (seed == 0) seed = 4357;
  st->s[0] = (uint32_t)(seed & 0xffffffffULL);
  for (int k = 1; k < 624; ++k) {
    uint32_t prev = st->s[k - 1];
    st->s[k] = (uint32_t)(1812433253u * (prev ^ (prev >> 30)) + (uint32_t)k);
  }
  st->idx = 624;
}

/* Matsumoto & Nishimura MT19937 (n=624, m=397), 32-bit output. */
extern "C" uint32_t orc_mt_next(orc_mt* st) {
  if (st->idx >= 624) {
    for (int k = 0; k < 624; ++k) {
      uint32_t hi = st->s[k] & 0x80000000u;
      uint32_t lo = st->s[(k + 1) % 624] & 0x7fffffffu;
      uint32_t v = hi | lo;
      uint32_t nxt = st->s[(k + 397) % 624] ^ (v >> 1);
      if (v & 1u) nxt ^= 0x9908b0dfu;
      st->s[k] = nxt;
    }
    st->idx = 0;
  }
  uint32_t y = st->s[st->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* gsl_rng_uniform_int for a generator with min 0 and max 2^32-1: integer
 * bucket width range/n, reject draws past the last full bucket. */
extern "C" uint32_t orc_uniform_int(orc_mt* st, uint32_t n) {
  const uint32_t range = 0xffffffffu;
  uint32_t width = range / n;
  uint32_t k;
  do {
    k = orc_mt_next(st) / width;
  } while (k >= n);
  return k;
}

/* gsl_ran_shuffle: for i = n-1 down to 1 swap i with uniform_int(i+1). */
extern "C" void orc_shuffle(orc_mt* st, int64_t* a, int64_t n) {
  for (int64_t i = n - 1; i > 0; --i) {
    int64_t j = (int64_t)orc_uniform_int(st, (uint32_t)(i + 1));
    int64_t t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
}

/* gsl_ran_flat(r, a, b) = a(1-u) + b u with u = next()/2^32. */
extern "C" double orc_flat(orc_mt* st, double a, double b) {
  double u = orc_mt_next(st) / 4294967296.0;
  return a * (1 - u) + b * u;
}

/* Index tables of a chunk (makePermutationsJoin, eqtlbma_bf.cpp:846-863 and
 * Gene::MakePermutationsJoin, gene.cpp:617-641): the stream is seeded once per
 * chunk; each analysed gene starts from the identity and takes nperm
 * successive (cumulative) shuffles.  table[k][p][i] = perm_p(i) for the k-th
 * analysed gene of the chunk. */
extern "C" void orc_perm_tables(uint64_t seed, int64_t n_samples, int64_t n_genes,
                                int64_t nperm, int32_t* table) {
  orc_mt st;
  orc_mt_seed(&st, seed);
  std::vector<int64_t> perm(n_samples);
  for (int64_t k = 0; k < n_genes; ++k) {
    for (int64_t i = 0; i < n_samples; ++i) perm[i] = i;
    for (int64_t p = 0; p < nperm; ++p) {
      orc_shuffle(&st, perm.data(), n_samples);
      int32_t* row = table + (k * nperm + p) * n_samples;
      for (int64_t i = 0; i < n_samples; ++i) row[i] = (int32_t)perm[i];
    }
  }
}

/* ======================================================================== */
/* 2. Special functions (GSL gsl_cdf_tdist_P/Q, gsl_cdf_gaussian_Pinv,       */
/*    gsl_cdf_fdist_Q, gsl_cdf_chisq_Qinv at utils_math.cpp:202,             */
/*    gene_snp_pair.cpp:274, MVLR.cpp:404-405).  Exact mathematics, pinned   */
/*    against scipy/mpmath; GSL's own approximation branches are not         */
/*    reproduced (parity to GSL bits unpinned).                              */
/* ======================================================================== */

/* Continued fraction of the incomplete beta (modified Lentz). */
static ldouble betacf(ldouble a, ldouble b, ldouble x) {
  const ldouble tiny = 1e-4000L;
  const ldouble eps = 1e-19L;
  ldouble qab = a + b, qap = a + 1, qam = a - 1;
  ldouble c = 1, d = 1 - qab * x / qap;
  if (fabsl(d) < tiny) d = tiny;
  d = 1 / d;
  ldouble h = d;
  for (int m = 1; m <= 100000; ++m) {
    int m2 = 2 * m;
    ldouble aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1 + aa * d;
    if (fabsl(d) < tiny) d = tiny;
    c = 1 + aa / c;
    if (fabsl(c) < tiny) c = tiny;
    d = 1 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1 + aa * d;
    if (fabsl(d) < tiny) d = tiny;
    c = 1 + aa / c;
    if (fabsl(c) < tiny) c = tiny;
    d = 1 / d;
    ldouble del = d * c;
    h *= del;
    if (fabsl(del - 1) < eps) break;
  }
  return h;
}

/* Regularised incomplete beta I_x(a,b) given x and y = 1-x computed
 * separately (avoids cancellation in 1-x). */
static ldouble ibeta_xy(ldouble a, ldouble b, ldouble x, ldouble y) {
  if (x <= 0) return 0;
  if (y <= 0) return 1;
  ldouble lfront = lgammal(a + b) - lgammal(a) - lgammal(b) + a * logl(x) + b * logl(y);
  if (x < (a + 1) / (a + b + 2)) return expl(lfront) * betacf(a, b, x) / a;
  return 1 - expl(lfront) * betacf(b, a, y) / b;
}

/* Lower tail P(T_nu <= -|t|) of Student's t, long double. */
ldouble orc::tdist_lower_neg(ldouble at, ldouble nu) {
  if (std::isnan((double)at) || std::isnan((double)nu)) return NAN;
  if (std::isinf((double)at)) return 0;
  ldouble t2 = at * at;
  ldouble x = nu / (nu + t2), y = t2 / (nu + t2);
  return 0.5L * ibeta_xy(nu / 2, 0.5L, x, y);
}

extern "C" double orc_tdist_P(double t, double nu) {
  if (std::isnan(t) || std::isnan(nu)) return kNaN;
  ldouble lo = tdist_lower_neg(fabsl((ldouble)t), nu);
  return (double)(t < 0 ? lo : 1 - lo);
}
extern "C" double orc_tdist_Q(double t, double nu) {
  if (std::isnan(t) || std::isnan(nu)) return kNaN;
  ldouble lo = tdist_lower_neg(fabsl((ldouble)t), nu);
  return (double)(t > 0 ? lo : 1 - lo);
}

/* Inverse standard normal CDF: Wichura's AS241 (Applied Statistics 37,
 * 1988) rational approximations, then one Halley step in long double. */
static ldouble as241(ldouble p) {
  ldouble q = p - 0.5L, r, val;
  if (fabsl(q) <= 0.425L) {
    r = 0.180625L - q * q;
    val = q *
          (((((((r * 2509.0809287301226727L + 33430.575583588128105L) * r + 67265.770927008700853L) * r +
               45921.953931549871457L) * r + 13731.693765509461125L) * r + 1971.5909503065514427L) * r +
            133.14166789178437745L) * r + 3.387132872796366608L) /
          (((((((r * 5226.495278852545561L + 28729.085735721942674L) * r + 39307.89580009271061L) * r +
               21213.794301586595867L) * r + 5394.1960214247511077L) * r + 687.1870074920579083L) * r +
            42.313330701600911252L) * r + 1.0L);
    return val;
  }
  r = (q < 0) ? p : 1 - p;
  if (r <= 0) return q < 0 ? -INFINITY : INFINITY;
  r = sqrtl(-logl(r));
  if (r <= 5) {
    r -= 1.6L;
    val = (((((((r * 7.7454501427834140764e-4L + 0.0227238449892691845833L) * r + 0.24178072517745061177L) * r +
               1.27045825245236838258L) * r + 3.64784832476320460504L) * r + 5.7694972214606914055L) * r +
            4.6303378461565452959L) * r + 1.42343711074968357734L) /
          (((((((r * 1.05075007164441684324e-9L + 5.475938084995344946e-4L) * r + 0.0151986665636164571966L) * r +
               0.14810397642748007459L) * r + 0.68976733498510000455L) * r + 1.6763848301838038494L) * r +
            2.05319162663775882187L) * r + 1.0L);
  } else {
    r -= 5;
    val = (((((((r * 2.01033439929228813265e-7L + 2.71155556874348757815e-5L) * r + 0.0012426609473880784386L) * r +
               0.026532189526576123093L) * r + 0.29656057182850489123L) * r + 1.7848265399172913358L) * r +
            5.4637849111641143699L) * r + 6.6579046435011037772L) /
          (((((((r * 2.04426310338993978564e-15L + 1.4215117583164458887e-7L) * r + 1.8463183175100546818e-5L) * r +
               7.868691311456132591e-4L) * r + 0.0148753612908506148525L) * r + 0.13692988092273580531L) * r +
            0.59983220655588793769L) * r + 1.0L);
  }
  return q < 0 ? -val : val;
}

ldouble orc::gauss_pinv_ld(ldouble p) {
  if (std::isnan((double)p)) return NAN;
  if (p <= 0) return -INFINITY;
  if (p >= 1) return INFINITY;
  ldouble z = as241(p);
  if (std::isinf((double)z)) return z;
  /* Halley refinement on Phi(z) - p. */
  const ldouble sqrt2 = 1.41421356237309504880168872420969808L;
  const ldouble sqrt2pi = 2.50662827463100050241576528481104525L;
  for (int it = 0; it < 2; ++it) {
    ldouble cdf = (z < 0) ? 0.5L * erfcl(-z / sqrt2) : 1 - 0.5L * erfcl(z / sqrt2);
    ldouble e = cdf - p;
    ldouble u = e * sqrt2pi * expl(z * z / 2);
    z = z - u / (1 + z * u / 2);
  }
  return z;
}

extern "C" double orc_gauss_Pinv(double p) { return (double)gauss_pinv_ld(p); }

/* gsl_cdf_fdist_Q(F; 1, nu) = 2 T_nu.Q(sqrt F). */
extern "C" double orc_fdist_Q_1(double F, double nu) {
  if (std::isnan(F)) return kNaN;
  if (F <= 0) return 1.0;
  return (double)(2 * tdist_lower_neg(sqrtl((ldouble)F), nu));
}
/* gsl_cdf_chisq_Qinv(q; 1) = Phi^-1(q/2)^2. */
extern "C" double orc_chisq_Qinv_1(double q) {
  if (std::isnan(q)) return kNaN;
  ldouble z = gauss_pinv_ld((ldouble)q / 2);
  return (double)(z * z);
}

/* ======================================================================== */
/* 3. OLS summary statistics (utils::FitSingleGeneWithSingleSnp,             */
/*    utils_math.cpp:166-209).  Householder QR in long double on the columns */
/*    [1, covariates, genotype] (the fit does not depend on column order).   */
/*    Full-rank results are pinned by closed forms; the rank-deficient       */
/*    output of GSL's SVD is GSL-specific -- parity unpinned: this oracle    */
/*    returns NaN for betahat, sebetahat and betapval when the genotype is   */
/*    collinear with the intercept and covariates.                           */
/* ======================================================================== */
/* X is row-major N x P with the genotype in the LAST column. */
extern "C" void orc_fit_ols(int64_t N, int64_t P, const double* X, const double* y, double* out5) {
  /* out5: pve, sigmahat, betahat_geno, sebetahat_geno, betapval_geno */
  if (N < P + 1) {
    for (int k = 0; k < 5; ++k) out5[k] = kNaN;
    return;
  }
  /* Thin QR by modified Gram-Schmidt with one re-orthogonalisation pass;
   * a column whose residual norm falls below 1e-9 of its own norm is
   * treated as linearly dependent and dropped (rank deficiency). */
  std::vector<std::vector<ldouble>> Qc;   /* kept orthonormal columns   */
  std::vector<int64_t> kc;                /* original index of each     */
  std::vector<std::vector<ldouble>> Rc;   /* R column for each kept col */
  for (int64_t j = 0; j < P; ++j) {
    std::vector<ldouble> v(N);
    ldouble n0 = 0;
    for (int64_t i = 0; i < N; ++i) {
      v[i] = X[i * P + j];
      n0 += v[i] * v[i];
    }
    n0 = sqrtl(n0);
    std::vector<ldouble> rcol(Qc.size() + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      for (size_t a = 0; a < Qc.size(); ++a) {
        ldouble d = 0;
        for (int64_t i = 0; i < N; ++i) d += Qc[a][i] * v[i];
        for (int64_t i = 0; i < N; ++i) v[i] -= d * Qc[a][i];
        rcol[a] += d;
      }
    }
    ldouble nv = 0;
    for (int64_t i = 0; i < N; ++i) nv += v[i] * v[i];
    nv = sqrtl(nv);
    if (nv <= 1e-9L * n0 || nv == 0) continue;
    for (int64_t i = 0; i < N; ++i) v[i] /= nv;
    rcol[Qc.size()] = nv;
    Qc.push_back(v);
    Rc.push_back(rcol);
    kc.push_back(j);
  }
  const int64_t rank = (int64_t)kc.size();
  /* coefficients: R beta = Q'y */
  std::vector<ldouble> qy(rank, 0), beta(rank, 0);
  for (int64_t a = 0; a < rank; ++a) {
    ldouble d = 0;
    for (int64_t i = 0; i < N; ++i) d += Qc[a][i] * (ldouble)y[i];
    qy[a] = d;
  }
  for (int64_t a = rank - 1; a >= 0; --a) {
    ldouble s = qy[a];
    for (int64_t c = a + 1; c < rank; ++c) s -= Rc[c][a] * beta[c];
    beta[a] = s / Rc[a][a];
  }
  /* residual computed directly from the fitted coefficients */
  ldouble rss = 0, ymean = 0, tss = 0;
  for (int64_t i = 0; i < N; ++i) {
    ldouble f = 0;
    for (int64_t a = 0; a < rank; ++a) f += (ldouble)X[i * P + kc[a]] * beta[a];
    ldouble e = (ldouble)y[i] - f;
    rss += e * e;
    ymean += (ldouble)y[i];
  }
  ymean /= N;
  for (int64_t i = 0; i < N; ++i) tss += ((ldouble)y[i] - ymean) * ((ldouble)y[i] - ymean);
  out5[0] = (double)(1 - rss / tss);
  out5[1] = (double)sqrtl(rss / (ldouble)(N - rank));
  if (rank == 0 || kc[rank - 1] != P - 1) {
    out5[2] = out5[3] = out5[4] = kNaN;
    return;
  }
  /* The genotype is the last kept column, so (R^{-1})_{last,last} =
   * 1/R_{last,last} and var(betahat_geno) = s2 / R_{last,last}^2 with
   * s2 = rss/(N-P) (GSL >= 2.3 scales the covariance by rss/(n-p)). */
  ldouble s2 = rss / (ldouble)(N - P);
  ldouble rl = Rc[rank - 1][rank - 1];
  ldouble se = sqrtl(s2) / rl;
  ldouble bh = beta[rank - 1];
  out5[2] = (double)bh;
  out5[3] = (double)se;
  out5[4] = (double)(2 * tdist_lower_neg(fabsl(bh / se), (ldouble)(N - rank)));
}

/* ======================================================================== */
/* 4. Bayes factors (gene_snp_pair.cpp:256-622, utils_math.cpp:100-159).     */
/*    Plain double arithmetic in the reference's operation order.            */
/* ======================================================================== */

/* utils::log10_weighted_sum with uniform weights (utils_math.cpp:100-131). */
extern "C" double orc_log10_weighted_sum(const double* vec, int64_t size) {
  double mx = vec[0];
  for (int64_t i = 0; i < size; ++i)
    if (vec[i] > mx) mx = vec[i];
  double w = 1 / ((double)size);
  double sum = 0.0;
  for (int64_t i = 0; i < size; ++i) {
    if (std::isnan(vec[i])) continue;
    sum += w * pow(10, vec[i] - mx);
  }
  double res = mx + log10(sum);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  return res;
}

/* utils::log10_weighted_sum with explicit weights (utils_math.cpp:135-159). */
extern "C" double orc_log10_weighted_sum_w(const double* vec, const double* w, int64_t size) {
  double mx = vec[0];
  for (int64_t i = 0; i < size; ++i)
    if (vec[i] > mx) mx = vec[i];
  double sum = 0.0;
  for (int64_t i = 0; i < size; ++i) {
    if (std::isnan(vec[i])) continue;
    sum += w[i] * pow(10, vec[i] - mx);
  }
  double res = mx + log10(sum);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  return res;
}

/* CalcLog10AbfUvlr (gene_snp_pair.cpp:297-356).  stdsstats is S x 3
 * (b, se, t) row-major; gamma[s] in {0,1}. */
extern "C" double orc_l10abf_uvlr(int32_t S, const int32_t* gamma, const double* stdsstats,
                                  double phi2, double oma2) {
  double l10AbfAll = 0.0, bbarhat_num = 0.0, bbarhat_denom = 0.0, varbbarhat = 0.0;
  std::vector<double> single;
  for (int32_t s = 0; s < S; ++s) {
    if (gamma[s] == 0) continue;
    double bhat = stdsstats[3 * s + 0];
    double varbhat = pow(stdsstats[3 * s + 1], 2);
    double t = stdsstats[3 * s + 2];
    double l;
    if (fabs(t) < 1e-8) {
      l = 0;
    } else {
      bbarhat_num += bhat / (varbhat + phi2);
      bbarhat_denom += 1 / (varbhat + phi2);
      varbbarhat += 1 / (varbhat + phi2);
      l = 0.5 * log10(varbhat) - 0.5 * log10(varbhat + phi2) +
          (0.5 * pow(t, 2) * phi2 / (varbhat + phi2)) / log(10);
    }
    single.push_back(l);
  }
  double bbarhat = (bbarhat_denom != 0.0) ? bbarhat_num / bbarhat_denom : 0.0;
  varbbarhat = (varbbarhat != 0.0) ? 1 / varbbarhat : kInf;
  if (bbarhat != 0.0 && varbbarhat < kInf) {
    double T2 = pow(bbarhat, 2.0) / varbbarhat;
    double lABF_bbar = (T2 != 0) ? 0.5 * log10(varbbarhat) - 0.5 * log10(varbbarhat + oma2) +
                                       (0.5 * T2 * oma2 / (varbbarhat + oma2)) / log(10)
                                 : 0;
    l10AbfAll = lABF_bbar;
    for (double l : single) l10AbfAll += l;
  } else {
    l10AbfAll = 0.0;
  }
  return l10AbfAll;
}

/* StandardizeSstatsAndCorrectSmallSampleSize for one subgroup
 * (gene_snp_pair.cpp:256-290): in = (N, Q, sigmahat, betahat, sebetahat),
 * out = (b, se, t), NaN when t is NaN. */
extern "C" void orc_standardize(double N, double Q, double sigmahat, double betahat, double sebetahat,
                                double* out3) {
  out3[0] = out3[1] = out3[2] = kNaN;
  double bhat = betahat / sigmahat, sebhat = sebetahat / sigmahat, t = bhat / sebhat;
  if (std::isnan(t)) return;
  double nu = N - 2 - Q;
  t = orc_gauss_Pinv(orc_tdist_P(-fabs(bhat / sebhat), nu));
  if (fabs(t) > 1e-8) {
    double sigmahat2 = fabs(betahat) / (fabs(t) * sebhat);
    bhat = betahat / sigmahat2;
    sebhat = fabs(bhat / t);
  } else {
    bhat = 0;
    sebhat = kInf;
  }
  out3[0] = bhat;
  out3[1] = sebhat;
  out3[2] = t;
}

/* gsl_combination order of all non-empty subsets: size 1..S, lexicographic
 * within a size (gene_snp_pair.cpp:469-550; confirmed by the hard-coded row
 * order "1","2","3","1-2","1-3","2-3","1-2-3" in tests/functional_tests.R:850). */
void orc::config_masks(int32_t bfs, int32_t S, std::vector<uint64_t>& masks) {
  masks.clear();
  if (bfs == EQBF_BFS_GEN) return;
  if (bfs == EQBF_BFS_SIN) {
    for (int32_t s = 0; s < S; ++s) masks.push_back(1ull << s);
    return;
  }
  for (int32_t k = 1; k <= S; ++k) {
    std::vector<int32_t> c(k);
    for (int32_t i = 0; i < k; ++i) c[i] = i;
    while (true) {
      uint64_t m = 0;
      for (int32_t i = 0; i < k; ++i) m |= 1ull << c[i];
      masks.push_back(m);
      int32_t i = k - 1;
      while (i >= 0 && c[i] == S - k + i) --i;
      if (i < 0) break;
      ++c[i];
      for (int32_t j = i + 1; j < k; ++j) c[j] = c[j - 1] + 1;
    }
  }
}

extern "C" int64_t orc_config_masks(int32_t bfs, int32_t S, uint64_t* out) {
  std::vector<uint64_t> m;
  config_masks(bfs, S, m);
  if (out)
    for (size_t i = 0; i < m.size(); ++i) out[i] = m[i];
  return (int64_t)m.size();
}

static double choose(int n, int k) {
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* Per-pair Bayes factors (CalcAbfsUvlr, gene_snp_pair.cpp:604-622 and the
 * functions it calls, :364-602).  has[s]: subgroup has results
 * (HasResults); std3: S x 3 standardised stats. raw: 3*GL + ncfg*GS values,
 * w: 3 + ncfg + 2 values (gen-sin, all at the end; NaN when not computed). */
void orc::pair_abfs(int32_t S, int32_t bfs, const std::vector<int>& has, const std::vector<double>& std3,
                      const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                      const double* omaS, int32_t GS, double* raw, double* w) {
  std::vector<int32_t> gamma(S);
  std::vector<double> st(3 * S);
  /* consistent configuration: subgroups with results (:371-383) */
  for (int32_t s = 0; s < S; ++s) {
    gamma[s] = has[s] ? 1 : 0;
    for (int k = 0; k < 3; ++k) st[3 * s + k] = has[s] ? std3[3 * s + k] : 0.0;
  }
  std::vector<double> gen(GL), fix(GL), maxh(GL);
  for (int32_t g = 0; g < GL; ++g) {
    gen[g] = orc_l10abf_uvlr(S, gamma.data(), st.data(), phiL[g], omaL[g]);
    fix[g] = orc_l10abf_uvlr(S, gamma.data(), st.data(), 0.0, phiL[g] + omaL[g]);
    maxh[g] = orc_l10abf_uvlr(S, gamma.data(), st.data(), phiL[g] + omaL[g], 0.0);
  }
  for (int32_t g = 0; g < GL; ++g) {
    raw[g] = gen[g];
    raw[GL + g] = fix[g];
    raw[2 * GL + g] = maxh[g];
  }
  w[0] = orc_log10_weighted_sum(gen.data(), GL);
  w[1] = orc_log10_weighted_sum(fix.data(), GL);
  w[2] = orc_log10_weighted_sum(maxh.data(), GL);
  std::vector<uint64_t> masks;
  config_masks(bfs, S, masks);
  int64_t nc = (int64_t)masks.size();
  w[3 + nc] = kNaN;
  w[3 + nc + 1] = kNaN;
  if (bfs == EQBF_BFS_GEN) return;
  std::vector<double> l(GS);
  for (int64_t c = 0; c < nc; ++c) {
    for (int32_t s = 0; s < S; ++s) {
      int on = ((masks[c] >> s) & 1) && has[s];
      gamma[s] = on;
      for (int k = 0; k < 3; ++k) st[3 * s + k] = on ? std3[3 * s + k] : 0.0;
    }
    bool any = false;
    for (int32_t s = 0; s < S; ++s) any |= gamma[s] != 0;
    for (int32_t g = 0; g < GS; ++g) {
      /* singletons of absent subgroups are all 0.0 (:436,458-459); for 'all',
       * configurations with no active subgroup also evaluate to 0. */
      l[g] = any ? orc_l10abf_uvlr(S, gamma.data(), st.data(), phiS[g], omaS[g]) : 0.0;
      raw[3 * GL + c * GS + g] = l[g];
    }
    w[3 + c] = orc_log10_weighted_sum(l.data(), GS);
  }
  orc_bma(S, bfs, w + 3, w[0], &w[3 + nc], bfs == EQBF_BFS_ALL ? &w[3 + nc + 1] : nullptr);
}

/* CalcBMAlite (gene_snp_pair.cpp:552-570): log10 of the weighted sum of the
 * singleton averages "1".."S" (weight 1/(2S) each) and gen (weight 1/2).
 * CalcBMA (:572-602): log10 of the weighted sum over every configuration of
 * size k (weight (1/S) / C(S, k)); w_cfg holds the configuration averages in
 * gsl_combination order (for 'all' the singletons are the first S).  Both go
 * through log10_weighted_sum, so a NaN first term poisons the result. */
extern "C" void orc_bma(int32_t S, int32_t bfs, const double* w_cfg, double gen, double* gen_sin, double* all) {
  {
    std::vector<double> v, wt;
    for (int32_t s = 0; s < S; ++s) {
      v.push_back(w_cfg[s]);
      wt.push_back((1.0 / 2.0) * (1.0 / S));
    }
    v.push_back(gen);
    wt.push_back(1.0 / 2.0);
    *gen_sin = orc_log10_weighted_sum_w(v.data(), wt.data(), (int64_t)v.size());
  }
  if (bfs == EQBF_BFS_ALL && all) {
    std::vector<uint64_t> masks;
    config_masks(bfs, S, masks);
    std::vector<double> v, wt;
    for (size_t c = 0; c < masks.size(); ++c) {
      int k = __builtin_popcountll(masks[c]);
      v.push_back(w_cfg[c]);
      wt.push_back((1.0 / (double)S) * (1.0 / choose(S, k)));
    }
    *all = orc_log10_weighted_sum_w(v.data(), wt.data(), (int64_t)v.size());
  }
}

/* ======================================================================== */
/* 5. One chunk of association tests, uvlr (Gene::TestForAssociations,       */
/*    gene.cpp:263-332; GeneSnpPair::FillStlContainers :79-170 with          */
/*    perm == NULL, CalcSstatsOneSbgrp :175-208, CalcAbfsUvlr :604-622).     */
/* ======================================================================== */
double orc::expl_at(const eqbf_chunk* c, int64_t g, int32_t s, int64_t i) {
  return c->explevels[(g * c->n_subgroups + s) * c->n_samples + i];
}
double orc::geno_at(const eqbf_chunk* c, int32_t s, int64_t i, int64_t j) {
  int64_t set = c->geno_set_of_sbgrp[s];
  int64_t off = set * c->geno_set_stride + i * c->geno_ld + j;
  if (c->geno_dtype == EQBF_F32) return ((const float*)c->genotypes)[off];
  return ((const double*)c->genotypes)[off];
}

/* Members of (gene g, subgroup s) under a permutation of expression indices
 * (perm == NULL: identity): sample i is used when its permuted expression
 * value exists and is not NaN and i has a genotype column in s
 * (gene_snp_pair.cpp:98-116).  Returns pairs (i, expression value). */
void orc::members(const eqbf_chunk* c, int64_t g, int32_t s, const int32_t* perm, std::vector<int64_t>& idx,
                    std::vector<double>& y) {
  idx.clear();
  y.clear();
  for (int64_t i = 0; i < c->n_samples; ++i) {
    int64_t e = perm ? perm[i] : i;
    double v = expl_at(c, g, s, e);
    if (std::isnan(v) || !c->geno_in_sbgrp[(int64_t)s * c->n_samples + i]) continue;
    idx.push_back(i);
    y.push_back(v);
  }
}

/* Summary statistics of one (pair, subgroup): returns n and fills out5.
 * Returns -1 on a data error (missing covariate of a used sample). */
int64_t orc::sstats_one(const eqbf_chunk* c, int64_t g, int32_t s, int64_t snp, const int32_t* perm,
                          double* out5) {
  std::vector<int64_t> idx;
  std::vector<double> y;
  members(c, g, s, perm, idx, y);
  int64_t N = (int64_t)idx.size();
  if (N == 0) return 0; /* reference dereferences an empty vector here */
  int64_t Q = c->covar_count[s];
  int64_t P = 2 + Q;
  std::vector<double> X(N * P);
  for (int64_t r = 0; r < N; ++r) {
    int64_t i = idx[r];
    X[r * P + 0] = 1.0;
    for (int64_t q = 0; q < Q; ++q) {
      if (!c->covar_in_sbgrp[(int64_t)s * c->n_samples + i]) return -1;
      X[r * P + 1 + q] = c->covars[((int64_t)s * c->n_covars_max + q) * c->n_samples + i];
    }
    X[r * P + P - 1] = geno_at(c, s, i, snp);
  }
  orc_fit_ols(N, P, X.data(), y.data(), out5);
  return N;
}

int64_t orc::n_configs(int32_t bfs, int32_t S) {
  if (bfs == EQBF_BFS_GEN) return 0;
  if (bfs == EQBF_BFS_SIN) return S;
  return (int64_t)((1ull << S) - 1);
}

/* Full association test of a chunk.  Outputs follow include/eqbf.h
 * (structure of arrays, leading dimension n_pairs); raw ABFs are double. */
extern "C" int orc_assoc_uvlr(const eqbf_chunk* c, int32_t analysis, int32_t bfs, const double* phiL,
                              const double* omaL, int32_t GL, const double* phiS, const double* omaS,
                              int32_t GS, int32_t* n_out, double* sstats, double* stdsstats, double* raw,
                              double* w_abf) {
  const int32_t S = c->n_subgroups;
  const int64_t NP = c->n_pairs;
  const int64_t nc = n_configs(bfs, S);
  const int64_t n_raw = 3 * (int64_t)GL + nc * GS;
  const int64_t n_w = 3 + nc + 2;
  std::vector<double> raw1(n_raw), w1(n_w);
  for (int64_t g = 0; g < c->n_genes; ++g) {
    for (int64_t p = c->pair_offset[g]; p < c->pair_offset[g + 1]; ++p) {
      int64_t snp = c->pair_snp[p];
      std::vector<int> has(S, 0);
      std::vector<double> std3(3 * S, kNaN);
      for (int32_t s = 0; s < S; ++s) {
        double o5[5] = {kNaN, kNaN, kNaN, kNaN, kNaN};
        int64_t n = 0;
        if (c->gene_in_sbgrp[g * S + s] && c->snp_in_sbgrp[snp * S + s]) {
          n = sstats_one(c, g, s, snp, nullptr, o5);
          if (n < 0) {
            snprintf(g_err, sizeof g_err, "missing covariate for a sample of subgroup %d", s);
            return EQBF_ERR_DATA;
          }
        }
        n_out[(int64_t)s * NP + p] = (int32_t)n;
        if (sstats)
          for (int k = 0; k < 5; ++k) sstats[((int64_t)k * S + s) * NP + p] = o5[k];
        has[s] = n > 0;
        if (n > 0) orc_standardize((double)n, (double)c->covar_count[s], o5[1], o5[2], o5[3], &std3[3 * s]);
        if (stdsstats)
          for (int k = 0; k < 3; ++k) stdsstats[((int64_t)k * S + s) * NP + p] = std3[3 * s + k];
      }
      if (analysis != EQBF_ANALYSIS_JOIN) continue;
      pair_abfs(S, bfs, has, std3, phiL, omaL, GL, phiS, omaS, GS, raw1.data(), w1.data());
      if (raw)
        for (int64_t k = 0; k < n_raw; ++k) raw[k * NP + p] = raw1[k];
      if (w_abf)
        for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = w1[k];
    }
  }
  return EQBF_OK;
}

/* ======================================================================== */
/* 6. Summary statistics as input (--inss): GeneSnpPair::SetSstats          */
/*    (gene_snp_pair.cpp:241-254) then CalcAbfsUvlr (:604-622) for pairs   */
/*    built by fillGeneSnpPairsWithSstats (data_loader.cpp:1250-1327).     */
/*    Arrays [S][NP]; n <= 0: no result (HasResults, :61-68).  The         */
/*    covariate count of given statistics is the map default, 0 (:273).    */
/*    nu = n - 2 <= 0 (n <= 2) is given a NaN z (GSL's t distribution at   */
/*    nu <= 0 is not restated).                                            */
/* ======================================================================== */
extern "C" int orc_abf_from_sstats(int32_t S, int64_t NP, const int64_t* n, const double* sigmahat,
                                   const double* betahat, const double* sebetahat, int32_t bfs,
                                   const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                                   const double* omaS, int32_t GS, double* stdsstats, double* raw, double* w_abf) {
  std::vector<uint64_t> masks;
  config_masks(bfs, S, masks);
  const int64_t nc = (int64_t)masks.size();
  const int64_t n_raw = 3 * (int64_t)GL + nc * GS, n_w = 3 + nc + 2;
  std::vector<double> raw1(n_raw), w1(n_w), std3(3 * S);
  std::vector<int> has(S);
  for (int64_t p = 0; p < NP; ++p) {
    for (int32_t s = 0; s < S; ++s) {
      const int64_t e = (int64_t)s * NP + p;
      has[s] = n[e] > 0;
      for (int k = 0; k < 3; ++k) std3[3 * s + k] = kNaN;
      if (has[s]) {
        if (n[e] > 2) {
          orc_standardize((double)n[e], 0.0, sigmahat[e], betahat[e], sebetahat[e], &std3[3 * s]);
        } else if (!std::isnan((betahat[e] / sigmahat[e]) / (sebetahat[e] / sigmahat[e]))) {
          std3[3 * s + 0] = 0.0; /* |z| > 1e-8 fails for a NaN z (:278-286) */
          std3[3 * s + 1] = kInf;
        }
      }
      if (stdsstats)
        for (int k = 0; k < 3; ++k) stdsstats[(int64_t)k * S * NP + e] = std3[3 * s + k];
    }
    pair_abfs(S, bfs, has, std3, phiL, omaL, GL, phiS, omaS, GS, raw1.data(), w1.data());
    if (raw)
      for (int64_t k = 0; k < n_raw; ++k) raw[k * NP + p] = raw1[k];
    if (w_abf)
      for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = w1[k];
  }
  return 0;
}
