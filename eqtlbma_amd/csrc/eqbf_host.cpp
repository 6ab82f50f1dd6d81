/*
 * Host-side pieces of the C ABI: the GSL-compatible random numbers that make
 * the permutation index tables, and the reference's permutation counting
 * loop.  Plain C++; no device code.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/eqbf.h"

/* gsl_rng_set for mt19937: seed 0 means 4357, then Knuth's initialisation
 * with multiplier 1812433253; the state index starts past the end so that
 * the first draw regenerates the state. */
extern "C" void eqbf_rng_seed(eqbf_rng* r, uint64_t seed) {
  if (seed == 0) seed = 4357;
  r->mt[0] = (uint32_t)(seed & 0xffffffffULL);
  for (int k = 1; k < 624; ++k) r->mt[k] = 1812433253u * (r->mt[k - 1] ^ (r->mt[k - 1] >> 30)) + (uint32_t)k;
  r->idx = 624;
}

/* MT19937 (Matsumoto & Nishimura 1998), 32-bit output. */
extern "C" uint32_t eqbf_rng_get(eqbf_rng* r) {
  static const uint32_t kMatrix = 0x9908b0dfu;
  if (r->idx >= 624) {
    uint32_t* m = r->mt;
    for (int k = 0; k < 624; ++k) {
      uint32_t y = (m[k] & 0x80000000u) | (m[(k + 1) % 624] & 0x7fffffffu);
      m[k] = m[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? kMatrix : 0u);
    }
    r->idx = 0;
  }
  uint32_t y = r->mt[r->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  return y ^ (y >> 18);
}

/* gsl_rng_uniform_int: bucket width floor((2^32-1)/n), reject past the last
 * whole bucket. */
extern "C" uint32_t eqbf_rng_uniform_int(eqbf_rng* r, uint32_t n) {
  const uint32_t width = 0xffffffffu / n;
  uint32_t k;
  do {
    k = eqbf_rng_get(r) / width;
  } while (k >= n);
  return k;
}

/* gsl_ran_flat: a (1 - u) + b u with u = next / 2^32. */
extern "C" double eqbf_rng_flat(eqbf_rng* r, double a, double b) {
  const double u = eqbf_rng_get(r) / 4294967296.0;
  return a * (1 - u) + b * u;
}

/* gsl_ran_shuffle: i = n-1 .. 1, swap a[i] with a[uniform_int(i+1)]. */
extern "C" void eqbf_rng_shuffle(eqbf_rng* r, int32_t* a, int64_t n) {
  for (int64_t i = n - 1; i > 0; --i) {
    const int64_t j = eqbf_rng_uniform_int(r, (uint32_t)(i + 1));
    std::swap(a[i], a[j]);
  }
}

extern "C" void eqbf_perm_tables(eqbf_rng* r, int64_t n_samples, int64_t n_genes, int64_t nperm, int32_t* out) {
  std::vector<int32_t> perm(n_samples);
  for (int64_t k = 0; k < n_genes; ++k) {
    for (int64_t i = 0; i < n_samples; ++i) perm[i] = (int32_t)i;
    for (int64_t p = 0; p < nperm; ++p) {
      eqbf_rng_shuffle(r, perm.data(), n_samples);
      std::copy(perm.begin(), perm.end(), out + (k * nperm + p) * n_samples);
    }
  }
}

/* utils::median (utils_math.hpp:61-80) over the valid statistics. */
static double median_of(std::vector<double>& v) {
  if (v.empty()) return std::numeric_limits<double>::quiet_NaN();
  const size_t n = v.size(), mid = n / 2;
  std::nth_element(v.begin(), v.begin() + mid, v.end());
  const double a = v[mid];
  if (n % 2) return a;
  std::nth_element(v.begin(), v.begin() + mid - 1, v.end());
  return (a + v[mid - 1]) / 2.0;
}

extern "C" int64_t eqbf_perm_pvalue(const double* stats, double true_stat, int32_t sep, int64_t* nb_permutations,
                                    int32_t trick, int64_t trick_cutoff, eqbf_rng* rng_trick, double* pval,
                                    int64_t* nperm_used, double* median) {
  int64_t sofar = 0, drawn = 0;
  double count = 1; /* the reference starts the count at 1 (gene.cpp:403,515,621) */
  bool shuffle_only = false;
  std::vector<double> kept;
  const int64_t n_all = *nb_permutations;
  for (int64_t p = 0; p < n_all; ++p) {
    ++drawn;
    if (shuffle_only) continue;
    const double st = stats[p];
    if (std::isnan(st)) {
      --(*nb_permutations);
      continue;
    }
    ++sofar;
    if (sep ? (st <= true_stat) : (st >= true_stat)) ++count;
    if (!sep) kept.push_back(st);
    if (trick != 0 && count == 1 + trick_cutoff) {
      if (trick == 1) break;
      shuffle_only = true;
    }
  }
  if (sofar == *nb_permutations)
    *pval = count / (double)(*nb_permutations + 1);
  else if (rng_trick)
    *pval = eqbf_rng_flat(rng_trick, (1 + trick_cutoff) / ((double)(sofar + 2)),
                          (1 + trick_cutoff) / ((double)(sofar + 1)));
  else
    *pval = std::numeric_limits<double>::quiet_NaN(); /* the reference uses a NULL generator here */
  if (nperm_used) *nperm_used = sofar;
  if (median) *median = sep ? std::numeric_limits<double>::quiet_NaN() : median_of(kept);
  return drawn;
}
