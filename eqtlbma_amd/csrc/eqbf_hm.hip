/*
 * EM of eqtlbma_hm's hierarchical model on raw log10 ABFs
 * (Controller::run_EM_classic / run_EM_square, eqtlbma_hm.cpp:925-1346,
 * over the per-SNP and per-gene log10 sums of hm_methods.cpp:74-781).
 *
 * One EM pass over the data evaluates the model at the current parameters
 * and yields everything an EM update needs, in three kernels:
 *  hm_pair_kernel  per (gene, SNP) pair: the grid-averaged BF of each
 *                  configuration (or type), the SNP's BF, the
 *                  configuration-averaged BF of each grid point and, for the
 *                  types model, each (type, subgroup)'s numerator of the
 *                  subgroup update (hm_methods.cpp:74-300).  Configs model
 *                  (hm_cfg_tile_kernel): persistent workgroups over tiles of
 *                  pairs, the next tile's raw ABFs loaded into registers
 *                  while the current one is computed from LDS; one thread
 *                  per configuration row (sum over the grid), then one per
 *                  grid column and one per SNP (sums over the
 *                  configurations), with one exponential per raw value where
 *                  the pair's range allows it.  Types model, and
 *                  configurations too many for a tile: one wave per pair;
 *  hm_gene_kernel  one workgroup per gene, lanes over the columns and the
 *                  four waves over the gene's SNPs: the gene's BF (weight
 *                  1/P), its log10 observed likelihood, and each column's
 *                  SNP average minus that likelihood -- the gene's
 *                  contribution to each weight's update (hm_methods.cpp:
 *                  392-704, gene_eQTL::em_update_*);
 *  hm_col_kernel   one workgroup per column: the log10 sum over genes of
 *                  those contributions, the likelihood and pi0's sum
 *                  (eqtlbma_hm.cpp:617-868).
 * The host then normalises the new weights exactly as the reference does
 * (log10 sums with the previous weight added, divided by their log10 sum)
 * and runs the reference's iteration logic; one pass per EM iteration, all
 * in FP64.  Every log10 sum keeps log10_weighted_sum's rules
 * (utils_math.cpp:133-159): the maximum starts from the first element (a
 * NaN first element makes the sum NaN), NaN terms are skipped, and a result
 * within DBL_EPSILON of zero is zero.
 */
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

constexpr int HM_TYPES_KMAX = 16; /* types model: types per wave in registers */
constexpr int HM_NT = 256;

__device__ inline double hm_finish(double m, double s) {
  const double r = m + log10(s);
  return fabs(r) <= DBL_EPSILON ? 0.0 : r;
}

/* log10 sum w 10^v over the wave, v given per lane (valid lanes only):
 * the first element's NaN decides, as the reference's sequential max */
__device__ inline double hm_wave_lws(double v, double w, bool valid, double first) {
  double m = (valid && !isnan(v)) ? v : -INFINITY;
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (isnan(first)) m = NAN;
  double s = (valid && !isnan(v)) ? w * exp10(v - m) : 0.0;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return hm_finish(m, s);
}

struct HmDev {
  const double* raw; /* [NP][D][L] */
  const int64_t* off;
  int64_t NP, G;
  int D, L, K, S;  /* D = K (configs) or S (types) */
  const double* par; /* pi0, grid[L], dimw[K], sub[K][S] */
  double* cfg;       /* [NP][K] grid-averaged BF per configuration / type  */
  double* grd;       /* [NP][L] configuration-averaged BF per grid point    */
  double* num;       /* [NP][K][S] types: subgroup-update numerators        */
  double* snp;       /* [NP] SNP BF                                         */
  double* rows;      /* [G][NC + 3]: per-gene columns, NC = K + L (+ K S), then
                        log10 obslik, pi0's term and the gene's BF          */
  double* red;       /* [NC + 2] over genes                                 */
  int stale;         /* 1: keep each gene's stored log10 obslik (row[NC])   */
  int lik_only;      /* 1: the likelihood only, no update columns           */
  int tp;            /* configs model: pairs per workgroup tile (0: none)   */
  double range;      /* configs model: widest pair span for factored sums  */
};

/* log10(q 10^b + 1 - q), a subgroup's term of the types model
 * (hm_methods.cpp:143-146) */
__device__ inline double hm_sub_term(double q, double b) { return log10(exp10(log10(q) + b) + 1 - q); }

template <bool TYPES>
__global__ void __launch_bounds__(HM_NT) hm_pair_kernel(const HmDev h) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * (HM_NT / 64) + (threadIdx.x >> 6);
  if (p >= h.NP) return;
  const int L = h.L, K = h.K, S = h.S, D = h.D;
  const double* R = h.raw + (size_t)p * D * L;
  const double* grid = h.par + 1;
  const double* dw = h.par + 1 + L;
  if (!TYPES) {
    /* per configuration: log10 sum_l grid_l 10^R[k][l] (hm_methods.cpp:81-84) */
    for (int k = lane; k < K; k += 64) {
      const double* r = R + (size_t)k * L;
      double m = r[0];
      for (int l = 1; l < L; ++l)
        if (r[l] > m) m = r[l];
      double s = 0.0;
      for (int l = 0; l < L; ++l)
        if (!isnan(r[l])) s += grid[l] * exp10(r[l] - m);
      h.cfg[(size_t)p * K + k] = hm_finish(m, s);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    /* the SNP's BF: log10 sum_k w_k 10^cfg[k] (hm_methods.cpp:102-104) */
    const double* c = h.cfg + (size_t)p * K;
    const double first = c[0];
    double m = -INFINITY;
    for (int k = lane; k < K; k += 64)
      if (!isnan(c[k])) m = fmax(m, c[k]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (isnan(first)) m = NAN;
    double s = 0.0;
    for (int k = lane; k < K; k += 64)
      if (!isnan(c[k])) s += dw[k] * exp10(c[k] - m);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) h.snp[p] = hm_finish(m, s);
    if (h.lik_only) return;
    /* per grid point: log10 sum_k w_k 10^R[k][l] (hm_methods.cpp:261-275) */
    for (int l = lane; l < L; l += 64) {
      double mm = R[l];
      for (int k = 1; k < K; ++k)
        if (R[(size_t)k * L + l] > mm) mm = R[(size_t)k * L + l];
      double ss = 0.0;
      for (int k = 0; k < K; ++k) {
        const double v = R[(size_t)k * L + l];
        if (!isnan(v)) ss += dw[k] * exp10(v - mm);
      }
      h.grd[(size_t)p * L + l] = hm_finish(mm, ss);
    }
    return;
  }
  /* types model, one lane per grid point (L <= 64): t[k] = sum_s
   * log10(q_ks 10^R[s][l] + 1 - q_ks) */
  const double* sub = h.par + 1 + L + K;
  const bool on = lane < L;
  const int l = on ? lane : 0;
  double t[HM_TYPES_KMAX];
#pragma unroll
  for (int k = 0; k < HM_TYPES_KMAX; ++k) {
    double acc = 0.0;
    if (k < K)
      for (int s = 0; s < S; ++s) acc += hm_sub_term(sub[(size_t)k * S + s], R[(size_t)s * L + l]);
    t[k] = acc;
  }
  /* per type: log10 sum_l grid_l 10^t[k][l] (hm_methods.cpp:131-160) */
  for (int k = 0; k < HM_TYPES_KMAX && k < K; ++k) {
    const double first = __shfl(t[k], 0, 64);
    const double v = hm_wave_lws(t[k], grid[l], on, first);
    if (lane == 0) h.cfg[(size_t)p * K + k] = v;
    /* subgroup-update numerators: R[s][l] + sum over the other subgroups
     * (hm_methods.cpp:236-247) */
    for (int s = 0; s < S && !h.lik_only; ++s) {
      double x = 0.0;
      for (int s2 = 0; s2 < S; ++s2)
        if (s2 != s) x += hm_sub_term(sub[(size_t)k * S + s2], R[(size_t)s2 * L + l]);
      x = R[(size_t)s * L + l] + x;
      const double f = __shfl(x, 0, 64);
      const double nv = hm_wave_lws(x, grid[l], on, f);
      if (lane == 0) h.num[((size_t)p * K + k) * S + s] = nv;
    }
  }
  /* per grid point: log10 sum_k w_k 10^t[k] (hm_methods.cpp:277-300) */
  if (on && !h.lik_only) {
    double mm = t[0];
    for (int k = 1; k < HM_TYPES_KMAX && k < K; ++k)
      if (t[k] > mm) mm = t[k];
    double ss = 0.0;
    for (int k = 0; k < HM_TYPES_KMAX && k < K; ++k)
      if (!isnan(t[k])) ss += dw[k] * exp10(t[k] - mm);
    h.grd[(size_t)p * L + l] = hm_finish(mm, ss);
  }
  /* the SNP's BF over the types */
  if (lane == 0) {
    const double* c = h.cfg + (size_t)p * K;
    double m = c[0];
    for (int k = 1; k < K; ++k)
      if (c[k] > m) m = c[k];
    double s = 0.0;
    for (int k = 0; k < K; ++k)
      if (!isnan(c[k])) s += dw[k] * exp10(c[k] - m);
    h.snp[p] = hm_finish(m, s);
  }
}

/* log10 sum_i w_i 10^v[i*stride] over n values, two passes as
 * log10_weighted_sum does (max from the first value, NaN terms skipped) */
__device__ inline double hm_lws_seq(const double* v, int n, int stride, const double* w) {
  double m = v[0];
  for (int i = 1; i < n; ++i)
    if (v[i * stride] > m) m = v[i * stride];
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const double x = v[i * stride];
    if (!isnan(x)) s += w[i] * exp10(x - m);
  }
  return hm_finish(m, s);
}

/* i / d for 0 <= i < 2^22 through a float reciprocal, corrected both ways */
__device__ inline int hm_div(int i, int d, float inv) {
  int q = (int)((float)i * inv);
  if (q * d > i)
    --q;
  else if ((q + 1) * d <= i)
    ++q;
  return q;
}

constexpr int HM_PRE = 16;                            /* raw values per thread in flight  */
constexpr int HM_LGKM0 = 15 | (7 << 4) | (3 << 14);  /* s_waitcnt lgkmcnt(0) only       */

/* Workgroup barrier that waits for this wave's LDS operations only: the
 * release fence of __syncthreads would also wait (vmcnt(0)) for the next
 * tile's loads, which are meant to stay in flight over the computation.
 * The empty asm statements keep the compiler from moving memory accesses
 * across it. */
__device__ __forceinline__ void hm_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(HM_LGKM0);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

/* Configs model: persistent workgroups over tiles of h.tp pairs; the next
 * tile's raw ABFs are loaded into registers (coalesced, HM_PRE values per
 * thread) while the current tile is computed from LDS.
 *
 * A pair whose raw ABFs are all finite and span at most h.range (log10
 * units) takes factored sums: with M the pair's maximum, each value becomes
 * T = 10^(R - M) once, and
 *   configuration k: M + log10 sum_l grid_l T[k][l],
 *   grid point l:    M + log10 sum_k w_k T[k][l],
 *   SNP:             M + log10 sum_k w_k sum_l grid_l T[k][l],
 * the same quantities as the reference's two-pass sums (each against its
 * own row or column maximum), one exponential per value instead of two.
 * h.range keeps every product of a weight and T a normal double (it is 290
 * plus the log10 of the smallest positive grid and configuration weights),
 * so no term underflows.  Any other pair takes the reference's two-pass
 * sums.
 *
 * LDS: grid and configuration weights [L + K], raw ABFs [tp*K][L|1] (odd
 * row stride), row maxima then configuration sums [tp*K], row minima then
 * row sums [tp*K], pair maxima [tp], pair flags [tp]. */
__global__ void __launch_bounds__(HM_NT) __attribute__((amdgpu_waves_per_eu(4, 8))) hm_cfg_tile_kernel(const HmDev h) {
  extern __shared__ double hm_lds[];
  const int K = h.K, L = h.L, TP = h.tp, tid = threadIdx.x;
  const int Lp = L | 1, KL = K * L;
  double* wg = hm_lds;
  double* wd = wg + L;
  double* rl = wd + K;
  double* rm = rl + (size_t)TP * K * Lp;
  double* rs = rm + (size_t)TP * K;
  double* pm = rs + (size_t)TP * K;
  double* pf = pm + TP;
  for (int i = tid; i < L + K; i += HM_NT) wg[i] = h.par[1 + i];
  const float invL = 1.0f / (float)L, invK = 1.0f / (float)K;
  const int64_t ntiles = (h.NP + TP - 1) / TP;
  double pre[HM_PRE];
  auto fetch = [&](int64_t t) {
    const int64_t q0 = t * TP;
    const int nv = (int)min((int64_t)TP, h.NP - q0) * KL;
    const double* src = h.raw + (size_t)q0 * KL;
#pragma unroll
    for (int j = 0; j < HM_PRE; ++j) {
      const int i = tid + j * HM_NT;
      pre[j] = i < nv ? src[i] : 0.0;
    }
  };
  int64_t tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  for (; tile < ntiles; tile += gridDim.x) {
    const int64_t p0 = tile * TP;
    const int np = (int)min((int64_t)TP, h.NP - p0);
    const int nv = np * KL, nrow = np * K;
    {
      /* element i = tid + j * HM_NT is row r, column c: stepped by
       * HM_NT = dq rows + dr columns */
      const int dq = HM_NT / L, dr = HM_NT - dq * L;
      int r = hm_div(tid, L, invL), c = tid - r * L;
#pragma unroll
      for (int j = 0; j < HM_PRE; ++j) {
        if (tid + j * HM_NT < nv) rl[r * Lp + c] = pre[j];
        r += dq;
        c += dr;
        if (c >= L) {
          c -= L;
          ++r;
        }
      }
    }
    hm_barrier();
    if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
    /* rows: maximum as the reference takes it (from the first value, NaN
     * never greater), minimum, finiteness */
    for (int r = tid; r < nrow; r += HM_NT) {
      const double* x = rl + r * Lp;
      double m = x[0], lo = x[0];
      bool fin = isfinite(x[0]);
      for (int l = 1; l < L; ++l) {
        const double v = x[l];
        if (v > m) m = v;
        lo = fmin(lo, v);
        fin = fin && isfinite(v);
      }
      rm[r] = m;
      rs[r] = fin ? lo : NAN;
    }
    hm_barrier();
    for (int q = tid; q < np; q += HM_NT) {
      double M = -INFINITY, lo = INFINITY;
      bool ok = true;
      for (int k = 0; k < K; ++k) {
        const double b = rs[q * K + k];
        M = fmax(M, rm[q * K + k]);
        if (isnan(b))
          ok = false;
        else
          lo = fmin(lo, b);
      }
      pm[q] = M;
      pf[q] = (ok && M - lo <= h.range) ? 1.0 : 0.0;
    }
    hm_barrier();
    /* configuration sums (hm_methods.cpp:81-84) */
    for (int r = tid; r < nrow; r += HM_NT) {
      const int q = hm_div(r, K, invK);
      double* x = rl + r * Lp;
      double c;
      if (pf[q] != 0.0) {
        const double M = pm[q];
        double s0 = 0.0, s1 = 0.0;
        int l = 0;
        for (; l + 1 < L; l += 2) {
          const double t0 = exp10(x[l] - M), t1 = exp10(x[l + 1] - M);
          x[l] = t0;
          x[l + 1] = t1;
          s0 += wg[l] * t0;
          s1 += wg[l + 1] * t1;
        }
        if (l < L) {
          const double t0 = exp10(x[l] - M);
          x[l] = t0;
          s0 += wg[l] * t0;
        }
        const double sum = s0 + s1;
        rs[r] = sum;
        c = hm_finish(M, sum);
      } else {
        const double m = rm[r];
        double sum = 0.0;
        for (int l = 0; l < L; ++l) {
          const double v = x[l];
          if (!isnan(v)) sum += wg[l] * exp10(v - m);
        }
        c = hm_finish(m, sum);
      }
      rm[r] = c;
      h.cfg[(size_t)p0 * K + r] = c;
    }
    hm_barrier();
    /* the SNP's BF (log10 sum_k w_k 10^cfg[k], hm_methods.cpp:102-104) and
     * the grid columns (log10 sum_k w_k 10^R[k][l], hm_methods.cpp:261-275),
     * one task per thread */
    const int ntask = np + (h.lik_only ? 0 : np * L);
    for (int t = tid; t < ntask; t += HM_NT) {
      const bool snp = t < np;
      const int q = snp ? t : hm_div(t - np, L, invL);
      const int l = snp ? 0 : t - np - q * L;
      double out;
      if (pf[q] != 0.0) {
        const double* a = snp ? rs + q * K : rl + (size_t)q * K * Lp + l;
        const int st = snp ? 1 : Lp;
        double s0 = 0.0, s1 = 0.0;
        int k = 0;
        for (; k + 1 < K; k += 2) {
          s0 += wd[k] * a[k * st];
          s1 += wd[k + 1] * a[(k + 1) * st];
        }
        if (k < K) s0 += wd[k] * a[k * st];
        out = hm_finish(pm[q], s0 + s1);
      } else {
        out = snp ? hm_lws_seq(rm + q * K, K, 1, wd) : hm_lws_seq(rl + (size_t)q * K * Lp + l, K, Lp, wd);
      }
      if (snp)
        h.snp[p0 + q] = out;
      else
        h.grd[(size_t)(p0 + q) * L + l] = out;
    }
    hm_barrier();
  }
}

/* block-wide log10 sum of w 10^v, v strided over the block's threads; the
 * first element decides NaN */
__device__ double hm_block_lws(const double* v, int64_t n, int64_t stride, double w, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double m = -INFINITY;
  for (int64_t i = tid; i < n; i += HM_NT) {
    const double x = v[i * stride];
    if (!isnan(x)) m = fmax(m, x);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __syncthreads();
  if (lane == 0) red[wv] = m;
  __syncthreads();
  m = red[0];
  for (int k = 1; k < HM_NT / 64; ++k) m = fmax(m, red[k]);
  if (n > 0 && isnan(v[0])) m = NAN;
  double s = 0.0;
  for (int64_t i = tid; i < n; i += HM_NT) {
    const double x = v[i * stride];
    if (!isnan(x)) s += w * exp10(x - m);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __syncthreads();
  if (lane == 0) red[wv] = s;
  __syncthreads();
  s = 0.0;
  for (int k = 0; k < HM_NT / 64; ++k) s += red[k];
  return hm_finish(m, s);
}

/* per gene: BF, observed likelihood and the contributions to the updates
 * (one thread per column, SNPs in order, as gene_eQTL::em_update_* sum) */
template <bool TYPES>
__global__ void __launch_bounds__(HM_NT) hm_gene_kernel(const HmDev h) {
  __shared__ double red[HM_NT / 64];
  __shared__ double s_ol;
  const int64_t g = blockIdx.x;
  const int64_t p0 = h.off[g], np = h.off[g + 1] - p0;
  const int K = h.K, L = h.L, S = h.S;
  const int NC = K + L + (TYPES ? K * S : 0);
  double* row = h.rows + (size_t)g * (NC + 3);
  const double pi0 = h.par[0];
  const double gbf = hm_block_lws(h.snp + p0, np, 1, 1.0 / (double)np, red);
  if (threadIdx.x == 0) {
    /* log10(pi0 10^0 + (1 - pi0) 10^gbf) (hm_methods.cpp:504-520), or the
     * likelihood stored by the last evaluation */
    double ol;
    if (h.stale) {
      ol = row[NC];
    } else {
      const double m = gbf > 0.0 ? gbf : 0.0;
      double s = pi0 * exp10(0.0 - m);
      if (!isnan(gbf)) s += (1 - pi0) * exp10(gbf - m);
      ol = hm_finish(m, s);
      row[NC] = ol;
    }
    s_ol = ol;
    row[NC + 1] = exp10(log10(pi0) - ol); /* em_update_pi0 (hm_methods.cpp:542-545) */
    row[NC + 2] = gbf;
  }
  __syncthreads();
  const double ol = s_ol;
  if (h.lik_only) return;
  /* lanes over columns, the four waves over the gene's SNPs: column max,
   * then the sum against it (the reference's two passes); the first SNP's
   * NaN makes the column NaN */
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, NW = HM_NT / 64;
  const double w = 1.0 / (double)np;
  __shared__ double part[HM_NT / 64][64];
  for (int c0 = 0; c0 < NC; c0 += 64) {
    const int c = c0 + lane;
    const bool on = c < NC;
    const double* v = h.cfg;
    int64_t stride = K;
    if (on) {
      if (c < K) {
        v = h.cfg + (size_t)p0 * K + c;
        stride = K;
      } else if (c < K + L) {
        v = h.grd + (size_t)p0 * L + (c - K);
        stride = L;
      } else {
        v = h.num + (size_t)p0 * K * S + (c - K - L);
        stride = (int64_t)K * S;
      }
    }
    double m = -INFINITY;
    if (on) {
      double m1 = -INFINITY;
      int64_t i = wv;
      for (; i + NW < np; i += 2 * NW) {
        const double x0 = v[i * stride], x1 = v[(i + NW) * stride];
        if (x0 > m) m = x0;
        if (x1 > m1) m1 = x1;
      }
      if (i < np && v[i * stride] > m) m = v[i * stride];
      m = fmax(m, m1);
    }
    part[wv][lane] = m;
    __syncthreads();
    m = part[0][lane];
    for (int k = 1; k < NW; ++k) m = fmax(m, part[k][lane]);
    if (on && isnan(v[0])) m = NAN;
    __syncthreads();
    double sum = 0.0;
    if (on) {
      double sum1 = 0.0;
      int64_t i = wv;
      for (; i + NW < np; i += 2 * NW) {
        const double x0 = v[i * stride], x1 = v[(i + NW) * stride];
        if (!isnan(x0)) sum += w * exp10(x0 - m);
        if (!isnan(x1)) sum1 += w * exp10(x1 - m);
      }
      if (i < np && !isnan(v[i * stride])) sum += w * exp10(v[i * stride] - m);
      sum += sum1;
    }
    part[wv][lane] = sum;
    __syncthreads();
    if (wv == 0 && on) {
      double t = 0.0;
      for (int k = 0; k < NW; ++k) t += part[k][lane];
      row[c] = hm_finish(m, t) - ol;
    }
    __syncthreads();
  }
}

/* over genes: log10 sum of each contribution column (weights 1), the sum
 * of the log10 likelihoods and pi0's sum */
__global__ void __launch_bounds__(HM_NT) hm_col_kernel(const HmDev h, int NC, int j0) {
  __shared__ double red[HM_NT / 64];
  const int j = j0 + blockIdx.x;
  const double* col = h.rows + j;
  const int64_t stride = NC + 3;
  if (j < NC) {
    const double v = hm_block_lws(col, h.G, stride, 1.0, red);
    if (threadIdx.x == 0) h.red[j] = v;
    return;
  }
  double s = 0.0;
  for (int64_t g = threadIdx.x; g < h.G; g += HM_NT) s += col[g * stride];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < HM_NT / 64; ++k) t += red[k];
    h.red[j] = t;
  }
}

/* ---- host side ------------------------------------------------------------ */

struct HmParams {
  double pi0 = NAN;
  std::vector<double> grid, dimw, sub; /* dimw: configs or types; sub [K][S] */
};

struct HmPass {
  double lik = NAN;    /* log10 likelihood at the parameters (a fresh pass)  */
  bool lik_ok = false; /* finite (compute_log10_obs_lik throws otherwise)    */
  bool upd_ok = false; /* no NaN unnormalised configuration or grid weight   */
  HmParams next;       /* one EM update from the parameters                  */
};

/* log10_weighted_sum, sequential (utils_math.cpp:133-159) */
double h_lws(const double* v, const double* w, size_t n) {
  double m = v[0];
  for (size_t i = 0; i < n; ++i)
    if (v[i] > m) m = v[i];
  double s = 0.0;
  for (size_t i = 0; i < n; ++i)
    if (!std::isnan(v[i])) s += w[i] * std::pow(10.0, v[i] - m);
  const double r = m + std::log10(s);
  return std::fabs(r) <= DBL_EPSILON ? 0.0 : r;
}

struct HmRun {
  eqbf_ctx* ctx;
  hipStream_t st;
  HmDev h;
  const eqbf_hm_opts* o;
  bool types;
  int NC;
  std::vector<double> hpar, hred;

  int upload(const HmParams& P) {
    const int K = h.K, L = h.L, S = h.S;
    hpar.assign(1 + L + K + (types ? K * S : 0), 0.0);
    hpar[0] = P.pi0;
    std::memcpy(&hpar[1], P.grid.data(), L * sizeof(double));
    std::memcpy(&hpar[1 + L], P.dimw.data(), K * sizeof(double));
    if (types) std::memcpy(&hpar[1 + L + K], P.sub.data(), (size_t)K * S * sizeof(double));
    /* factored sums only while every weight x 10^-range product stays a
     * normal double: 290 + log10 of the smallest positive weights */
    double lo_g = 1.0, lo_d = 1.0;
    bool bad = false;
    for (double g : P.grid) {
      if (std::isnan(g) || g < 0) bad = true;
      if (g > 0) lo_g = std::min(lo_g, g);
    }
    for (double d : P.dimw) {
      if (std::isnan(d) || d < 0) bad = true;
      if (d > 0) lo_d = std::min(lo_d, d);
    }
    h.range = bad ? -1.0 : 290.0 + std::log10(lo_g) + std::log10(lo_d);
    EQBF_HIP(ctx, hipMemcpyAsync((void*)h.par, hpar.data(), hpar.size() * sizeof(double), hipMemcpyHostToDevice, st));
    return EQBF_OK;
  }

  int launch() {
    const unsigned pg = (unsigned)((h.NP + HM_NT / 64 - 1) / (HM_NT / 64));
    {
      ProfScope ps(ctx, "hm_pair_kernel", st);
      if (types) {
        hipLaunchKernelGGL(hm_pair_kernel<true>, dim3(pg), dim3(HM_NT), 0, st, h);
      } else if (h.tp > 0) {
        const int64_t nt = (h.NP + h.tp - 1) / h.tp;
        const size_t lds = ((size_t)h.tp * (h.K * (h.L | 1) + 2 * h.K + 2) + h.K + h.L) * sizeof(double);
        /* persistent: as many workgroups as fit the CUs at once (LDS-bound) */
        const int per_cu = std::max(1, std::min(8, (int)((160u << 10) / lds)));
        const unsigned nb = (unsigned)std::min<int64_t>(nt, (int64_t)std::max(ctx->n_cu, 1) * per_cu);
        hipLaunchKernelGGL(hm_cfg_tile_kernel, dim3(nb), dim3(HM_NT), lds, st, h);
      } else {
        hipLaunchKernelGGL(hm_pair_kernel<false>, dim3(pg), dim3(HM_NT), 0, st, h);
      }
    }
    {
      ProfScope ps(ctx, "hm_gene_kernel", st);
      if (types)
        hipLaunchKernelGGL(hm_gene_kernel<true>, dim3((unsigned)h.G), dim3(HM_NT), 0, st, h);
      else
        hipLaunchKernelGGL(hm_gene_kernel<false>, dim3((unsigned)h.G), dim3(HM_NT), 0, st, h);
    }
    EQBF_HIP(ctx, hipGetLastError());
    return EQBF_OK;
  }

  /* one pass: the likelihood at P and the EM update from P
   * (Controller::em_update_*, eqtlbma_hm.cpp:659-868); stale: the update
   * reads the genes' likelihoods stored by the previous pass */
  int pass(const HmParams& P, HmPass& out, bool stale = false) {
    h.stale = stale ? 1 : 0;
    h.lik_only = 0;
    int rc = upload(P);
    if (rc) return rc;
    if ((rc = launch())) return rc;
    {
      ProfScope ps(ctx, "hm_col_kernel", st);
      hipLaunchKernelGGL(hm_col_kernel, dim3((unsigned)(NC + 2)), dim3(HM_NT), 0, st, h, NC, 0);
    }
    EQBF_HIP(ctx, hipGetLastError());
    hred.resize(NC + 2);
    EQBF_HIP(ctx, hipMemcpyAsync(hred.data(), h.red, (NC + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st));
    const int K = h.K, L = h.L, S = h.S;
    out.lik = hred[NC];
    out.lik_ok = std::isfinite(out.lik);
    HmParams& Q = out.next;
    Q = P;
    bool ok = true;
    if (!o->fixed[0]) Q.pi0 = hred[NC + 1] / (double)h.G;
    std::vector<double> ones(std::max(K, L), 1.0), t;
    auto normalise = [&](const double* contrib, const std::vector<double>& old, std::vector<double>& nw,
                         bool check) {
      const size_t D = old.size();
      t.assign(D, 0.0);
      for (size_t k = 0; k < D; ++k) {
        t[k] = contrib[k] + std::log10(old[k]);
        /* the reference throws here for configurations and grid points
         * (eqtlbma_hm.cpp:697-700, 844-847), not for types */
        if (check && std::isnan(t[k])) ok = false;
      }
      const double den = h_lws(t.data(), ones.data(), D);
      nw.resize(D);
      for (size_t k = 0; k < D; ++k) nw[k] = std::pow(10.0, t[k] - den);
    };
    if (!types) {
      if (!o->fixed[2] && K > 1) normalise(&hred[0], P.dimw, Q.dimw, true);
    } else {
      if (!o->fixed[3]) normalise(&hred[0], P.dimw, Q.dimw, false);
      if (!o->fixed[4])
        for (int k = 0; k < K; ++k)
          for (int s = 0; s < S; ++s)
            Q.sub[(size_t)k * S + s] =
                std::pow(10.0, hred[K + L + k * S + s] + std::log10(P.sub[(size_t)k * S + s]) - hred[k]);
    }
    if (!o->fixed[1]) normalise(&hred[K], P.grid, Q.grid, true);
    out.upd_ok = ok;
    return EQBF_OK;
  }
};

/* compute_log10_obs_lik(P) alone (eqtlbma_hm.cpp:617-650) */
int hm_loglik(HmRun& R, const HmParams& P, double* lik) {
  HmDev& h = R.h;
  h.stale = 0;
  h.lik_only = 1;
  int rc = R.upload(P);
  if (!rc) rc = R.launch();
  if (rc) return rc;
  {
    ProfScope ps(R.ctx, "hm_col_kernel", R.st);
    hipLaunchKernelGGL(hm_col_kernel, dim3(1), dim3(HM_NT), 0, R.st, h, R.NC, R.NC);
  }
  EQBF_HIP(R.ctx, hipGetLastError());
  EQBF_HIP(R.ctx, hipMemcpyAsync(lik, h.red + R.NC, sizeof(double), hipMemcpyDeviceToHost, R.st));
  EQBF_HIP(R.ctx, hipStreamSynchronize(R.st));
  return EQBF_OK;
}

double sq(double x) { return x * x; }

int hm_throw(HmRun& R) {
  return ctx_fail(R.ctx, EQBF_ERR_DATA, "log10(obslik) or an unnormalised weight is NaN or infinite");
}

/* run_EM_fixedpoint from cur's parameters: the update cur.next and, in
 * nxt, its likelihood and the update after it */
int hm_fixedpoint(HmRun& R, const HmPass& cur, HmPass& nxt) {
  if (!cur.upd_ok) return hm_throw(R);
  const int rc = R.pass(cur.next, nxt);
  if (rc) return rc;
  return nxt.lik_ok ? EQBF_OK : hm_throw(R);
}

bool hm_stop(const HmRun& R, double a, double b, int64_t iter) {
  return std::fabs(a - b) < R.o->thresh || (R.o->max_iters > 0 && iter == R.o->max_iters - 1);
}

/* run_EM_classic (eqtlbma_hm.cpp:1076-1107) with passes: pass(P_t) gives
 * lik(P_t) and P_{t+1} */
int hm_classic(HmRun& R, HmParams& P, double* lik_out, int64_t* iters_out) {
  HmPass cur;
  int rc = R.pass(P, cur);
  if (rc) return rc;
  if (!cur.lik_ok) return hm_throw(R);
  int64_t iter = 0;
  while (true) {
    ++iter;
    HmPass nxt;
    if ((rc = hm_fixedpoint(R, cur, nxt))) return rc;
    if (nxt.lik < cur.lik)
      return ctx_fail(R.ctx, EQBF_ERR_DATA, "observed log-likelihood is decreasing (%f < %f)", nxt.lik, cur.lik);
    const bool stop = hm_stop(R, nxt.lik, cur.lik, iter);
    P = cur.next;
    cur = nxt;
    if (stop) break;
  }
  /* the last update_params, one more iteration and update_params */
  ++iter;
  HmPass fin;
  if ((rc = hm_fixedpoint(R, cur, fin))) return rc;
  P = cur.next;
  *lik_out = fin.lik;
  *iters_out = iter;
  return EQBF_OK;
}

/* run_EM_square (eqtlbma_hm.cpp:1215-1324): step length rule 3
 * (alpha = sqrt(|r|^2 / |v|^2), Varadhan and Roland 2008), stepmin0 =
 * stepmax0 = 1, mstep = 4, maxdist = 1.  After an extrapolation rejected for
 * falling too far below the start, the parameters return to the second EM
 * iterate while the genes keep the likelihoods of the last evaluation, which
 * the next update reads (a stale pass). */
int hm_square(HmRun& R, HmParams& P, double* lik_out, int64_t* iters_out) {
  const double stepmin0 = 1, stepmax0 = 1, mstep = 4, maxdist = 1, stepmax_user = R.o->stepmax;
  const bool dimv = R.types || R.h.K > 1;
  int64_t iter = 0;
  double steplen = 1.0, stepmin = stepmin0, stepmax = stepmax0;
  HmPass cur;
  int rc = R.pass(P, cur);
  if (rc) return rc;
  if (!cur.lik_ok) return hm_throw(R);
  double lik = cur.lik;
  while (true) {
    const double lik0 = lik;
    const HmParams P0 = P;
    ++iter;
    HmPass c1;
    if ((rc = hm_fixedpoint(R, cur, c1))) return rc;
    P = cur.next;
    lik = c1.lik;
    cur = c1;
    const HmParams P1 = P;
    const double lik1 = lik;
    if (hm_stop(R, lik1, lik0, iter)) break;
    ++iter;
    HmPass c2;
    if ((rc = hm_fixedpoint(R, cur, c2))) return rc;
    P = cur.next;
    lik = c2.lik;
    cur = c2;
    const HmParams P2 = P;
    const double lik2 = lik;
    if (hm_stop(R, lik2, lik1, iter)) break;
    /* compute_steplength */
    double sr2 = sq(P1.pi0 - P0.pi0), sv2 = sq(P2.pi0 - 2 * P1.pi0 + P0.pi0);
    auto acc = [&](const std::vector<double>& a0, const std::vector<double>& a1, const std::vector<double>& a2) {
      for (size_t k = 0; k < a0.size(); ++k) {
        sr2 += sq(a1[k] - a0[k]);
        sv2 += sq(a2[k] - 2 * a1[k] + a0[k]);
      }
    };
    if (dimv) acc(P0.dimw, P1.dimw, P2.dimw);
    if (R.types) acc(P0.sub, P1.sub, P2.sub);
    acc(P0.grid, P1.grid, P2.grid);
    double alpha = std::sqrt(sr2 / sv2);
    alpha = std::max(stepmin, std::min(stepmax, alpha));
    alpha = std::min(stepmax_user, alpha);
    steplen = alpha;
    /* proposals_squarem */
    HmParams Pn = P2;
    auto prop = [&](double a0, double a1, double a2) {
      return a0 + 2.0 * alpha * (a1 - a0) + std::pow(alpha, 2) * (a2 - 2.0 * a1 + a0);
    };
    auto propv = [&](const std::vector<double>& a0, const std::vector<double>& a1, const std::vector<double>& a2,
                     std::vector<double>& out) {
      for (size_t k = 0; k < a0.size(); ++k) out[k] = prop(a0[k], a1[k], a2[k]);
    };
    Pn.pi0 = prop(P0.pi0, P1.pi0, P2.pi0);
    if (dimv) propv(P0.dimw, P1.dimw, P2.dimw, Pn.dimw);
    if (R.types) propv(P0.sub, P1.sub, P2.sub, Pn.sub);
    propv(P0.grid, P1.grid, P2.grid, Pn.grid);
    HmPass cn;
    if ((rc = R.pass(Pn, cn))) return rc;
    if (!cn.lik_ok) return hm_throw(R);
    P = Pn;
    lik = cn.lik;
    cur = cn;
    bool extrap = true;
    if (std::fabs(steplen - 1) > 0.01) {
      /* consistency check: one EM iteration from the proposal, which on
       * failure goes back to the second EM iterate */
      ++iter;
      HmPass cc;
      bool failed = !cn.upd_ok;
      if (!failed) {
        if ((rc = R.pass(cn.next, cc))) return rc;
        failed = !cc.lik_ok;
      }
      if (failed) {
        P = P2;
        HmPass cb;
        if ((rc = R.pass(P, cb))) return rc;
        if (!cb.lik_ok) return hm_throw(R);
        lik = cb.lik;
        cur = cb;
        if (steplen == stepmax) stepmax = std::max(stepmax0, stepmax / mstep);
        steplen = 1;
        if (steplen == stepmax) stepmax = mstep * stepmax;
        if (stepmin < 0 && steplen == stepmin) stepmin = mstep * stepmin;
        continue;
      }
      P = cn.next;
      lik = cc.lik;
      cur = cc;
    }
    if (extrap && lik < lik0 - maxdist) {
      lik = lik2;
      P = P2;
      if ((rc = R.pass(P, cur, true))) return rc;
      cur.lik = lik2;
      if (steplen == stepmax) stepmax = std::max(stepmax0, stepmax / mstep);
      steplen = 1;
    }
    if (steplen == stepmax) stepmax = mstep * stepmax;
    if (stepmin < 0 && steplen == stepmin) stepmin = mstep * stepmin;
  }
  ++iter;
  HmPass fin;
  if ((rc = hm_fixedpoint(R, cur, fin))) return rc;
  P = cur.next;
  *lik_out = fin.lik;
  *iters_out = iter;
  return EQBF_OK;
}

int hm_setup(eqbf_ctx* ctx, HmRun& R, const double* raw_dev, int64_t n_pairs, const int64_t* off_dev, int64_t G,
             const eqbf_hm_opts* o, hipStream_t st) {
  if (!o || G < 1 || n_pairs < 1 || o->n_grid < 1 || o->dim < 1)
    return ctx_fail(ctx, EQBF_ERR_ARG, "eqbf_hm: empty data or invalid options");
  const bool types = o->model == EQBF_HM_TYPES;
  if (!types && o->model != EQBF_HM_CONFIGS) return ctx_fail(ctx, EQBF_ERR_ARG, "eqbf_hm: unknown model %d", o->model);
  if (types && (o->dim > HM_TYPES_KMAX || o->n_grid > 64 || o->n_subgroups < 1))
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "types model: at most %d types and 64 grid points", HM_TYPES_KMAX);
  R.ctx = ctx;
  R.st = st;
  R.o = o;
  R.types = types;
  HmDev& h = R.h;
  memset(&h, 0, sizeof h);
  h.raw = raw_dev;
  h.off = off_dev;
  h.NP = n_pairs;
  h.G = G;
  h.K = o->dim;
  h.L = o->n_grid;
  h.S = types ? o->n_subgroups : 0;
  h.D = types ? h.S : h.K;
  R.NC = h.K + h.L + (types ? h.K * h.S : 0);
  /* configs model: tiles of up to 16 pairs within 40 KB of LDS (four
   * workgroups per CU) and HM_PRE raw values per thread; a pair past that
   * takes the wave kernel */
  if (!types) {
    const size_t per = (size_t)(h.K * (h.L | 1) + 2 * h.K + 2) * sizeof(double);
    const size_t room = (40u << 10) - std::min<size_t>(40u << 10, (size_t)(h.K + h.L) * sizeof(double));
    h.tp = (int)std::min<size_t>(std::min<size_t>(16, room / per), (size_t)(HM_PRE * HM_NT) / (h.K * h.L));
  }
  h.par = (const double*)ctx_buf(ctx, "hm_par", (size_t)(1 + h.L + h.K + h.K * h.S) * sizeof(double));
  h.cfg = (double*)ctx_buf(ctx, "hm_cfg", (size_t)n_pairs * h.K * sizeof(double));
  h.grd = (double*)ctx_buf(ctx, "hm_grd", (size_t)n_pairs * h.L * sizeof(double));
  h.num = types ? (double*)ctx_buf(ctx, "hm_num", (size_t)n_pairs * h.K * h.S * sizeof(double)) : nullptr;
  h.snp = (double*)ctx_buf(ctx, "hm_snp", (size_t)n_pairs * sizeof(double));
  h.rows = (double*)ctx_buf(ctx, "hm_rows", (size_t)G * (R.NC + 3) * sizeof(double));
  h.red = (double*)ctx_buf(ctx, "hm_red", (size_t)(R.NC + 2) * sizeof(double));
  if (!h.par || !h.cfg || !h.grd || (types && !h.num) || !h.snp || !h.rows || !h.red) return EQBF_ERR_HIP;
  return EQBF_OK;
}

void hm_load(const eqbf_hm_opts* o, const eqbf_hm_params* p, HmParams& P) {
  const bool types = o->model == EQBF_HM_TYPES;
  P.pi0 = p->pi0;
  P.grid.assign(p->grid, p->grid + o->n_grid);
  const double* d = types ? p->type : p->config;
  P.dimw.assign(d, d + o->dim);
  if (types) P.sub.assign(p->subgroup, p->subgroup + (size_t)o->dim * o->n_subgroups);
}

void hm_store(const eqbf_hm_opts* o, const HmParams& P, eqbf_hm_params* p) {
  const bool types = o->model == EQBF_HM_TYPES;
  p->pi0 = P.pi0;
  std::memcpy(p->grid, P.grid.data(), P.grid.size() * sizeof(double));
  std::memcpy(types ? p->type : p->config, P.dimw.data(), P.dimw.size() * sizeof(double));
  if (types) std::memcpy(p->subgroup, P.sub.data(), P.sub.size() * sizeof(double));
}

int hm_check_params(eqbf_ctx* ctx, const eqbf_hm_opts* o, const eqbf_hm_params* p) {
  if (!p || !p->grid || (o->model == EQBF_HM_TYPES ? (!p->type || !p->subgroup) : !p->config))
    return ctx_fail(ctx, EQBF_ERR_ARG, "eqbf_hm: missing parameter arrays");
  return EQBF_OK;
}

int hm_em_run(eqbf_ctx* ctx, const double* raw_dev, int64_t n_pairs, const int64_t* off_dev, int64_t G,
              const eqbf_hm_opts* o, eqbf_hm_params* p, double* lik, int64_t* iters, hipStream_t st) {
  HmRun R;
  int rc = hm_setup(ctx, R, raw_dev, n_pairs, off_dev, G, o, st);
  if (rc) return rc;
  if ((rc = hm_check_params(ctx, o, p))) return rc;
  HmParams P;
  hm_load(o, p, P);
  double l = NAN;
  int64_t it = 0;
  rc = o->stepmax == 1.0 ? hm_classic(R, P, &l, &it) : hm_square(R, P, &l, &it);
  if (rc) return rc;
  hm_store(o, P, p);
  if (lik) *lik = l;
  if (iters) *iters = it;
  return EQBF_OK;
}

/* raw log10 ABFs and gene offsets from host arrays into the context's
 * buffers */
int hm_upload(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset, int64_t n_genes,
              const eqbf_hm_opts* opts, const double** d_raw_out, const int64_t** d_off_out) {
  const int D = opts->model == EQBF_HM_TYPES ? opts->n_subgroups : opts->dim;
  if (n_pairs < 1 || n_genes < 1 || D < 1 || opts->n_grid < 1)
    return ctx_fail(ctx, EQBF_ERR_ARG, "eqbf_hm: empty data or invalid options");
  const size_t rb = (size_t)n_pairs * D * opts->n_grid * sizeof(double);
  double* d_raw = (double*)ctx_buf(ctx, "hm_raw", rb);
  int64_t* d_off = (int64_t*)ctx_buf(ctx, "hm_off", (size_t)(n_genes + 1) * sizeof(int64_t));
  if (!d_raw || !d_off) return EQBF_ERR_HIP;
  EQBF_HIP(ctx, hipMemcpyAsync(d_raw, raw, rb, hipMemcpyHostToDevice, ctx->stream));
  EQBF_HIP(ctx, hipMemcpyAsync(d_off, gene_offset, (size_t)(n_genes + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                               ctx->stream));
  *d_raw_out = d_raw;
  *d_off_out = d_off;
  return EQBF_OK;
}

/* Controller::estimate_profile_ci (eqtlbma_hm.cpp:1348-1546): from the
 * estimates, each of pi0, the configuration weights and the grid weights
 * walks down and up in steps of 0.001 (the other weights of a simplex
 * rescaled to keep their sum) while the natural-log likelihood stays
 * within 2 of its maximum; types and subgroup probabilities keep their
 * estimate as both bounds. */
int hm_ci(HmRun& R, const HmParams& P, double* left, double* right) {
  const double tick = 0.001, le = std::log10(std::exp(1.0));
  double mx;
  int rc = hm_loglik(R, P, &mx);
  if (rc) return rc;
  if (!std::isfinite(mx)) return hm_throw(R);
  HmParams Q = P;
  auto eval = [&](double* v) -> int {
    const int e = hm_loglik(R, Q, v);
    if (e) return e;
    return std::isfinite(*v) ? EQBF_OK : hm_throw(R);
  };
  auto below = [&](double v) { return v / le < mx / le - 2.0; };
  {
    double l = P.pi0, r = P.pi0, v;
    while (l >= 0) {
      l -= tick;
      if (l < 0) {
        l = 0;
        break;
      }
      Q.pi0 = l;
      if ((rc = eval(&v))) return rc;
      if (below(v)) {
        l += tick;
        break;
      }
    }
    while (r <= 1) {
      r += tick;
      if (r > 1) {
        r = 1;
        break;
      }
      Q.pi0 = r;
      if ((rc = eval(&v))) return rc;
      if (below(v)) {
        r -= tick;
        break;
      }
    }
    Q.pi0 = P.pi0;
    left[0] = l;
    right[0] = r;
  }
  /* one weight of a simplex, the others rescaled by (mle_i - x) / (1 - mle_i) */
  auto walk = [&](std::vector<double>& w, const std::vector<double>& mle, double* lo, double* hi) -> int {
    for (size_t i = 0; i < mle.size(); ++i) {
      double l = mle[i], r = mle[i], v;
      const double cp = mle[i], st = 1 - cp;
      auto set = [&](double x) {
        const double diff = cp - x;
        for (size_t j = 0; j < mle.size(); ++j)
          if (j != i) w[j] = mle[j] + diff * mle[j] / st;
        w[i] = x;
      };
      while (l >= 0) {
        l -= tick;
        if (l < 0) {
          l = 0;
          break;
        }
        set(l);
        if ((rc = eval(&v))) return rc;
        if (below(v)) {
          l += tick;
          break;
        }
      }
      while (r <= 1) {
        r += tick;
        if (r > 1) {
          r = 1;
          break;
        }
        set(r);
        if ((rc = eval(&v))) return rc;
        if (below(v)) {
          r -= tick;
          break;
        }
      }
      lo[i] = l;
      hi[i] = r;
    }
    w = mle;
    return EQBF_OK;
  };
  const int K = R.h.K, S = R.h.S, L = R.h.L;
  if (!R.types) {
    if ((rc = walk(Q.dimw, P.dimw, left + 1, right + 1))) return rc;
  } else {
    for (int k = 0; k < K; ++k) left[1 + k] = right[1 + k] = P.dimw[k];
    for (int i = 0; i < K * S; ++i) left[1 + K + i] = right[1 + K + i] = P.sub[i];
  }
  const int g0 = 1 + K + (R.types ? K * S : 0);
  return walk(Q.grid, P.grid, left + g0, right + g0);
}

}  // namespace

extern "C" int eqbf_hm_em_device(eqbf_ctx* ctx, const double* raw_dev, int64_t n_pairs, const int64_t* gene_offset_dev,
                                 int64_t n_genes, const eqbf_hm_opts* opts, eqbf_hm_params* params, double* log10_lik,
                                 int64_t* n_iters) {
  if (!ctx) return EQBF_ERR_ARG;
  return hm_em_run(ctx, raw_dev, n_pairs, gene_offset_dev, n_genes, opts, params, log10_lik, n_iters, ctx->stream);
}

extern "C" int eqbf_hm_em(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset,
                          int64_t n_genes, const eqbf_hm_opts* opts, eqbf_hm_params* params, double* log10_lik,
                          int64_t* n_iters) {
  if (!ctx || !raw || !gene_offset || !opts) return EQBF_ERR_ARG;
  const double* d_raw;
  const int64_t* d_off;
  const int rc = hm_upload(ctx, raw, n_pairs, gene_offset, n_genes, opts, &d_raw, &d_off);
  if (rc) return rc;
  return hm_em_run(ctx, d_raw, n_pairs, d_off, n_genes, opts, params, log10_lik, n_iters, ctx->stream);
}

extern "C" int eqbf_hm_bfs(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset,
                           int64_t n_genes, const eqbf_hm_opts* opts, const eqbf_hm_params* params, double* gene_post,
                           double* gene_l10bf, double* snp_l10bf, double* dim_l10bf) {
  if (!ctx || !raw || !gene_offset || !opts) return EQBF_ERR_ARG;
  int rc = hm_check_params(ctx, opts, params);
  if (rc) return rc;
  const double* d_raw;
  const int64_t* d_off;
  if ((rc = hm_upload(ctx, raw, n_pairs, gene_offset, n_genes, opts, &d_raw, &d_off))) return rc;
  hipStream_t st = ctx->stream;
  HmRun R;
  if ((rc = hm_setup(ctx, R, d_raw, n_pairs, d_off, n_genes, opts, st))) return rc;
  HmParams P;
  hm_load(opts, params, P);
  if ((rc = R.upload(P))) return rc;
  if ((rc = R.launch())) return rc;
  const int NC = R.NC;
  std::vector<double> rows((size_t)n_genes * (NC + 3));
  EQBF_HIP(ctx, hipMemcpyAsync(rows.data(), R.h.rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  if (snp_l10bf) EQBF_HIP(ctx, hipMemcpyAsync(snp_l10bf, R.h.snp, n_pairs * sizeof(double), hipMemcpyDeviceToHost, st));
  if (dim_l10bf)
    EQBF_HIP(ctx, hipMemcpyAsync(dim_l10bf, R.h.cfg, (size_t)n_pairs * opts->dim * sizeof(double), hipMemcpyDeviceToHost,
                                 st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  for (int64_t g = 0; g < n_genes; ++g) {
    const double ol = rows[(size_t)g * (NC + 3) + NC], gb = rows[(size_t)g * (NC + 3) + NC + 2];
    if (gene_l10bf) gene_l10bf[g] = gb;
    if (gene_post) {
      const double pp = std::pow(10.0, std::log10(1.0 - P.pi0) + gb - ol);
      gene_post[g] = pp > 1.0 ? 1.0 : pp;
    }
  }
  return EQBF_OK;
}

extern "C" int eqbf_hm_ci(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset,
                          int64_t n_genes, const eqbf_hm_opts* opts, const eqbf_hm_params* params, double* left,
                          double* right) {
  if (!ctx || !raw || !gene_offset || !opts || !left || !right) return EQBF_ERR_ARG;
  int rc = hm_check_params(ctx, opts, params);
  if (rc) return rc;
  const double* d_raw;
  const int64_t* d_off;
  if ((rc = hm_upload(ctx, raw, n_pairs, gene_offset, n_genes, opts, &d_raw, &d_off))) return rc;
  HmRun R;
  if ((rc = hm_setup(ctx, R, d_raw, n_pairs, d_off, n_genes, opts, ctx->stream))) return rc;
  HmParams P;
  hm_load(opts, params, P);
  return hm_ci(R, P, left, right);
}
