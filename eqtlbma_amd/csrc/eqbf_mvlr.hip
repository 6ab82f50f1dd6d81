/*
 * Multivariate (MVLR) Bayes factors, --error mvlr, in FP64.
 *
 * Reference: GeneSnpPair::CalcAbfsMvlr and CalcAbfsMvlrFor*
 * (gene_snp_pair.cpp:624-758) over class MVLR (MVLR.cpp:26-764) with one
 * SNP per pair.  With T the projection off [1, covariates of the first
 * subgroup], x~ = T x, y~_s = T y_s, k = x~'x~ and b_s = y~_s'x~:
 *   Sigma0 = Y~'Y~ / n                      (compute_Sigma_null :176-195, m = q+s+1)
 *   residual of y_s on [Xc, x] (gamma_s = 1) or on Xc (gamma_s = 0):
 *     (E'E)_ij = (Y~'Y~)_ij - [gamma_i or gamma_j] b_i b_j / k
 *   F -> chi-square factor for gamma_s = 1 (compute_residual :366-410):
 *     T2 = (b_s^2/k) / sigma1^2 with sigma1 = rss_s/(n-q-1) (the reference
 *     divides by the square of a variance), F = T2,
 *     factor = F / chisq_Qinv(fdist_Q(F; 1, n-q-1); 1)
 *            = T2 / Phi^-1(P(T_{n-q-1} <= -sqrt T2))^2  (1 when F < 1e-8,
 *              0 when the F tail probability underflows in FP64)
 *   Sigma = fiterr [ (H m/(m+n) + E'E/(m+n)) C ] + (1 - fiterr) Sigma0,
 *     C = diag(sqrt(factor)), H = 1e-4 I (Sigma C, not C Sigma C: :268-273)
 *   eVb = Sigma^-1 b, t = gamma o sqrt(diag Sigma), W = oma2 t t' + phi2 diag(t^2),
 *   M = I + k Sigma^-1 W,  log10 ABF = (eVb' W M^-1 eVb - ln|det M|) / (2 ln 10)
 *                                                    (:497-512, :559-605)
 * so no n x n matrix is ever formed.  M is factorised with partial pivoting
 * (as gsl_linalg_LU_decomp); ln|det M| is summed from the pivots and
 * M^-1 eVb comes from one solve instead of an explicit inverse.
 *
 * Kernels (one chunk):
 *   mvlr_prep_kernel   one wave per gene: common members, basis of
 *                      [1, covariates], y~_s, Y~'Y~, Sigma0 and Sigma0^-1
 *   mvlr_sums_kernel   one thread per pair: k and b (FP64 sums over members)
 *   mvlr_cfg_kernel    one thread per (pair, configuration): Sigma, Sigma^-1,
 *                      Sigma^-1 t, t, eVb (structure-of-arrays, coalesced)
 *   mvlr_grid_kernel   one thread per (pair, raw row): M, LU, log10 ABF
 *   mvlr_wavg_kernel   one thread per (pair, weighted row): grid average
 *                      (utils::log10_weighted_sum)
 * followed by the BMA kernel of the univariate path (gen-sin, all).
 */
#include <algorithm>
#include <type_traits>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

enum { MV_OK = 0, MV_SKIP = 1, MV_MISMATCH = 2, MV_FEW = 3 };
struct MvRec {
  int32_t n, flags, qcov, rank;
};

__device__ inline double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* LU factorisation with partial pivoting of an SM x SM register matrix
 * (largest |a| of the column, first one on ties).  piv[j] = row swapped
 * with j at step j.  Returns ln|det A|. */
template <int SM>
__device__ inline double lu_factor(double (&A)[SM][SM], int (&piv)[SM]) {
  double lnd = 0.0;
#pragma unroll
  for (int j = 0; j < SM; ++j) {
    int p = j;
    double best = fabs(A[j][j]);
#pragma unroll
    for (int i = j + 1; i < SM; ++i)
      if (fabs(A[i][j]) > best) {
        best = fabs(A[i][j]);
        p = i;
      }
    piv[j] = p;
#pragma unroll
    for (int i = j + 1; i < SM; ++i) {
      const bool sw = p == i;
#pragma unroll
      for (int c = 0; c < SM; ++c) {
        const double a = A[j][c], b = A[i][c];
        A[j][c] = sw ? b : a;
        A[i][c] = sw ? a : b;
      }
    }
    const double d = A[j][j];
    lnd += log(fabs(d));
#pragma unroll
    for (int i = j + 1; i < SM; ++i) {
      const double f = A[i][j] / d;
      A[i][j] = f;
#pragma unroll
      for (int c = j + 1; c < SM; ++c) A[i][c] -= f * A[j][c];
    }
  }
  return lnd;
}

/* Solves A x = rhs in place from the factorisation of lu_factor. */
template <int SM>
__device__ inline void lu_apply(const double (&A)[SM][SM], const int (&piv)[SM], double (&x)[SM]) {
#pragma unroll
  for (int j = 0; j < SM; ++j) {
#pragma unroll
    for (int i = j + 1; i < SM; ++i) {
      const bool sw = piv[j] == i;
      const double a = x[j], b = x[i];
      x[j] = sw ? b : a;
      x[i] = sw ? a : b;
    }
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int c = 0; c < i; ++c) x[i] -= A[i][c] * x[c];
#pragma unroll
  for (int i = SM - 1; i >= 0; --i) {
    double s = x[i];
#pragma unroll
    for (int c = i + 1; c < SM; ++c) s -= A[i][c] * x[c];
    x[i] = s / A[i][i];
  }
}

__global__ void mv_zero_kernel(int* p) { *p = 0; }

/* ---- per-gene preparation: one wave per gene ------------------------------------ */
struct MvPrep {
  MvRec* mrec;
  int32_t* midx; /* [G][N] member sample indices            */
  double* Ut;    /* [G][KU][N] orthonormal basis columns     */
  double* Yt;    /* [G][S][N] y~_s over the members          */
  double* YtY;   /* [G][S][S]                                */
  double* S0;    /* [G][S][S] Sigma0                         */
  double* S0inv; /* [G][S][S]                                */
  int* flag;
  int KU;
};

template <int SM>
__global__ void __launch_bounds__(64) mvlr_prep_kernel(const eqbf_chunk c, const MvPrep pp) {
  const int S = c.n_subgroups, N = c.n_samples, g = blockIdx.x, lane = threadIdx.x;
  const int KU = pp.KU;
  MvRec r;
  r.n = 0;
  r.flags = MV_OK;
  r.qcov = min(max(c.covar_count[0], 0), c.n_covars_max);
  r.rank = 0;
  for (int s = 0; s < S; ++s)
    if (!c.gene_in_sbgrp[(int64_t)g * S + s]) r.flags = MV_SKIP;
  if (r.flags != MV_OK) {
    if (lane == 0) pp.mrec[g] = r;
    return;
  }
  /* members of subgroup 0 (expression level not NaN, genotyped); every other
   * subgroup must have the same ones */
  int32_t* idx = pp.midx + (int64_t)g * N;
  int n = 0;
  bool mismatch = false;
  for (int base = 0; base < N; base += 64) {
    const int i = base + lane;
    bool m0 = false, diff = false;
    if (i < N) {
      m0 = !isnan(c.explevels[((int64_t)g * S) * N + i]) && c.geno_in_sbgrp[i];
      for (int s = 1; s < S; ++s) {
        const bool ms = !isnan(c.explevels[((int64_t)g * S + s) * N + i]) && c.geno_in_sbgrp[(int64_t)s * N + i];
        diff |= ms != m0;
      }
    }
    mismatch |= __ballot(diff) != 0ull;
    const uint64_t bal = __ballot(m0);
    if (m0) idx[n + __popcll(bal & ((1ull << lane) - 1ull))] = i;
    n += __popcll(bal);
  }
  __syncthreads();
  r.n = n;
  const int Q = r.qcov;
  if (mismatch)
    r.flags = MV_MISMATCH;
  else if (n < Q + 3) /* no residual degree of freedom for the F statistic */
    r.flags = MV_FEW;
  if (r.flags != MV_OK) {
    if (lane == 0) pp.mrec[g] = r;
    return;
  }
  if (Q > 0) { /* a member without covariates is a data error (gene_snp_pair.cpp:137-144) */
    int miss = 0;
    for (int k = lane; k < n; k += 64) miss |= !c.covar_in_sbgrp[idx[k]];
    if (__ballot(miss) != 0ull) {
      r.flags = MV_SKIP;
      if (lane == 0) {
        atomicOr(pp.flag, FLAG_MISSING_COVAR);
        pp.mrec[g] = r;
      }
      return;
    }
  }
  double* U = pp.Ut + (int64_t)g * KU * N;
  double* Y = pp.Yt + (int64_t)g * S * N;
  /* orthonormal basis of [1, covariates] by two-pass modified Gram-Schmidt;
   * each lane only touches its own rows k = lane + 64 m */
  uint32_t kept = 0;
  for (int j = 0; j <= Q; ++j) {
    double* cj = U + (int64_t)j * N;
    for (int k = lane; k < n; k += 64)
      cj[k] = (j == 0) ? 1.0 : c.covars[((int64_t)(j - 1)) * N + idx[k]];
    double nn = 0;
    for (int k = lane; k < n; k += 64) nn += cj[k] * cj[k];
    const double n0 = sqrt(wsum64(nn));
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < j; ++l) {
        if (!((kept >> l) & 1u)) continue;
        const double* cl = U + (int64_t)l * N;
        double d = 0;
        for (int k = lane; k < n; k += 64) d += cl[k] * cj[k];
        d = wsum64(d);
        for (int k = lane; k < n; k += 64) cj[k] -= d * cl[k];
      }
    double nv = 0;
    for (int k = lane; k < n; k += 64) nv += cj[k] * cj[k];
    nv = sqrt(wsum64(nv));
    const bool keep = nv > 1e-9 * n0 && nv > 0.0;
    const double sc = keep ? 1.0 / nv : 0.0;
    for (int k = lane; k < n; k += 64) cj[k] *= sc;
    if (keep) {
      kept |= 1u << j;
      ++r.rank;
    }
  }
  for (int j = Q + 1; j < KU; ++j)
    for (int k = lane; k < n; k += 64) U[(int64_t)j * N + k] = 0.0;
  for (int s = 0; s < S; ++s) {
    double* ys = Y + (int64_t)s * N;
    for (int k = lane; k < n; k += 64) ys[k] = c.explevels[((int64_t)g * S + s) * N + idx[k]];
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l <= Q; ++l) {
        if (!((kept >> l) & 1u)) continue;
        const double* cl = U + (int64_t)l * N;
        double d = 0;
        for (int k = lane; k < n; k += 64) d += cl[k] * ys[k];
        d = wsum64(d);
        for (int k = lane; k < n; k += 64) ys[k] -= d * cl[k];
      }
  }
  double* yty = pp.YtY + (int64_t)g * S * S;
  for (int a = 0; a < S; ++a)
    for (int b = a; b < S; ++b) {
      const double* ya = Y + (int64_t)a * N;
      const double* yb = Y + (int64_t)b * N;
      double d = 0;
      for (int k = lane; k < n; k += 64) d += ya[k] * yb[k];
      d = wsum64(d);
      if (lane == 0) {
        yty[a * S + b] = d;
        yty[b * S + a] = d;
      }
    }
  if (lane == 0) {
    /* Sigma0 = Y'TY / (n + m - q - s - 1) = Y'TY / n, m = q + s + 1 */
    double A[SM][SM];
    int piv[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SM; ++j) A[i][j] = (i < S && j < S) ? yty[i * S + j] / (double)n : (i == j ? 1.0 : 0.0);
    for (int i = 0; i < S; ++i)
      for (int j = 0; j < S; ++j) pp.S0[((int64_t)g * S + i) * S + j] = yty[i * S + j] / (double)n;
    (void)lu_factor<SM>(A, piv);
    for (int col = 0; col < S; ++col) {
      double e[SM];
#pragma unroll
      for (int i = 0; i < SM; ++i) e[i] = (i == col) ? 1.0 : 0.0;
      lu_apply<SM>(A, piv, e);
#pragma unroll
      for (int i = 0; i < SM; ++i)
        if (i < S) pp.S0inv[((int64_t)g * S + i) * S + col] = e[i];
    }
    pp.mrec[g] = r;
  }
}

/* ---- per pair: k and b over the members ------------------------------------------ */
struct MvSums {
  const MvRec* mrec;
  const int32_t* midx;
  const double* Ut;
  const double* Yt;
  double* kb;         /* [S + 1][NP]: k, b_0 .. b_{S-1}          */
  int32_t* pgene;     /* [NP]: gene of the pair, -1 = no result    */
  int32_t* n_out;
  double* sstats;
  double* stdsstats;
  int* flag;
  int KU;
};

__device__ inline int gene_of_pair(const int64_t* __restrict__ off, int G, int64_t p) {
  int lo = 0, hi = G - 1; /* largest g with off[g] <= p */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int SM>
__global__ void __launch_bounds__(256) mvlr_sums_kernel(const eqbf_chunk c, const MvSums ms) {
  const int64_t NP = c.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const int S = c.n_subgroups, N = c.n_samples;
  const int g = gene_of_pair(c.pair_offset, c.n_genes, p);
  const int64_t snp = c.pair_snp[p];
  const MvRec r = ms.mrec[g];
  bool all_geno = true;
  for (int s = 0; s < S; ++s) all_geno &= c.snp_in_sbgrp[snp * S + s] != 0;
  if (r.flags == MV_MISMATCH && all_geno) atomicOr(ms.flag, FLAG_MVLR_SAMPLES);
  const bool has = r.flags != MV_SKIP && r.flags != MV_MISMATCH && all_geno;
  /* the reference records the sample size of each subgroup and NaN summary
   * statistics for --error mvlr (gene_snp_pair.cpp:154-158) */
  for (int s = 0; s < S; ++s) {
    ms.n_out[(int64_t)s * NP + p] = has ? r.n : 0;
    if (ms.sstats)
      for (int k = 0; k < 5; ++k) ms.sstats[((int64_t)k * S + s) * NP + p] = NAN;
    if (ms.stdsstats)
      for (int k = 0; k < 3; ++k) ms.stdsstats[((int64_t)k * S + s) * NP + p] = NAN;
  }
  if (!has || r.flags == MV_FEW) {
    ms.pgene[p] = -1;
    return;
  }
  const int n = r.n;
  const int32_t* idx = ms.midx + (int64_t)g * N;
  const double* U = ms.Ut + (int64_t)g * ms.KU * N;
  const double* Y = ms.Yt + (int64_t)g * S * N;
  const int64_t gbase = (int64_t)c.geno_set_of_sbgrp[0] * c.geno_set_stride + snp;
  const bool f32 = c.geno_dtype == EQBF_F32;
  double xx = 0, b[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) b[s] = 0;
  for (int k = 0; k < n; ++k) {
    const int64_t off = gbase + (int64_t)idx[k] * c.geno_ld;
    const double x = f32 ? (double)((const float*)c.genotypes)[off] : ((const double*)c.genotypes)[off];
    xx += x * x;
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < S) b[s] += x * Y[(int64_t)s * N + k];
  }
  double proj = 0;
  for (int q = 0; q < ms.KU; ++q) {
    const double* uq = U + (int64_t)q * N;
    double a = 0;
    for (int k = 0; k < n; ++k) {
      const int64_t off = gbase + (int64_t)idx[k] * c.geno_ld;
      const double x = f32 ? (double)((const float*)c.genotypes)[off] : ((const double*)c.genotypes)[off];
      a += x * uq[k];
    }
    proj += a * a;
  }
  ms.kb[p] = xx - proj;
#pragma unroll
  for (int s = 0; s < SM; ++s)
    if (s < S) ms.kb[(int64_t)(1 + s) * NP + p] = b[s];
  ms.pgene[p] = g;
}

/* ---- per (pair, configuration): Sigma, Sigma^-1, Sigma^-1 t, t, eVb -------------- */
struct MvCfg {
  const MvRec* mrec;
  const double* YtY;
  const double* S0;
  const double* S0inv;
  const double* kb;
  const int32_t* pgene;
  const double* tz;
  const uint64_t* masks; /* [1 + n_cfg]: the consistent configuration first */
  double* par;           /* [cfg in batch][PSZ][NP]                          */
  double fiterr;
  int cfg0, PSZ;
};

/* Calls launch(std::integral_constant<int, E>) with E the exact subgroup
 * count where SM pads it (5..7 in 8, 9..11 in 12), else E = SM: the
 * per-(pair, configuration) kernels then hold E x E matrices, fewer
 * registers and no identity-padded LU. */
template <int SM, class Launch>
void with_exact_sm(int S, Launch&& launch) {
  if constexpr (SM == 8) {
    if (S == 5) return launch(std::integral_constant<int, 5>{});
    if (S == 6) return launch(std::integral_constant<int, 6>{});
    if (S == 7) return launch(std::integral_constant<int, 7>{});
  } else if constexpr (SM == 12) {
    if (S == 9) return launch(std::integral_constant<int, 9>{});
    if (S == 10) return launch(std::integral_constant<int, 10>{});
    if (S == 11) return launch(std::integral_constant<int, 11>{});
  }
  launch(std::integral_constant<int, SM>{});
}

template <int SM>
__global__ void __launch_bounds__(64) mvlr_cfg_kernel(const eqbf_chunk c, const MvCfg mc) {
  const int64_t NP = c.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const int ci = blockIdx.y;
  const int S = c.n_subgroups;
  const int PSZ = mc.PSZ;
  double* out = mc.par + (int64_t)ci * PSZ * NP + p;
  const int g = mc.pgene[p];
  if (g < 0) {
    out[(int64_t)(PSZ - 1) * NP] = 1.0;
    return;
  }
  const MvRec r = mc.mrec[g];
  const uint64_t gm = mc.masks[mc.cfg0 + ci];
  const double kk = mc.kb[p];
  double b[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) b[s] = (s < S) ? mc.kb[(int64_t)(1 + s) * NP + p] : 0.0;
  const double* YY = mc.YtY + (int64_t)g * S * S;
  const double* S0 = mc.S0 + (int64_t)g * S * S;
  double Sg[SM][SM];
  const bool fit = !(mc.fiterr < 1e-6);
  if (!fit) {
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SM; ++j) Sg[i][j] = (i < S && j < S) ? S0[i * S + j] : (i == j ? 1.0 : 0.0);
  } else {
    const double nn = (double)r.n;
    const double m = (double)(r.qcov + 1 + S + 1);
    const int nu = r.n - (r.qcov + 1) - 1;
    double cf[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      cf[i] = 1.0;
      if (i < S && ((gm >> i) & 1ull)) {
        const double b2k = b[i] * b[i] / kk;
        const double s1 = (YY[i * S + i] - b2k) / (double)nu;
        const double T2 = b2k / (s1 * s1);
        if (!(T2 < 1e-8)) {
          const double w = dm::tz_eval(mc.tz, nu, sqrt(T2)); /* -ln P(T_nu <= -sqrt T2) */
          if (2.0 * exp(-w) == 0.0) {
            cf[i] = 0.0; /* the F tail underflows: chi-square quantile +inf */
          } else {
            const double z = dm::gauss_pinv_from_w(w);
            cf[i] = sqrt(T2 / (z * z));
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SM; ++j) {
        if (i < S && j < S) {
          const bool any = ((gm >> i) & 1ull) || ((gm >> j) & 1ull);
          const double ee = YY[i * S + j] - (any ? b[i] * b[j] / kk : 0.0);
          const double h = (i == j) ? 1e-4 : 0.0;
          const double sm = (h * m / (m + nn) + ee / (m + nn)) * cf[j];
          Sg[i][j] = mc.fiterr * sm + (1.0 - mc.fiterr) * S0[i * S + j];
        } else {
          Sg[i][j] = (i == j) ? 1.0 : 0.0;
        }
      }
  }
  double tv[SM];
#pragma unroll
  for (int j = 0; j < SM; ++j) tv[j] = (j < S && ((gm >> j) & 1ull)) ? sqrt(Sg[j][j]) : 0.0;
  double u[SM], eVb[SM];
#pragma unroll
  for (int i = 0; i < SM; ++i) u[i] = eVb[i] = 0.0;
  if (!fit) {
    const double* Si = mc.S0inv + (int64_t)g * S * S;
#pragma unroll
    for (int i = 0; i < SM; ++i)
      if (i < S) {
#pragma unroll
        for (int j = 0; j < SM; ++j)
          if (j < S) {
            const double v = Si[i * S + j];
            out[(int64_t)(i * S + j) * NP] = v;
            u[i] += v * tv[j];
            eVb[i] += v * b[j];
          }
      }
  } else {
    int piv[SM];
    (void)lu_factor<SM>(Sg, piv);
    for (int col = 0; col < S; ++col) {
      double e[SM];
#pragma unroll
      for (int i = 0; i < SM; ++i) e[i] = (i == col) ? 1.0 : 0.0;
      lu_apply<SM>(Sg, piv, e);
      double tc = 0, bc = 0;
#pragma unroll
      for (int j = 0; j < SM; ++j)
        if (j == col) {
          tc = tv[j];
          bc = b[j];
        }
#pragma unroll
      for (int i = 0; i < SM; ++i)
        if (i < S) {
          out[(int64_t)(i * S + col) * NP] = e[i];
          u[i] += e[i] * tc;
          eVb[i] += e[i] * bc;
        }
    }
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
    if (i < S) {
      out[(int64_t)(S * S + i) * NP] = u[i];
      out[(int64_t)(S * S + S + i) * NP] = tv[i];
      out[(int64_t)(S * S + 2 * S + i) * NP] = eVb[i];
    }
  out[(int64_t)(S * S + 3 * S) * NP] = kk;
  out[(int64_t)(PSZ - 1) * NP] = 0.0;
}

/* ---- per (pair, raw row): one log10 ABF ------------------------------------------ */
struct MvGrid {
  const double* par;
  const double* grid; /* phiL, omaL, phiS, omaS */
  double* raw;        /* [n_raw][NP]            */
  int GL, GS, cfg0, PSZ;
  int64_t row0, nrows, pair_blocks;
};

template <int SM>
__global__ void __launch_bounds__(64) mvlr_grid_kernel(const eqbf_chunk c, const MvGrid mg) {
  const int64_t NP = c.n_pairs;
  /* rows fastest: the blocks in flight share one tile of 64 pairs, whose
   * parameters stay in L2 across the rows */
  const int64_t pb = blockIdx.x / mg.nrows;
  const int64_t rr = blockIdx.x - pb * mg.nrows;
  const int64_t p = pb * 64 + threadIdx.x;
  if (p >= NP || rr >= mg.nrows) return;
  const int64_t row = mg.row0 + rr;
  const int S = c.n_subgroups;
  int cfg, gi, type = 0;
  double p2, o2;
  if (row < 3 * (int64_t)mg.GL) {
    cfg = 0;
    type = (int)(row / mg.GL);
    gi = (int)(row - (int64_t)type * mg.GL);
    p2 = mg.grid[gi];
    o2 = mg.grid[mg.GL + gi];
    if (type == 1) { /* gen-fix: phi2 = 0, oma2 = phi2 + oma2 */
      o2 = p2 + o2;
      p2 = 0.0;
    } else if (type == 2) { /* gen-maxh: phi2 = phi2 + oma2, oma2 = 0 */
      p2 = p2 + o2;
      o2 = 0.0;
    }
  } else {
    const int64_t rs = row - 3 * (int64_t)mg.GL;
    cfg = 1 + (int)(rs / mg.GS);
    gi = (int)(rs - (int64_t)(cfg - 1) * mg.GS);
    p2 = mg.grid[2 * mg.GL + gi];
    o2 = mg.grid[2 * mg.GL + mg.GS + gi];
  }
  const int PSZ = mg.PSZ;
  const double* pr = mg.par + (int64_t)(cfg - mg.cfg0) * PSZ * NP + p;
  double l = NAN;
  if (pr[(int64_t)(PSZ - 1) * NP] == 0.0) {
    const double kk = pr[(int64_t)(S * S + 3 * S) * NP];
    double t[SM], z[SM], u[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      t[i] = (i < S) ? pr[(int64_t)(S * S + S + i) * NP] : 0.0;
      u[i] = (i < S) ? pr[(int64_t)(S * S + i) * NP] : 0.0;
      z[i] = (i < S) ? pr[(int64_t)(S * S + 2 * S + i) * NP] : 0.0;
    }
    /* M = I + k Sigma^-1 (oma2 t t' + phi2 diag(t^2)) */
    double M[SM][SM];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SM; ++j) {
        double v = (i == j) ? 1.0 : 0.0;
        if (i < S && j < S) v += kk * (o2 * u[i] * t[j] + p2 * pr[(int64_t)(i * S + j) * NP] * t[j] * t[j]);
        M[i][j] = v;
      }
    double e[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) e[i] = z[i];
    int piv[SM];
    const double lnd = lu_factor<SM>(M, piv);
    lu_apply<SM>(M, piv, z); /* z = M^-1 eVb */
    /* eVb' W z with W = oma2 t t' + phi2 diag(t^2) */
    double et = 0, tz = 0, q2 = 0;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      et += e[i] * t[i];
      tz += t[i] * z[i];
      q2 += e[i] * t[i] * t[i] * z[i];
    }
    const double quad = o2 * et * tz + p2 * q2;
    l = (0.5 * quad - 0.5 * lnd) / 2.302585092994045684;
  }
  mg.raw[row * NP + p] = l;
}

/* ---- per (pair, group of rows): Hessenberg form, every grid point ----------------
 * With B = k Sigma^-1 T^2 (T = diag t) and the rank-one term o2 k u t'
 * (u = Sigma^-1 t), the matrix of every row is M = I + p2 B + o2 (k u) t'.
 * An orthogonal similarity Q'BQ = H (Householder reflections, upper
 * Hessenberg) is computed once per (pair, configuration); then for each grid
 * point A = I + p2 H is factorised with partial pivoting in O(S^2) (only
 * adjacent rows can swap), the rank-one term is added by the matrix
 * determinant lemma and Sherman-Morrison, and
 *   ln|det M| = ln|det A| + ln|1 + o2 t~'A^-1 u~|,
 *   M^-1 eVb  = Q (A^-1 z~ - o2 A^-1 u~ (t~'A^-1 z~) / (1 + o2 t~'A^-1 u~)),
 * with u~ = k Q'u, z~ = Q'eVb, t~ = Q't; the quadratic form eVb'W M^-1 eVb
 * = o2 (eVb't)(t~'x~) + p2 (Q'T^2 eVb)'x~.  Rows of the gen-fix type
 * (p2 = 0) reduce to scalars.  The row kernel above builds and factorises
 * each M from scratch (O(S^3) per row); this one gives the same values to
 * rounding.  Groups: for the consistent configuration gen, gen-fix and
 * gen-maxh (GL rows each), for the others their GS rows. */
template <int SM>
struct HessIdx {
  /* position of entry (i, j), j >= i - 1, of an upper Hessenberg SM x SM matrix */
  static constexpr int at(int i, int j) { return i == 0 ? j : SM + (i - 1) * SM - (i - 1) * (i - 2) / 2 + (j - (i - 1)); }
  static constexpr int size = SM == 1 ? 1 : SM + (SM - 1) * SM - (SM - 1) * (SM - 2) / 2;
};

/* per (pair, configuration of the batch): the Hessenberg form and the
 * carried vectors, [cfg][E][NP] with E = HessIdx::size + 4 SM + 1:
 * H's entries, u~ = k Q'u, z~ = Q'eVb, t~ = Q't, v~ = Q'T^2 eVb, eVb't */
template <int SM>
__global__ void __launch_bounds__(64) mvlr_hess_kernel(const eqbf_chunk c, const MvGrid mg, double* __restrict__ hess) {
  const int64_t NP = c.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)64 + threadIdx.x;
  if (p >= NP) return;
  const int cfg = blockIdx.y;
  const int PSZ = mg.PSZ, S = SM;
  constexpr int HS = HessIdx<SM>::size, E = HS + 4 * SM + 1;
  const double* pr = mg.par + (int64_t)cfg * PSZ * NP + p;
  double* out = hess + (int64_t)cfg * E * NP + p;
  if (pr[(int64_t)(PSZ - 1) * NP] != 0.0) return; /* no result: the row kernel writes NaN */
  const double kk = pr[(int64_t)(S * S + 3 * S) * NP];
  double t[SM], ut[SM], zt[SM], tt[SM], vt[SM], B[SM][SM];
  double et = 0.0;
#pragma unroll
  for (int i = 0; i < SM; ++i) t[i] = pr[(int64_t)(S * S + S + i) * NP];
#pragma unroll
  for (int i = 0; i < SM; ++i) {
    const double u = pr[(int64_t)(S * S + i) * NP], z = pr[(int64_t)(S * S + 2 * S + i) * NP];
#pragma unroll
    for (int j = 0; j < SM; ++j) B[i][j] = kk * pr[(int64_t)(i * S + j) * NP] * t[j] * t[j];
    ut[i] = kk * u;
    zt[i] = z;
    tt[i] = t[i];
    vt[i] = t[i] * t[i] * z;
    et += z * t[i];
  }
#pragma unroll
  for (int k = 0; k + 2 < SM; ++k) {
    double nrm2 = 0.0;
#pragma unroll
    for (int i = k + 1; i < SM; ++i) nrm2 += B[i][k] * B[i][k];
    const double x0 = B[k + 1][k];
    double alpha = sqrt(nrm2);
    if (x0 > 0) alpha = -alpha;
    double v[SM];
    double vv = 0.0;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      v[i] = i == k + 1 ? x0 - alpha : (i > k + 1 ? B[i][k] : 0.0);
      vv += v[i] * v[i];
    }
    const double beta = vv > 0.0 ? 2.0 / vv : 0.0;
#pragma unroll
    for (int j = k; j < SM; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = k + 1; i < SM; ++i) s += v[i] * B[i][j];
      s *= beta;
#pragma unroll
      for (int i = k + 1; i < SM; ++i) B[i][j] -= v[i] * s;
    }
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = k + 1; j < SM; ++j) s += B[i][j] * v[j];
      s *= beta;
#pragma unroll
      for (int j = k + 1; j < SM; ++j) B[i][j] -= s * v[j];
    }
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
    for (int i = k + 1; i < SM; ++i) {
      s1 += v[i] * ut[i];
      s2 += v[i] * zt[i];
      s3 += v[i] * tt[i];
      s4 += v[i] * vt[i];
    }
#pragma unroll
    for (int i = k + 1; i < SM; ++i) {
      ut[i] -= beta * v[i] * s1;
      zt[i] -= beta * v[i] * s2;
      tt[i] -= beta * v[i] * s3;
      vt[i] -= beta * v[i] * s4;
    }
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = (i == 0 ? 0 : i - 1); j < SM; ++j) out[(int64_t)HessIdx<SM>::at(i, j) * NP] = B[i][j];
#pragma unroll
  for (int i = 0; i < SM; ++i) {
    out[(int64_t)(HS + i) * NP] = ut[i];
    out[(int64_t)(HS + SM + i) * NP] = zt[i];
    out[(int64_t)(HS + 2 * SM + i) * NP] = tt[i];
    out[(int64_t)(HS + 3 * SM + i) * NP] = vt[i];
  }
  out[(int64_t)(HS + 4 * SM) * NP] = et;
}

/* per (pair, group of rows): every grid point of the group from the
 * Hessenberg form (H staged in LDS, the carried vectors in registers) */
template <int SM>
__global__ void __launch_bounds__(64) mvlr_rows_kernel(const eqbf_chunk c, const MvGrid mg,
                                                       const double* __restrict__ hess) {
  const int64_t NP = c.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)64 + threadIdx.x;
  const int lane = threadIdx.x;
  const int g = blockIdx.y;
  /* group g -> configuration (relative to the batch) and row type */
  int cfg, type;
  if (mg.cfg0 == 0) {
    cfg = g < 3 ? 0 : g - 2;
    type = g < 3 ? g : 3;
  } else {
    cfg = g;
    type = 3;
  }
  constexpr int HS = HessIdx<SM>::size, E = HS + 4 * SM + 1;
  __shared__ double Hs[HS][64];
  if (p >= NP) return; /* no barrier below: each lane uses its own LDS column */
  const int PSZ = mg.PSZ, S = SM;
  const double* pr = mg.par + (int64_t)cfg * PSZ * NP + p;
  const bool ok = pr[(int64_t)(PSZ - 1) * NP] == 0.0;
  const int n_rows = type == 3 ? mg.GS : mg.GL;
  const int64_t row_base = type < 3 ? (int64_t)type * mg.GL : 3 * (int64_t)mg.GL + (int64_t)(mg.cfg0 + cfg - 1) * mg.GS;
  const double* gp = type == 3 ? mg.grid + 2 * mg.GL : mg.grid;
  const double* go = type == 3 ? mg.grid + 2 * mg.GL + mg.GS : mg.grid + mg.GL;
  const double ln10 = 2.302585092994045684;
  double* raw = mg.raw + row_base * NP + p;
  if (!ok) {
    for (int gi = 0; gi < n_rows; ++gi) raw[(int64_t)gi * NP] = NAN;
    return;
  }
  if (type == 1) { /* gen-fix: M = I + o2 k u t' */
    const double kk = pr[(int64_t)(S * S + 3 * S) * NP];
    double tu = 0.0, tz = 0.0, et = 0.0;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      const double t = pr[(int64_t)(S * S + S + i) * NP];
      const double u = pr[(int64_t)(S * S + i) * NP], z = pr[(int64_t)(S * S + 2 * S + i) * NP];
      tu += t * u;
      tz += t * z;
      et += z * t;
    }
    for (int gi = 0; gi < n_rows; ++gi) {
      const double o2 = gp[gi] + go[gi];
      const double sig = 1.0 + o2 * kk * tu;
      const double quad = o2 * et * (tz / sig);
      raw[(int64_t)gi * NP] = (0.5 * quad - 0.5 * log(fabs(sig))) / ln10;
    }
    return;
  }
  const double* hv = hess + (int64_t)cfg * E * NP + p;
  for (int e = 0; e < HS; ++e) Hs[e][lane] = hv[(int64_t)e * NP];
  double ut[SM], zt[SM], tt[SM], vt[SM];
#pragma unroll
  for (int i = 0; i < SM; ++i) {
    ut[i] = hv[(int64_t)(HS + i) * NP];
    zt[i] = hv[(int64_t)(HS + SM + i) * NP];
    tt[i] = hv[(int64_t)(HS + 2 * SM + i) * NP];
    vt[i] = hv[(int64_t)(HS + 3 * SM + i) * NP];
  }
  const double et = hv[(int64_t)(HS + 4 * SM) * NP];
  for (int gi = 0; gi < n_rows; ++gi) {
    double p2, o2;
    if (type == 2) { /* gen-maxh */
      p2 = gp[gi] + go[gi];
      o2 = 0.0;
    } else {
      p2 = gp[gi];
      o2 = go[gi];
    }
    double W[SM][SM], r1[SM], r2[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) {
#pragma unroll
      for (int j = (i == 0 ? 0 : i - 1); j < SM; ++j) W[i][j] = (i == j ? 1.0 : 0.0) + p2 * Hs[HessIdx<SM>::at(i, j)][lane];
      r1[i] = ut[i];
      r2[i] = zt[i];
    }
    /* LU of the Hessenberg matrix, partial pivoting between rows j and j+1 */
    double mant = 1.0;
    int ex = 0;
#pragma unroll
    for (int j = 0; j + 1 < SM; ++j) {
      const bool sw = fabs(W[j + 1][j]) > fabs(W[j][j]);
#pragma unroll
      for (int col = j; col < SM; ++col) {
        const double a = W[j][col], b = W[j + 1][col];
        W[j][col] = sw ? b : a;
        W[j + 1][col] = sw ? a : b;
      }
      {
        const double a = r1[j], b = r1[j + 1];
        r1[j] = sw ? b : a;
        r1[j + 1] = sw ? a : b;
        const double c2 = r2[j], d2 = r2[j + 1];
        r2[j] = sw ? d2 : c2;
        r2[j + 1] = sw ? c2 : d2;
      }
      const double piv = W[j][j];
      const double f = W[j + 1][j] / piv;
#pragma unroll
      for (int col = j + 1; col < SM; ++col) W[j + 1][col] -= f * W[j][col];
      r1[j + 1] -= f * r1[j];
      r2[j + 1] -= f * r2[j];
      int e;
      mant = frexp(mant * piv, &e);
      ex += e;
    }
    {
      int e;
      mant = frexp(mant * W[SM - 1][SM - 1], &e);
      ex += e;
    }
    double lnd = log(fabs(mant)) + (double)ex * 0.69314718055994530942;
#pragma unroll
    for (int i = SM - 1; i >= 0; --i) {
      double a = r1[i], b = r2[i];
#pragma unroll
      for (int col = i + 1; col < SM; ++col) {
        a -= W[i][col] * r1[col];
        b -= W[i][col] * r2[col];
      }
      r1[i] = a / W[i][i];
      r2[i] = b / W[i][i];
    }
    if (o2 != 0.0) {
      double tau = 0.0, tz = 0.0;
#pragma unroll
      for (int i = 0; i < SM; ++i) {
        tau += tt[i] * r1[i];
        tz += tt[i] * r2[i];
      }
      const double sig = 1.0 + o2 * tau;
      const double coef = o2 * tz / sig;
#pragma unroll
      for (int i = 0; i < SM; ++i) r2[i] -= coef * r1[i];
      lnd += log(fabs(sig));
    }
    double txd = 0.0, vxd = 0.0;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
      txd += tt[i] * r2[i];
      vxd += vt[i] * r2[i];
    }
    const double quad = o2 * et * txd + p2 * vxd;
    raw[(int64_t)gi * NP] = (0.5 * quad - 0.5 * lnd) / ln10;
  }
}

/* ---- per (pair, weighted row): utils::log10_weighted_sum over the grid ------------ */
__global__ void __launch_bounds__(256) mvlr_wavg_kernel(const double* __restrict__ raw, double* __restrict__ w,
                                                         int64_t NP, int GL, int GS, int64_t wrow0, int64_t nw) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nw * NP) return;
  const int64_t wr = wrow0 + i / NP, p = i % NP;
  const int G = wr < 3 ? GL : GS;
  const int64_t r0 = wr < 3 ? wr * GL : 3 * (int64_t)GL + (wr - 3) * GS;
  const double* v = raw + r0 * NP + p;
  double mx = v[0];
  for (int k = 1; k < G; ++k) {
    const double x = v[(int64_t)k * NP];
    if (x > mx) mx = x;
  }
  const double wt = 1.0 / (double)G;
  double sum = 0;
  for (int k = 0; k < G; ++k) {
    const double x = v[(int64_t)k * NP];
    if (!isnan(x)) sum += wt * exp10(x - mx);
  }
  double res = mx + log10(sum);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  w[wr * NP + p] = res;
}

/* Raw and grid-averaged rows of every configuration (the consistent one
 * first), then the BMA rows.  fill(c0, nb, par, PSZ, pair_blocks) enqueues
 * the kernel writing the parameters of configurations [c0, c0 + nb) in the
 * layout mvlr_grid_kernel reads: [cfg - c0][PSZ][NP] with Sigma^-1 [S][S],
 * Sigma^-1 t, t, eVb, k and a flag (0 = computed).  Configurations go in
 * batches bounded by the parameter buffer and by the 65535 limit of the
 * grid's y dimension. */
/* the row kernel at the exact subgroup count (no padded rows in the LU) */
void launch_grid_kernel(int S, unsigned blocks, hipStream_t st, const eqbf_chunk& c, const MvGrid& mg) {
#define EQBF_GRID_CASE(n) \
  case n:                 \
    hipLaunchKernelGGL(mvlr_grid_kernel<n>, dim3(blocks), dim3(64), 0, st, c, mg); \
    break;
  switch (S) {
    EQBF_GRID_CASE(1)
    EQBF_GRID_CASE(2)
    EQBF_GRID_CASE(3)
    EQBF_GRID_CASE(4)
    EQBF_GRID_CASE(5)
    EQBF_GRID_CASE(6)
    EQBF_GRID_CASE(7)
    EQBF_GRID_CASE(8)
    EQBF_GRID_CASE(9)
    EQBF_GRID_CASE(10)
    EQBF_GRID_CASE(11)
    EQBF_GRID_CASE(12)
    EQBF_GRID_CASE(13)
    EQBF_GRID_CASE(14)
    EQBF_GRID_CASE(15)
    default:
      hipLaunchKernelGGL(mvlr_grid_kernel<16>, dim3(blocks), dim3(64), 0, st, c, mg);
      break;
  }
#undef EQBF_GRID_CASE
}

/* the Hessenberg kernels at the exact subgroup count, 2 <= S <= 10 */
template <int n>
void launch_hess_rows(unsigned pair_blocks, unsigned nb, unsigned groups, hipStream_t st, const eqbf_chunk& c,
                      const MvGrid& mg, double* hess, eqbf_ctx* ctx) {
  {
    ProfScope ps(ctx, "mvlr_hess_kernel", st);
    hipLaunchKernelGGL(mvlr_hess_kernel<n>, dim3(pair_blocks, nb), dim3(64), 0, st, c, mg, hess);
  }
  ProfScope ps(ctx, "mvlr_rows_kernel", st);
  hipLaunchKernelGGL(mvlr_rows_kernel<n>, dim3(pair_blocks, groups), dim3(64), 0, st, c, mg, (const double*)hess);
}

constexpr bool hess_ok(int S) { return S >= 2 && S <= 10; }
constexpr int hess_entries(int S) { return (S == 1 ? 1 : S + (S - 1) * S - (S - 1) * (S - 2) / 2) + 4 * S + 1; }

void launch_rows_kernel(int S, unsigned pair_blocks, unsigned nb, unsigned groups, hipStream_t st,
                        const eqbf_chunk& c, const MvGrid& mg, double* hess, eqbf_ctx* ctx) {
  switch (S) {
    case 2: launch_hess_rows<2>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 3: launch_hess_rows<3>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 4: launch_hess_rows<4>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 5: launch_hess_rows<5>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 6: launch_hess_rows<6>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 7: launch_hess_rows<7>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 8: launch_hess_rows<8>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    case 9: launch_hess_rows<9>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
    default: launch_hess_rows<10>(pair_blocks, nb, groups, st, c, mg, hess, ctx); break;
  }
}

template <int SM, class Fill>
int mv_rows_run(eqbf_ctx* ctx, const eqbf_chunk& c, int bfs, eqbf_assoc_out* out, hipStream_t st, Fill fill) {
  const int S = c.n_subgroups;
  const int64_t NP = c.n_pairs;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const int64_t n_cfg = eqbf_n_configs(bfs, S);
  const int PSZ = S * S + 3 * S + 2;
  const int64_t total = 1 + n_cfg;
  int64_t per = std::max<int64_t>(1, ((int64_t)1 << 30) / std::max<int64_t>(1, NP * PSZ * 8));
  /* 2 to 10 subgroups: one thread per (pair, group of rows), the grid
   * points of a group from one Hessenberg reduction (mvlr_rows_kernel);
   * otherwise one thread per (pair, row) */
  per = std::min<int64_t>(std::min<int64_t>(per, hess_ok(S) ? 65533 : 65535), total);
  double* par = (double*)ctx_buf(ctx, "mv_par", (size_t)per * NP * PSZ * sizeof(double));
  double* hess = hess_ok(S) ? (double*)ctx_buf(ctx, "mv_hess", (size_t)per * NP * hess_entries(S) * sizeof(double))
                            : nullptr;
  if (!par || (hess_ok(S) && !hess)) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating mvlr buffers failed");
  const int64_t pair_blocks = (NP + 63) / 64;
  for (int64_t c0 = 0; c0 < total; c0 += per) {
    const int nb = (int)std::min<int64_t>(per, total - c0);
    fill(c0, nb, par, PSZ, pair_blocks);
    /* raw rows of configurations [c0, c0 + nb) and their weighted rows */
    const int64_t c1 = c0 + nb;
    const int64_t row0 = c0 == 0 ? 0 : 3 * (int64_t)GL + (c0 - 1) * GS;
    const int64_t row1 = 3 * (int64_t)GL + (c1 - 1) * GS;
    const int64_t wrow0 = c0 == 0 ? 0 : 3 + (c0 - 1);
    const int64_t wrow1 = 3 + (c1 - 1);
    if (hess_ok(S)) {
      MvGrid mg;
      mg.par = par;
      mg.grid = (const double*)ctx->d_grid.ptr;
      mg.raw = (double*)out->raw_abf;
      mg.GL = GL;
      mg.GS = GS;
      mg.cfg0 = (int)c0;
      mg.PSZ = PSZ;
      mg.row0 = row0;
      mg.nrows = row1 - row0;
      mg.pair_blocks = pair_blocks;
      const unsigned groups = (unsigned)(nb + (c0 == 0 ? 2 : 0));
      launch_rows_kernel(S, (unsigned)pair_blocks, (unsigned)nb, groups, st, c, mg, hess, ctx);
    }
    for (int64_t r0 = row0; !hess_ok(S) && r0 < row1;) {
      const int64_t nr = std::min<int64_t>(row1 - r0, std::max<int64_t>(1, ((int64_t)1 << 30) / pair_blocks));
      MvGrid mg;
      mg.par = par;
      mg.grid = (const double*)ctx->d_grid.ptr;
      mg.raw = (double*)out->raw_abf;
      mg.GL = GL;
      mg.GS = GS;
      mg.cfg0 = (int)c0;
      mg.PSZ = PSZ;
      mg.row0 = r0;
      mg.nrows = nr;
      mg.pair_blocks = pair_blocks;
      {
        ProfScope ps(ctx, "mvlr_grid_kernel", st);
        launch_grid_kernel(S, (unsigned)(nr * pair_blocks), st, c, mg);
      }
      r0 += nr;
    }
    const int64_t nw = wrow1 - wrow0;
    hipLaunchKernelGGL(mvlr_wavg_kernel, dim3((unsigned)((nw * NP + 255) / 256)), dim3(256), 0, st,
                       (const double*)out->raw_abf, out->w_abf, NP, GL, GS, wrow0, nw);
    EQBF_HIP(ctx, hipGetLastError());
  }
  return launch_bma(ctx, out->w_abf, NP, S, bfs, st);
}

template <int SM>
int mvlr_run_sm(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_mvlr_opts* o, eqbf_assoc_out* out, hipStream_t st) {
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  const int64_t NP = c.n_pairs;
  const int64_t n_cfg = eqbf_n_configs(o->bfs, S);
  const int KU = 1 + c.n_covars_max;
  MvPrep pp;
  pp.mrec = (MvRec*)ctx_buf(ctx, "mv_rec", (size_t)G * sizeof(MvRec));
  pp.midx = (int32_t*)ctx_buf(ctx, "mv_idx", (size_t)G * N * sizeof(int32_t));
  pp.Ut = (double*)ctx_buf(ctx, "mv_U", (size_t)G * KU * N * sizeof(double));
  pp.Yt = (double*)ctx_buf(ctx, "mv_Y", (size_t)G * S * N * sizeof(double));
  pp.YtY = (double*)ctx_buf(ctx, "mv_YtY", (size_t)G * S * S * sizeof(double));
  pp.S0 = (double*)ctx_buf(ctx, "mv_S0", (size_t)G * S * S * sizeof(double));
  pp.S0inv = (double*)ctx_buf(ctx, "mv_S0i", (size_t)G * S * S * sizeof(double));
  pp.KU = KU;
  pp.flag = (int*)ctx->d_flag.ptr;
  double* kb = (double*)ctx_buf(ctx, "mv_kb", (size_t)(S + 1) * NP * sizeof(double));
  int32_t* pgene = (int32_t*)ctx_buf(ctx, "mv_pgene", (size_t)NP * sizeof(int32_t));
  /* configuration masks: the consistent one (all subgroups) first */
  std::vector<uint64_t> hm(1 + n_cfg);
  hm[0] = (S >= 64) ? ~0ull : ((1ull << S) - 1ull);
  if (n_cfg > 0) eqbf_config_masks(o->bfs, S, hm.data() + 1);
  uint64_t* dmask = (uint64_t*)ctx_buf(ctx, "mv_masks", hm.size() * sizeof(uint64_t));
  if (!pp.mrec || !pp.midx || !pp.Ut || !pp.Yt || !pp.YtY || !pp.S0 || !pp.S0inv || !kb || !pgene || !dmask)
    return ctx_fail(ctx, EQBF_ERR_HIP, "allocating mvlr buffers failed");
  EQBF_HIP(ctx, hipMemcpyAsync(dmask, hm.data(), hm.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st)); /* the host vector goes out of scope */
  int rc = ctx_ensure_tz(ctx, N, st);
  if (rc) return rc;
  hipLaunchKernelGGL(mv_zero_kernel, dim3(1), dim3(1), 0, st, (int*)ctx->d_flag.ptr);
  {
    ProfScope ps(ctx, "mvlr_prep_kernel", st);
    with_exact_sm<SM>(c.n_subgroups, [&](auto e) {
      hipLaunchKernelGGL(mvlr_prep_kernel<decltype(e)::value>, dim3(G), dim3(64), 0, st, c, pp);
    });
  }
  MvSums ms;
  ms.mrec = pp.mrec;
  ms.midx = pp.midx;
  ms.Ut = pp.Ut;
  ms.Yt = pp.Yt;
  ms.kb = kb;
  ms.pgene = pgene;
  ms.n_out = out->n;
  ms.sstats = out->sstats;
  ms.stdsstats = out->stdsstats;
  ms.flag = (int*)ctx->d_flag.ptr;
  ms.KU = KU;
  {
    ProfScope ps(ctx, "mvlr_sums_kernel", st);
    hipLaunchKernelGGL(mvlr_sums_kernel<SM>, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, st, c, ms);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return mv_rows_run<SM>(ctx, c, o->bfs, out, st, [&](int64_t c0, int nb, double* par, int PSZ, int64_t pair_blocks) {
    MvCfg mc;
    mc.mrec = pp.mrec;
    mc.YtY = pp.YtY;
    mc.S0 = pp.S0;
    mc.S0inv = pp.S0inv;
    mc.kb = kb;
    mc.pgene = pgene;
    mc.tz = ctx_tz_table(ctx);
    mc.masks = dmask;
    mc.par = par;
    mc.fiterr = o->fiterr;
    mc.cfg0 = (int)c0;
    mc.PSZ = PSZ;
    ProfScope ps(ctx, "mvlr_cfg_kernel", st);
    with_exact_sm<SM>(c.n_subgroups, [&](auto e) {
      hipLaunchKernelGGL(mvlr_cfg_kernel<decltype(e)::value>, dim3((unsigned)pair_blocks, (unsigned)nb), dim3(64), 0,
                         st, c, mc);
    });
  });
}

/* ==== --error hybrid ================================================================
 * GeneSnpPair::CalcAbfsHybrid (gene_snp_pair.cpp:1384-1423) and what it
 * calls (:760-1163), with the Bayes factor of CalcLog10AbfMvlr
 * (:1165-1256).  Per pair of a gene expressed in every subgroup and a SNP
 * genotyped in every subgroup:
 *   b_s, Sigma_ss, Vg_ss from the subgroup's own regression over its members
 *     (present: genotyped, expression level not NaN) on [1, x, covariates],
 *     sigma2 = rss / n, Sigma_ss = f s2_full + (1 - f) s2_null and
 *     Vg_ss = Sigma_ss [(X'X)^-1]_xx;
 *   for s1 < s2, over the individuals present in both (rows [1, x, covariates
 *     of s1 read at the individual's index]): Sigma_12 = f full_01 + (1 - f)
 *     null_01 from the ML error covariances Y'(I - H)Y / n12, and
 *     Vg_12 = Sigma_12 [M1^-1 G12 M2^-1]_xx with G12 = X12'X12 and M_s the
 *     same sums over every individual present in s (the reference's
 *     pinv(G12 + Xs'Xs) X12' products, :977-1131);
 *   log10 ABF = (b'Vg^-1 W (I + Vg^-1 W)^-1 Vg^-1 b - ln|det(I + Vg^-1 W)|)
 *     / (2 ln 10), W = (oma2 1 1' + phi2 I) o gamma gamma' o sd sd',
 *     sd_s = sqrt(Sigma_ss): the MVLR row kernel with k = 1 and Vg in place
 *     of Sigma.
 * Every sum over individuals splits into a covariate block that does not
 * depend on the SNP (per gene: A = sum c c', its inverse, the null
 * regression coefficients psi = A^-1 sum c y and the null residual
 * cross-product, hyb_gene_kernel) and SNP terms (u = sum c x, sum x^2,
 * sum x y, hyb_pair_kernel).  With w = A^-1 u and k = sum x^2 - u'w (the
 * Schur complement, x~'x~ of Frisch-Waugh), the genotype column of M^-1 is
 * (-w, 1) / k, so
 *   [(X'X)^-1]_xx = 1 / k,  x~'y~ = sum x y - u'psi,
 *   rss_full = rss_null - (x~'y~)^2 / k,
 *   [M1^-1 G12 M2^-1]_xx = (w1'A12 w2 - w1'u12 - u12'w2 + sum_12 x^2) / (k1 k2)
 * and no (Q+2) x (Q+2) system is solved per pair.  The reference's
 * pseudo-inverses invert every singular value, so on full-rank designs they
 * are these inverses; the covariate blocks are inverted by Gauss-Jordan
 * elimination with partial pivoting. */
struct HybGene {
  const double* cp;  /* pairwise covariates [S][Qmax][N] (by index in the samples) */
  const double* yp;  /* pairwise expression levels [G][S][N]                       */
  double* rec;       /* [G][NSET][RSZ]                                             */
  int NSET, RSZ, Q1M;
};

/* set j of a gene: 0 <= j < S: members of s with its own covariates and
 * levels; S + s1: members of s1 with the pairwise covariates of s1;
 * 2S + pi: present in s1 and s2 (pair pi = (s1, s2), s1 < s2 in
 * lexicographic order) with the pairwise covariates and levels;
 * 2S + npr + pi: present in s2 with the pairwise covariates of s1. */
__device__ inline void hyb_pair_of(int pi, int S, int& s1, int& s2) {
  s1 = 0;
  while (pi >= S - 1 - s1) {
    pi -= S - 1 - s1;
    ++s1;
  }
  s2 = s1 + 1 + pi;
}

/* present in subgroup s: genotyped there and a level (not NaN) in lev
 * [G][S][N]; the per-subgroup sets read the chunk's levels, the pairwise
 * sets the pairwise levels, which permutations leave as loaded
 * (samples.cpp:148-192) */
__device__ inline bool hyb_present(const eqbf_chunk& c, const double* lev, int g, int s, int i) {
  const int S = c.n_subgroups, N = c.n_samples;
  return c.geno_in_sbgrp[(int64_t)s * N + i] && !isnan(lev[((int64_t)g * S + s) * N + i]);
}

/* record of a set: A^-1 [Q1M][Q1M], A [Q1M][Q1M], psi1 [Q1M], psi2 [Q1M],
 * r (null residual cross-product y1'(I - Hc)y2), n */
__global__ void __launch_bounds__(256) hyb_gene_kernel(const eqbf_chunk c, const HybGene hg) {
  const int g = blockIdx.x, j = blockIdx.y, t = threadIdx.x;
  const int S = c.n_subgroups, N = c.n_samples, QM = c.n_covars_max;
  for (int s = 0; s < S; ++s)
    if (!c.gene_in_sbgrp[(int64_t)g * S + s]) return;
  const int npr = S * (S - 1) / 2;
  int type, sa, sb = -1;
  if (j < S) {
    type = 0;
    sa = j;
  } else if (j < 2 * S) {
    type = 1;
    sa = j - S;
  } else if (j < 2 * S + npr) {
    type = 2;
    hyb_pair_of(j - 2 * S, S, sa, sb);
  } else {
    type = 3;
    hyb_pair_of(j - 2 * S - npr, S, sa, sb);
  }
  const int Q1 = min(max(c.covar_count[sa], 0), QM) + 1, Q1M = hg.Q1M;
  const double* cov = (type == 0) ? c.covars + (int64_t)sa * QM * N : hg.cp + (int64_t)sa * QM * N;
  const bool has_y = type == 0 || type == 2;
  const double* y1 = (type == 0) ? c.explevels + ((int64_t)g * S + sa) * N : hg.yp + ((int64_t)g * S + sa) * N;
  const double* y2 = (type == 0) ? y1 : (type == 2 ? hg.yp + ((int64_t)g * S + sb) * N : nullptr);
  /* entries: the upper triangle of A, then sum c y1, sum c y2, sum y1 y2, n */
  const int nA = Q1 * (Q1 + 1) / 2;
  const int nE = nA + (has_y ? 2 * Q1 + 1 : 0) + 1;
  __shared__ double ent[17 * 18 / 2 + 2 * 17 + 2];
  __shared__ double M[17][34];
  __shared__ double fac[17];
  __shared__ int pivrow;
  /* one wave per entry, lanes striding the individuals */
  const int lane = t & 63, nwv = blockDim.x >> 6;
  const double* lev = (type == 0) ? c.explevels : hg.yp;
  for (int e = t >> 6; e < nE; e += nwv) {
    int kind, a = 0, b = 0;
    if (e < nA) {
      kind = 0;
      int r = e;
      while (r >= Q1 - a) {
        r -= Q1 - a;
        ++a;
      }
      b = a + r;
    } else if (has_y && e < nA + Q1) {
      kind = 1;
      a = e - nA;
    } else if (has_y && e < nA + 2 * Q1) {
      kind = 2;
      a = e - nA - Q1;
    } else if (has_y && e == nA + 2 * Q1) {
      kind = 3;
    } else {
      kind = 4;
    }
    double acc = 0.0;
    for (int i = lane; i < N; i += 64) {
      bool in;
      if (type == 0 || type == 1)
        in = hyb_present(c, lev, g, sa, i);
      else if (type == 2)
        in = hyb_present(c, lev, g, sa, i) && hyb_present(c, lev, g, sb, i);
      else
        in = hyb_present(c, lev, g, sb, i);
      if (!in) continue;
      const double va = a == 0 ? 1.0 : cov[(int64_t)(a - 1) * N + i];
      switch (kind) {
        case 0: acc += va * (b == 0 ? 1.0 : cov[(int64_t)(b - 1) * N + i]); break;
        case 1: acc += va * y1[i]; break;
        case 2: acc += va * y2[i]; break;
        case 3: acc += y1[i] * y2[i]; break;
        default: acc += 1.0; break;
      }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) ent[e] = acc;
  }
  __syncthreads();
  /* [A | I] */
  for (int e = t; e < Q1 * 2 * Q1; e += blockDim.x) {
    const int r = e / (2 * Q1), col = e - r * 2 * Q1;
    double v;
    if (col < Q1) {
      const int a = min(r, col), b = max(r, col);
      v = ent[a * Q1 - a * (a - 1) / 2 + (b - a)];
    } else {
      v = (col - Q1 == r) ? 1.0 : 0.0;
    }
    M[r][col] = v;
  }
  __syncthreads();
  for (int jc = 0; jc < Q1; ++jc) {
    if (t == 0) {
      int p = jc;
      double best = fabs(M[jc][jc]);
      for (int r = jc + 1; r < Q1; ++r)
        if (fabs(M[r][jc]) > best) {
          best = fabs(M[r][jc]);
          p = r;
        }
      pivrow = p;
    }
    __syncthreads();
    const int p = pivrow;
    if (p != jc && t < 2 * Q1) {
      const double tmp = M[jc][t];
      M[jc][t] = M[p][t];
      M[p][t] = tmp;
    }
    __syncthreads();
    if (t < Q1) fac[t] = (t == jc) ? M[jc][jc] : M[t][jc];
    __syncthreads();
    if (t < 2 * Q1) M[jc][t] /= fac[jc];
    __syncthreads();
    for (int e = t; e < Q1 * 2 * Q1; e += blockDim.x) {
      const int r = e / (2 * Q1), col = e - r * 2 * Q1;
      if (r != jc) M[r][col] -= fac[r] * M[jc][col];
    }
    __syncthreads();
  }
  const int RSZ = hg.RSZ;
  double* out = hg.rec + ((int64_t)g * hg.NSET + j) * RSZ;
  for (int e = t; e < Q1M * Q1M; e += blockDim.x) {
    const int r = e / Q1M, col = e - r * Q1M;
    const bool in = r < Q1 && col < Q1;
    out[e] = in ? M[r][Q1 + col] : 0.0;
    if (in) {
      const int a = min(r, col), b = max(r, col);
      out[Q1M * Q1M + e] = ent[a * Q1 - a * (a - 1) / 2 + (b - a)];
    } else {
      out[Q1M * Q1M + e] = 0.0;
    }
  }
  if (t < Q1M) {
    double p1 = 0.0, p2 = 0.0;
    if (has_y && t < Q1)
      for (int k = 0; k < Q1; ++k) {
        p1 += M[t][Q1 + k] * ent[nA + k];
        p2 += M[t][Q1 + k] * ent[nA + Q1 + k];
      }
    out[2 * Q1M * Q1M + t] = p1;
    out[2 * Q1M * Q1M + Q1M + t] = p2;
  }
  if (t == 0) {
    double r = 0.0;
    if (has_y) {
      r = ent[nA + 2 * Q1];
      for (int k = 0; k < Q1; ++k) {
        double p2 = 0.0;
        for (int l = 0; l < Q1; ++l) p2 += M[k][Q1 + l] * ent[nA + Q1 + l];
        r -= ent[nA + k] * p2;
      }
    }
    out[2 * Q1M * Q1M + 2 * Q1M] = r;
    out[2 * Q1M * Q1M + 2 * Q1M + 1] = ent[nE - 1];
  }
}

struct HybTile {
  int64_t p0;
  int32_t g, n;
};

struct HybPair {
  const HybTile* tiles;
  const uint64_t* pres;  /* [G][S][W] presence bits from the chunk's levels  */
  const uint64_t* presp; /* the same from the pairwise levels (may be pres) */
  const double* cp;
  const double* yp;
  const double* rec;
  double* b;      /* [S][NP]    */
  double* Sig;    /* [S*S][NP]  */
  double* Vg;     /* [S*S][NP]  */
  int32_t* pgene; /* [NP]       */
  int32_t* n_out;
  double* sstats;
  int* flag;
  double fiterr;
  int NSET, RSZ;
};

template <int Q1M>
__device__ inline double hyb_quad(const double* __restrict__ Ai, const double (&u)[Q1M], const double (&v)[Q1M],
                                  int Q1) {
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < Q1M; ++a) {
    if (a >= Q1) break;
    double r = 0.0;
#pragma unroll
    for (int b = 0; b < Q1M; ++b)
      if (b < Q1) r += Ai[a * Q1M + b] * v[b];
    s += u[a] * r;
  }
  return s;
}

/* presence bits: bit i of word w of (gene, subgroup) = individual 64 w + i
 * is genotyped in the subgroup and has a level (not NaN) in lev [G][S][N] */
__global__ void __launch_bounds__(64) hyb_pres_kernel(const eqbf_chunk c, const double* __restrict__ lev,
                                                      uint64_t* __restrict__ pres, int W) {
  const int g = blockIdx.x, s = blockIdx.y, lane = threadIdx.x;
  const int S = c.n_subgroups, N = c.n_samples;
  for (int w = 0; w < W; ++w) {
    const int i = w * 64 + lane;
    const bool in = i < N && c.geno_in_sbgrp[(int64_t)s * N + i] && !isnan(lev[((int64_t)g * S + s) * N + i]);
    const uint64_t m = __ballot(in);
    if (lane == 0) pres[((int64_t)g * S + s) * W + w] = m;
  }
}

/* one lane per pair; the 64 pairs of a block belong to one gene.  The
 * individuals go by chunks of 64: the block stages the chunk's covariates
 * and levels in LDS (every lane reads them by broadcast), the presence bits
 * come as one uniform word a chunk, and each lane loads its genotypes eight
 * individuals at a time so that the loads overlap.  SPLIT: the pairwise
 * sets have presence bits of their own (hp.presp != hp.pres), so each
 * subgroup's first pass keeps its per-subgroup and pairwise sums apart. */
template <int Q1M, class GT, bool SPLIT>
__global__ void __launch_bounds__(64) hyb_pair_kernel(const eqbf_chunk c, const HybPair hp) {
  __shared__ double st[(2 * (Q1M - 1) + 2) * 64];
  const HybTile tl = hp.tiles[blockIdx.x];
  const int g = tl.g, lane = threadIdx.x;
  const bool act = lane < tl.n;
  const int64_t p = tl.p0 + (act ? lane : 0);
  const int64_t NP = c.n_pairs;
  const int S = c.n_subgroups, N = c.n_samples, QM = c.n_covars_max;
  const int W = (N + 63) >> 6;
  const int64_t snp = c.pair_snp[p];
  bool gene_ok = true, ok = act;
  for (int s = 0; s < S; ++s) {
    gene_ok &= c.gene_in_sbgrp[(int64_t)g * S + s] != 0;
    ok &= c.snp_in_sbgrp[snp * S + s] != 0;
  }
  ok &= gene_ok;
  if (gene_ok) {
    const uint64_t* pres = hp.pres + (int64_t)g * S * W;
    const uint64_t* presp = hp.presp + (int64_t)g * S * W;
    const int npr = S * (S - 1) / 2;
    const int RSZ = hp.RSZ;
    const double* R = hp.rec + (int64_t)g * hp.NSET * RSZ;
    const double f = hp.fiterr;
    int pi = 0;
    for (int s1 = 0; s1 < S; ++s1) {
      const int Q1 = min(max(c.covar_count[s1], 0), QM) + 1, Q = Q1 - 1;
      const double* cov0 = c.covars + (int64_t)s1 * QM * N;
      const double* covp = hp.cp + (int64_t)s1 * QM * N;
      const double* y = c.explevels + ((int64_t)g * S + s1) * N;
      const GT* xb = (const GT*)c.genotypes + (int64_t)c.geno_set_of_sbgrp[s1] * c.geno_set_stride + snp;
      const int64_t ld = c.geno_ld;
      double u0[Q1M], u1[Q1M], sxx = 0.0, sxxp = 0.0, sxy = 0.0;
#pragma unroll
      for (int a = 0; a < Q1M; ++a) u0[a] = u1[a] = 0.0;
      for (int c0 = 0; c0 < N; c0 += 64) {
        const int n = min(64, N - c0);
        __syncthreads();
        if (lane < n) {
          for (int q = 0; q < Q; ++q) {
            st[q * 64 + lane] = cov0[(int64_t)q * N + c0 + lane];
            st[(Q + q) * 64 + lane] = covp[(int64_t)q * N + c0 + lane];
          }
          st[2 * Q * 64 + lane] = y[c0 + lane];
        }
        __syncthreads();
        const uint64_t m = pres[(int64_t)s1 * W + (c0 >> 6)];
        const uint64_t mp = SPLIT ? presp[(int64_t)s1 * W + (c0 >> 6)] : m;
        for (int j0 = 0; j0 < n; j0 += 8) {
          double xv[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) xv[k] = (ok && j0 + k < n) ? (double)xb[(int64_t)(c0 + j0 + k) * ld] : 0.0;
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int j = j0 + k;
            const bool in0 = (m >> j) & 1ull, inp = (mp >> j) & 1ull;
            if (j >= n || !(in0 || inp)) continue;
            const double x = xv[k];
            if (!SPLIT) { /* one set: the per-subgroup and pairwise sums together */
              sxx += x * x;
              sxy += x * st[2 * Q * 64 + j];
              u0[0] += x;
              u1[0] += x;
#pragma unroll
              for (int a = 1; a < Q1M; ++a)
                if (a < Q1) {
                  u0[a] += st[(a - 1) * 64 + j] * x;
                  u1[a] += st[(Q + a - 1) * 64 + j] * x;
                }
              continue;
            }
            if (in0) {
              sxx += x * x;
              sxy += x * st[2 * Q * 64 + j];
              u0[0] += x;
#pragma unroll
              for (int a = 1; a < Q1M; ++a)
                if (a < Q1) u0[a] += st[(a - 1) * 64 + j] * x;
            }
            if (inp) {
              sxxp += x * x;
              u1[0] += x;
#pragma unroll
              for (int a = 1; a < Q1M; ++a)
                if (a < Q1) u1[a] += st[(Q + a - 1) * 64 + j] * x;
            }
          }
        }
      }
      const double* R0 = R + (int64_t)s1 * RSZ;
      const double* R1 = R + (int64_t)(S + s1) * RSZ;
      const double k0 = sxx - hyb_quad<Q1M>(R0, u0, u0, Q1);
      double xty = sxy;
#pragma unroll
      for (int a = 0; a < Q1M; ++a)
        if (a < Q1) xty -= u0[a] * R0[2 * Q1M * Q1M + a];
      const double rss0 = R0[2 * Q1M * Q1M + 2 * Q1M], n0 = R0[2 * Q1M * Q1M + 2 * Q1M + 1];
      const double rss1 = rss0 - xty * xty / k0;
      const double sg = f * rss1 / n0 + (1.0 - f) * rss0 / n0;
      if (ok) {
        hp.b[(int64_t)s1 * NP + p] = xty / k0;
        hp.Sig[(int64_t)(s1 * S + s1) * NP + p] = sg;
        hp.Vg[(int64_t)(s1 * S + s1) * NP + p] = sg / k0;
      }
      /* w1 = A1^-1 u1, k1: the genotype column of M1^-1 is (-w1, 1) / k1 */
      double w1[Q1M];
#pragma unroll
      for (int a = 0; a < Q1M; ++a) {
        double r = 0.0;
#pragma unroll
        for (int b2 = 0; b2 < Q1M; ++b2)
          if (a < Q1 && b2 < Q1) r += R1[a * Q1M + b2] * u1[b2];
        w1[a] = r;
      }
      double k1 = SPLIT ? sxxp : sxx;
#pragma unroll
      for (int a = 0; a < Q1M; ++a)
        if (a < Q1) k1 -= u1[a] * w1[a];
      const double* yp1 = hp.yp + ((int64_t)g * S + s1) * N;
      /* PW second subgroups per pass over the individuals (the genotypes
       * are read once for them); a missing second one repeats the first
       * and is not written */
      constexpr int PW = Q1M <= 8 ? 2 : 1;
      for (int s2 = s1 + 1; s2 < S; s2 += PW) {
        const int nb = min(PW, S - s2);
        const double* yp2[PW];
        int sb[PW];
#pragma unroll
        for (int b = 0; b < PW; ++b) {
          sb[b] = s2 + min(b, nb - 1);
          yp2[b] = hp.yp + ((int64_t)g * S + sb[b]) * N;
        }
        double u12[PW][Q1M], u2[PW][Q1M], sxx12[PW], sxx2[PW], sy1[PW], sy2[PW];
#pragma unroll
        for (int b = 0; b < PW; ++b) {
          sxx12[b] = sxx2[b] = sy1[b] = sy2[b] = 0.0;
#pragma unroll
          for (int a = 0; a < Q1M; ++a) u12[b][a] = u2[b][a] = 0.0;
        }
        for (int c0 = 0; c0 < N; c0 += 64) {
          const int n = min(64, N - c0);
          __syncthreads();
          if (lane < n) {
            for (int q = 0; q < Q; ++q) st[q * 64 + lane] = covp[(int64_t)q * N + c0 + lane];
            st[Q * 64 + lane] = yp1[c0 + lane];
#pragma unroll
            for (int b = 0; b < PW; ++b) st[(Q + 1 + b) * 64 + lane] = yp2[b][c0 + lane];
          }
          __syncthreads();
          const uint64_t m1 = presp[(int64_t)s1 * W + (c0 >> 6)];
          uint64_t m2[PW], any = 0ull;
#pragma unroll
          for (int b = 0; b < PW; ++b) {
            m2[b] = presp[(int64_t)sb[b] * W + (c0 >> 6)];
            any |= m2[b];
          }
          for (int j0 = 0; j0 < n; j0 += 8) {
            double xv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) xv[k] = (ok && j0 + k < n) ? (double)xb[(int64_t)(c0 + j0 + k) * ld] : 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              const int j = j0 + k;
              if (j >= n || !((any >> j) & 1ull)) continue;
              const double x = xv[k];
              double cv[Q1M];
              cv[0] = 1.0;
#pragma unroll
              for (int a = 1; a < Q1M; ++a) cv[a] = (a < Q1) ? st[(a - 1) * 64 + j] : 0.0;
              const bool in1 = (m1 >> j) & 1ull;
#pragma unroll
              for (int b = 0; b < PW; ++b) {
                if (!((m2[b] >> j) & 1ull)) continue;
                sxx2[b] += x * x;
#pragma unroll
                for (int a = 0; a < Q1M; ++a) u2[b][a] += cv[a] * x;
                if (in1) {
                  sxx12[b] += x * x;
                  sy1[b] += x * st[Q * 64 + j];
                  sy2[b] += x * st[(Q + 1 + b) * 64 + j];
#pragma unroll
                  for (int a = 0; a < Q1M; ++a) u12[b][a] += cv[a] * x;
                }
              }
            }
          }
        }
#pragma unroll
        for (int b = 0; b < PW; ++b) {
          if (b >= nb) break;
          const int s2b = s2 + b, pib = pi + b;
          const double* R2 = R + (int64_t)(2 * S + pib) * RSZ;
          const double* R3 = R + (int64_t)(2 * S + npr + pib) * RSZ;
          const double n12 = R2[2 * Q1M * Q1M + 2 * Q1M + 1];
          if (ok && !(n12 > 0.0)) atomicOr(hp.flag, FLAG_HYB_NO_COMMON);
          const double k12 = sxx12[b] - hyb_quad<Q1M>(R2, u12[b], u12[b], Q1);
          double xty1 = sy1[b], xty2 = sy2[b];
#pragma unroll
          for (int a = 0; a < Q1M; ++a)
            if (a < Q1) {
              xty1 -= u12[b][a] * R2[2 * Q1M * Q1M + a];
              xty2 -= u12[b][a] * R2[2 * Q1M * Q1M + Q1M + a];
            }
          const double s12 = R2[2 * Q1M * Q1M + 2 * Q1M] / n12 - f * xty1 * xty2 / (k12 * n12);
          double w2[Q1M];
#pragma unroll
          for (int a = 0; a < Q1M; ++a) {
            double r = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < Q1M; ++b2)
              if (a < Q1 && b2 < Q1) r += R3[a * Q1M + b2] * u2[b][b2];
            w2[a] = r;
          }
          double k2 = sxx2[b], wu = 0.0;
#pragma unroll
          for (int a = 0; a < Q1M; ++a)
            if (a < Q1) {
              k2 -= u2[b][a] * w2[a];
              wu += w1[a] * u12[b][a] + u12[b][a] * w2[a];
            }
          const double aw = hyb_quad<Q1M>(R2 + Q1M * Q1M, w1, w2, Q1);
          const double a11 = (aw - wu + sxx12[b]) / (k1 * k2);
          if (ok) {
            hp.Sig[(int64_t)(s1 * S + s2b) * NP + p] = s12;
            hp.Sig[(int64_t)(s2b * S + s1) * NP + p] = s12;
            hp.Vg[(int64_t)(s1 * S + s2b) * NP + p] = s12 * a11;
            hp.Vg[(int64_t)(s2b * S + s1) * NP + p] = s12 * a11;
          }
        }
        pi += nb;
      }
    }
  }
  if (!act) return;
  if (ok) {
    hp.pgene[p] = g;
    return;
  }
  /* gene.cpp:314-321 skips the pair: no sample size, no statistics */
  for (int s = 0; s < S; ++s) {
    hp.n_out[(int64_t)s * NP + p] = 0;
    if (hp.sstats)
      for (int k = 0; k < 5; ++k) hp.sstats[((int64_t)k * S + s) * NP + p] = NAN;
  }
  hp.pgene[p] = -1;
}

/* per (pair, configuration): Vg^-1, Vg^-1 t, t = gamma o sqrt(diag Sigma),
 * Vg^-1 b and k = 1, in the layout of mvlr_grid_kernel */
struct HybCfg {
  const double* b;
  const double* Sig;
  const double* Vg;
  const int32_t* pgene;
  const uint64_t* masks;
  double* par;
  int cfg0, PSZ;
};

template <int SM>
__global__ void __launch_bounds__(64) hyb_cfg_kernel(const eqbf_chunk c, const HybCfg hc) {
  const int64_t NP = c.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const int ci = blockIdx.y;
  const int S = c.n_subgroups, PSZ = hc.PSZ;
  double* out = hc.par + (int64_t)ci * PSZ * NP + p;
  if (hc.pgene[p] < 0) {
    out[(int64_t)(PSZ - 1) * NP] = 1.0;
    return;
  }
  const uint64_t gm = hc.masks[hc.cfg0 + ci];
  double A[SM][SM], b[SM], tv[SM];
#pragma unroll
  for (int i = 0; i < SM; ++i) {
    b[i] = (i < S) ? hc.b[(int64_t)i * NP + p] : 0.0;
    tv[i] = (i < S && ((gm >> i) & 1ull)) ? sqrt(hc.Sig[(int64_t)(i * S + i) * NP + p]) : 0.0;
#pragma unroll
    for (int j = 0; j < SM; ++j) A[i][j] = (i < S && j < S) ? hc.Vg[(int64_t)(i * S + j) * NP + p] : (i == j ? 1.0 : 0.0);
  }
  int piv[SM];
  (void)lu_factor<SM>(A, piv);
  double u[SM], eVb[SM];
#pragma unroll
  for (int i = 0; i < SM; ++i) u[i] = eVb[i] = 0.0;
  for (int col = 0; col < S; ++col) {
    double e[SM];
#pragma unroll
    for (int i = 0; i < SM; ++i) e[i] = (i == col) ? 1.0 : 0.0;
    lu_apply<SM>(A, piv, e);
    double tc = 0, bc = 0;
#pragma unroll
    for (int j = 0; j < SM; ++j)
      if (j == col) {
        tc = tv[j];
        bc = b[j];
      }
#pragma unroll
    for (int i = 0; i < SM; ++i)
      if (i < S) {
        out[(int64_t)(i * S + col) * NP] = e[i];
        u[i] += e[i] * tc;
        eVb[i] += e[i] * bc;
      }
  }
#pragma unroll
  for (int i = 0; i < SM; ++i)
    if (i < S) {
      out[(int64_t)(S * S + i) * NP] = u[i];
      out[(int64_t)(S * S + S + i) * NP] = tv[i];
      out[(int64_t)(S * S + 2 * S + i) * NP] = eVb[i];
    }
  out[(int64_t)(S * S + 3 * S) * NP] = 1.0;
  out[(int64_t)(PSZ - 1) * NP] = 0.0;
}

template <int SM, int Q1M>
int hybrid_run_sm(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_hybrid_opts* o, eqbf_assoc_out* out,
                  hipStream_t st) {
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  const int64_t NP = c.n_pairs;
  const int64_t n_cfg = eqbf_n_configs(o->bfs, S);
  /* sample sizes and summary statistics: the separate analysis; the batched
   * path (FP64 sums when summary statistics are asked for) where it applies,
   * else the FP64 plain path */
  int KU = 1;
  while (KU < c.n_covars_max + 1) KU = (KU == 1) ? 4 : KU * 2;
  const int prec = uvlr_fast_eligible(c, KU) ? EQBF_PREC_MIXED : EQBF_PREC_FP64_BATCHED;
  eqbf_assoc_opts so{EQBF_ANALYSIS_SEP, EQBF_BFS_GEN, prec, EQBF_F64};
  eqbf_assoc_out sout;
  memset(&sout, 0, sizeof sout);
  sout.n = out->n;
  sout.sstats = out->sstats;
  int rc = assoc_uvlr_run(ctx, &c, &so, &sout, st);
  if (rc) return rc;
  const int npr = S * (S - 1) / 2;
  const int NSET = 2 * S + 2 * npr;
  const int RSZ = 2 * Q1M * Q1M + 2 * Q1M + 2;
  /* tiles of at most 64 pairs of one gene */
  std::vector<int64_t> off(G + 1);
  EQBF_HIP(ctx, hipMemcpyAsync(off.data(), c.pair_offset, (G + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  std::vector<HybTile> tiles;
  for (int g = 0; g < G; ++g)
    for (int64_t p0 = off[g]; p0 < off[g + 1]; p0 += 64)
      tiles.push_back(HybTile{p0, g, (int32_t)std::min<int64_t>(64, off[g + 1] - p0)});
  if (tiles.empty()) return EQBF_OK;
  HybTile* dtiles = (HybTile*)ctx_buf(ctx, "hyb_tiles", tiles.size() * sizeof(HybTile));
  double* rec = (double*)ctx_buf(ctx, "hyb_rec", (size_t)G * NSET * RSZ * sizeof(double));
  double* bb = (double*)ctx_buf(ctx, "hyb_b", (size_t)S * NP * sizeof(double));
  double* sig = (double*)ctx_buf(ctx, "hyb_sig", (size_t)S * S * NP * sizeof(double));
  double* vg = (double*)ctx_buf(ctx, "hyb_vg", (size_t)S * S * NP * sizeof(double));
  int32_t* pgene = (int32_t*)ctx_buf(ctx, "mv_pgene", (size_t)NP * sizeof(int32_t));
  std::vector<uint64_t> hm(1 + n_cfg);
  hm[0] = (S >= 64) ? ~0ull : ((1ull << S) - 1ull);
  if (n_cfg > 0) eqbf_config_masks(o->bfs, S, hm.data() + 1);
  uint64_t* dmask = (uint64_t*)ctx_buf(ctx, "mv_masks", hm.size() * sizeof(uint64_t));
  if (!dtiles || !rec || !bb || !sig || !vg || !pgene || !dmask)
    return ctx_fail(ctx, EQBF_ERR_HIP, "allocating hybrid buffers failed");
  EQBF_HIP(ctx, hipMemcpyAsync(dtiles, tiles.data(), tiles.size() * sizeof(HybTile), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(dmask, hm.data(), hm.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st)); /* the host vectors go out of scope */
  HybGene hg;
  hg.cp = o->covars_by_position ? o->covars_by_position : c.covars;
  hg.yp = o->explevels_pairs ? o->explevels_pairs : c.explevels;
  hg.rec = rec;
  hg.NSET = NSET;
  hg.RSZ = RSZ;
  hg.Q1M = Q1M;
  {
    ProfScope ps(ctx, "hyb_gene_kernel", st);
    hipLaunchKernelGGL(hyb_gene_kernel, dim3(G, NSET), dim3(256), 0, st, c, hg);
  }
  const int W = (N + 63) / 64;
  const bool split = hg.yp != c.explevels;
  uint64_t* pres = (uint64_t*)ctx_buf(ctx, "hyb_pres", (size_t)(split ? 2 : 1) * G * S * W * sizeof(uint64_t));
  if (!pres) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating hybrid buffers failed");
  uint64_t* presp = split ? pres + (size_t)G * S * W : pres;
  hipLaunchKernelGGL(hyb_pres_kernel, dim3(G, S), dim3(64), 0, st, c, c.explevels, pres, W);
  if (split) hipLaunchKernelGGL(hyb_pres_kernel, dim3(G, S), dim3(64), 0, st, c, hg.yp, presp, W);
  HybPair hp;
  hp.tiles = dtiles;
  hp.pres = pres;
  hp.presp = presp;
  hp.cp = hg.cp;
  hp.yp = hg.yp;
  hp.rec = rec;
  hp.b = bb;
  hp.Sig = sig;
  hp.Vg = vg;
  hp.pgene = pgene;
  hp.n_out = out->n;
  hp.sstats = out->sstats;
  hp.flag = (int*)ctx->d_flag.ptr;
  hp.fiterr = o->fiterr;
  hp.NSET = NSET;
  hp.RSZ = RSZ;
  {
    ProfScope ps(ctx, "hyb_pair_kernel", st);
    const dim3 grid((unsigned)tiles.size());
    if (c.geno_dtype == EQBF_F32 && split)
      hipLaunchKernelGGL((hyb_pair_kernel<Q1M, float, true>), grid, dim3(64), 0, st, c, hp);
    else if (c.geno_dtype == EQBF_F32)
      hipLaunchKernelGGL((hyb_pair_kernel<Q1M, float, false>), grid, dim3(64), 0, st, c, hp);
    else if (split)
      hipLaunchKernelGGL((hyb_pair_kernel<Q1M, double, true>), grid, dim3(64), 0, st, c, hp);
    else
      hipLaunchKernelGGL((hyb_pair_kernel<Q1M, double, false>), grid, dim3(64), 0, st, c, hp);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return mv_rows_run<SM>(ctx, c, o->bfs, out, st, [&](int64_t c0, int nb, double* par, int PSZ, int64_t pair_blocks) {
    HybCfg hc;
    hc.b = bb;
    hc.Sig = sig;
    hc.Vg = vg;
    hc.pgene = pgene;
    hc.masks = dmask;
    hc.par = par;
    hc.cfg0 = (int)c0;
    hc.PSZ = PSZ;
    ProfScope ps(ctx, "hyb_cfg_kernel", st);
    with_exact_sm<SM>(c.n_subgroups, [&](auto e) {
      hipLaunchKernelGGL(hyb_cfg_kernel<decltype(e)::value>, dim3((unsigned)pair_blocks, (unsigned)nb), dim3(64), 0,
                         st, c, hc);
    });
  });
}

template <int SM>
int hybrid_run_q(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_hybrid_opts* o, eqbf_assoc_out* out,
                 hipStream_t st) {
  const int Q1 = c.n_covars_max + 1;
  if (Q1 <= 4) return hybrid_run_sm<SM, 4>(ctx, c, o, out, st);
  if (Q1 <= 8) return hybrid_run_sm<SM, 8>(ctx, c, o, out, st);
  return hybrid_run_sm<SM, 17>(ctx, c, o, out, st);
}

}  // namespace

int mvlr_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_mvlr_opts* o, eqbf_assoc_out* out, hipStream_t st) {
  const eqbf_chunk& c = *cp;
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  if (S < 1 || S > 16) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--error mvlr supports 1 to 16 subgroups (S=%d)", S);
  if (N < 1 || G < 0 || c.n_pairs < 0 || c.n_covars_max < 0 || c.n_covars_max > 16)
    return ctx_fail(ctx, EQBF_ERR_ARG, "chunk sizes out of range (N=%d G=%d Q=%d)", N, G, c.n_covars_max);
  if (o->bfs != EQBF_BFS_GEN && o->bfs != EQBF_BFS_SIN && o->bfs != EQBF_BFS_ALL)
    return ctx_fail(ctx, EQBF_ERR_ARG, "unknown --bfs value %d", o->bfs);
  if (!(o->fiterr >= 0.0 && o->fiterr <= 1.0)) return ctx_fail(ctx, EQBF_ERR_ARG, "--fiterr must be in [0, 1]");
  if (GL < 1 || (o->bfs != EQBF_BFS_GEN && GS < 1))
    return ctx_fail(ctx, EQBF_ERR_ARG, "--error mvlr needs the grids (eqbf_set_grids)");
  if (!out || !out->n || !out->w_abf || !out->raw_abf) return ctx_fail(ctx, EQBF_ERR_ARG, "missing output buffers");
  if (G == 0 || c.n_pairs == 0) return EQBF_OK;
  if (S <= 4) return mvlr_run_sm<4>(ctx, c, o, out, st);
  if (S <= 8) return mvlr_run_sm<8>(ctx, c, o, out, st);
  if (S <= 12) return mvlr_run_sm<12>(ctx, c, o, out, st);
  return mvlr_run_sm<16>(ctx, c, o, out, st);
}

int hybrid_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_hybrid_opts* o, eqbf_assoc_out* out, hipStream_t st) {
  const eqbf_chunk& c = *cp;
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  if (S < 1 || S > 16) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--error hybrid supports 1 to 16 subgroups (S=%d)", S);
  if (N < 1 || G < 0 || c.n_pairs < 0 || c.n_covars_max < 0 || c.n_covars_max > 16)
    return ctx_fail(ctx, EQBF_ERR_ARG, "chunk sizes out of range (N=%d G=%d Q=%d)", N, G, c.n_covars_max);
  if (o->bfs != EQBF_BFS_GEN && o->bfs != EQBF_BFS_SIN && o->bfs != EQBF_BFS_ALL)
    return ctx_fail(ctx, EQBF_ERR_ARG, "unknown --bfs value %d", o->bfs);
  if (!(o->fiterr >= 0.0 && o->fiterr <= 1.0)) return ctx_fail(ctx, EQBF_ERR_ARG, "--fiterr must be in [0, 1]");
  if (GL < 1 || (o->bfs != EQBF_BFS_GEN && GS < 1))
    return ctx_fail(ctx, EQBF_ERR_ARG, "--error hybrid needs the grids (eqbf_set_grids)");
  if (!out || !out->n || !out->w_abf || !out->raw_abf) return ctx_fail(ctx, EQBF_ERR_ARG, "missing output buffers");
  if (G == 0 || c.n_pairs == 0) return EQBF_OK;
  if (S <= 4) return hybrid_run_q<4>(ctx, c, o, out, st);
  if (S <= 8) return hybrid_run_q<8>(ctx, c, o, out, st);
  if (S <= 12) return hybrid_run_q<12>(ctx, c, o, out, st);
  return hybrid_run_q<16>(ctx, c, o, out, st);
}
