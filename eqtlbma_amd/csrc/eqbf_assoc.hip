/*
 * Association tests with the uvlr error model: per-subgroup OLS summary
 * statistics and Wen-Stephens Bayes factors for every gene-SNP pair of a
 * chunk.
 *
 * Reference path (eqtlbma 1.3.3):
 *   Gene::TestForAssociations           gene.cpp:263-332
 *   GeneSnpPair::FillStlContainers      gene_snp_pair.cpp:79-170
 *   GeneSnpPair::CalcSstatsOneSbgrp     gene_snp_pair.cpp:175-208
 *   utils::FitSingleGeneWithSingleSnp   utils_math.cpp:166-209
 *   StandardizeSstatsAndCorrectSmallSampleSize  gene_snp_pair.cpp:256-290
 *   CalcLog10AbfUvlr / CalcAbfsUvlr*    gene_snp_pair.cpp:297-622
 *   utils::log10_weighted_sum           utils_math.cpp:100-159
 *
 * Algebra.  For subgroup s of gene g the members are the samples with a
 * genotype column in s and a non-missing expression value.  With U an
 * orthonormal basis of [1, covariates] over the members and y~ = y - U U'y,
 * the OLS fit of y on [1, covariates, x] has (Frisch-Waugh-Lovell)
 *   x~'x~ = x'x - |U'x|^2,  betahat = x'y~ / x~'x~,
 *   rss = y~'y~ - (x'y~)^2 / x~'x~,  se^2 = rss/(N-P) / x~'x~,
 * so a pair needs only the sums x'x, x'y~ and U'x over the members.  The prep
 * kernel builds U and y~ once per (gene, subgroup); the pair kernels stream
 * the genotypes once and accumulate those sums in FP64.
 *
 * Two paths:
 *   mixed (EQBF_PREC_MIXED): one persistent launch over a packed list of
 *     64-pair tiles of every gene of the chunk, XCD-aware; FP64 sums and
 *     residual sums of squares; t-tail by tabulated cubic Hermite in FP64;
 *     Bayes-factor terms in FP32; grid averages with FP64 accumulation.
 *   plain (EQBF_PREC_FP64_PLAIN): one launch per gene, one thread per pair,
 *     every quantity in FP64 with the reference's loop structure;
 *     EQBF_PREC_FP64_BATCHED runs the same per-pair code for every gene of
 *     the call in one launch.
 */
#include <algorithm>
#include <cstring>
#include <mutex>

#include "eqbf_internal.h"

namespace {

constexpr int TILE = 64;  /* pairs per tile (one per lane)           */
constexpr float kHalfInvLn10 = 0.21714724095162590f; /* 1 / (2 ln 10) */


struct AssocParams {
  const PrepRec* rec;
  const double* B;       /* [G][N][S][KU+1]: y~, U columns              */
  const uint64_t* memb;  /* [G][N] bit s = sample is a member of s      */
  const int32_t* idx;    /* [G][S][N] compact member lists (plain path) */
  const int4* tiles;     /* (gene, first pair, count, 0)                */
  const int* n_tiles;
  const double* tz;
  const double* grid;    /* phiL, omaL, phiS, omaS                      */
  const uint64_t* masks; /* configurations                              */
  int32_t* n_out;
  double* sstats;
  double* stdsstats;
  void* raw;
  double* w;
  int* flag;
  int GL, GS, n_cfg, n_w, KU, bfs, join, raw_f64, raw_tiled, multiset, CH;
};

__device__ inline double geno_load(const eqbf_chunk& c, int set, int64_t i, int64_t j) {
  int64_t off = (int64_t)set * c.geno_set_stride + i * c.geno_ld + j;
  if (c.geno_dtype == EQBF_F32) return (double)((const float*)c.genotypes)[off];
  return ((const double*)c.genotypes)[off];
}

/* ---- block helpers ----------------------------------------------------------- */
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
  return t;
}

__global__ void zero_u64_kernel(uint64_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

/* ---- prep: one workgroup per (gene, subgroup) ---------------------------------- */
__global__ void __launch_bounds__(256) prep_kernel(const eqbf_chunk c, PrepRec* __restrict__ rec,
                                                   double* __restrict__ B, uint64_t* __restrict__ memb,
                                                   int32_t* __restrict__ idxs, int* __restrict__ flag, int KU) {
  __shared__ double red[8];
  __shared__ int wcnt[4];
  __shared__ int kept[17];
  const int S = c.n_subgroups, N = c.n_samples;
  const int g = blockIdx.x / S, s = blockIdx.x % S;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  PrepRec r;
  r.n = 0;
  r.flags = 0;
  r.q = c.covar_count[s];
  r.rank = 0;
  r.yy = 0;
  r.tss = 0;
  if (!c.gene_in_sbgrp[(int64_t)g * S + s]) {
    if (tid == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  const double* y = c.explevels + ((int64_t)g * S + s) * N;
  const uint8_t* gin = c.geno_in_sbgrp + (int64_t)s * N;
  int32_t* idx = idxs + ((int64_t)g * S + s) * N;
  /* membership and compaction in sample order (gene_snp_pair.cpp:98-116) */
  int count = 0;
  for (int base = 0; base < N; base += 256) {
    int i = base + tid;
    bool m = (i < N) && !isnan(y[i]) && gin[i];
    uint64_t bal = __ballot(m);
    int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int off = count;
    for (int k = 0; k < wv; ++k) off += wcnt[k];
    if (m) idx[off + pre] = i;
    int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
    count += tot;
  }
  __syncthreads();
  const int n = count, Q = r.q;
  r.n = n;
  if (n == 0) {
    if (tid == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  if (Q > 0) {
    int miss = 0;
    for (int k = tid; k < n; k += 256) miss |= !c.covar_in_sbgrp[(int64_t)s * N + idx[k]];
    if (__syncthreads_or(miss)) {
      if (tid == 0) {
        atomicOr(flag, FLAG_MISSING_COVAR);
        r.n = 0;
        rec[(int64_t)g * S + s] = r;
      }
      return;
    }
  }
  if (n < Q + 3) { /* N < P + 1 with P = Q + 2 (utils_math.cpp:175-178) */
    if (tid == 0) {
      r.flags = PREP_TOO_FEW;
      rec[(int64_t)g * S + s] = r;
    }
    return;
  }
  const int KC = KU + 1;
  auto Bat = [&](int i, int col) -> double& { return B[(((int64_t)g * N + i) * S + s) * KC + col]; };
  /* orthonormal basis of [1, covariates] by two-pass modified Gram-Schmidt */
  int rank = 0;
  for (int j = 0; j <= Q; ++j) {
    double nn = 0;
    for (int k = tid; k < n; k += 256) {
      int i = idx[k];
      double v = (j == 0) ? 1.0 : c.covars[((int64_t)s * c.n_covars_max + (j - 1)) * N + i];
      Bat(i, 1 + j) = v;
      nn += v * v;
    }
    double n0 = sqrt(block_sum(nn, red));
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < j; ++l) {
        if (!kept[l]) continue;
        double d = 0;
        for (int k = tid; k < n; k += 256) {
          int i = idx[k];
          d += Bat(i, 1 + l) * Bat(i, 1 + j);
        }
        d = block_sum(d, red);
        for (int k = tid; k < n; k += 256) {
          int i = idx[k];
          Bat(i, 1 + j) -= d * Bat(i, 1 + l);
        }
      }
    double nv = 0;
    for (int k = tid; k < n; k += 256) {
      double v = Bat(idx[k], 1 + j);
      nv += v * v;
    }
    nv = sqrt(block_sum(nv, red));
    bool keep = nv > 1e-9 * n0 && nv > 0;
    double sc = keep ? 1.0 / nv : 0.0;
    for (int k = tid; k < n; k += 256) Bat(idx[k], 1 + j) *= sc;
    if (tid == 0) kept[j] = keep;
    rank += keep;
    __syncthreads();
  }
  for (int j = Q + 1; j < KU; ++j)
    for (int k = tid; k < n; k += 256) Bat(idx[k], 1 + j) = 0.0;
  /* y~ = y - U U'y (two passes), y~'y~, and the total sum of squares */
  double sy = 0;
  for (int k = tid; k < n; k += 256) {
    int i = idx[k];
    Bat(i, 0) = y[i];
    sy += y[i];
  }
  double ybar = block_sum(sy, red) / n;
  double ts = 0;
  for (int k = tid; k < n; k += 256) {
    double d = y[idx[k]] - ybar;
    ts += d * d;
  }
  double tss = block_sum(ts, red);
  for (int pass = 0; pass < 2; ++pass)
    for (int l = 0; l <= Q; ++l) {
      if (!kept[l]) continue;
      double d = 0;
      for (int k = tid; k < n; k += 256) {
        int i = idx[k];
        d += Bat(i, 1 + l) * Bat(i, 0);
      }
      d = block_sum(d, red);
      for (int k = tid; k < n; k += 256) {
        int i = idx[k];
        Bat(i, 0) -= d * Bat(i, 1 + l);
      }
    }
  double yy = 0;
  for (int k = tid; k < n; k += 256) {
    double v = Bat(idx[k], 0);
    yy += v * v;
  }
  yy = block_sum(yy, red);
  for (int k = tid; k < n; k += 256) atomicOr((unsigned long long*)&memb[(int64_t)g * N + idx[k]], 1ull << s);
  if (tid == 0) {
    r.rank = rank;
    r.yy = yy;
    r.tss = tss;
    rec[(int64_t)g * S + s] = r;
  }
}

/* ---- prep, one wave per (gene, subgroup): columns held in LDS ------------------ */
/* Same computation as prep_kernel (which remains for very large N): member
 * list, orthonormal basis of [1, covariates] by two-pass modified
 * Gram-Schmidt, y~ = y - U U'y, y~'y~ and the total sum of squares; all
 * reductions are wave-level, the coefficient rows are written once. */
__global__ void __launch_bounds__(64) prep_wave_kernel(const eqbf_chunk c, PrepRec* __restrict__ rec,
                                                       double* __restrict__ B, uint64_t* __restrict__ memb,
                                                       int32_t* __restrict__ idxs, int* __restrict__ flag, int KU) {
  extern __shared__ __align__(16) unsigned char psm[];
  const int S = c.n_subgroups, N = c.n_samples;
  const int g = blockIdx.x / S, s = blockIdx.x % S;
  const int lane = threadIdx.x;
  const int KC = KU + 1;
  double* col = (double*)psm;                 /* [KC][N]: y~, U columns over members */
  int* lidx = (int*)(col + (size_t)KC * N);   /* [N] member sample indices          */
  PrepRec r;
  r.n = 0;
  r.flags = 0;
  r.q = c.covar_count[s];
  r.rank = 0;
  r.yy = 0;
  r.tss = 0;
  if (!c.gene_in_sbgrp[(int64_t)g * S + s]) {
    if (lane == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  const double* y = c.explevels + ((int64_t)g * S + s) * N;
  const uint8_t* gin = c.geno_in_sbgrp + (int64_t)s * N;
  int32_t* idx = idxs + ((int64_t)g * S + s) * N;
  int n = 0;
  for (int base = 0; base < N; base += 64) {
    const int i = base + lane;
    const bool m = (i < N) && !isnan(y[i]) && gin[i];
    const uint64_t bal = __ballot(m);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (m) {
      lidx[n + pre] = i;
      idx[n + pre] = i;
    }
    n += __popcll(bal);
  }
  __syncthreads();
  const int Q = r.q;
  r.n = n;
  if (n == 0) {
    if (lane == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  if (Q > 0) {
    int miss = 0;
    for (int k = lane; k < n; k += 64) miss |= !c.covar_in_sbgrp[(int64_t)s * N + lidx[k]];
    if (__ballot(miss)) {
      if (lane == 0) {
        atomicOr(flag, FLAG_MISSING_COVAR);
        r.n = 0;
        rec[(int64_t)g * S + s] = r;
      }
      return;
    }
  }
  if (n < Q + 3) { /* N < P + 1 with P = Q + 2 (utils_math.cpp:175-178) */
    if (lane == 0) {
      r.flags = PREP_TOO_FEW;
      rec[(int64_t)g * S + s] = r;
    }
    return;
  }
  /* load y and the columns [1, covariates] of the members */
  double sy = 0;
  for (int k = lane; k < n; k += 64) {
    const int i = lidx[k];
    col[k] = y[i];
    sy += y[i];
    col[(size_t)1 * N + k] = 1.0;
    for (int j = 1; j <= Q; ++j) col[(size_t)(1 + j) * N + k] = c.covars[((int64_t)s * c.n_covars_max + (j - 1)) * N + i];
    for (int j = Q + 1; j < KU; ++j) col[(size_t)(1 + j) * N + k] = 0.0;
  }
  const double ybar = wave_sum(sy) / n;
  double ts = 0;
  for (int k = lane; k < n; k += 64) {
    const double d = col[k] - ybar;
    ts += d * d;
  }
  const double tss = wave_sum(ts);
  uint32_t kept = 0;
  int rank = 0;
  for (int j = 0; j <= Q; ++j) {
    double* cj = col + (size_t)(1 + j) * N;
    double nn = 0;
    for (int k = lane; k < n; k += 64) nn += cj[k] * cj[k];
    const double n0 = sqrt(wave_sum(nn));
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < j; ++l) {
        if (!((kept >> l) & 1u)) continue;
        const double* cl = col + (size_t)(1 + l) * N;
        double d = 0;
        for (int k = lane; k < n; k += 64) d += cl[k] * cj[k];
        d = wave_sum(d);
        for (int k = lane; k < n; k += 64) cj[k] -= d * cl[k];
      }
    double nv = 0;
    for (int k = lane; k < n; k += 64) nv += cj[k] * cj[k];
    nv = sqrt(wave_sum(nv));
    const bool keep = nv > 1e-9 * n0 && nv > 0;
    const double sc = keep ? 1.0 / nv : 0.0;
    for (int k = lane; k < n; k += 64) cj[k] *= sc;
    if (keep) {
      kept |= 1u << j;
      ++rank;
    }
  }
  for (int pass = 0; pass < 2; ++pass)
    for (int l = 0; l <= Q; ++l) {
      if (!((kept >> l) & 1u)) continue;
      const double* cl = col + (size_t)(1 + l) * N;
      double d = 0;
      for (int k = lane; k < n; k += 64) d += cl[k] * col[k];
      d = wave_sum(d);
      for (int k = lane; k < n; k += 64) col[k] -= d * cl[k];
    }
  double yy = 0;
  for (int k = lane; k < n; k += 64) yy += col[k] * col[k];
  yy = wave_sum(yy);
  for (int k = lane; k < n; k += 64) {
    const int i = lidx[k];
    double* dst = B + (((int64_t)g * N + i) * S + s) * KC;
    for (int cc = 0; cc < KC; ++cc) dst[cc] = col[(size_t)cc * N + k];
    atomicOr((unsigned long long*)&memb[(int64_t)g * N + i], 1ull << s);
  }
  if (lane == 0) {
    r.rank = rank;
    r.yy = yy;
    r.tss = tss;
    rec[(int64_t)g * S + s] = r;
  }
}

/* The same computation with the columns in registers: member k of the
 * (gene, subgroup) is slot k / 64 of lane k % 64, for N <= 64 NPL samples and
 * at most PREP_REG_KU basis columns.  Slots past the members hold zeros,
 * which every update keeps at zero, so the sums add the same terms in the
 * same order as prep_wave_kernel and the results are identical; only the
 * member list stays in LDS, so many more waves fit a CU. */
constexpr int PREP_REG_KU = 4;
template <int NPL>
__global__ void __launch_bounds__(64) prep_reg_kernel(const eqbf_chunk c, PrepRec* __restrict__ rec,
                                                      double* __restrict__ B, uint64_t* __restrict__ memb,
                                                      int32_t* __restrict__ idxs, int* __restrict__ flag, int KU) {
  constexpr int KCM = PREP_REG_KU + 1;
  __shared__ int lidx[NPL * 64];
  const int S = c.n_subgroups, N = c.n_samples;
  const int g = blockIdx.x / S, s = blockIdx.x % S;
  const int lane = threadIdx.x;
  const int KC = KU + 1;
  PrepRec r;
  r.n = 0;
  r.flags = 0;
  r.q = c.covar_count[s];
  r.rank = 0;
  r.yy = 0;
  r.tss = 0;
  if (!c.gene_in_sbgrp[(int64_t)g * S + s]) {
    if (lane == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  const double* y = c.explevels + ((int64_t)g * S + s) * N;
  const uint8_t* gin = c.geno_in_sbgrp + (int64_t)s * N;
  int32_t* idx = idxs + ((int64_t)g * S + s) * N;
  int n = 0;
  for (int base = 0; base < N; base += 64) {
    const int i = base + lane;
    const bool m = (i < N) && !isnan(y[i]) && gin[i];
    const uint64_t bal = __ballot(m);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (m) {
      lidx[n + pre] = i;
      idx[n + pre] = i;
    }
    n += __popcll(bal);
  }
  __syncthreads();
  const int Q = r.q;
  r.n = n;
  if (n == 0) {
    if (lane == 0) rec[(int64_t)g * S + s] = r;
    return;
  }
  if (Q > 0) {
    int miss = 0;
    for (int k = lane; k < n; k += 64) miss |= !c.covar_in_sbgrp[(int64_t)s * N + lidx[k]];
    if (__ballot(miss)) {
      if (lane == 0) {
        atomicOr(flag, FLAG_MISSING_COVAR);
        r.n = 0;
        rec[(int64_t)g * S + s] = r;
      }
      return;
    }
  }
  if (n < Q + 3) { /* N < P + 1 with P = Q + 2 (utils_math.cpp:175-178) */
    if (lane == 0) {
      r.flags = PREP_TOO_FEW;
      rec[(int64_t)g * S + s] = r;
    }
    return;
  }
  /* y and the columns [1, covariates] of the members */
  int li[NPL];
  double cv[KCM][NPL];
  double sy = 0;
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    const int k = t * 64 + lane;
    const bool in = k < n;
    const int i = in ? lidx[k] : 0;
    li[t] = i;
    const double yv = in ? y[i] : 0.0;
    cv[0][t] = yv;
    sy += yv;
    cv[1][t] = in ? 1.0 : 0.0;
#pragma unroll
    for (int j = 1; j < KCM - 1; ++j)
      cv[1 + j][t] = (in && j <= Q) ? c.covars[((int64_t)s * c.n_covars_max + (j - 1)) * N + i] : 0.0;
  }
  const double ybar = wave_sum(sy) / n;
  double ts = 0;
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    if (t * 64 + lane < n) {
      const double d = cv[0][t] - ybar;
      ts += d * d;
    }
  }
  const double tss = wave_sum(ts);
  uint32_t kept = 0;
  int rank = 0;
#pragma unroll
  for (int j = 0; j < KCM - 1; ++j) {
    if (j > Q) continue;
    double nn = 0;
#pragma unroll
    for (int t = 0; t < NPL; ++t) nn += cv[1 + j][t] * cv[1 + j][t];
    const double n0 = sqrt(wave_sum(nn));
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
      for (int l = 0; l < j; ++l) {
        if (!((kept >> l) & 1u)) continue;
        double d = 0;
#pragma unroll
        for (int t = 0; t < NPL; ++t) d += cv[1 + l][t] * cv[1 + j][t];
        d = wave_sum(d);
#pragma unroll
        for (int t = 0; t < NPL; ++t) cv[1 + j][t] -= d * cv[1 + l][t];
      }
    }
    double nv = 0;
#pragma unroll
    for (int t = 0; t < NPL; ++t) nv += cv[1 + j][t] * cv[1 + j][t];
    nv = sqrt(wave_sum(nv));
    const bool keep = nv > 1e-9 * n0 && nv > 0;
    const double sc = keep ? 1.0 / nv : 0.0;
#pragma unroll
    for (int t = 0; t < NPL; ++t) cv[1 + j][t] *= sc;
    if (keep) {
      kept |= 1u << j;
      ++rank;
    }
  }
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int l = 0; l < KCM - 1; ++l) {
      if (l > Q || !((kept >> l) & 1u)) continue;
      double d = 0;
#pragma unroll
      for (int t = 0; t < NPL; ++t) d += cv[1 + l][t] * cv[0][t];
      d = wave_sum(d);
#pragma unroll
      for (int t = 0; t < NPL; ++t) cv[0][t] -= d * cv[1 + l][t];
    }
  }
  double yy = 0;
#pragma unroll
  for (int t = 0; t < NPL; ++t) yy += cv[0][t] * cv[0][t];
  yy = wave_sum(yy);
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    if (t * 64 + lane < n) {
      const int i = li[t];
      double* dst = B + (((int64_t)g * N + i) * S + s) * KC;
#pragma unroll
      for (int cc = 0; cc < KCM; ++cc)
        if (cc < KC) dst[cc] = cv[cc][t];
      atomicOr((unsigned long long*)&memb[(int64_t)g * N + i], 1ull << s);
    }
  }
  if (lane == 0) {
    r.rank = rank;
    r.yy = yy;
    r.tss = tss;
    rec[(int64_t)g * S + s] = r;
  }
}

template <int NPL>
void launch_prep_reg(int blocks, hipStream_t st, const eqbf_chunk& c, PrepRec* rec, double* B, uint64_t* memb,
                     int32_t* idx, int* flag, int KU) {
  hipLaunchKernelGGL((prep_reg_kernel<NPL>), dim3(blocks), dim3(64), 0, st, c, rec, B, memb, idx, flag, KU);
}

/* ---- packed work list: tiles of up to 64 pairs of one gene ---------------------- */
__global__ void __launch_bounds__(1024) tiles_kernel(const int64_t* __restrict__ pair_offset, int G,
                                                     int4* __restrict__ tiles, int* __restrict__ n_tiles) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < G; base += 1024) {
    int g = base + threadIdx.x;
    int nt = 0;
    if (g < G) nt = (int)((pair_offset[g + 1] - pair_offset[g] + TILE - 1) / TILE);
    part[threadIdx.x] = nt;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) { /* inclusive scan */
      int v = (threadIdx.x >= (unsigned)o) ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    int start = carry + part[threadIdx.x] - nt;
    if (g < G) {
      int64_t p0 = pair_offset[g], p1 = pair_offset[g + 1];
      for (int k = 0; k < nt; ++k) {
        int64_t a = p0 + (int64_t)k * TILE;
        int cnt = (int)min((int64_t)TILE, p1 - a);
        tiles[start + k] = make_int4(g, (int)a, cnt, 0);
      }
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_tiles = carry;
}

/* ---- Bayes-factor phase helpers (FP32 terms, FP64 grid averages) ------------ */
__device__ inline float lg10(float x) { return __log2f(x) * 0.30102999566398120f; }

/* Last step of CalcLog10AbfUvlr (gene_snp_pair.cpp:331-355) from the sums over
 * the active subgroups: L = sum of single-subgroup terms, A = sum b/(v+phi2),
 * Bs = sum 1/(v+phi2).  bbar = A/Bs and V = 1/Bs; the ABF is 0 when no
 * subgroup contributes (Bs == 0), when bbar == 0, or when V is not finite. */
__device__ inline float abf_finish(float L, float A, float Bs, float oma2) {
  if (!(Bs > 0.f)) return 0.f;
  const float bbar = A / Bs;
  const float V = 1.f / Bs;
  if (bbar == 0.f || !(V < INFINITY)) return 0.f;
  const float T2 = bbar * bbar / V;
  float lbar = 0.f;
  if (T2 != 0.f) {
    const float Vo = V + oma2;
    lbar = 0.5f * (lg10(V) - lg10(Vo)) + kHalfInvLn10 * T2 * oma2 / Vo;
  }
  return lbar + L;
}

/* ABFs of configuration `mask` at up to BLK grid points (p2[k], o2[k]).
 * Subgroup-outer loop: each subgroup's statistics are read once from LDS and
 * applied to every point of the block.  Subgroups with |t| < 1e-8 are
 * skipped and absent ones are inactive; an active subgroup with NaN
 * statistics makes the variance sum NaN, so the ABF is 0 (:336-354). */
template <int BLK>
__device__ __noinline__ void abf_block(uint64_t mask, const float* sv, const float* sbp, const float* sz2,
                                 const float* slv, const uint8_t* sfl, int lane, const float* p2,
                                 const float* o2, int nb, float* out) {
  float Lr[BLK], Ar[BLK], Br[BLK];
#pragma unroll
  for (int k = 0; k < BLK; ++k) Lr[k] = Ar[k] = Br[k] = 0.f;
  bool nan = false;
  uint64_t m = mask;
  while (m) {
    const int s = __ffsll((unsigned long long)m) - 1;
    m &= m - 1ull;
    const int f = sfl[s * TILE + lane];
    if (f == F_NAN) nan = true;
    if (f != F_NORMAL) continue;
    const float v = sv[s * TILE + lane], bp = sbp[s * TILE + lane];
    const float z2 = sz2[s * TILE + lane], lv = slv[s * TILE + lane];
#pragma unroll
    for (int k = 0; k < BLK; ++k) {
      if (k < nb) {
        const float vp = v + p2[k];
        const float r = __builtin_amdgcn_rcpf(vp);
        Lr[k] += 0.5f * (lv - lg10(vp)) + kHalfInvLn10 * z2 * p2[k] * r;
        Ar[k] += bp * r;
        Br[k] += r;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < BLK; ++k)
    if (k < nb) out[k] = nan ? 0.f : abf_finish(Lr[k], Ar[k], Br[k], o2[k]);
}

/* Online log10 of the mean of 10^v (utils::log10_weighted_sum semantics:
 * NaN when the first value is NaN, later NaNs skipped). */
struct L10Acc {
  float m;
  double s;
  bool first, nan0;
  __device__ void init() {
    m = 0.f;
    s = 0.0;
    first = true;
    nan0 = false;
  }
  __device__ void add(float v) {
    if (first) {
      first = false;
      if (isnan(v)) {
        nan0 = true;
        return;
      }
      m = v;
      s = 1.0;
      return;
    }
    if (isnan(v)) return;
    if (v > m) {
      s = s * (double)exp2f((m - v) * 3.32192809488736235f) + 1.0;
      m = v;
    } else {
      s += (double)exp2f((v - m) * 3.32192809488736235f);
    }
  }
};

__device__ inline double l10_finish(float m, double s, bool nan, double weight) {
  if (nan) return NAN;
  double res = (double)m + log10(s * weight);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  return res;
}

__device__ inline void store_raw(const AssocParams& kp, int64_t row, int64_t NP, int64_t p, float l) {
  const int64_t ro = raw_index(row, p, NP, 3 * (int64_t)kp.GL + (int64_t)kp.n_cfg * kp.GS, kp.raw_tiled);
  if (kp.raw_f64)
    ((double*)kp.raw)[ro] = l;
  else
    ((float*)kp.raw)[ro] = l;
}

/* Summary statistics of one (pair, subgroup) from its sums x'x, x'y~ and
 * |U'x|^2 (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209, then
 * StandardizeSstatsAndCorrectSmallSampleSize, gene_snp_pair.cpp:256-290);
 * writes the outputs and the tile's LDS copy of the standardised values.
 * Kept out of line so that its temporaries do not stack on the live sums. */
__device__ __noinline__ void pair_stats(const eqbf_chunk& c, const AssocParams& kp, double sxx, double sxy,
                                        double proj, int g, int s, int64_t p, int64_t snp, bool valid, int lane,
                                        float* sv, float* sbp, float* sz2, float* slv, uint8_t* sfl) {
  const int S = c.n_subgroups;
  const int64_t NP = c.n_pairs;
  const PrepRec r = kp.rec[(int64_t)g * S + s];
  const bool has = r.n > 0 && c.snp_in_sbgrp[snp * S + s];
  double o5[5], st3[3];
  const int f = uvlr_stats(r, has, sxx, sxy, proj, kp.tz, o5, st3);
  if (valid) {
    kp.n_out[(int64_t)s * NP + p] = has ? r.n : 0;
    if (kp.sstats)
      for (int k = 0; k < 5; ++k) kp.sstats[((int64_t)k * S + s) * NP + p] = o5[k];
    if (kp.stdsstats)
      for (int k = 0; k < 3; ++k) kp.stdsstats[((int64_t)k * S + s) * NP + p] = st3[k];
  }
  const float v = (float)(st3[1] * st3[1]);
  sv[s * TILE + lane] = v;
  sbp[s * TILE + lane] = (float)st3[0];
  sz2[s * TILE + lane] = (float)(st3[2] * st3[2]);
  slv[s * TILE + lane] = (f == F_NORMAL) ? lg10(v) : 0.f;
  sfl[s * TILE + lane] = (uint8_t)f;
}

/* ---- mixed path: persistent kernel over tiles ----------------------------------- */
constexpr int CHMAX = 8;   /* samples per staged chunk                 */
constexpr int UPRE = 4;    /* staged doubles prefetched per thread     */
constexpr int XPRE = 2;    /* genotype samples prefetched per lane     */
constexpr int BLK = 8;     /* grid points per register block           */

template <int NWV, int SPW, int KU>
__global__ void __launch_bounds__(NWV * 64) assoc_mixed_kernel(const eqbf_chunk c, const AssocParams kp) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NT = NWV * 64;
  constexpr int KC = KU + 1;
  const int S = c.n_subgroups, N = c.n_samples;
  const int64_t NP = c.n_pairs;
  const int CH = kp.CH;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  /* LDS, sum phase: two staging buffers, each CH samples of B rows followed
   * by the CH membership words.  Afterwards: standardised statistics of the
   * tile [S][64] and the partial grid sums of the consistent configuration. */
  const int rows_d = CH * S * KC;
  const int stage_d = rows_d + CH + CH * TILE / 2; /* + genotypes as floats */
  double* bst = (double*)smem;
  float* sv = (float*)smem;
  float* sbp = sv + S * TILE;
  float* sz2 = sbp + S * TILE;
  float* slv = sz2 + S * TILE;
  uint8_t* sfl = (uint8_t*)(slv + S * TILE);
  double* part = (double*)(smem + (((size_t)S * TILE * 17 + 15) / 16) * 16); /* [3][4][2][64] */

  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const int T = *kp.n_tiles;
  const int span = (T + 7) / 8;
  const int t0 = xcd * span, t1 = min(T, t0 + span);
  const double* grid = kp.grid;
  const double *phiL = grid, *omaL = grid + kp.GL, *phiS = grid + 2 * kp.GL, *omaS = phiS + kp.GS;

  for (int t = t0 + slot; t < t1; t += per_xcd) {
    const int4 tile = kp.tiles[t];
    const int g = tile.x;
    const int64_t p0 = tile.y;
    const int cnt = tile.z;
    const bool valid = lane < cnt;
    const int64_t p = p0 + (valid ? lane : 0);
    const int64_t snp = c.pair_snp[p];
    uint64_t wmask = 0;
#pragma unroll
    for (int sl = 0; sl < SPW; ++sl) {
      const int s = w + NWV * sl;
      if (s < S) wmask |= 1ull << s;
    }
    double acc[SPW][KU + 2];
#pragma unroll
    for (int sl = 0; sl < SPW; ++sl)
#pragma unroll
      for (int k = 0; k < KU + 2; ++k) acc[sl][k] = 0.0;
    const double* Bg = kp.B + (int64_t)g * N * S * KC;
    const uint64_t* mg = kp.memb + (int64_t)g * N;
    const int nchunk = (N + CH - 1) / CH;
    /* stage chunk 0: B rows, membership words and the tile's genotypes
     * (sample k of the chunk is loaded by wave k mod NWV) */
    {
      const int ni = min(CH, N);
      double* b0 = bst;
      for (int e = tid; e < ni * S * KC; e += NT) b0[e] = Bg[e];
      for (int e = tid; e < ni; e += NT) ((uint64_t*)(b0 + rows_d))[e] = mg[e];
      float* x0 = (float*)(b0 + rows_d + CH);
      for (int k = w; k < ni; k += NWV) x0[k * TILE + lane] = (float)geno_load(c, 0, k, snp);
    }
    __syncthreads();
    for (int ck = 0; ck < nchunk; ++ck) {
      const int i0 = ck * CH;
      const int ni = min(CH, N - i0);
      const double* cur = bst + (ck & 1) * stage_d;
      double* nxt = bst + ((ck + 1) & 1) * stage_d;
      /* prefetch the next chunk into registers */
      double pre[UPRE];
      float xpre[XPRE];
      int nn = 0, nrow = 0, ni2 = 0;
      if (ck + 1 < nchunk) {
        ni2 = min(CH, N - i0 - CH);
        nrow = ni2 * S * KC;
        nn = nrow + ni2;
        const double* src = Bg + (int64_t)(i0 + CH) * S * KC;
#pragma unroll
        for (int u = 0; u < UPRE; ++u) {
          const int e = tid + u * NT;
          pre[u] = (e < nrow) ? src[e] : (e < nn ? __longlong_as_double((long long)mg[i0 + CH + e - nrow]) : 0.0);
        }
#pragma unroll
        for (int u = 0; u < XPRE; ++u) {
          const int k = w + u * NWV;
          xpre[u] = (k < ni2) ? (float)geno_load(c, 0, i0 + CH + k, snp) : 0.f;
        }
      }
      const uint64_t* mcur = (const uint64_t*)(cur + rows_d);
      const float* xcur = (const float*)(cur + rows_d + CH);
#pragma unroll 1
      for (int ii = 0; ii < ni; ++ii) {
        const uint64_t m = mcur[ii] & wmask;
        if (!m) continue;
        const double x = (double)xcur[ii * TILE + lane];
        const int i = i0 + ii;
#pragma unroll
        for (int sl = 0; sl < SPW; ++sl) {
          const int s = w + NWV * sl;
          if ((m >> s) & 1ull) {
            double xs = x;
            if (kp.multiset) xs = geno_load(c, c.geno_set_of_sbgrp[s], i, snp);
            const double* b = cur + (ii * S + s) * KC;
            acc[sl][0] += xs * xs;
            acc[sl][1] += xs * b[0];
#pragma unroll
            for (int q = 0; q < KU; ++q) acc[sl][2 + q] += xs * b[1 + q];
          }
        }
      }
      __syncthreads();
      if (ck + 1 < nchunk) {
        /* rows at [0, nrow), membership words at [rows_d, rows_d + ni2) */
#pragma unroll
        for (int u = 0; u < UPRE; ++u) {
          const int e = tid + u * NT;
          if (e < nn) nxt[e < nrow ? e : rows_d + (e - nrow)] = pre[u];
        }
        const double* src = Bg + (int64_t)(i0 + CH) * S * KC;
        for (int e = tid + UPRE * NT; e < nn; e += NT) {
          if (e < nrow)
            nxt[e] = src[e];
          else
            nxt[rows_d + (e - nrow)] = __longlong_as_double((long long)mg[i0 + CH + e - nrow]);
        }
        float* xn = (float*)(nxt + rows_d + CH);
#pragma unroll
        for (int u = 0; u < XPRE; ++u) {
          const int k = w + u * NWV;
          if (k < ni2) xn[k * TILE + lane] = xpre[u];
        }
      }
      __syncthreads();
    }
    /* ---- per (pair, subgroup) statistics (FP64), one call per subgroup ---- */
#pragma unroll
    for (int sl = 0; sl < SPW; ++sl) {
      const int s = w + NWV * sl;
      if (s >= S) continue;
      double proj = 0;
#pragma unroll
      for (int q = 0; q < KU; ++q) proj += acc[sl][2 + q] * acc[sl][2 + q];
      pair_stats(c, kp, acc[sl][0], acc[sl][1], proj, g, s, p, snp, valid, lane, sv, sbp, sz2, slv, sfl);
    }
    __syncthreads();
    if (kp.join) {
      /* ---- Bayes factors (CalcAbfsUvlrFor*, gene_snp_pair.cpp:364-550):
       * 12 jobs for the consistent configuration (gen, gen-fix, gen-maxh x 4
       * grid ranges) then one job per configuration. ---- */
      const uint64_t all = (S >= 64) ? ~0ull : ((1ull << S) - 1ull);
      const int njobs = 12 + kp.n_cfg;
      for (int j = w; j < njobs; j += NWV) {
        L10Acc a;
        a.init();
        if (j < 12) {
          const int type = j >> 2, qtr = j & 3;
          const int ga = (kp.GL * qtr) >> 2, gb = (kp.GL * (qtr + 1)) >> 2;
          if (type == 1) {
            /* gen-fix: phi2 = 0 makes every single-subgroup term 0 and the
             * sums independent of the grid point */
            float A = 0.f, Bs = 0.f;
            bool nan = false;
            for (int s = 0; s < S; ++s) {
              const int f = sfl[s * TILE + lane];
              if (f == F_NAN) nan = true;
              if (f != F_NORMAL) continue;
              const float r = __builtin_amdgcn_rcpf(sv[s * TILE + lane]);
              A += sbp[s * TILE + lane] * r;
              Bs += r;
            }
            for (int gi = ga; gi < gb; ++gi) {
              const float l = nan ? 0.f : abf_finish(0.f, A, Bs, (float)(phiL[gi] + omaL[gi]));
              if (valid && kp.raw) store_raw(kp, kp.GL + gi, NP, p, l);
              a.add(l);
            }
          } else {
            for (int g0 = ga; g0 < gb; g0 += BLK) {
              const int nb = min(BLK, gb - g0);
              float p2[BLK], o2[BLK], out[BLK];
#pragma unroll
              for (int k = 0; k < BLK; ++k) {
                const int gi = min(g0 + k, kp.GL - 1);
                p2[k] = type == 0 ? (float)phiL[gi] : (float)(phiL[gi] + omaL[gi]);
                o2[k] = type == 0 ? (float)omaL[gi] : 0.f;
              }
              abf_block<BLK>(all, sv, sbp, sz2, slv, sfl, lane, p2, o2, nb, out);
#pragma unroll
              for (int k = 0; k < BLK; ++k)
                if (k < nb) {
                  if (valid && kp.raw) store_raw(kp, type * kp.GL + g0 + k, NP, p, out[k]);
                  a.add(out[k]);
                }
            }
          }
          double* pp = part + ((type * 4 + qtr) * 2) * TILE;
          pp[lane] = a.first ? -INFINITY : (a.nan0 ? NAN : (double)a.m);
          pp[TILE + lane] = a.s;
        } else {
          const int cfg = j - 12;
          const uint64_t mk = kp.masks[cfg];
          const int64_t row0 = 3 * (int64_t)kp.GL + (int64_t)cfg * kp.GS;
          if (__popcll(mk) == 1) {
            /* singleton: with one active subgroup bbar = b, V = v + phi2 and
             * the two log terms telescope:
             * ABF = (log10 v - log10(v+phi2+oma2))/2
             *       + z^2 (phi2/(v+phi2) + v oma2/((v+phi2)(v+phi2+oma2)))/(2 ln 10) */
            const int s = __ffsll((unsigned long long)mk) - 1;
            const int f = sfl[s * TILE + lane];
            const float v = sv[s * TILE + lane], z2 = sz2[s * TILE + lane], lv = slv[s * TILE + lane];
            for (int gi = 0; gi < kp.GS; ++gi) {
              float l = 0.f;
              if (f == F_NORMAL) {
                const float p2 = (float)phiS[gi], o2 = (float)omaS[gi];
                const float vp = v + p2, va = vp + o2;
                const float r1 = __builtin_amdgcn_rcpf(vp), r2 = __builtin_amdgcn_rcpf(va);
                l = 0.5f * (lv - lg10(va)) + kHalfInvLn10 * z2 * (p2 * r1 + v * o2 * r1 * r2);
              }
              if (valid && kp.raw) store_raw(kp, row0 + gi, NP, p, l);
              a.add(l);
            }
          } else {
            for (int g0 = 0; g0 < kp.GS; g0 += BLK) {
              const int nb = min(BLK, kp.GS - g0);
              float p2[BLK], o2[BLK], out[BLK];
#pragma unroll
              for (int k = 0; k < BLK; ++k) {
                const int gi = min(g0 + k, kp.GS - 1);
                p2[k] = (float)phiS[gi];
                o2[k] = (float)omaS[gi];
              }
              abf_block<BLK>(mk, sv, sbp, sz2, slv, sfl, lane, p2, o2, nb, out);
#pragma unroll
              for (int k = 0; k < BLK; ++k)
                if (k < nb) {
                  if (valid && kp.raw) store_raw(kp, row0 + g0 + k, NP, p, out[k]);
                  a.add(out[k]);
                }
            }
          }
          if (valid) kp.w[(int64_t)(3 + cfg) * NP + p] = l10_finish(a.m, a.s, a.nan0, 1.0 / kp.GS);
        }
      }
      __syncthreads();
      if (w < 3) { /* combine the four grid ranges of gen / gen-fix / gen-maxh */
        const int type = w;
        double M = -INFINITY;
        bool nan = false;
        for (int q = 0; q < 4; ++q) {
          const double m = part[((type * 4 + q) * 2) * TILE + lane];
          if (isnan(m))
            nan = true;
          else if (m > M)
            M = m;
        }
        double ssum = 0;
        for (int q = 0; q < 4; ++q) {
          const double m = part[((type * 4 + q) * 2) * TILE + lane];
          const double sq = part[((type * 4 + q) * 2 + 1) * TILE + lane];
          if (!isnan(m) && m > -INFINITY) ssum += sq * (double)exp2f((float)(m - M) * 3.32192809488736235f);
        }
        if (valid) kp.w[(int64_t)type * NP + p] = l10_finish((float)M, ssum, nan, 1.0 / kp.GL);
      }
    }
    __syncthreads(); /* the next tile's staging reuses this LDS */
  }
}

/* gen-sin (CalcBMAlite, gene_snp_pair.cpp:552-570) and all (CalcBMA
 * :572-602) from the grid-averaged values; one thread per pair, FP64. */
__global__ void bma_kernel(double* __restrict__ w, int64_t NP, int S, int n_cfg, int bfs,
                           const uint64_t* __restrict__ masks, const double* __restrict__ inv_choose) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  if (bfs == EQBF_BFS_GEN) { /* not computed with --bfs gen */
    w[(int64_t)3 * NP + p] = NAN;
    w[(int64_t)4 * NP + p] = NAN;
    return;
  }
  /* BMAlite: singletons 1..S then gen, weights 1/(2S) and 1/2 */
  {
    double v0 = w[3 * NP + p];
    double m = v0;
    for (int s = 0; s < S; ++s) {
      double v = w[(int64_t)(3 + s) * NP + p];
      if (v > m) m = v;
    }
    double vg = w[p];
    if (vg > m) m = vg;
    double sum = 0;
    for (int s = 0; s < S; ++s) {
      double v = w[(int64_t)(3 + s) * NP + p];
      if (!isnan(v)) sum += (0.5 / S) * exp10(v - m);
    }
    if (!isnan(vg)) sum += 0.5 * exp10(vg - m);
    double res = m + log10(sum);
    if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
    w[(int64_t)(3 + n_cfg) * NP + p] = res;
  }
  if (bfs == EQBF_BFS_ALL) {
    double m = w[3 * NP + p];
    for (int k = 0; k < n_cfg; ++k) {
      double v = w[(int64_t)(3 + k) * NP + p];
      if (v > m) m = v;
    }
    double sum = 0;
    for (int k = 0; k < n_cfg; ++k) {
      double v = w[(int64_t)(3 + k) * NP + p];
      if (!isnan(v)) sum += (1.0 / S) * inv_choose[__popcll(masks[k])] * exp10(v - m);
    }
    double res = m + log10(sum);
    if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
    w[(int64_t)(3 + n_cfg + 1) * NP + p] = res;
  } else {
    w[(int64_t)(3 + n_cfg + 1) * NP + p] = NAN;
  }
}

/* Configuration masks (eqbf_config_masks: gsl_combination order, sizes 1..S,
 * lexicographic within a size) and 1 / C(S, k), on the device (the same
 * arithmetic as the host's upload, so the same bits). */
__global__ void config_masks_kernel(int bfs, int S, uint64_t* __restrict__ masks, double* __restrict__ inv_choose) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t k = 0;
  if (bfs == EQBF_BFS_SIN) {
    for (int s = 0; s < S; ++s) masks[k++] = 1ull << s;
  } else if (bfs == EQBF_BFS_ALL) {
    int c[64];
    for (int sz = 1; sz <= S; ++sz) {
      for (int i = 0; i < sz; ++i) c[i] = i;
      while (true) {
        uint64_t m = 0;
        for (int i = 0; i < sz; ++i) m |= 1ull << c[i];
        masks[k++] = m;
        int i = sz - 1;
        while (i >= 0 && c[i] == S - sz + i) --i;
        if (i < 0) break;
        ++c[i];
        for (int j = i + 1; j < sz; ++j) c[j] = c[j - 1] + 1;
      }
    }
  }
  for (int kk = 0; kk < 65; ++kk) inv_choose[kk] = 0.0;
  for (int kk = 1; kk <= S; ++kk) {
    double ch = 1;
    for (int i = 1; i <= kk; ++i) ch = ch * (S - kk + i) / i;
    inv_choose[kk] = 1.0 / ch;
  }
}

/* ---- plain FP64 path: one launch per gene, one thread per pair ------------------ */
__device__ double abf_plain(uint64_t mask, int S, const double* st_b, const double* st_se, const double* st_t,
                            const uint8_t* has, double phi2, double oma2) {
  /* CalcLog10AbfUvlr with the reference's expressions */
  double bn = 0, bd = 0, vb = 0, single = 0;
  bool any_single = false;
  for (int s = 0; s < S; ++s) {
    if (!((mask >> s) & 1ull) || !has[s]) continue;
    double bhat = st_b[s], varbhat = pow(st_se[s], 2), t = st_t[s];
    double l;
    if (fabs(t) < 1e-8) {
      l = 0;
    } else {
      bn += bhat / (varbhat + phi2);
      bd += 1 / (varbhat + phi2);
      vb += 1 / (varbhat + phi2);
      l = 0.5 * log10(varbhat) - 0.5 * log10(varbhat + phi2) + (0.5 * pow(t, 2) * phi2 / (varbhat + phi2)) / log(10.0);
    }
    single += l;
    any_single = true;
  }
  (void)any_single;
  double bbar = (bd != 0.0) ? bn / bd : 0.0;
  vb = (vb != 0.0) ? 1 / vb : INFINITY;
  if (bbar != 0.0 && vb < INFINITY) {
    double T2 = pow(bbar, 2.0) / vb;
    double lb = (T2 != 0) ? 0.5 * log10(vb) - 0.5 * log10(vb + oma2) + (0.5 * T2 * oma2 / (vb + oma2)) / log(10.0) : 0;
    return lb + single;
  }
  return 0.0;
}

__device__ double l10ws_plain(const double* v, int n) {
  double mx = v[0];
  for (int i = 0; i < n; ++i)
    if (v[i] > mx) mx = v[i];
  double w = 1 / ((double)n), sum = 0;
  for (int i = 0; i < n; ++i)
    if (!isnan(v[i])) sum += w * pow(10.0, v[i] - mx);
  double res = mx + log10(sum);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  return res;
}

#define PLAIN_SMAX 64
#define PLAIN_GMAX 128

/* The consistent configuration over the large grid (gen, gen-fix, gen-maxh)
 * and the configurations over the small grid of pair p, FP64
 * (CalcAbfsUvlrForConsistentConfiguration / ForSingletons /
 * ForEachConfiguration, gene_snp_pair.cpp:364-550); gen-sin and all follow in
 * bma_kernel. */
__device__ __forceinline__ void plain_pair_abfs(const AssocParams& kp, int S, int64_t NP, int64_t p,
                                                const double* st_b, const double* st_se, const double* st_t,
                                                const uint8_t* has) {
  const double* phiL = kp.grid;
  const double* omaL = phiL + kp.GL;
  const double* phiS = phiL + 2 * kp.GL;
  const double* omaS = phiS + kp.GS;
  const uint64_t all = (S >= 64) ? ~0ull : ((1ull << S) - 1ull);
  double buf[PLAIN_GMAX];
  for (int type = 0; type < 3; ++type) {
    for (int gi = 0; gi < kp.GL; ++gi) {
      double p2 = type == 0 ? phiL[gi] : (type == 1 ? 0.0 : phiL[gi] + omaL[gi]);
      double o2 = type == 0 ? omaL[gi] : (type == 1 ? phiL[gi] + omaL[gi] : 0.0);
      buf[gi] = abf_plain(all, S, st_b, st_se, st_t, has, p2, o2);
      if (kp.raw) {
        const int64_t ro = raw_index(type * kp.GL + gi, p, NP, 3 * (int64_t)kp.GL + (int64_t)kp.n_cfg * kp.GS,
                                     kp.raw_tiled);
        if (kp.raw_f64)
          ((double*)kp.raw)[ro] = buf[gi];
        else
          ((float*)kp.raw)[ro] = (float)buf[gi];
      }
    }
    kp.w[(int64_t)type * NP + p] = l10ws_plain(buf, kp.GL);
  }
  for (int cfg = 0; cfg < kp.n_cfg; ++cfg) {
    const uint64_t mk = kp.masks[cfg];
    bool any = false;
    for (int s = 0; s < S; ++s) any |= ((mk >> s) & 1ull) && has[s];
    for (int gi = 0; gi < kp.GS; ++gi) {
      buf[gi] = any ? abf_plain(mk, S, st_b, st_se, st_t, has, phiS[gi], omaS[gi]) : 0.0;
      if (kp.raw) {
        const int64_t ro = raw_index(3 * kp.GL + cfg * kp.GS + gi, p, NP,
                                     3 * (int64_t)kp.GL + (int64_t)kp.n_cfg * kp.GS, kp.raw_tiled);
        if (kp.raw_f64)
          ((double*)kp.raw)[ro] = buf[gi];
        else
          ((float*)kp.raw)[ro] = (float)buf[gi];
      }
    }
    kp.w[(int64_t)(3 + cfg) * NP + p] = l10ws_plain(buf, kp.GS);
  }
}

/* g >= 0: pairs of gene g (the plain path, one launch per gene); g < 0:
 * every pair of the chunk in one launch, thread = pair, its gene found by
 * binary search of pair_offset (EQBF_PREC_FP64_BATCHED).  Each pair's
 * arithmetic is the same either way. */
__global__ void __launch_bounds__(64) assoc_plain_kernel(const eqbf_chunk c, const AssocParams kp, int g) {
  const int S = c.n_subgroups, N = c.n_samples;
  const int64_t NP = c.n_pairs;
  int64_t p;
  if (g >= 0) {
    p = c.pair_offset[g] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= c.pair_offset[g + 1]) return;
  } else {
    p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= NP) return;
    int lo = 0, hi = c.n_genes; /* pair_offset[lo] <= p < pair_offset[hi] */
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (c.pair_offset[mid] <= p)
        lo = mid;
      else
        hi = mid;
    }
    g = lo;
  }
  const int64_t snp = c.pair_snp[p];
  const int KC = kp.KU + 1;
  double st_b[PLAIN_SMAX], st_se[PLAIN_SMAX], st_t[PLAIN_SMAX];
  uint8_t has[PLAIN_SMAX];
  for (int s = 0; s < S; ++s) {
    const PrepRec r = kp.rec[(int64_t)g * S + s];
    has[s] = r.n > 0 && c.snp_in_sbgrp[snp * S + s];
    st_b[s] = st_se[s] = st_t[s] = NAN;
    double o5[5] = {NAN, NAN, NAN, NAN, NAN};
    if (has[s] && !(r.flags & PREP_TOO_FEW)) {
      const int32_t* idx = kp.idx + ((int64_t)g * S + s) * N;
      double sxx = 0, sxy = 0, a[17];
      for (int q = 0; q < kp.KU; ++q) a[q] = 0;
      const int set = c.geno_set_of_sbgrp[s];
      for (int k = 0; k < r.n; ++k) {
        const int i = idx[k];
        const double x = geno_load(c, set, i, snp);
        const double* b = kp.B + (((int64_t)g * N + i) * S + s) * KC;
        sxx += x * x;
        sxy += x * b[0];
        for (int q = 0; q < kp.KU; ++q) a[q] += x * b[1 + q];
      }
      double proj = 0;
      for (int q = 0; q < kp.KU; ++q) proj += a[q] * a[q];
      const double xx = sxx - proj;
      const int n = r.n, P = r.q + 2;
      if (xx > 1e-12 * sxx) {
        const double beta = sxy / xx;
        const double rss = r.yy - sxy * beta;
        const int rank = r.rank + 1;
        const double se = sqrt(rss / (double)(n - P) / xx);
        const double sig = sqrt(rss / (double)(n - rank));
        o5[0] = 1.0 - rss / r.tss;
        o5[1] = sig;
        o5[2] = beta;
        o5[3] = se;
        o5[4] = 2.0 * dm::t_lower_neg(fabs(beta / se), (double)(n - rank));
        /* StandardizeSstatsAndCorrectSmallSampleSize (gene_snp_pair.cpp:256-290) */
        double bhat = beta / sig, sebhat = se / sig, t = bhat / sebhat;
        if (!isnan(t)) {
          double nu = (double)n - 2 - r.q;
          t = dm::gauss_pinv(dm::t_lower_neg(fabs(bhat / sebhat), nu));
          if (fabs(t) > 1e-8) {
            double sig2 = fabs(beta) / (fabs(t) * sebhat);
            bhat = beta / sig2;
            sebhat = fabs(bhat / t);
          } else {
            bhat = 0;
            sebhat = INFINITY;
          }
          st_b[s] = bhat;
          st_se[s] = sebhat;
          st_t[s] = t;
        }
      } else {
        o5[0] = 1.0 - r.yy / r.tss;
        o5[1] = sqrt(r.yy / (double)(n - r.rank));
      }
    }
    kp.n_out[(int64_t)s * NP + p] = has[s] ? r.n : 0;
    if (kp.sstats)
      for (int k = 0; k < 5; ++k) kp.sstats[((int64_t)k * S + s) * NP + p] = o5[k];
    if (kp.stdsstats) {
      kp.stdsstats[((int64_t)0 * S + s) * NP + p] = st_b[s];
      kp.stdsstats[((int64_t)1 * S + s) * NP + p] = st_se[s];
      kp.stdsstats[((int64_t)2 * S + s) * NP + p] = st_t[s];
    }
  }
  if (!kp.join) return;
  plain_pair_abfs(kp, S, NP, p, st_b, st_se, st_t, has);
}

/* --inss: summary statistics given per (pair, subgroup) instead of computed
 * (GeneSnpPair::SetSstats, gene_snp_pair.cpp:241-254), standardised with
 * nu = n - 2 since no covariate count is known for them (:273, the map entry
 * defaults to 0), then the Bayes factors as on the plain path; one thread per
 * pair, FP64.  n <= 0: no result in that subgroup (HasResults, :61-68). */
__global__ void __launch_bounds__(64) inss_abf_kernel(const eqbf_sstats_in in, const AssocParams kp) {
  const int S = in.n_subgroups;
  const int64_t NP = in.n_pairs;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= NP) return;
  double st_b[PLAIN_SMAX], st_se[PLAIN_SMAX], st_t[PLAIN_SMAX];
  uint8_t has[PLAIN_SMAX];
  for (int s = 0; s < S; ++s) {
    const int64_t e = (int64_t)s * NP + p;
    const int64_t n = in.n[e];
    has[s] = n > 0;
    st_b[s] = st_se[s] = st_t[s] = NAN;
    if (has[s]) {
      /* StandardizeSstatsAndCorrectSmallSampleSize (:256-290) */
      const double beta = in.betahat[e];
      double bhat = beta / in.sigmahat[e], sebhat = in.sebetahat[e] / in.sigmahat[e], t = bhat / sebhat;
      if (!isnan(t)) {
        const double nu = (double)n - 2;
        t = nu > 0 ? dm::gauss_pinv(dm::t_lower_neg(fabs(bhat / sebhat), nu)) : NAN;
        if (fabs(t) > 1e-8) {
          const double sig2 = fabs(beta) / (fabs(t) * sebhat);
          bhat = beta / sig2;
          sebhat = fabs(bhat / t);
        } else {
          bhat = 0;
          sebhat = INFINITY;
        }
        st_b[s] = bhat;
        st_se[s] = sebhat;
        st_t[s] = t;
      }
    }
    if (kp.n_out) kp.n_out[e] = has[s] ? (int32_t)(n < 2147483647 ? n : 2147483647) : 0;
    if (kp.stdsstats) {
      kp.stdsstats[e] = st_b[s];
      kp.stdsstats[(int64_t)S * NP + e] = st_se[s];
      kp.stdsstats[(int64_t)2 * S * NP + e] = st_t[s];
    }
  }
  plain_pair_abfs(kp, S, NP, p, st_b, st_se, st_t, has);
}

template <int NWV, int SPW, int KU>
void launch_mixed(const eqbf_chunk& c, const AssocParams& kp, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((assoc_mixed_kernel<NWV, SPW, KU>), dim3(grid), dim3(NWV * 64), lds, st, c, kp);
}

/* Kernel variant: 4 waves per workgroup for S <= 8, else 8 waves, each wave
 * summing ceil(S / waves) subgroups. */
template <int KU>
int dispatch_mixed(int S, const eqbf_chunk& c, const AssocParams& kp, int grid, size_t lds, hipStream_t st) {
  if (S <= 4) return launch_mixed<4, 1, KU>(c, kp, grid, lds, st), 0;
  if (S <= 8) return launch_mixed<4, 2, KU>(c, kp, grid, lds, st), 0;
  if (S <= 16) return launch_mixed<8, 2, KU>(c, kp, grid, lds, st), 0;
  if (S <= 24) return launch_mixed<8, 3, KU>(c, kp, grid, lds, st), 0;
  if (S <= 32) return launch_mixed<8, 4, KU>(c, kp, grid, lds, st), 0;
  if (S <= 48) return launch_mixed<8, 6, KU>(c, kp, grid, lds, st), 0;
  if (S <= 64) return launch_mixed<8, 8, KU>(c, kp, grid, lds, st), 0;
  return -1;
}

}  // namespace

/* ---- orchestration ------------------------------------------------------------- */
int launch_bma(eqbf_ctx* ctx, double* w, int64_t NP, int S, int bfs, hipStream_t st) {
  const int64_t n_cfg = eqbf_n_configs(bfs, S);
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", S);
  const int64_t nm = std::max<int64_t>(n_cfg, 1);
  uint64_t* d_masks = (uint64_t*)ctx_buf(ctx, "bma_tab", (size_t)nm * sizeof(uint64_t) + 65 * sizeof(double));
  if (!d_masks) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating the BMA tables failed");
  double* d_inv_choose = (double*)(d_masks + nm);
  std::vector<uint64_t> hm(nm, 0);
  if (n_cfg > 0) eqbf_config_masks(bfs, S, hm.data());
  std::vector<double> ic(65, 0.0);
  for (int k = 1; k <= S; ++k) {
    double ch = 1;
    for (int i = 1; i <= k; ++i) ch = ch * (S - k + i) / i;
    ic[k] = 1.0 / ch;
  }
  EQBF_HIP(ctx, hipMemcpyAsync(d_masks, hm.data(), nm * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipMemcpyAsync(d_inv_choose, ic.data(), 65 * sizeof(double), hipMemcpyHostToDevice, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st)); /* host vectors go out of scope */
  {
    ProfScope ps(ctx, "bma_kernel", st);
    hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, st, w, NP, S, (int)n_cfg,
                       n_cfg > 0 ? bfs : EQBF_BFS_GEN, d_masks, d_inv_choose);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}

int assoc_uvlr_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_assoc_opts* o, eqbf_assoc_out* out,
                   hipStream_t st, ScratchArena* arena) {
  const eqbf_chunk& c = *cp;
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  if (S < 1 || S > 64 || N < 1 || G < 0 || c.n_covars_max < 0 || c.n_covars_max > 16)
    return ctx_fail(ctx, EQBF_ERR_ARG, "chunk sizes out of range (S=%d N=%d G=%d Q=%d)", S, N, G, c.n_covars_max);
  if (!out || !out->n || (o->analysis == EQBF_ANALYSIS_JOIN && !out->w_abf))
    return ctx_fail(ctx, EQBF_ERR_ARG, "missing output buffers");
  if (o->raw_dtype != EQBF_F32 && o->raw_dtype != EQBF_F64 && o->raw_dtype != EQBF_F32_TILED)
    return ctx_fail(ctx, EQBF_ERR_ARG, "invalid raw_dtype %d", o->raw_dtype);
  const bool join = o->analysis == EQBF_ANALYSIS_JOIN;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  if (join && GL < 1) return ctx_fail(ctx, EQBF_ERR_ARG, "--analys join needs the large grid (eqbf_set_grids)");
  if (join && o->bfs != EQBF_BFS_GEN && GS < 1)
    return ctx_fail(ctx, EQBF_ERR_ARG, "--bfs sin/all needs the small grid (eqbf_set_grids)");
  int64_t n_cfg = join ? eqbf_n_configs(o->bfs, S) : 0;
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", S);
  if (G == 0 || c.n_pairs == 0) return EQBF_OK;
  const bool plain = o->precision == EQBF_PREC_FP64_PLAIN || o->precision == EQBF_PREC_FP64_BATCHED;
  if (!plain && o->precision != EQBF_PREC_MIXED) return ctx_fail(ctx, EQBF_ERR_ARG, "invalid precision %d", o->precision);
  if (plain && (GL > PLAIN_GMAX || GS > PLAIN_GMAX))
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "plain path supports grids of at most %d points", PLAIN_GMAX);
  int KU = 1;
  while (KU < c.n_covars_max + 1) KU = (KU == 1) ? 4 : KU * 2;
  if (KU > 16) KU = 17;
  if (o->precision == EQBF_PREC_MIXED && KU > 8)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "mixed path supports at most 7 covariates");
  const int KC = KU + 1;

  if (arena && !plain)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "the entry point without allocation runs the FP64 precisions only");
  /* scratch of the call: the context's named buffers, or (arena) slices of
   * the caller-less static scratch array, sized by assoc_scratch_bytes */
  auto buf = [&](const char* name, size_t bytes) -> void* { return arena ? arena->take(bytes) : ctx_buf(ctx, name, bytes); };
  PrepRec* rec = (PrepRec*)buf("rec", (size_t)G * S * sizeof(PrepRec));
  double* B = (double*)buf("B", (size_t)G * N * S * KC * sizeof(double));
  uint64_t* memb = (uint64_t*)buf("memb", (size_t)G * N * sizeof(uint64_t));
  int32_t* idx = (int32_t*)buf("idx", (size_t)G * S * N * sizeof(int32_t));
  const size_t max_tiles = (size_t)(c.n_pairs / TILE) + (size_t)G + 1;
  int4* tiles = arena ? nullptr : (int4*)buf("tiles", max_tiles * sizeof(int4));
  int* n_tiles = arena ? nullptr : (int*)buf("ntiles", 256);
  uint64_t* d_masks = (uint64_t*)buf("masks", (size_t)std::max<int64_t>(n_cfg, 1) * sizeof(uint64_t) + 65 * sizeof(double));
  if (!rec || !B || !memb || !idx || (!arena && (!tiles || !n_tiles)) || !d_masks)
    return arena ? ctx_fail(ctx, EQBF_ERR_ARG, "the call needs more scratch than the static array holds") : EQBF_ERR_HIP;
  double* d_inv_choose = (double*)(d_masks + std::max<int64_t>(n_cfg, 1));
  /* the masks stay on the device until (bfs, S) or the buffer changes, so
   * that a repeated call enqueues kernels only (hipGraph capture) */
  char mkey[96];
  snprintf(mkey, sizeof mkey, "%d/%d/%p/%llu", (int)o->bfs, S, (void*)d_masks, (unsigned long long)ctx->buf_epoch);
  if (arena) {
    hipLaunchKernelGGL(config_masks_kernel, dim3(1), dim3(64), 0, st, n_cfg > 0 ? (int)o->bfs : (int)EQBF_BFS_GEN, S,
                       d_masks, d_inv_choose);
  } else if (n_cfg > 0 && ctx->masks_key != mkey) {
    ctx->masks_key = mkey;
    std::vector<uint64_t> hm(n_cfg);
    eqbf_config_masks(o->bfs, S, hm.data());
    std::vector<double> ic(65, 0.0);
    for (int k = 1; k <= S; ++k) {
      double ch = 1;
      for (int i = 1; i <= k; ++i) ch = ch * (S - k + i) / i;
      ic[k] = 1.0 / ch;
    }
    EQBF_HIP(ctx, hipMemcpyAsync(d_masks, hm.data(), n_cfg * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_inv_choose, ic.data(), 65 * sizeof(double), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st)); /* host vectors go out of scope */
  }
  int rc = arena ? EQBF_OK : ctx_ensure_tz(ctx, N, st); /* the FP64 kernels do not read the tables */
  if (rc) return rc;

  const bool fast = o->precision == EQBF_PREC_MIXED && uvlr_fast_eligible(c, KU);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1024), dim3(256), 0, st, memb, (int64_t)G * N);
  /* the batched path multiplies over every sample: rows of non-members are zero */
  if (fast) hipLaunchKernelGGL(zero_u64_kernel, dim3(4096), dim3(256), 0, st, (uint64_t*)B, (int64_t)G * N * S * KC);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, (uint64_t*)ctx->d_flag.ptr, (int64_t)1);
  {
    ProfScope ps(ctx, "prep_kernel", st);
    const size_t plds = (size_t)N * (KC * sizeof(double) + sizeof(int));
    int* fl = (int*)ctx->d_flag.ptr;
    const int npl = (N + 63) / 64;
    if (KU <= PREP_REG_KU && npl <= 8) {
      switch (npl) {
        case 1: launch_prep_reg<1>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 2: launch_prep_reg<2>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 3: launch_prep_reg<3>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 4: launch_prep_reg<4>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 5: launch_prep_reg<5>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 6: launch_prep_reg<6>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        case 7: launch_prep_reg<7>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
        default: launch_prep_reg<8>(G * S, st, c, rec, B, memb, idx, fl, KU); break;
      }
    } else if (plds <= 64 * 1024)
      hipLaunchKernelGGL(prep_wave_kernel, dim3(G * S), dim3(64), plds, st, c, rec, B, memb, idx,
                         (int*)ctx->d_flag.ptr, KU);
    else
      hipLaunchKernelGGL(prep_kernel, dim3(G * S), dim3(256), 0, st, c, rec, B, memb, idx, (int*)ctx->d_flag.ptr, KU);
  }
  EQBF_HIP(ctx, hipGetLastError());

  AssocParams kp;
  memset(&kp, 0, sizeof kp);
  kp.rec = rec;
  kp.B = B;
  kp.memb = memb;
  kp.idx = idx;
  kp.tiles = tiles;
  kp.n_tiles = n_tiles;
  kp.tz = ctx_tz_table(ctx);
  kp.grid = (const double*)ctx->d_grid.ptr;
  kp.masks = d_masks;
  kp.n_out = out->n;
  kp.sstats = out->sstats;
  kp.stdsstats = out->stdsstats;
  kp.raw = join ? out->raw_abf : nullptr;
  kp.w = out->w_abf;
  kp.flag = (int*)ctx->d_flag.ptr;
  kp.GL = GL;
  kp.GS = GS;
  kp.n_cfg = (int)n_cfg;
  kp.n_w = (int)(3 + n_cfg + 2);
  kp.KU = KU;
  kp.bfs = o->bfs;
  kp.join = join;
  kp.raw_f64 = o->raw_dtype == EQBF_F64;
  kp.raw_tiled = o->raw_dtype == EQBF_F32_TILED;
  kp.multiset = c.n_geno_sets > 1;

  if (plain) {
    if (S > PLAIN_SMAX) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "plain path supports at most 64 subgroups");
    if (o->precision == EQBF_PREC_FP64_BATCHED) {
      ProfScope ps(ctx, "assoc_plain_kernel", st);
      hipLaunchKernelGGL(assoc_plain_kernel, dim3((unsigned)((c.n_pairs + 63) / 64)), dim3(64), 0, st, c, kp, -1);
    } else {
      std::vector<int64_t> off(G + 1);
      EQBF_HIP(ctx, hipMemcpyAsync(off.data(), c.pair_offset, (G + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      EQBF_HIP(ctx, hipStreamSynchronize(st));
      for (int g = 0; g < G; ++g) {
        int64_t np = off[g + 1] - off[g];
        if (np <= 0) continue;
        ProfScope ps(ctx, "assoc_plain_kernel", st);
        hipLaunchKernelGGL(assoc_plain_kernel, dim3((unsigned)((np + 63) / 64)), dim3(64), 0, st, c, kp, g);
      }
    }
    EQBF_HIP(ctx, hipGetLastError());
    if (join && n_cfg > 0) {
      hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((c.n_pairs + 255) / 256)), dim3(256), 0, st, out->w_abf,
                         c.n_pairs, S, (int)n_cfg, o->bfs, d_masks, d_inv_choose);
    } else if (join) {
      hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((c.n_pairs + 255) / 256)), dim3(256), 0, st, out->w_abf,
                         c.n_pairs, S, 0, EQBF_BFS_GEN, d_masks, d_inv_choose);
    }
    EQBF_HIP(ctx, hipGetLastError());
    return EQBF_OK;
  }

  /* mixed path */
  if (join && n_cfg > 60000) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "too many configurations");
  {
    ProfScope ps(ctx, "tiles_kernel", st);
    hipLaunchKernelGGL(tiles_kernel, dim3(1), dim3(1024), 0, st, c.pair_offset, G, tiles, n_tiles);
  }
  if (fast) {
    rc = uvlr_fast_run(ctx, c, o, out, st, rec, B, memb, idx, tiles, n_tiles, KU, d_masks, (int)n_cfg);
    if (rc) return rc;
    if (join) {
      ProfScope ps(ctx, "bma_kernel", st);
      hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((c.n_pairs + 255) / 256)), dim3(256), 0, st, out->w_abf,
                         c.n_pairs, S, (int)n_cfg, n_cfg > 0 ? o->bfs : EQBF_BFS_GEN, d_masks, d_inv_choose);
      EQBF_HIP(ctx, hipGetLastError());
    }
    return EQBF_OK;
  }
  /* LDS: max(two staging buffers of CH samples, statistics + partial sums) */
  int CH = CHMAX;
  while (CH > 1 && (size_t)2 * (CH * S * KC + CH + CH * TILE / 2) * sizeof(double) > 48 * 1024) CH >>= 1;
  kp.CH = CH;
  size_t lds_stage = (size_t)2 * (CH * S * KC + CH + CH * TILE / 2) * sizeof(double);
  size_t lds_stats = (((size_t)S * TILE * 17 + 15) / 16) * 16 + (size_t)3 * 4 * 2 * TILE * sizeof(double);
  size_t lds = std::max(lds_stage, lds_stats);
  int grid = ctx->n_cu * 2;
  grid = (grid + 7) / 8 * 8;
  int r2;
  {
    ProfScope ps(ctx, "assoc_mixed_kernel", st);
    if (KU == 1)
      r2 = dispatch_mixed<1>(S, c, kp, grid, lds, st);
    else if (KU == 4)
      r2 = dispatch_mixed<4>(S, c, kp, grid, lds, st);
    else
      r2 = dispatch_mixed<8>(S, c, kp, grid, lds, st);
  }
  if (r2) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "no kernel variant for S=%d", S);
  EQBF_HIP(ctx, hipGetLastError());
  if (join) {
    ProfScope ps(ctx, "bma_kernel", st);
    hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((c.n_pairs + 255) / 256)), dim3(256), 0, st, out->w_abf,
                       c.n_pairs, S, (int)n_cfg, n_cfg > 0 ? o->bfs : EQBF_BFS_GEN, d_masks, d_inv_choose);
    EQBF_HIP(ctx, hipGetLastError());
  }
  return EQBF_OK;
}

/* --inss: Bayes factors from given summary statistics (inss_abf_kernel, then
 * gen-sin / all in bma_kernel).  in and out are device pointers. */
int abf_from_sstats_run(eqbf_ctx* ctx, const eqbf_sstats_in* in, int bfs, int raw_dtype, eqbf_assoc_out* out,
                        hipStream_t st) {
  const int S = in->n_subgroups;
  const int64_t NP = in->n_pairs;
  if (S < 1 || NP < 0) return ctx_fail(ctx, EQBF_ERR_ARG, "invalid summary statistics sizes");
  if (S > PLAIN_SMAX) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "at most %d subgroups", PLAIN_SMAX);
  if (bfs != EQBF_BFS_GEN && bfs != EQBF_BFS_SIN && bfs != EQBF_BFS_ALL)
    return ctx_fail(ctx, EQBF_ERR_ARG, "invalid bfs %d", bfs);
  if (raw_dtype != EQBF_F32 && raw_dtype != EQBF_F64 && raw_dtype != EQBF_F32_TILED)
    return ctx_fail(ctx, EQBF_ERR_ARG, "invalid raw_dtype");
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  if (GL < 1) return ctx_fail(ctx, EQBF_ERR_ARG, "the large grid is empty (eqbf_set_grids)");
  if (bfs != EQBF_BFS_GEN && GS < 1) return ctx_fail(ctx, EQBF_ERR_ARG, "the small grid is empty (eqbf_set_grids)");
  if (GL > PLAIN_GMAX || GS > PLAIN_GMAX)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "grids of at most %d points", PLAIN_GMAX);
  const int64_t n_cfg = eqbf_n_configs(bfs, S);
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", S);
  if (!out->w_abf) return ctx_fail(ctx, EQBF_ERR_ARG, "w_abf is required");
  if (NP == 0) return EQBF_OK;
  if (!in->n || !in->sigmahat || !in->betahat || !in->sebetahat) return ctx_fail(ctx, EQBF_ERR_ARG, "null input");
  const int64_t nm = std::max<int64_t>(n_cfg, 1);
  uint64_t* d_masks = (uint64_t*)ctx_buf(ctx, "inss_tab", (size_t)nm * sizeof(uint64_t) + 65 * sizeof(double));
  if (!d_masks) return EQBF_ERR_HIP; /* ctx_buf has set the message */
  double* d_inv_choose = (double*)(d_masks + nm);
  {
    std::vector<uint64_t> hm(nm, 0);
    if (n_cfg > 0) eqbf_config_masks(bfs, S, hm.data());
    std::vector<double> ic(65, 0.0);
    for (int k = 1; k <= S; ++k) {
      double ch = 1;
      for (int i = 1; i <= k; ++i) ch = ch * (S - k + i) / i;
      ic[k] = 1.0 / ch;
    }
    EQBF_HIP(ctx, hipMemcpyAsync(d_masks, hm.data(), nm * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_inv_choose, ic.data(), 65 * sizeof(double), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st)); /* host vectors go out of scope */
  }
  AssocParams kp;
  memset(&kp, 0, sizeof kp);
  kp.grid = (const double*)ctx->d_grid.ptr;
  kp.masks = d_masks;
  kp.n_out = out->n;
  kp.stdsstats = out->stdsstats;
  kp.raw = out->raw_abf;
  kp.w = out->w_abf;
  kp.GL = GL;
  kp.GS = GS;
  kp.n_cfg = (int)n_cfg;
  kp.n_w = (int)(3 + n_cfg + 2);
  kp.bfs = bfs;
  kp.join = 1;
  kp.raw_f64 = raw_dtype == EQBF_F64;
  kp.raw_tiled = raw_dtype == EQBF_F32_TILED;
  {
    ProfScope ps(ctx, "inss_abf_kernel", st);
    hipLaunchKernelGGL(inss_abf_kernel, dim3((unsigned)((NP + 63) / 64)), dim3(64), 0, st, *in, kp);
  }
  EQBF_HIP(ctx, hipGetLastError());
  {
    ProfScope ps(ctx, "bma_kernel", st);
    hipLaunchKernelGGL(bma_kernel, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, st, out->w_abf, NP, S,
                       (int)n_cfg, n_cfg > 0 ? bfs : EQBF_BFS_GEN, d_masks, d_inv_choose);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}

/* ---- association that allocates nothing ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): eqbf_assoc_uvlr_device with the FP64 precisions, its
 * per-call scratch (preparation records, basis and residual columns, member
 * lists, configuration masks) carved from one static device array instead of
 * context buffers.  Each device holds its own copy of the array from the
 * first use of the library until the process ends; every call on a device
 * shares it, so calls that overlap on one device race. */
#define NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

/* Called by eqbf_create with the context's device current: the device
 * address of `scratch` on that device (the array's name in host code is a
 * host-side stand-in that no kernel may be given). */
int assoc_noalloc_setup(eqbf_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return EQBF_OK; /* the entry point reports it */
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return EQBF_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return ctx_fail(ctx, EQBF_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  g_scratch_ptr[ctx->device] = p;
  return EQBF_OK;
}

/* Scratch bytes one call needs (the slices assoc_uvlr_run takes, each
 * rounded up to 256 bytes). */
static size_t assoc_scratch_bytes(const eqbf_chunk& c, int64_t n_cfg) {
  const size_t G = (size_t)c.n_genes, S = (size_t)c.n_subgroups, N = (size_t)c.n_samples;
  int KU = 1;
  while (KU < c.n_covars_max + 1) KU = (KU == 1) ? 4 : KU * 2;
  if (KU > 16) KU = 17;
  const size_t KC = (size_t)KU + 1;
  auto r = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; };
  return r(G * S * sizeof(PrepRec)) + r(G * N * S * KC * sizeof(double)) + r(G * N * sizeof(uint64_t)) +
         r(G * S * N * sizeof(int32_t)) + r((size_t)std::max<int64_t>(n_cfg, 1) * sizeof(uint64_t) + 65 * sizeof(double));
}

extern "C" int eqbf_assoc_uvlr_device_noalloc(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_assoc_opts* o,
                                              eqbf_assoc_out* out, void* stream) {
  if (!ctx || !c || !o || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  if (o->precision != EQBF_PREC_FP64_PLAIN && o->precision != EQBF_PREC_FP64_BATCHED)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "the entry point without allocation runs the FP64 precisions only");
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[ctx->device])
    return ctx_fail(ctx, EQBF_ERR_HIP, "no static scratch address for device %d", ctx->device);
  if (c->n_subgroups < 1 || c->n_subgroups > 64 || c->n_samples < 1 || c->n_genes < 0 || c->n_covars_max < 0 ||
      c->n_covars_max > 16)
    return ctx_fail(ctx, EQBF_ERR_ARG, "chunk sizes out of range (S=%d N=%d G=%d Q=%d)", c->n_subgroups, c->n_samples,
                    c->n_genes, c->n_covars_max);
  const int64_t n_cfg = o->analysis == EQBF_ANALYSIS_JOIN ? eqbf_n_configs(o->bfs, c->n_subgroups) : 0;
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", c->n_subgroups);
  const size_t need = assoc_scratch_bytes(*c, n_cfg);
  if (need > NOALLOC_SCRATCH_BYTES)
    return ctx_fail(ctx, EQBF_ERR_ARG,
                    "the call needs %zu bytes of scratch; the static array holds %zu (fewer genes per call)", need,
                    (size_t)NOALLOC_SCRATCH_BYTES);
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  eqbf_assoc_opts ob = *o;
  ob.precision = EQBF_PREC_FP64_BATCHED; /* one launch: no host copy of the pair offsets */
  ScratchArena arena;
  arena.base = (uint8_t*)g_scratch_ptr[ctx->device];
  arena.cap = NOALLOC_SCRATCH_BYTES;
  return assoc_uvlr_run(ctx, c, &ob, out, st, &arena);
}
