/*
 * Permutation statistics, contraction path.
 *
 * Applies to the genes of a chunk whose every sample is a member of every
 * subgroup (expression present everywhere, one genotype file, hard-call
 * genotypes, no covariates).  Then the members of a subgroup do not change
 * under a permutation, x'x and sum(x) are fixed per SNP, y'y is fixed per
 * subgroup, and the only quantity that moves is x'y_perm:
 *
 *   XY[snp][s][p] = sum_i x[i][snp] * yc_s[perm_p(i)]        (yc = y - ybar)
 *
 * for all SNPs of the gene, all subgroups and all permutations: a dense
 * contraction G^T . Y_perm on the matrix cores (MakePermutationsJoin /
 * MakePermutationsSep*, gene.cpp:380-717, recompute this one fit at a time
 * through GeneSnpPair::CalcSstatsOneSbgrp).
 *
 * Exactness: genotypes are small integers and are held exactly in int8.
 * yc is held in 32-bit fixed point (scaled by 2^E per gene and subgroup so
 * that max|yc| maps to ~2^30) and split into four signed 8-bit digits;
 * v_mfma_i32_16x16x64_i8 gives each digit's dot product exactly in int32,
 * and the digits are recombined exactly in FP64.  x'yc is therefore exact up
 * to the fixed-point rounding of yc (2^-31 of its range), far below FP64's
 * own rounding of the reference's SVD fit.
 *
 * From XY the fit (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209) is
 * closed form in FP64 (the residual sum of squares y'y - (x'y)^2/x'x included);
 * the standardised z (gene_snp_pair.cpp:256-290) comes from a table of
 * z(asinh(|t|/sqrt(nu))) = z(atanh(R)) in FP32, R = |x'y| / sqrt(x'x y'y),
 * or from the FP64 tail functions for |t| < 1e-5, where the reference's
 * |z| > 1e-8 test is decided.
 *
 * Because every subgroup has the same members, the standardised variance
 * v = 1/x'x is the same in every subgroup, and the log10 ABF of a
 * configuration gamma with m = |gamma| active subgroups at grid point
 * (phi2, omega2) (CalcLog10AbfUvlr, gene_snp_pair.cpp:297-356) reduces to
 *   l = Cabf_m + q * Z2 + D_m * Z1^2,
 *   Z1 = sum_{s in gamma} z_s,  Z2 = sum_{s in gamma} z_s^2,
 *   r = 1/(v+phi2), q = phi2 r/(2 ln10), Cabf_m = m/2 log10(v r) -
 *   1/2 log10(1 + m r omega2), D_m = omega2 v r^2 / (2 ln10 (1 + m r omega2)),
 * with per-SNP coefficients and per-permutation Z1, Z2.  A subgroup with
 * |z| <= 1e-8 drops out of A and B in the reference; here it contributes
 * z = 0 to Z1 and Z2 and m counts only the others.  The statistic
 * (BMA over configurations, BMAlite, or the consistent configuration,
 * gene_snp_pair.cpp:552-602, then max or log10 average over SNPs,
 * gene.cpp:575-596) is a sum of 2^(l log2 10) terms per SNP with the weights
 * folded into the exponents.
 *
 * Anything the closed forms do not cover for a permutation (a residual sum of
 * squares <= 0, |t| past the tail table, an FP32 sum that overflows) marks
 * the gene, and the caller recomputes that gene on the general path, which
 * is exact for every input.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eqbf_internal.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int PF_NMAX = 2048;        /* samples (LDS holds the digit fragments) */
constexpr int ZT_STRIDE = 3;                             /* every third t-tail node   */
constexpr int ZT_NODES = (TZ_NODES - 1) / ZT_STRIDE + 1; /* nodes of the z table      */
constexpr int ZT_ROW = 4 + 2 * ZT_NODES;                 /* floats per z-table row    */
constexpr int PF_SB = 64;            /* SNPs per Bayes-factor work item         */
constexpr float kLog2_10f = 3.32192809488736235f;
constexpr double kLog2_10 = 3.32192809488736235;

enum { PF_JOIN_ALL = 0, PF_JOIN_GEN = 1, PF_JOIN_GENSIN = 2, PF_SEP = 3 };

__device__ inline double pf_geno(const eqbf_chunk& c, int64_t i, int64_t j) {
  const int64_t off = i * c.geno_ld + j;
  if (c.geno_dtype == EQBF_F32) return (double)((const float*)c.genotypes)[off];
  return ((const double*)c.genotypes)[off];
}

/* ---- which genes qualify ------------------------------------------------------ */
__global__ void __launch_bounds__(256) pf_check_kernel(const eqbf_chunk c, uint8_t* __restrict__ ok) {
  const int g = blockIdx.x, S = c.n_subgroups, N = c.n_samples;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  int b = 0;
  for (int e = threadIdx.x; e < S * N; e += blockDim.x) {
    const int s = e / N, i = e - s * N;
    if (!c.geno_in_sbgrp[(int64_t)s * N + i]) b = 1;
    const double y = c.explevels[((int64_t)g * S + s) * N + i];
    if (!(fabs(y) < INFINITY)) b = 1;
  }
  for (int s = threadIdx.x; s < S; s += blockDim.x)
    if (!c.gene_in_sbgrp[(int64_t)g * S + s]) b = 1;
  const int64_t p0 = c.pair_offset[g], np = c.pair_offset[g + 1] - p0;
  if (np <= 0) b = 1;
  for (int64_t e = threadIdx.x; e < np * N; e += blockDim.x) {
    const int64_t i = e / np, q = e - i * np;
    const double x = pf_geno(c, i, c.pair_snp[p0 + q]);
    if (!(x == rint(x)) || x < -127.0 || x > 127.0) b = 1;
  }
  for (int64_t e = threadIdx.x; e < np * S; e += blockDim.x) {
    const int64_t q = e / S;
    const int s = (int)(e - q * S);
    if (!c.snp_in_sbgrp[(int64_t)c.pair_snp[p0 + q] * S + s]) b = 1;
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) ok[g] = bad ? 0 : 1;
}

/* ---- per (gene, subgroup): centred y in four signed 8-bit digits ---------------- */
struct YStat {
  double yy;    /* sum (y - ybar)^2                        */
  double scale; /* 2^-E: value of one fixed-point unit      */
  double tsum;  /* sum of the fixed-point values (integer)  */
};

__device__ inline double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}
__device__ inline double block_max(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

/* One workgroup per (batch gene, subgroup). */
__global__ void __launch_bounds__(256) pf_prep_y_kernel(const eqbf_chunk c, const int32_t* __restrict__ bgene,
                                                        int Npad, uint32_t* __restrict__ W, YStat* __restrict__ ys) {
  __shared__ double red[256];
  const int S = c.n_subgroups, N = c.n_samples;
  const int bi = blockIdx.x / S, s = blockIdx.x - bi * S;
  const int g = bgene[bi];
  const double* y = c.explevels + ((int64_t)g * S + s) * N;
  double a = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a += y[i];
  const double ybar = block_sum(a, red) / (double)N;
  double q = 0, mx = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double d = y[i] - ybar;
    q += d * d;
    mx = fmax(mx, fabs(d));
  }
  const double yy = block_sum(q, red);
  mx = block_max(mx, red);
  int ex = 0;
  if (mx > 0) {
    int e2;
    (void)frexp(mx, &e2); /* mx < 2^e2 */
    ex = 30 - e2;         /* max |yc| 2^ex < 2^30 */
  }
  const double up = ldexp(1.0, ex);
  double ts = 0;
  uint32_t* Wg = W + ((int64_t)bi * S + s) * Npad;
  for (int i = threadIdx.x; i < Npad; i += blockDim.x) {
    int64_t v = 0;
    if (i < N) v = (int64_t)llrint((y[i] - ybar) * up);
    ts += (double)v;
    uint32_t word = 0;
    for (int j = 0; j < 3; ++j) {
      int d = (int)(v & 0xff);
      if (d >= 128) d -= 256;
      v = (v - d) >> 8;
      word |= (uint32_t)(d & 0xff) << (8 * j);
    }
    word |= (uint32_t)((int)v & 0xff) << 24;
    Wg[i] = word;
  }
  ts = block_sum(ts, red);
  if (threadIdx.x == 0) ys[(int64_t)bi * S + s] = YStat{yy, ldexp(1.0, -ex), ts};
}

/* Genotypes of the batch pairs, SNP-major int8 [pair][Npad], and per pair
 * x'x - (sum x)^2 / N, 1 / that, sum x and the not-estimable flag
 * (uvlr_stats: x'x - |U'x|^2 <= 1e-12 x'x). */
struct PStat {
  double xx, inv_xx, sx, rs_xx; /* rs_xx = 1/sqrt(xx) */
};
__global__ void __launch_bounds__(64) pf_prep_x_kernel(const eqbf_chunk c, const int32_t* __restrict__ bsrc,
                                                       int64_t nbp, int Npad, int8_t* __restrict__ G8,
                                                       PStat* __restrict__ ps, uint8_t* __restrict__ pflag) {
  const int64_t b = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (b >= nbp) return;
  const int N = c.n_samples;
  const int64_t snp = c.pair_snp[bsrc[b]];
  uint32_t* row = (uint32_t*)(G8 + b * Npad);
  double sx = 0, sxx = 0;
  for (int i0 = 0; i0 < Npad; i0 += 4) {
    uint32_t word = 0;
    for (int j = 0; j < 4; ++j) {
      const int i = i0 + j;
      const int xi = i < N ? (int)pf_geno(c, i, snp) : 0;
      sx += xi;
      sxx += (double)xi * xi;
      word |= (uint32_t)(xi & 0xff) << (8 * j);
    }
    row[i0 >> 2] = word;
  }
  const double xx = sxx - sx * sx / (double)N;
  ps[b] = PStat{xx, 1.0 / xx, sx, 1.0 / sqrt(xx)};
  pflag[b] = !(xx > 1e-12 * sxx);
}

/* ---- z as a function of s = asinh(|t| / sqrt(nu)) for one nu -------------------- */
/* From every third node of the t-tail table row of nu (segment A, s in
 * [0, sA]): value and derivative of |z| = -Phi^-1(exp(-w)), d|z|/dw =
 * exp(-w + z^2/2) sqrt(2 pi).  11 KB, so that it sits in LDS. */
__global__ void pf_zeta_table_kernel(const double* __restrict__ tz, int nu, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const double* row = tz + (size_t)nu * TZ_ROW;
  const double h = row[0] * ZT_STRIDE, sA = row[1];
  if (k == 0) {
    out[0] = (float)h;
    out[1] = (float)sA;
    out[2] = (float)(1.0 / h);
    out[3] = 0.f;
  }
  if (k >= ZT_NODES) return;
  const double w = row[4 + 2 * ZT_STRIDE * k], dw = row[4 + 2 * ZT_STRIDE * k + 1];
  const double z = -dm::gauss_pinv_from_w(w);
  const double dz = exp(-w + 0.5 * z * z) * 2.5066282746310002 * dw;
  out[4 + 2 * k] = (float)z;
  out[4 + 2 * k + 1] = (float)dz;
}

/* |z| at s (< sA) by cubic Hermite interpolation of the table. */
__device__ inline float pf_zeta_abs(const float* __restrict__ zt, float s) {
  const float h = zt[0], inv_h = zt[2];
  const float u = s * inv_h;
  int k = (int)u;
  if (k > ZT_NODES - 2) k = ZT_NODES - 2;
  const float f = u - (float)k;
  const float2 a = ((const float2*)(zt + 4))[k], b = ((const float2*)(zt + 4))[k + 1];
  const float f2 = f * f, f3 = f2 * f;
  const float h00 = 2.f * f3 - 3.f * f2 + 1.f, h10 = f3 - 2.f * f2 + f, h01 = -2.f * f3 + 3.f * f2, h11 = f3 - f2;
  return h00 * a.x + h10 * (a.y * h) + h01 * b.x + h11 * (b.y * h);
}

/* z where the reference decides |z| > 1e-8: the FP64 tail functions that the
 * association path uses (rarely taken, kept out of line). */
__device__ __noinline__ float pf_zeta_exact(const double* __restrict__ tz, int nu, double tt, double sxy) {
  const double z = fabs(dm::gauss_pinv_from_w(dm::tz_eval(tz, nu, sqrt(tt))));
  return z > 1e-8 ? (float)copysign(z, sxy) : 0.f;
}

/* ---- the contraction ---------------------------------------------------------------
 * One workgroup per (batch gene, 16 permutations, subgroup).  Phase 1 gathers
 * yc_s[perm(i)] for its 16 permutations into LDS as the B fragments of every
 * 64-sample step (4 digits x 64 lanes x 16 bytes per step).  Phase 2: each
 * wave takes 16-SNP tiles; per step one A fragment (16 samples of one SNP per
 * lane) and four MFMAs (one per digit).  Lane l's 16 bytes of A and of B cover
 * the same 16 samples, in the same order, so the product is the sum over
 * samples whatever the hardware's k order within an operand. */
struct PfGemmArgs {
  const int8_t* G8;
  const uint32_t* W;
  const YStat* ys;
  const PStat* ps;
  const uint8_t* pflag;
  const int4* items;
  const int64_t* bpoff;
  const int32_t* tables;
  const int32_t* btable;
  const float* zt;
  const double* tz;
  float* zeta;   /* join: [nbp][S][NPP]  */
  double* tmax;  /* sep:  [nb][S][NPP] max t^2 over SNPs (-1: none) */
  int* fallback; /* [nb] */
  int S, N, Npad, nks, nu, n_items;
  int64_t nperm, NPP;
};

__device__ inline uint32_t perm_b(uint32_t hi, uint32_t lo, uint32_t sel) { return __builtin_amdgcn_perm(hi, lo, sel); }

template <int MODE>
__global__ void __launch_bounds__(256) pf_gemm_kernel(const PfGemmArgs a) {
  extern __shared__ __align__(16) unsigned char sm[];
  uint32_t* Ws = (uint32_t*)sm;
  v4i* Bf = (v4i*)(sm + (size_t)a.Npad * 4);
  __shared__ double redt[4][16];
  __shared__ __align__(16) float zts[MODE == PF_SEP ? 4 : ZT_ROW];
  /* XCD-aware order (blockIdx & 7 = XCD): contiguous item ranges per XCD, so
   * that neighbouring permutation blocks write their z rows through one L2 */
  const int nit = a.n_items, per = (nit + 7) / 8;
  const int itx = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (itx >= nit) return;
  const int4 it = a.items[itx];
  const int bi = it.x, pb = it.y, s = it.z;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int N = a.N;
  {
    const uint32_t* Wg = a.W + ((int64_t)bi * a.S + s) * a.Npad;
    for (int i = tid; i < a.Npad; i += 256) Ws[i] = Wg[i];
    if (MODE != PF_SEP)
      for (int i = tid; i < ZT_ROW; i += 256) zts[i] = a.zt[i];
  }
  __syncthreads();
  {
    const int32_t* tab = a.tables + (int64_t)a.btable[bi] * a.nperm * N;
    for (int e = tid; e < a.nks * 64; e += 256) {
      const int ks = e >> 6, ln = e & 63;
      const int64_t p = min((int64_t)pb * 16 + (ln & 15), a.nperm - 1);
      const int k0 = ks * 64 + (ln >> 4) * 16;
      const int32_t* row = tab + p * N;
      uint32_t wd[16];
      if ((N & 3) == 0 && k0 + 16 <= N) { /* 16-byte aligned rows: four vector loads */
        const int4* r4 = (const int4*)(row + k0);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int4 v = r4[b];
          wd[4 * b] = Ws[min(max(v.x, 0), N - 1)];
          wd[4 * b + 1] = Ws[min(max(v.y, 0), N - 1)];
          wd[4 * b + 2] = Ws[min(max(v.z, 0), N - 1)];
          wd[4 * b + 3] = Ws[min(max(v.w, 0), N - 1)];
        }
      } else {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
          const int k = k0 + b;
          int idx = k < N ? row[k] : 0;
          idx = min(max(idx, 0), N - 1);
          wd[b] = Ws[idx];
        }
      }
      uint32_t o[4][4];
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        const uint32_t w0 = wd[4 * gq], w1 = wd[4 * gq + 1], w2 = wd[4 * gq + 2], w3 = wd[4 * gq + 3];
        const uint32_t t0 = perm_b(w1, w0, 0x05010400u), t1 = perm_b(w1, w0, 0x07030602u);
        const uint32_t u0 = perm_b(w3, w2, 0x05010400u), u1 = perm_b(w3, w2, 0x07030602u);
        o[0][gq] = perm_b(u0, t0, 0x05040100u);
        o[1][gq] = perm_b(u0, t0, 0x07060302u);
        o[2][gq] = perm_b(u1, t1, 0x05040100u);
        o[3][gq] = perm_b(u1, t1, 0x07060302u);
      }
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        v4i f;
        f.x = (int)o[d][0];
        f.y = (int)o[d][1];
        f.z = (int)o[d][2];
        f.w = (int)o[d][3];
        Bf[(ks * 4 + d) * 64 + ln] = f;
      }
    }
  }
  __syncthreads();

  const int64_t P0 = a.bpoff[bi], np = a.bpoff[bi + 1] - P0;
  const int ntile = (int)((np + 15) / 16);
  const YStat y = a.ys[(int64_t)bi * a.S + s];
  const double invN = 1.0 / (double)N, nu = (double)a.nu, inv_yy = 1.0 / y.yy, rs_yy = sqrt(inv_yy);
  const float sA = zts[MODE == PF_SEP ? 0 : 1];
  const int64_t perm = (int64_t)pb * 16 + (lane & 15);
  const bool pvalid = perm < a.nperm;
  double r2mx = -1.0; /* sep: max R^2 = (x'y)^2 / (x'x y'y) over SNPs; t^2 = nu R^2 / (1 - R^2) */
  bool fb = false;
  for (int t = w; t < ntile; t += 4) {
    v4i acc[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) acc[d] = v4i{0, 0, 0, 0};
    const int64_t arow = P0 + min((int64_t)t * 16 + (lane & 15), np - 1);
    const v4i* ap = (const v4i*)(a.G8 + arow * a.Npad) + (lane >> 4);
    /* the epilogue's per-SNP inputs of this lane's four rows, loaded ahead of
     * the products so that their latency overlaps the matrix work */
    uint8_t fl4[4];
    double sx4[4], sc4[4]; /* sx, and rs_xx (join) or inv_xx (sep) */
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t pr = P0 + min((int64_t)t * 16 + 4 * (lane >> 4) + r, np - 1);
      fl4[r] = a.pflag[pr];
      sx4[r] = a.ps[pr].sx;
      sc4[r] = MODE == PF_SEP ? a.ps[pr].inv_xx : a.ps[pr].rs_xx;
    }
    /* four 64-sample steps per round; the next round's A loads are issued
     * before this round's products */
    v4i af[4], an[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) af[u] = u < a.nks ? ap[u * 4] : v4i{0, 0, 0, 0};
    for (int ks0 = 0; ks0 < a.nks; ks0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) an[u] = ks0 + 4 + u < a.nks ? ap[(ks0 + 4 + u) * 4] : v4i{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ks0 + u < a.nks) {
#pragma unroll
          for (int d = 0; d < 4; ++d)
            acc[d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[u], Bf[((ks0 + u) * 4 + d) * 64 + lane], acc[d], 0,
                                                           0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) af[u] = an[u];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t q = (int64_t)t * 16 + 4 * (lane >> 4) + r;
      if (q >= np || !pvalid) continue;
      const int64_t pr = P0 + q;
      if (fl4[r]) continue; /* no p-value, no z (the statistic's kernels handle these SNPs) */
      const double xy = (double)acc[0][r] + 256.0 * (double)acc[1][r] + 65536.0 * (double)acc[2][r] +
                        16777216.0 * (double)acc[3][r];
      const double sxy = (xy - sx4[r] * y.tsum * invN) * y.scale;
      if (MODE == PF_SEP) {
        /* R^2 = 1 - rss / y'y; the residual sum of squares stays in FP64 */
        const double r2 = sxy * sxy * sc4[r] * inv_yy;
        if (!(1.0 - r2 >= 0.0)) { /* rss < 0: NaN statistic */
          if (q == 0) fb = true;  /* a NaN first statistic: general path */
          continue;
        }
        r2mx = fmax(r2mx, sxy != 0.0 ? r2 : 0.0);
      } else {
        /* R = |x'y| / sqrt(x'x y'y); |t| / sqrt(nu) = R / sqrt(1 - R^2), so
         * s = asinh(|t| / sqrt(nu)) = atanh(R); 1 - R in FP64 */
        const double R = fabs(sxy) * sc4[r] * rs_yy;
        const double omR = 1.0 - R;
        float zeta;
        if (!(omR > 0.0)) {
          fb = true; /* rss <= 0 */
          zeta = 0.f;
        } else {
          const float Rf = (float)R;
          if (Rf * Rf * (float)a.nu < 1e-10f) {
            zeta = pf_zeta_exact(a.tz, a.nu, nu * R * R / (omR * (1.0 + R)), sxy);
          } else {
            float sv;
            if (Rf < 0.3f) {
              const float r2 = Rf * Rf;
              sv = Rf * (1.f + r2 * (1.f / 3.f + r2 * (1.f / 5.f + r2 * (1.f / 7.f + r2 * (1.f / 9.f + r2 * (1.f / 11.f +
                                                                                                          r2 * (1.f / 13.f + r2 * (1.f / 15.f))))))));
            } else {
              sv = 0.34657359027997264f * __log2f((1.f + Rf) * __builtin_amdgcn_rcpf((float)omR));
            }
            if (!(sv < sA)) {
              fb = true;
              zeta = 0.f;
            } else {
              zeta = copysignf(pf_zeta_abs(zts, sv), (float)sxy);
            }
          }
        }
        a.zeta[(pr * a.S + s) * a.NPP + perm] = zeta;
      }
    }
  }
  if (fb) a.fallback[bi] = 1;
  if (MODE == PF_SEP) {
    r2mx = fmax(r2mx, __shfl_xor(r2mx, 16, 64));
    r2mx = fmax(r2mx, __shfl_xor(r2mx, 32, 64));
    if (lane < 16) redt[w][lane] = r2mx;
    __syncthreads();
    if (tid < 16) {
      const int64_t p = (int64_t)pb * 16 + tid;
      const double m = fmax(fmax(redt[0][tid], redt[1][tid]), fmax(redt[2][tid], redt[3][tid]));
      if (p < a.nperm) a.tmax[((int64_t)bi * a.S + s) * a.NPP + p] = m < 0.0 ? -1.0 : nu * m / (1.0 - m);
    }
  }
}

/* ---- per-SNP coefficients of the Bayes factors ---------------------------------------
 * coef[pair][j][0..2S]: q, Cabf_1..S, D_1..S at grid point j of the grids
 * the statistic uses (large grid first when used, then small), scaled by
 * log2(10) so that the terms are 2^l. */
__global__ void pf_coef_kernel(const PStat* __restrict__ ps, const uint8_t* __restrict__ pflag, int64_t nbp,
                               const double* __restrict__ phi, const double* __restrict__ oma, int ngrid, int S,
                               float* __restrict__ coef) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nbp * ngrid) return;
  const int64_t b = e / ngrid;
  const int j = (int)(e - b * ngrid);
  const int nc = 1 + 2 * S;
  float* out = coef + (b * ngrid + j) * nc;
  if (pflag[b]) {
    for (int k = 0; k < nc; ++k) out[k] = 0.f;
    return;
  }
  const double v = ps[b].inv_xx, p2 = phi[j], o2 = oma[j];
  const double r = 1.0 / (v + p2);
  const double c2 = 1.0 / (2.0 * 0.69314718055994531); /* log2(10) / (2 ln 10) */
  out[0] = (float)(p2 * r * c2);
  const double lvr = log2(v * r);
  for (int m = 1; m <= S; ++m) {
    const double den = 1.0 + m * r * o2;
    out[m] = (float)(0.5 * m * lvr - 0.5 * log2(den));
    out[S + m] = (float)(o2 * v * r * r * c2 / den);
  }
}

/* ---- Bayes-factor statistic per (SNP, permutation) --------------------------------------
 * Work item = (batch gene, 64 permutations, block of SNPs); one wave, lane =
 * permutation, so the SNP's coefficients are uniform across the wave.  Writes
 * the block's max (or sum) over SNPs of the linear-scale statistic. */
struct PfBmaArgs {
  const float* zeta;
  const float* coef;
  const uint8_t* pflag;
  const int4* items; /* (bi, pb, sb, slot) */
  const int64_t* bpoff;
  float* pmax;  /* [slot][NPP] */
  double* psum; /* [slot][NPP] */
  int* fallback;
  float lw[2][9]; /* log2 weights: [0] large-grid terms by m, [1] small-grid terms by m */
  float lwc[9];   /* log2 C(S, m) (bounds of pf_bma_max_kernel) */
  int GL, GS, use_max, n_items;
  int64_t NPP, nperm;
};

/* Some |z| <= 1e-8: those subgroups drop out of their configurations
 * (rare; kept out of line). */
template <int S>
struct PfZ {
  float v[S];
};
template <int S>
__device__ __noinline__ float pf_sum_all_zero(const PfZ<S> z, const float* __restrict__ cf, int GS, const PfZ<9> lw,
                                              int nzm) {
  constexpr int NC = (1 << S) - 1;
  constexpr int NCF = 1 + 2 * S;
  float sum = 0.f;
  for (int g = 1; g <= NC; ++g) {
    float Z1 = 0.f, Z2 = 0.f;
#pragma unroll
    for (int s = 0; s < S; ++s)
      if ((g >> s) & 1) {
        Z1 += z.v[s];
        Z2 = fmaf(z.v[s], z.v[s], Z2);
      }
    Z1 *= Z1;
    const int m = __popc(g), me = __popc(g & nzm);
    for (int j = 0; j < GS; ++j) {
      const float* row = cf + j * NCF;
      float l = lw.v[m];
      if (me > 0) l += fmaf(row[S + me], Z1, fmaf(row[0], Z2, row[me]));
      sum += __builtin_amdgcn_exp2f(l);
    }
  }
  return sum;
}

template <int S>
__device__ inline float pf_sum_all(const float (&z)[S], const float* __restrict__ cf, int GS, const float* lw,
                                   int nzm) {
  constexpr int NC = (1 << S) - 1;
  constexpr int NCF = 1 + 2 * S;
  if (nzm != NC) {
    PfZ<S> zz;
    PfZ<9> ww;
#pragma unroll
    for (int s = 0; s < S; ++s) zz.v[s] = z[s];
#pragma unroll
    for (int m = 0; m < 9; ++m) ww.v[m] = lw[m];
    return pf_sum_all_zero<S>(zz, cf, GS, ww, nzm);
  }
  float Z1[NC + 1], Z2[NC + 1];
  Z1[0] = 0.f;
  Z2[0] = 0.f;
#pragma unroll
  for (int g = 1; g <= NC; ++g) {
    const int lo = g & (g - 1), s = __builtin_ctz(g);
    Z1[g] = Z1[lo] + z[s];
    Z2[g] = fmaf(z[s], z[s], Z2[lo]);
  }
#pragma unroll
  for (int g = 1; g <= NC; ++g) Z1[g] = Z1[g] * Z1[g];
  float part[4] = {0.f, 0.f, 0.f, 0.f}; /* four partial sums: independent add chains */
  {
    for (int j = 0; j < GS; ++j) {
      const float* row = cf + j * NCF;
      const float q = row[0];
      float C[S + 1], D[S + 1];
#pragma unroll
      for (int m = 1; m <= S; ++m) {
        C[m] = row[m] + lw[m];
        D[m] = row[S + m];
      }
#pragma unroll
      for (int g = 1; g <= NC; ++g) {
        const int m = __builtin_popcount(g);
        part[g & 3] += __builtin_amdgcn_exp2f(fmaf(D[m], Z1[g], fmaf(q, Z2[g], C[m])));
      }
    }
  }
  return (part[0] + part[1]) + (part[2] + part[3]);
}

template <int S>
__device__ inline float pf_sum_gen(const float (&z)[S], const float* __restrict__ cf, int G, float lw, int nzm) {
  constexpr int NCF = 1 + 2 * S;
  float Z1 = 0.f, Z2 = 0.f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    Z1 += z[s];
    Z2 = fmaf(z[s], z[s], Z2);
  }
  Z1 *= Z1;
  const int me = __popc(nzm);
  float sum = 0.f;
  if (me == 0) return (float)G * __builtin_amdgcn_exp2f(lw);
  for (int j = 0; j < G; ++j) {
    const float* row = cf + j * NCF;
    sum += __builtin_amdgcn_exp2f(fmaf(row[S + me], Z1, fmaf(row[0], Z2, row[me] + lw)));
  }
  return sum;
}

template <int S>
__device__ inline float pf_sum_sin(const float (&z)[S], const float* __restrict__ cf, int G, float lw) {
  constexpr int NCF = 1 + 2 * S;
  float sum = 0.f;
  for (int j = 0; j < G; ++j) {
    const float* row = cf + j * NCF;
    const float c1 = row[1] + lw, k1 = row[0] + row[S + 1];
#pragma unroll
    for (int s = 0; s < S; ++s) sum += __builtin_amdgcn_exp2f(z[s] != 0.f ? fmaf(k1, z[s] * z[s], c1) : lw);
  }
  return sum;
}

template <int S, int MODE>
__global__ void __launch_bounds__(256) pf_bma_kernel(const PfBmaArgs a) {
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int4 it = a.items[item];
  const int bi = it.x, pb = it.y, sb = it.z, slot = it.w;
  const int64_t P0 = a.bpoff[bi], np = a.bpoff[bi + 1] - P0;
  const int64_t q0 = (int64_t)sb * PF_SB, q1 = min(np, q0 + PF_SB);
  const int64_t p = (int64_t)pb * 64 + lane;
  const bool valid = p < a.nperm;
  const int64_t pp = valid ? p : 0;
  const int NCF = 1 + 2 * S;
  const int ngrid = (MODE == PF_JOIN_ALL) ? a.GS : (MODE == PF_JOIN_GEN ? a.GL : a.GL + a.GS);
  float mx = 0.f;
  double sm = 0.0;
  bool bad = false;
  for (int64_t q = q0; q < q1; ++q) {
    const int64_t pr = P0 + q;
    float val;
    if (a.pflag[pr]) {
      val = 1.f; /* every ABF of a non-estimable SNP is 1 */
    } else {
      float z[S];
      int nzm = 0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        z[s] = a.zeta[(pr * S + s) * a.NPP + pp];
        nzm |= (z[s] != 0.f) << s;
      }
      const float* cf = a.coef + pr * (int64_t)ngrid * NCF;
      if (MODE == PF_JOIN_ALL) {
        val = pf_sum_all<S>(z, cf, a.GS, a.lw[1], nzm);
      } else if (MODE == PF_JOIN_GEN) {
        val = pf_sum_gen<S>(z, cf, a.GL, a.lw[0][S], nzm);
      } else {
        val = pf_sum_gen<S>(z, cf, a.GL, a.lw[0][S], nzm) + pf_sum_sin<S>(z, cf + a.GL * NCF, a.GS, a.lw[1][1]);
      }
    }
    bad |= !(val < INFINITY) || !(val > 1e-30f);
    mx = fmaxf(mx, val);
    sm += (double)val;
  }
  if (valid) {
    a.pmax[(int64_t)slot * a.NPP + p] = mx;
    a.psum[(int64_t)slot * a.NPP + p] = sm;
    if (bad) a.fallback[bi] = 1;
  }
}

/* ---- statistic per (gene, permutation) --------------------------------------------- */
struct PfFinalArgs {
  const float* pmax;
  const double* psum;
  const double* tmax;
  const int32_t* bgene;
  const int64_t* bpoff;
  const int32_t* bslot; /* first slot of each batch gene */
  const uint8_t* pflag;
  const double* tz;
  double* stats; /* [G][nperm] */
  int nb, S, nu, mode, use_max, sep_sbgrp;
  int one_slot; /* the maximum of each (gene, permutation) is in its first slot */
  int64_t nperm, NPP;
};

__global__ void pf_final_kernel(const PfFinalArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)a.nb * a.nperm) return;
  const int bi = (int)(e / a.nperm);
  const int64_t p = e - (int64_t)bi * a.nperm;
  const int64_t np = a.bpoff[bi + 1] - a.bpoff[bi];
  double res;
  if (a.mode == PF_SEP) {
    double m = -1.0;
    if (a.sep_sbgrp >= 0) {
      m = a.tmax[((int64_t)bi * a.S + a.sep_sbgrp) * a.NPP + p];
    } else {
      for (int s = 0; s < a.S; ++s) m = fmax(m, a.tmax[((int64_t)bi * a.S + s) * a.NPP + p]);
    }
    if (a.sep_sbgrp >= 0 && a.pflag[a.bpoff[bi]]) {
      res = NAN; /* std::min_element: the first SNP's p-value is NaN */
    } else if (m < 0) {
      res = 1.0;
    } else {
      const double pv = 2.0 * exp(-dm::tz_eval(a.tz, a.nu, sqrt(m)));
      res = pv < 1.0 ? pv : 1.0;
    }
  } else {
    const int nsb = a.one_slot ? 1 : (int)((np + PF_SB - 1) / PF_SB);
    const int slot0 = a.bslot[bi];
    if (a.use_max) {
      float m = 0.f;
      for (int k = 0; k < nsb; ++k) m = fmaxf(m, a.pmax[(int64_t)(slot0 + k) * a.NPP + p]);
      res = log10((double)m);
    } else {
      double sm = 0.0;
      for (int k = 0; k < nsb; ++k) sm += a.psum[(int64_t)(slot0 + k) * a.NPP + p];
      res = log10(sm / (double)np);
    }
    if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  }
  a.stats[(int64_t)a.bgene[bi] * a.nperm + p] = res;
}

/* ---- --pbf all --maxbf: the maximum over SNPs without evaluating every SNP --------------
 * Work item = (batch gene, 4 x 64 permutations), one wave per 64
 * permutations (lane = permutation), the gene's SNPs in blocks of 64 whose
 * Bayes-factor coefficients are staged in LDS.  Every configuration of
 * size m and grid point j has an exponent (log2) of at most
 *   E_mj = lw_m + C_mj + q_j Z2max_m + D_mj Z1max_m^2,
 * Z2max_m = the sum of the m largest z^2, Z1max_m^2 = the larger square of
 * the sums of the m largest and of the m smallest z (q, D >= 0).  Two bounds
 * of log2(statistic) follow: log2(number of terms) + max E_mj (a few
 * operations per grid point, every SNP), and log2 sum_m,j C(S, m) 2^E_mj (one
 * exponential per (m, j), the SNPs that pass the first).  Only a SNP whose
 * bounds reach the lane's running maximum can raise it: those go through two
 * queues in LDS, 64 entries at a time, and are evaluated exactly by
 * pf_sum_all (the arithmetic of pf_bma_kernel), so the running maxima, and the
 * final maximum, are exactly pf_bma_kernel's.  A SNP with some z = 0 (its
 * subgroup drops out of the configurations) skips the bounds. */
constexpr int PFQ = 128;  /* entries per queue and wave      */
constexpr int PF_WG = 4;  /* waves (permutation blocks) per item */

template <int S>
struct PfSorted {
  float Z2[S + 1], Q1[S + 1];
};

template <int S>
__device__ inline PfSorted<S> pf_sorted(const float (&z)[S]) {
  float a[S], b[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    a[s] = z[s];
    b[s] = z[s] * z[s];
  }
#pragma unroll
  for (int i = 1; i < S; ++i)
#pragma unroll
    for (int k = i; k > 0; --k) {
      const float lo = fminf(a[k - 1], a[k]), hi = fmaxf(a[k - 1], a[k]);
      a[k - 1] = lo;
      a[k] = hi;
      const float bl = fmaxf(b[k - 1], b[k]), bs = fminf(b[k - 1], b[k]);
      b[k - 1] = bl;
      b[k] = bs;
    }
  PfSorted<S> r;
  float top = 0.f, bot = 0.f, sq = 0.f;
#pragma unroll
  for (int m = 1; m <= S; ++m) {
    top += a[S - m];
    bot += a[m - 1];
    sq += b[m - 1];
    r.Z2[m] = sq;
    r.Q1[m] = fmaxf(top * top, bot * bot);
  }
  return r;
}

/* first bound: max over (m, j) of E_mj (without the log2 of the term count) */
template <int S>
__device__ inline float pf_bound_max(const PfSorted<S>& t, const float* cf, int GS, const float* lw) {
  constexpr int NCF = 1 + 2 * S;
  float u = -INFINITY;
  for (int j = 0; j < GS; ++j) {
    const float* row = cf + j * NCF;
    const float q = row[0];
#pragma unroll
    for (int m = 1; m <= S; ++m) u = fmaxf(u, fmaf(row[S + m], t.Q1[m], fmaf(q, t.Z2[m], row[m] + lw[m])));
  }
  return u;
}

/* second bound: log2 sum over (m, j) of C(S, m) 2^E_mj; lwc_m = lw_m + log2 C(S, m) */
template <int S>
__device__ inline float pf_bound_sum(const PfSorted<S>& t, const float* cf, int GS, const float* lwc) {
  constexpr int NCF = 1 + 2 * S;
  float sum = 0.f;
  for (int j = 0; j < GS; ++j) {
    const float* row = cf + j * NCF;
    const float q = row[0];
#pragma unroll
    for (int m = 1; m <= S; ++m)
      sum += __builtin_amdgcn_exp2f(fmaf(row[S + m], t.Q1[m], fmaf(q, t.Z2[m], row[m] + lwc[m])));
  }
  return __builtin_amdgcn_logf(sum);
}

template <int S>
__global__ void __launch_bounds__(64 * PF_WG) pf_bma_max_kernel(const PfBmaArgs a) {
  extern __shared__ float cfs[]; /* [64 SNPs][GS][1 + 2S] */
  __shared__ uint32_t q1s[PF_WG][PFQ], q2s[PF_WG][PFQ], mbs[PF_WG][64];
  constexpr int NC = (1 << S) - 1;
  constexpr int NCF = 1 + 2 * S;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int4 it = a.items[blockIdx.x];
  const int bi = it.x, slot = it.w;
  const int64_t P0 = a.bpoff[bi], np = a.bpoff[bi + 1] - P0;
  const int64_t pbase = ((int64_t)it.y + wv) * 64;
  const int64_t p = pbase + lane;
  const bool valid = p < a.nperm;
  const int GS = a.GS, rowf = GS * NCF;
  uint32_t* q1 = q1s[wv];
  uint32_t* q2 = q2s[wv];
  uint32_t* mb = mbs[wv];
  float lw[S + 1], lwc[S + 1];
#pragma unroll
  for (int m = 1; m <= S; ++m) {
    lw[m] = a.lw[1][m];
    lwc[m] = a.lw[1][m] + a.lwc[m];
  }
  const float lg_terms = __builtin_amdgcn_logf((float)(NC * GS));
  mb[lane] = 0u; /* +0.0f: statistics are > 0 */
  float lm = -INFINITY; /* log2 of this lane's running maximum */
  int n1 = 0, n2 = 0;
  bool bad = false;
  int64_t b0 = 0; /* first SNP of the staged block */

  auto zload = [&](int64_t pr, int64_t pp, float (&z)[S]) {
#pragma unroll
    for (int s = 0; s < S; ++s) z[s] = a.zeta[(pr * S + s) * a.NPP + pp];
  };
  /* exact statistics of the first k entries of queue 2 */
  auto drain2 = [&](int k) {
    if (lane < k) {
      const uint32_t e = q2[lane];
      const int ql = (int)(e >> 6), le = (int)(e & 63u);
      float z[S];
      zload(P0 + b0 + ql, pbase + le, z);
      int nzm = 0;
#pragma unroll
      for (int s = 0; s < S; ++s) nzm |= (z[s] != 0.f) << s;
      const float val = pf_sum_all<S>(z, cfs + ql * rowf, GS, a.lw[1], nzm);
      bad |= !(val < INFINITY) || !(val > 1e-30f);
      atomicMax(&mb[le], __float_as_uint(val));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const int rest = n2 - k;
    if (lane < rest) {
      const uint32_t v = q2[k + lane];
      q2[lane] = v;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    n2 = rest;
    lm = __builtin_amdgcn_logf(__uint_as_float(mb[lane]));
  };
  /* second bound of the first k entries of queue 1; survivors to queue 2 */
  auto drain1 = [&](int k) {
    bool pass = false;
    const uint32_t e = lane < k ? q1[lane] : 0u;
    const int le = lane < k ? (int)(e & 63u) : lane;
    const float lmq = __shfl(lm, le, 64); /* every lane takes part */
    if (lane < k) {
      const int ql = (int)(e >> 6);
      float z[S];
      zload(P0 + b0 + ql, pbase + le, z);
      bool zero = false;
#pragma unroll
      for (int s = 0; s < S; ++s) zero |= z[s] == 0.f;
      pass = zero || !(pf_bound_sum<S>(pf_sorted<S>(z), cfs + ql * rowf, GS, lwc) + 0.00390625f < lmq);
    }
    const uint64_t bal = __ballot(pass);
    const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
    if (pass) q2[n2 + pos] = e;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    n2 += __popcll(bal);
    if (n2 >= 64) drain2(64);
    const int rest = n1 - k;
    if (lane < rest) {
      const uint32_t v = q1[k + lane];
      q1[lane] = v;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    n1 = rest;
  };

  for (b0 = 0; b0 < np; b0 += 64) {
    const int nq = (int)min((int64_t)64, np - b0);
    __syncthreads(); /* every wave is done with the previous block */
    {
      const float* src = a.coef + (P0 + b0) * (int64_t)rowf;
      for (int e = threadIdx.x; e < nq * rowf; e += 64 * PF_WG) cfs[e] = src[e];
    }
    __syncthreads();
    if (pbase >= a.nperm) continue; /* a wave past the last permutation only joins the barriers */
    float zn[S];
    zload(P0 + b0, valid ? p : pbase, zn);
    for (int ql = 0; ql < nq; ++ql) {
      float z[S];
#pragma unroll
      for (int s = 0; s < S; ++s) z[s] = zn[s];
      if (ql + 1 < nq) zload(P0 + b0 + ql + 1, valid ? p : pbase, zn);
      if (a.pflag[P0 + b0 + ql]) { /* every ABF of a non-estimable SNP is 1 */
        if (valid) atomicMax(&mb[lane], __float_as_uint(1.f));
        __builtin_amdgcn_s_waitcnt(0);
        lm = __builtin_amdgcn_logf(__uint_as_float(mb[lane]));
        continue;
      }
      bool zero = false;
#pragma unroll
      for (int s = 0; s < S; ++s) zero |= z[s] == 0.f;
      const float u = pf_bound_max<S>(pf_sorted<S>(z), cfs + ql * rowf, GS, lw) + lg_terms;
      /* a margin of 2^-8 for the FP32 rounding of the bound and the maximum */
      const bool cand = valid && (zero || !(u + 0.00390625f < lm));
      const uint64_t bal = __ballot(cand);
      if (bal) {
        const int pos =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
        if (cand) q1[n1 + pos] = ((uint32_t)ql << 6) | (uint32_t)lane;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        n1 += __popcll(bal);
        if (n1 >= 64) drain1(64);
      }
    }
    while (n1 > 0) drain1(min(n1, 64));
    while (n2 > 0) drain2(min(n2, 64));
  }
  if (bad) atomicOr(&a.fallback[bi], 1);
  if (valid) a.pmax[(int64_t)slot * a.NPP + p] = __uint_as_float(mb[lane]);
}

template <int MODE>
void launch_bma_s(int S, dim3 grid, hipStream_t st, const PfBmaArgs& a) {
  switch (S) {
#define PF_CASE(SS) \
  case SS: hipLaunchKernelGGL((pf_bma_kernel<SS, MODE>), grid, dim3(256), 0, st, a); break;
    PF_CASE(1) PF_CASE(2) PF_CASE(3) PF_CASE(4) PF_CASE(5) PF_CASE(6)
#undef PF_CASE
    default: break;
  }
}
template <>
void launch_bma_s<PF_JOIN_GEN>(int S, dim3 grid, hipStream_t st, const PfBmaArgs& a) {
  switch (S) {
#define PF_CASE(SS) \
  case SS: hipLaunchKernelGGL((pf_bma_kernel<SS, PF_JOIN_GEN>), grid, dim3(256), 0, st, a); break;
    PF_CASE(1) PF_CASE(2) PF_CASE(3) PF_CASE(4) PF_CASE(5) PF_CASE(6) PF_CASE(7) PF_CASE(8)
#undef PF_CASE
    default: break;
  }
}
template <>
void launch_bma_s<PF_JOIN_GENSIN>(int S, dim3 grid, hipStream_t st, const PfBmaArgs& a) {
  switch (S) {
#define PF_CASE(SS) \
  case SS: hipLaunchKernelGGL((pf_bma_kernel<SS, PF_JOIN_GENSIN>), grid, dim3(256), 0, st, a); break;
    PF_CASE(1) PF_CASE(2) PF_CASE(3) PF_CASE(4) PF_CASE(5) PF_CASE(6) PF_CASE(7) PF_CASE(8)
#undef PF_CASE
    default: break;
  }
}

void launch_bma_max(int S, dim3 grid, hipStream_t st, const PfBmaArgs& a) {
  const size_t lds = (size_t)64 * a.GS * (1 + 2 * S) * sizeof(float);
  switch (S) {
#define PF_CASE(SS) \
  case SS: hipLaunchKernelGGL((pf_bma_max_kernel<SS>), grid, dim3(64 * PF_WG), lds, st, a); break;
    PF_CASE(1) PF_CASE(2) PF_CASE(3) PF_CASE(4) PF_CASE(5) PF_CASE(6)
#undef PF_CASE
    default: break;
  }
}

double lg2(double x) { return std::log2(x); }

}  // namespace

/* Eligible genes of the chunk go through the contraction path; done[g] = 1
 * for each gene whose statistics it wrote.  The caller runs the rest (and
 * any gene marked for fallback) on the general path. */
int perm_fast_run(eqbf_ctx* ctx, const eqbf_chunk* cp, const eqbf_perm_opts* po, const int32_t* tables_dev,
                  int32_t n_tables, const int32_t* gene_table, int64_t nperm, double* stats_dev, hipStream_t st,
                  std::vector<uint8_t>& done) {
  const eqbf_chunk& c = *cp;
  const int S = c.n_subgroups, N = c.n_samples, G = c.n_genes;
  done.assign(G, 0);
  const bool join = po->analysis == EQBF_ANALYSIS_JOIN;
  if (po->precision != EQBF_PREC_MIXED || c.n_geno_sets != 1 || N > PF_NMAX || N < 4 || nperm <= 0) return EQBF_OK;
  int mode;
  if (!join)
    mode = PF_SEP;
  else
    mode = po->pbf == EQBF_PBF_ALL ? PF_JOIN_ALL : (po->pbf == EQBF_PBF_GEN ? PF_JOIN_GEN : PF_JOIN_GENSIN);
  if (S < 1 || S > (mode == PF_JOIN_ALL ? 6 : 8)) return EQBF_OK;
  if (!join && po->sep_sbgrp >= S) return EQBF_OK;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  if (join && ((mode != PF_JOIN_ALL && GL < 1) || (mode != PF_JOIN_GEN && GS < 1))) return EQBF_OK;
  {
    std::vector<int32_t> cc(S);
    EQBF_HIP(ctx, hipMemcpyAsync(cc.data(), c.covar_count, S * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s)
      if (cc[s] != 0) return EQBF_OK;
  }
  /* eligibility per gene */
  uint8_t* d_ok = (uint8_t*)ctx_buf(ctx, "pf_ok", (size_t)G + 64);
  if (!d_ok) return EQBF_ERR_HIP; /* ctx_buf has set the message */
  hipLaunchKernelGGL(pf_check_kernel, dim3(G), dim3(256), 0, st, c, d_ok);
  EQBF_HIP(ctx, hipGetLastError());
  std::vector<uint8_t> ok(G);
  std::vector<int64_t> off(G + 1);
  EQBF_HIP(ctx, hipMemcpyAsync(ok.data(), d_ok, G, hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipMemcpyAsync(off.data(), c.pair_offset, (G + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  std::vector<int> genes;
  for (int g = 0; g < G; ++g)
    if (gene_table[g] >= 0 && gene_table[g] < n_tables && ok[g] && off[g + 1] > off[g]) genes.push_back(g);
  if (genes.empty()) return EQBF_OK;

  const int nu = N - 2;
  int rc = ctx_ensure_tz(ctx, N, st);
  if (rc) return rc;
  float* d_zt = nullptr;
  if (join) {
    d_zt = (float*)ctx_buf(ctx, "pf_zt", (size_t)ZT_ROW * sizeof(float));
    if (!d_zt) return EQBF_ERR_HIP; /* ctx_buf has set the message */
    hipLaunchKernelGGL(pf_zeta_table_kernel, dim3((ZT_NODES + 255) / 256), dim3(256), 0, st, ctx_tz_table(ctx), nu,
                       d_zt);
    EQBF_HIP(ctx, hipGetLastError());
  }
  const int Npad = (N + 63) / 64 * 64, nks = Npad / 64;
  const int64_t NPP = (nperm + 63) / 64 * 64;
  const int npb16 = (int)((nperm + 15) / 16), npb64 = (int)((nperm + 63) / 64);
  const int ngrid = mode == PF_JOIN_ALL ? GS : (mode == PF_JOIN_GEN ? GL : GL + GS);
  const int NCF = 1 + 2 * S;
  /* batch: z values of at most 16 GB */
  size_t zeta_budget = (size_t)16 << 30;
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) zeta_budget = std::min(zeta_budget, fr / 4);
    zeta_budget = std::max(zeta_budget, (size_t)256 << 20);
  }
  const int s_lo = (!join && po->sep_sbgrp >= 0) ? po->sep_sbgrp : 0;
  const int s_hi = (!join && po->sep_sbgrp >= 0) ? po->sep_sbgrp + 1 : S;

  /* log2 weights of the terms (gene_snp_pair.cpp:552-602, utils_math.cpp:100-159) */
  PfBmaArgs ba;
  memset(&ba, 0, sizeof ba);
  for (int m = 0; m <= 8; ++m) {
    if (mode == PF_JOIN_ALL && m >= 1 && m <= S) {
      double ch = 1;
      for (int k = 0; k < m; ++k) ch = ch * (S - k) / (k + 1);
      ba.lw[1][m] = (float)lg2(1.0 / ((double)S * ch * GS));
      ba.lwc[m] = (float)lg2(ch);
    }
    if (mode == PF_JOIN_GEN) ba.lw[0][m] = (float)lg2(1.0 / GL);
    if (mode == PF_JOIN_GENSIN) {
      ba.lw[0][m] = (float)lg2(0.5 / GL);
      ba.lw[1][m] = (float)lg2(0.5 / ((double)S * GS));
    }
  }

  std::vector<int> fb_all;
  size_t gi = 0;
  while (gi < genes.size()) {
    /* gather the batch */
    std::vector<int32_t> bgene, btable, bsrc, bslot;
    std::vector<int64_t> bpoff(1, 0);
    size_t zbytes = 0;
    int nslots = 0;
    while (gi < genes.size()) {
      const int g = genes[gi];
      const int64_t np = off[g + 1] - off[g];
      const size_t zb = join ? (size_t)np * S * NPP * sizeof(float) : 0;
      if (!bgene.empty() && (zbytes + zb > zeta_budget || bgene.size() >= 4096)) break;
      zbytes += zb;
      bgene.push_back(g);
      btable.push_back(gene_table[g]);
      bslot.push_back(nslots);
      nslots += (int)((np + PF_SB - 1) / PF_SB);
      for (int64_t p = off[g]; p < off[g + 1]; ++p) bsrc.push_back((int32_t)p);
      bpoff.push_back(bpoff.back() + np);
      ++gi;
    }
    const int nb = (int)bgene.size();
    const int64_t nbp = bpoff.back();
    /* work lists */
    std::vector<int4> gitems;
    for (int b = 0; b < nb; ++b)
      for (int pb = 0; pb < npb16; ++pb)
        for (int s = s_lo; s < s_hi; ++s) gitems.push_back(make_int4(b, pb, s, 0));
    std::vector<int4> bitems;
    /* --pbf all --maxbf: one item per (gene, 64 permutations) over all the
     * gene's SNPs (pf_bma_max_kernel); otherwise blocks of PF_SB SNPs */
    const bool prune = join && mode == PF_JOIN_ALL && po->use_max_bf && S <= 6 &&
                       (size_t)64 * GS * (1 + 2 * S) * sizeof(float) <= 48 * 1024;
    if (join)
      for (int b = 0; b < nb; ++b) {
        if (prune) {
          for (int pb = 0; pb < npb64; pb += PF_WG) bitems.push_back(make_int4(b, pb, 0, bslot[b]));
          continue;
        }
        const int nsb = (int)((bpoff[b + 1] - bpoff[b] + PF_SB - 1) / PF_SB);
        for (int pb = 0; pb < npb64; ++pb)
          for (int sb = 0; sb < nsb; ++sb) bitems.push_back(make_int4(b, pb, sb, bslot[b] + sb));
      }
    /* device buffers */
    const size_t sz_meta = (size_t)nb * 3 * sizeof(int32_t) + (size_t)(nb + 1) * sizeof(int64_t) +
                           (size_t)nbp * sizeof(int32_t) + (gitems.size() + bitems.size()) * sizeof(int4) + 256;
    uint8_t* meta = (uint8_t*)ctx_buf(ctx, "pf_meta", sz_meta);
    int8_t* d_G8 = (int8_t*)ctx_buf(ctx, "pf_g8", (size_t)nbp * Npad);
    uint32_t* d_W = (uint32_t*)ctx_buf(ctx, "pf_w", (size_t)nb * S * Npad * sizeof(uint32_t));
    YStat* d_ys = (YStat*)ctx_buf(ctx, "pf_ys", (size_t)nb * S * sizeof(YStat));
    PStat* d_ps = (PStat*)ctx_buf(ctx, "pf_ps", (size_t)nbp * sizeof(PStat));
    uint8_t* d_pf = (uint8_t*)ctx_buf(ctx, "pf_pflag", (size_t)nbp + 64);
    int* d_fb = (int*)ctx_buf(ctx, "pf_fb", (size_t)nb * sizeof(int) + 64);
    if (!meta || !d_G8 || !d_W || !d_ys || !d_ps || !d_pf || !d_fb)
      return EQBF_ERR_HIP; /* ctx_buf has set the message */
    float* d_zeta = nullptr;
    float* d_coef = nullptr;
    float* d_pmax = nullptr;
    double* d_psum = nullptr;
    double* d_tmax = nullptr;
    if (join) {
      d_zeta = (float*)ctx_buf(ctx, "pf_zeta", zbytes + 64);
      d_coef = (float*)ctx_buf(ctx, "pf_coef", (size_t)nbp * ngrid * NCF * sizeof(float) + 64);
      d_pmax = (float*)ctx_buf(ctx, "pf_pmax", (size_t)nslots * NPP * sizeof(float) + 64);
      d_psum = (double*)ctx_buf(ctx, "pf_psum", (size_t)nslots * NPP * sizeof(double) + 64);
      if (!d_zeta || !d_coef || !d_pmax || !d_psum)
        return EQBF_ERR_HIP; /* ctx_buf has set the message */
    } else {
      d_tmax = (double*)ctx_buf(ctx, "pf_tmax", (size_t)nb * S * NPP * sizeof(double) + 64);
      if (!d_tmax) return EQBF_ERR_HIP; /* ctx_buf has set the message */
      /* subgroups that are not computed (--permsep 2) are never read */
    }
    int32_t* d_bgene = (int32_t*)meta;
    int32_t* d_btable = d_bgene + nb;
    int32_t* d_bslot = d_btable + nb;
    int64_t* d_bpoff = (int64_t*)(((uintptr_t)(d_bslot + nb) + 15) & ~(uintptr_t)15);
    int32_t* d_bsrc = (int32_t*)(d_bpoff + nb + 1);
    int4* d_gitems = (int4*)(((uintptr_t)(d_bsrc + nbp) + 15) & ~(uintptr_t)15);
    int4* d_bitems = d_gitems + gitems.size();
    EQBF_HIP(ctx, hipMemcpyAsync(d_bgene, bgene.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_btable, btable.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_bslot, bslot.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_bpoff, bpoff.data(), (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_bsrc, bsrc.data(), nbp * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(d_gitems, gitems.data(), gitems.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    if (!bitems.empty())
      EQBF_HIP(ctx, hipMemcpyAsync(d_bitems, bitems.data(), bitems.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemsetAsync(d_fb, 0, (size_t)nb * sizeof(int), st));
    {
      ProfScope ps(ctx, "pf_prep_kernel", st);
      hipLaunchKernelGGL(pf_prep_y_kernel, dim3(nb * S), dim3(256), 0, st, c, d_bgene, Npad, d_W, d_ys);
      hipLaunchKernelGGL(pf_prep_x_kernel, dim3((unsigned)((nbp + 63) / 64)), dim3(64), 0, st, c, d_bsrc, nbp, Npad,
                         d_G8, d_ps, d_pf);
    }
    EQBF_HIP(ctx, hipGetLastError());
    if (join) {
      const int64_t ne = nbp * ngrid;
      std::vector<double> gphi, goma;
      if (mode != PF_JOIN_ALL) {
        gphi.insert(gphi.end(), ctx->phiL.begin(), ctx->phiL.end());
        goma.insert(goma.end(), ctx->omaL.begin(), ctx->omaL.end());
      }
      if (mode != PF_JOIN_GEN) {
        gphi.insert(gphi.end(), ctx->phiS.begin(), ctx->phiS.end());
        goma.insert(goma.end(), ctx->omaS.begin(), ctx->omaS.end());
      }
      double* d_g = (double*)ctx_buf(ctx, "pf_grid", 2 * gphi.size() * sizeof(double) + 64);
      if (!d_g) return EQBF_ERR_HIP; /* ctx_buf has set the message */
      EQBF_HIP(ctx, hipMemcpyAsync(d_g, gphi.data(), gphi.size() * sizeof(double), hipMemcpyHostToDevice, st));
      EQBF_HIP(ctx, hipMemcpyAsync(d_g + gphi.size(), goma.data(), goma.size() * sizeof(double), hipMemcpyHostToDevice,
                                   st));
      ProfScope ps(ctx, "pf_coef_kernel", st);
      hipLaunchKernelGGL(pf_coef_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, d_ps, d_pf, nbp, d_g,
                         d_g + gphi.size(), ngrid, S, d_coef);
      EQBF_HIP(ctx, hipGetLastError());
    }
    PfGemmArgs ga;
    memset(&ga, 0, sizeof ga);
    ga.G8 = d_G8;
    ga.W = d_W;
    ga.ys = d_ys;
    ga.ps = d_ps;
    ga.pflag = d_pf;
    ga.items = d_gitems;
    ga.bpoff = d_bpoff;
    ga.tables = tables_dev;
    ga.btable = d_btable;
    ga.zt = d_zt;
    ga.tz = ctx_tz_table(ctx);
    ga.zeta = d_zeta;
    ga.tmax = d_tmax;
    ga.fallback = d_fb;
    ga.S = S;
    ga.N = N;
    ga.Npad = Npad;
    ga.nks = nks;
    ga.nu = nu;
    ga.nperm = nperm;
    ga.NPP = NPP;
    ga.n_items = (int)gitems.size();
    const size_t lds = (size_t)Npad * 4 + (size_t)nks * 4 * 64 * 16;
    {
      ProfScope ps(ctx, "pf_gemm_kernel", st);
      if (join)
        hipLaunchKernelGGL(pf_gemm_kernel<PF_JOIN_ALL>, dim3((unsigned)((gitems.size() + 7) / 8 * 8)), dim3(256), lds, st, ga);
      else
        hipLaunchKernelGGL(pf_gemm_kernel<PF_SEP>, dim3((unsigned)((gitems.size() + 7) / 8 * 8)), dim3(256), lds, st, ga);
    }
    EQBF_HIP(ctx, hipGetLastError());
    if (join) {
      ba.zeta = d_zeta;
      ba.coef = d_coef;
      ba.pflag = d_pf;
      ba.items = d_bitems;
      ba.bpoff = d_bpoff;
      ba.pmax = d_pmax;
      ba.psum = d_psum;
      ba.fallback = d_fb;
      ba.GL = GL;
      ba.GS = GS;
      ba.use_max = po->use_max_bf;
      ba.n_items = (int)bitems.size();
      ba.NPP = NPP;
      ba.nperm = nperm;
      const dim3 grid((unsigned)((bitems.size() + 3) / 4));
      ProfScope ps(ctx, "pf_bma_kernel", st);
      if (prune)
        launch_bma_max(S, dim3((unsigned)bitems.size()), st, ba);
      else if (mode == PF_JOIN_ALL)
        launch_bma_s<PF_JOIN_ALL>(S, grid, st, ba);
      else if (mode == PF_JOIN_GEN)
        launch_bma_s<PF_JOIN_GEN>(S, grid, st, ba);
      else
        launch_bma_s<PF_JOIN_GENSIN>(S, grid, st, ba);
      EQBF_HIP(ctx, hipGetLastError());
    }
    {
      PfFinalArgs fa;
      memset(&fa, 0, sizeof fa);
      fa.pmax = d_pmax;
      fa.psum = d_psum;
      fa.tmax = d_tmax;
      fa.bgene = d_bgene;
      fa.bpoff = d_bpoff;
      fa.bslot = d_bslot;
      fa.pflag = d_pf;
      fa.tz = ctx_tz_table(ctx);
      fa.stats = stats_dev;
      fa.nb = nb;
      fa.S = S;
      fa.nu = nu;
      fa.mode = mode;
      fa.use_max = po->use_max_bf;
      fa.sep_sbgrp = po->sep_sbgrp;
      fa.one_slot = prune ? 1 : 0;
      fa.nperm = nperm;
      fa.NPP = NPP;
      const int64_t ne = (int64_t)nb * nperm;
      ProfScope ps(ctx, "pf_final_kernel", st);
      hipLaunchKernelGGL(pf_final_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, fa);
      EQBF_HIP(ctx, hipGetLastError());
    }
    std::vector<int> fb(nb);
    EQBF_HIP(ctx, hipMemcpyAsync(fb.data(), d_fb, nb * sizeof(int), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b)
      if (!fb[b]) done[bgene[b]] = 1;
  }
  return EQBF_OK;
}
