/*
 * Mixed-precision uvlr association, batched path for one genotype set:
 * two kernels after the prep.
 *
 * uvlr_sums_kernel -- the sums of every (pair, subgroup) as a dense FP64
 *   matrix product on the matrix cores.  For the 64 pairs (SNPs) of a tile
 *   and the N samples of the chunk:
 *     C_x[snp][col] = sum_i x[i][snp] * B[i][col]   col = (s, k): y~_s, U_s,k
 *     C_m[snp][s]   = sum_i x[i][snp]^2 * m_s[i]    m_s = membership of s
 *   B holds zeros outside the members (the prep writes only members' rows
 *   into a zeroed buffer), so the products over all N samples are the sums
 *   over each subgroup's members.  v_mfma_f64_16x16x4f64: A = 16 SNPs x 4
 *   samples, B = 4 samples x 16 columns, C = 16 x 16 in four FP64 per lane
 *   (lane l: A[l % 16][l / 16], B[l / 16][l % 16], D[l / 16 + 4 r][l % 16]).
 *   Eight waves: wave w owns SNP rows 16*(w&3).. and half of the column
 *   tiles.  Samples are staged through LDS in chunks (double buffered, the
 *   next chunk prefetched into registers).  The epilogue moves the sums of a
 *   range of subgroups through LDS and computes the summary statistics
 *   (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209, and the
 *   standardisation of gene_snp_pair.cpp:256-290) one (pair, subgroup) per
 *   thread, writing the outputs and the standardised values the Bayes
 *   factors need.
 *
 * uvlr_abf_kernel -- the Bayes factors (CalcAbfsUvlrFor*,
 *   gene_snp_pair.cpp:364-550) for 64 pairs per workgroup, one lane per
 *   pair, FP32 terms and FP64 grid sums.  The consistent configuration is
 *   evaluated once per distinct phi2 of the gen / gen-fix / gen-maxh grids:
 *   the single-subgroup terms and the sums sum b/(v+phi2), sum 1/(v+phi2)
 *   depend on phi2 only, and every grid point with that phi2 is finished from
 *   them with its own omega2.  Singletons use the closed form; other
 *   configurations sweep their active subgroups per grid point.
 */
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "eqbf_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NT_SUMS = 512; /* threads per sums workgroup (8 waves) */
constexpr int BPRE = 8;      /* B elements prefetched per thread      */
constexpr float kHalfInvLn10f = 0.21714724095162590f;

struct SumsParams {
  const PrepRec* rec;
  const double* B;      /* [G][N][CX], CX = S * KC                      */
  const uint64_t* memb; /* [G][N]                                       */
  const int4* tiles;
  const int* n_tiles;
  const double* tz;
  int32_t* n_out;
  double* sstats;
  double* stdsstats;
  float* st_v;   /* [S][NP] standardised variance s'^2 (join only)      */
  float* st_b;   /* [S][NP] standardised effect b'                      */
  float* st_z2;  /* [S][NP] z^2                                         */
  uint8_t* st_f; /* [S][NP] role of the subgroup (F_*)                  */
  double* sx2;   /* [2][S][NP] x~'x~ (NaN: not estimable) and x'y~, when the
                    statistics are a kernel of their own                  */
  const int32_t* idx; /* [G][S][N] members of (g, s), for the FP64 sums  */
  uint32_t* redo;     /* [S * NP bits] items sums_repair_kernel recomputes */
  int KU, KC, CX, NXT, KCH, SPP;
};

__device__ inline double geno_at(const eqbf_chunk& c, int64_t i, int64_t j) {
  const int64_t off = i * c.geno_ld + j;
  if (c.geno_dtype == EQBF_F32) return (double)((const float*)c.genotypes)[off];
  return ((const double*)c.genotypes)[off];
}

/* Epilogue of one tile: the sums of SPP subgroups at a time (one pass when
 * MULTI is false) move through LDS, then one (pair, subgroup) per thread
 * gets its summary statistics (FitSingleGeneWithSingleSnp,
 * utils_math.cpp:166-209, and the standardisation of gene_snp_pair.cpp:
 * 256-290) and the Bayes-factor inputs. */
template <int NXW, int NMW, bool MULTI, bool SPLIT>
__device__ __forceinline__ void sums_epilogue(const eqbf_chunk& c, const SumsParams& sp, double* E, const d4 (&ax)[NXW],
                                              const d4 (&am)[NMW], int g, int64_t p0, int cnt) {
  const int S = c.n_subgroups, KC = sp.KC, CX = sp.CX;
  const int64_t NP = c.n_pairs;
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, kq = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rb = w & 3, half = w >> 2;
  const int tx0 = half * NXW, tm0 = half * NMW;
  /* lane holds D[(lane >> 4) + 4 r][lane & 15] */
  for (int sa = 0; sa < S; sa += MULTI ? sp.SPP : S) {
    const int sb = MULTI ? min(S, sa + sp.SPP) : S, ns = sb - sa, NL = ns * KC, ecols = NL + ns;
#pragma unroll
    for (int k = 0; k < NXW; ++k) {
      const int gc = (tx0 + k) * 16 + l16;
      const int s = gc / KC;
      if (gc < CX && s >= sa && s < sb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) E[(rb * 16 + kq + 4 * r) * ecols + (s - sa) * KC + (gc - s * KC)] = ax[k][r];
      }
    }
#pragma unroll
    for (int k = 0; k < NMW; ++k) {
      const int s = (tm0 + k) * 16 + l16;
      if (s >= sa && s < sb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) E[(rb * 16 + kq + 4 * r) * ecols + NL + (s - sa)] = am[k][r];
      }
    }
    __syncthreads();
    for (int it = tid; it < 64 * ns; it += NT_SUMS) {
      const int j = it & 63, s = sa + (it >> 6);
      if (j >= cnt) continue;
      const double* e = E + j * ecols + (s - sa) * KC;
      double proj = 0;
      for (int q = 0; q < sp.KU; ++q) proj += e[1 + q] * e[1 + q];
      const int64_t p = p0 + j, o = (int64_t)s * NP + p;
      if (SPLIT) {
        sp.sx2[o] = uvlr_xx(E[j * ecols + NL + (s - sa)], proj);
        sp.sx2[(int64_t)S * NP + o] = e[0];
        /* a NaN genotype somewhere in the SNP's column reaches every
         * subgroup through the zero rows of non-members */
        if (isnan(e[0])) atomicOr(&sp.redo[o >> 5], 1u << (o & 31));
        continue;
      }
      const int64_t snp = c.pair_snp[p];
      const PrepRec r = sp.rec[(int64_t)g * S + s];
      const bool has = r.n > 0 && c.snp_in_sbgrp[snp * S + s];
      double o5[5], st3[3];
      const int f = uvlr_stats(r, has, E[j * ecols + NL + (s - sa)], e[0], proj, sp.tz, o5, st3);
      sp.n_out[o] = has ? r.n : 0;
      if (sp.sstats)
        for (int k = 0; k < 5; ++k) sp.sstats[(int64_t)k * S * NP + o] = o5[k];
      if (sp.stdsstats)
        for (int k = 0; k < 3; ++k) sp.stdsstats[(int64_t)k * S * NP + o] = st3[k];
      if (sp.st_f) {
        sp.st_v[o] = (float)(st3[1] * st3[1]);
        sp.st_b[o] = (float)st3[0];
        sp.st_z2[o] = (float)(st3[2] * st3[2]);
        sp.st_f[o] = (uint8_t)f;
      }
    }
    __syncthreads();
  }
}

/* Each wave owns NXW x-tiles (columns y~, U) and NMW m-tiles (membership)
 * of the tile's 64 x (columns) product; the column tiles of the two halves
 * of the workgroup are padded to those counts (surplus tiles read padding
 * and are discarded), so the product loop has no branches. */
template <int NXW, int NMW, bool MULTI>
__global__ void __launch_bounds__(NT_SUMS) uvlr_sums_kernel(const eqbf_chunk c, const SumsParams sp) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int CXPP = 2 * NXW * 16; /* staged B row length */
  const int S = c.n_subgroups, N = c.n_samples;
  const int64_t NP = c.n_pairs;
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, kq = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rb = w & 3, half = w >> 2;
  const int KCH = sp.KCH, CX = sp.CX, KC = sp.KC;
  const int stage_d = KCH * 64 + KCH * CXPP + KCH; /* doubles per buffer */
  double* const buf = (double*)smem;
  const int nchunk = (N + KCH - 1) / KCH;
  const int belem = KCH * CX;
  const int tx0 = half * NXW, tm0 = half * NMW;

  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const int T = *sp.n_tiles;
  const int span = (T + 7) / 8;
  const int t0 = xcd * span, t1 = min(T, t0 + span);
  for (int t = t0 + slot; t < t1; t += per_xcd) {
    const int4 tile = sp.tiles[t];
    const int g = tile.x;
    const int64_t p0 = tile.y;
    const int cnt = tile.z;
    /* thread tid stages x of SNP column xj for samples (tid >> 6) + 8u */
    const int xj = tid & 63;
    const int64_t xsnp = c.pair_snp[p0 + min(xj, cnt - 1)];
    const double* Bg = sp.B + (int64_t)g * N * CX;
    const uint64_t* mg = sp.memb + (int64_t)g * N;
    d4 ax[NXW], am[NMW];
#pragma unroll
    for (int k = 0; k < NXW; ++k) ax[k] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NMW; ++k) am[k] = d4{0.0, 0.0, 0.0, 0.0};

    {
      double* xs = buf;
      double* bs = buf + KCH * 64;
      uint64_t* mb = (uint64_t*)(bs + KCH * CXPP);
      for (int kk = tid >> 6; kk < KCH; kk += NT_SUMS / 64) xs[kk * 64 + xj] = (kk < N) ? geno_at(c, kk, xsnp) : 0.0;
      for (int e = tid; e < belem; e += NT_SUMS) {
        const int r = e / CX;
        bs[r * CXPP + (e - r * CX)] = (r < N) ? Bg[e] : 0.0;
      }
      if (tid < KCH) mb[tid] = (tid < N) ? mg[tid] : 0ull;
    }
    __syncthreads();
    /* per-thread walk of the B chunk: element e = tid + u * NT_SUMS */
    const int q512 = NT_SUMS / CX, r512 = NT_SUMS - q512 * CX;
    const int r_first = tid / CX, c_first = tid - r_first * CX;
    for (int ck = 0; ck < nchunk; ++ck) {
      const double* cur = buf + (ck & 1) * stage_d;
      double* nxt = buf + ((ck + 1) & 1) * stage_d;
      const bool more = ck + 1 < nchunk;
      const int i1 = (ck + 1) * KCH;
      double bpre[BPRE], xpre[2];
      uint64_t mpre = 0;
      if (more) {
        const double* src = Bg + (int64_t)i1 * CX;
        int r = r_first, cc = c_first;
#pragma unroll
        for (int u = 0; u < BPRE; ++u) {
          bpre[u] = (r < KCH && i1 + r < N) ? src[r * CX + cc] : 0.0;
          r += q512;
          cc += r512;
          if (cc >= CX) {
            cc -= CX;
            ++r;
          }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int kk = (tid >> 6) + u * (NT_SUMS / 64);
          xpre[u] = (kk < KCH && i1 + kk < N) ? geno_at(c, i1 + kk, xsnp) : 0.0;
        }
        if (tid < KCH) mpre = (i1 + tid < N) ? mg[i1 + tid] : 0ull;
      }
      /* products of the current chunk: four samples per step */
      const double* xs = cur;
      const double* bs = cur + KCH * 64;
      const uint64_t* mb = (const uint64_t*)(bs + KCH * CXPP);
      for (int kk = 0; kk < KCH; kk += 4) {
        const int krow = kk + kq;
        const double a = xs[krow * 64 + rb * 16 + l16];
        const double a2 = a * a;
        const uint64_t mbits = mb[krow];
        double bx[NXW];
#pragma unroll
        for (int k = 0; k < NXW; ++k) bx[k] = bs[krow * CXPP + (tx0 + k) * 16 + l16];
#pragma unroll
        for (int k = 0; k < NXW; ++k) ax[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bx[k], ax[k], 0, 0, 0);
#pragma unroll
        for (int k = 0; k < NMW; ++k) {
          const int s = (tm0 + k) * 16 + l16;
          const double bm = (s < S && ((mbits >> (s & 63)) & 1ull)) ? 1.0 : 0.0;
          am[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bm, am[k], 0, 0, 0);
        }
      }
      __syncthreads();
      if (more) {
        double* xs2 = nxt;
        double* bs2 = nxt + KCH * 64;
        uint64_t* mb2 = (uint64_t*)(bs2 + KCH * CXPP);
        int r = r_first, cc = c_first;
#pragma unroll
        for (int u = 0; u < BPRE; ++u) {
          if (r < KCH) bs2[r * CXPP + cc] = bpre[u];
          r += q512;
          cc += r512;
          if (cc >= CX) {
            cc -= CX;
            ++r;
          }
        }
        for (int e = tid + BPRE * NT_SUMS; e < belem; e += NT_SUMS) {
          const int rr = e / CX;
          bs2[rr * CXPP + (e - rr * CX)] = (i1 + rr < N) ? Bg[(int64_t)i1 * CX + e] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int kk = (tid >> 6) + u * (NT_SUMS / 64);
          if (kk < KCH) xs2[kk * 64 + xj] = xpre[u];
        }
        for (int kk = (tid >> 6) + 2 * (NT_SUMS / 64); kk < KCH; kk += NT_SUMS / 64)
          xs2[kk * 64 + xj] = (i1 + kk < N) ? geno_at(c, i1 + kk, xsnp) : 0.0;
        if (tid < KCH) mb2[tid] = mpre;
      }
      __syncthreads();
    }

    sums_epilogue<NXW, NMW, MULTI, true>(c, sp, buf, ax, am, g, p0, cnt);
  }
}

/* The same product with LDS-direct staging (global_load_lds): a chunk of
 * KCH samples is copied from HBM straight into LDS by the load unit
 * (B rows as 16-byte units of a contiguous block, genotypes gathered one
 * sample row per wave instruction, the membership words) while the waves
 * multiply the previous chunk; one barrier per chunk, no staging registers.
 * Needs CX even (16-byte units never straddle the gene's block).  Rows past
 * N in the last chunk are zero-filled so the padded products add nothing. */
typedef __attribute__((address_space(3))) void lds_void_t;

template <int NXW, int NMW, bool MULTI, bool XF64>
__global__ void __launch_bounds__(NT_SUMS) __attribute__((amdgpu_waves_per_eu(4)))
uvlr_sums_dma_kernel(const eqbf_chunk c, const SumsParams sp) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename std::conditional<XF64, double, float>::type xt;
  const int N = c.n_samples;
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, kq = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rb = w & 3, half = w >> 2;
  const int KCH = sp.KCH, CX = sp.CX, S = c.n_subgroups;
  const int bbytes = KCH * CX * 8, xbytes = KCH * 64 * (int)sizeof(xt);
  const int stage = bbytes + xbytes + KCH * 8;
  const int nchunk = (N + KCH - 1) / KCH;
  const int tx0 = half * NXW, tm0 = half * NMW;

  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const int T = *sp.n_tiles;
  const int span = (T + 7) / 8;
  const int t0 = xcd * span, t1 = min(T, t0 + span);
  for (int t = t0 + slot; t < t1; t += per_xcd) {
    const int4 tile = sp.tiles[t];
    const int g = tile.x;
    const int64_t p0 = tile.y;
    const int cnt = tile.z;
    /* genotype column of this lane: SNP lane (FP32), or SNP lane / 2 + 32 h
     * and 4-byte half lane & 1 (FP64, two instructions per sample row) */
    const int64_t xs0 = c.pair_snp[p0 + min(XF64 ? (lane >> 1) : lane, cnt - 1)];
    const int64_t xs1 = XF64 ? c.pair_snp[p0 + min(32 + (lane >> 1), cnt - 1)] : 0;
    const char* Bg = (const char*)(sp.B + (int64_t)g * N * CX);
    const char* mg = (const char*)(sp.memb + (int64_t)g * N);
    d4 ax[NXW], am[NMW];
#pragma unroll
    for (int k = 0; k < NXW; ++k) ax[k] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NMW; ++k) am[k] = d4{0.0, 0.0, 0.0, 0.0};

    auto issue = [&](int ck, unsigned char* st) {
      const int i1 = ck * KCH, rows = min(KCH, N - i1);
      const int n16 = rows * CX / 2;
      const char* src = Bg + (int64_t)i1 * CX * 8;
      for (int u0 = w * 64; u0 < n16; u0 += NT_SUMS)
        if (u0 + lane < n16)
          __builtin_amdgcn_global_load_lds((const void*)(src + (int64_t)(u0 + lane) * 16), (lds_void_t*)(st + u0 * 16),
                                           16, 0, 0);
      unsigned char* xs = st + bbytes;
      for (int r = w; r < rows; r += NT_SUMS / 64) {
        const int64_t row = (int64_t)(i1 + r) * c.geno_ld;
        if (!XF64) {
          __builtin_amdgcn_global_load_lds((const void*)((const float*)c.genotypes + row + xs0),
                                           (lds_void_t*)(xs + r * 256), 4, 0, 0);
        } else {
          const char* gp = (const char*)((const double*)c.genotypes + row);
          __builtin_amdgcn_global_load_lds((const void*)(gp + xs0 * 8 + (lane & 1) * 4),
                                           (lds_void_t*)(xs + r * 512), 4, 0, 0);
          __builtin_amdgcn_global_load_lds((const void*)(gp + xs1 * 8 + (lane & 1) * 4),
                                           (lds_void_t*)(xs + r * 512 + 256), 4, 0, 0);
        }
      }
      if (w == NT_SUMS / 64 - 1 && lane < 2 * rows)
        __builtin_amdgcn_global_load_lds((const void*)(mg + (int64_t)i1 * 8 + lane * 4),
                                         (lds_void_t*)(st + bbytes + xbytes), 4, 0, 0);
      if (rows < KCH) { /* last chunk: zero rows past N */
        double* bz = (double*)st;
        for (int e = rows * CX + tid; e < KCH * CX; e += NT_SUMS) bz[e] = 0.0;
        xt* xz = (xt*)(st + bbytes);
        for (int e = rows * 64 + tid; e < KCH * 64; e += NT_SUMS) xz[e] = (xt)0;
        uint64_t* mz = (uint64_t*)(st + bbytes + xbytes);
        for (int e = rows + tid; e < KCH; e += NT_SUMS) mz[e] = 0ull;
      }
    };

    issue(0, smem);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int ck = 0; ck < nchunk; ++ck) {
      const unsigned char* cur = smem + (ck & 1) * stage;
      if (ck + 1 < nchunk) issue(ck + 1, smem + ((ck + 1) & 1) * stage);
      const double* bs = (const double*)cur;
      const xt* xs = (const xt*)(cur + bbytes);
      const uint64_t* mb = (const uint64_t*)(cur + bbytes + xbytes);
      for (int kk = 0; kk < KCH; kk += 4) {
        const int krow = kk + kq;
        const double a = (double)xs[krow * 64 + rb * 16 + l16];
        const double a2 = a * a;
        const uint64_t mbits = mb[krow];
        double bx[NXW];
#pragma unroll
        for (int k = 0; k < NXW; ++k) bx[k] = bs[krow * CX + (tx0 + k) * 16 + l16];
#pragma unroll
        for (int k = 0; k < NXW; ++k) ax[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bx[k], ax[k], 0, 0, 0);
#pragma unroll
        for (int k = 0; k < NMW; ++k) {
          const int s = (tm0 + k) * 16 + l16;
          const double bm = (s < S && ((mbits >> (s & 63)) & 1ull)) ? 1.0 : 0.0;
          am[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bm, am[k], 0, 0, 0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    sums_epilogue<NXW, NMW, MULTI, true>(c, sp, (double*)smem, ax, am, g, p0, cnt);
  }
}

/* Summary statistics of every (pair, subgroup) of a tile from x~'x~ and
 * x'y~ (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209, and
 * StandardizeSstatsAndCorrectSmallSampleSize, gene_snp_pair.cpp:256-290),
 * with the outputs and the Bayes-factor inputs; one workgroup per tile. */
constexpr int NT_STATS = 256;

/* AS241 rational approximations evaluated in FP32 (the coefficients of
 * dm::as241_central / dm::as241_tail): relative error ~1e-7. */
__device__ inline float as241_central_f(float q) {
  const float r = 0.180625f - q * q;
  return q * (((((((r * 2509.0809287301226727f + 33430.575583588128105f) * r + 67265.770927008700853f) * r +
                  45921.953931549871457f) * r + 13731.693765509461125f) * r + 1971.5909503065514427f) * r +
               133.14166789178437745f) * r + 3.387132872796366608f) /
         (((((((r * 5226.495278852545561f + 28729.085735721942674f) * r + 39307.89580009271061f) * r +
              21213.794301586595867f) * r + 5394.1960214247511077f) * r + 687.1870074920579083f) * r +
           42.313330701600911252f) * r + 1.0f);
}
__device__ inline float as241_tail_f(float r) {
  if (r <= 5.0f) {
    r -= 1.6f;
    return (((((((r * 7.7454501427834140764e-4f + 0.0227238449892691845833f) * r + 0.24178072517745061177f) * r +
                1.27045825245236838258f) * r + 3.64784832476320460504f) * r + 5.7694972214606914055f) * r +
             4.6303378461565452959f) * r + 1.42343711074968357734f) /
           (((((((r * 1.05075007164441684324e-9f + 5.475938084995344946e-4f) * r + 0.0151986665636164571966f) * r +
                0.14810397642748007459f) * r + 0.68976733498510000455f) * r + 1.6763848301838038494f) * r +
             2.05319162663775882187f) * r + 1.0f);
  }
  r -= 5.0f;
  return (((((((r * 2.01033439929228813265e-7f + 2.71155556874348757815e-5f) * r + 0.0012426609473880784386f) * r +
              0.026532189526576123093f) * r + 0.29656057182850489123f) * r + 1.7848265399172913358f) * r +
           5.4637849111641143699f) * r + 6.6579046435011037772f) /
         (((((((r * 2.04426310338993978564e-15f + 1.4215117583164458887e-7f) * r + 1.8463183175100546818e-5f) * r +
              7.868691311456132591e-4f) * r + 0.0148753612908506148525f) * r + 0.13692988092273580531f) * r +
           0.59983220655588793769f) * r + 1.0f);
}

/* The FP32 Bayes-factor inputs without the FP64 outputs: beta and the
 * residual sum of squares in FP64, then (exact identities of the reference's
 * expressions, gene_snp_pair.cpp:256-290) s'^2 = se^2 / sigma^2 =
 * (n - rank) / ((n - P) x~'x~), b' = sign(beta) |z| s', and z from the t tail
 * table with the normal quantile in FP32.  Returns false where the FP64
 * path must decide: residual sum of squares <= 0, |t| < 1e-3 (the |z| > 1e-8
 * test), or a tail past the table. */
__device__ inline bool uvlr_bf_inputs_fast(const PrepRec& r, double xx, double sxy, const double* __restrict__ tz,
                                           float& v, float& b, float& z2) {
  const int n = r.n, P = r.q + 2, rank = r.rank + 1;
  const double beta = sxy / xx;
  const double rss = r.yy - sxy * beta;
  if (!(rss > 0.0)) return false;
  const float at = fabsf((float)beta) * sqrtf((float)xx * (float)(n - P) / (float)rss);
  if (!(at >= 1e-3f)) return false;
  const double w = dm::tz_eval(tz, n - P, (double)at);
  if (!(w < INFINITY)) return false;
  const float wf = (float)w;
  const float az = wf < 2.5902671654458267f ? -as241_central_f(__expf(-wf) - 0.5f) : as241_tail_f(sqrtf(wf));
  v = (float)((double)(n - rank) / ((double)(n - P) * xx));
  b = copysignf(az * sqrtf(v), (float)beta);
  z2 = az * az;
  return true;
}

/* x~'x~ and x'y~ in FP64 of the items the sums kernel marked, one wave per
 * item: the uncentred sums over the members, as the FP64 product forms
 * them (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209).  A genotype
 * that is NaN for a member makes them NaN; one that is NaN for a
 * non-member is not read (GeneSnpPair::FillStlContainers reads the members'
 * genotypes only, gene_snp_pair.cpp:79-170). */
constexpr int NT_REPAIR = 256;
__global__ void redo_zero_kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0u;
}
__device__ inline double repair_wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__global__ void __launch_bounds__(NT_REPAIR) sums_repair_kernel(const eqbf_chunk c, const SumsParams sp) {
  const int N = c.n_samples, S = c.n_subgroups, G = c.n_genes;
  const int64_t NP = c.n_pairs, W = ((int64_t)S * NP + 31) / 32;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (NT_REPAIR / 64) + (threadIdx.x >> 6);
  const int64_t n_waves = (int64_t)gridDim.x * (NT_REPAIR / 64);
  for (int64_t word = wave; word < W; word += n_waves) {
    uint32_t bits = sp.redo[word];
    while (bits) {
      const int b = __builtin_ctz(bits);
      bits &= bits - 1u;
      const int64_t o = word * 32 + b;
      const int s = (int)(o / NP);
      const int64_t p = o - (int64_t)s * NP;
      int lo = 0, hi = G - 1; /* gene of pair p: pair_offset[g] <= p < pair_offset[g + 1] */
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c.pair_offset[mid] <= p)
          lo = mid;
        else
          hi = mid - 1;
      }
      const int g = lo;
      const int64_t snp = c.pair_snp[p];
      const PrepRec r = sp.rec[(int64_t)g * S + s];
      if (r.n == 0 || (r.flags & PREP_TOO_FEW) || !c.snp_in_sbgrp[snp * S + s]) continue; /* no statistics */
      const int n = r.n;
      const int32_t* idx = sp.idx + ((int64_t)g * S + s) * N;
      double sxx = 0.0, sy = 0.0, u[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) u[k] = 0.0;
      for (int m = lane; m < n; m += 64) {
        const int i = idx[m];
        const double x = geno_at(c, i, snp);
        const double* bb = sp.B + ((int64_t)g * N + i) * sp.CX + (int64_t)s * sp.KC;
        sxx += x * x;
        sy += x * bb[0];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < sp.KU) u[k] += x * bb[1 + k];
      }
      sxx = repair_wave_sum(sxx);
      sy = repair_wave_sum(sy);
      double proj = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < sp.KU) {
          const double t = repair_wave_sum(u[k]);
          proj += t * t;
        }
      if (lane == 0) {
        sp.sx2[o] = uvlr_xx(sxx, proj);
        sp.sx2[(int64_t)S * NP + o] = sy;
      }
    }
  }
}

/* FULL: the FP64 summary statistics are written too (sstats or stdsstats
 * requested), so every item takes the FP64 path. */
template <bool FULL>
__global__ void __launch_bounds__(NT_STATS) __attribute__((amdgpu_waves_per_eu(4, 8)))
uvlr_stats_kernel(const eqbf_chunk c, const SumsParams sp) {
  /* XCD-aware order (blockIdx & 7 = XCD): contiguous tile ranges per XCD */
  const int T = *sp.n_tiles, per = (T + 7) / 8;
  const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (t >= T) return;
  const int4 tile = sp.tiles[t];
  const int g = tile.x, cnt = tile.z;
  const int64_t p0 = tile.y;
  const int S = c.n_subgroups;
  const int64_t NP = c.n_pairs;
  for (int it = threadIdx.x; it < 64 * S; it += NT_STATS) {
    const int j = it & 63, s = it >> 6;
    if (j >= cnt) continue;
    const int64_t p = p0 + j, o = (int64_t)s * NP + p;
    const int64_t snp = c.pair_snp[p];
    const PrepRec r = sp.rec[(int64_t)g * S + s];
    const bool has = r.n > 0 && c.snp_in_sbgrp[snp * S + s];
    const double xx = sp.sx2[o], sxy = sp.sx2[(int64_t)S * NP + o];
    sp.n_out[o] = has ? r.n : 0;
    if (!FULL && sp.st_f && has && !(r.flags & PREP_TOO_FEW) && !isnan(xx)) {
      float v, b, z2;
      if (uvlr_bf_inputs_fast(r, xx, sxy, sp.tz, v, b, z2)) {
        sp.st_v[o] = v;
        sp.st_b[o] = b;
        sp.st_z2[o] = z2;
        sp.st_f[o] = (uint8_t)F_NORMAL;
        continue;
      }
    }
    double o5[5], st3[3];
    const int f = uvlr_stats_xx(r, has, xx, sxy, sp.tz, o5, st3);
    if (FULL && sp.sstats)
      for (int k = 0; k < 5; ++k) sp.sstats[(int64_t)k * S * NP + o] = o5[k];
    if (FULL && sp.stdsstats)
      for (int k = 0; k < 3; ++k) sp.stdsstats[(int64_t)k * S * NP + o] = st3[k];
    if (sp.st_f) {
      sp.st_v[o] = (float)(st3[1] * st3[1]);
      sp.st_b[o] = (float)st3[0];
      sp.st_z2[o] = (float)(st3[2] * st3[2]);
      sp.st_f[o] = (uint8_t)f;
    }
  }
}

/* ---- Bayes factors ------------------------------------------------------------ */
struct PhiJob {
  float phi2;
  int e0, e1; /* entries [e0, e1) */
};
struct PhiEntry {
  int type, gi;
  float oma2;
};

constexpr int ABF_GMAX = 256; /* small-grid points of the batched Bayes factors */
struct AbfParams {
  const float* st_v;
  const float* st_b;
  const float* st_z2;
  const uint8_t* st_f;
  const PhiJob* jobs;
  const PhiEntry* ents;
  const uint64_t* masks;
  const double* grid; /* phiL, omaL, phiS, omaS */
  void* raw;
  double* w;
  int n_jobs, n_cfg, GL, GS, raw_f64, raw_tiled;
  int64_t n_raw;
  /* small grid as FP32 in the kernel arguments (read through the scalar
   * cache): phi2 at [0, GS), omega2 at [ABF_GMAX, ABF_GMAX + GS),
   * phi2 + omega2 at [2 ABF_GMAX, 2 ABF_GMAX + GS) */
  float gS[3 * ABF_GMAX];
};
static_assert(sizeof(AbfParams) + 16 <= 4096, "kernel arguments of uvlr_abf_kernel exceed 4 KiB");

__device__ inline float lg10f(float x) { return __log2f(x) * 0.30102999566398120f; }
/* 2^x on the transcendental unit, without the library's rescaling of
 * results below 2^-126: such terms are invisible in the log-sums here */
__device__ inline float ex2f(float x) { return __builtin_amdgcn_exp2f(x); }

/* Finishes one ABF from the single-subgroup sum L and A = sum b/(v+phi2),
 * Bs = sum 1/(v+phi2) (CalcLog10AbfUvlr, gene_snp_pair.cpp:336-354): 0
 * unless Bs > 0, b-bar != 0 and its variance V is finite; the average's
 * term only when b-bar^2 / V != 0.  Evaluated without branches (the terms
 * of the cases that return 0 are computed and dropped). */
__device__ inline float abf_fin(float L, float A, float Bs, float oma2) {
  const float V = __builtin_amdgcn_rcpf(Bs), bbar = A * V;
  const float T2 = A * bbar; /* bbar^2 / V */
  const float rVo = __builtin_amdgcn_rcpf(V + oma2);
  const float lbar = 0.5f * lg10f(V * rVo) + kHalfInvLn10f * T2 * oma2 * rVo;
  const bool ok = Bs > 0.f && bbar != 0.f && V < INFINITY;
  return ok ? (T2 != 0.f ? lbar : 0.f) + L : 0.f;
}

/* log10 of the mean of 10^v over a grid: running maximum m and the sum of
 * 10^(v - m) rescaled when the maximum moves (FP32, branch-free). */
constexpr float kLog2_10 = 3.32192809488736235f;
struct Acc {
  float m, s;
  __device__ void init() {
    m = -INFINITY;
    s = 0.f;
  }
  __device__ void add(float v) {
    const bool up = v > m;
    const float e = ex2f(-fabsf(v - m) * kLog2_10);
    s = up ? fmaf(s, e, 1.f) : s + e;
    m = up ? v : m;
  }
  /* merges another accumulator into this one */
  __device__ void merge(const Acc& o) {
    const float M = fmaxf(m, o.m);
    if (M == -INFINITY) return;
    s = s * ex2f((m - M) * kLog2_10) + o.s * ex2f((o.m - M) * kLog2_10);
    m = M;
  }
};

__device__ inline double acc_finish(float m, double s, double weight) {
  double res = (double)m + log10(s * weight);
  if (fabs(res) <= 2.220446049250313e-16) res = 0.0;
  return res;
}

/* Where one lane stores its pair's raw ABFs: row r lies at base + r * stride
 * bytes (row-major: stride NP elements; 64-pair tiles: stride the tile's
 * width, base at the tile's start plus the pair's column).  One 32 x 32 ->
 * 64-bit multiply-add per store. */
template <bool RAW64>
struct RawRow {
  char* base;
  uint32_t stride;
  __device__ RawRow(const AbfParams& ap, int64_t p, int64_t NP) {
    const int64_t es = RAW64 ? 8 : 4;
    if (!RAW64 && ap.raw_tiled) {
      const int64_t tb = p & ~(int64_t)63, tw = NP - tb < 64 ? NP - tb : 64;
      base = (char*)ap.raw + (tb * ap.n_raw + (p & 63)) * es;
      stride = (uint32_t)(tw * es);
    } else {
      base = (char*)ap.raw + p * es;
      stride = (uint32_t)(NP * es);
    }
  }
  __device__ void put(int row, float l) const {
    char* a = base + (uint64_t)(uint32_t)row * stride;
    /* streaming stores: the raw ABFs are not read again by this launch */
    if (RAW64)
      __builtin_nontemporal_store((double)l, (double*)a);
    else
      __builtin_nontemporal_store(l, (float*)a);
  }
};

constexpr int NT_ABF = 512;

/* Bayes factors of 64 pairs, one lane per pair (eight waves share the jobs).
 * Per (subgroup, pair) the LDS holds {v, z^2 / (2 ln 10)} and b for a
 * subgroup that enters the Bayes factors, and {2^100, 0}, 0 otherwise, so
 * the sums over subgroups run without branches: such a subgroup has
 * v / (v + phi2) = 1 and b / (v + phi2) = 0, adding exact zeros to the log
 * terms and to sum b/(v+phi2), and adds 2^-100 to sum 1/(v+phi2), below
 * FP32's resolution of any present subgroup's term (with no subgroup
 * present, b-bar = 0 and the ABF is 0 either way).  12 + 1 bytes per
 * (subgroup, pair) keep four workgroups (32 waves) per CU at 44 subgroups.  Lanes past
 * the end of the pairs repeat the last pair (same inputs, same values, same
 * addresses).  WITH_RAW: the raw ABFs are stored (FP64 when RAW64). */
template <bool WITH_RAW, bool RAW64>
__global__ void __launch_bounds__(NT_ABF) __attribute__((amdgpu_waves_per_eu(8))) uvlr_abf_kernel(const int S, const int64_t NP, const AbfParams ap) {
  extern __shared__ __align__(16) unsigned char sm2[];
  float2* lvz = (float2*)sm2;                     /* [S][64] {v, z^2 / (2 ln 10)}    */
  float* lbv = (float*)(lvz + S * 64);            /* [S][64] b                       */
  uint8_t* lf = (uint8_t*)(lbv + S * 64);         /* [S][64] role of the subgroup    */
  double* part = (double*)sm2;                    /* [8 waves][3][2][64], after jobs */
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  /* XCD-aware order: the workgroups of one XCD (blockIdx & 7) take one
   * contiguous range of 64-pair blocks, so that the raw rows each L2 writes
   * back are contiguous runs rather than every eighth 256-byte piece */
  const int nb = (int)((NP + 63) / 64), per = (nb + 7) / 8;
  const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (blk >= nb) return;
  const int64_t p0 = (int64_t)blk * 64;
  const int cnt = (int)min((int64_t)64, NP - p0);
  for (int e = tid; e < S * 64; e += NT_ABF) {
    const int s = e >> 6, j = e & 63;
    const int64_t o = (int64_t)s * NP + p0 + min(j, cnt - 1);
    const uint8_t f = ap.st_f[o];
    lf[e] = f;
    const bool on = f == F_NORMAL;
    lvz[e] = on ? make_float2(ap.st_v[o], kHalfInvLn10f * ap.st_z2[o]) : make_float2(0x1p100f, 0.f);
    lbv[e] = on ? ap.st_b[o] : 0.f;
  }
  __syncthreads();
  const int64_t p = p0 + min(lane, cnt - 1);
  /* subgroups with NaN statistics: an active one sets every ABF to 0 */
  uint64_t nanm = 0;
  for (int s = 0; s < S; ++s)
    if (lf[s * 64 + lane] == F_NAN) nanm |= 1ull << s;
  const bool any_nan = nanm != 0;
  const float2* lq = lvz + lane;
  const float* lqb = lbv + lane;
  const RawRow<RAW64> rr(ap, p, NP);

  Acc ta0, ta1, ta2;
  ta0.init();
  ta1.init();
  ta2.init();
  const int njobs = ap.n_jobs + ap.n_cfg;
  for (int j = w; j < njobs; j += NT_ABF / 64) {
    if (j < ap.n_jobs) {
      /* consistent configuration at one phi2 */
      const PhiJob jb = ap.jobs[j];
      const float p2 = jb.phi2;
      float L = 0.f, A = 0.f, Bs = 0.f;
      if (p2 != 0.f) {
#pragma unroll 4
        for (int s = 0; s < S; ++s) {
          const float2 q = lq[s * 64];
          const float r = __builtin_amdgcn_rcpf(q.x + p2);
          L += 0.5f * lg10f(q.x * r) + q.y * p2 * r;
          A += lqb[s * 64] * r;
          Bs += r;
        }
      } else {
#pragma unroll 4
        for (int s = 0; s < S; ++s) {
          const float r = __builtin_amdgcn_rcpf(lq[s * 64].x);
          A += lqb[s * 64] * r;
          Bs += r;
        }
      }
      for (int e = jb.e0; e < jb.e1; ++e) {
        const PhiEntry en = ap.ents[e];
        const float l = any_nan ? 0.f : abf_fin(L, A, Bs, en.oma2);
        if (WITH_RAW) rr.put(en.type * ap.GL + en.gi, l);
        if (en.type == 0)
          ta0.add(l);
        else if (en.type == 1)
          ta1.add(l);
        else
          ta2.add(l);
      }
    } else {
      const int cfg = j - ap.n_jobs;
      const uint64_t mk = ap.masks[cfg];
      const int row0 = 3 * ap.GL + cfg * ap.GS;
      float am = -INFINITY, as = 0.f;
      if (__popcll(mk) == 1) {
        /* singleton: bbar = b, V = v + phi2; the log terms telescope to
         * log10(v / (v + w)) and the exponents add up to z^2 w / (v + w) / 2
         * with w = phi2 + omega2 (the subgroup's and the average's terms of
         * CalcLog10AbfUvlr); blocks of 8 grid points share one rescaling of
         * the running sum.  A subgroup outside the Bayes factors (v = 2^100,
         * z = 0) gives log10(1) + 0 = 0 at every grid point. */
        const int s = __ffsll((unsigned long long)mk) - 1;
        const float2 q = lq[s * 64];
        const float v = q.x, hz = q.y;
        auto block = [&](const int g0, const int n) {
          float lb[8];
          float bm = -INFINITY;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            float l = -INFINITY;
            if (i < n) {
              const float wg = ap.gS[2 * ABF_GMAX + g0 + i];
              const float r = __builtin_amdgcn_rcpf(v + wg);
              l = 0.5f * lg10f(v * r) + hz * wg * r;
              if (WITH_RAW) rr.put(row0 + g0 + i, l);
            }
            lb[i] = l;
            bm = fmaxf(bm, l);
          }
          const float M = fmaxf(am, bm);
          float acc = as * ex2f((am - M) * kLog2_10);
#pragma unroll
          for (int i = 0; i < 8; ++i) acc += ex2f((lb[i] - M) * kLog2_10);
          as = acc;
          am = M;
        };
        int g0 = 0;
        for (; g0 + 8 <= ap.GS; g0 += 8) block(g0, 8);
        if (g0 < ap.GS) block(g0, ap.GS - g0);
      } else {
        const bool nan = (mk & nanm) != 0ull;
        Acc a;
        a.init();
        for (int gi = 0; gi < ap.GS; ++gi) {
          const float p2 = ap.gS[gi], o2 = ap.gS[ABF_GMAX + gi];
          float L = 0.f, A = 0.f, Bs = 0.f;
          uint64_t m = mk;
          while (m) {
            const int s = __ffsll((unsigned long long)m) - 1;
            m &= m - 1ull;
            const float2 q = lq[s * 64];
            const float r = __builtin_amdgcn_rcpf(q.x + p2);
            if (p2 != 0.f) L += 0.5f * lg10f(q.x * r) + q.y * p2 * r;
            A += lqb[s * 64] * r;
            Bs += r;
          }
          const float l = nan ? 0.f : abf_fin(L, A, Bs, o2);
          if (WITH_RAW) rr.put(row0 + gi, l);
          a.add(l);
        }
        am = a.m;
        as = a.s;
      }
      ap.w[(int64_t)(3 + cfg) * NP + p] = acc_finish(am, (double)as, 1.0 / ap.GS);
    }
  }
  /* combine the consistent-configuration partial sums of the eight waves
   * (part lies over the subgroup records, which are no longer read) */
  __syncthreads();
  part[((w * 3 + 0) * 2) * 64 + lane] = (double)ta0.m;
  part[((w * 3 + 0) * 2 + 1) * 64 + lane] = (double)ta0.s;
  part[((w * 3 + 1) * 2) * 64 + lane] = (double)ta1.m;
  part[((w * 3 + 1) * 2 + 1) * 64 + lane] = (double)ta1.s;
  part[((w * 3 + 2) * 2) * 64 + lane] = (double)ta2.m;
  part[((w * 3 + 2) * 2 + 1) * 64 + lane] = (double)ta2.s;
  __syncthreads();
  if (w < 3) {
    const int k = w;
    double M = -INFINITY;
    for (int ww = 0; ww < NT_ABF / 64; ++ww) M = fmax(M, part[((ww * 3 + k) * 2) * 64 + lane]);
    double ssum = 0;
    for (int ww = 0; ww < NT_ABF / 64; ++ww) {
      const double m = part[((ww * 3 + k) * 2) * 64 + lane];
      if (m > -INFINITY) ssum += part[((ww * 3 + k) * 2 + 1) * 64 + lane] * exp10(m - M);
    }
    ap.w[(int64_t)k * NP + p] = acc_finish((float)M, ssum, 1.0 / ap.GL);
  }
}

template <int NXW, int NMW, bool MULTI>
void launch_sums_m(const eqbf_chunk& c, const SumsParams& sp, int grid, size_t lds, hipStream_t st, bool dma) {
  if (!dma)
    hipLaunchKernelGGL((uvlr_sums_kernel<NXW, NMW, MULTI>), dim3(grid), dim3(NT_SUMS), lds, st, c, sp);
  else if (c.geno_dtype == EQBF_F32)
    hipLaunchKernelGGL((uvlr_sums_dma_kernel<NXW, NMW, MULTI, false>), dim3(grid), dim3(NT_SUMS), lds, st, c, sp);
  else
    hipLaunchKernelGGL((uvlr_sums_dma_kernel<NXW, NMW, MULTI, true>), dim3(grid), dim3(NT_SUMS), lds, st, c, sp);
}

template <int NXW, int NMW>
void launch_sums(const eqbf_chunk& c, const SumsParams& sp, int grid, size_t lds, hipStream_t st, bool dma) {
  if (sp.SPP >= c.n_subgroups)
    launch_sums_m<NXW, NMW, false>(c, sp, grid, lds, st, dma);
  else
    launch_sums_m<NXW, NMW, true>(c, sp, grid, lds, st, dma);
}

/* column tiles per wave: x-tiles padded to one of a few counts */
template <int NMW>
int dispatch_sums(int nxw, const eqbf_chunk& c, const SumsParams& sp, int grid, size_t lds, hipStream_t st, bool dma) {
  if (nxw <= 2) launch_sums<2, NMW>(c, sp, grid, lds, st, dma);
  else if (nxw <= 4) launch_sums<4, NMW>(c, sp, grid, lds, st, dma);
  else if (nxw <= 7) launch_sums<7, NMW>(c, sp, grid, lds, st, dma);
  else if (nxw <= 10) launch_sums<10, NMW>(c, sp, grid, lds, st, dma);
  else if (nxw <= 14) launch_sums<14, NMW>(c, sp, grid, lds, st, dma);
  else if (nxw <= 18) launch_sums<18, NMW>(c, sp, grid, lds, st, dma);
  else return 1;
  return 0;
}

int sums_nxw(int nxw) { return nxw <= 2 ? 2 : nxw <= 4 ? 4 : nxw <= 7 ? 7 : nxw <= 10 ? 10 : nxw <= 14 ? 14 : 18; }

}  // namespace

bool uvlr_fast_eligible(const eqbf_chunk& c, int KU) {
  return c.n_geno_sets == 1 && c.n_subgroups <= 64 && KU <= 8;
}


int uvlr_fast_run(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_assoc_opts* o, eqbf_assoc_out* out, hipStream_t st,
                  const PrepRec* rec, const double* B, const uint64_t* memb, const int32_t* idx, const int4* tiles,
                  const int* n_tiles, int KU, const uint64_t* d_masks, int n_cfg) {
  const int S = c.n_subgroups, G = c.n_genes;
  const int64_t NP = c.n_pairs;
  const bool join = o->analysis == EQBF_ANALYSIS_JOIN;
  SumsParams sp;
  memset(&sp, 0, sizeof sp);
  sp.rec = rec;
  sp.B = B;
  sp.memb = memb;
  sp.idx = idx;
  sp.tiles = tiles;
  sp.n_tiles = n_tiles;
  sp.KU = KU;
  sp.KC = KU + 1;
  sp.CX = S * sp.KC;
  sp.NXT = (sp.CX + 15) / 16;
  sp.tz = ctx_tz_table(ctx);
  sp.n_out = out->n;
  sp.sstats = out->sstats;
  sp.stdsstats = out->stdsstats;
  if (join) {
    uint8_t* base = (uint8_t*)ctx_buf(ctx, "abf_in", (size_t)S * NP * 13 + 64);
    if (!base) return ctx_fail(ctx, EQBF_ERR_HIP, "allocating the Bayes-factor inputs failed");
    sp.st_v = (float*)base;
    sp.st_b = sp.st_v + (size_t)S * NP;
    sp.st_z2 = sp.st_b + (size_t)S * NP;
    sp.st_f = (uint8_t*)(sp.st_z2 + (size_t)S * NP);
  }
  const int nxw = sums_nxw((sp.NXT + 1) / 2), nmw = ((S + 15) / 16 + 1) / 2;
  if ((sp.NXT + 1) / 2 > 18) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "too many subgroup columns for the batched path");
  const int CXPP = 2 * nxw * 16;
  /* Launch shape: samples per LDS stage (KCH, two stages within 80 KB),
   * subgroups per epilogue pass (SPP; all at once needs 512 (KC + 1) bytes
   * per subgroup and may leave room for one workgroup per CU), workgroups per
   * CU.  With the autotuner on, the first call of each shape times the
   * candidates on the chunk itself (every candidate makes the same sums in
   * the same order) and keeps the fastest for the context's lifetime. */
  /* LDS-direct staging when the coefficient rows split into 16-byte units;
   * the statistics are then a kernel of their own (the sums kernel keeps
   * its registers for the products) */
  /* int8-digit sums when only Bayes factors are written (their inputs are
   * FP32); EQBF_I8_SUMS=0 in the environment keeps the FP64 product */
  const char* i8env = getenv("EQBF_I8_SUMS");
  const bool i8 = join && !sp.sstats && !sp.stdsstats && uvlr_i8_eligible(c, KU) && !(i8env && i8env[0] == '0');
  const bool dma = i8 || (sp.CX % 2) == 0;
  const size_t max_tiles = (size_t)(NP / 64) + (size_t)G + 1;
  /* the sums go through sx2 to the statistics kernel; items the sums
   * kernel marks in the redo bitmap are recomputed in FP64 in between */
  const int64_t redo_words = ((int64_t)S * NP + 31) / 32;
  sp.sx2 = (double*)ctx_buf(ctx, "sums_x2", (size_t)2 * S * std::max<int64_t>(NP, 1) * sizeof(double));
  sp.redo = (uint32_t*)ctx_buf(ctx, "sums_redo", (size_t)(redo_words + 2) * sizeof(uint32_t));
  if (!sp.sx2 || !sp.redo) return EQBF_ERR_HIP; /* ctx_buf has set the message */
  hipLaunchKernelGGL(redo_zero_kernel, dim3(256), dim3(256), 0, st, sp.redo, redo_words);
  const size_t xsz = c.geno_dtype == EQBF_F32 ? 4 : 8;
  auto stage_bytes = [&](int kch) {
    return dma ? (size_t)kch * ((size_t)sp.CX * 8 + 64 * xsz + 8) : (size_t)8 * kch * (65 + CXPP);
  };
  int kch_max = 4;
  while (kch_max < 32 && 2 * stage_bytes(kch_max + 4) <= (dma ? 144 : 80) * 1024) kch_max += 4;
  if (!dma) kch_max = std::min(16, kch_max);
  if (2 * stage_bytes(kch_max) > 160 * 1024)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "too many subgroup columns for the batched path");
  const size_t epi_unit = (size_t)64 * 8 * (sp.KC + 1);
  auto shape_lds = [&](int kch, int spp) { return std::max(2 * stage_bytes(kch), epi_unit * spp); };
  eqbf_ctx::TuneChoice choice;
  choice.kch = kch_max;
  choice.spp = (int)std::min<size_t>(S, std::max(shape_lds(kch_max, 0), std::min<size_t>(epi_unit * S, 160 * 1024)) /
                                            epi_unit);
  choice.grid_mult = 2;
  if (dma) {
    /* default: two workgroups per CU (the kernel is built for 4 waves per
     * SIMD), i.e. at most 80 KB of LDS each */
    int k2 = 4;
    while (k2 < 32 && 2 * stage_bytes(k2 + 4) <= 80 * 1024) k2 += 4;
    const int s2 = (int)std::min<size_t>(S, (80 * 1024) / epi_unit);
    if (2 * stage_bytes(k2) <= 80 * 1024 && s2 >= 1) {
      choice.kch = k2;
      choice.spp = s2;
    }
  }
  auto launch = [&](const eqbf_ctx::TuneChoice& ch) -> int {
    SumsParams q = sp;
    q.KCH = ch.kch;
    q.SPP = ch.spp;
    const size_t lds = shape_lds(ch.kch, ch.spp);
    const int grid = (ctx->n_cu * ch.grid_mult + 7) / 8 * 8;
    return nmw <= 1 ? dispatch_sums<1>(nxw, c, q, grid, lds, st, dma) : dispatch_sums<2>(nxw, c, q, grid, lds, st, dma);
  };
  char key[96];
  snprintf(key, sizeof key, "sums/%d/%d/%d/%d/%d/%s", S, sp.KC, nxw, nmw, c.n_samples,
           dma ? (xsz == 4 ? "dma32" : "dma64") : "regs");
  auto tk = ctx->tuned.find(key);
  if (i8) {
    const int rc = uvlr_i8_sums_run(ctx, c, st, rec, B, memb, tiles, n_tiles, KU, sp.sx2, sp.redo);
    if (rc) return rc;
  } else if (tk != ctx->tuned.end()) {
    choice = tk->second;
  } else if (ctx->autotune_on && !ctx->capturing) {
    std::vector<eqbf_ctx::TuneChoice> cands;
    for (int kch : {kch_max, std::max(4, (kch_max / 2) & ~3), std::max(4, (kch_max / 4) & ~3)})
      for (int spp : {choice.spp, (S + 1) / 2, (S + 3) / 4})
        for (int gm : {1, 2}) {
          eqbf_ctx::TuneChoice ch;
          ch.kch = kch;
          ch.spp = std::max(1, std::min(S, spp));
          ch.grid_mult = gm;
          if (shape_lds(ch.kch, ch.spp) > 160 * 1024) continue;
          bool dup = false;
          for (auto& d : cands) dup |= d.kch == ch.kch && d.spp == ch.spp && d.grid_mult == ch.grid_mult;
          if (!dup) cands.push_back(ch);
        }
    hipEvent_t e0, e1;
    EQBF_HIP(ctx, hipEventCreate(&e0));
    EQBF_HIP(ctx, hipEventCreate(&e1));
    float best = 1e30f;
    for (auto& ch : cands) {
      if (launch(ch)) continue; /* warm-up */
      EQBF_HIP(ctx, hipEventRecord(e0, st));
      launch(ch);
      EQBF_HIP(ctx, hipEventRecord(e1, st));
      EQBF_HIP(ctx, hipEventSynchronize(e1));
      float ms = 0.f;
      EQBF_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        choice = ch;
        choice.ms = ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tuned[key] = choice;
  }
  if (!i8) {
    ProfScope ps(ctx, "uvlr_sums_kernel", st);
    if (launch(choice)) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "too many subgroup columns for the batched path");
  }
  EQBF_HIP(ctx, hipGetLastError());
  {
    ProfScope ps(ctx, "sums_repair_kernel", st);
    hipLaunchKernelGGL(sums_repair_kernel, dim3(1024), dim3(NT_REPAIR), 0, st, c, sp);
  }
  EQBF_HIP(ctx, hipGetLastError());
  {
    ProfScope ps(ctx, "uvlr_stats_kernel", st);
    const dim3 sgrid((unsigned)((max_tiles + 7) / 8 * 8));
    if (sp.sstats || sp.stdsstats)
      hipLaunchKernelGGL(uvlr_stats_kernel<true>, sgrid, dim3(NT_STATS), 0, st, c, sp);
    else
      hipLaunchKernelGGL(uvlr_stats_kernel<false>, sgrid, dim3(NT_STATS), 0, st, c, sp);
    EQBF_HIP(ctx, hipGetLastError());
  }
  if (!join) return EQBF_OK;

  /* jobs of the consistent configuration: distinct phi2 of gen, gen-fix,
   * gen-maxh with the grid points each one finishes.  They depend on the
   * large grid only and stay on the device until a buffer or the grids
   * change, so that a repeated call enqueues kernels only. */
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const size_t jobs_cap = (size_t)3 * GL * (sizeof(PhiJob) + sizeof(PhiEntry)) + 64;
  uint8_t* jt = (uint8_t*)ctx_buf(ctx, "abf_jobs", jobs_cap);
  if (!jt) return EQBF_ERR_HIP;
  if (ctx->jobs_epoch != ctx->buf_epoch) {
    std::vector<PhiJob> jobs;
    std::vector<PhiEntry> ents;
    std::vector<std::pair<float, PhiEntry>> all;
    for (int gi = 0; gi < GL; ++gi) {
      const double p = ctx->phiL[gi], om = ctx->omaL[gi];
      all.push_back({(float)p, PhiEntry{0, gi, (float)om}});
      all.push_back({0.f, PhiEntry{1, gi, (float)(p + om)}});
      all.push_back({(float)(p + om), PhiEntry{2, gi, 0.f}});
    }
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<float, PhiEntry>& a, const std::pair<float, PhiEntry>& b) { return a.first < b.first; });
    for (size_t k = 0; k < all.size(); ++k) {
      if (jobs.empty() || jobs.back().phi2 != all[k].first) jobs.push_back(PhiJob{all[k].first, (int)k, (int)k});
      ents.push_back(all[k].second);
      jobs.back().e1 = (int)k + 1;
    }
    const size_t jb_bytes = jobs.size() * sizeof(PhiJob), en_bytes = ents.size() * sizeof(PhiEntry);
    EQBF_HIP(ctx, hipMemcpyAsync(jt, jobs.data(), jb_bytes, hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipMemcpyAsync(jt + jb_bytes, ents.data(), en_bytes, hipMemcpyHostToDevice, st));
    EQBF_HIP(ctx, hipStreamSynchronize(st)); /* host vectors go out of scope */
    ctx->jobs_bytes = jb_bytes;
    ctx->jobs_n = (int)jobs.size();
    ctx->jobs_epoch = ctx->buf_epoch;
  }
  if (GS > ABF_GMAX)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "mixed path supports small grids of at most %d points", ABF_GMAX);
  if (out->raw_abf && o->raw_dtype != EQBF_F32_TILED && NP * (o->raw_dtype == EQBF_F64 ? 8 : 4) > (int64_t)UINT32_MAX)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "row-major raw ABFs of more than %lld pairs per call",
                    (long long)(UINT32_MAX / (o->raw_dtype == EQBF_F64 ? 8 : 4)));
  AbfParams ap;
  memset(&ap, 0, sizeof ap);
  ap.st_v = sp.st_v;
  ap.st_b = sp.st_b;
  ap.st_z2 = sp.st_z2;
  ap.st_f = sp.st_f;
  ap.jobs = (const PhiJob*)jt;
  ap.ents = (const PhiEntry*)(jt + ctx->jobs_bytes);
  ap.masks = d_masks;
  ap.grid = (const double*)ctx->d_grid.ptr;
  ap.raw = out->raw_abf;
  ap.w = out->w_abf;
  ap.n_jobs = ctx->jobs_n;
  ap.n_cfg = n_cfg;
  ap.GL = GL;
  ap.GS = GS;
  ap.raw_f64 = o->raw_dtype == EQBF_F64;
  ap.raw_tiled = o->raw_dtype == EQBF_F32_TILED;
  ap.n_raw = 3 * (int64_t)GL + (int64_t)n_cfg * GS;
  for (int k = 0; k < GS; ++k) {
    ap.gS[k] = (float)ctx->phiS[k];
    ap.gS[ABF_GMAX + k] = (float)ctx->omaS[k];
    ap.gS[2 * ABF_GMAX + k] = (float)(ctx->phiS[k] + ctx->omaS[k]);
  }
  const size_t lds2 = std::max((size_t)S * 64 * 13, (size_t)(NT_ABF / 64) * 3 * 2 * 64 * 8);
  {
    ProfScope ps(ctx, "uvlr_abf_kernel", st);
    const dim3 grid((unsigned)(((NP + 63) / 64 + 7) / 8 * 8));
    if (!out->raw_abf)
      hipLaunchKernelGGL((uvlr_abf_kernel<false, false>), grid, dim3(NT_ABF), lds2, st, S, NP, ap);
    else if (ap.raw_f64)
      hipLaunchKernelGGL((uvlr_abf_kernel<true, true>), grid, dim3(NT_ABF), lds2, st, S, NP, ap);
    else
      hipLaunchKernelGGL((uvlr_abf_kernel<true, false>), grid, dim3(NT_ABF), lds2, st, S, NP, ap);
  }
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}
