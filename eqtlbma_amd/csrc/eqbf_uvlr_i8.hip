/*
 * Batched uvlr sums on the int8 matrix cores, for the association calls
 * whose outputs are Bayes factors only (no FP64 summary statistic written).
 *
 * The sums of a tile of pairs (SNPs) of gene g over the N samples,
 *   C_x[snp][col] = sum_i x'[i][snp] * B[i][col]   col = (s, k): y~_s, U_s,k
 *   C_m[snp][s]   = sum_i x'[i][snp]^2 * m_s[i]   m_s = membership of s
 * (the same products as the FP64 sums kernel; B is zero outside the
 * members), are formed exactly in integer arithmetic from fixed-point
 * digits:
 *  - x' = x - c_snp, c_snp the mean of the first 64 samples of the SNP
 *    (rounded to the fixed-point grid); the intercept is a column of U, so
 *    x~'x~ = x''x' - |U'x'|^2 and x'y~ = x''y~ (y~ is orthogonal to U)
 *    equal the uncentred sums.  x' is held as X = rint(x' 2^e) with |X| <=
 *    2^30 and split into four balanced 8-bit digits, x'^2 as Y = rint(X^2
 *    2^-30), four digits;
 *  - B as Bq = rint(b 2^E) per column (E = 30 for the unit-norm U columns,
 *    30 - ceil(log2 ||y~||) for y~), four digits, precomputed once per
 *    chunk in the layout the matrix cores read (b8_convert_kernel);
 *  - v_mfma_i32_16x16x64_i8 multiplies 16 SNPs x 64 samples by 64 samples x
 *    16 columns exactly in int32; the digit pairs of weights 2^48 .. 2^24
 *    (ten pairs) are kept, the rest (< 2^-30 of the product's range) are
 *    dropped.  The int32 sums of one weight class are exact over every
 *    chunk (N <= 32768), and the recombination in FP64 is exact too.
 * So the only errors are the roundings of x' (2^-31 of the range the scale
 * covers around the centre, 4 for dosages) and of B (2^-31 of the column's
 * bound),
 * ~1e-8 relative in x~'x~ and x'y~: far below the FP32 arithmetic of the
 * Bayes-factor terms that consume them.  Where the subtraction x''x' -
 * |U'x'|^2 loses more than three bits (x~'x~ <= x''x' / 8: near-collinear
 * or monomorphic SNPs), where R^2 >= 7/8 (the residual sum of squares
 * would lose bits), or where a genotype is not finite or outside the range
 * the scale covers, the tile marks the (pair, subgroup) in a bitmap and
 * sums_repair_kernel recomputes it in FP64 over its members, as the FP64
 * path does (FitSingleGeneWithSingleSnp, utils_math.cpp:166-209).
 *
 * Launch: one 512-thread workgroup per CU (LDS: two B stages of 64
 * samples, two digit buffers of the tile's genotypes), persistent over half
 * tiles of 32 pairs with contiguous ranges per XCD.  Each wave owns both
 * 16-row blocks of the half tile for NBX y~/U column blocks of 16 (ten
 * products per row block and step) and one x'^2 block (four); the slots are
 * static, so the int32 accumulators stay in place (an empty slot multiplies
 * block 0 and is not written).  Per 64-sample step: the load unit copies
 * the next step's B digits HBM -> LDS (global_load_lds) and the genotypes
 * of three steps ahead are loaded into registers while the matrix cores
 * work; then the next step's genotypes are centred, scaled and split into
 * the other digit buffer; one barrier per step.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "eqbf_internal.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void_t;

constexpr int I8_NT = 512;          /* threads per sums workgroup (8 waves)       */
constexpr int I8_ROWS = 32;         /* SNP rows per pass (half a tile)             */
constexpr int I8_PLANE = I8_ROWS * 64; /* bytes of one digit plane: 32 rows x 64 samples */
constexpr int I8_ABUF = 8 * I8_PLANE;  /* the eight planes of one step              */
/* s_waitcnt vmcnt(n) lgkmcnt(0) (expcnt not waited for) */
constexpr int I8_VMCNT4 = 4 | (7 << 4);
constexpr int I8_LGKM0 = 15 | (7 << 4) | (3 << 14); /* lgkmcnt(0) only */
constexpr int I8_VMCNT8 = 8 | (7 << 4);

/* 16 bytes per lane from global memory to LDS (M0 + 16 lane), by the load
 * unit.  Written out rather than through the builtin: the compiler treats
 * the builtin's LDS write as aliasing every later LDS read and waits for
 * every vector-memory operation (vmcnt(0)) before the next one, which
 * serialises the copy of the next step with the reads of this one.  The
 * instruction has no register output; its completion is awaited by the
 * explicit vmcnt of i8_barrier. */
__device__ __forceinline__ void dma16(const void* src, uint32_t lds_addr) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds_addr) : "memory", "m0");
}

/* Workgroup barrier after the given wait: LDS writes done (lgkmcnt(0)) and
 * the vector-memory operations older than the newest n done (the copies
 * into LDS among them), without the release fence of __syncthreads, whose
 * vmcnt(0) would also wait for the genotype loads still in flight.  The
 * empty asm statements keep the compiler from moving memory accesses
 * across it. */
template <int WAIT>
__device__ __forceinline__ void i8_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(WAIT);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
constexpr int I8_NBMAX = 3;         /* y~/U column blocks per wave                */
constexpr double I8_TAU = 0.125;    /* x~'x~ <= I8_TAU x''x': recomputed in FP64  */
constexpr double I8_RSS = 0.125;    /* rss <= I8_RSS y~'y~ (R^2 >= 7/8): recomputed */

struct I8Params {
  const PrepRec* rec;
  const uint8_t* B8;  /* [G][NCK][CKB] digit planes of B and the membership    */
  const int4* tiles;
  const int* n_tiles;
  double* sx2;        /* [2][S][NP]: x~'x~ and x'y~                              */
  uint32_t* redo;     /* [S * NP bits]: items to recompute in FP64               */
  int KC, KU, CX, CXP, SXP, NXB, NCK, CKB, ES;
  int8_t bx[8 * I8_NBMAX]; /* y~/U column blocks of each wave (-1: none)      */
  int8_t bm[8];             /* x'^2 column block of each wave (-1: none)       */
};

/* the 16-byte unit u of row r of a 64-byte digit row sits at u ^ swz(r), so
 * that any 8 lanes reading one unit of 8 consecutive rows cover 128
 * consecutive bank bytes */
__device__ __forceinline__ int swz(int r) { return (r >> 1) & 3; }

/* 4 x 4 byte transpose: out[p] holds byte 3 - p of in[0..3] (digit p of four
 * samples whose digit words hold d3, d2, d1, d0 in bytes 0..3) */
__device__ __forceinline__ void transpose4(const uint32_t (&in)[4], uint32_t (&out)[4]) {
  const uint32_t t0 = __builtin_amdgcn_perm(in[1], in[0], 0x05010400u), t1 = __builtin_amdgcn_perm(in[1], in[0], 0x07030602u);
  const uint32_t u0 = __builtin_amdgcn_perm(in[3], in[2], 0x05010400u), u1 = __builtin_amdgcn_perm(in[3], in[2], 0x07030602u);
  out[3] = __builtin_amdgcn_perm(u0, t0, 0x05040100u);
  out[2] = __builtin_amdgcn_perm(u0, t0, 0x07060302u);
  out[1] = __builtin_amdgcn_perm(u1, t1, 0x05040100u);
  out[0] = __builtin_amdgcn_perm(u1, t1, 0x07060302u);
}

/* balanced base-256 digits: X = d0 2^24 + d1 2^16 + d2 2^8 + d3 with d1..d3
 * in [-128, 127] and d0 in [-64, 64] for |X| <= 2^30 */
__device__ __forceinline__ void digits4(int X, int (&d)[4]) {
  d[3] = (int)(int8_t)(X & 255);
  X = (X - d[3]) >> 8;
  d[2] = (int)(int8_t)(X & 255);
  X = (X - d[2]) >> 8;
  d[1] = (int)(int8_t)(X & 255);
  d[0] = (X - d[1]) >> 8;
}

/* scale exponent of the y~ column of (g, s): |y~_i| <= ||y~|| < 2^k */
__device__ __forceinline__ int ytilde_exp(double yy) {
  if (!(yy > 0.0)) return 30;
  int k;
  (void)frexp(sqrt(yy), &k);
  return 30 - k;
}

/* Digit planes of B and of the membership for chunk ck of gene g: plane p
 * (0..3) holds digit p of every column (CXP rows of 64 samples), plane 4
 * the membership bytes of every subgroup (SXP rows); units swizzled. */
__global__ void __launch_bounds__(256) b8_convert_kernel(const PrepRec* __restrict__ rec, const double* __restrict__ B,
                                                         const uint64_t* __restrict__ memb, uint8_t* __restrict__ B8,
                                                         int N, int S, int KC, int CX, int CXP, int SXP, int NCK,
                                                         int CKB) {
  const int g = blockIdx.x / NCK, ck = blockIdx.x % NCK;
  const int CT = CXP + SXP;
  uint8_t* dst = B8 + ((int64_t)g * NCK + ck) * CKB;
  /* the four 16-sample units of a column row on adjacent lanes: each row
   * is one 64-byte write */
  for (int it = threadIdx.x; it < 4 * CT; it += 256) {
    const int u = it & 3, col = it >> 2;
    const int i0 = ck * 64 + 16 * u;
    if (col < CXP) {
      uint32_t w[4][4];
      if (col < CX) {
        const int s = col / KC, k = col - s * KC;
        const double sc = ldexp(1.0, k == 0 ? ytilde_exp(rec[(int64_t)g * S + s].yy) : 30);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
          uint32_t pw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int i = i0 + 4 * t4 + t;
            const double b = i < N ? B[((int64_t)g * N + i) * CX + col] : 0.0;
            int d[4];
            digits4((int)rint(b * sc), d);
#pragma unroll
            for (int p = 0; p < 4; ++p) pw[p] |= ((uint32_t)d[p] & 255u) << (8 * t);
          }
#pragma unroll
          for (int p = 0; p < 4; ++p) w[p][t4] = pw[p];
        }
      } else {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int t4 = 0; t4 < 4; ++t4) w[p][t4] = 0u;
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
        *(uint4*)(dst + ((int64_t)p * CXP + col) * 64 + 16 * (u ^ swz(col))) =
            make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
    } else {
      const int s = col - CXP;
      uint32_t w[4] = {0u, 0u, 0u, 0u};
      if (s < S) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const int i = i0 + t;
          const uint32_t bit = i < N ? (uint32_t)((memb[(int64_t)g * N + i] >> s) & 1ull) : 0u;
          w[t >> 2] |= bit << (8 * (t & 3));
        }
      }
      *(uint4*)(dst + ((int64_t)4 * CXP + s) * 64 + 16 * (u ^ swz(s))) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

template <int NBX, bool XF64>
__global__ void __launch_bounds__(I8_NT) uvlr_sums_i8_kernel(const eqbf_chunk c, const I8Params q) {
  extern __shared__ __align__(16) unsigned char smem[];
  typedef typename std::conditional<XF64, double, float>::type xt;
  const int N = c.n_samples, S = c.n_subgroups;
  const int64_t NP = c.n_pairs;
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, kq = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  /* conversion: SNP row jx, samples 4 gq .. 4 gq + 3 of a 64-sample step */
  const int jx = tid & 31, gq = tid >> 5;
  const int CKB = q.CKB, NCK = q.NCK, CXP = q.CXP, NXB = q.NXB;
  unsigned char* Abase = smem + 2 * CKB;
  const int small_off = max(2 * CKB + 2 * I8_ABUF, I8_ROWS * q.ES * 8);
  double* ryy_l = (double*)(smem + small_off); /* y~'y~ of each subgroup of the gene      */
  int* rn_l = (int*)(ryy_l + 64);          /* members of each subgroup                  */
  int* rf_l = rn_l + 64;                   /* prep flags of each subgroup               */
  int* yexp_l = rf_l + 64;                 /* scale exponent of the y~ column           */
  int* cx_l = yexp_l + 64;                 /* centre of each SNP row, in units of 2^-e  */
  int* ex_l = cx_l + I8_ROWS;              /* scale exponent e of each SNP row          */
  int* bad_l = ex_l + I8_ROWS;             /* row to recompute in FP64                  */
  uint8_t* snpin_l = (uint8_t*)(bad_l + I8_ROWS); /* [32][64]: SNP of the row kept in s */
  const int sw = swz(l16);
  /* every wave runs NBX y~/U slots and one x'^2 slot without branches (an
   * empty slot multiplies block 0 and is not written), so that the
   * accumulators stay in place */
  int bxs[NBX];
#pragma unroll
  for (int b = 0; b < NBX; ++b) bxs[b] = max(0, (int)q.bx[w * I8_NBMAX + b]);
  const int bms = max(0, (int)q.bm[w]);

  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const int T = 2 * *q.n_tiles; /* half tiles of 32 pairs */
  const int span = (T + 7) / 8;
  const int t0 = xcd * span, t1 = min(T, t0 + span);
  /* the SNP of this thread's row in half tile t (a valid pair of the tile
   * when the half is empty) */
  auto row_snp = [&](int t, const int4& tl) -> int64_t {
    const int cn = tl.z - 32 * (t & 1);
    return c.pair_snp[cn > 0 ? tl.y + 32 * (t & 1) + min(jx, cn - 1) : tl.y];
  };
  /* the next half tile's descriptor and SNP are loaded one half tile ahead */
  int4 tile_n = make_int4(0, 0, 0, 0);
  int64_t snp_n = 0;
  if (t0 + slot < t1) {
    tile_n = q.tiles[(t0 + slot) >> 1];
    snp_n = row_snp(t0 + slot, tile_n);
  }
  for (int t = t0 + slot; t < t1; t += per_xcd) {
    const int4 tile = tile_n;
    const int64_t snp = snp_n;
    if (t + per_xcd < t1) {
      tile_n = q.tiles[(t + per_xcd) >> 1];
      snp_n = row_snp(t + per_xcd, tile_n);
    }
    const int g = tile.x;
    const int64_t p0 = tile.y + 32 * (t & 1);
    const int cnt = tile.z - 32 * (t & 1);
    if (cnt <= 0) continue; /* uniform over the workgroup */
    const xt* gcol = (const xt*)c.genotypes + snp;
    const uint8_t* Bg = q.B8 + (int64_t)g * NCK * CKB;

    auto issue_b = [&](int ck) {
      const unsigned char* src = Bg + (int64_t)ck * CKB;
      const uint32_t dst = (uint32_t)(uintptr_t)(lds_void_t*)(smem + (ck & 1) * CKB);
      const int n16 = CKB / 16; /* a multiple of 64: CXP and SXP are multiples of 16 */
      for (int u0 = w * 64; u0 < n16; u0 += I8_NT) dma16(src + (int64_t)(u0 + lane) * 16, dst + u0 * 16);
    };
    /* genotypes of steps ck + 1 (xr), ck + 2 (xn) and ck + 3 (loaded into
     * xl during step ck): every load is issued, so that the step's wait for
     * its B copy can leave exactly the four newest loads in flight.  Rows
     * past N read row N - 1: their B rows and membership bytes are zero, so
     * whatever digits they get add nothing. */
    xt xr[4], xn[4], xl[4];
    auto load_x = [&](int ck, xt (&dst)[4]) {
#pragma unroll
      for (int k = 0; k < 4; ++k) dst[k] = gcol[(int64_t)min(ck * 64 + 4 * gq + k, N - 1) * c.geno_ld];
    };
    /* centre, scale and split the step's genotypes into digit planes 0..3
     * (x') and 4..7 (x'^2) of A buffer ck & 1.  X = rint(x 2^e) - C with C =
     * rint(c 2^e) is rint((x - c') 2^e) for the centre c' = C 2^-e; the
     * balanced digits of X are the bytes of X + 0x808080 with the low three
     * flipped by 0x80, and four samples' digit words are transposed into the
     * planes with byte permutes. */
    float sxf = 0.f;
    double sxd = 0.0;
    int Cx = 0;
    auto convert = [&](int ck, const xt (&xv)[4]) {
      uint32_t wx[4], wy[4];
      bool bad = false;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        int D;
        if (XF64) {
          const double r = rint((double)xv[k] * sxd);
          bad |= r != r;
          D = (int)fmin(fmax(r, -2147483648.0), 2147483520.0) - Cx;
        } else {
          const float r = rintf((float)xv[k] * sxf);
          bad |= r != r;
          D = (int)__builtin_amdgcn_fmed3f(r, -2147483648.f, 2147483520.f) - Cx;
        }
        /* |X| <= 2^30; a saturated conversion lands outside too (|C| < 2^30) */
        bad |= (uint32_t)(D + (1 << 30)) > (1u << 31);
        const double t = (double)D * (1.0 / 32768.0);
        const int Y = (int)rint(t * t); /* X^2 2^-30, 0 <= Y <= 2^30 */
        wx[k] = (uint32_t)(D + 0x808080) ^ 0x808080u;
        wy[k] = (uint32_t)(Y + 0x808080) ^ 0x808080u;
      }
      if (bad) bad_l[jx] = 1;
      uint32_t px[4], py[4];
      transpose4(wx, px);
      transpose4(wy, py);
      unsigned char* A = Abase + (ck & 1) * I8_ABUF;
      const int off = jx * 64 + 16 * ((gq >> 2) ^ swz(jx)) + 4 * (gq & 3);
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        *(uint32_t*)(A + p * I8_PLANE + off) = px[p];
        *(uint32_t*)(A + (4 + p) * I8_PLANE + off) = py[p];
      }
    };

    v4i acc[NBX + 1][2][4]; /* [NBX]: the x'^2 slot */
#pragma unroll
    for (int b = 0; b < NBX + 1; ++b)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[b][rb][k] = v4i{0, 0, 0, 0};

    issue_b(0);
    /* what the epilogue reads, staged in LDS with the first step: the
     * subgroup records of the gene and whether each row's SNP is kept in
     * each subgroup */
    double r_yy = 0.0;
    int r_n = 0, r_fl = 0;
    if (tid < S) {
      const PrepRec* rp = q.rec + (int64_t)g * S + tid;
      r_yy = rp->yy;
      r_n = rp->n;
      r_fl = rp->flags;
    }
    uint32_t sin4 = 0u; /* byte k: row jx's SNP kept in subgroup gq + 16 k */
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (gq + 16 * k < S) sin4 |= (uint32_t)c.snp_in_sbgrp[snp * S + gq + 16 * k] << (8 * k);
    load_x(0, xr);
    load_x(1, xn);
    load_x(2, xl);
    /* centre and scale of each SNP from the first step: c = mean of the
     * finite values, e = 29 - k with 2^k > max(max |x - c|, 1): |x - c| up
     * to 2^(k+1) fits in 2^30 (4 for dosages, whatever the centre) */
    {
      double sm = 0.0, mn = INFINITY, mx = -INFINITY;
      int nf = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double v = (double)xr[k];
        if (4 * gq + k < N && isfinite(v)) {
          sm += v;
          ++nf;
          mn = fmin(mn, v);
          mx = fmax(mx, v);
        }
      }
      double* rs = (double*)(Abase + I8_ABUF); /* partials in A buffer 1, not yet in use */
      rs[(0 * 16 + gq) * 32 + jx] = sm;
      rs[(1 * 16 + gq) * 32 + jx] = mn;
      rs[(2 * 16 + gq) * 32 + jx] = mx;
      rs[(3 * 16 + gq) * 32 + jx] = (double)nf;
      if (tid < S) {
        ryy_l[tid] = r_yy;
        rn_l[tid] = r_n;
        rf_l[tid] = r_fl;
        yexp_l[tid] = ytilde_exp(r_yy);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (gq + 16 * k < S) snpin_l[jx * 64 + gq + 16 * k] = (uint8_t)(sin4 >> (8 * k));
      i8_barrier<I8_LGKM0>();
      if (tid < 32) {
        double s2 = 0.0, n2 = 0.0, lo = INFINITY, hi = -INFINITY;
        for (int k = 0; k < 16; ++k) {
          s2 += rs[(0 * 16 + k) * 32 + tid];
          lo = fmin(lo, rs[(1 * 16 + k) * 32 + tid]);
          hi = fmax(hi, rs[(2 * 16 + k) * 32 + tid]);
          n2 += rs[(3 * 16 + k) * 32 + tid];
        }
        const double cj = n2 > 0.0 ? s2 / n2 : 0.0;
        const double dev = n2 > 0.0 ? fmax(hi - cj, cj - lo) : 0.0;
        int kx;
        (void)frexp(fmax(dev, 1.0), &kx);
        const double cs = rint(ldexp(cj, 29 - kx));
        cx_l[tid] = fabs(cs) < 1073741824.0 ? (int)cs : 0;
        ex_l[tid] = 29 - kx;
        bad_l[tid] = fabs(cs) < 1073741824.0 ? 0 : 1; /* centre beyond the fixed-point range */
      }
      i8_barrier<I8_LGKM0>();
    }
    Cx = cx_l[jx];
    sxf = ldexpf(1.f, ex_l[jx]);
    sxd = ldexp(1.0, ex_l[jx]);
    convert(0, xr);
    /* B of step 0 (issued before the genotype loads) and the digit writes */
    i8_barrier<I8_VMCNT8>();

    /* step ck, with P = genotypes of step ck + 1 and Q = those of ck + 2 in
     * registers: copy the B digits of step ck + 1, load the genotypes of
     * step ck + 3 into R, multiply step ck, split P into the other digit
     * buffer (a wave done with its products converts while the other wave
     * of its SIMD still multiplies); one barrier per step.  The loop is
     * unrolled three times so that P, Q, R rotate without register copies
     * (a copy of a register a load is still writing waits for every load). */
    auto step = [&](int ck, const xt (&P)[4], xt (&R)[4]) {
      const bool more = ck + 1 < NCK;
      if (more) issue_b(ck + 1);
      __builtin_amdgcn_sched_barrier(0); /* keep the B copy ahead of the genotype loads */
      load_x(ck + 3, R);
      __builtin_amdgcn_sched_barrier(0);
      const unsigned char* Bs = smem + (ck & 1) * CKB;
      const unsigned char* A = Abase + (ck & 1) * I8_ABUF;
      v4i ax[2][4];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const unsigned char* ap = A + (rb * 16 + l16) * 64 + 16 * (kq ^ sw);
#pragma unroll
        for (int p = 0; p < 4; ++p) ax[rb][p] = *(const v4i*)(ap + p * I8_PLANE);
      }
#pragma unroll
      for (int b = 0; b < NBX; ++b) {
        const unsigned char* bp = Bs + (bxs[b] * 16 + l16) * 64 + 16 * (kq ^ sw);
        v4i bf[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) bf[p] = *(const v4i*)(bp + (size_t)p * CXP * 64);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const v4i* af = ax[rb];
          acc[b][rb][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[0], bf[0], acc[b][rb][0], 0, 0, 0);
          acc[b][rb][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[0], bf[1], acc[b][rb][1], 0, 0, 0);
          acc[b][rb][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[1], bf[0], acc[b][rb][1], 0, 0, 0);
          acc[b][rb][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[0], bf[2], acc[b][rb][2], 0, 0, 0);
          acc[b][rb][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[1], bf[1], acc[b][rb][2], 0, 0, 0);
          acc[b][rb][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[2], bf[0], acc[b][rb][2], 0, 0, 0);
          acc[b][rb][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[0], bf[3], acc[b][rb][3], 0, 0, 0);
          acc[b][rb][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[1], bf[2], acc[b][rb][3], 0, 0, 0);
          acc[b][rb][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[2], bf[1], acc[b][rb][3], 0, 0, 0);
          acc[b][rb][3] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[3], bf[0], acc[b][rb][3], 0, 0, 0);
        }
      }
      {
        const v4i bmf = *(const v4i*)(Bs + (4 * CXP + bms * 16 + l16) * 64 + 16 * (kq ^ sw));
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const unsigned char* ap = A + (rb * 16 + l16) * 64 + 16 * (kq ^ sw) + 4 * I8_PLANE;
#pragma unroll
          for (int p = 0; p < 4; ++p)
            acc[NBX][rb][p] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(*(const v4i*)(ap + p * I8_PLANE), bmf, acc[NBX][rb][p], 0, 0, 0);
        }
      }
      convert(ck + 1, P); /* past the last step: digits nobody reads */
      /* the next step's B copy has landed (only the four genotype loads
       * issued after it may be in flight) and the digit writes are done;
       * every wave is done with the buffers of step ck */
      i8_barrier<I8_VMCNT4>();
    };
    /* after step 0's conversion: xn = step 1, xl = step 2, xr free; whole
     * rounds of three steps in the loop (every path through its body leaves
     * the same loads in flight), the remaining one or two after it */
    int ck = 0;
    for (; ck + 3 <= NCK; ck += 3) {
      step(ck, xn, xr);
      step(ck + 1, xl, xn);
      step(ck + 2, xr, xl);
    }
    if (ck < NCK) step(ck, xn, xr);
    if (ck + 1 < NCK) step(ck + 1, xl, xn);

    /* fold the weight classes (exact in FP64) and stage the sums in LDS:
     * E[row][col], col < CX: x'b; CX + s: x''x' over the members of s */
    double* E = (double*)smem;
    const int ES = q.ES;
#pragma unroll
    for (int b = 0; b < NBX + 1; ++b) {
      const bool xb = b < NBX;
      const int bk = xb ? q.bx[w * I8_NBMAX + (b < NBX ? b : 0)] : q.bm[w];
      if (bk < 0) continue;
      int ecol, eb;
      if (xb) {
        ecol = bk * 16 + l16;
        if (ecol >= q.CX) continue;
        const int s = ecol / q.KC, k = ecol - s * q.KC;
        eb = 24 - (k == 0 ? yexp_l[s] : 30);
      } else {
        const int s = bk * 16 + l16;
        if (s >= S) continue;
        ecol = q.CX + s;
        eb = 30;
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rb * 16 + 4 * kq + r;
          const double v = (double)acc[b][rb][0][r] * 16777216.0 + (double)acc[b][rb][1][r] * 65536.0 +
                           (double)acc[b][rb][2][r] * 256.0 + (double)acc[b][rb][3][r];
          E[row * ES + ecol] = ldexp(v, xb ? eb - ex_l[row] : eb - 2 * ex_l[row]);
        }
    }
    i8_barrier<I8_LGKM0>();
    for (int it = tid; it < I8_ROWS * S; it += I8_NT) {
      const int j = it & 31, s = it >> 5;
      if (j >= cnt) continue;
      const double* e = E + j * ES + s * q.KC;
      double proj = 0.0;
      for (int k = 0; k < q.KU; ++k) proj += e[1 + k] * e[1 + k];
      const double sxx = E[j * ES + q.CX + s];
      const double xx = sxx - proj, sxy = e[0];
      const int64_t o = (int64_t)s * NP + p0 + j;
      q.sx2[o] = xx;
      q.sx2[(int64_t)S * NP + o] = sxy;
      const double yy = ryy_l[s];
      const bool ok = !bad_l[j] && xx > I8_TAU * sxx && yy - sxy * (sxy / xx) > I8_RSS * yy;
      if (!ok && rn_l[s] > 0 && !(rf_l[s] & PREP_TOO_FEW) && snpin_l[j * 64 + s])
        atomicOr(&q.redo[o >> 5], 1u << (o & 31));
    }
    i8_barrier<I8_LGKM0>(); /* E is read; the stores need not land */
  }
}

template <int NB>
void launch_i8(bool xf64, const eqbf_chunk& c, const I8Params& q, int grid, size_t lds, hipStream_t st) {
  if (xf64)
    hipLaunchKernelGGL((uvlr_sums_i8_kernel<NB, true>), dim3(grid), dim3(I8_NT), lds, st, c, q);
  else
    hipLaunchKernelGGL((uvlr_sums_i8_kernel<NB, false>), dim3(grid), dim3(I8_NT), lds, st, c, q);
}

struct I8Shape {
  int NXB, NMB, CXP, SXP, NCK, CKB, ES, NB;
  size_t lds;
};

I8Shape i8_shape(const eqbf_chunk& c, int KU) {
  I8Shape sh;
  const int S = c.n_subgroups, KC = KU + 1, CX = S * KC;
  sh.NXB = (CX + 15) / 16;
  sh.NMB = (S + 15) / 16;
  sh.CXP = sh.NXB * 16;
  sh.SXP = sh.NMB * 16;
  sh.NCK = (c.n_samples + 63) / 64;
  sh.CKB = (4 * sh.CXP + sh.SXP) * 64;
  sh.ES = (CX + S) | 1;
  sh.NB = (sh.NXB + 7) / 8; /* y~/U slots per wave; the x'^2 blocks take one slot each (NMB <= 4) */
  const size_t body = std::max<size_t>((size_t)2 * sh.CKB + 2 * I8_ABUF, (size_t)I8_ROWS * sh.ES * 8);
  sh.lds = body + 64 * (8 + 3 * 4) + I8_ROWS * 3 * 4 + I8_ROWS * 64;
  return sh;
}

}  // namespace

bool uvlr_i8_eligible(const eqbf_chunk& c, int KU) {
  if (c.n_geno_sets != 1 || c.n_samples > 32768 || c.n_samples < 1 || KU > 8) return false;
  const I8Shape sh = i8_shape(c, KU);
  return sh.NB <= 2 && sh.lds <= 160 * 1024; /* three slots per wave spill registers */
}

int uvlr_i8_sums_run(eqbf_ctx* ctx, const eqbf_chunk& c, hipStream_t st, const PrepRec* rec, const double* B,
                     const uint64_t* memb, const int4* tiles, const int* n_tiles, int KU, double* sx2,
                     uint32_t* redo) {
  const int S = c.n_subgroups, G = c.n_genes, N = c.n_samples;
  const I8Shape sh = i8_shape(c, KU);
  uint8_t* B8 = (uint8_t*)ctx_buf(ctx, "sums_b8", (size_t)G * sh.NCK * sh.CKB);
  if (!B8) return EQBF_ERR_HIP;
  {
    ProfScope ps(ctx, "b8_convert_kernel", st);
    hipLaunchKernelGGL(b8_convert_kernel, dim3((unsigned)(G * sh.NCK)), dim3(256), 0, st, rec, B, memb, B8, N, S,
                       KU + 1, S * (KU + 1), sh.CXP, sh.SXP, sh.NCK, sh.CKB);
  }
  EQBF_HIP(ctx, hipGetLastError());
  I8Params q;
  memset(&q, 0, sizeof q);
  q.rec = rec;
  q.B8 = B8;
  q.tiles = tiles;
  q.n_tiles = n_tiles;
  q.sx2 = sx2;
  q.redo = redo;
  q.KC = KU + 1;
  q.KU = KU;
  q.CX = S * (KU + 1);
  q.CXP = sh.CXP;
  q.SXP = sh.SXP;
  q.NXB = sh.NXB;
  q.NCK = sh.NCK;
  q.CKB = sh.CKB;
  q.ES = sh.ES;
  /* column blocks to waves round robin: y~/U block k to wave k % 8, slot
   * k / 8; x'^2 block m to wave 7 - m */
  const int NB = sh.NB;
  for (int k = 0; k < 8 * I8_NBMAX; ++k) q.bx[k] = -1;
  for (int w = 0; w < 8; ++w) q.bm[w] = -1;
  for (int bk = 0; bk < sh.NXB; ++bk) q.bx[(bk % 8) * I8_NBMAX + bk / 8] = (int8_t)bk;
  for (int m = 0; m < sh.NMB; ++m) q.bm[7 - m] = (int8_t)m;
  const int grid = std::max(8, ctx->n_cu / 8 * 8);
  ProfScope ps(ctx, "uvlr_sums_kernel", st);
  const bool xf64 = c.geno_dtype == EQBF_F64;
  if (NB == 1)
    launch_i8<1>(xf64, c, q, grid, sh.lds, st);
  else if (NB == 2)
    launch_i8<2>(xf64, c, q, grid, sh.lds, st);
  else
    launch_i8<3>(xf64, c, q, grid, sh.lds, st);
  EQBF_HIP(ctx, hipGetLastError());
  return EQBF_OK;
}
