/*
 * Internal declarations of libeqbf_hip: the context, device buffers, error
 * plumbing and the device special functions shared by the kernels.
 */
#ifndef EQBF_INTERNAL_H
#define EQBF_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/eqbf.h"

/* ---- context ------------------------------------------------------------ */
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct eqbf_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  std::string err;
  /* grids (host copies; device copy packed as phiL, omaL, phiS, omaS) */
  std::vector<double> phiL, omaL, phiS, omaS;
  DevBuf d_grid;
  /* growable named scratch buffers owned by the context */
  std::map<std::string, DevBuf> bufs;
  /* t-tail tables: built for nu = 1 .. tz_nu_max */
  int tz_nu_max = 0;
  /* device error word written by kernels (EQBF_ERR_DATA conditions) */
  DevBuf d_flag;
  /* bumped whenever a buffer that a kernel argument points to is replaced
   * (ctx_buf growth, new grids): captured graphs older than it are stale */
  uint64_t buf_epoch = 1;
  /* per-call tables kept on the device between calls */
  std::string masks_key;   /* (bfs, S) of the configuration masks uploaded */
  uint64_t jobs_epoch = 0; /* buf_epoch of the uploaded Bayes-factor jobs   */
  size_t jobs_bytes = 0;   /* bytes of the job records before the entries  */
  int jobs_n = 0;
  /* hipGraph replay of the batched association (eqbf_set_graphs) */
  bool graph_on = false;
  bool capturing = false; /* a stream capture is open: enqueue work only */
  struct GraphEntry {
    uint64_t epoch = 0;
    int seen = 0;
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::string, GraphEntry> graphs;
  int64_t graph_launches = 0;
  /* launch-shape autotuner of the batched sums kernel (eqbf_set_autotune) */
  bool autotune_on = false;
  struct TuneChoice {
    int kch = 0, spp = 0, grid_mult = 0;
    float ms = 0.f;
  };
  std::map<std::string, TuneChoice> tuned;
  /* live kernel timing (eqbf_profile) */
  bool prof_on = false;
  struct ProfEntry {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0;
    int64_t n = 0;
  };
  std::map<std::string, ProfEntry> prof;
};

/* Brackets one kernel launch with events when profiling is on. */
struct ProfScope {
  eqbf_ctx* ctx;
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(eqbf_ctx* c, const char* n, hipStream_t s) : ctx(c), name(n), st(s) {
    if (ctx->prof_on && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, st);
  }
  ~ProfScope() {
    if (ctx->prof_on && a && b) {
      (void)hipEventRecord(b, st);
      ctx->prof[name].pending.push_back({a, b});
    }
  }
};

int ctx_fail(eqbf_ctx* ctx, int code, const char* fmt, ...);
#define EQBF_HIP(ctx, call)                                                              \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess)                                                                \
      return ctx_fail((ctx), EQBF_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(_e)); \
  } while (0)

/* Returns a device buffer of at least `bytes` (contents undefined). */
void* ctx_buf(eqbf_ctx* ctx, const std::string& name, size_t bytes);
/* Ensures t-tail tables for nu in [1, nu_max] exist on the device. */
int ctx_ensure_tz(eqbf_ctx* ctx, int nu_max, hipStream_t st);
double* ctx_tz_table(eqbf_ctx* ctx);

/* ---- t-tail tables ---------------------------------------------------------
 * For each integer nu the function w(s) = -ln P(T_nu <= -|t|) is tabulated
 * in s = asinh(|t| / sqrt(nu)), value and derivative, and evaluated by cubic
 * Hermite interpolation.  Two segments of TZ_NODES equally spaced nodes:
 * A = [0, sA] with sA = min(s_max, 8) (the curved part) and B = [sA, s_max]
 * (nearly linear, only for small nu), where w(s_max) > 800 (P underflows).
 * Row layout per nu: [hA, sA, hB, s_max, A nodes (w, dw/ds)..., B nodes...]. */
#define TZ_NODES 4096
#define TZ_ROW (4 + 4 * TZ_NODES)

/* ---- per (gene, subgroup) preparation record ----------------------------- */
struct PrepRec {
  int32_t n;      /* members; 0 = no result in this subgroup               */
  int32_t flags;  /* PREP_TOO_FEW: N < P + 1, every statistic NaN          */
  int32_t q;      /* covariates of the subgroup                            */
  int32_t rank;   /* kept columns of [1, covariates]                       */
  double yy;      /* ||y~||^2, residual of y on [1, covariates]            */
  double tss;     /* sum (y - ybar)^2                                      */
};
#define PREP_TOO_FEW 1
#define FLAG_MISSING_COVAR 1
#define FLAG_MVLR_SAMPLES 2  /* --error mvlr: subgroups with different samples */
#define FLAG_HYB_NO_COMMON 4 /* --error hybrid: two subgroups without a common individual */

/* ---- device special functions (FP64) ------------------------------------- */
namespace dm {

/* Incomplete beta continued fraction (modified Lentz). */
__device__ __host__ inline double betacf(double a, double b, double x) {
  const double tiny = 1e-300, eps = 1e-16;
  double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (fabs(d) < tiny) d = tiny;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 5000; ++m) {
    double m2 = 2.0 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < eps) break;
  }
  return h;
}

/* P(T_nu <= -|t|), FP64, via the regularised incomplete beta
 * I_{nu/(nu+t^2)}(nu/2, 1/2) / 2. */
__device__ __host__ inline double t_lower_neg(double at, double nu) {
  if (isnan(at) || isnan(nu)) return NAN;
  if (isinf(at)) return 0.0;
  double t2 = at * at;
  double x = nu / (nu + t2), y = t2 / (nu + t2);
  double a = 0.5 * nu, b = 0.5;
  if (x <= 0.0) return 0.0;
  if (y <= 0.0) return 0.5;
  double lfront = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log(y);
  double I;
  if (x < (a + 1.0) / (a + b + 2.0))
    I = exp(lfront) * betacf(a, b, x) / a;
  else
    I = 1.0 - exp(lfront) * betacf(b, a, y) / b;
  return 0.5 * I;
}

/* ln P(T_nu <= -|t|) without underflow (the tail branch stays in logs). */
__device__ __host__ inline double log_t_lower_neg(double at, double nu) {
  if (isnan(at) || isnan(nu)) return NAN;
  if (isinf(at)) return -INFINITY;
  double t2 = at * at;
  double x = nu / (nu + t2), y = t2 / (nu + t2);
  double a = 0.5 * nu, b = 0.5;
  if (y <= 0.0) return log(0.5);
  if (x <= 0.0) return -INFINITY;
  double lfront = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log(y);
  if (x < (a + 1.0) / (a + b + 2.0)) return log(0.5) + lfront + log(betacf(a, b, x) / a);
  return log(0.5 * (1.0 - exp(lfront) * betacf(b, a, y) / b));
}

/* Wichura's AS241 (PPND16) rational approximations of the normal quantile.
 * as241_lower(p, w): quantile of p <= 0.5 given w = -ln p (avoids forming
 * p in the tail). */
__device__ __host__ inline double as241_central(double q) {
  double r = 0.180625 - q * q;
  return q *
         (((((((r * 2509.0809287301226727 + 33430.575583588128105) * r + 67265.770927008700853) * r +
              45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
           133.14166789178437745) * r + 3.387132872796366608) /
         (((((((r * 5226.495278852545561 + 28729.085735721942674) * r + 39307.89580009271061) * r +
              21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
           42.313330701600911252) * r + 1.0);
}
__device__ __host__ inline double as241_tail(double r) { /* r = sqrt(-ln p); returns |z| */
  double val;
  if (r <= 5.0) {
    r -= 1.6;
    val = (((((((r * 7.7454501427834140764e-4 + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
               1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
            4.6303378461565452959) * r + 1.42343711074968357734) /
          (((((((r * 1.05075007164441684324e-9 + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
               0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
            2.05319162663775882187) * r + 1.0);
  } else {
    r -= 5.0;
    val = (((((((r * 2.01033439929228813265e-7 + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
               0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
            5.4637849111641143699) * r + 6.6579046435011037772) /
          (((((((r * 2.04426310338993978564e-15 + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
               7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
            0.59983220655588793769) * r + 1.0);
  }
  return val;
}
/* Normal quantile of p in (0, 1]. */
__device__ __host__ inline double gauss_pinv(double p) {
  if (isnan(p)) return NAN;
  if (p <= 0.0) return -INFINITY;
  if (p >= 1.0) return INFINITY;
  double q = p - 0.5;
  if (fabs(q) <= 0.425) return as241_central(q);
  double r = (q < 0) ? p : 1.0 - p;
  double v = as241_tail(sqrt(-log(r)));
  return q < 0 ? -v : v;
}
/* Normal quantile of P = exp(-w) <= 0.5 from w (tail-safe). */
__device__ __host__ inline double gauss_pinv_from_w(double w) {
  if (isnan(w)) return NAN;
  if (isinf(w)) return -INFINITY;
  /* p < 0.075  <=>  w > -ln(0.075) */
  if (w < 2.5902671654458267) return as241_central(exp(-w) - 0.5);
  return -as241_tail(sqrt(w));
}

/* Cubic Hermite evaluation of the tail table of nu at |t|.  Returns
 * w = -ln P(T_nu <= -|t|) (+inf past the table end, NaN for NaN). */
__device__ inline double tz_eval(const double* __restrict__ tz, int nu, double at) {
  if (isnan(at)) return NAN;
  const double* row = tz + (size_t)nu * TZ_ROW;
  double hA = row[0], sA = row[1], hB = row[2], smax = row[3];
  double s = asinh(at * rsqrt((double)nu));
  if (!(s < smax)) return INFINITY;
  double u, h;
  const double* seg;
  if (s < sA) {
    u = s / hA;
    h = hA;
    seg = row + 4;
  } else {
    u = (s - sA) / hB;
    h = hB;
    seg = row + 4 + 2 * TZ_NODES;
  }
  int k = (int)u;
  if (k > TZ_NODES - 2) k = TZ_NODES - 2;
  double f = u - k;
  const double* nd = seg + 2 * k;
  double w0 = nd[0], d0 = nd[1] * h, w1 = nd[2], d1 = nd[3] * h;
  double f2 = f * f, f3 = f2 * f;
  double h00 = 2 * f3 - 3 * f2 + 1, h10 = f3 - 2 * f2 + f, h01 = -2 * f3 + 3 * f2, h11 = f3 - f2;
  return h00 * w0 + h10 * d0 + h01 * w1 + h11 * d1;
}

}  // namespace dm

/* Index of raw ABF (row, p) of n_raw rows and NP pairs: [n_raw][NP], or in
 * tiles of 64 pairs (EQBF_F32_TILED). */
__device__ inline int64_t raw_index(int64_t row, int64_t p, int64_t NP, int64_t n_raw, bool tiled) {
  if (!tiled) return row * NP + p;
  const int64_t tb = p & ~(int64_t)63, tw = NP - tb < 64 ? NP - tb : 64;
  return tb * n_raw + row * tw + (p & 63);
}

/* ---- per (pair, subgroup) summary statistics -------------------------------
 * From the sums x'x, x'y~ and |U'x|^2 over the members:
 * FitSingleGeneWithSingleSnp (utils_math.cpp:166-209) then
 * StandardizeSstatsAndCorrectSmallSampleSize (gene_snp_pair.cpp:256-290).
 * o5: pve, sigmahat, betahat, sebetahat, p; st3: standardised b, se, z.
 * Returns the subgroup's role in the Bayes factors: */
enum { F_ABSENT = 0, F_NORMAL = 1, F_ZERO = 2, F_NAN = 3 };
/* xx: x~'x~ = x'x - |U'x|^2 over the members, NaN when the genotype is
 * collinear with [1, covariates] (not estimable). */
__device__ inline int uvlr_stats_xx(const PrepRec& r, bool has, double xx, double sxy, const double* __restrict__ tz,
                                    double (&o5)[5], double (&st3)[3]) {
  for (int k = 0; k < 5; ++k) o5[k] = NAN;
  for (int k = 0; k < 3; ++k) st3[k] = NAN;
  if (!has) return F_ABSENT;
  if (r.flags & PREP_TOO_FEW) return F_NAN;
  const int n = r.n, P = r.q + 2;
  if (isnan(xx)) { /* genotype collinear with [1, covariates]: not estimable */
    o5[0] = 1.0 - r.yy / r.tss;
    o5[1] = sqrt(r.yy / (double)(n - r.rank));
    return F_NAN;
  }
  const double beta = sxy / xx;
  const double rss = r.yy - sxy * beta;
  const int rank = r.rank + 1;
  const double se = sqrt(rss / (double)(n - P) / xx);
  const double sig = sqrt(rss / (double)(n - rank));
  const double tstat = beta / se;
  const double wv = dm::tz_eval(tz, n - rank, fabs(tstat));
  const double wz = (n - rank == n - P) ? wv : dm::tz_eval(tz, n - P, fabs(tstat));
  o5[0] = 1.0 - rss / r.tss;
  o5[1] = sig;
  o5[2] = beta;
  o5[3] = se;
  o5[4] = 2.0 * exp(-wv);
  const double z = dm::gauss_pinv_from_w(wz);
  const double bhat = beta / sig, sebhat = se / sig;
  if (isnan(bhat / sebhat)) return F_NAN;
  if (fabs(z) > 1e-8) {
    const double sig2 = fabs(beta) / (fabs(z) * sebhat);
    const double b2 = beta / sig2;
    st3[0] = b2;
    st3[1] = fabs(b2 / z);
    st3[2] = z;
    return F_NORMAL;
  }
  st3[0] = 0.0;
  st3[1] = INFINITY;
  st3[2] = z;
  return F_ZERO;
}
/* x~'x~ from the sums, NaN when not estimable (below 1e-12 of x'x). */
__device__ inline double uvlr_xx(double sxx, double proj) {
  const double xx = sxx - proj;
  return xx > 1e-12 * sxx ? xx : NAN;
}
__device__ inline int uvlr_stats(const PrepRec& r, bool has, double sxx, double sxy, double proj,
                                 const double* __restrict__ tz, double (&o5)[5], double (&st3)[3]) {
  return uvlr_stats_xx(r, has, uvlr_xx(sxx, proj), sxy, tz, o5, st3);
}

/* ---- kernels launched from the C-ABI layer ------------------------------- */
/* Bump allocator over a fixed device region (the static scratch array of
 * eqbf_assoc_uvlr_device_noalloc): 256-byte aligned slices, nullptr when
 * the region is exhausted. */
struct ScratchArena {
  uint8_t* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t off = (used + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (off + bytes > cap) return nullptr;
    used = off + bytes;
    return base + off;
  }
};
int assoc_uvlr_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_assoc_opts* o, eqbf_assoc_out* out,
                   hipStream_t st, ScratchArena* arena = nullptr);
/* Address of the static scratch array on the current device (set up once
 * per device by eqbf_create). */
int assoc_noalloc_setup(eqbf_ctx* ctx);
/* packed work list of 64-pair tiles of every gene (one block scan) */
void launch_tiles(const int64_t* pair_offset, int G, int4* tiles, int* n_tiles, hipStream_t st);
void launch_zero_u64(uint64_t* p, int64_t n, hipStream_t st);
/* batched mixed-precision uvlr path (one genotype set, matrix-core sums) */
bool uvlr_fast_eligible(const eqbf_chunk& c, int KU);
int uvlr_fast_run(eqbf_ctx* ctx, const eqbf_chunk& c, const eqbf_assoc_opts* o, eqbf_assoc_out* out, hipStream_t st,
                  const PrepRec* rec, const double* B, const uint64_t* memb, const int32_t* idx, const int4* tiles,
                  const int* n_tiles, int KU, const uint64_t* d_masks, int n_cfg);
/* int8-digit sums of the batched path (eqbf_uvlr_i8.hip): x~'x~ and x'y~
 * into sx2, and a bit in redo for each item to recompute in FP64 */
bool uvlr_i8_eligible(const eqbf_chunk& c, int KU);
int uvlr_i8_sums_run(eqbf_ctx* ctx, const eqbf_chunk& c, hipStream_t st, const PrepRec* rec, const double* B,
                     const uint64_t* memb, const int4* tiles, const int* n_tiles, int KU, double* sx2,
                     uint32_t* redo);
int mvlr_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_mvlr_opts* o, eqbf_assoc_out* out, hipStream_t st);
int hybrid_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_hybrid_opts* o, eqbf_assoc_out* out, hipStream_t st);
/* eqtlbma_avg_bfs kernels (device pointers; weight tables already staged) */
int avg_bfs_run(eqbf_ctx* ctx, const double* raw, int64_t NP, const int64_t* off, int64_t G, const eqbf_avg_opts* o,
                const double* d_gw, const int32_t* d_keep, const double* d_dw, const uint64_t* d_cs,
                const double* d_sw, eqbf_avg_out* out, hipStream_t st);
int avg_grid_run(eqbf_ctx* ctx, const double* raw, int64_t n_rows, int L, const double* gw, const int32_t* keep,
                 int nk, double* out, hipStream_t st);
/* --inss: Bayes factors from given summary statistics (device pointers) */
int abf_from_sstats_run(eqbf_ctx* ctx, const eqbf_sstats_in* in, int bfs, int raw_dtype, eqbf_assoc_out* out,
                        hipStream_t st);
/* gen-sin and all rows of out->w_abf from the grid-averaged rows. */
int launch_bma(eqbf_ctx* ctx, double* w, int64_t n_pairs, int S, int bfs, hipStream_t st);
int perm_stats_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po, const int32_t* tables_dev,
                   int32_t n_tables, const int32_t* gene_table, int64_t nperm, double* stats_dev, hipStream_t st);
/* perm_stats_run, then for every gene with a permutation statistic within
 * tie_bound of its true statistic (true_dev, the caller's): the true
 * statistic in FP64 (overwriting true_dev[g]) and those permutation
 * statistics in FP64 */
int perm_ties_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po, const int32_t* tables_dev,
                  int32_t n_tables, const int32_t* gene_table, int64_t nperm, double tie_bound, double* stats_dev,
                  double* true_dev, int64_t* n_rechecked, hipStream_t st);
/* contraction path for the genes with full membership; done[g] = 1 for the
 * genes it finished (the others go to the general path) */
int perm_fast_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po, const int32_t* tables_dev,
                  int32_t n_tables, const int32_t* gene_table, int64_t nperm, double* stats_dev, hipStream_t st,
                  std::vector<uint8_t>& done);

#endif
