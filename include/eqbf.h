/*
 * eqbf.h -- C ABI of the MI355X Bayes-factor engine behind eqtlbma_bf.
 *
 * The reference (eqtlbma 1.3.3) has no plugin API: its only public surface is
 * the eqtlbma_bf command line.  The seam this library replaces is the pair of
 * per-chunk calls inside run():
 *   testForAssociations(...)   src/eqtlbma_bf.cpp:707-771   (called :1551-1555)
 *   makePermutations(...)      src/eqtlbma_bf.cpp:867-917   (called :1558-1561)
 * and the GeneSnpPair / Gene getters the writers read afterwards
 * (src/eqtlbma/quantgen/gene_snp_pair.cpp:1425-1470, gene.cpp:719-725).
 *
 * Conventions
 *  - Every function returns EQBF_OK (0) or an error code; the message is in
 *    eqbf_last_error(ctx).  The reference prints to stderr and exits with
 *    EXIT_FAILURE instead (utils_io.cpp:214-219); the host program maps a
 *    non-zero code to exactly that.
 *  - Numerical failure is data, not an error: it comes back as NaN, as in the
 *    reference (utils_math.cpp:175-178).
 *  - The caller owns every buffer it passes.  A context owns its device memory,
 *    its stream and its cached tables.  One context per host thread.
 *  - "_device" entry points take device pointers (already resident in HBM) and
 *    enqueue on the given stream (NULL = the context's stream) without
 *    synchronising; the other entry points take host pointers, stage them and
 *    return when the results are back in host memory.
 *  - All per-pair outputs are structure-of-arrays with leading dimension
 *    n_pairs, so that consecutive pairs are consecutive in memory.
 *
 * Layout of the inputs (eqbf_chunk)
 *  - subgroups are indexed 0..S-1 in lexicographic name order and samples
 *    0..N-1 in lexicographic name order, as the reference orders them
 *    (data_loader.cpp:141-143, samples.cpp:67).
 *  - explevels[g][s][i] is the expression level of gene g in subgroup s for
 *    sample i, NaN when the sample has no column in that subgroup's expression
 *    file or the value is NA (gene.cpp:104-118).
 *  - genotypes: element (set, i, j) is at genotypes[set*geno_set_stride +
 *    i*geno_ld + j] for sample i and SNP j; subgroup s reads set
 *    geno_set_of_sbgrp[s] (subgroups sharing one genotype file share a set,
 *    data_loader.cpp:691-715).
 *  - the pairs of gene g are pair_offset[g] .. pair_offset[g+1]-1 and pair p
 *    tests SNP pair_snp[p] (cis SNPs in position order, gene.cpp:140-157).
 */
#ifndef EQBF_H
#define EQBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define EQBF_OK               0
#define EQBF_ERR_ARG          1  /* invalid argument or size                  */
#define EQBF_ERR_HIP          2  /* HIP runtime failure                       */
#define EQBF_ERR_DATA         3  /* input the reference rejects with exit(1)  */
#define EQBF_ERR_UNSUPPORTED  4  /* valid option this build does not run      */

/* ---- option values (eqtlbma_bf.cpp:262-299, defaults :1586-1597) ------- */
enum { EQBF_ANALYSIS_SEP = 0, EQBF_ANALYSIS_JOIN = 1 };         /* --analys */
enum { EQBF_BFS_GEN = 0, EQBF_BFS_SIN = 1, EQBF_BFS_ALL = 2 };  /* --bfs    */
enum { EQBF_PBF_GEN = 0, EQBF_PBF_GENSIN = 1, EQBF_PBF_ALL = 2 };/* --pbf   */
/* precision / batching mode:
 *  FP64_PLAIN:   every quantity in FP64, one kernel launch per gene, reference
 *                loop structure (the "plain GPU" comparison path);
 *  MIXED:        batched persistent launch over all genes of the chunk, FP64
 *                sums and residual sums of squares, FP32 Bayes-factor terms,
 *                FP64 log10-sum-exp;
 *  FP64_BATCHED: the FP64_PLAIN arithmetic, bit for bit, with every gene of
 *                the call in one launch. */
enum { EQBF_PREC_FP64_PLAIN = 0, EQBF_PREC_MIXED = 1, EQBF_PREC_FP64_BATCHED = 2 };
enum { EQBF_F32 = 0, EQBF_F64 = 1 };
/* raw_dtype only: FP32 raw ABFs in tiles of 64 pairs (no counterpart in the
 * reference, whose raw ABFs live in per-pair maps): element (row, p) at
 * (p & ~63) * n_raw + row * w + (p & 63), w = min(64, n_pairs - (p & ~63)),
 * so that one workgroup's 64 pairs write one contiguous block. */
enum { EQBF_F32_TILED = 2 };

/* ---- opaque context ------------------------------------------------------ */
typedef struct eqbf_ctx eqbf_ctx;

/* Create a context on HIP device `device` (ordinal). */
int eqbf_create(eqbf_ctx** out, int device);
void eqbf_destroy(eqbf_ctx* ctx);
const char* eqbf_last_error(const eqbf_ctx* ctx);
/* Block until every operation enqueued on the context's stream is done. */
int eqbf_sync(eqbf_ctx* ctx);

/* Grids of prior variances (phi2, oma2); replaces Grid::Grid (grid.cpp:28-67).
 * The large grid is used for the consistent configuration (gen, gen-fix,
 * gen-maxh), the small one for singletons / all configurations. */
int eqbf_set_grids(eqbf_ctx* ctx,
                   const double* phi2_large, const double* oma2_large, int32_t n_large,
                   const double* phi2_small, const double* oma2_small, int32_t n_small);

/* ---- inputs ----------------------------------------------------------------*/
typedef struct eqbf_chunk {
  int32_t n_genes;
  int32_t n_subgroups;          /* S  (<= 64)                                 */
  int32_t n_samples;            /* N  (union of all samples)                  */
  int32_t n_covars_max;         /* Q_max covariate slots per subgroup (<= 16) */
  int64_t n_pairs;
  int64_t n_snps;               /* columns of the genotype matrices           */
  const double*  explevels;     /* [n_genes][S][N]                            */
  const uint8_t* gene_in_sbgrp; /* [n_genes][S] gene has expression in s      */
  const uint8_t* geno_in_sbgrp; /* [S][N] sample has a genotype column in s   */
  const int32_t* covar_count;   /* [S] number of covariates of s              */
  const double*  covars;        /* [S][Q_max][N]                              */
  const uint8_t* covar_in_sbgrp;/* [S][N] sample has covariate values in s    */
  int32_t        n_geno_sets;
  const int32_t* geno_set_of_sbgrp; /* [S]                                    */
  const void*    genotypes;     /* see layout note above                      */
  int32_t        geno_dtype;    /* EQBF_F32 or EQBF_F64                       */
  int64_t        geno_ld;       /* >= n_snps                                  */
  int64_t        geno_set_stride;
  const uint8_t* snp_in_sbgrp;  /* [n_snps][S] SNP kept in s (no NA, MAF ok)  */
  const int64_t* pair_offset;   /* [n_genes+1]                                */
  const int32_t* pair_snp;      /* [n_pairs]                                  */
} eqbf_chunk;

typedef struct eqbf_assoc_opts {
  int32_t analysis;    /* EQBF_ANALYSIS_*                                     */
  int32_t bfs;         /* EQBF_BFS_*  (join only)                             */
  int32_t precision;   /* EQBF_PREC_*                                         */
  int32_t raw_dtype;   /* out->raw_abf: EQBF_F32 / EQBF_F64 / EQBF_F32_TILED  */
} eqbf_assoc_opts;

/* Number of configurations written after the consistent ones:
 * gen -> 0, sin -> S, all -> 2^S - 1 (gsl_combination order: size 1..S,
 * lexicographic within a size; gene_snp_pair.cpp:469-550). */
int64_t eqbf_n_configs(int32_t bfs, int32_t n_subgroups);
/* Bit masks (bit s = subgroup s active) of those configurations, in order. */
int eqbf_config_masks(int32_t bfs, int32_t n_subgroups, uint64_t* masks);
/* Raw Bayes factors per pair: 3*n_large (gen, gen-fix, gen-maxh) followed by
 * n_configs*n_small. */
int64_t eqbf_n_raw(int32_t bfs, int32_t n_subgroups, int32_t n_large, int32_t n_small);
/* Grid-averaged Bayes factors per pair: gen, gen-fix, gen-maxh, then one per
 * configuration, then gen-sin (BMAlite) and all (BMA); NaN when not computed. */
int64_t eqbf_n_weighted(int32_t bfs, int32_t n_subgroups);

/* ---- outputs (structure of arrays, leading dimension n_pairs) ------------- */
typedef struct eqbf_assoc_out {
  int32_t* n;         /* [S][n_pairs] sample size, 0 = no result in s
                         (GeneSnpPair::GetSampleSize, gene_snp_pair.cpp:1425) */
  double*  sstats;    /* [5][S][n_pairs] pve, sigmahat, betahat.geno,
                         sebetahat.geno, betapval.geno (:1430-1453); may be NULL */
  double*  stdsstats; /* [3][S][n_pairs] standardised b, se, t after the
                         small-sample correction (:256-290); may be NULL       */
  void*    raw_abf;   /* [n_raw][n_pairs] log10 ABFs (:1455-1465); may be NULL */
  double*  w_abf;     /* [n_weighted][n_pairs] (:1467-1470); join only         */
} eqbf_assoc_out;

/* Association test of every pair of the chunk with the uvlr error model:
 * per-subgroup OLS summary statistics (GeneSnpPair::CalcSstatsOneSbgrp,
 * gene_snp_pair.cpp:175-239 -> utils::FitSingleGeneWithSingleSnp,
 * utils_math.cpp:166-209), and for --analys join the Bayes factors
 * (GeneSnpPair::CalcAbfsUvlr, gene_snp_pair.cpp:604-622).
 * Replaces testForAssociations, eqtlbma_bf.cpp:707-771, for uvlr/sep. */
int eqbf_assoc_uvlr(eqbf_ctx* ctx, const eqbf_chunk* chunk,
                    const eqbf_assoc_opts* opts, eqbf_assoc_out* out);
int eqbf_assoc_uvlr_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev,
                           const eqbf_assoc_opts* opts, eqbf_assoc_out* out_dev,
                           void* hip_stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  eqbf_assoc_uvlr_device with an FP64 precision
 * (FP64_PLAIN or FP64_BATCHED, run as FP64_BATCHED: the same results bit for
 * bit) whose per-call scratch comes from one static device array of 64 MiB
 * instead of context buffers: no hipMalloc / hipFree, no host copies, no
 * synchronisation.  The scratch of a call is about
 *   n_genes * (n_samples * (8 * S * (1 + KU) + 4 * S + 8) + 32 * S) + 8 * n_cfg
 * bytes (KU = 1, 4, 8, 16 or 17 basis columns for 0, <= 3, <= 7, <= 15, 16
 * covariates); a call that needs more than the array holds fails with
 * EQBF_ERR_ARG (split the chunk).  Every call on one device shares the
 * array, which keeps the last call's contents: calls that overlap on one
 * device (two streams, two contexts, two threads) race.  The grids must be
 * set (eqbf_set_grids) beforehand; other precisions give
 * EQBF_ERR_UNSUPPORTED. */
int eqbf_assoc_uvlr_device_noalloc(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, const eqbf_assoc_opts* opts,
                                   eqbf_assoc_out* out_dev, void* hip_stream);

/* ---- multivariate (MVLR) Bayes factors ------------------------------------------
 * --error mvlr: GeneSnpPair::CalcAbfsMvlr (gene_snp_pair.cpp:730-758) and
 * class MVLR (MVLR.cpp:26-764) for one SNP per pair.  Every subgroup must use
 * the same samples (the reference requires one genotype file and reads the
 * phenotype vectors by position, data_loader.cpp:133-140, MVLR.cpp:29-41);
 * the covariates are those of the first subgroup (gene_snp_pair.cpp:120-151).
 * Genes without expression in every subgroup and SNPs without genotypes in
 * every subgroup are skipped (n = 0, NaN Bayes factors), as gene.cpp:314-321
 * and eqtlbma_bf.cpp:755-762 do.  Computed in FP64.  Outputs use the layout
 * of eqbf_assoc_out with raw_abf in FP64 (required, like n and w_abf); as in
 * the reference, n holds the sample size of each subgroup and the optional
 * sstats / stdsstats are NaN (gene_snp_pair.cpp:154-158).  Up to 16
 * subgroups and 16 covariates.  Errors: EQBF_ERR_UNSUPPORTED when the
 * subgroups of an analysed pair have different samples (reported by the
 * host-pointer entry), EQBF_ERR_DATA for a missing covariate. */
typedef struct eqbf_mvlr_opts {
  int32_t bfs;       /* EQBF_BFS_*                                        */
  double fiterr;     /* --fiterr, proportion of the ML error covariance   */
} eqbf_mvlr_opts;
int eqbf_assoc_mvlr(eqbf_ctx* ctx, const eqbf_chunk* chunk, const eqbf_mvlr_opts* opts, eqbf_assoc_out* out);
int eqbf_assoc_mvlr_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, const eqbf_mvlr_opts* opts,
                           eqbf_assoc_out* out_dev, void* hip_stream);

/* ---- hybrid error model (--error hybrid) -----------------------------------------
 * GeneSnpPair::CalcAbfsHybrid (gene_snp_pair.cpp:1384-1423): per-subgroup
 * regressions give the effects and the diagonals of the error and effect
 * covariances (CalcBetahatsAndDiagsPerSubgroup :760-880), each pair of
 * subgroups the off-diagonal terms over the individuals present in both and
 * in either (CalcOffDiagCovarsFromPairsOfSubgroups :1035-1131), and the Bayes
 * factors take the MVLR form (CalcLog10AbfMvlr :1165-1256).  Subgroups may
 * have different samples (expression levels missing per subgroup).  Genes
 * without expression in every subgroup and SNPs without genotypes in every
 * subgroup are skipped (gene.cpp:314-321): n = 0, NaN statistics and Bayes
 * factors.  n and sstats are those of the separate analysis, as the
 * reference records them (:835-870).  Computed in FP64; raw_abf in FP64,
 * required like n and w_abf.  Up to 16 subgroups and 16 covariates.
 * EQBF_ERR_DATA when two subgroups of an analysed pair share no individual
 * (the reference exits, :895-900) or a member lacks a covariate. */
typedef struct eqbf_hybrid_opts {
  int32_t bfs;                       /* EQBF_BFS_*                                      */
  double fiterr;                     /* --fiterr                                        */
  /* covariates of the pairwise designs [S][Qmax][N]: the reference reads the
   * first subgroup's covariate at the individual's index in the union of
   * samples, not at its column of the covariate file (:920-927); NULL: the
   * chunk's covariates (the same when the files list the samples in the
   * same order) */
  const double* covars_by_position;
  /* expression levels of the pairwise step [G][S][N]: the reference reads
   * them unnormalised under --qnorm (Gene::GetExplevel, :908-915); NULL: the
   * chunk's */
  const double* explevels_pairs;
} eqbf_hybrid_opts;
int eqbf_assoc_hybrid(eqbf_ctx* ctx, const eqbf_chunk* chunk, const eqbf_hybrid_opts* opts, eqbf_assoc_out* out);
/* Device-pointer variant: chunk, opts' arrays and outputs on the device. */
int eqbf_assoc_hybrid_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, const eqbf_hybrid_opts* opts,
                             eqbf_assoc_out* out_dev, void* hip_stream);

/* ---- summary statistics as input (--inss) ----------------------------------------
 * Bayes factors of pairs whose summary statistics are given rather than
 * computed: replaces, for the pairs loadSummaryStats / fillGeneSnpPairsWith-
 * Sstats build (data_loader.cpp:1250-1343), GeneSnpPair::SetSstats followed
 * by CalcAbfsUvlr (gene_snp_pair.cpp:241-254, 604-622) inside
 * Gene::TestForAssociations (gene.cpp:263-332).  Inputs are structure of
 * arrays [S][n_pairs]; n <= 0 means no result in that subgroup (HasResults,
 * gene_snp_pair.cpp:61-68).  The small-sample correction uses nu = n - 2: no
 * covariate count is known for given statistics (:273).  n <= 2 (nu <= 0)
 * gives a NaN z.  FP64 throughout, one thread per pair.  Outputs as in
 * eqbf_assoc_uvlr: n (optional, int32), stdsstats (optional), raw_abf (FP64
 * or FP32 as raw_dtype, optional), w_abf (required); sstats is not written.
 * At most 64 subgroups and grids of at most 128 points. */
typedef struct eqbf_sstats_in {
  int64_t n_pairs;
  int32_t n_subgroups;
  const int64_t* n;         /* [S][n_pairs] sample size ("n" column)          */
  const double*  sigmahat;  /* [S][n_pairs] "sigmahat"                        */
  const double*  betahat;   /* [S][n_pairs] "betahat.geno"                    */
  const double*  sebetahat; /* [S][n_pairs] "sebetahat.geno"                  */
} eqbf_sstats_in;
int eqbf_abf_from_sstats(eqbf_ctx* ctx, const eqbf_sstats_in* in, int32_t bfs, int32_t raw_dtype,
                         eqbf_assoc_out* out);
int eqbf_abf_from_sstats_device(eqbf_ctx* ctx, const eqbf_sstats_in* in_dev, int32_t bfs, int32_t raw_dtype,
                                eqbf_assoc_out* out_dev, void* hip_stream);

/* ---- permutations ------------------------------------------------------------
 * Replaces makePermutations (eqtlbma_bf.cpp:867-917) and the permutation
 * loops of Gene::MakePermutationsJoin (gene.cpp:598-717),
 * MakePermutationsSepAllSubgroups (gene.cpp:493-570) and
 * MakePermutationsSepPerSubgroup (gene.cpp:380-450).
 *
 * The permutation indices are made on the host exactly as the reference
 * draws them (GSL mt19937, gsl_ran_shuffle; eqbf_perm_tables below); the
 * device computes the test statistic of every permutation; the host replays
 * the reference's counting loop (eqbf_perm_pvalue) so that the --trick
 * rules and the NaN bookkeeping follow the reference step by step. */
typedef struct eqbf_perm_opts {
  int32_t analysis;    /* EQBF_ANALYSIS_JOIN: Bayes-factor statistic;
                          EQBF_ANALYSIS_SEP: minimum p-value               */
  int32_t pbf;         /* join: EQBF_PBF_* (--pbf)                         */
  int32_t use_max_bf;  /* join: 1 = max over SNPs (--maxbf), 0 = average   */
  int32_t sep_sbgrp;   /* sep: subgroup (--permsep 2) or -1 (--permsep 1)  */
  int32_t precision;   /* EQBF_PREC_*                                      */
  /* join: EQBF_ERROR_UVLR (0) or EQBF_ERROR_MVLR (--error mvlr,
   * gene.cpp:662-672): every permutation's Bayes factors come from the MVLR
   * engine (FP64, eqbf_assoc_mvlr) on the permuted levels; a SNP without
   * genotypes in every subgroup contributes 0 to the statistic (its initial
   * value in the reference's loop, gene.cpp:650, 663-664); the subgroups of
   * a gene must keep the same samples under every permutation. */
  int32_t error_model;
  double fiterr;       /* --fiterr with EQBF_ERROR_MVLR / EQBF_ERROR_HYBRID */
  /* EQBF_ERROR_HYBRID (--error hybrid, gene.cpp:669-672): every
   * permutation's Bayes factors come from the hybrid engine
   * (eqbf_assoc_hybrid): the per-subgroup fits see the permuted levels, the
   * pairwise step the levels as loaded, and its sets of individuals from
   * them (CalcOffDiagCovarsFromPairsOfSubgroups takes no permutation,
   * gene_snp_pair.cpp:1157-1159, samples.cpp:148-192); a SNP without
   * genotypes in every subgroup contributes 0 as with EQBF_ERROR_MVLR.
   * covars_by_position and explevels_pairs as in eqbf_hybrid_opts (NULL:
   * the chunk's); host arrays for the host-pointer entries, device arrays
   * for the _device ones. */
  const double* covars_by_position; /* [S][Qmax][N]                          */
  const double* explevels_pairs;    /* [G][S][N]                             */
} eqbf_perm_opts;
enum { EQBF_ERROR_UVLR = 0, EQBF_ERROR_MVLR = 1, EQBF_ERROR_HYBRID = 2 };

/* Statistic of every permutation of every gene of the chunk.
 * perm_idx:   [n_tables][nperm][n_samples]; sample i of permutation p of
 *             table t takes the expression level of sample perm_idx[t][p][i]
 *             (gene_snp_pair.cpp:104-115; genotypes and covariates stay with i)
 * gene_table: [n_genes] table used by each gene, -1 to skip the gene
 * stats:      [n_genes][nperm] output; NaN where the reference's statistic
 *             is NaN (those permutations decrement nb_permutations). */
int eqbf_perm_stats(eqbf_ctx* ctx, const eqbf_chunk* chunk, const eqbf_perm_opts* opts,
                    const int32_t* perm_idx, int32_t n_tables, const int32_t* gene_table,
                    int64_t nperm, double* stats);
int eqbf_perm_stats_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, const eqbf_perm_opts* opts,
                           const int32_t* perm_idx_dev, int32_t n_tables, const int32_t* gene_table,
                           int64_t nperm, double* stats_dev, void* hip_stream);

/* Near-ties in FP64.  The reference computes every statistic in FP64 and
 * counts a permutation when its statistic is >= the true one (join) or <= it
 * (sep) (gene.cpp:685, 697, 435, 555).  With the MIXED precision a statistic
 * carries a small error (about 1e-6 in log10 BF), so a permutation whose
 * statistic lies that close to the true one could be counted differently.
 * This entry point is eqbf_perm_stats followed by, for every gene g with a
 * permutation statistic within tie_bound of true_stats[g] (absolute for
 * join, tie_bound * |true| for sep; tie_bound must exceed the error of the
 * statistic plus that of the caller's true statistic):
 *  - true_stats[g] replaced by the true statistic in FP64 (the FP64_PLAIN
 *    arithmetic) with the reference's semantics: Gene::FindMaxTrueL10Abf /
 *    AvgTrueL10Abfs (gene.cpp:575-596: NaN never wins the maximum, the
 *    average runs over the non-NaN SNPs) or FindMinTruePvalue*
 *    (gene.cpp:369-378, 475-488: starts from 1);
 *  - those permutation statistics recomputed in FP64.
 * true_stats (in/out, [n_genes]): on entry the caller's true statistics,
 * as its counting loop would use them (e.g. from eqbf_assoc_uvlr's
 * grid-averaged Bayes factors); the other genes keep them.  The counting
 * loop (eqbf_perm_pvalue) then compares FP64 values wherever the precision
 * could decide the count.  n_rechecked (may be NULL): permutation
 * statistics recomputed.  With an FP64 precision nothing is rechecked.
 * No counterpart in the reference (which has no mixed precision). */
int eqbf_perm_stats_fp64_ties(eqbf_ctx* ctx, const eqbf_chunk* chunk, const eqbf_perm_opts* opts,
                              const int32_t* perm_idx, int32_t n_tables, const int32_t* gene_table,
                              int64_t nperm, double tie_bound, double* stats, double* true_stats,
                              int64_t* n_rechecked);
int eqbf_perm_stats_fp64_ties_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, const eqbf_perm_opts* opts,
                                     const int32_t* perm_idx_dev, int32_t n_tables, const int32_t* gene_table,
                                     int64_t nperm, double tie_bound, double* stats_dev, double* true_stats_dev,
                                     int64_t* n_rechecked, void* hip_stream);

/* ---- host random numbers with GSL's semantics ----------------------------------
 * gsl_rng_mt19937 + gsl_rng_set / gsl_rng_uniform_int / gsl_ran_shuffle /
 * gsl_ran_flat as called at eqtlbma_bf.cpp:889-916, gene.cpp:358-362,410,
 * 523,639.  GSL_RNG_TYPE is not consulted: the generator is mt19937, the
 * reference's default (gsl_rng_env_setup, eqtlbma_bf.cpp:890). */
typedef struct eqbf_rng {
  uint32_t mt[624];
  int32_t idx;
} eqbf_rng;
void eqbf_rng_seed(eqbf_rng* r, uint64_t seed);
uint32_t eqbf_rng_get(eqbf_rng* r);
uint32_t eqbf_rng_uniform_int(eqbf_rng* r, uint32_t n);
double eqbf_rng_flat(eqbf_rng* r, double a, double b);
void eqbf_rng_shuffle(eqbf_rng* r, int32_t* a, int64_t n);
/* Index tables of n_genes analysed genes continuing the stream r: each gene
 * starts from the identity and takes nperm cumulative shuffles
 * (gene.cpp:617,638-639).  out: [n_genes][nperm][n_samples]. */
void eqbf_perm_tables(eqbf_rng* r, int64_t n_samples, int64_t n_genes, int64_t nperm, int32_t* out);

/* The reference's counting loop over one gene's permutation statistics
 * (gene.cpp:637-716 join, :409-449 / :522-569 sep) and
 * CalcPermutationPvalue (gene.cpp:348-364).  A permutation counts when its
 * statistic is >= the true one (join) or <= (sep).  nb_permutations is the
 * caller's shared counter, decremented on NaN statistics as the reference
 * does; rng_trick may be NULL when trick == 0.  Returns the number of
 * permutations drawn.  median (join only, may be NULL) is taken over the
 * valid statistics: the reference reads one element past them
 * (gene.cpp:713-714). */
int64_t eqbf_perm_pvalue(const double* stats, double true_stat, int32_t sep, int64_t* nb_permutations,
                         int32_t trick, int64_t trick_cutoff, eqbf_rng* rng_trick, double* pval,
                         int64_t* nperm_used, double* median);

/* ---- live kernel timing -------------------------------------------------------
 * No counterpart in the reference.  When enabled, every kernel launch of the
 * context is bracketed by HIP events on the stream it is launched on; the
 * totals per kernel name are read back with eqbf_profile_read (which
 * synchronises the context's streams).  names: n_max entries of 64 chars. */
int eqbf_profile(eqbf_ctx* ctx, int enable);
int eqbf_profile_read(eqbf_ctx* ctx, char* names, double* total_ms, int64_t* launches, int32_t n_max,
                      int32_t* n_out);

/* ---- launch graphs and launch-shape tuning -------------------------------------
 * No counterpart in the reference.
 * eqbf_set_graphs: when enabled, eqbf_assoc_uvlr_device calls with the mixed
 * precision are replayed from a hipGraph: the first call with a given set of
 * arguments (chunk, options, output pointers, stream) runs normally, the
 * second is captured and launched, later ones launch the graph.  The inputs'
 * contents may change between calls; their addresses and sizes may not
 * (other arguments are another graph).  Ignored while eqbf_profile is on.
 * eqbf_set_autotune: when enabled, the first batched association of each
 * shape (subgroups, coefficient columns, samples) times the candidate launch
 * shapes of the sums kernel on that call's own data and keeps the fastest;
 * results do not depend on the shape.  eqbf_autotune_read lists the choices
 * (keys: n_max entries of 64 chars). */
int eqbf_set_graphs(eqbf_ctx* ctx, int enable);
/* graphs currently held and graph launches so far */
int eqbf_graph_stats(eqbf_ctx* ctx, int64_t* n_graphs, int64_t* n_launches);
int eqbf_set_autotune(eqbf_ctx* ctx, int enable);
int eqbf_autotune_read(eqbf_ctx* ctx, char* keys, int32_t* kch, int32_t* spp, int32_t* grid_mult, float* ms,
                       int32_t n_max, int32_t* n_out);

/* ---- diagnostics ------------------------------------------------------------
 * No counterpart in the reference.  Evaluates, on the device, the Student-t
 * tail w = -ln P(T_nu <= -|t|) from the tabulated interpolant (w_tab) and
 * from the direct continued fraction (w_cf), and the corresponding normal
 * quantiles z (as GSL's gsl_cdf_gaussian_Pinv(gsl_cdf_tdist_P(-|t|, nu)),
 * gene_snp_pair.cpp:274).  Host arrays of length n. */
int eqbf_debug_ttail(eqbf_ctx* ctx, const double* t, const int32_t* nu, int32_t n,
                     double* w_tab, double* w_cf, double* z_tab, double* z_cf);

/* ---- learned permutation allocator: its network ------------------------------
 * No counterpart in the reference (eqtlbma_bf gives every gene the same
 * --nperm).  A multilayer perceptron n_in -> n_hidden (n_layers, GELU) ->
 * n_out with BF16 weights, an FP32 master copy and FP32 accumulation,
 * trained with Adam on the device.  It predicts the null distribution of a
 * gene's permutation statistic (GEV location, log scale, shape) from the
 * gene's features; the host decides how many permutations each gene gets.
 * Parameters are one flat FP32 vector: per layer W[out][in] then b[out].
 * All arrays are host arrays; the calls synchronise. */
typedef struct eqbf_mlp eqbf_mlp;
int eqbf_mlp_create(eqbf_ctx* ctx, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                    uint64_t seed, eqbf_mlp** out);
void eqbf_mlp_destroy(eqbf_mlp* mlp);
int64_t eqbf_mlp_n_params(const eqbf_mlp* mlp);
int eqbf_mlp_set_params(eqbf_ctx* ctx, eqbf_mlp* mlp, const float* params);
int eqbf_mlp_get_params(eqbf_ctx* ctx, eqbf_mlp* mlp, float* params);
/* y[B][n_out] = network(x[B][n_in]) */
int eqbf_mlp_forward(eqbf_ctx* ctx, eqbf_mlp* mlp, const float* x, int64_t B, float* y);
/* loss = sum_b sum_j w_j (y_bj - target_bj)^2 / B and its gradient with
 * respect to every parameter (grad: n_params floats, may be NULL) */
int eqbf_mlp_loss_grad(eqbf_ctx* ctx, eqbf_mlp* mlp, const float* x, const float* target,
                       const float* out_weight, int64_t B, double* loss, float* grad);
/* epochs of minibatch Adam over the n_rows examples (shuffled per epoch with
 * the library's mt19937 seeded by seed); epoch_loss: mean minibatch loss of
 * each epoch (may be NULL) */
int eqbf_mlp_train(eqbf_ctx* ctx, eqbf_mlp* mlp, const float* x, const float* target, const float* out_weight,
                   int64_t n_rows, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* epoch_loss);

/* Inputs and targets of the network (no counterpart in the reference).
 * eqbf_ld_eigenvalues: the n_eig (<= 28) largest eigenvalues of the
 * correlation matrix of each gene's cis genotypes over the chunk's samples
 * (N <= 512), descending; eig: host [n_genes][n_eig].  The _device variant
 * takes a chunk whose pointers are device addresses.
 * eqbf_gev_fit: location, scale and shape (xi; GEV tail
 * 1 - exp(-(1 + xi (x - mu) / sigma)^(-1/xi))) of each gene's permutation
 * statistics stats[n_genes][nperm] (host, NaN draws ignored, nperm <= 16384)
 * by probability-weighted moments; params: host [n_genes][3]. */
int eqbf_ld_eigenvalues(eqbf_ctx* ctx, const eqbf_chunk* chunk, int32_t n_eig, float* eig);
int eqbf_ld_eigenvalues_device(eqbf_ctx* ctx, const eqbf_chunk* chunk_dev, int32_t n_eig, float* eig);
int eqbf_gev_fit(eqbf_ctx* ctx, const double* stats, int64_t n_genes, int64_t nperm, double* params);

/* ---- eqtlbma_hm: EM of the hierarchical model on raw log10 ABFs --------
 * Replaces Controller::run_EM (eqtlbma_hm.cpp:925-1346) and the per-gene /
 * per-SNP log10 sums it calls (hm_methods.cpp:74-781).
 * raw[pair][d][l]: the log10 ABFs of the file eqtlbma_bf writes
 * (_l10abfs_raw), one row d per kept configuration (configs model, d <
 * dim) or per subgroup (types model, d < n_subgroups), n_grid values each;
 * the pairs of gene g are gene_offset[g] .. gene_offset[g+1]-1 (SNP order
 * irrelevant, each SNP weighted 1/P as set_snp_prior does).
 * The parameters are the starting point on input (Controller::init_params)
 * and the estimates on output.  FP64 throughout.  log10_lik: the log10
 * likelihood of the output parameters (the last EM iterate, as the reference
 * keeps it); n_iters: the reference's iteration count (the iteration after
 * the stopping rule included).  EQBF_ERR_DATA where the reference stops
 * with an error: the likelihood decreases under classic EM, the log10
 * likelihood is NaN or infinite, or an unnormalised configuration or grid
 * weight is NaN (eqtlbma_hm.cpp:640-647, 697-700, 844-847, 1088-1093).
 * stepmax (--msl): 1 = classic EM (run_EM_classic); any other value =
 * SQUAREM with step-length rule 3 (run_EM_square, eqtlbma_hm.cpp:1109-1324;
 * run_EM selects classic only for exactly 1). */
enum { EQBF_HM_CONFIGS = 0, EQBF_HM_TYPES = 1 };
typedef struct eqbf_hm_opts {
  int32_t model;       /* EQBF_HM_CONFIGS (--model configs) or EQBF_HM_TYPES  */
  int32_t dim;         /* --dim: configurations kept, or types                */
  int32_t n_subgroups; /* --nsubgrp                                           */
  int32_t n_grid;      /* --ngrid                                             */
  double thresh;       /* --thresh (stopping rule on the log10 likelihood)    */
  int64_t max_iters;   /* --maxit; <= 0: no limit                             */
  double stepmax;      /* --msl                                               */
  int32_t fixed[5];    /* held fixed: pi0, grid-points, configs, types,
                          subgroups-per-type (the --init file's third column) */
} eqbf_hm_opts;
typedef struct eqbf_hm_params {
  double pi0;
  double* grid;     /* [n_grid]                                  */
  double* config;   /* [dim]          configs model (else NULL)  */
  double* type;     /* [dim]          types model (else NULL)    */
  double* subgroup; /* [dim][n_subgroups] types model (else NULL) */
} eqbf_hm_params;
int eqbf_hm_em(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset, int64_t n_genes,
               const eqbf_hm_opts* opts, eqbf_hm_params* params, double* log10_lik, int64_t* n_iters);
/* raw_dev and gene_offset_dev in HBM (e.g. eqtlbma_bf's raw ABFs left on the
 * device); the parameters stay host arrays */
int eqbf_hm_em_device(eqbf_ctx* ctx, const double* raw_dev, int64_t n_pairs, const int64_t* gene_offset_dev,
                      int64_t n_genes, const eqbf_hm_opts* opts, eqbf_hm_params* params, double* log10_lik,
                      int64_t* n_iters);
/* --getbf outputs (gene_eQTL::compute_posterior, hm_methods.cpp:752-781;
 * Controller::save_result, eqtlbma_hm.cpp:1683-1740) for the given
 * parameters: gene_post[g] = min(1, (1 - pi0) 10^(BF_g - obslik_g)),
 * gene_l10bf[g], snp_l10bf[pair], dim_l10bf[pair][dim] (per configuration
 * or type, averaged over the grid).  Host arrays. */
int eqbf_hm_bfs(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset, int64_t n_genes,
                const eqbf_hm_opts* opts, const eqbf_hm_params* params, double* gene_post, double* gene_l10bf,
                double* snp_l10bf, double* dim_l10bf);
/* --getci (Controller::estimate_profile_ci, eqtlbma_hm.cpp:1348-1546):
 * profile-likelihood bounds in steps of 0.001 where the natural-log
 * likelihood stays within 2 of its value at the given parameters; pi0,
 * configuration and grid weights are walked (a simplex's other weights
 * rescaled to keep their sum), types and subgroup probabilities keep their
 * value as both bounds.  left/right: [1 + dim + (types model: dim *
 * n_subgroups) + n_grid] in the order pi0, configurations or types,
 * subgroups per type, grid points.  Host arrays.  EQBF_ERR_DATA if a log10
 * likelihood on the way is NaN or infinite. */
int eqbf_hm_ci(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset, int64_t n_genes,
               const eqbf_hm_opts* opts, const eqbf_hm_params* params, double* left, double* right);

/* ---- eqtlbma_avg_bfs: averaging raw ABFs, posteriors ------------------------------
 * Replaces the per-gene loop of averageBFs (src/eqtlbma_avg_bfs.cpp:1025-1101)
 * over Snp::avg_raw_bfs (:418-475), Gene::avg_raw_bfs (:598-622) and the
 * posteriors (:634-796); the host program keeps the reference's file
 * parsing, filters, SNP ordering and writers.  Raw log10 ABFs [n_pairs][R]
 * [n_grid] in FP64: R = dim for the configs model (the j-th configuration
 * row read for the SNP goes with the j-th weight, as in the reference),
 * R = n_subgroups for the types model (subgroup-specific rows).  Every sum
 * is utils::log10_weighted_sum (utils_math.cpp:100-159).  Outputs (FP64,
 * [n_genes] or [n_pairs][...]): gene_bf, snp_bf, dim_bf [n_pairs][dim],
 * best_dim (configs: first strict maximum of dim_bf, replaced by that of
 * post_dim when computed; types: -1 unless post_dim is computed), and the
 * posteriors selected by want: EQBF_AVG_POST_A gene_post, _B post_the,
 * _C post_an (never capped: the reference tests post_the, :677-678), _D
 * post_sub [n_pairs][n_subgroups] (configs: sums of post_dim over the
 * configurations whose subgroup bit is set in cfg_subgroups), _DIMS
 * post_dim [n_pairs][dim] (--bestdim / --alldim; the configs model computes
 * it also for _D).  Outputs not selected may be NULL. */
enum { EQBF_AVG_CONFIGS = 0, EQBF_AVG_TYPES = 1 };
enum { EQBF_AVG_POST_A = 1, EQBF_AVG_POST_B = 2, EQBF_AVG_POST_C = 4, EQBF_AVG_POST_D = 8, EQBF_AVG_POST_DIMS = 16 };
typedef struct eqbf_avg_opts {
  int32_t model;               /* EQBF_AVG_CONFIGS / EQBF_AVG_TYPES (--model)          */
  int32_t dim;                 /* --dim: configurations or types                       */
  int32_t n_subgroups;         /* types: raw rows per pair; configs: post_sub columns  */
  int32_t n_grid;              /* raw values per row (grid weights read, --gwts)       */
  int32_t n_keep;              /* grid points kept (--gtk)                             */
  const int32_t* keep;         /* [n_keep] indices < n_grid                            */
  const double* grid_weights;  /* [n_grid]                                             */
  const double* dim_weights;   /* [dim] configuration (--cwts) or type (--tswts) weights */
  const uint64_t* cfg_subgroups; /* [dim] configs model: bit s = subgroup s+1 in the name */
  const double* sub_weights;   /* [dim][n_subgroups] types model                       */
  double pi0;                  /* --pi0 (EQBF_AVG_POST_A)                              */
  int32_t want;                /* EQBF_AVG_POST_* bits                                 */
} eqbf_avg_opts;
typedef struct eqbf_avg_out {
  double* gene_bf;
  double* gene_post;
  double* snp_bf;
  double* dim_bf;
  double* post_the;
  double* post_an;
  double* post_dim;
  double* post_sub;
  int32_t* best_dim;
} eqbf_avg_out;
int eqbf_avg_bfs(eqbf_ctx* ctx, const double* raw, int64_t n_pairs, const int64_t* gene_offset, int64_t n_genes,
                 const eqbf_avg_opts* opts, eqbf_avg_out* out);
/* raw_dev, gene_offset_dev and out_dev in HBM; opts' arrays stay host arrays */
int eqbf_avg_bfs_device(eqbf_ctx* ctx, const double* raw_dev, int64_t n_pairs, const int64_t* gene_offset_dev,
                        int64_t n_genes, const eqbf_avg_opts* opts, eqbf_avg_out* out_dev, void* hip_stream);
/* Without configuration or type weights (averageRawBFsOverGrid, :834-879):
 * each raw row [n_rows][n_grid] averaged over the kept grid points. */
int eqbf_avg_grid(eqbf_ctx* ctx, const double* raw, int64_t n_rows, int32_t n_grid, const double* grid_weights,
                  const int32_t* keep, int32_t n_keep, double* out);

#ifdef __cplusplus
}
#endif
#endif /* EQBF_H */
