/*
 * C-ABI entry points of the association test (include/eqbf.h): the
 * device-pointer variant enqueues on a stream; the host-pointer variant stages
 * the chunk into context-owned device buffers, runs, and copies back.
 */
#include <cstring>
#include <string>

#include "eqbf_internal.h"

static int check_flag(eqbf_ctx* ctx, hipStream_t st) {
  int flag = 0;
  EQBF_HIP(ctx, hipMemcpyAsync(&flag, ctx->d_flag.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  if (flag & FLAG_MISSING_COVAR)
    return ctx_fail(ctx, EQBF_ERR_DATA,
                    "missing covariate for a sample with genotype and expression level "
                    "(gene_snp_pair.cpp:138-144)");
  if (flag & FLAG_MVLR_SAMPLES)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED,
                    "--error mvlr needs the same samples (genotyped, with expression levels) in every subgroup");
  if (flag & FLAG_HYB_NO_COMMON)
    return ctx_fail(ctx, EQBF_ERR_DATA, "two subgroups have no individuals in common (gene_snp_pair.cpp:895-900)");
  return EQBF_OK;
}

/* hipGraph replay of the batched association.  A call is identified by
 * its argument bytes (chunk, options, output pointers, stream).  The first
 * time it runs eagerly (sizing every buffer and uploading the per-grid
 * tables); the second time, if nothing was reallocated in between, the
 * enqueued kernels are captured into a graph and launched; later calls
 * launch the graph.  A reallocation or new grids (buf_epoch) retire it. */
static int assoc_graph_run(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_assoc_opts* o, eqbf_assoc_out* out,
                           hipStream_t st) {
  std::string key((const char*)c, sizeof *c);
  key.append((const char*)o, sizeof *o);
  key.append((const char*)out, sizeof *out);
  key.append((const char*)&st, sizeof st);
  eqbf_ctx::GraphEntry& e = ctx->graphs[key];
  if (e.exec && e.epoch == ctx->buf_epoch) {
    EQBF_HIP(ctx, hipGraphLaunch(e.exec, st));
    ++ctx->graph_launches;
    return EQBF_OK;
  }
  if (e.exec) {
    (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
    e.seen = 0;
  }
  if (!e.seen || e.epoch != ctx->buf_epoch) {
    const int rc = assoc_uvlr_run(ctx, c, o, out, st);
    e.seen = 1;
    e.epoch = ctx->buf_epoch;
    return rc;
  }
  hipGraph_t g = nullptr;
  EQBF_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const uint64_t ep = ctx->buf_epoch;
  ctx->capturing = true;
  const int rc = assoc_uvlr_run(ctx, c, o, out, st);
  ctx->capturing = false;
  const hipError_t ec = hipStreamEndCapture(st, &g);
  if (rc == EQBF_OK && ec == hipSuccess && g && ctx->buf_epoch == ep) {
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
      (void)hipGraphDestroy(g);
      e.exec = ex;
      e.epoch = ep;
      EQBF_HIP(ctx, hipGraphLaunch(ex, st));
      ++ctx->graph_launches;
      return EQBF_OK;
    }
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  /* nothing ran during the capture: run the call eagerly */
  e.seen = 0;
  return assoc_uvlr_run(ctx, c, o, out, st);
}

extern "C" int eqbf_assoc_uvlr_device(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_assoc_opts* o,
                                      eqbf_assoc_out* out, void* stream) {
  if (!ctx || !c || !o || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  if (ctx->graph_on && !ctx->prof_on && o->precision == EQBF_PREC_MIXED) return assoc_graph_run(ctx, c, o, out, st);
  return assoc_uvlr_run(ctx, c, o, out, st);
}

extern "C" int eqbf_set_graphs(eqbf_ctx* ctx, int enable) {
  if (!ctx) return EQBF_ERR_ARG;
  ctx->graph_on = enable != 0;
  if (!ctx->graph_on) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->graphs)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    ctx->graphs.clear();
  }
  return EQBF_OK;
}

extern "C" int eqbf_graph_stats(eqbf_ctx* ctx, int64_t* n_graphs, int64_t* n_launches) {
  if (!ctx) return EQBF_ERR_ARG;
  int64_t n = 0;
  for (auto& kv : ctx->graphs) n += kv.second.exec != nullptr;
  if (n_graphs) *n_graphs = n;
  if (n_launches) *n_launches = ctx->graph_launches;
  return EQBF_OK;
}

extern "C" int eqbf_set_autotune(eqbf_ctx* ctx, int enable) {
  if (!ctx) return EQBF_ERR_ARG;
  ctx->autotune_on = enable != 0;
  return EQBF_OK;
}

extern "C" int eqbf_autotune_read(eqbf_ctx* ctx, char* keys, int32_t* kch, int32_t* spp, int32_t* grid_mult,
                                  float* ms, int32_t n_max, int32_t* n_out) {
  if (!ctx || !n_out) return EQBF_ERR_ARG;
  int32_t k = 0;
  for (auto& kv : ctx->tuned) {
    if (k >= n_max) break;
    if (keys) {
      strncpy(keys + 64 * k, kv.first.c_str(), 63);
      keys[64 * k + 63] = 0;
    }
    if (kch) kch[k] = kv.second.kch;
    if (spp) spp[k] = kv.second.spp;
    if (grid_mult) grid_mult[k] = kv.second.grid_mult;
    if (ms) ms[k] = kv.second.ms;
    ++k;
  }
  *n_out = k;
  return EQBF_OK;
}

namespace {
struct Stager {
  eqbf_ctx* ctx;
  hipStream_t st;
  int rc = EQBF_OK;
  template <class T>
  T* put(const char* name, const T* host, size_t count) {
    if (!host || count == 0) return nullptr;
    T* d = (T*)ctx_buf(ctx, name, count * sizeof(T));
    if (!d) {
      rc = EQBF_ERR_HIP;
      return nullptr;
    }
    if (hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) rc = EQBF_ERR_HIP;
    return d;
  }
  template <class T>
  T* alloc(const char* name, size_t count, bool want) {
    if (!want || count == 0) return nullptr;
    T* d = (T*)ctx_buf(ctx, name, count * sizeof(T));
    if (!d) rc = EQBF_ERR_HIP;
    return d;
  }
};
}  // namespace

/* Copies a host chunk into context-owned device buffers. */
static int stage_chunk(eqbf_ctx* ctx, Stager& sg, const eqbf_chunk* c, eqbf_chunk* d) {
  const int64_t G = c->n_genes, S = c->n_subgroups, N = c->n_samples, NP = c->n_pairs;
  const int64_t Qm = c->n_covars_max;
  if (G < 0 || S < 1 || N < 1 || NP < 0 || c->n_geno_sets < 1)
    return ctx_fail(ctx, EQBF_ERR_ARG, "invalid chunk sizes");
  const size_t esz = c->geno_dtype == EQBF_F32 ? 4 : 8;
  const size_t geno_elems = (size_t)(c->n_geno_sets - 1) * c->geno_set_stride + (size_t)N * c->geno_ld;
  *d = *c;
  d->explevels = sg.put("h_expl", c->explevels, (size_t)(G * S * N));
  d->gene_in_sbgrp = sg.put("h_genein", c->gene_in_sbgrp, (size_t)(G * S));
  d->geno_in_sbgrp = sg.put("h_genoin", c->geno_in_sbgrp, (size_t)(S * N));
  d->covar_count = sg.put("h_covcnt", c->covar_count, (size_t)S);
  d->covars = Qm > 0 ? sg.put("h_covars", c->covars, (size_t)(S * Qm * N)) : nullptr;
  d->covar_in_sbgrp = sg.put("h_covin", c->covar_in_sbgrp, (size_t)(S * N));
  d->geno_set_of_sbgrp = sg.put("h_gset", c->geno_set_of_sbgrp, (size_t)S);
  d->genotypes = sg.put("h_geno", (const uint8_t*)c->genotypes, geno_elems * esz);
  d->snp_in_sbgrp = sg.put("h_snpin", c->snp_in_sbgrp, (size_t)(c->n_snps * S));
  d->pair_offset = sg.put("h_poff", c->pair_offset, (size_t)(G + 1));
  d->pair_snp = sg.put("h_psnp", c->pair_snp, (size_t)NP);
  if (sg.rc) return ctx_fail(ctx, sg.rc, "staging the chunk failed");
  return EQBF_OK;
}

extern "C" int eqbf_perm_stats_device(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po,
                                      const int32_t* perm_idx_dev, int32_t n_tables, const int32_t* gene_table,
                                      int64_t nperm, double* stats_dev, void* stream) {
  if (!ctx || !c || !po || !perm_idx_dev || !gene_table || !stats_dev || n_tables < 0)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return perm_stats_run(ctx, c, po, perm_idx_dev, n_tables, gene_table, nperm, stats_dev, st);
}

/* the hybrid permutation options' arrays on the device */
static eqbf_perm_opts stage_perm_opts(Stager& sg, const eqbf_chunk* c, const eqbf_perm_opts* po) {
  eqbf_perm_opts d = *po;
  if (po->error_model == EQBF_ERROR_HYBRID) {
    const int64_t G = c->n_genes, S = c->n_subgroups, N = c->n_samples;
    d.covars_by_position =
        c->n_covars_max > 0 ? sg.put("h_pcp", po->covars_by_position, (size_t)(S * c->n_covars_max * N)) : nullptr;
    d.explevels_pairs = sg.put("h_pyp", po->explevels_pairs, (size_t)(G * S * N));
  } else {
    d.covars_by_position = nullptr;
    d.explevels_pairs = nullptr;
  }
  return d;
}

extern "C" int eqbf_perm_stats(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po, const int32_t* perm_idx,
                               int32_t n_tables, const int32_t* gene_table, int64_t nperm, double* stats) {
  if (!ctx || !c || !po || !perm_idx || !gene_table || !stats || n_tables < 0)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  Stager sg{ctx, st};
  eqbf_chunk d;
  int rc = stage_chunk(ctx, sg, c, &d);
  if (rc) return rc;
  const int32_t* dtab = sg.put("h_ptab", perm_idx, (size_t)n_tables * nperm * c->n_samples);
  double* dstats = sg.alloc<double>("o_pstats", (size_t)c->n_genes * nperm, true);
  const eqbf_perm_opts pd = stage_perm_opts(sg, c, po);
  if (sg.rc || (n_tables > 0 && !dtab)) return ctx_fail(ctx, EQBF_ERR_HIP, "staging the tables failed");
  rc = perm_stats_run(ctx, &d, &pd, dtab, n_tables, gene_table, nperm, dstats, st);
  if (rc) return rc;
  rc = check_flag(ctx, st);
  if (rc) return rc;
  EQBF_HIP(ctx, hipMemcpyAsync(stats, dstats, (size_t)c->n_genes * nperm * sizeof(double), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_perm_stats_fp64_ties_device(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po,
                                                const int32_t* perm_idx_dev, int32_t n_tables,
                                                const int32_t* gene_table, int64_t nperm, double tie_bound,
                                                double* stats_dev, double* true_stats_dev, int64_t* n_rechecked,
                                                void* stream) {
  if (!ctx || !c || !po || !perm_idx_dev || !gene_table || !stats_dev || !true_stats_dev || n_tables < 0)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return perm_ties_run(ctx, c, po, perm_idx_dev, n_tables, gene_table, nperm, tie_bound, stats_dev, true_stats_dev,
                       n_rechecked, st);
}

extern "C" int eqbf_perm_stats_fp64_ties(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_perm_opts* po,
                                         const int32_t* perm_idx, int32_t n_tables, const int32_t* gene_table,
                                         int64_t nperm, double tie_bound, double* stats, double* true_stats,
                                         int64_t* n_rechecked) {
  if (!ctx || !c || !po || !perm_idx || !gene_table || !stats || !true_stats || n_tables < 0)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  Stager sg{ctx, st};
  eqbf_chunk d;
  int rc = stage_chunk(ctx, sg, c, &d);
  if (rc) return rc;
  const int32_t* dtab = sg.put("h_ptab", perm_idx, (size_t)n_tables * std::max<int64_t>(nperm, 0) * c->n_samples);
  double* dstats = sg.alloc<double>("o_pstats", (size_t)c->n_genes * std::max<int64_t>(nperm, 1), true);
  double* dtrue = c->n_genes > 0 ? sg.put("h_ptrue", true_stats, (size_t)c->n_genes) : nullptr;
  const eqbf_perm_opts pd = stage_perm_opts(sg, c, po);
  if (sg.rc || (n_tables > 0 && nperm > 0 && !dtab) || !dstats || (c->n_genes > 0 && !dtrue))
    return ctx_fail(ctx, EQBF_ERR_HIP, "staging the tables failed");
  rc = perm_ties_run(ctx, &d, &pd, dtab ? dtab : (const int32_t*)dstats, n_tables, gene_table, nperm, tie_bound, dstats,
                     dtrue, n_rechecked, st);
  if (rc) return rc;
  rc = check_flag(ctx, st);
  if (rc) return rc;
  if (nperm > 0)
    EQBF_HIP(ctx, hipMemcpyAsync(stats, dstats, (size_t)c->n_genes * nperm * sizeof(double), hipMemcpyDeviceToHost, st));
  if (c->n_genes > 0)
    EQBF_HIP(ctx, hipMemcpyAsync(true_stats, dtrue, (size_t)c->n_genes * sizeof(double), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_assoc_mvlr_device(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_mvlr_opts* o,
                                      eqbf_assoc_out* out, void* stream) {
  if (!ctx || !c || !o || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return mvlr_run(ctx, c, o, out, st);
}

extern "C" int eqbf_assoc_mvlr(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_mvlr_opts* o, eqbf_assoc_out* out) {
  if (!ctx || !c || !o || !out || !out->n || !out->raw_abf || !out->w_abf)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t S = c->n_subgroups, NP = c->n_pairs;
  Stager sg{ctx, st};
  eqbf_chunk d;
  int rc = stage_chunk(ctx, sg, c, &d);
  if (rc) return rc;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const int64_t n_cfg = eqbf_n_configs(o->bfs, (int)S);
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", (int)S);
  const int64_t n_raw = eqbf_n_raw(o->bfs, (int)S, GL, GS), n_w = eqbf_n_weighted(o->bfs, (int)S);
  eqbf_assoc_out dout;
  memset(&dout, 0, sizeof dout);
  dout.n = sg.alloc<int32_t>("o_n", (size_t)(S * NP), true);
  dout.sstats = sg.alloc<double>("o_ss", (size_t)(5 * S * NP), out->sstats != nullptr);
  dout.stdsstats = sg.alloc<double>("o_std", (size_t)(3 * S * NP), out->stdsstats != nullptr);
  dout.raw_abf = sg.alloc<double>("o_raw", (size_t)(n_raw * NP), true);
  dout.w_abf = sg.alloc<double>("o_w", (size_t)(n_w * NP), true);
  if (sg.rc) return ctx_fail(ctx, sg.rc, "allocating outputs failed");
  rc = mvlr_run(ctx, &d, o, &dout, st);
  if (rc) return rc;
  rc = check_flag(ctx, st);
  if (rc) return rc;
  if (NP > 0) {
    EQBF_HIP(ctx, hipMemcpyAsync(out->n, dout.n, S * NP * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (out->sstats)
      EQBF_HIP(ctx, hipMemcpyAsync(out->sstats, dout.sstats, 5 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    if (out->stdsstats)
      EQBF_HIP(ctx, hipMemcpyAsync(out->stdsstats, dout.stdsstats, 3 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipMemcpyAsync(out->raw_abf, dout.raw_abf, n_raw * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipMemcpyAsync(out->w_abf, dout.w_abf, n_w * NP * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_assoc_hybrid_device(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_hybrid_opts* o,
                                        eqbf_assoc_out* out, void* stream) {
  if (!ctx || !c || !o || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return hybrid_run(ctx, c, o, out, st);
}

extern "C" int eqbf_assoc_hybrid(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_hybrid_opts* o, eqbf_assoc_out* out) {
  if (!ctx || !c || !o || !out || !out->n || !out->raw_abf || !out->w_abf)
    return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t G = c->n_genes, S = c->n_subgroups, N = c->n_samples, NP = c->n_pairs;
  Stager sg{ctx, st};
  eqbf_chunk d;
  int rc = stage_chunk(ctx, sg, c, &d);
  if (rc) return rc;
  eqbf_hybrid_opts od = *o;
  od.covars_by_position =
      c->n_covars_max > 0 ? sg.put("h_hyb_cp", o->covars_by_position, (size_t)(S * c->n_covars_max * N)) : nullptr;
  od.explevels_pairs = sg.put("h_hyb_yp", o->explevels_pairs, (size_t)(G * S * N));
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const int64_t n_cfg = eqbf_n_configs(o->bfs, (int)S);
  if (n_cfg < 0) return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs all with %d subgroups", (int)S);
  const int64_t n_raw = eqbf_n_raw(o->bfs, (int)S, GL, GS), n_w = eqbf_n_weighted(o->bfs, (int)S);
  eqbf_assoc_out dout;
  memset(&dout, 0, sizeof dout);
  dout.n = sg.alloc<int32_t>("o_n", (size_t)(S * NP), true);
  dout.sstats = sg.alloc<double>("o_ss", (size_t)(5 * S * NP), out->sstats != nullptr);
  dout.raw_abf = sg.alloc<double>("o_raw", (size_t)(n_raw * NP), true);
  dout.w_abf = sg.alloc<double>("o_w", (size_t)(n_w * NP), true);
  if (sg.rc) return ctx_fail(ctx, sg.rc, "allocating outputs failed");
  rc = hybrid_run(ctx, &d, &od, &dout, st);
  if (rc) return rc;
  rc = check_flag(ctx, st);
  if (rc) return rc;
  if (NP > 0) {
    EQBF_HIP(ctx, hipMemcpyAsync(out->n, dout.n, S * NP * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (out->sstats)
      EQBF_HIP(ctx, hipMemcpyAsync(out->sstats, dout.sstats, 5 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipMemcpyAsync(out->raw_abf, dout.raw_abf, n_raw * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    EQBF_HIP(ctx, hipMemcpyAsync(out->w_abf, dout.w_abf, n_w * NP * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_assoc_uvlr(eqbf_ctx* ctx, const eqbf_chunk* c, const eqbf_assoc_opts* o, eqbf_assoc_out* out) {
  if (!ctx || !c || !o || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t S = c->n_subgroups, NP = c->n_pairs;
  Stager sg{ctx, st};
  eqbf_chunk d;
  int rc0 = stage_chunk(ctx, sg, c, &d);
  if (rc0) return rc0;
  const bool join = o->analysis == EQBF_ANALYSIS_JOIN;
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  int64_t n_raw = join ? eqbf_n_raw(o->bfs, (int)S, GL, GS) : 0;
  int64_t n_w = join ? eqbf_n_weighted(o->bfs, (int)S) : 0;
  const size_t rsz = o->raw_dtype == EQBF_F64 ? 8 : 4;
  eqbf_assoc_out dout;
  memset(&dout, 0, sizeof dout);
  dout.n = sg.alloc<int32_t>("o_n", (size_t)(S * NP), true);
  dout.sstats = sg.alloc<double>("o_ss", (size_t)(5 * S * NP), out->sstats != nullptr);
  dout.stdsstats = sg.alloc<double>("o_std", (size_t)(3 * S * NP), out->stdsstats != nullptr);
  dout.raw_abf = sg.alloc<uint8_t>("o_raw", (size_t)(n_raw * NP) * rsz, join && out->raw_abf != nullptr);
  dout.w_abf = sg.alloc<double>("o_w", (size_t)(n_w * NP), join);
  if (sg.rc) return ctx_fail(ctx, sg.rc, "allocating outputs failed");
  int rc = assoc_uvlr_run(ctx, &d, o, &dout, st);
  if (rc) return rc;
  rc = check_flag(ctx, st);
  if (rc) return rc;
  if (NP > 0) {
    EQBF_HIP(ctx, hipMemcpyAsync(out->n, dout.n, S * NP * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (out->sstats)
      EQBF_HIP(ctx, hipMemcpyAsync(out->sstats, dout.sstats, 5 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    if (out->stdsstats)
      EQBF_HIP(ctx, hipMemcpyAsync(out->stdsstats, dout.stdsstats, 3 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
    if (join && out->raw_abf)
      EQBF_HIP(ctx, hipMemcpyAsync(out->raw_abf, dout.raw_abf, n_raw * NP * rsz, hipMemcpyDeviceToHost, st));
    if (join && out->w_abf)
      EQBF_HIP(ctx, hipMemcpyAsync(out->w_abf, dout.w_abf, n_w * NP * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}

extern "C" int eqbf_abf_from_sstats_device(eqbf_ctx* ctx, const eqbf_sstats_in* in, int32_t bfs, int32_t raw_dtype,
                                           eqbf_assoc_out* out, void* stream) {
  if (!ctx || !in || !out) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  return abf_from_sstats_run(ctx, in, bfs, raw_dtype, out, st);
}

extern "C" int eqbf_abf_from_sstats(eqbf_ctx* ctx, const eqbf_sstats_in* in, int32_t bfs, int32_t raw_dtype,
                                    eqbf_assoc_out* out) {
  if (!ctx || !in || !out || !out->w_abf) return ctx_fail(ctx, EQBF_ERR_ARG, "null argument");
  EQBF_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t S = in->n_subgroups, NP = in->n_pairs;
  if (S < 1 || NP < 0) return ctx_fail(ctx, EQBF_ERR_ARG, "invalid summary statistics sizes");
  if (NP == 0) return EQBF_OK;
  if (!in->n || !in->sigmahat || !in->betahat || !in->sebetahat) return ctx_fail(ctx, EQBF_ERR_ARG, "null input");
  const int GL = (int)ctx->phiL.size(), GS = (int)ctx->phiS.size();
  const int64_t n_raw = eqbf_n_raw(bfs, (int)S, GL, GS), n_w = eqbf_n_weighted(bfs, (int)S);
  if (eqbf_n_configs(bfs, (int)S) < 0 || n_raw < 0)
    return ctx_fail(ctx, EQBF_ERR_UNSUPPORTED, "--bfs %d with %d subgroups", bfs, (int)S);
  const size_t rsz = raw_dtype == EQBF_F64 ? 8 : 4;
  Stager sg{ctx, st};
  eqbf_sstats_in d = *in;
  d.n = sg.put("h_inss_n", in->n, (size_t)(S * NP));
  d.sigmahat = sg.put("h_inss_sig", in->sigmahat, (size_t)(S * NP));
  d.betahat = sg.put("h_inss_beta", in->betahat, (size_t)(S * NP));
  d.sebetahat = sg.put("h_inss_se", in->sebetahat, (size_t)(S * NP));
  eqbf_assoc_out dout;
  memset(&dout, 0, sizeof dout);
  dout.n = sg.alloc<int32_t>("o_n", (size_t)(S * NP), out->n != nullptr);
  dout.stdsstats = sg.alloc<double>("o_std", (size_t)(3 * S * NP), out->stdsstats != nullptr);
  dout.raw_abf = sg.alloc<uint8_t>("o_raw", (size_t)(n_raw * NP) * rsz, out->raw_abf != nullptr);
  dout.w_abf = sg.alloc<double>("o_w", (size_t)(n_w * NP), true);
  if (sg.rc) return ctx_fail(ctx, sg.rc, "staging the summary statistics failed");
  int rc = abf_from_sstats_run(ctx, &d, bfs, raw_dtype, &dout, st);
  if (rc) return rc;
  if (out->n) EQBF_HIP(ctx, hipMemcpyAsync(out->n, dout.n, S * NP * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (out->stdsstats)
    EQBF_HIP(ctx, hipMemcpyAsync(out->stdsstats, dout.stdsstats, 3 * S * NP * sizeof(double), hipMemcpyDeviceToHost, st));
  if (out->raw_abf)
    EQBF_HIP(ctx, hipMemcpyAsync(out->raw_abf, dout.raw_abf, n_raw * NP * rsz, hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipMemcpyAsync(out->w_abf, dout.w_abf, n_w * NP * sizeof(double), hipMemcpyDeviceToHost, st));
  EQBF_HIP(ctx, hipStreamSynchronize(st));
  return EQBF_OK;
}
