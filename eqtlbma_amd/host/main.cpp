/*
 * eqtlbma_bf drop-in: loads the inputs like the reference, then processes
 * genes in chunks of --wrtsize (in gene-name order) through the MI355X
 * engine (include/eqbf.h) and appends each chunk's results to the gzipped
 * output files.  Mirrors run() in src/eqtlbma_bf.cpp:1449-1582:
 * testForAssociations (:707-771) -> eqbf_assoc_uvlr / eqbf_assoc_mvlr;
 * makePermutations (:867-917) -> host index tables with the reference's
 * generator and reseeding, device statistics with FP64 true statistics and
 * near-ties (eqbf_perm_stats_fp64_ties), host counting (eqbf_perm_pvalue); writeRes (:1401-1447) and its writers
 * (:919-1399) for the file formats.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <ctime>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/eqbf.h"
#include "data.hpp"
#include "options.hpp"
#include "util.hpp"

using namespace host;

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

/* ---- grids (grid.cpp:28-67) ------------------------------------------------ */
void load_grid(const std::string& file, std::vector<double>& phi2, std::vector<double>& oma2, int verbose) {
  if (file.empty()) return;
  if (verbose > 0) std::cout << "load grid in " << file << " ..." << std::endl;
  LineReader r(file);
  std::string line;
  std::vector<std::string> tok;
  while (r.next(line)) {
    tokenize(line, " \t", tok);
    if (tok.size() != 2) fail("format of file " + file + " should be phi2<space/tab>oma2");
    phi2.push_back(std::atof(tok[0].c_str()));
    oma2.push_back(std::atof(tok[1].c_str()));
  }
  if (verbose > 0) std::cout << "grid size: " << phi2.size() << std::endl;
}

/* ---- normal quantile (Wichura 1988, AS241 PPND16) for --qnorm ------------------ */
double normal_quantile(double p) {
  const double q = p - 0.5;
  if (std::fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    return q * (((((((r * 2509.0809287301226727 + 33430.575583588128105) * r + 67265.770927008700853) * r +
                    45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
                 133.14166789178437745) * r + 3.387132872796366608) /
           (((((((r * 5226.495278852545561 + 28729.085735721942674) * r + 39307.89580009271061) * r +
                21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
             42.313330701600911252) * r + 1.0);
  }
  double r = q < 0 ? p : 1.0 - p;
  r = std::sqrt(-std::log(r));
  double v;
  if (r <= 5.0) {
    r -= 1.6;
    v = (((((((r * 7.7454501427834140764e-4 + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
             1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
          4.6303378461565452959) * r + 1.42343711074968357734) /
        (((((((r * 1.05075007164441684324e-9 + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
             0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
          2.05319162663775882187) * r + 1.0);
  } else {
    r -= 5.0;
    v = (((((((r * 2.01033439929228813265e-7 + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
             0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
          5.4637849111641143699) * r + 6.6579046435011037772) /
        (((((((r * 2.04426310338993978564e-15 + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
             7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
          0.59983220655588793769) * r + 1.0);
  }
  return q < 0 ? -v : v;
}

/* utils::qqnorm (utils_math.cpp:80-96) on the values at idx (ties keep the
 * sample order; GSL's heap-sort order of ties is not reproduced). */
void qqnorm(double* y, const std::vector<int32_t>& idx) {
  const size_t n = idx.size();
  if (n == 0) return;
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return y[idx[a]] < y[idx[b]]; });
  const double a = n <= 10 ? 0.375 : 0.5;
  std::vector<double> out(n);
  for (size_t i = 0; i < n; ++i) out[ord[i]] = normal_quantile((i + 1 - a) / (n + 1 - 2 * a));
  for (size_t i = 0; i < n; ++i) y[idx[i]] = out[i];
}

/* ---- engine ---------------------------------------------------------------- */
struct Engine {
  eqbf_ctx* ctx = nullptr;
  explicit Engine(int device) {
    if (eqbf_create(&ctx, device) != EQBF_OK) fail("can't create the GPU context on device " + std::to_string(device));
  }
  ~Engine() { eqbf_destroy(ctx); }
  void check(int rc) const {
    if (rc != EQBF_OK) fail(eqbf_last_error(ctx));
  }
};

/* Subgroup-level arrays shared by every chunk. */
struct Shared {
  int S = 0, N = 0, Qmax = 0;
  std::vector<uint8_t> geno_in, covar_in;
  std::vector<int32_t> covar_count, gset_of_sbgrp;
  std::vector<double> covars;
  /* --error hybrid's pairwise designs read covariate q of subgroup s at the
   * sample's index in the sorted union of samples, not at its column of the
   * covariate file (gene_snp_pair.cpp:920-927, covariates.cpp:79-84): the
   * value in column i, NaN past the file's columns */
  std::vector<double> covars_pos;
  /* genotype sets of the chunk: (loaded set, column map of its subgroups) */
  std::vector<int32_t> set_src;                   /* loaded set per chunk set */
  std::vector<std::vector<int32_t>> set_cols;     /* [chunk set][union sample] */
};

Shared make_shared(const Data& d) {
  Shared sh;
  sh.S = (int)d.sbgrps.size();
  sh.N = (int)d.samples.size();
  for (auto& sg : d.sbgrps) sh.Qmax = std::max(sh.Qmax, (int)sg.covar_names.size());
  sh.geno_in.assign((size_t)sh.S * sh.N, 0);
  sh.covar_in.assign((size_t)sh.S * sh.N, 0);
  sh.covar_count.assign(sh.S, 0);
  sh.covars.assign((size_t)sh.S * std::max(sh.Qmax, 1) * sh.N, 0.0);
  std::map<std::pair<int, std::vector<int32_t>>, int> key;
  for (int s = 0; s < sh.S; ++s) {
    const Subgroup& sg = d.sbgrps[s];
    const GenoSet& gs = d.sets[sg.geno_set];
    std::vector<int32_t> cols(sh.N, -1);
    for (int i = 0; i < sh.N; ++i) {
      const int32_t c = sg.geno_col[i];
      if (c >= 0 && c < gs.ncols) cols[i] = c;
      sh.geno_in[(size_t)s * sh.N + i] = cols[i] >= 0;
    }
    auto k = std::make_pair(sg.geno_set, cols);
    auto it = key.find(k);
    if (it == key.end()) {
      it = key.emplace(k, (int)sh.set_src.size()).first;
      sh.set_src.push_back(sg.geno_set);
      sh.set_cols.push_back(cols);
    }
    sh.gset_of_sbgrp.push_back(it->second);
    sh.covar_count[s] = (int)sg.covar_names.size();
    if (sh.covars_pos.empty()) sh.covars_pos.assign(sh.covars.size(), kNaN);
    for (int q = 0; q < sh.covar_count[s]; ++q)
      for (int i = 0; i < sh.N && i < (int)sg.covar_vals[q].size(); ++i)
        sh.covars_pos[((size_t)s * std::max(sh.Qmax, 1) + q) * sh.N + i] = sg.covar_vals[q][i];
    for (int i = 0; i < sh.N; ++i) {
      const int32_t c = sg.covar_col[i];
      if (c < 0) continue;
      sh.covar_in[(size_t)s * sh.N + i] = 1;
      for (int q = 0; q < sh.covar_count[s]; ++q)
        sh.covars[((size_t)s * std::max(sh.Qmax, 1) + q) * sh.N + i] = sg.covar_vals[q][c];
    }
  }
  return sh;
}

/* One chunk of genes in the engine's layout. */
struct Chunk {
  std::vector<const Gene*> genes;
  std::vector<std::vector<const Snp*>> cis;  /* pairs of each gene           */
  std::vector<const Snp*> snps;              /* distinct SNPs of the chunk   */
  std::vector<double> expl, geno;
  std::vector<double> expl_raw; /* --error hybrid with --qnorm: levels before normalisation */
  std::vector<uint8_t> gene_in, snp_in;
  std::vector<int64_t> pair_offset;
  std::vector<int32_t> pair_snp;
  eqbf_chunk c{};
  int64_t n_pairs() const { return pair_offset.back(); }
};

/* Cis SNPs (Gene::SetCisSnps + Snp::IsInCis, gene.cpp:140-157,
 * snp.cpp:274-297) over the position-sorted SNPs of the gene's chromosome. */
std::vector<const Snp*> cis_snps(const Data& d, const Gene& g, const std::string& anchor, uint64_t radius) {
  std::vector<const Snp*> out;
  auto it = d.chr_snps.find(g.chr);
  if (it == d.chr_snps.end()) return out;
  const int64_t r = (int64_t)radius;
  const int64_t lo = g.start >= r ? g.start - r : std::numeric_limits<int64_t>::min();
  const int64_t hi = (anchor == "TSS+TES" ? g.end : g.start) + r;
  for (const Snp* s : it->second) {
    if (s->pos > hi) break;
    if (s->pos < lo) continue;
    out.push_back(s);
  }
  return out;
}

void build_chunk(const Data& d, const Shared& sh, const Options& o, std::vector<const Gene*> genes, Chunk& ch) {
  const int S = sh.S, N = sh.N;
  const int G = (int)genes.size();
  ch.genes = genes;
  ch.cis.assign(G, {});
  ch.pair_offset.assign(G + 1, 0);
  ch.gene_in.assign((size_t)G * S, 0);
  ch.expl.assign((size_t)G * S * N, kNaN);
  const bool multi = o.analysis == "join" && o.error_model != "uvlr";
  const bool keep_raw = o.analysis == "join" && o.error_model == "hybrid" && o.need_qnorm;
  ch.expl_raw.clear();
  if (keep_raw) ch.expl_raw.assign((size_t)G * S * N, kNaN);
  std::map<const Snp*, int32_t> col;
  ch.snps.clear();
  ch.pair_snp.clear();
  for (int g = 0; g < G; ++g) {
    const Gene& gn = *genes[g];
    for (int s = 0; s < S; ++s) {
      const int32_t row = gn.exp_row[s];
      if (row < 0) continue;
      ch.gene_in[(size_t)g * S + s] = 1;
      const Subgroup& sg = d.sbgrps[s];
      double* y = &ch.expl[((size_t)g * S + s) * N];
      for (int i = 0; i < N; ++i)
        if (sg.exp_col[i] >= 0) y[i] = sg.exp_vals[(size_t)row * sg.exp_ncols + sg.exp_col[i]];
      if (keep_raw) std::copy(y, y + N, &ch.expl_raw[((size_t)g * S + s) * N]);
      if (o.need_qnorm) {
        std::vector<int32_t> mem;
        for (int i = 0; i < N; ++i)
          if (!std::isnan(y[i]) && sh.geno_in[(size_t)s * N + i]) mem.push_back(i);
        qqnorm(y, mem);
      }
    }
    std::vector<const Snp*> cis = cis_snps(d, gn, o.anchor, o.radius);
    /* a gene is tested only when a cis SNP has genotypes in a subgroup where
     * the gene has expression levels (eqtlbma_bf.cpp:747-753); --error mvlr
     * and hybrid also need expression levels in every subgroup (:755-762) */
    bool any = false, all_expr = true;
    for (int s = 0; s < S; ++s) all_expr &= gn.exp_row[s] >= 0;
    for (const Snp* sn : cis) {
      for (int s = 0; s < S && !any; ++s) any = gn.exp_row[s] >= 0 && sn->row[d.sbgrps[s].geno_set] >= 0;
      if (any) break;
    }
    if (!any || (multi && !all_expr)) cis.clear();
    for (const Snp* sn : cis) {
      auto it = col.find(sn);
      if (it == col.end()) {
        it = col.emplace(sn, (int32_t)ch.snps.size()).first;
        ch.snps.push_back(sn);
      }
      ch.pair_snp.push_back(it->second);
    }
    ch.cis[g] = cis;
    ch.pair_offset[g + 1] = (int64_t)ch.pair_snp.size();
  }
  const int64_t P = (int64_t)ch.snps.size();
  const int nsets = (int)sh.set_src.size();
  ch.geno.assign((size_t)nsets * N * std::max<int64_t>(P, 1), 0.0);
  ch.snp_in.assign((size_t)std::max<int64_t>(P, 1) * S, 0);
  for (int64_t j = 0; j < P; ++j) {
    const Snp& sn = *ch.snps[j];
    for (int s = 0; s < S; ++s) ch.snp_in[(size_t)j * S + s] = sn.row[d.sbgrps[s].geno_set] >= 0;
    for (int k = 0; k < nsets; ++k) {
      const int32_t row = sn.row[sh.set_src[k]];
      if (row < 0) continue;
      const GenoSet& gs = d.sets[sh.set_src[k]];
      const double* v = &gs.vals[(size_t)row * gs.ncols];
      for (int i = 0; i < N; ++i) {
        const int32_t c = sh.set_cols[k][i];
        if (c >= 0) ch.geno[((size_t)k * N + i) * P + j] = v[c];
      }
    }
  }
  eqbf_chunk& c = ch.c;
  c.n_genes = G;
  c.n_subgroups = S;
  c.n_samples = N;
  c.n_covars_max = sh.Qmax;
  c.n_pairs = ch.n_pairs();
  c.n_snps = P;
  c.explevels = ch.expl.data();
  c.gene_in_sbgrp = ch.gene_in.data();
  c.geno_in_sbgrp = sh.geno_in.data();
  c.covar_count = sh.covar_count.data();
  c.covars = sh.covars.data();
  c.covar_in_sbgrp = sh.covar_in.data();
  c.n_geno_sets = nsets;
  c.geno_set_of_sbgrp = sh.gset_of_sbgrp.data();
  c.genotypes = ch.geno.data();
  c.geno_dtype = EQBF_F64;
  c.geno_ld = std::max<int64_t>(P, 1);
  c.geno_set_stride = (int64_t)N * std::max<int64_t>(P, 1);
  c.snp_in_sbgrp = ch.snp_in.data();
  c.pair_offset = ch.pair_offset.data();
  c.pair_snp = ch.pair_snp.data();
}

/* Association results of a chunk. */
struct Results {
  int S = 0;
  int64_t NP = 0, n_raw = 0, n_w = 0;
  std::vector<int32_t> n;
  std::vector<double> sstats, raw, w;
  bool has_abf = false;
  /* --inss: subgroups with a line per pair (GeneSnpPair::GetNbSubgroups,
   * gene_snp_pair.hpp:251, counts n = 0 lines too); empty: count n > 0 */
  std::vector<int32_t> nsub;
};

/* ---- writers (eqtlbma_bf.cpp:919-1447) ---------------------------------------- */
struct Writer {
  const Options& o;
  const Data& d;
  std::vector<std::string> cfg_labels; /* "1", "1-2", ... in configuration order */
  int GL = 0, GS = 0;
  std::string mode;

  void sumstats(const Chunk* ch, const Results* r) {
    for (int s = 0; s < (int)d.sbgrps.size(); ++s) {
      GzWriter f(o.out + "_sumstats_" + d.sbgrps[s].name + ".txt.gz", mode.c_str());
      std::string t;
      if (!ch) {
        f.put("gene\tsnp\tmaf\tn\tpve\tsigmahat\tbetahat.geno\tsebetahat.geno\tbetapval.geno\n");
        continue;
      }
      for (size_t g = 0; g < ch->genes.size(); ++g)
        for (int64_t p = ch->pair_offset[g]; p < ch->pair_offset[g + 1]; ++p) {
          const int32_t n = r->n[(size_t)s * r->NP + p];
          if (n <= 0) continue;
          const Snp& sn = *ch->snps[ch->pair_snp[p]];
          t += ch->genes[g]->name;
          t += '\t';
          t += sn.name;
          t += '\t';
          put_sci(t, sn.maf[d.sbgrps[s].geno_set]);
          t += '\t';
          put_int(t, n);
          for (int k = 0; k < 5; ++k) {
            t += '\t';
            put_sci(t, r->sstats[((size_t)k * r->S + s) * r->NP + p]);
          }
          t += '\n';
        }
      f.put(t);
    }
  }

  void raw(const Chunk* ch, const Results* r) {
    GzWriter f(o.out + "_l10abfs_raw.txt.gz", mode.c_str());
    std::string t;
    if (!ch) {
      t = "gene\tsnp\tconfig";
      for (int i = 0; i < GL; ++i) t += "\tl10abf.grid" + std::to_string(i + 1);
      f.put(t + "\n");
      return;
    }
    static const char* kTypes[3] = {"gen", "gen-fix", "gen-maxh"};
    for (size_t g = 0; g < ch->genes.size(); ++g)
      for (int64_t p = ch->pair_offset[g]; p < ch->pair_offset[g + 1]; ++p) {
        const std::string head = ch->genes[g]->name + "\t" + ch->snps[ch->pair_snp[p]]->name + "\t";
        for (int type = 0; type < 3; ++type) {
          t += head;
          t += kTypes[type];
          for (int k = 0; k < GL; ++k) {
            t += '\t';
            put_sci(t, r->raw[((size_t)type * GL + k) * r->NP + p]);
          }
          t += '\n';
        }
        for (size_t c = 0; c < cfg_labels.size(); ++c) {
          t += head;
          t += cfg_labels[c];
          /* |gridL| columns: the small-grid values, then NaN (:1207-1212) */
          for (int k = 0; k < GL; ++k) {
            t += '\t';
            put_sci(t, k < GS ? r->raw[((size_t)3 * GL + c * GS + k) * r->NP + p] : kNaN);
          }
          t += '\n';
        }
      }
    f.put(t);
  }

  void avg(const Chunk* ch, const Results* r) {
    GzWriter f(o.out + "_l10abfs_avg-grids.txt.gz", mode.c_str());
    std::string t;
    const size_t nc = cfg_labels.size();
    if (!ch) {
      t = "gene\tsnp\tnb.subgroups\tl10abf.gen\tl10abf.gen.fix\tl10abf.gen.maxh";
      if (o.bfs == "sin" || o.bfs == "all") t += "\tl10abf.gen.sin";
      if (o.bfs == "all") t += "\tl10abf.all";
      for (auto& l : cfg_labels) t += "\tl10abf." + l;
      f.put(t + "\n");
      return;
    }
    for (size_t g = 0; g < ch->genes.size(); ++g)
      for (int64_t p = ch->pair_offset[g]; p < ch->pair_offset[g + 1]; ++p) {
        int nsub = 0;
        if (!r->nsub.empty())
          nsub = r->nsub[p];
        else
          for (int s = 0; s < r->S; ++s) nsub += r->n[(size_t)s * r->NP + p] > 0;
        t += ch->genes[g]->name + "\t" + ch->snps[ch->pair_snp[p]]->name + "\t";
        put_int(t, nsub);
        auto w = [&](size_t row) { return r->w[row * r->NP + p]; };
        for (int k = 0; k < 3; ++k) {
          t += '\t';
          put_sci(t, w(k));
        }
        if (o.bfs == "sin" || o.bfs == "all") {
          t += '\t';
          put_sci(t, w(3 + nc));
        }
        if (o.bfs == "all") {
          t += '\t';
          put_sci(t, w(3 + nc + 1));
        }
        for (size_t c = 0; c < nc; ++c) {
          t += '\t';
          put_sci(t, w(3 + c));
        }
        t += '\n';
      }
    f.put(t);
  }
};

/* --ktimes: GPU time of every kernel of the run, on stderr. */
void print_kernel_times(Engine& eng) {
  const int32_t kmax = 256;
  std::vector<char> names((size_t)kmax * 64);
  std::vector<double> ms(kmax);
  std::vector<int64_t> launches(kmax);
  int32_t nk = 0;
  eng.check(eqbf_profile_read(eng.ctx, names.data(), ms.data(), launches.data(), kmax, &nk));
  for (int32_t k = 0; k < std::min(nk, kmax); ++k)
    std::cerr << "kernel\t" << &names[(size_t)k * 64] << "\t" << launches[k] << "\t" << ms[k] << " ms\n";
}

/* Per-gene permutation results. */
struct PermRes {
  double pval = 1, true_stat = kNaN, median = kNaN;
  int64_t used = 0;
  bool done = false;
};

int pbf_code(const std::string& p) { return p == "gen" ? EQBF_PBF_GEN : (p == "gen-sin" ? EQBF_PBF_GENSIN : EQBF_PBF_ALL); }

int bfs_code(const std::string& b) { return b == "gen" ? EQBF_BFS_GEN : (b == "sin" ? EQBF_BFS_SIN : EQBF_BFS_ALL); }

/* log10_weighted_sum with uniform weights (utils_math.cpp:100-131). */
double l10_avg(const std::vector<double>& v) {
  if (v.empty()) return kNaN;
  double mx = v[0];
  for (double x : v)
    if (x > mx) mx = x;
  double sum = 0;
  for (double x : v)
    if (!std::isnan(x)) sum += (1.0 / v.size()) * std::pow(10.0, x - mx);
  double res = mx + std::log10(sum);
  if (std::fabs(res) <= std::numeric_limits<double>::epsilon()) res = 0;
  return res;
}

/* Index tables of consecutive genes continuing one generator: each gene
 * starts from the identity and takes nperm cumulative shuffles
 * (gene.cpp:617,638-639).  The reference reseeds the generator once per
 * chunk (eqtlbma_bf.cpp:889-916), so every chunk draws the same tables as
 * long as no gene of it stops early (--trick 1) or loses permutations to NaN
 * statistics: the tables and the generator state after each gene are kept
 * across chunks, keyed by the generator state they start from, up to
 * kMaxBytes of tables. */
struct PermTableCache {
  static constexpr size_t kMaxBytes = (size_t)1 << 30;
  eqbf_rng start{};
  int64_t N = -1, nperm = -1;
  std::vector<int32_t> tab;    /* [count][nperm][N] */
  std::vector<eqbf_rng> after; /* generator state after each gene */
  static size_t cap_genes(int64_t n, int64_t np) {
    return std::max<size_t>(1, kMaxBytes / ((size_t)np * n * sizeof(int32_t)));
  }
};

bool same_rng(const eqbf_rng& a, const eqbf_rng& b) { return a.idx == b.idx && !std::memcmp(a.mt, b.mt, sizeof a.mt); }

/* Tables for `n` genes from generator state `r`: a pointer to [n][nperm][N]
 * and the state after each gene (from the cache when r is its start or the
 * state after one of its genes; generated otherwise, into `local`). */
const int32_t* tables_from(PermTableCache& cache, const eqbf_rng& r, int64_t N, int64_t nperm, size_t n,
                           bool chunk_start, std::vector<int32_t>& local, std::vector<eqbf_rng>& states) {
  const size_t per = (size_t)nperm * N;
  long pos = -1;
  if (cache.N == N && cache.nperm == nperm) {
    if (same_rng(r, cache.start))
      pos = 0;
    else
      for (size_t j = 0; j < cache.after.size() && pos < 0; ++j)
        if (same_rng(r, cache.after[j])) pos = (long)j + 1;
  }
  if (pos < 0 && chunk_start) {
    cache.start = r;
    cache.N = N;
    cache.nperm = nperm;
    cache.tab.clear();
    cache.after.clear();
    pos = 0;
  }
  if (pos >= 0 && (size_t)pos + n <= PermTableCache::cap_genes(N, nperm)) {
    while (cache.after.size() < (size_t)pos + n) {
      eqbf_rng g = cache.after.empty() ? cache.start : cache.after.back();
      const size_t k = cache.after.size();
      cache.tab.resize((k + 1) * per);
      eqbf_perm_tables(&g, N, 1, nperm, cache.tab.data() + k * per);
      cache.after.push_back(g);
    }
    states.assign(cache.after.begin() + pos, cache.after.begin() + pos + (long)n);
    return cache.tab.data() + (size_t)pos * per;
  }
  local.resize(n * per);
  states.clear();
  eqbf_rng g = r;
  for (size_t j = 0; j < n; ++j) {
    eqbf_perm_tables(&g, N, 1, nperm, local.data() + j * per);
    states.push_back(g);
  }
  return local.data();
}

/* Near-tie bounds of the FP64 re-check: absolute on the log10 BF statistic
 * (about 40 times the mixed path's largest measured error, 1.3e-6), relative
 * on the sep p-value statistic. */
constexpr double kTieJoin = 5e-5, kTieSep = 1e-6;

/* Statistics of `genes` (chunk indices, in order) under the reference's
 * sequential use of one generator: each gene starts from the identity and
 * draws its shuffles from where the previous gene stopped; --trick 1 and
 * NaN statistics change how many draws a gene takes and how many the next
 * gene is allowed (nb_permutations is shared), so tables are made for the
 * rest of the chunk (in batches of at most the cache's size) and made again
 * from the first gene whose use differed.  truth[g] (in/out): the true
 * statistics from the association results; the library replaces them by FP64
 * values for the genes with a permutation statistic within tie_bound of them
 * and recomputes those statistics in FP64.  count(g, stats, nb) runs the
 * counting loop of gene g and returns the number of shuffles it drew. */
template <class Count>
void run_perms(Engine& eng, const Chunk& ch, const eqbf_perm_opts& po, const std::vector<int>& genes, int64_t N,
               eqbf_rng& rng, int64_t& nb_perm, PermTableCache& cache, double tie_bound, std::vector<double>& truth,
               Count count) {
  size_t k = 0;
  std::vector<int32_t> local;
  std::vector<eqbf_rng> states;
  std::vector<double> tr(ch.genes.size());
  while (k < genes.size()) {
    const int64_t nb_cur = nb_perm;
    if (nb_cur <= 0) {
      for (size_t j = k; j < genes.size(); ++j) count(genes[j], nullptr, nb_cur);
      return;
    }
    const size_t rest = genes.size() - k;
    const size_t n = std::min(rest, PermTableCache::cap_genes(N, nb_cur));
    const int32_t* tables = tables_from(cache, rng, N, nb_cur, n, k == 0, local, states);
    std::vector<int32_t> gene_table(ch.genes.size(), -1);
    for (size_t j = 0; j < n; ++j) gene_table[genes[k + j]] = (int32_t)j;
    std::vector<double> stats((size_t)ch.genes.size() * nb_cur, kNaN);
    int64_t n_re = 0;
    tr = truth;
    eng.check(eqbf_perm_stats_fp64_ties(eng.ctx, &ch.c, &po, tables, (int32_t)n, gene_table.data(), nb_cur,
                                       tie_bound, stats.data(), tr.data(), &n_re));
    for (size_t j = 0; j < n; ++j) truth[genes[k + j]] = tr[genes[k + j]];
    size_t j = 0;
    for (; j < n; ++j) {
      const int64_t before = nb_perm;
      const eqbf_rng at = rng;
      const int64_t drawn = count(genes[k + j], &stats[(size_t)genes[k + j] * nb_cur], nb_cur);
      if (drawn == nb_cur) {
        rng = states[j];
      } else { /* stopped early: the stream advanced by the shuffles drawn */
        rng = at;
        std::vector<int32_t> perm(N);
        std::iota(perm.begin(), perm.end(), 0);
        for (int64_t p = 0; p < drawn; ++p) eqbf_rng_shuffle(&rng, perm.data(), N);
      }
      if (drawn != nb_cur || nb_perm != before) {
        ++j;
        break;
      }
    }
    k += j;
  }
}

/* --inss (run(), eqtlbma_bf.cpp:1496-1503 with loadSummaryStats): the pairs
 * of the summary-statistics files, genes in name order, chunks of --wrtsize
 * genes through eqbf_abf_from_sstats, then the same writers.  The sumstats
 * files get their header only: the reference writes a row only for a gene
 * with expression levels and a SNP with genotypes in the subgroup
 * (Gene::HasAtLeastOneCisSnp, gene.cpp:218-232), which given statistics never
 * provide. */
int run_inss(const Options& o, time_t t_start, const char* prog) {
  const int verbose = o.verbose;
  InssData in;
  if (!load_inss(o.inss, verbose, in)) return EXIT_SUCCESS;
  std::vector<double> phiL, omaL, phiS, omaS;
  load_grid(o.gridL, phiL, omaL, verbose);
  load_grid(o.gridS, phiS, omaS, verbose);
  const int S = (int)in.sbgrps.size();
  const int bfs = bfs_code(o.bfs);
  Engine eng(o.device);
  if (o.kernel_times) eng.check(eqbf_profile(eng.ctx, 1));
  eng.check(eqbf_set_grids(eng.ctx, phiL.data(), omaL.data(), (int32_t)phiL.size(), phiS.data(), omaS.data(),
                           (int32_t)phiS.size()));
  Data d; /* subgroup names for the writers */
  for (auto& n : in.sbgrps) {
    d.sbgrps.emplace_back();
    d.sbgrps.back().name = n;
  }
  Writer wr{o, d, {}, (int)phiL.size(), (int)phiS.size(), "wb"};
  if (bfs != EQBF_BFS_GEN) {
    const int64_t nc = eqbf_n_configs(bfs, S);
    if (nc < 0) fail("--bfs all with " + std::to_string(S) + " subgroups is not supported");
    std::vector<uint64_t> masks(nc);
    eqbf_config_masks(bfs, S, masks.data());
    for (uint64_t m : masks) {
      std::string l;
      for (int s = 0; s < S; ++s)
        if ((m >> s) & 1ull) l += (l.empty() ? "" : "-") + std::to_string(s + 1);
      wr.cfg_labels.push_back(l);
    }
  }
  if (o.save_sstats) wr.sumstats(nullptr, nullptr);
  wr.raw(nullptr, nullptr);
  if (o.save_weighted_abfs) wr.avg(nullptr, nullptr);
  wr.mode = "ab";
  if (verbose > 0)
    std::cout << "test for association between each pair gene-SNP ..." << std::endl
              << "analysis=" << o.analysis << " likelihood=" << o.likelihood << " error_model=" << o.error_model
              << " precision=fp64" << std::endl;
  std::vector<const std::pair<const std::string, InssGene>*> all;
  for (auto& kv : in.genes) all.push_back(&kv);
  size_t n_genes_done = 0, n_pairs_done = 0;
  for (size_t g0 = 0; g0 < all.size(); g0 += (size_t)o.write_group_size) {
    const size_t g1 = std::min(all.size(), g0 + (size_t)o.write_group_size);
    const int G = (int)(g1 - g0);
    std::vector<Gene> genes(G);
    std::vector<Snp> snps;
    Chunk ch;
    ch.pair_offset.assign(G + 1, 0);
    for (int g = 0; g < G; ++g) {
      genes[g].name = all[g0 + g]->first;
      ch.pair_offset[g + 1] = ch.pair_offset[g] + (int64_t)all[g0 + g]->second.snps.size();
    }
    const int64_t NP = ch.pair_offset[G];
    snps.resize(NP);
    Results r;
    r.S = S;
    r.NP = NP;
    r.n_raw = eqbf_n_raw(bfs, S, (int32_t)phiL.size(), (int32_t)phiS.size());
    r.n_w = eqbf_n_weighted(bfs, S);
    r.n.assign((size_t)S * NP, 0);
    r.nsub.assign(NP, 0);
    r.raw.assign((size_t)r.n_raw * NP, kNaN);
    r.w.assign((size_t)r.n_w * NP, kNaN);
    std::vector<int64_t> n((size_t)S * NP, 0);
    std::vector<double> sig((size_t)S * NP, kNaN), beta((size_t)S * NP, kNaN), se((size_t)S * NP, kNaN);
    for (int g = 0; g < G; ++g) {
      ch.genes.push_back(&genes[g]);
      const InssGene& ig = all[g0 + g]->second;
      for (size_t k = 0; k < ig.snps.size(); ++k) {
        const int64_t p = ch.pair_offset[g] + (int64_t)k;
        snps[p].name = ig.snps[k];
        ch.snps.push_back(&snps[p]);
        ch.pair_snp.push_back((int32_t)p);
        for (int s = 0; s < S; ++s) {
          const SstatsEntry& e = ig.st[k][s];
          if (!e.set) continue;
          ++r.nsub[p];
          const size_t i = (size_t)s * NP + p;
          n[i] = e.n;
          sig[i] = e.sigmahat;
          beta[i] = e.betahat;
          se[i] = e.sebetahat;
        }
      }
      ++n_genes_done;
      n_pairs_done += ig.snps.size();
    }
    if (NP > 0) {
      eqbf_sstats_in si{NP, S, n.data(), sig.data(), beta.data(), se.data()};
      eqbf_assoc_out out{};
      out.n = r.n.data();
      out.raw_abf = r.raw.data();
      out.w_abf = r.w.data();
      eng.check(eqbf_abf_from_sstats(eng.ctx, &si, bfs, EQBF_F64, &out));
    }
    wr.raw(&ch, &r);
    if (o.save_weighted_abfs) wr.avg(&ch, &r);
  }
  if (o.kernel_times) print_kernel_times(eng);
  if (verbose > 0) {
    std::cout << "nb of analyzed gene-SNP pairs: " << n_pairs_done << " (" << n_genes_done << " genes)" << std::endl;
    std::cout << "END " << prog << "\nelapsed -> " << (long)(time(nullptr) - t_start) << " sec" << std::endl;
  }
  return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse_options(argc, argv);
  const int verbose = o.verbose;
  time_t t_start = time(nullptr);
  if (verbose > 0) std::cout << "START " << argv[0] << "\ncmd-line: " << o.cmdline << std::endl;

  if (!o.inss.empty()) return run_inss(o, t_start, argv[0]);

  LoadOptions lo;
  lo.geno_list = o.geno;
  lo.snp_coords = o.scoord;
  lo.exp_list = o.exp;
  lo.gene_coords = o.gcoord;
  lo.covar_list = o.covar;
  lo.snps_to_keep = o.snps_to_keep;
  lo.subgroups_to_keep = o.subgroups_to_keep;
  lo.error_model = o.analysis == "join" ? o.error_model : "uvlr";
  lo.min_maf = o.min_maf;
  lo.anchor = o.anchor;
  lo.radius = o.radius;
  lo.verbose = verbose;
  Data d;
  if (!load_data(lo, d)) return EXIT_SUCCESS;

  std::vector<double> phiL, omaL, phiS, omaS;
  load_grid(o.gridL, phiL, omaL, verbose);
  load_grid(o.gridS, phiS, omaS, verbose);
  const bool join = o.analysis == "join";
  const bool mvlr = join && o.error_model == "mvlr";
  const bool hybrid = join && o.error_model == "hybrid";
  const int S = (int)d.sbgrps.size();
  const int bfs = bfs_code(o.bfs);
  Engine eng(o.device);
  if (o.kernel_times) eng.check(eqbf_profile(eng.ctx, 1));
  eng.check(eqbf_set_grids(eng.ctx, phiL.data(), omaL.data(), (int32_t)phiL.size(), phiS.data(), omaS.data(),
                           (int32_t)phiS.size()));
  Shared sh = make_shared(d);
  if (o.need_qnorm && o.nb_permutations > 0) {
    /* the reference normalises every permuted phenotype vector over its own
     * members; this build normalises once, which is the same only when every
     * sample has an expression column and a genotype column in each subgroup */
    for (int s = 0; s < S; ++s)
      for (int i = 0; i < sh.N; ++i)
        if (d.sbgrps[s].exp_col[i] < 0 || !sh.geno_in[(size_t)s * sh.N + i])
          fail("--qnorm with permutations needs every sample in every subgroup in this build");
  }

  Writer wr{o, d, {}, (int)phiL.size(), (int)phiS.size(), "wb"};
  if (join && bfs != EQBF_BFS_GEN) {
    const int64_t nc = eqbf_n_configs(bfs, S);
    if (nc < 0) fail("--bfs all with " + std::to_string(S) + " subgroups is not supported");
    std::vector<uint64_t> masks(nc);
    eqbf_config_masks(bfs, S, masks.data());
    for (uint64_t m : masks) {
      std::string l;
      for (int s = 0; s < S; ++s)
        if ((m >> s) & 1ull) l += (l.empty() ? "" : "-") + std::to_string(s + 1);
      wr.cfg_labels.push_back(l);
    }
  }
  const bool perm_join = join && o.nb_permutations > 0 && o.permbf != "none";
  const bool perm_sep = !join && o.nb_permutations > 0 && o.perm_sep != 0;
  auto write_all = [&](const Chunk* ch, const Results* r, const std::vector<PermRes>* pj,
                       const std::vector<std::vector<PermRes>>* ps) {
    if (!join || (o.save_sstats && !mvlr)) wr.sumstats(ch, r);
    if (perm_sep) {
      for (int s = 0; s < (o.perm_sep == 1 ? 1 : S); ++s) {
        const std::string name = o.out + "_sepPermPvals" + (o.perm_sep == 1 ? "" : "_" + d.sbgrps[s].name) + ".txt.gz";
        GzWriter f(name, wr.mode.c_str());
        std::string t;
        if (!ch) {
          f.put("# seed=" + std::to_string(o.seed) + "\ngene\tnb.snps\tsep.perm.pval\tnb.permutations\ttrue.min.pval\n");
          continue;
        }
        for (size_t g = 0; g < ch->genes.size(); ++g) {
          const int64_t np = ch->pair_offset[g + 1] - ch->pair_offset[g];
          if (np <= 0) continue;
          int64_t nsnp = np;
          if (o.perm_sep == 2) {
            nsnp = 0;
            for (int64_t p = ch->pair_offset[g]; p < ch->pair_offset[g + 1]; ++p) nsnp += r->n[(size_t)s * r->NP + p] > 0;
          }
          const PermRes& pr = (*ps)[s][g];
          t += ch->genes[g]->name + "\t";
          put_int(t, nsnp);
          t += '\t';
          put_gen(t, pr.pval);
          t += '\t';
          put_int(t, pr.used);
          t += '\t';
          put_sci(t, pr.true_stat);
          t += '\n';
        }
        f.put(t);
      }
    }
    if (join) {
      wr.raw(ch, r);
      if (o.save_weighted_abfs) wr.avg(ch, r);
    }
    if (join && o.nb_permutations > 0) {
      GzWriter f(o.out + "_joinPermPvals.txt.gz", wr.mode.c_str());
      if (!ch) {
        f.put("# perm.bf=" + o.permbf + " seed=" + std::to_string(o.seed) +
              "\ngene\tnb.snps\tjoin.perm.pval\tnb.permutations\ttrue.l10abf\tmed.perm.l10abf\n");
      } else {
        std::string t;
        for (size_t g = 0; g < ch->genes.size(); ++g) {
          const int64_t np = ch->pair_offset[g + 1] - ch->pair_offset[g];
          if (np <= 0) continue;
          const PermRes& pr = (*pj)[g];
          t += ch->genes[g]->name + "\t";
          put_int(t, np);
          t += '\t';
          put_gen(t, pr.pval);
          t += '\t';
          put_int(t, pr.used);
          t += '\t';
          put_sci(t, pr.true_stat);
          t += '\t';
          put_sci(t, pr.median);
          t += '\n';
        }
        f.put(t);
      }
    }
  };
  write_all(nullptr, nullptr, nullptr, nullptr);
  wr.mode = "ab";

  if (verbose > 0) {
    std::cout << "test for association between each pair gene-SNP ..." << std::endl
              << "analysis=" << o.analysis << " likelihood=" << o.likelihood << " error_model=" << o.error_model;
    if (o.error_model != "uvlr") std::cout << " prop_cov_errors=" << o.prop_cov_errors;
    std::cout << " anchor=" << o.anchor << " radius=" << o.radius << " precision=" << o.precision << std::endl;
  }
  std::vector<const Gene*> all;
  for (auto& kv : d.genes) all.push_back(&kv.second);
  int64_t nb_perm = (int64_t)o.nb_permutations;
  PermTableCache tcache;
  size_t n_genes_done = 0, n_pairs_done = 0;
  /* fp64: the plain path's arithmetic, all genes of a chunk in one launch */
  const int prec = o.precision == "fp64" ? EQBF_PREC_FP64_BATCHED : EQBF_PREC_MIXED;
  for (size_t g0 = 0; g0 < all.size(); g0 += (size_t)o.write_group_size) {
    const size_t g1 = std::min(all.size(), g0 + (size_t)o.write_group_size);
    Chunk ch;
    build_chunk(d, sh, o, std::vector<const Gene*>(all.begin() + g0, all.begin() + g1), ch);
    const int G = (int)ch.genes.size();
    Results r;
    r.S = S;
    r.NP = ch.n_pairs();
    r.n.assign((size_t)S * std::max<int64_t>(r.NP, 1), 0);
    r.sstats.assign((size_t)5 * S * std::max<int64_t>(r.NP, 1), kNaN);
    if (join) {
      r.n_raw = eqbf_n_raw(bfs, S, (int32_t)phiL.size(), (int32_t)phiS.size());
      r.n_w = eqbf_n_weighted(bfs, S);
      r.raw.assign((size_t)r.n_raw * std::max<int64_t>(r.NP, 1), kNaN);
      r.w.assign((size_t)r.n_w * std::max<int64_t>(r.NP, 1), kNaN);
    }
    if (r.NP > 0) {
      eqbf_assoc_out out{};
      out.n = r.n.data();
      out.sstats = r.sstats.data();
      out.raw_abf = join ? r.raw.data() : nullptr;
      out.w_abf = join ? r.w.data() : nullptr;
      if (mvlr) {
        eqbf_mvlr_opts mo{bfs, o.prop_cov_errors};
        out.sstats = nullptr;
        eng.check(eqbf_assoc_mvlr(eng.ctx, &ch.c, &mo, &out));
      } else if (hybrid) {
        eqbf_hybrid_opts ho{bfs, o.prop_cov_errors, sh.Qmax > 0 ? sh.covars_pos.data() : nullptr,
                            ch.expl_raw.empty() ? nullptr : ch.expl_raw.data()};
        eng.check(eqbf_assoc_hybrid(eng.ctx, &ch.c, &ho, &out));
      } else {
        eqbf_assoc_opts ao{join ? EQBF_ANALYSIS_JOIN : EQBF_ANALYSIS_SEP, bfs, prec, EQBF_F64};
        eng.check(eqbf_assoc_uvlr(eng.ctx, &ch.c, &ao, &out));
      }
    }
    for (int g = 0; g < G; ++g)
      if (ch.pair_offset[g + 1] > ch.pair_offset[g]) {
        ++n_genes_done;
        n_pairs_done += (size_t)(ch.pair_offset[g + 1] - ch.pair_offset[g]);
      }

    std::vector<PermRes> pj(G);
    std::vector<std::vector<PermRes>> ps(S, std::vector<PermRes>(G));
    if (perm_join || perm_sep) {
      const int64_t N = sh.N;
      /* makePermutationsJoin (eqtlbma_bf.cpp:851-858): genes with a cis SNP;
       * with --error mvlr or hybrid also expression in every subgroup */
      std::vector<int> genes;
      for (int g = 0; g < G; ++g) {
        if (ch.pair_offset[g + 1] <= ch.pair_offset[g]) continue;
        bool all_expr = true;
        for (int s = 0; s < S; ++s) all_expr &= ch.genes[g]->exp_row[s] >= 0;
        if ((mvlr || hybrid) && !all_expr) continue;
        genes.push_back(g);
      }
      eqbf_rng rng, rngT;
      auto reseed = [&]() {
        eqbf_rng_seed(&rng, o.seed);
        if (o.trick != 0) eqbf_rng_seed(&rngT, o.seed);
      };
      eqbf_perm_opts po{};
      po.precision = prec;
      po.error_model = mvlr ? EQBF_ERROR_MVLR : (hybrid ? EQBF_ERROR_HYBRID : EQBF_ERROR_UVLR);
      po.fiterr = o.prop_cov_errors;
      if (hybrid) {
        po.covars_by_position = sh.Qmax > 0 ? sh.covars_pos.data() : nullptr;
        po.explevels_pairs = ch.expl_raw.empty() ? nullptr : ch.expl_raw.data();
      }
      std::vector<double> truth;
      if (perm_join) {
        reseed();
        po.analysis = EQBF_ANALYSIS_JOIN;
        po.pbf = pbf_code(o.permbf);
        po.use_max_bf = o.use_max_bf;
        po.sep_sbgrp = -1;
        const int64_t nc = (int64_t)wr.cfg_labels.size();
        const size_t row = o.permbf == "gen" ? 0 : (o.permbf == "gen-sin" ? 3 + nc : 3 + nc + 1);
        for (int g : genes) {
          double t;
          if (o.use_max_bf) { /* Gene::FindMaxTrueL10Abf (gene.cpp:575-583) */
            t = -std::numeric_limits<double>::infinity();
            for (int64_t p = ch.pair_offset[g]; p < ch.pair_offset[g + 1]; ++p)
              if (r.w[row * r.NP + p] > t) t = r.w[row * r.NP + p];
          } else { /* Gene::AvgTrueL10Abfs (gene.cpp:587-596) */
            std::vector<double> v;
            for (int64_t p = ch.pair_offset[g]; p < ch.pair_offset[g + 1]; ++p)
              if (!std::isnan(r.w[row * r.NP + p])) v.push_back(r.w[row * r.NP + p]);
            t = l10_avg(v);
          }
          pj[g].true_stat = t;
        }
        truth.assign(G, kNaN);
        for (int g : genes) truth[g] = pj[g].true_stat;
        run_perms(eng, ch, po, genes, N, rng, nb_perm, tcache, kTieJoin, truth,
                  [&](int g, const double* st, int64_t n_all) -> int64_t {
          PermRes& pr = pj[g];
          pr.done = true;
          (void)n_all;
          if (st) pr.true_stat = truth[g]; /* FP64 where a permutation came near it */
          static const double none = 0.0; /* no statistic is read when no permutation is left */
          return eqbf_perm_pvalue(st ? st : &none, pr.true_stat, 0, &nb_perm, o.trick, (int64_t)o.trick_cutoff,
                                  o.trick ? &rngT : nullptr, &pr.pval, &pr.used, &pr.median);
        });
      } else {
        po.analysis = EQBF_ANALYSIS_SEP;
        const int passes = o.perm_sep == 1 ? 1 : S;
        for (int pass = 0; pass < passes; ++pass) {
          reseed();
          po.sep_sbgrp = o.perm_sep == 1 ? -1 : pass;
          for (int g : genes) { /* Gene::FindMinTruePvalue* (gene.cpp:369-378, 475-488) */
            double t = 1.0;
            for (int64_t p = ch.pair_offset[g]; p < ch.pair_offset[g + 1]; ++p)
              for (int s = 0; s < S; ++s) {
                if (o.perm_sep == 2 && s != pass) continue;
                if (o.perm_sep == 1 && ch.genes[g]->exp_row[s] < 0) continue;
                const double pv = r.sstats[((size_t)4 * S + s) * r.NP + p];
                if (r.n[(size_t)s * r.NP + p] > 0 && pv < t) t = pv;
              }
            ps[pass][g].true_stat = t;
          }
          truth.assign(G, kNaN);
          for (int g : genes) truth[g] = ps[pass][g].true_stat;
          run_perms(eng, ch, po, genes, N, rng, nb_perm, tcache, kTieSep, truth,
                    [&](int g, const double* st, int64_t) -> int64_t {
            PermRes& pr = ps[pass][g];
            pr.done = true;
            if (st) pr.true_stat = truth[g]; /* FP64 where a permutation came near it */
            static const double none = 0.0; /* no statistic is read when no permutation is left */
            return eqbf_perm_pvalue(st ? st : &none, pr.true_stat, 1, &nb_perm, o.trick, (int64_t)o.trick_cutoff,
                                    o.trick ? &rngT : nullptr, &pr.pval, &pr.used, nullptr);
          });
        }
      }
    }
    write_all(&ch, &r, &pj, &ps);
  }
  if (o.kernel_times) print_kernel_times(eng);
  if (verbose > 0) {
    std::cout << "nb of analyzed gene-SNP pairs: " << n_pairs_done << " (" << n_genes_done << " genes)" << std::endl;
    std::cout << "END " << argv[0] << "\nelapsed -> " << (long)(time(nullptr) - t_start) << " sec" << std::endl;
  }
  return EXIT_SUCCESS;
}
