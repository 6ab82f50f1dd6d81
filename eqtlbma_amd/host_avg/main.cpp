/*
 * eqtlbma_avg_bfs: averages the raw log10 ABFs that eqtlbma_bf writes
 * (--outraw) over the grid only, or over the grid and the configurations or
 * types, and computes posteriors, on the GPU through libeqbf_hip.so
 * (eqbf_avg_grid, eqbf_avg_bfs).
 *
 * Mirrors the reference program (src/eqtlbma_avg_bfs.cpp): the same
 * options, checks and messages (parseCmdLine, :112-365), the weight loaders
 * (:880-1010: grid weights one per line, first token; configuration weights
 * "name<sep>value" with the subgroups of a name split at '-'; type and
 * subgroup weights from lines whose name contains "type" / "subgroup"),
 * the line rules (parseLines, :1103-1193: header "gene snp config", only the
 * configurations of the weights file, or "1".."S" for the types model; the
 * first |grid weights| values of a row; a gene's and a SNP's rows are
 * consecutive), the grid indices of --gtk used as 0-based positions (as the
 * code does), SNPs sorted by decreasing BF per gene with --bestsnp 0/1/2,
 * and the writers (:834-879, :1195-1295): "%.6e" for the grid-only file,
 * "%.4e" and a "#pi0=" line for the other.
 * Where the reference's behaviour is undefined this program prints NaN: a
 * SNP with fewer configuration rows than weights (its missing rows are NaN),
 * --bestdim without --save post (no posterior of the best dimension),
 * subgroup posterior columns beyond log2(dim + 1) for the configs model.
 * Ties and NaNs in the SNP ordering: a stable sort, NaN last.
 */
#include <getopt.h>
#include <glob.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "../host/util.hpp"
#include "eqbf.h"

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();
const size_t kUnset = (size_t)-1;

struct Options {
  int verbose = 1, threads = 1, best_snps = 0;
  size_t nsubgrp = kUnset, dim = kUnset;
  std::string in, gwts, model = "configs", cwts, tswts, out, genes, snps, gene_snp;
  std::vector<size_t> gtk;
  std::vector<std::string> save, post;
  double pi0 = kNaN;
  bool best_dim = false, all_dims = false;
};

std::vector<std::string> split(const std::string& s, const char* delims) {
  std::vector<std::string> out;
  host::tokenize(s, delims, out);
  return out;
}

/* split at one character keeping empty tokens (std::getline, utils_io.cpp:108-118) */
std::vector<std::string> split_char(const std::string& s, char d) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i <= s.size()) {
    const size_t j = s.find(d, i);
    if (j == std::string::npos) {
      if (i < s.size()) out.push_back(s.substr(i));
      break;
    }
    out.push_back(s.substr(i, j - i));
    i = j + 1;
  }
  return out;
}

bool has(const std::vector<std::string>& v, const std::string& x) { return std::find(v.begin(), v.end(), x) != v.end(); }

std::string cmd_line(int argc, char** argv) {
  std::string s;
  for (int i = 0; i < argc; ++i) s += (i ? " " : "") + std::string(argv[i]);
  return s;
}

void help(const char* prog) {
  std::cout << "`" << prog << "' averages the raw BFs over the grid only,\n"
            << "or over both the grid and the configurations,\n"
            << "and it can also compute posteriors (on the GPU).\n\n"
            << "Usage: " << prog << " [OPTIONS] ...\n\n"
            << "Options:\n"
            << "  -h, --help\tdisplay the help and exit\n"
            << "  -V, --version\toutput version information and exit\n"
            << "  -v, --verbose\tverbosity level (0/default=1/2/3)\n"
            << "      --in\tpattern to glob '_l10abfs_raw' files from 'eqtlbma_bf'\n"
            << "      --gwts\tfile with grid weights (one per line, only the value)\n"
            << "      --gtk\tind-ex/icies of grid weights to keep (all by default), e.g. '1+3+5+7+9'\n"
            << "      --model\twhich model (default=configs/types)\n"
            << "      --nsubgrp\tnumber of subgroups\n"
            << "      --dim\tdimension of the model (nb of active configs or types)\n"
            << "      --cwts\tfile with configuration weights (one per line, name<sep>value)\n"
            << "      --tswts\tfile with type and subgroup weights (one per line, name<sep>value)\n"
            << "      --save\tprecise what to save (bf/post/bf+post)\n"
            << "      --pi0\tproba for a gene to have no eQTL in any subgroup\n"
            << "      --post\tsave various kinds of posterior probabilities (e.g. 'a+b')\n"
            << "      --gene\tfile with subset of gene(s) to keep (one per line)\n"
            << "      --snp\tfile with subset of snp(s) to keep (one per line)\n"
            << "      --gene-snp\tfile with subset of gene-snp pai(s) to keep (gene<tab>snp, one per line)\n"
            << "      --bestsnp\treport the best SNP(s) per gene (0: all, 1: the best, 2: up to 0.95 of post.the)\n"
            << "      --bestdim\treport the best config/type per SNP (and its posterior)\n"
            << "      --alldim\treport also BF and/or posterior for all dimensions (configs or types)\n"
            << "      --out\tname of the output file (gzipped)\n"
            << "      --thread\tnumber of threads (default=1; the GPU does the averaging)\n"
            << std::endl;
}

[[noreturn]] void bad(int argc, char** argv, const std::string& msg) {
  std::cerr << "cmd-line: " << cmd_line(argc, argv) << std::endl << std::endl
            << "ERROR: " << msg << std::endl << std::endl;
  help(argv[0]);
  std::exit(EXIT_FAILURE);
}

Options parse(int argc, char** argv) {
  Options o;
  static const struct option longopts[] = {
      {"help", no_argument, 0, 'h'},          {"version", no_argument, 0, 'V'},
      {"verbose", required_argument, 0, 'v'}, {"in", required_argument, 0, 0},
      {"gwts", required_argument, 0, 0},      {"gtk", required_argument, 0, 0},
      {"model", required_argument, 0, 0},     {"nsubgrp", required_argument, 0, 0},
      {"dim", required_argument, 0, 0},       {"cwts", required_argument, 0, 0},
      {"tswts", required_argument, 0, 0},     {"save", required_argument, 0, 0},
      {"pi0", required_argument, 0, 0},       {"post", required_argument, 0, 0},
      {"gene", required_argument, 0, 0},      {"snp", required_argument, 0, 0},
      {"gene-snp", required_argument, 0, 0},  {"bestsnp", required_argument, 0, 0},
      {"bestdim", no_argument, 0, 0},         {"alldim", no_argument, 0, 0},
      {"out", required_argument, 0, 0},       {"thread", required_argument, 0, 0},
      {0, 0, 0, 0}};
  while (true) {
    int idx = 0;
    const int c = getopt_long(argc, argv, "hVv:", longopts, &idx);
    if (c == -1) break;
    if (c == 'h') {
      help(argv[0]);
      std::exit(0);
    }
    if (c == 'V') {
      std::cout << argv[0] << " (GPU build)" << std::endl;
      std::exit(0);
    }
    if (c == 'v') {
      o.verbose = std::atoi(optarg);
      continue;
    }
    if (c != 0) {
      std::printf("\n");
      help(argv[0]);
      std::exit(EXIT_FAILURE);
    }
    const std::string n = longopts[idx].name;
    const char* a = optarg;
    if (n == "in") o.in = a;
    else if (n == "gwts") o.gwts = a;
    else if (n == "gtk") {
      for (const std::string& t : split(a, "+")) o.gtk.push_back((size_t)std::atoi(t.c_str()));
    } else if (n == "model") o.model = a;
    else if (n == "nsubgrp") o.nsubgrp = (size_t)std::atol(a);
    else if (n == "dim") o.dim = (size_t)std::atoi(a);
    else if (n == "cwts") o.cwts = a;
    else if (n == "tswts") o.tswts = a;
    else if (n == "save") o.save = split(a, "+");
    else if (n == "pi0") o.pi0 = std::atof(a);
    else if (n == "post") o.post = split(a, "+");
    else if (n == "gene") o.genes = a;
    else if (n == "snp") o.snps = a;
    else if (n == "gene-snp") o.gene_snp = a;
    else if (n == "bestsnp") o.best_snps = std::atoi(a);
    else if (n == "bestdim") o.best_dim = true;
    else if (n == "alldim") o.all_dims = true;
    else if (n == "out") o.out = a;
    else if (n == "thread") o.threads = std::atoi(a);
  }
  /* checks in the reference's order and words (:281-365) */
  if (o.in.empty()) bad(argc, argv, "missing compulsory option --bf-out");
  if (o.gwts.empty()) bad(argc, argv, "missing compulsory option --gwts");
  if (!host::file_exists(o.gwts)) bad(argc, argv, "can't find file " + o.gwts);
  if (!o.model.empty() && o.model != "configs" && o.model != "types")
    bad(argc, argv, "--model " + o.model + " is unvalid");
  if (!o.model.empty() && o.nsubgrp == kUnset) bad(argc, argv, "missing compulsory option --nsubgrp");
  if (!o.model.empty() && o.dim == kUnset) bad(argc, argv, "missing compulsory option --dim");
  if (!o.cwts.empty() && !host::file_exists(o.cwts)) bad(argc, argv, "can't find file " + o.cwts);
  if (!o.tswts.empty() && !host::file_exists(o.tswts)) bad(argc, argv, "can't find file " + o.tswts);
  if (o.out.empty()) bad(argc, argv, "missing compulsory option --hm-in");
  if (has(o.save, "post") && o.post.empty()) bad(argc, argv, "missing compulsory option --post with --save post");
  if (!std::isnan(o.pi0) && o.post.empty()) bad(argc, argv, "missing compulsory option --post with --pi0");
  if (std::isnan(o.pi0) && !o.post.empty()) bad(argc, argv, "missing compulsory option --pi0 with --post");
  if (o.threads <= 0) bad(argc, argv, "--thread " + std::to_string(o.threads) + " is invalid");
  if (o.best_snps != 0 && o.best_snps != 1 && o.best_snps != 2)
    bad(argc, argv, "--bestsnp " + std::to_string(o.best_snps) + " is invalid");
  if (o.best_snps == 2 && !has(o.save, "post")) bad(argc, argv, "--bestsnp 2 can only be used with --save post");
  if (!o.genes.empty() && !host::file_exists(o.genes)) bad(argc, argv, "can't find file " + o.genes);
  if (!o.snps.empty() && !host::file_exists(o.snps)) bad(argc, argv, "can't find file " + o.snps);
  if (!o.snps.empty()) std::cout << "WARNING: using --snp can change the gene-level BFs and posteriors" << std::endl;
  if (!o.gene_snp.empty() && !host::file_exists(o.gene_snp)) bad(argc, argv, "can't find file " + o.gene_snp);
  if (!o.gene_snp.empty())
    std::cout << "WARNING: using --gene-snp can change the gene-level BFs and posteriors" << std::endl;
  return o;
}

std::string sci(double v, int prec) {
  char b[64];
  std::snprintf(b, sizeof b, "%.*e", prec, v);
  return b;
}

/* ---- weights (:880-1010) -------------------------------------------------------- */
struct Weights {
  std::vector<double> grid;
  std::vector<int32_t> keep;
  std::vector<std::string> cfg_names;
  std::vector<double> cfg;
  std::vector<std::vector<std::string>> cfg_subs;
  std::vector<double> type;
  std::vector<std::vector<double>> sub;
};

void load_grid_weights(const Options& o, Weights& w) {
  if (o.verbose > 0) std::cout << "load grid weights ..." << std::endl;
  host::LineReader r(o.gwts);
  std::string line;
  while (r.next(line)) {
    const std::vector<std::string> t = split(line, " \t");
    w.grid.push_back(t.empty() ? 0.0 : std::atof(t[0].c_str()));
  }
  if (!r.complete()) host::fail("can't read successfully file " + o.gwts + " up to the end");
  if (o.verbose > 0)
    for (size_t i = 0; i < w.grid.size(); ++i)
      std::cout << "grid weight " << i + 1 << ": " << sci(w.grid[i], 4) << std::endl;
  if (o.gtk.empty()) {
    if (o.verbose > 0) std::cout << "use all grid weights" << std::endl;
    for (size_t i = 0; i < w.grid.size(); ++i) w.keep.push_back((int32_t)i);
  } else {
    if (*std::max_element(o.gtk.begin(), o.gtk.end()) >= w.grid.size())
      host::fail("--gtk doesn't correspond to --gwts");
    for (size_t i : o.gtk) w.keep.push_back((int32_t)i);
    if (o.verbose > 0) std::cout << "use only " << w.keep.size() << " grid weights" << std::endl;
  }
}

void load_config_weights(const Options& o, Weights& w) {
  if (!o.cwts.empty()) {
    if (o.verbose > 0) std::cout << "load config weights ..." << std::endl;
    host::LineReader r(o.cwts);
    std::string line;
    while (r.next(line)) {
      const std::vector<std::string> t = split(line, " \t,");
      w.cfg_names.push_back(t.size() > 0 ? t[0] : std::string());
      w.cfg.push_back(t.size() > 1 ? std::atof(t[1].c_str()) : 0.0);
      w.cfg_subs.push_back(split_char(w.cfg_names.back(), '-'));
    }
    if (!r.complete()) host::fail("can't read successfully file " + o.cwts + " up to the end");
    if (w.cfg.size() != o.dim)
      host::fail("file " + o.cwts + " should contain " + std::to_string(o.dim) + " config weights");
    if (o.verbose > 0)
      for (size_t k = 0; k < o.dim; ++k)
        std::cout << "config weight " << k + 1 << " (" << w.cfg_names[k] << "): " << sci(w.cfg[k], 4) << std::endl;
  } else {
    for (size_t s = 0; s < o.nsubgrp && o.nsubgrp != kUnset; ++s) w.cfg_names.push_back(std::to_string(s + 1));
  }
}

void load_type_weights(const Options& o, Weights& w) {
  if (o.tswts.empty()) return;
  if (o.verbose > 0) std::cout << "load type and subgroup weights ..." << std::endl;
  host::LineReader r(o.tswts);
  std::string line;
  size_t it = 0, is = 0;
  while (r.next(line)) {
    const std::vector<std::string> t = split(line, " \t,");
    if (t.empty()) continue;
    const double v = t.size() > 1 ? std::atof(t[1].c_str()) : 0.0;
    if (t[0].find("type") != std::string::npos) {
      w.type.push_back(v);
    } else if (t[0].find("subgroup") != std::string::npos) {
      if (is == 0) w.sub.emplace_back();
      if (it >= w.sub.size()) w.sub.resize(it + 1);
      w.sub[it].push_back(v);
      if (++is == o.nsubgrp) {
        ++it;
        is = 0;
      }
    }
  }
  if (!r.complete()) host::fail("can't read successfully file " + o.tswts + " up to the end");
  if (w.type.size() != o.dim)
    host::fail("file " + o.tswts + " should contain " + std::to_string(o.dim) + " type weights");
  if (w.sub.size() != o.dim)
    host::fail("file " + o.tswts + " should contain " + std::to_string(o.dim) + " types for subgroups-per-type weights");
  for (size_t k = 0; k < o.dim; ++k)
    if (w.sub[k].size() != o.nsubgrp)
      host::fail("file " + o.tswts + " should contain " + std::to_string(o.nsubgrp) +
                 " subgroups for subgroups-per-type " + std::to_string(k + 1) + " weights");
  if (o.verbose > 0)
    for (size_t k = 0; k < o.dim; ++k) {
      std::cout << "type weight " << k + 1 << ": " << sci(w.type[k], 4) << std::endl;
      for (size_t s = 0; s < o.nsubgrp; ++s)
        std::cout << "subgroup " << s + 1 << " per type " << k + 1 << ": " << sci(w.sub[k][s], 4) << std::endl;
    }
}

std::vector<std::string> glob_files(const std::string& pattern) {
  glob_t g;
  std::memset(&g, 0, sizeof g);
  glob(pattern.c_str(), 0, nullptr, &g);
  std::vector<std::string> files(g.gl_pathv, g.gl_pathv + g.gl_pathc);
  globfree(&g);
  return files;
}

void check_header(const std::string& path, const std::vector<std::string>& lines) {
  /* the reference's test needs all three names wrong to fail (:1039-1044) */
  const std::vector<std::string> t = split(lines[0], " \t,");
  if (t.size() >= 3 && t[0] != "gene" && t[1] != "snp" && t[2] != "config")
    host::fail("wrong header line in file " + path);
}

struct Engine {
  eqbf_ctx* ctx = nullptr;
  Engine() {
    if (eqbf_create(&ctx, 0) != EQBF_OK) host::fail("can't create the GPU context");
  }
  ~Engine() { eqbf_destroy(ctx); }
  void check(int rc) const {
    if (rc != EQBF_OK) host::fail(eqbf_last_error(ctx));
  }
};

/* ---- grid only (averageRawBFsOverGrid, :834-879) ------------------------------ */
void average_over_grid(const Options& o, const Weights& w, const std::vector<std::string>& files, Engine& eng) {
  if (o.verbose > 0) std::cout << "average raw BFs over grid only and save them ..." << std::endl;
  host::GzWriter out(o.out, "wb");
  out.put("gene\tsnp\tconfig\tl10abf.grid.avg\n");
  const size_t L = w.grid.size();
  for (const std::string& f : files) {
    const std::vector<std::string> lines = host::read_file_lines(f);
    if (lines.empty()) continue;
    check_header(f, lines);
    const size_t n = lines.size() - 1;
    std::vector<std::string> names(n);
    std::vector<double> raw(n * L, 0.0), avg(n);
    for (size_t i = 1; i < lines.size(); ++i) {
      const std::vector<std::string> t = split(lines[i], " \t,");
      names[i - 1] = (t.size() > 0 ? t[0] : "") + "\t" + (t.size() > 1 ? t[1] : "") + "\t" + (t.size() > 2 ? t[2] : "");
      for (size_t j = 0; j < L; ++j) raw[(i - 1) * L + j] = 3 + j < t.size() ? std::atof(t[3 + j].c_str()) : kNaN;
    }
    eng.check(eqbf_avg_grid(eng.ctx, raw.data(), (int64_t)n, (int32_t)L, w.grid.data(), w.keep.data(),
                            (int32_t)w.keep.size(), avg.data()));
    std::string buf;
    for (size_t i = 0; i < n; ++i) {
      buf += names[i];
      buf += '\t';
      buf += sci(avg[i], 6);
      buf += '\n';
      if (buf.size() > (1u << 20)) {
        out.put(buf);
        buf.clear();
      }
    }
    out.put(buf);
  }
}

/* ---- grid and configurations / types (:1195-1475) -------------------------------- */
struct Parsed {
  std::vector<std::string> gene_names, snp_names;
  std::vector<int64_t> off{0};
  std::vector<double> raw; /* [pairs][R][L] */
};

void parse_lines(const std::string& path, const std::vector<std::string>& lines, const Options& o, const Weights& w,
                 const std::set<std::string>& genes_keep, const std::set<std::string>& snps_keep,
                 const std::map<std::string, std::set<std::string>>& pairs_keep, size_t R, Parsed& d) {
  check_header(path, lines);
  const size_t L = w.grid.size();
  std::string cur_gene, cur_snp;
  bool gene_open = false;
  std::vector<std::vector<double>> rows; /* configuration rows of the current SNP */
  auto push_snp = [&]() {
    d.snp_names.push_back(cur_snp);
    const size_t base = d.raw.size();
    d.raw.resize(base + R * L, kNaN);
    for (size_t j = 0; j < rows.size() && j < R; ++j) std::copy(rows[j].begin(), rows[j].end(), d.raw.begin() + base + j * L);
  };
  for (size_t i = 1; i < lines.size(); ++i) {
    const std::vector<std::string> t = split(lines[i], " \t,");
    if (t.size() < 3) continue;
    if (!has(w.cfg_names, t[2])) continue;
    if (!genes_keep.empty() && !genes_keep.count(t[0])) continue;
    if (!snps_keep.empty() && !snps_keep.count(t[1])) continue;
    if (!pairs_keep.empty()) {
      auto it = pairs_keep.find(t[0]);
      if (it == pairs_keep.end() || !it->second.count(t[1])) continue;
    }
    if (t[0] != cur_gene) {
      if (gene_open) {
        push_snp();
        d.off.push_back((int64_t)d.snp_names.size());
      }
      cur_gene = t[0];
      d.gene_names.push_back(cur_gene);
      gene_open = true;
      rows.clear();
    }
    if (t[1] != cur_snp) {
      if (!rows.empty()) push_snp();
      cur_snp = t[1];
      rows.clear();
    }
    rows.emplace_back(L, kNaN);
    for (size_t j = 0; j < L; ++j) rows.back()[j] = 3 + j < t.size() ? std::atof(t[3 + j].c_str()) : kNaN;
  }
  if (gene_open && !cur_snp.empty()) {
    push_snp();
    d.off.push_back((int64_t)d.snp_names.size());
  } else if (gene_open) {
    d.gene_names.pop_back();
  }
}

void average_over_grid_and_dims(const Options& o, const Weights& w, const std::vector<std::string>& files,
                                Engine& eng) {
  if (o.verbose > 0) std::cout << "average raw BFs over hyperparameters and save them ..." << std::endl;
  const bool types = o.model == "types";
  const size_t D = o.dim, S = o.nsubgrp, L = w.grid.size();
  const size_t R = types ? S : D;
  const bool save_bf = has(o.save, "bf"), save_post = has(o.save, "post");
  const bool pa = has(o.post, "a"), pb = has(o.post, "b"), pc = has(o.post, "c"), pd = has(o.post, "d");
  int want = 0;
  if (save_post) {
    if (pa) want |= EQBF_AVG_POST_A;
    if (pb || o.best_snps == 2) want |= EQBF_AVG_POST_B;
    if (pc) want |= EQBF_AVG_POST_C;
    if (pd) want |= EQBF_AVG_POST_D;
    if (o.best_dim || o.all_dims) want |= EQBF_AVG_POST_DIMS;
  }
  const bool dims_post = (want & EQBF_AVG_POST_DIMS) || (!types && (want & EQBF_AVG_POST_D));
  /* subgroup posteriors of the configs model: as many as log2(dim + 1) (:725) */
  const size_t S_cfg = types ? S : (size_t)std::log2((double)D + 1);
  std::vector<uint64_t> cs(D, 0);
  if (!types)
    for (size_t j = 0; j < D && j < w.cfg_subs.size(); ++j)
      for (size_t s = 0; s < S_cfg && s < 64; ++s)
        if (has(w.cfg_subs[j], std::to_string(s + 1))) cs[j] |= 1ull << s;
  std::vector<double> sw;
  if (types)
    for (size_t k = 0; k < D; ++k) sw.insert(sw.end(), w.sub[k].begin(), w.sub[k].end());

  std::set<std::string> genes_keep, snps_keep;
  std::map<std::string, std::set<std::string>> pairs_keep;
  if (!o.genes.empty()) {
    for (const std::string& l : host::read_file_lines(o.genes)) genes_keep.insert(l);
    if (!genes_keep.empty()) std::cout << "genes to keep: " << genes_keep.size() << std::endl;
  }
  if (!o.snps.empty()) {
    for (const std::string& l : host::read_file_lines(o.snps)) snps_keep.insert(l);
    if (!snps_keep.empty()) std::cout << "snps to keep: " << snps_keep.size() << std::endl;
  }
  if (!o.gene_snp.empty()) {
    size_t n = 0;
    for (const std::string& l : host::read_file_lines(o.gene_snp)) {
      const std::vector<std::string> t = split(l, "\t");
      if (t.size() < 2) host::fail("filw with gene-snp pairs to keep should have two columns gene<tab>snp");
      pairs_keep[t[0]].insert(t[1]);
      ++n;
    }
    if (o.verbose > 0 && !pairs_keep.empty())
      std::cout << "nb of gene-snp pairs to keep: " << n << " (" << pairs_keep.size() << " genes)" << std::endl;
  }

  host::GzWriter out(o.out, "wb");
  std::string hdr;
  if (save_post) hdr += "#pi0=" + sci(o.pi0, 4) + "\n";
  hdr += "gene\tsnp";
  if (save_bf) {
    hdr += "\tgene.log10.bf\tsnp.log10.bf";
    if (o.all_dims)
      for (size_t k = 0; k < D; ++k) hdr += types ? "\tlog10.bf.type." + std::to_string(k + 1) : "\tlog10.bf." + w.cfg_names[k];
  }
  if (save_post) {
    if (pa) hdr += "\tgene.post";
    if (pb) hdr += "\tsnp.post.the";
    if (pc) hdr += "\tsnp.post.an";
    if (pd)
      for (size_t s = 0; s < S; ++s) hdr += "\tsnp.post.s" + std::to_string(s + 1);
    if (o.all_dims)
      for (size_t k = 0; k < D; ++k) hdr += types ? "\tpost.type." + std::to_string(k + 1) : "\tpost.config." + w.cfg_names[k];
  }
  if (o.best_dim) hdr += types ? "\tbest.type\tpost.best.type" : "\tbest.config\tpost.best.config";
  hdr += "\n";
  out.put(hdr);

  eqbf_avg_opts ao;
  std::memset(&ao, 0, sizeof ao);
  ao.model = types ? EQBF_AVG_TYPES : EQBF_AVG_CONFIGS;
  ao.dim = (int32_t)D;
  ao.n_subgroups = (int32_t)(types ? S : S_cfg);
  ao.n_grid = (int32_t)L;
  ao.n_keep = (int32_t)w.keep.size();
  ao.keep = w.keep.data();
  ao.grid_weights = w.grid.data();
  ao.dim_weights = types ? w.type.data() : w.cfg.data();
  ao.cfg_subgroups = types ? nullptr : cs.data();
  ao.sub_weights = types ? sw.data() : nullptr;
  ao.pi0 = o.pi0;
  ao.want = want;
  const size_t SP = std::max<size_t>(ao.n_subgroups, 1);
  for (const std::string& f : files) {
    if (o.verbose > 1) std::cout << "read file " << f << std::endl;
    const std::vector<std::string> lines = host::read_file_lines(f);
    if (lines.empty()) continue;
    Parsed d;
    parse_lines(f, lines, o, w, genes_keep, snps_keep, pairs_keep, R, d);
    const int64_t G = (int64_t)d.gene_names.size(), NP = (int64_t)d.snp_names.size();
    if (o.verbose > 1) std::cout << "nb of genes: " << G << " (" << NP << " gene-SNP pairs)" << std::endl;
    if (G == 0) continue;
    std::vector<double> gene_bf(G, kNaN), gene_post(G, kNaN), snp_bf(NP, kNaN), dim_bf(NP * D, kNaN),
        post_the(NP, kNaN), post_an(NP, kNaN), post_dim(NP * D, kNaN), post_sub(NP * SP, kNaN);
    std::vector<int32_t> best(NP, -1);
    eqbf_avg_out r{gene_bf.data(), gene_post.data(), snp_bf.data(), dim_bf.data(), post_the.data(),
                   post_an.data(), post_dim.data(), post_sub.data(), best.data()};
    eng.check(eqbf_avg_bfs(eng.ctx, d.raw.data(), NP, d.off.data(), G, &ao, &r));
    std::string buf;
    for (int64_t g = 0; g < G; ++g) {
      /* identify_best_snps (:800-829): SNPs by decreasing BF */
      std::vector<int64_t> ord((size_t)(d.off[g + 1] - d.off[g]));
      std::iota(ord.begin(), ord.end(), d.off[g]);
      std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        const double x = snp_bf[a], y = snp_bf[b];
        if (std::isnan(x) || std::isnan(y)) return !std::isnan(x) && std::isnan(y);
        return x > y;
      });
      size_t last = ord.size() - 1;
      if (o.best_snps == 1) {
        last = 0;
      } else if (o.best_snps == 2) {
        double cum = 0.0;
        for (size_t q = 0; q < ord.size(); ++q) {
          cum += post_the[ord[q]];
          if (cum >= 0.95) {
            last = q;
            break;
          }
        }
      }
      for (size_t q = 0; q <= last; ++q) {
        const int64_t p = ord[q];
        buf += d.gene_names[g] + "\t" + d.snp_names[p];
        if (save_bf) {
          buf += "\t" + sci(gene_bf[g], 4) + "\t" + sci(snp_bf[p], 4);
          if (o.all_dims)
            for (size_t k = 0; k < D; ++k) buf += "\t" + sci(dim_bf[p * D + k], 4);
        }
        if (save_post) {
          if (pa) buf += "\t" + sci(gene_post[g], 4);
          if (pb) buf += "\t" + sci(post_the[p], 4);
          if (pc) buf += "\t" + sci(post_an[p], 4);
          if (pd)
            for (size_t s = 0; s < S; ++s)
              buf += "\t" + sci(s < (size_t)ao.n_subgroups ? post_sub[p * SP + s] : kNaN, 4);
          if (o.all_dims)
            for (size_t k = 0; k < D; ++k) buf += "\t" + sci(post_dim[p * D + k], 4);
        }
        if (o.best_dim) {
          const int32_t b = best[p];
          if (types)
            buf += "\t" + std::to_string(b + 1);
          else
            buf += "\t" + (b >= 0 && (size_t)b < w.cfg_names.size() ? w.cfg_names[b] : std::string("nan"));
          buf += "\t" + sci(dims_post && b >= 0 ? post_dim[p * D + b] : kNaN, 4);
        }
        buf += "\n";
      }
      if (buf.size() > (1u << 20)) {
        out.put(buf);
        buf.clear();
      }
    }
    out.put(buf);
  }
}

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse(argc, argv);
  const time_t t0 = time(nullptr);
  if (o.verbose > 0)
    std::cout << "START " << argv[0] << "\ncmd-line: " << cmd_line(argc, argv) << std::endl;
  if (o.verbose > 0 && !std::isnan(o.pi0)) std::cout << "pi0: " << sci(o.pi0, 4) << std::endl;
  Weights w;
  load_grid_weights(o, w);
  load_config_weights(o, w);
  load_type_weights(o, w);
  const std::vector<std::string> files = glob_files(o.in);
  if (o.verbose > 0) std::cerr << "nb of files with raw BFs: " << files.size() << std::endl;
  if (!(o.cwts.empty() && o.tswts.empty())) {
    /* the reference reads no weight of the other model and averages over
     * an empty set of dimensions: refused here */
    if (o.model == "types" && w.type.empty()) host::fail("--model types needs --tswts");
    if (o.model == "configs" && w.cfg.empty()) host::fail("--model configs needs --cwts");
  }
  Engine eng;
  if (o.cwts.empty() && o.tswts.empty())
    average_over_grid(o, w, files, eng);
  else
    average_over_grid_and_dims(o, w, files, eng);
  if (o.verbose > 0)
    std::cout << "END " << argv[0] << "\nelapsed -> " << (long)(time(nullptr) - t0) << " sec" << std::endl;
  return EXIT_SUCCESS;
}
