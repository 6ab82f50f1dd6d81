/*
 * eqtlbma_hm: fits the hierarchical model of eQtlBma by EM on the raw log10
 * ABFs that eqtlbma_bf writes (--outraw), on the GPU through
 * libeqbf_hip.so (eqbf_hm_em, eqbf_hm_bfs, eqbf_hm_ci).
 *
 * Mirrors the reference program (eqtlbma_hm.cpp): the same options and
 * checks (parseCmdLine, :1800-2065), the data loader's line rules
 * (Controller::load_data_one_file, :286-369: header "gene snp config", rows
 * of a gene and of a SNP are consecutive, configurations named "gen..."
 * skipped unless --keepgen, which still skips gen-fix and gen-maxh,
 * --configs keeps a subset, the types model keeps the configurations
 * "1".."S", the first --ngrid values of a row are read and missing ones are
 * NaN), the initialisation (init_params, :430-560 and :600-615; --seed or
 * --rand draw pi0 and the weights from GSL's mt19937, randomize_parameter_sp
 * :539-598), the run (run, :2067-2100) and the output file (save_result,
 * :1611-1740: iostream flags set once and kept, so every number after
 * "#pi0" prints as %.4e and booleans as true/false).
 *
 * Files are read whole as the reference's readFile does (utils_io.cpp:
 * 290-376): the last byte of the file is dropped (normally its final
 * newline) and empty lines vanish.
 */
#include <getopt.h>
#include <glob.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "../host/util.hpp"
#include "eqbf.h"

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();
const size_t kUnset = (size_t)-1;

struct Options {
  std::string data, model = "configs", out, init, ci_file;
  size_t nsubgrp = kUnset, dim = kUnset, ngrid = kUnset, seed = kUnset, maxit = kUnset;
  double thresh = 0.05, stepmax = 1.0, fixed_pi0 = kNaN;
  std::vector<std::string> configs;
  bool rand_init = false, keepgen = false, getci = false, getbf = false;
  int verbose = 1, threads = 1;
};

void help(const char* prog) {
  std::cout << "`" << prog << "' fits the hierarchical model of eQtlBma with an EM algorithm (on the GPU).\n\n"
            << "Usage: " << prog << " [OPTIONS] ...\n\n"
            << "Options:\n"
            << "  -h, --help\tdisplay the help and exit\n"
            << "  -V, --version\toutput version information and exit\n"
            << "  -v, --verbose\tverbosity level (0/default=1/2/3)\n"
            << "      --data\tinput data (usually output files from eqtlbma_bf)\n"
            << "      --nsubgrp\tnumber of subgroups\n"
            << "      --model\twhich model to fit (default=configs/types)\n"
            << "      --dim\tdimension of the model (nb of active configs or types)\n"
            << "      --ngrid\tnumber of grid points\n"
            << "      --out\toutput file (gzipped)\n"
            << "      --init\tfile for initialization\n"
            << "\t\t3 columns: param<tab>value<tab>fixed (TRUE or FALSE)\n"
            << "      --rand\trandom initialization\n"
            << "      --seed\tseed used with --rand, otherwise use time\n"
            << "      --thresh\tthreshold to stop the EM (default=0.05)\n"
            << "      --maxit\tmaximum number of iterations (optional)\n"
            << "      --msl\tmaximum step length for SQUAREM\n"
            << "\t\tdefault=1 (meaning classical EM), around 3 is a good option\n"
            << "      --thread\tnumber of threads (accepted, unused: the EM runs on the GPU)\n"
            << "      --configs\tsubset of configurations to keep (e.g. \"1|3|1-3\")\n"
            << "      --keepgen\tkeep 'general' ABFs (useful for BMAlite)\n"
            << "      --getci\tcompute the confidence intervals\n"
            << "      --getbf\tcompute the Bayes Factors using the estimated weights\n"
            << "      --pi0\tfixed value for pi0 (pi0 hence won't be updated in the EM)\n"
            << "      --ci\tfile with estimates of hyperparameters to only compute confidence intervals\n"
            << std::endl;
}

std::string cmd_line(int argc, char** argv) {
  std::string s;
  for (int i = 0; i < argc; ++i) s += (i ? " " : "") + std::string(argv[i]);
  return s;
}

[[noreturn]] void usage_error(int argc, char** argv, const std::string& msg) {
  std::cerr << "cmd-line: " << cmd_line(argc, argv) << std::endl
            << std::endl
            << "ERROR: " << msg << std::endl
            << std::endl;
  help(argv[0]);
  std::exit(EXIT_FAILURE);
}

/* split on any of the delimiters, empty tokens dropped (strtok) */
std::vector<std::string> split(const std::string& s, const char* delims) {
  std::vector<std::string> out;
  host::tokenize(s, delims, out);
  return out;
}

Options parse(int argc, char** argv) {
  Options o;
  static const struct option longopts[] = {
      {"help", no_argument, 0, 'h'},          {"version", no_argument, 0, 'V'},
      {"verbose", required_argument, 0, 'v'}, {"data", required_argument, 0, 0},
      {"nsubgrp", required_argument, 0, 0},   {"model", required_argument, 0, 0},
      {"dim", required_argument, 0, 0},       {"ngrid", required_argument, 0, 0},
      {"out", required_argument, 0, 0},       {"init", required_argument, 0, 0},
      {"rand", no_argument, 0, 0},            {"seed", required_argument, 0, 0},
      {"thresh", required_argument, 0, 0},    {"maxit", required_argument, 0, 0},
      {"msl", required_argument, 0, 0},       {"thread", required_argument, 0, 0},
      {"ci", required_argument, 0, 0},        {"configs", required_argument, 0, 0},
      {"keepgen", no_argument, 0, 0},         {"getci", no_argument, 0, 0},
      {"getbf", no_argument, 0, 0},           {"pi0", required_argument, 0, 0},
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
      std::cout << argv[0] << " (eqtlbma_amd)" << std::endl;
      std::exit(0);
    }
    if (c == 'v') {
      o.verbose = std::atoi(optarg);
      continue;
    }
    if (c != 0) {
      std::printf("\n");
      help(argv[0]);
      std::abort();
    }
    const std::string n = longopts[idx].name;
    if (n == "data") o.data = optarg;
    else if (n == "nsubgrp") o.nsubgrp = (size_t)std::atol(optarg);
    else if (n == "model") o.model = optarg;
    else if (n == "dim") o.dim = (size_t)std::atoi(optarg);
    else if (n == "ngrid") o.ngrid = (size_t)std::atoi(optarg);
    else if (n == "out") o.out = optarg;
    else if (n == "init") o.init = optarg;
    else if (n == "rand") o.rand_init = true;
    else if (n == "seed") o.seed = (size_t)std::atol(optarg);
    else if (n == "thresh") o.thresh = std::atof(optarg);
    else if (n == "maxit") o.maxit = (size_t)std::atol(optarg);
    else if (n == "msl") o.stepmax = std::atof(optarg);
    else if (n == "thread") o.threads = std::atoi(optarg);
    else if (n == "ci") o.ci_file = optarg;
    else if (n == "configs") o.configs = split(optarg, "|");
    else if (n == "keepgen") o.keepgen = true;
    else if (n == "getci") o.getci = true;
    else if (n == "getbf") o.getbf = true;
    else if (n == "pi0") o.fixed_pi0 = std::atof(optarg);
  }
  if (o.data.empty()) usage_error(argc, argv, "missing compulsory option --data");
  if (o.model.empty()) usage_error(argc, argv, "missing compulsory option --model");
  if (o.model != "configs" && o.model != "types") usage_error(argc, argv, "--model " + o.model + " is invalid");
  if (o.nsubgrp == kUnset) usage_error(argc, argv, "missing compulsory option --nsubgrp");
  if (o.dim == kUnset) usage_error(argc, argv, "missing compulsory option --dim");
  if (o.model == "configs" && o.dim >= (size_t)std::pow(2, (double)o.nsubgrp))
    usage_error(argc, argv,
                "--model configs --nsubgrp " + std::to_string(o.nsubgrp) + " --dim " + std::to_string(o.dim) +
                    " is invalid");
  if (o.ngrid == kUnset) usage_error(argc, argv, "missing compulsory option --ngrid");
  if (o.out.empty()) usage_error(argc, argv, "missing compulsory option --out");
  if (!o.init.empty() && !host::file_exists(o.init)) usage_error(argc, argv, "can't find " + o.init);
  if (o.rand_init && o.seed == kUnset) o.seed = (size_t)std::time(nullptr) ^ ((size_t)getpid() << 16);
  if (o.thresh <= 0.0) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", o.thresh);
    usage_error(argc, argv, std::string("--thresh ") + b + " is invalid");
  }
  if (!o.ci_file.empty() && !host::file_exists(o.ci_file)) usage_error(argc, argv, "can't find " + o.ci_file);
  if (!std::isnan(o.fixed_pi0) && (o.fixed_pi0 <= 0.0 || o.fixed_pi0 >= 1.0)) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", o.fixed_pi0);
    usage_error(argc, argv, std::string("--pi0 ") + b + " is invalid");
  }
  if (!o.ci_file.empty() && !o.getci) usage_error(argc, argv, "--getci should be given with --ci");
  if (o.model == "types") {
    o.configs.clear(); /* --configs is ignored: one row per subgroup */
    for (size_t s = 1; s <= o.nsubgrp; ++s) o.configs.push_back(std::to_string(s));
  }
  return o;
}

/* ---- data ------------------------------------------------------------------ */

struct Data {
  std::vector<std::string> genes, snps, config_names;
  std::vector<int64_t> off{0};
  std::vector<double> raw; /* [pairs][D][L] */
};

std::vector<std::string> read_lines(const std::string& path) { return host::read_file_lines(path); }

void load_file(const std::string& path, const Options& o, size_t D, Data& d) {
  const std::vector<std::string> lines = read_lines(path);
  if (lines.empty()) return;
  const std::vector<std::string> head = split(lines[0], " \t,");
  if (head.size() < 3 || head[0] != "gene" || head[1] != "snp" || head[2] != "config")
    host::fail("file " + path + " has wrong header line");
  const size_t L = o.ngrid;
  const bool want_names = o.model == "configs";
  std::string cur_gene, cur_snp;
  std::vector<double> rows; /* the current SNP's kept rows, L values each */
  bool have_gene = false;
  auto push_snp = [&]() {
    if (rows.empty()) return;
    const size_t nr = rows.size() / L;
    if (nr != D)
      host::fail("gene " + cur_gene + " snp " + cur_snp + " has " + std::to_string(nr) +
                 " kept rows, the model needs " + std::to_string(D));
    d.raw.insert(d.raw.end(), rows.begin(), rows.end());
    d.snps.push_back(cur_snp);
    rows.clear();
  };
  auto close_gene = [&]() {
    if (have_gene && (int64_t)d.snps.size() > d.off.back()) {
      d.genes.push_back(cur_gene);
      d.off.push_back((int64_t)d.snps.size());
    }
  };
  for (size_t i = 1; i < lines.size(); ++i) {
    const std::vector<std::string> t = split(lines[i], " \t,");
    if (t.size() < 3) host::fail("file " + path + " line " + std::to_string(i + 1) + " has fewer than 3 fields");
    const std::string& config = t[2];
    if ((!o.keepgen && config.find("gen") != std::string::npos) ||
        (!o.configs.empty() && std::find(o.configs.begin(), o.configs.end(), config) == o.configs.end()))
      continue;
    if (o.keepgen && (config == "gen-fix" || config == "gen-maxh")) continue;
    if (want_names && d.config_names.size() < o.dim) d.config_names.push_back(config);
    if (t[0] != cur_gene) { /* new gene */
      push_snp();
      close_gene();
      cur_gene = t[0];
      have_gene = true;
    }
    if (t[1] != cur_snp) { /* new SNP of the same gene */
      push_snp();
      cur_snp = t[1];
    }
    const size_t base = rows.size();
    rows.resize(base + L, kNaN);
    for (size_t j = 0; j < L && 3 + j < t.size(); ++j) rows[base + j] = std::atof(t[3 + j].c_str());
  }
  push_snp();
  close_gene();
}

Data load_data(const Options& o, size_t D) {
  glob_t g;
  std::memset(&g, 0, sizeof g);
  glob(o.data.c_str(), GLOB_TILDE, nullptr, &g);
  std::vector<std::string> files(g.gl_pathv, g.gl_pathv + g.gl_pathc);
  globfree(&g);
  if (files.empty()) host::fail("no input file was found from pattern " + o.data);
  if (o.verbose > 0) std::cout << "nb of input files: " << files.size() << std::endl;
  Data d;
  for (const std::string& f : files) load_file(f, o, D, d);
  if (d.genes.empty()) host::fail("no gene-SNP pair was kept from " + o.data);
  if (o.verbose > 0)
    std::fprintf(stderr, "finish loading %zu genes and %zu gene-snp pairs\n", d.genes.size(), d.snps.size());
  return d;
}

/* ---- parameters -------------------------------------------------------------- */

struct Params {
  double pi0 = kNaN;
  std::vector<double> grid, dimw, sub; /* sub: [dim][S] */
  int fixed[5] = {0, 0, 0, 0, 0};      /* pi0, grid-points, configs, types, subgroups-per-type */
};

bool is_true(const std::string& s) { return s == "TRUE" || s == "true"; }

/* init_params(init_file) (eqtlbma_hm.cpp:430-500) */
void init_from_file(const std::string& path, bool types, size_t S, Params& P) {
  size_t ig = 0, ic = 0, it = 0, s1 = 0, s2 = 0;
  for (const std::string& line : read_lines(path)) {
    if (line.find('#') == 0) continue;
    const std::vector<std::string> t = split(line, "\t, ");
    if (t.size() < 2) continue;
    if (t[0] == "param" && t[1] == "value") continue;
    const double v = std::atof(t[1].c_str());
    const bool fx = t.size() == 3 && is_true(t[2]);
    if (t[0].find("pi0") != std::string::npos) {
      P.pi0 = v;
      if (fx) P.fixed[0] = 1;
    } else if (t[0].find("grid") != std::string::npos) {
      if (ig >= P.grid.size()) host::fail("file " + path + " has more grid weights than --ngrid");
      P.grid[ig++] = v;
      if (fx) P.fixed[1] = 1;
    } else if (t[0].find("config") != std::string::npos) {
      if (types) continue; /* no configuration weights in the types model */
      if (ic >= P.dimw.size()) host::fail("file " + path + " has more configuration weights than --dim");
      P.dimw[ic++] = v;
      if (fx) P.fixed[2] = 1;
    } else if (t[0].find("type") != std::string::npos) {
      if (!types) continue;
      if (it >= P.dimw.size()) host::fail("file " + path + " has more type weights than --dim");
      P.dimw[it++] = v;
      if (fx) P.fixed[3] = 1;
    } else if (t[0].find("subgroup") != std::string::npos) {
      if (!types) continue;
      if (s2 == S) {
        s2 = 0;
        ++s1;
      }
      if (s1 * S + s2 >= P.sub.size()) host::fail("file " + path + " has too many subgroup weights");
      P.sub[s1 * S + s2] = v;
      ++s2;
      if (fx) P.fixed[4] = 1;
    }
  }
}

/* GSL's gsl_rng_uniform and gsl_ran_exponential (u in [0,1), -mu log1p(-u)) */
double rng_uniform(eqbf_rng* r) { return eqbf_rng_flat(r, 0.0, 1.0); }
double rng_exponential(eqbf_rng* r, double mu) { return -mu * std::log1p(-rng_uniform(r)); }

/* randomize_parameter_sp (eqtlbma_hm.cpp:539-598) */
void init_random(size_t seed, bool types, size_t S, Params& P) {
  eqbf_rng r;
  eqbf_rng_seed(&r, (uint64_t)seed);
  if (!P.fixed[0]) P.pi0 = rng_uniform(&r);
  double sum = 0.0;
  for (double& g : P.grid) sum += (g = rng_exponential(&r, 1.0));
  for (double& g : P.grid) g /= sum;
  if (!types) {
    if (P.dimw.size() == 1) {
      P.dimw[0] = 1.0;
    } else {
      double s = 0.0;
      for (double& w : P.dimw) s += (w = rng_exponential(&r, 1.0));
      for (double& w : P.dimw) w /= s;
    }
  } else {
    double s = 0.0;
    for (double& w : P.dimw) s += (w = rng_exponential(&r, 1.0));
    for (size_t k = 0; k < P.dimw.size(); ++k) {
      P.dimw[k] /= s;
      for (size_t j = 0; j < S; ++j) P.sub[k * S + j] = rng_uniform(&r);
    }
  }
}

/* init_params(seed) without a seed (eqtlbma_hm.cpp:600-615) */
void init_default(bool types, size_t S, Params& P) {
  if (!P.fixed[0]) P.pi0 = 0.5;
  for (double& g : P.grid) g = 1.0 / (double)P.grid.size();
  for (double& w : P.dimw) w = P.dimw.size() == 1 && !types ? 1.0 : 1.0 / (double)P.dimw.size();
  if (types)
    for (double& q : P.sub) q = 0.5;
  (void)S;
}

/* ---- output ------------------------------------------------------------------ */

void put4(std::string& out, double v) {
  if (std::isnan(v)) {
    out += "nan";
    return;
  }
  if (std::isinf(v)) {
    out += v > 0 ? "inf" : "-inf";
    return;
  }
  char b[64];
  std::snprintf(b, sizeof b, "%.4e", v);
  out += b;
}

void put_row(std::string& t, const std::string& name, double mle, double lo, double hi, bool fixed) {
  t += name;
  t += '\t';
  put4(t, mle);
  t += '\t';
  put4(t, lo);
  t += '\t';
  put4(t, hi);
  t += fixed ? "\ttrue\n" : "\tfalse\n";
}

struct Bfs {
  std::vector<double> gene_post, gene_bf, snp_bf, dim_bf;
};

/* save_result (eqtlbma_hm.cpp:1611-1740) */
void save(const Options& o, const Data& d, const Params& P, const std::vector<double>& left,
          const std::vector<double>& right, const Bfs* bf) {
  const bool types = o.model == "types";
  const size_t K = o.dim, S = o.nsubgrp, L = o.ngrid;
  host::GzWriter f(o.out, "wb");
  std::string t = "#param\tmle\tleft.ci\tright.ci\tfixed\n";
  put_row(t, "#pi0", P.pi0, left[0], right[0], P.fixed[0]);
  if (!types) {
    for (size_t k = 0; k < K; ++k)
      put_row(t, "#config." + (k < d.config_names.size() ? d.config_names[k] : std::string()), P.dimw[k],
              left[1 + k], right[1 + k], P.fixed[2]);
  } else {
    for (size_t k = 0; k < K; ++k)
      put_row(t, "#type." + std::to_string(k + 1), P.dimw[k], left[1 + k], right[1 + k], P.fixed[3]);
    for (size_t k = 0; k < K; ++k)
      for (size_t s = 0; s < S; ++s)
        put_row(t, "#subgroup." + std::to_string(k + 1) + "-" + std::to_string(s + 1), P.sub[k * S + s],
                left[1 + K + k * S + s], right[1 + K + k * S + s], P.fixed[4]);
  }
  const size_t g0 = 1 + K + (types ? K * S : 0);
  for (size_t l = 0; l < L; ++l)
    put_row(t, "#grid." + std::to_string(l + 1), P.grid[l], left[g0 + l], right[g0 + l], P.fixed[1]);
  f.put(t);
  if (!bf) return;
  t = "gene\tgene.posterior.prob\tgene.log10.bf\tsnp\tsnp.log10.bf";
  for (size_t k = 0; k < K; ++k)
    t += "\tlog10.bf." + (types ? std::to_string(k + 1) : (k < d.config_names.size() ? d.config_names[k] : ""));
  t += '\n';
  for (size_t g = 0; g < d.genes.size(); ++g) {
    for (int64_t p = d.off[g]; p < d.off[g + 1]; ++p) {
      t += d.genes[g];
      t += '\t';
      put4(t, bf->gene_post[g]);
      t += '\t';
      put4(t, bf->gene_bf[g]);
      t += '\t';
      t += d.snps[p];
      t += '\t';
      put4(t, bf->snp_bf[p]);
      for (size_t k = 0; k < K; ++k) {
        t += '\t';
        put4(t, bf->dim_bf[(size_t)p * K + k]);
      }
      t += '\n';
    }
    if (t.size() > (1u << 22)) {
      f.put(t);
      t.clear();
    }
  }
  f.put(t);
}

[[noreturn]] void lib_fail(eqbf_ctx* ctx, int rc) {
  const char* msg = ctx ? eqbf_last_error(ctx) : nullptr;
  std::cerr << "ERROR: " << (msg && *msg ? msg : "GPU library call failed") << " (code " << rc << ")" << std::endl;
  std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse(argc, argv);
  const bool types = o.model == "types";
  const size_t K = o.dim, S = o.nsubgrp, L = o.ngrid;
  const size_t D = types ? S : K;
  if (o.verbose > 0) std::cout << "START " << argv[0] << "\ncmd-line: " << cmd_line(argc, argv) << std::endl;
  const time_t t0 = std::time(nullptr);

  Data d = load_data(o, D);

  Params P;
  P.grid.assign(L, kNaN);
  P.dimw.assign(K, kNaN);
  if (types) P.sub.assign(K * S, kNaN);
  if (!std::isnan(o.fixed_pi0)) {
    P.pi0 = o.fixed_pi0;
    P.fixed[0] = 1;
  }
  const std::string init_file = !o.ci_file.empty() ? o.ci_file : o.init;
  if (!init_file.empty()) {
    if (o.verbose > 0) std::cout << "loading initialization file " << init_file << " ..." << std::endl;
    init_from_file(init_file, types, S, P);
  } else if (o.seed != kUnset) {
    init_random(o.seed, types, S, P);
  } else {
    init_default(types, S, P);
  }

  eqbf_ctx* ctx = nullptr;
  int rc = eqbf_create(&ctx, 0);
  if (rc) lib_fail(ctx, rc);
  eqbf_hm_opts ho;
  std::memset(&ho, 0, sizeof ho);
  ho.model = types ? EQBF_HM_TYPES : EQBF_HM_CONFIGS;
  ho.dim = (int32_t)K;
  ho.n_subgroups = (int32_t)S;
  ho.n_grid = (int32_t)L;
  ho.thresh = o.thresh;
  ho.max_iters = o.maxit == kUnset ? 0 : (int64_t)o.maxit;
  ho.stepmax = o.stepmax;
  for (int i = 0; i < 5; ++i) ho.fixed[i] = P.fixed[i];
  eqbf_hm_params hp;
  hp.pi0 = P.pi0;
  hp.grid = P.grid.data();
  hp.config = types ? nullptr : P.dimw.data();
  hp.type = types ? P.dimw.data() : nullptr;
  hp.subgroup = types ? P.sub.data() : nullptr;
  const int64_t NP = (int64_t)d.snps.size(), G = (int64_t)d.genes.size();

  const size_t n_ci = 1 + K + (types ? K * S : 0) + L;
  std::vector<double> left(n_ci, kNaN), right(n_ci, kNaN);
  Bfs bf;
  bool have_bf = false;
  if (!o.ci_file.empty()) {
    /* run() with --ci: posteriors and intervals at parameters the
     * Controller never copies into the weights the intervals read, so
     * its first likelihood is NaN (eqtlbma_hm.cpp:2083-2088, 1548-1552) */
    std::cerr << "ERROR: log10(obslik) is NaN" << std::endl;
    eqbf_destroy(ctx);
    return EXIT_FAILURE;
  }
  if (o.verbose > 0)
    std::cout << "run EM algorithm (" << (o.stepmax > 1.0 ? "square" : "classic") << ") ..." << std::endl;
  double lik = kNaN;
  int64_t iters = 0;
  rc = eqbf_hm_em(ctx, d.raw.data(), NP, d.off.data(), G, &ho, &hp, &lik, &iters);
  if (rc) lib_fail(ctx, rc);
  P.pi0 = hp.pi0;
  if (o.verbose > 0) std::printf("EM: %lld iterations, log10(obslik) %f\n", (long long)iters, lik);
  if (o.getbf) {
    bf.gene_post.resize(G);
    bf.gene_bf.resize(G);
    bf.snp_bf.resize(NP);
    bf.dim_bf.resize((size_t)NP * K);
    rc = eqbf_hm_bfs(ctx, d.raw.data(), NP, d.off.data(), G, &ho, &hp, bf.gene_post.data(), bf.gene_bf.data(),
                     bf.snp_bf.data(), bf.dim_bf.data());
    if (rc) lib_fail(ctx, rc);
    have_bf = true;
  }
  if (o.getci) {
    rc = eqbf_hm_ci(ctx, d.raw.data(), NP, d.off.data(), G, &ho, &hp, left.data(), right.data());
    if (rc) lib_fail(ctx, rc);
  }
  if (o.verbose > 0) std::cout << "save the results in " << o.out << " ..." << std::endl;
  save(o, d, P, left, right, have_bf ? &bf : nullptr);
  eqbf_destroy(ctx);
  if (o.verbose > 0)
    std::cout << "END " << argv[0] << "\nelapsed -> " << (long long)(std::time(nullptr) - t0) << " sec" << std::endl;
  return EXIT_SUCCESS;
}
