#include "options.hpp"

#include <getopt.h>
#include <sys/time.h>

#include <cstdlib>
#include <cstring>
#include <iostream>

#include "util.hpp"

namespace host {
namespace {

void help(const char* prog) {
  std::cout
      << "`" << prog << "' computes eQTL summary statistics and Bayes factors in several subgroups on one\n"
      << "MI355X GPU, with the command line and output files of eqtlbma_bf.\n\n"
      << "Usage: " << prog << " [OPTIONS] ...\n\n"
      << "Options:\n"
      << "  -h, --help\tshow this help and exit\n"
      << "  -V, --version\tshow the version and exit\n"
      << "  -v, --verbose\tverbosity (0/default=1/2/3)\n"
      << "      --geno\tlist file 'subgroup<space/tab>path' of genotype files (custom dose, VCF or IMPUTE)\n"
      << "      --scoord\tSNP coordinates in BED format (required with custom dose genotypes)\n"
      << "      --exp\tlist file 'subgroup<space/tab>path' of expression level files (row 1 = samples)\n"
      << "      --gcoord\tgene coordinates in BED format\n"
      << "      --anchor\tgene boundary of the cis region: TSS (default) or TSS+TES\n"
      << "      --cis\thalf-length of the cis region in bp (default=100000)\n"
      << "      --out\tprefix of the gzipped output files\n"
      << "      --outss\twrite the summary statistics (with --analys join)\n"
      << "      --outw\twrite the Bayes factors averaged over the grids\n"
      << "      --lik\tlikelihood: normal (default)\n"
      << "      --analys\tsep (each subgroup separately) or join (all subgroups jointly)\n"
      << "      --qnorm\tquantile-normalise the expression levels to N(0,1)\n"
      << "      --maf\tminimum minor allele frequency (default=0.0)\n"
      << "      --covar\tlist file 'subgroup<space/tab>path' of covariate files\n"
      << "      --gridL\tlarge grid of (phi^2, omega^2), one pair per line (required with --analys join)\n"
      << "      --gridS\tsmall grid (required with --analys join and --bfs sin or all)\n"
      << "      --bfs\tBayes factors to compute: gen (default), sin or all\n"
      << "      --error\terror model for --analys join: uvlr (default) or mvlr\n"
      << "      --fiterr\tweight of the fitted error covariance with mvlr, in [0,1] (default=0.5)\n"
      << "      --nperm\tnumber of permutations (default=0)\n"
      << "      --seed\tseed of the random number generators (default: from the clock)\n"
      << "      --trick\tpermutation trick: 0 (default), 1 or 2\n"
      << "      --tricut\tcutoff of the trick (default=10)\n"
      << "      --permsep\tpermutations of --analys sep: 0 (default), 1 (min P over SNPs and subgroups)\n"
      << "\t\tor 2 (min P over SNPs, each subgroup separately)\n"
      << "      --pbf\tBayes factor of the --analys join permutations: none (default), gen, gen-sin, all\n"
      << "      --maxbf\tuse the maximum Bayes factor over SNPs as the permutation statistic\n"
      << "\t\t(default: their average)\n"
      << "      --thread\taccepted for compatibility (the GPU does the work)\n"
      << "      --snp\tfile with the names of the SNPs to analyse\n"
      << "      --sbgrp\tsubgroup(s) to analyse, e.g. 's1+s3'\n"
      << "      --wrtsize\tnumber of genes processed and written at once (default=10)\n"
      << "      --precision\tdevice path: mixed (default) or fp64\n"
      << "      --device\tHIP device index (default=0)\n"
      << "      --ktimes\tprint the GPU time of every kernel to stderr at the end\n"
      << std::endl;
}

[[noreturn]] void bad(const Options& o, const char* prog, const std::string& msg) {
  std::cerr << "cmd-line: " << o.cmdline << std::endl << std::endl << "ERROR: " << msg << std::endl << std::endl;
  help(prog);
  std::exit(EXIT_FAILURE);
}

uint64_t clock_seed() {
  timeval t;
  gettimeofday(&t, nullptr);
  return (uint64_t)t.tv_usec * (uint64_t)t.tv_sec;
}

}  // namespace

Options parse_options(int argc, char** argv) {
  Options o;
  for (int i = 0; i < argc; ++i) o.cmdline += (i ? " " : "") + std::string(argv[i]);
  static const struct option longopts[] = {
      {"help", no_argument, 0, 'h'},         {"version", no_argument, 0, 'V'},
      {"verbose", required_argument, 0, 'v'}, {"geno", required_argument, 0, 0},
      {"scoord", required_argument, 0, 0},    {"exp", required_argument, 0, 0},
      {"gcoord", required_argument, 0, 0},    {"anchor", required_argument, 0, 0},
      {"cis", required_argument, 0, 0},       {"inss", required_argument, 0, 0},
      {"out", required_argument, 0, 0},       {"outss", no_argument, 0, 0},
      {"outm", no_argument, 0, 0},            {"outw", no_argument, 0, 0},
      {"lik", required_argument, 0, 0},       {"analys", required_argument, 0, 0},
      {"qnorm", no_argument, 0, 0},           {"maf", required_argument, 0, 0},
      {"covar", required_argument, 0, 0},     {"gridL", required_argument, 0, 0},
      {"gridS", required_argument, 0, 0},     {"bfs", required_argument, 0, 0},
      {"error", required_argument, 0, 0},     {"fiterr", required_argument, 0, 0},
      {"nperm", required_argument, 0, 0},     {"seed", required_argument, 0, 0},
      {"trick", required_argument, 0, 0},     {"tricut", required_argument, 0, 0},
      {"permsep", required_argument, 0, 0},   {"pbf", required_argument, 0, 0},
      {"maxbf", no_argument, 0, 0},           {"thread", required_argument, 0, 0},
      {"snp", required_argument, 0, 0},       {"sbgrp", required_argument, 0, 0},
      {"wrtsize", required_argument, 0, 0},   {"precision", required_argument, 0, 0},
      {"device", required_argument, 0, 0},    {"ktimes", no_argument, 0, 0},
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
      std::cout << argv[0] << " 1.3.3 (MI355X engine)" << std::endl;
      std::exit(0);
    }
    if (c == 'v') {
      o.verbose = std::atoi(optarg);
      continue;
    }
    if (c != 0) {
      std::cerr << "ERROR: unknown option" << std::endl;
      help(argv[0]);
      std::exit(EXIT_FAILURE);
    }
    const std::string n = longopts[idx].name;
    const char* a = optarg;
    if (n == "geno") o.geno = a;
    else if (n == "scoord") o.scoord = a;
    else if (n == "exp") o.exp = a;
    else if (n == "gcoord") o.gcoord = a;
    else if (n == "anchor") o.anchor = a;
    else if (n == "cis") o.radius = (uint64_t)std::atol(a);
    else if (n == "inss") o.inss = a;
    else if (n == "out") o.out = a;
    else if (n == "outss") o.save_sstats = true;
    else if (n == "outw") o.save_weighted_abfs = true;
    else if (n == "lik") o.likelihood = a;
    else if (n == "analys") o.analysis = a;
    else if (n == "qnorm") o.need_qnorm = true;
    else if (n == "maf") o.min_maf = (float)std::atof(a);
    else if (n == "covar") o.covar = a;
    else if (n == "gridL") o.gridL = a;
    else if (n == "gridS") o.gridS = a;
    else if (n == "bfs") o.bfs = a;
    else if (n == "error") o.error_model = a;
    else if (n == "fiterr") o.prop_cov_errors = (float)std::atof(a);
    else if (n == "nperm") o.nb_permutations = (uint64_t)std::atol(a);
    else if (n == "seed") o.seed = (uint64_t)std::atol(a);
    else if (n == "trick") o.trick = std::atoi(a);
    else if (n == "tricut") o.trick_cutoff = (uint64_t)std::atol(a);
    else if (n == "permsep") o.perm_sep = std::atoi(a);
    else if (n == "pbf") o.permbf = a;
    else if (n == "maxbf") o.use_max_bf = true;
    else if (n == "thread") o.nb_threads = std::atoi(a);
    else if (n == "snp") o.snps_to_keep = a;
    else if (n == "sbgrp") tokenize(a, "+", o.subgroups_to_keep);
    else if (n == "wrtsize") o.write_group_size = std::atoi(a);
    else if (n == "precision") o.precision = a;
    else if (n == "device") o.device = std::atoi(a);
    else if (n == "ktimes") o.kernel_times = true;
  }
  const char* prog = argv[0];
  if (o.inss.empty()) {
    if (o.geno.empty()) bad(o, prog, "missing compulsory option --geno");
    if (!file_exists(o.geno)) bad(o, prog, "can't find " + o.geno);
    if (!o.scoord.empty() && !file_exists(o.scoord)) bad(o, prog, "can't find " + o.scoord);
    if (o.exp.empty()) bad(o, prog, "missing compulsory option --exp");
    if (!file_exists(o.exp)) bad(o, prog, "can't find " + o.exp);
    if (o.gcoord.empty()) bad(o, prog, "missing compulsory option --gcoord");
    if (!file_exists(o.gcoord)) bad(o, prog, "can't find " + o.gcoord);
    if (o.anchor.empty()) bad(o, prog, "SNPs in trans not yet implemented, see --anchor and --cis");
    if (o.anchor != "TSS" && o.anchor != "TSS+TES") bad(o, prog, "--anchor should be TSS or TSS+TES");
  } else {
    if (!file_exists(o.inss)) bad(o, prog, "can't find " + o.inss);
    if (o.analysis != "join") bad(o, prog, "--inss requires --analys join");
    if (o.error_model != "uvlr") bad(o, prog, "--inss requires --error uvlr");
    /* the reference permutes no gene loaded from summary statistics and then
     * prints the genes' uninitialised permutation results */
    if (o.nb_permutations > 0) bad(o, prog, "--nperm with --inss is not supported by this build");
  }
  if (o.out.empty()) bad(o, prog, "missing compulsory option --out");
  if (o.write_group_size < 1) bad(o, prog, "--wrtsize should be greater than 1");
  {
    const std::string test = o.out + "_test.txt.gz";
    { GzWriter w(test, "wb"); }
    std::remove(test.c_str());
  }
  if (o.likelihood.empty()) bad(o, prog, "missing compulsory option --lik");
  if (o.likelihood != "normal" && o.likelihood != "poisson" && o.likelihood != "quasipoisson")
    bad(o, prog, "--lik " + o.likelihood + " is not valid");
  if (o.likelihood != "normal") bad(o, prog, "--lik " + o.likelihood + " is not supported by this build");
  if (o.analysis.empty()) bad(o, prog, "missing compulsory option --analys");
  if (o.analysis != "sep" && o.analysis != "join") bad(o, prog, "--analys " + o.analysis + " is not valid");
  /* the reference's check reads "analysis != join" (eqtlbma_bf.cpp:591); the
   * grid is needed by the joint analysis, so that is what is required here */
  if (o.analysis == "join" && o.gridL.empty()) bad(o, prog, "missing compulsory option --gridL with --analys join");
  if (!o.gridL.empty() && !file_exists(o.gridL)) bad(o, prog, "can't find " + o.gridL);
  if (o.analysis == "join" && o.gridS.empty() && (o.bfs == "sin" || o.bfs == "all"))
    bad(o, prog, "--gridS is required with --analys join and --bfs " + o.bfs);
  if (!o.gridS.empty() && !file_exists(o.gridS)) bad(o, prog, "can't find " + o.gridS);
  if (o.bfs != "gen" && o.bfs != "sin" && o.bfs != "all") bad(o, prog, "--bfs " + o.bfs + " is not valid");
  if (o.error_model != "uvlr" && o.error_model != "mvlr" && o.error_model != "hybrid")
    bad(o, prog, "--error " + o.error_model + " is not valid");
  if (o.trick != 0 && o.trick != 1 && o.trick != 2) bad(o, prog, "--trick " + std::to_string(o.trick) + " is not valid");
  if (o.trick != 0 && o.trick_cutoff > o.nb_permutations)
    bad(o, prog, "--tricut " + std::to_string(o.trick_cutoff) + " is larger than --nperm " +
                     std::to_string(o.nb_permutations));
  if (o.perm_sep != 0 && o.perm_sep != 1 && o.perm_sep != 2)
    bad(o, prog, "--permsep " + std::to_string(o.perm_sep) + " is not valid");
  if (o.analysis == "sep" && o.nb_permutations > 0 && o.perm_sep != 1 && o.perm_sep != 2)
    bad(o, prog, "if --type sep --nperm > 0, --permsep should be '1' or '2'");
  if (o.permbf != "none" && o.permbf != "gen" && o.permbf != "gen-sin" && o.permbf != "all")
    bad(o, prog, "--pbf " + o.permbf + " is unvalid");
  if (o.analysis == "join" && o.nb_permutations > 0 && o.permbf == "none")
    bad(o, prog, "if --analys join --nperm > 0, --pbf should be different than 'none'");
  if (o.analysis == "join" && o.nb_permutations > 0 && o.permbf != "none" && o.bfs == "gen" && o.permbf != "gen")
    bad(o, prog, "if --analys join --bfs gen --nperm > 0, --pbf should be 'gen'");
  if (o.analysis == "join" && o.nb_permutations > 0 && o.permbf != "none" && o.bfs == "sin" && o.permbf == "all")
    bad(o, prog, "if --analys join --bfs sin --nperm > 0, --pbf should be 'gen' or 'gen-sin'");
  if (o.nb_threads <= 0) bad(o, prog, "--thread " + std::to_string(o.nb_threads) + " is invalid");
  if (!o.snps_to_keep.empty() && !file_exists(o.snps_to_keep)) bad(o, prog, "can't find " + o.snps_to_keep);
  if (o.precision != "mixed" && o.precision != "fp64") bad(o, prog, "--precision should be mixed or fp64");
  if (o.seed == UINT64_MAX) o.seed = clock_seed();
  return o;
}

}  // namespace host
