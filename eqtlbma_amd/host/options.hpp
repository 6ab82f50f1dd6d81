/*
 * Command line of the eqtlbma_bf drop-in: the reference's flags, defaults
 * and validation rules (src/eqtlbma_bf.cpp:262-693, defaults :1586-1597),
 * plus --precision to choose the device path, --device and --ktimes.
 */
#ifndef EQBF_HOST_OPTIONS_HPP
#define EQBF_HOST_OPTIONS_HPP

#include <cstdint>
#include <string>
#include <vector>

namespace host {

struct Options {
  int verbose = 1, trick = 0, perm_sep = 0, nb_threads = 1, write_group_size = 10;
  uint64_t radius = 100000, nb_permutations = 0, seed = UINT64_MAX, trick_cutoff = 10;
  /* float, as the reference parses them (eqtlbma_bf.cpp:1590): --maf 0.05
   * compares against 0.05f, --fiterr 0.3 is 0.3f */
  float min_maf = 0.0f, prop_cov_errors = 0.5f;
  bool save_sstats = false, save_weighted_abfs = false, need_qnorm = false, use_max_bf = false;
  std::string geno, scoord, exp, gcoord, anchor = "TSS", inss, out, likelihood = "normal", analysis, covar,
      gridL, gridS, bfs = "gen", error_model = "uvlr", permbf = "none", snps_to_keep;
  std::vector<std::string> subgroups_to_keep;
  /* device path: "mixed" (batched, FP64 sums and FP32 Bayes-factor terms)
   * or "fp64" (every quantity in FP64) -- no counterpart in the reference */
  std::string precision = "mixed";
  int device = 0;
  /* print the GPU time of every kernel to stderr at the end (no counterpart
   * in the reference) */
  bool kernel_times = false;
  std::string cmdline;
};

/* Parses and validates; prints the help or the version and exits when asked,
 * and exits with the reference's messages on invalid input. */
Options parse_options(int argc, char** argv);

}  // namespace host

#endif
