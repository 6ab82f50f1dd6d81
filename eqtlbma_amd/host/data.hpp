/*
 * Input data of eqtlbma_bf held for the device engine.
 *
 * Follows the reference's loaders (src/eqtlbma/quantgen/data_loader.cpp):
 * list files of subgroup -> path (:37-79, :81-163), the sorted union of
 * sample names and each subgroup's column of every sample (:165-395,
 * samples.cpp:60-146), gene coordinates in BED (:396-435), expression
 * levels (:437-527), SNP coordinates in BED with the custom dose format
 * (:845-1010) or VCF / IMPUTE genotypes (:570-843), covariates
 * (:1097-1159), the --snp list (:529-568) and the MAF filter.
 */
#ifndef EQBF_HOST_DATA_HPP
#define EQBF_HOST_DATA_HPP

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace host {

struct Subgroup {
  std::string name, geno_file, exp_file, covar_file;
  int geno_set = -1;              /* index into Data::sets                 */
  std::vector<int32_t> exp_col;   /* per union sample: column or -1        */
  std::vector<int32_t> geno_col;
  std::vector<int32_t> covar_col;
  std::vector<std::string> covar_names;        /* sorted by name            */
  std::vector<std::vector<double>> covar_vals; /* [covariate][file column]  */
  std::vector<double> exp_vals;                /* [row][file column]        */
  int32_t exp_ncols = 0;
};

/* One genotype file (subgroups naming the same file share it). */
struct GenoSet {
  std::string file;
  int32_t ncols = 0;
  std::vector<double> vals; /* [row][file column]                         */
};

struct Snp {
  std::string name, chr;
  int64_t pos = 0;
  std::vector<int32_t> row; /* per genotype set: row in GenoSet::vals, -1  */
  std::vector<double> maf;  /* per genotype set                           */
};

struct Gene {
  std::string name, chr;
  int64_t start = 0, end = 0; /* start = BED start + 1 (gene.cpp:39-46)  */
  std::vector<int32_t> exp_row; /* per subgroup: row in exp_vals, or -1   */
};

struct LoadOptions {
  std::string geno_list, snp_coords, exp_list, gene_coords, covar_list, snps_to_keep;
  std::vector<std::string> subgroups_to_keep;
  std::string error_model = "uvlr";
  std::string anchor = "TSS"; /* cis anchor and radius: the regions of a   */
  uint64_t radius = 100000;   /* tabix-indexed --scoord are queried by them */
  float min_maf = 0.0f; /* compared as the reference does, maf < (float) cutoff */
  int verbose = 1;
};

struct Data {
  std::vector<Subgroup> sbgrps;  /* sorted by name                         */
  std::vector<std::string> samples; /* sorted union                        */
  std::vector<GenoSet> sets;
  std::map<std::string, Gene> genes;
  std::map<std::string, Snp> snps;
  std::map<std::string, std::vector<const Snp*>> chr_snps; /* by position */
};

/* Loads everything; exits with the reference's messages on bad input.
 * Returns false when there is nothing to analyse (no gene or no SNP). */
bool load_data(const LoadOptions& o, Data& d);

/* --inss: summary statistics per (gene, SNP, subgroup) given as input
 * (loadSummaryStats, data_loader.cpp:1212-1343). */
struct SstatsEntry {
  bool set = false; /* a line was read (GeneSnpPair::GetNbSubgroups counts it) */
  int64_t n = 0;
  double sigmahat = 0, betahat = 0, sebetahat = 0;
};
struct InssGene {
  std::vector<std::string> snps;           /* pairs, in order of first appearance */
  std::map<std::string, size_t> idx;
  std::vector<std::vector<SstatsEntry>> st; /* [pair][subgroup]                   */
};
struct InssData {
  std::vector<std::string> sbgrps;          /* sorted by name                     */
  std::map<std::string, InssGene> genes;
  size_t n_snps = 0;                        /* distinct SNP names                 */
};
/* Returns false when there is nothing to analyse (no gene or no SNP). */
bool load_inss(const std::string& list, int verbose, InssData& d);

}  // namespace host

#endif
