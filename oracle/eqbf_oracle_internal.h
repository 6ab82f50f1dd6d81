/* Internal declarations shared by the oracle's translation units
 * (test infrastructure only). */
#ifndef EQBF_ORACLE_INTERNAL_H
#define EQBF_ORACLE_INTERNAL_H
#include <cstdint>
#include <vector>

#include "../include/eqbf.h"

namespace orc {
typedef long double ldouble;
extern char g_err[512];
ldouble tdist_lower_neg(ldouble at, ldouble nu);
ldouble gauss_pinv_ld(ldouble p);
double expl_at(const eqbf_chunk* c, int64_t g, int32_t s, int64_t i);
double geno_at(const eqbf_chunk* c, int32_t s, int64_t i, int64_t j);
void config_masks(int32_t bfs, int32_t S, std::vector<uint64_t>& masks);
int64_t n_configs(int32_t bfs, int32_t S);
void members(const eqbf_chunk* c, int64_t g, int32_t s, const int32_t* perm, std::vector<int64_t>& idx,
             std::vector<double>& y);
int64_t sstats_one(const eqbf_chunk* c, int64_t g, int32_t s, int64_t snp, const int32_t* perm, double* out5);
void pair_abfs(int32_t S, int32_t bfs, const std::vector<int>& has, const std::vector<double>& std3,
               const double* phiL, const double* omaL, int32_t GL, const double* phiS, const double* omaS,
               int32_t GS, double* raw, double* w);
}  // namespace orc

extern "C" {
struct orc_mt;
void orc_standardize(double N, double Q, double sigmahat, double betahat, double sebetahat, double* out3);
double orc_log10_weighted_sum(const double* vec, int64_t size);
double orc_log10_weighted_sum_w(const double* vec, const double* w, int64_t size);
double orc_flat(orc_mt* st, double a, double b);
double orc_tdist_P(double t, double nu);
double orc_gauss_Pinv(double p);
double orc_fdist_Q_1(double F, double nu);
double orc_chisq_Qinv_1(double q);
void orc_bma(int32_t S, int32_t bfs, const double* w_cfg, double gen, double* gen_sin, double* all);
}
#endif
