/*
 * eqbf_oracle_mvlr.cpp -- CPU oracle for the multivariate (MVLR) Bayes
 * factors of eqtlbma_bf (--error mvlr).
 *
 * TEST INFRASTRUCTURE ONLY (see eqbf_oracle.cpp).
 *
 * Restates class MVLR (src/eqtlbma/quantgen/MVLR.cpp:26-764) for one SNP
 * (p = 1) with the same dense steps: the n x n projection T
 * (compute_common :102-144), the null covariance (compute_Sigma_null
 * :176-195), the mixed maximum-likelihood covariance with its
 * F -> chi-square correction (compute_Sigma_mle :201-298, compute_residual
 * :323-437), the meta prior Gamma (:473-494), W (:497-512), the summary
 * statistics (:517-554) and the per-grid-point log10 ABF with an LU
 * factorisation, log-determinant and inverse (:559-605).  Reference quirks
 * kept: m = q+s+1 (:66), Sigma := Sigma*C rather than C*Sigma*C (:268-273),
 * Hotelling T2 divided by sigma1 squared (:397).
 *
 * Pinned by numpy FP64 linear algebra on the same formulas and by the
 * identity det(I + k S^-1 W) = det(S + k W) / det(S)
 * (tests/test_oracle_mvlr.py).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <vector>

#include "eqbf_oracle_internal.h"

using namespace orc;
typedef std::vector<ldouble> Mat; /* row-major */
static const double kNaN = std::numeric_limits<double>::quiet_NaN();

static Mat matmul(const Mat& A, const Mat& B, int r, int k, int c, bool tA = false, bool tB = false) {
  /* (r x k) * (k x c); tA: A stored k x r; tB: B stored c x k */
  Mat C(r * c, 0);
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) {
      ldouble s = 0;
      for (int l = 0; l < k; ++l) {
        ldouble a = tA ? A[l * r + i] : A[i * k + l];
        ldouble b = tB ? B[j * k + l] : B[l * c + j];
        s += a * b;
      }
      C[i * c + j] = s;
    }
  return C;
}

/* LU with partial pivoting (largest |a| in the column, as gsl_linalg_LU_decomp). */
static bool lu_decomp(Mat& A, int n, std::vector<int>& piv, int& sign) {
  piv.resize(n);
  sign = 1;
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int j = 0; j < n - 1; ++j) {
    int p = j;
    ldouble mx = fabsl(A[j * n + j]);
    for (int i = j + 1; i < n; ++i)
      if (fabsl(A[i * n + j]) > mx) {
        mx = fabsl(A[i * n + j]);
        p = i;
      }
    if (p != j) {
      for (int k = 0; k < n; ++k) std::swap(A[j * n + k], A[p * n + k]);
      std::swap(piv[j], piv[p]);
      sign = -sign;
    }
    ldouble d = A[j * n + j];
    if (d == 0) continue;
    for (int i = j + 1; i < n; ++i) {
      ldouble f = A[i * n + j] / d;
      A[i * n + j] = f;
      for (int k = j + 1; k < n; ++k) A[i * n + k] -= f * A[j * n + k];
    }
  }
  return true;
}

static Mat lu_invert(const Mat& LU, int n, const std::vector<int>& piv) {
  Mat inv(n * n, 0);
  for (int col = 0; col < n; ++col) {
    std::vector<ldouble> b(n, 0);
    for (int i = 0; i < n; ++i) b[i] = (piv[i] == col) ? 1 : 0;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < i; ++k) b[i] -= LU[i * n + k] * b[k];
    for (int i = n - 1; i >= 0; --i) {
      for (int k = i + 1; k < n; ++k) b[i] -= LU[i * n + k] * b[k];
      b[i] /= LU[i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[i * n + col] = b[i];
  }
  return inv;
}

static Mat invert(const Mat& A, int n) {
  Mat LU = A;
  std::vector<int> piv;
  int sg;
  lu_decomp(LU, n, piv, sg);
  return lu_invert(LU, n, piv);
}

/* Symmetric eigen-decomposition by cyclic Jacobi; for the symmetric positive
 * semi-definite X'X this equals its SVD (MVLR.cpp:332). */
static void jacobi_eig(Mat A, int n, std::vector<ldouble>& w, Mat& V) {
  V.assign(n * n, 0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    ldouble off = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-36L) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        ldouble apq = A[p * n + q];
        if (fabsl(apq) < 1e-300L) continue;
        ldouble theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        ldouble t = (theta >= 0 ? 1 : -1) / (fabsl(theta) + sqrtl(theta * theta + 1));
        ldouble c = 1 / sqrtl(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          ldouble akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          ldouble apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          ldouble vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

namespace {
struct Mvlr {
  int n, s, q, m;
  Mat Y, x, Xc, T, Sigma0, Sigma0_inv, H;
  double fiterr;

  /* compute_residual (MVLR.cpp:323-437) for columns X (n x size); the first
   * q columns are Xc.  Returns the residual and sets factor. */
  std::vector<ldouble> residual(const Mat& X, int size, const std::vector<ldouble>& y, ldouble& factor) {
    Mat XtX = matmul(X, X, size, n, size, true, false);
    std::vector<ldouble> ev;
    Mat V;
    jacobi_eig(XtX, size, ev, V);
    Mat inv(size * size, 0);
    for (int i = 0; i < size; ++i)
      for (int j = 0; j < size; ++j) {
        ldouble acc = 0;
        for (int k = 0; k < size; ++k)
          if (ev[k] > 1e-8L) acc += V[i * size + k] * V[j * size + k] / ev[k];
        inv[i * size + j] = acc;
      }
    std::vector<ldouble> Xty(size, 0), hB(size, 0), res(n);
    for (int j = 0; j < size; ++j)
      for (int i = 0; i < n; ++i) Xty[j] += X[i * size + j] * y[i];
    for (int i = 0; i < size; ++i)
      for (int j = 0; j < size; ++j) hB[i] += inv[i * size + j] * Xty[j];
    for (int i = 0; i < n; ++i) {
      ldouble f = 0;
      for (int j = 0; j < size; ++j) f += X[i * size + j] * hB[j];
      res[i] = y[i] - f;
    }
    factor = 1;
    if (size > q) { /* one SNP: size - q == 1 */
      ldouble rr = 0;
      for (int i = 0; i < n; ++i) rr += res[i] * res[i];
      ldouble sigma1 = rr / (n - size);
      /* t7 = t5' T t5 with t5 the genotype column */
      std::vector<ldouble> t5(n);
      for (int i = 0; i < n; ++i) t5[i] = X[i * size + q];
      ldouble t7 = 0;
      for (int i = 0; i < n; ++i) {
        ldouble Tt = 0;
        for (int j = 0; j < n; ++j) Tt += T[i * n + j] * t5[j];
        t7 += t5[i] * Tt;
      }
      ldouble t8 = hB[q];
      ldouble t10 = t8 * t7 * t8;
      ldouble T2 = t10 / (sigma1 * sigma1);
      double v1 = size - q, v2 = n - size;
      double F = (double)((v2 - v1 + 1) * T2 / (v1 * v2));
      double qv = orc_fdist_Q_1(F, v2 - v1 + 1);
      double newF = orc_chisq_Qinv_1(qv) / v1;
      factor = (F < 1e-8) ? 1 : F / newF;
    }
    return res;
  }

  /* compute_Sigma (MVLR.cpp:148-173) for configuration gamma (p = 1). */
  void sigma(const std::vector<int>& gamma, Mat& Sig, Mat& Sig_inv) {
    if (fiterr < 1e-6) {
      Sig = Sigma0;
      Sig_inv = Sigma0_inv;
      return;
    }
    Mat E(n * s, 0);
    std::vector<ldouble> fac(s);
    for (int i = 0; i < s; ++i) {
      int size = q + (gamma[i] ? 1 : 0);
      Mat Xv(n * size);
      for (int r = 0; r < n; ++r) {
        for (int j = 0; j < q; ++j) Xv[r * size + j] = Xc[r * q + j];
        if (gamma[i]) Xv[r * size + q] = x[r];
      }
      std::vector<ldouble> yv(n);
      for (int r = 0; r < n; ++r) yv[r] = Y[r * s + i];
      ldouble f;
      std::vector<ldouble> res = residual(Xv, size, yv, f);
      fac[i] = sqrtl(f);
      for (int r = 0; r < n; ++r) E[r * s + i] = res[r];
    }
    Mat Sm = matmul(E, E, s, n, s, true, false);
    Sig.assign(s * s, 0);
    for (int i = 0; i < s * s; ++i) Sig[i] = H[i] * (ldouble)m / (ldouble)(m + n) + Sm[i] / (ldouble)(m + n);
    /* t2 = Sigma * C (C diagonal): scale column j by fac[j] */
    for (int i = 0; i < s; ++i)
      for (int j = 0; j < s; ++j) Sig[i * s + j] *= fac[j];
    for (int i = 0; i < s * s; ++i) Sig[i] = fiterr * Sig[i] + (1.0L - fiterr) * Sigma0[i];
    Sig_inv = invert(Sig, s);
  }

  void init(int n_, int s_, int qcov, const double* Yin, const double* xin, const double* Xcin, double fe) {
    n = n_;
    s = s_;
    q = qcov + 1;
    fiterr = fe;
    Y.assign(n * s, 0);
    for (int i = 0; i < s; ++i)
      for (int r = 0; r < n; ++r) Y[r * s + i] = Yin[i * n + r];
    x.assign(n, 0);
    for (int r = 0; r < n; ++r) x[r] = xin[r];
    Xc.assign(n * q, 0);
    for (int r = 0; r < n; ++r) {
      Xc[r * q] = 1.0;
      for (int j = 1; j < q; ++j) Xc[r * q + j] = Xcin[(j - 1) * n + r];
    }
    m = q + s + 1;
    H.assign(s * s, 0);
    for (int i = 0; i < s; ++i) H[i * s + i] = 1e-4L;
    /* compute_common: T = I - Xc (Xc'Xc)^-1 Xc' */
    Mat XtX = matmul(Xc, Xc, q, n, q, true, false);
    Mat XtXi(q * q);
    if (q == 1)
      XtXi[0] = 1.0L / XtX[0];
    else
      XtXi = invert(XtX, q);
    Mat t1 = matmul(XtXi, Xc, q, q, n, false, true);
    Mat t2 = matmul(Xc, t1, n, q, n);
    T.assign(n * n, 0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) T[i * n + j] = (i == j ? 1 : 0) - t2[i * n + j];
    /* compute_Sigma_null */
    Mat TY = matmul(T, Y, n, n, s);
    Sigma0 = matmul(Y, TY, s, n, s, true, false);
    for (auto& v : Sigma0) v /= (ldouble)(n + m - q - s - 1);
    Sigma0_inv = invert(Sigma0, s);
  }

  /* compute_log10_ABF_vec (MVLR.cpp:715-764) for one configuration. */
  void abfs(const std::vector<int>& gamma, const double* phi2, const double* oma2, int K, double* out) {
    bool any = false;
    for (int v : gamma) any |= v != 0;
    if (!any) {
      for (int k = 0; k < K; ++k) out[k] = kNaN;
      return;
    }
    Mat Sig, Sig_inv;
    sigma(gamma, Sig, Sig_inv);
    /* compute_stats */
    std::vector<ldouble> G(n, 0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) G[i] += T[i * n + j] * x[j];
    ldouble kk = 0;
    for (int i = 0; i < n; ++i) kk += G[i] * G[i];
    Mat Vinv(s * s);
    for (int i = 0; i < s * s; ++i) Vinv[i] = kk * Sig_inv[i];
    std::vector<ldouble> YtG(s, 0), eVb(s, 0);
    for (int j = 0; j < s; ++j)
      for (int i = 0; i < n; ++i) YtG[j] += Y[i * s + j] * G[i];
    for (int i = 0; i < s; ++i)
      for (int j = 0; j < s; ++j) eVb[i] += Sig_inv[i * s + j] * YtG[j];
    /* construct_meta_Gamma */
    std::vector<ldouble> tv(s);
    for (int j = 0; j < s; ++j) tv[j] = gamma[j] * sqrtl(Sig[j * s + j]);
    for (int g = 0; g < K; ++g) {
      Mat W(s * s);
      for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) W[i * s + j] = oma2[g] * tv[i] * tv[j];
      for (int j = 0; j < s; ++j) W[j * s + j] = (oma2[g] + phi2[g]) * tv[j] * tv[j];
      Mat M = matmul(Vinv, W, s, s, s);
      for (int i = 0; i < s; ++i) M[i * s + i] += 1;
      std::vector<int> piv;
      int sg;
      lu_decomp(M, s, piv, sg);
      ldouble lndet = 0;
      for (int i = 0; i < s; ++i) lndet += logl(fabsl(M[i * s + i]));
      Mat Minv = lu_invert(M, s, piv);
      Mat t3 = matmul(W, Minv, s, s, s);
      ldouble quad = 0;
      for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) quad += eVb[i] * t3[i * s + j] * eVb[j];
      out[g] = (double)((0.5L * quad - 0.5L * lndet) / logl(10.0L));
    }
  }
};
}  // namespace

/* log10 ABFs of one SNP for one configuration over K grid points.
 * Y: [s][n]; x: [n]; Xc: [qcov][n]. */
extern "C" void orc_mvlr_abfs(int32_t n, int32_t s, int32_t qcov, const double* Y, const double* x,
                              const double* Xc, double fiterr, const int32_t* gamma, const double* phi2,
                              const double* oma2, int32_t K, double* out) {
  Mvlr mv;
  mv.init(n, s, qcov, Y, x, Xc, fiterr);
  std::vector<int> gm(gamma, gamma + s);
  mv.abfs(gm, phi2, oma2, K, out);
}

/* Sigma and Sigma^-1 for one configuration (for tests of the intermediate). */
extern "C" void orc_mvlr_sigma(int32_t n, int32_t s, int32_t qcov, const double* Y, const double* x,
                               const double* Xc, double fiterr, const int32_t* gamma, double* Sig_out,
                               double* Sig0_out) {
  Mvlr mv;
  mv.init(n, s, qcov, Y, x, Xc, fiterr);
  std::vector<int> gm(gamma, gamma + s);
  Mat Sig, Sig_inv;
  mv.sigma(gm, Sig, Sig_inv);
  for (int i = 0; i < s * s; ++i) {
    Sig_out[i] = (double)Sig[i];
    Sig0_out[i] = (double)mv.Sigma0[i];
  }
}

/* Association test of a chunk with the mvlr error model
 * (GeneSnpPair::CalcAbfsMvlr gene_snp_pair.cpp:730-758 and the functions it
 * calls :624-728).  Pairs whose SNP lacks genotypes in some subgroup, and
 * genes without expression in every subgroup, are skipped (n = 0 for every
 * subgroup and NaN ABFs), as gene.cpp:314-321 and eqtlbma_bf.cpp:755-762 do.
 * Every subgroup must use the same samples (the reference requires one
 * genotype file and reads Y by position, MVLR.cpp:29-41). */
extern "C" int orc_assoc_mvlr(const eqbf_chunk* c, int32_t bfs, double fiterr, const double* phiL,
                              const double* omaL, int32_t GL, const double* phiS, const double* omaS,
                              int32_t GS, int32_t* n_out, double* raw, double* w_abf) {
  const int32_t S = c->n_subgroups;
  const int64_t NP = c->n_pairs;
  const int64_t nc = n_configs(bfs, S);
  const int64_t n_raw = 3 * (int64_t)GL + nc * GS, n_w = 3 + nc + 2;
  std::vector<uint64_t> masks;
  config_masks(bfs, S, masks);
  std::vector<double> phiFix(GL), omaFix(GL), phiMax(GL), omaMax(GL);
  for (int g = 0; g < GL; ++g) {
    phiFix[g] = 0.0;
    omaFix[g] = phiL[g] + omaL[g];
    phiMax[g] = phiL[g] + omaL[g];
    omaMax[g] = 0.0;
  }
  for (int64_t g = 0; g < c->n_genes; ++g) {
    bool all_expr = true;
    for (int32_t s = 0; s < S; ++s) all_expr &= c->gene_in_sbgrp[g * S + s] != 0;
    for (int64_t p = c->pair_offset[g]; p < c->pair_offset[g + 1]; ++p) {
      int64_t snp = c->pair_snp[p];
      bool all_geno = true;
      for (int32_t s = 0; s < S; ++s) all_geno &= c->snp_in_sbgrp[snp * S + s] != 0;
      for (int32_t s = 0; s < S; ++s) n_out[(int64_t)s * NP + p] = 0;
      for (int64_t k = 0; k < n_raw; ++k) raw[k * NP + p] = kNaN;
      for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = kNaN;
      if (!all_expr || !all_geno) continue;
      std::vector<int64_t> idx0, idx;
      std::vector<double> y0, ys, Ymat;
      members(c, g, 0, nullptr, idx0, y0);
      int n = (int)idx0.size();
      Ymat.assign((size_t)S * n, 0);
      for (int32_t s = 0; s < S; ++s) {
        members(c, g, s, nullptr, idx, ys);
        if (idx != idx0) {
          snprintf(g_err, sizeof g_err, "--error mvlr needs the same samples in every subgroup");
          return EQBF_ERR_UNSUPPORTED;
        }
        for (int r = 0; r < n; ++r) Ymat[(size_t)s * n + r] = ys[r];
        n_out[(int64_t)s * NP + p] = n;
      }
      int qc = c->covar_count[0];
      std::vector<double> xv(n), Xc((size_t)qc * n);
      for (int r = 0; r < n; ++r) {
        xv[r] = geno_at(c, 0, idx0[r], snp);
        for (int j = 0; j < qc; ++j) Xc[(size_t)j * n + r] = c->covars[((int64_t)0 * c->n_covars_max + j) * c->n_samples + idx0[r]];
      }
      Mvlr mv;
      mv.init(n, S, qc, Ymat.data(), xv.data(), Xc.data(), fiterr);
      std::vector<int> gm(S, 1);
      std::vector<double> gen(GL), fix(GL), mx(GL);
      mv.abfs(gm, phiL, omaL, GL, gen.data());
      mv.abfs(gm, phiFix.data(), omaFix.data(), GL, fix.data());
      mv.abfs(gm, phiMax.data(), omaMax.data(), GL, mx.data());
      for (int k = 0; k < GL; ++k) {
        raw[(int64_t)k * NP + p] = gen[k];
        raw[(int64_t)(GL + k) * NP + p] = fix[k];
        raw[(int64_t)(2 * GL + k) * NP + p] = mx[k];
      }
      std::vector<double> wv(n_w, kNaN);
      wv[0] = orc_log10_weighted_sum(gen.data(), GL);
      wv[1] = orc_log10_weighted_sum(fix.data(), GL);
      wv[2] = orc_log10_weighted_sum(mx.data(), GL);
      if (bfs != EQBF_BFS_GEN) {
        std::vector<double> l(GS);
        for (int64_t ci = 0; ci < nc; ++ci) {
          for (int32_t s = 0; s < S; ++s) gm[s] = (masks[ci] >> s) & 1;
          Mvlr mc; /* the reference builds a new MVLR per configuration */
          mc.init(n, S, qc, Ymat.data(), xv.data(), Xc.data(), fiterr);
          mc.abfs(gm, phiS, omaS, GS, l.data());
          for (int k = 0; k < GS; ++k) raw[(3 * GL + ci * GS + k) * NP + p] = l[k];
          wv[3 + ci] = orc_log10_weighted_sum(l.data(), GS);
        }
        std::vector<double> v, wt;
        for (int32_t s = 0; s < S; ++s) {
          v.push_back(wv[3 + s]);
          wt.push_back(0.5 / S);
        }
        v.push_back(wv[0]);
        wt.push_back(0.5);
        wv[3 + nc] = orc_log10_weighted_sum_w(v.data(), wt.data(), (int64_t)v.size());
        if (bfs == EQBF_BFS_ALL) {
          v.clear();
          wt.clear();
          for (int64_t ci = 0; ci < nc; ++ci) {
            int k = __builtin_popcountll(masks[ci]);
            double ch = 1;
            for (int i = 1; i <= k; ++i) ch = ch * (S - k + i) / i;
            v.push_back(wv[3 + ci]);
            wt.push_back((1.0 / S) * (1.0 / ch));
          }
          wv[3 + nc + 1] = orc_log10_weighted_sum_w(v.data(), wt.data(), (int64_t)v.size());
        }
      }
      for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = wv[k];
    }
  }
  return EQBF_OK;
}

/* ======================================================================== */
/* --error hybrid (GeneSnpPair::CalcAbfsHybrid, gene_snp_pair.cpp:1384-1423, */
/* and what it calls, :760-1382; utils_math.cpp:249-348;                     */
/* samples.cpp:148-192).  Same dense steps as the reference:                 */
/*  - per subgroup, X = [1, genotype, covariates] over the members           */
/*    (FillStlContainers): pseudo-inverses from the eigen-decomposition of   */
/*    X'X (X's squared singular values, every one inverted as the reference  */
/*    does), rank = singular values > DBL_EPSILON, betahat, se, p, pve,      */
/*    sigmahat, sigma2 = rss / N for the full and the null ([1, covariates]) */
/*    model, Sigma_ss = f s2_full + (1 - f) s2_null, Vg_ss = Sigma_ss        */
/*    [(X'X)^-1]_11 (CalcBetahatsAndDiagsPerSubgroup :760-880);              */
/*  - per pair of subgroups, individuals in both / only the first / only the */
/*    second (present: genotyped and a non-NaN expression level), rows       */
/*    [1, genotype, covariates of the FIRST subgroup read at the             */
/*    individual's position in the union of samples] (:882-975),            */
/*    A_s = pinv(X12'X12 + Xs'Xs) X12' (:977-1013), the ML error covariance  */
/*    Y'(I - X pinv(X'X) X')Y / n12 of the full and null model               */
/*    (CalcMleErrorCovariance utils_math.cpp:306-348), Sigma_12 = f full_01 */
/*    + (1 - f) null_01, Vg_12 = Sigma_12 (A_1 A_2')_11 (:1035-1131);        */
/*  - CalcLog10AbfMvlr (:1165-1256) with explicit inverses.                  */
/* covars_pos: [S][Qmax][N], the covariate value at file column i of each   */
/* subgroup (NaN past the file's columns); NULL: the chunk's covariates.    */
/* expl_pair: [G][S][N] expression levels of the pairwise step, which the   */
/* reference reads without --qnorm's normalisation (Gene::GetExplevel) and */
/* without permutation; their NaN pattern also decides who is present in   */
/* the pairwise step.  NULL: the chunk's.                                   */
/* Pinned in tests/test_oracle_hybrid.py: per-subgroup quantities by numpy  */
/* least squares, the pairwise covariances by residual cross-products, Vg   */
/* off the diagonal by a different algebra, the Bayes factor by the ratio   */
/* of two multivariate normal densities (scipy).                            */
/* ======================================================================== */
namespace {

/* pseudo-inverse of a symmetric positive semi-definite matrix (its SVD is
 * its eigen-decomposition), every eigenvalue inverted */
static Mat pinv_sym(const Mat& A, int n, std::vector<ldouble>* evs = nullptr) {
  std::vector<ldouble> ev;
  Mat V;
  jacobi_eig(A, n, ev, V);
  Mat inv(n * n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      ldouble acc = 0;
      for (int k = 0; k < n; ++k) acc += V[i * n + k] * V[j * n + k] / ev[k];
      inv[i * n + j] = acc;
    }
  if (evs) *evs = ev;
  return inv;
}

/* [Y'(I - X pinv(X'X) X')Y / n]_01 for Y n x 2, X n x P */
static ldouble mle_cov01(const Mat& Y, const Mat& X, int n, int P) {
  Mat XtX = matmul(X, X, P, n, P, true, false);
  Mat Xi = pinv_sym(XtX, P);
  Mat t1 = matmul(X, Xi, n, P, P);              /* n x P */
  Mat t2 = matmul(t1, X, n, P, n, false, true); /* n x n */
  ldouble acc = 0;
  for (int i = 0; i < n; ++i) {
    ldouble Ty = Y[i * 2 + 1];
    for (int j = 0; j < n; ++j) Ty -= t2[i * n + j] * Y[j * 2 + 1];
    acc += Y[i * 2 + 0] * Ty;
  }
  return acc / n;
}

/* CalcLog10AbfMvlr (gene_snp_pair.cpp:1165-1256) */
static double hybrid_l10abf(int S, const std::vector<int>& gamma, const std::vector<ldouble>& b, const Mat& Sig,
                            const Mat& Vg, double phi2, double oma2) {
  Mat Wg(S * S);
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) Wg[i * S + j] = ((i == j) ? phi2 + oma2 : oma2) * gamma[i] * gamma[j];
  Mat Vi = invert(Vg, S);
  std::vector<ldouble> bV(S, 0);
  for (int j = 0; j < S; ++j)
    for (int k = 0; k < S; ++k) bV[j] += b[k] * Vi[k * S + j];
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) Wg[i * S + j] *= sqrtl(Sig[i * S + i]) * sqrtl(Sig[j * S + j]);
  Mat ivw = matmul(Vi, Wg, S, S, S);
  for (int i = 0; i < S; ++i) ivw[i * S + i] += 1;
  Mat LU = ivw;
  std::vector<int> piv;
  int sg;
  lu_decomp(LU, S, piv, sg);
  ldouble lnd = 0;
  for (int i = 0; i < S; ++i) lnd += logl(fabsl(LU[i * S + i]));
  Mat ivi = lu_invert(LU, S, piv);
  std::vector<ldouble> t4(S, 0), t5(S, 0);
  for (int j = 0; j < S; ++j)
    for (int k = 0; k < S; ++k) t4[j] += bV[k] * Wg[k * S + j];
  for (int j = 0; j < S; ++j)
    for (int k = 0; k < S; ++k) t5[j] += t4[k] * ivi[k * S + j];
  ldouble t6 = 0;
  for (int j = 0; j < S; ++j) t6 += t5[j] * bV[j];
  return (double)((-0.5L * lnd + 0.5L * t6) / logl(10.0L));
}

/* b, Sigma and Vg of one pair (CalcSstatsHybrid, :1133-1163); writes n and
 * the summary statistics of the pair when n_out / sstats are given.
 * Returns 0, or EQBF_ERR_DATA (two subgroups without a common individual,
 * where the reference exits). */
static int hybrid_pair(const eqbf_chunk* c, int64_t g, int64_t p, double fiterr, const double* covars_pos,
                       const double* expl_pair, std::vector<ldouble>& b, Mat& Sig, Mat& Vg, int32_t* n_out, double* sstats) {
  const int32_t S = c->n_subgroups;
  const int64_t NP = c->n_pairs, N = c->n_samples, QM = c->n_covars_max;
  const int64_t snp = c->pair_snp[p];
  const ldouble f = fiterr;
  /* the pairwise step picks its individuals from the levels it reads, which
   * permutations leave as loaded (samples.cpp:148-192) */
  auto present = [&](int32_t s, int64_t i) {
    const double y = expl_pair ? expl_pair[(g * S + s) * N + i] : expl_at(c, g, s, i);
    return c->geno_in_sbgrp[(int64_t)s * N + i] && !std::isnan(y);
  };
  b.assign(S, 0);
  std::vector<ldouble> sd(S), vd(S);
  /* per subgroup (CalcBetahatsAndDiagsPerSubgroup) */
  for (int32_t s = 0; s < S; ++s) {
    std::vector<int64_t> idx;
    std::vector<double> y;
    members(c, g, s, nullptr, idx, y);
    const int n = (int)idx.size();
    const int Q = c->covar_count[s], P = Q + 2;
    Mat X((size_t)n * P), Xc((size_t)n * (P - 1));
    for (int r = 0; r < n; ++r) {
      X[r * P + 0] = 1;
      X[r * P + 1] = geno_at(c, s, idx[r], snp);
      Xc[r * (P - 1) + 0] = 1;
      for (int q = 0; q < Q; ++q) {
        const ldouble v = c->covars[((int64_t)s * QM + q) * N + idx[r]];
        X[r * P + 2 + q] = v;
        Xc[r * (P - 1) + 1 + q] = v;
      }
    }
    std::vector<ldouble> ev;
    Mat XtX = matmul(X, X, P, n, P, true, false);
    Mat Xi = pinv_sym(XtX, P, &ev);
    int rank = 0;
    for (int k = 0; k < P; ++k)
      if (sqrtl(ev[k]) > 2.220446049250313e-16L) ++rank;
    std::vector<ldouble> Xty(P, 0), B(P, 0);
    for (int j = 0; j < P; ++j)
      for (int r = 0; r < n; ++r) Xty[j] += X[r * P + j] * y[r];
    for (int i = 0; i < P; ++i)
      for (int j = 0; j < P; ++j) B[i] += Xi[i * P + j] * Xty[j];
    ldouble rss = 0, ym = 0, tss = 0;
    for (int r = 0; r < n; ++r) {
      ldouble fit = 0;
      for (int j = 0; j < P; ++j) fit += X[r * P + j] * B[j];
      rss += (y[r] - fit) * (y[r] - fit);
      ym += y[r];
    }
    ym /= n;
    for (int r = 0; r < n; ++r) tss += (y[r] - ym) * (y[r] - ym);
    const ldouble sig = sqrtl(rss / (n - rank));
    const ldouble se = sig * sqrtl(Xi[1 * P + 1]);
    if (sstats) {
      double* o = sstats + (int64_t)s * NP + p;
      o[0] = (double)(1 - rss / tss);
      o[(int64_t)1 * S * NP] = (double)sig;
      o[(int64_t)2 * S * NP] = (double)B[1];
      o[(int64_t)3 * S * NP] = (double)se;
      o[(int64_t)4 * S * NP] = (double)(2 * tdist_lower_neg(fabsl(B[1] / se), (ldouble)(n - rank)));
    }
    if (n_out) n_out[(int64_t)s * NP + p] = n;
    /* null model */
    Mat XcX = matmul(Xc, Xc, P - 1, n, P - 1, true, false);
    Mat Xci = pinv_sym(XcX, P - 1);
    std::vector<ldouble> Xcy(P - 1, 0), Bn(P - 1, 0);
    for (int j = 0; j < P - 1; ++j)
      for (int r = 0; r < n; ++r) Xcy[j] += Xc[r * (P - 1) + j] * y[r];
    for (int i = 0; i < P - 1; ++i)
      for (int j = 0; j < P - 1; ++j) Bn[i] += Xci[i * (P - 1) + j] * Xcy[j];
    ldouble rn = 0;
    for (int r = 0; r < n; ++r) {
      ldouble fit = 0;
      for (int j = 0; j < P - 1; ++j) fit += Xc[r * (P - 1) + j] * Bn[j];
      rn += (y[r] - fit) * (y[r] - fit);
    }
    b[s] = B[1];
    sd[s] = f * rss / n + (1 - f) * rn / n;
    vd[s] = sd[s] * Xi[1 * P + 1];
  }
  /* pairs of subgroups (CalcOffDiagCovarsFromPairsOfSubgroups) */
  Sig.assign(S * S, 0);
  Vg.assign(S * S, 0);
  for (int32_t s = 0; s < S; ++s) {
    Sig[s * S + s] = sd[s];
    Vg[s * S + s] = vd[s];
  }
  for (int32_t s1 = 0; s1 + 1 < S; ++s1)
    for (int32_t s2 = s1 + 1; s2 < S; ++s2) {
      const int Q = c->covar_count[s1], P = Q + 2;
      std::vector<int64_t> i12, i1, i2;
      for (int64_t i = 0; i < N; ++i) {
        const bool a = present(s1, i), bb = present(s2, i);
        if (a && bb)
          i12.push_back(i);
        else if (a)
          i1.push_back(i);
        else if (bb)
          i2.push_back(i);
      }
      if (i12.empty()) {
        snprintf(g_err, sizeof g_err, "subgroups %d and %d have no individuals in common", s1, s2);
        return EQBF_ERR_DATA;
      }
      auto rows = [&](const std::vector<int64_t>& ii, Mat& X) {
        X.assign(ii.size() * P, 0);
        for (size_t r = 0; r < ii.size(); ++r) {
          const int64_t i = ii[r];
          X[r * P + 0] = 1;
          X[r * P + 1] = c->geno_in_sbgrp[(int64_t)s1 * N + i] ? geno_at(c, s1, i, snp) : kNaN;
          for (int q = 0; q < Q; ++q)
            X[r * P + 2 + q] = covars_pos ? covars_pos[((int64_t)s1 * QM + q) * N + i]
                                          : c->covars[((int64_t)s1 * QM + q) * N + i];
        }
      };
      Mat X12, X1, X2;
      rows(i12, X12);
      rows(i1, X1);
      rows(i2, X2);
      const int n12 = (int)i12.size();
      Mat Y12((size_t)n12 * 2);
      for (int r = 0; r < n12; ++r) {
        if (expl_pair) {
          Y12[r * 2 + 0] = expl_pair[(g * S + s1) * N + i12[r]];
          Y12[r * 2 + 1] = expl_pair[(g * S + s2) * N + i12[r]];
        } else {
          Y12[r * 2 + 0] = expl_at(c, g, s1, i12[r]);
          Y12[r * 2 + 1] = expl_at(c, g, s2, i12[r]);
        }
      }
      Mat G12 = matmul(X12, X12, P, n12, P, true, false);
      auto amat = [&](const Mat& Xu, int nu) {
        Mat Gs = G12;
        if (nu > 0) {
          Mat Gu = matmul(Xu, Xu, P, nu, P, true, false);
          for (int k = 0; k < P * P; ++k) Gs[k] += Gu[k];
        }
        return matmul(pinv_sym(Gs, P), X12, P, P, n12, false, true); /* P x n12 */
      };
      Mat A1 = amat(X1, (int)i1.size()), A2 = amat(X2, (int)i2.size());
      const ldouble full = mle_cov01(Y12, X12, n12, P);
      Mat Xc12((size_t)n12 * (P - 1));
      for (int r = 0; r < n12; ++r) {
        Xc12[r * (P - 1)] = 1;
        for (int j = 2; j < P; ++j) Xc12[r * (P - 1) + j - 1] = X12[r * P + j];
      }
      const ldouble null = mle_cov01(Y12, Xc12, n12, P - 1);
      const ldouble s12 = f * full + (1 - f) * null;
      ldouble a = 0;
      for (int r = 0; r < n12; ++r) a += A1[1 * n12 + r] * A2[1 * n12 + r];
      Sig[s1 * S + s2] = Sig[s2 * S + s1] = s12;
      Vg[s1 * S + s2] = Vg[s2 * S + s1] = s12 * a;
    }
  return 0;
}

}  // namespace

extern "C" int orc_assoc_hybrid(const eqbf_chunk* c, int32_t bfs, double fiterr, const double* covars_pos,
                                const double* expl_pair, const double* phiL, const double* omaL, int32_t GL, const double* phiS,
                                const double* omaS, int32_t GS, int32_t* n_out, double* sstats, double* raw,
                                double* w_abf) {
  const int32_t S = c->n_subgroups;
  const int64_t NP = c->n_pairs;
  const int64_t nc = n_configs(bfs, S);
  const int64_t n_raw = 3 * (int64_t)GL + nc * GS, n_w = 3 + nc + 2;
  std::vector<uint64_t> masks;
  config_masks(bfs, S, masks);
  for (int64_t g = 0; g < c->n_genes; ++g) {
    bool all_expr = true;
    for (int32_t s = 0; s < S; ++s) all_expr &= c->gene_in_sbgrp[g * S + s] != 0;
    for (int64_t p = c->pair_offset[g]; p < c->pair_offset[g + 1]; ++p) {
      const int64_t snp = c->pair_snp[p];
      bool all_geno = true;
      for (int32_t s = 0; s < S; ++s) all_geno &= c->snp_in_sbgrp[snp * S + s] != 0;
      for (int32_t s = 0; s < S; ++s) {
        n_out[(int64_t)s * NP + p] = 0;
        for (int k = 0; k < 5; ++k) sstats[((int64_t)k * S + s) * NP + p] = kNaN;
      }
      for (int64_t k = 0; k < n_raw; ++k) raw[k * NP + p] = kNaN;
      for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = kNaN;
      if (!all_expr || !all_geno) continue;
      std::vector<ldouble> b;
      Mat Sig, Vg;
      const int rc = hybrid_pair(c, g, p, fiterr, covars_pos, expl_pair, b, Sig, Vg, n_out, sstats);
      if (rc) return rc;
      /* Bayes factors (CalcAbfsHybridFor*) */
      std::vector<int> gm(S, 1);
      std::vector<double> gen(GL), fix(GL), mx(GL);
      for (int k = 0; k < GL; ++k) {
        gen[k] = hybrid_l10abf(S, gm, b, Sig, Vg, phiL[k], omaL[k]);
        fix[k] = hybrid_l10abf(S, gm, b, Sig, Vg, 0.0, phiL[k] + omaL[k]);
        mx[k] = hybrid_l10abf(S, gm, b, Sig, Vg, phiL[k] + omaL[k], 0.0);
        raw[(int64_t)k * NP + p] = gen[k];
        raw[(int64_t)(GL + k) * NP + p] = fix[k];
        raw[(int64_t)(2 * GL + k) * NP + p] = mx[k];
      }
      std::vector<double> wv(n_w, kNaN);
      wv[0] = orc_log10_weighted_sum(gen.data(), GL);
      wv[1] = orc_log10_weighted_sum(fix.data(), GL);
      wv[2] = orc_log10_weighted_sum(mx.data(), GL);
      if (bfs != EQBF_BFS_GEN) {
        std::vector<double> l(GS);
        for (int64_t ci = 0; ci < nc; ++ci) {
          for (int32_t s = 0; s < S; ++s) gm[s] = (masks[ci] >> s) & 1;
          for (int k = 0; k < GS; ++k) {
            l[k] = hybrid_l10abf(S, gm, b, Sig, Vg, phiS[k], omaS[k]);
            raw[(3 * GL + ci * GS + k) * NP + p] = l[k];
          }
          wv[3 + ci] = orc_log10_weighted_sum(l.data(), GS);
        }
        double gs, al;
        orc_bma(S, bfs, wv.data() + 3, wv[0], &gs, &al);
        wv[3 + nc] = gs;
        if (bfs == EQBF_BFS_ALL) wv[3 + nc + 1] = al;
      }
      for (int64_t k = 0; k < n_w; ++k) w_abf[k * NP + p] = wv[k];
    }
  }
  return 0;
}

/* b [S], Sigma [S][S] and Vg [S][S] of pair p of gene g (test access). */
extern "C" int orc_hybrid_pair_stats(const eqbf_chunk* c, int64_t g, int64_t p, double fiterr,
                                     const double* covars_pos, const double* expl_pair, double* b, double* Sig,
                                     double* Vg) {
  std::vector<ldouble> bb;
  Mat Sg, V;
  const int rc = hybrid_pair(c, g, p, fiterr, covars_pos, expl_pair, bb, Sg, V, nullptr, nullptr);
  if (rc) return rc;
  const int S = c->n_subgroups;
  for (int i = 0; i < S; ++i) {
    b[i] = (double)bb[i];
    for (int j = 0; j < S; ++j) {
      Sig[i * S + j] = (double)Sg[i * S + j];
      Vg[i * S + j] = (double)V[i * S + j];
    }
  }
  return 0;
}
