"""Pins the oracle's --error hybrid restatement (orc_assoc_hybrid in
oracle/eqbf_oracle_mvlr.cpp; the reference's GeneSnpPair::CalcAbfsHybrid,
quantgen/gene_snp_pair.cpp:1384-1423, and what it calls).

* per subgroup: b, Sigma_ss and Vg_ss from numpy least squares of the full
  ([1, genotype, covariates]) and null ([1, covariates]) models over the
  subgroup's members, sigma2 = rss / n;
* per pair of subgroups: Sigma_12 from the cross-product of the two
  subgroups' least-squares residuals over the individuals present in both,
  and Vg_12 = Sigma_12 e1' M1^-1 G12 M2^-1 e1 (M_s = G12 + X_s'X_s), an
  algebra different from the oracle's A_1 A_2' product of pseudo-inverses;
* the Bayes factor as the log10 ratio of two multivariate normal densities
  N(b; 0, Vg + W) / N(b; 0, Vg) (scipy);
* with fiterr = 0, every individual in every subgroup and the same
  covariates, the hybrid model is the multivariate regression at fiterr = 0
  (Sigma = null residual covariance, Vg = Sigma / x~'x~ by Frisch-Waugh):
  the whole chunk output equals the oracle's MVLR output, itself pinned by
  tests/test_oracle_mvlr.py;
* the summary statistics are those of the univariate separate analysis;
* permutations (Gene::MakePermutationsJoin, gene.cpp:669-672) permute the
  per-subgroup fits only: the pairwise step reads the levels and picks the
  individuals as loaded (gene_snp_pair.cpp:1157-1159, samples.cpp:148-192);
  each permuted statistic from numpy fits and scipy densities.
"""
import numpy as np
import pytest
from scipy import stats

from eqtlbma_amd import synth
from eqtlbma_amd._abi import config_masks

LN10 = np.log(10.0)


def hybrid_chunk(seed, **kw):
    return synth.hybrid(seed=seed, **kw)


def _lsq_resid(X, y):
    coef, *_ = np.linalg.lstsq(X, y, rcond=None)
    return coef, y - X @ coef


def numpy_pair(hc, g, p, f, covars_pos=None, expl_pair=None, perm=None):
    """perm: individual i of the per-subgroup fits takes the level of
    individual perm[i]; the pairwise step reads the levels as loaded."""
    S, N = hc.S, hc.N
    x = hc.genotypes[0][:, hc.pair_snp[p]]
    yp = hc.explevels[g] if expl_pair is None else expl_pair[g]
    y = hc.explevels[g] if perm is None else hc.explevels[g][:, perm]
    pres = ~np.isnan(y) & (hc.geno_in_sbgrp == 1)
    presp = ~np.isnan(yp) & (hc.geno_in_sbgrp == 1)
    b, Sig, Vg = np.zeros(S), np.zeros((S, S)), np.zeros((S, S))
    for s in range(S):
        Q = hc.covar_count[s]
        m = pres[s]
        Xc = np.column_stack([np.ones(m.sum())] + [hc.covars[s, q, m] for q in range(Q)])
        X = np.column_stack([Xc[:, :1], x[m], Xc[:, 1:]])
        coef, e1 = _lsq_resid(X, y[s, m])
        _, e0 = _lsq_resid(Xc, y[s, m])
        n = m.sum()
        b[s] = coef[1]
        Sig[s, s] = f * (e1 @ e1) / n + (1 - f) * (e0 @ e0) / n
        Vg[s, s] = Sig[s, s] * np.linalg.inv(X.T @ X)[1, 1]
    cov = hc.covars if covars_pos is None else covars_pos
    for s1 in range(S):
        for s2 in range(s1 + 1, S):
            Q = hc.covar_count[s1]

            def design(mask):
                return np.column_stack([np.ones(mask.sum()), x[mask]] + [cov[s1, q, mask] for q in range(Q)])
            both, only1, only2 = presp[s1] & presp[s2], presp[s1] & ~presp[s2], ~presp[s1] & presp[s2]
            X12 = design(both)
            n12 = both.sum()
            Y12 = np.column_stack([yp[s1, both], yp[s2, both]])
            _, E1 = _lsq_resid(X12, Y12)
            _, E0 = _lsq_resid(np.delete(X12, 1, axis=1), Y12)
            s12 = f * (E1[:, 0] @ E1[:, 1]) / n12 + (1 - f) * (E0[:, 0] @ E0[:, 1]) / n12
            G12 = X12.T @ X12
            M1 = G12 + design(only1).T @ design(only1)
            M2 = G12 + design(only2).T @ design(only2)
            e1 = np.eye(X12.shape[1])[1]
            a = np.linalg.solve(M1, e1) @ G12 @ np.linalg.solve(M2, e1)
            Sig[s1, s2] = Sig[s2, s1] = s12
            Vg[s1, s2] = Vg[s2, s1] = s12 * a
    return b, Sig, Vg


def mvn_l10abf(b, Sig, Vg, gamma, phi2, oma2):
    sd = np.sqrt(np.diag(Sig)) * gamma
    W = oma2 * np.outer(sd, sd) + phi2 * np.diag(sd * sd)
    z = np.zeros(len(b))
    return (stats.multivariate_normal(z, Vg + W).logpdf(b) - stats.multivariate_normal(z, Vg).logpdf(b)) / LN10


@pytest.mark.parametrize("f", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("seed", [1, 2])
def test_pair_quantities_vs_numpy(orc, seed, f):
    hc = hybrid_chunk(seed)
    for g in range(hc.G):
        for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
            b, Sig, Vg = orc.hybrid_pair_stats(hc, g, p, f)
            wb, wSig, wVg = numpy_pair(hc, g, p, f)
            assert np.allclose(b, wb, rtol=1e-10, atol=1e-12)
            assert np.allclose(Sig, wSig, rtol=1e-10, atol=1e-13)
            assert np.allclose(Vg, wVg, rtol=1e-9, atol=1e-14)


def test_pairwise_covariates_by_file_column(orc):
    """The pairwise designs read the first subgroup's covariates at the
    individual's index in the union of samples (covars_pos)."""
    hc = hybrid_chunk(5, Q=1)
    cp = np.random.default_rng(6).normal(size=hc.covars.shape)
    b, Sig, Vg = orc.hybrid_pair_stats(hc, 0, 1, 0.5, covars_pos=cp)
    wb, wSig, wVg = numpy_pair(hc, 0, 1, 0.5, covars_pos=cp)
    assert np.allclose(Vg, wVg, rtol=1e-9, atol=1e-14) and np.allclose(Sig, wSig, rtol=1e-10, atol=1e-13)
    _, _, Vg0 = orc.hybrid_pair_stats(hc, 0, 1, 0.5)
    assert not np.allclose(Vg[0, 1], Vg0[0, 1], rtol=1e-6)
    assert np.allclose(np.diag(Vg), np.diag(Vg0), rtol=1e-12)


def test_pairwise_levels_before_normalisation(orc):
    """Under --qnorm the reference normalises the levels of the per-subgroup
    fits only; the pairwise step reads the levels as loaded (expl_pair)."""
    hc = hybrid_chunk(9)
    ep = hc.explevels * 2.5 + np.sin(np.arange(hc.N))[None, None, :]
    for p in range(hc.pair_offset[1]):
        b, Sig, Vg = orc.hybrid_pair_stats(hc, 0, p, 0.5, expl_pair=ep)
        wb, wSig, wVg = numpy_pair(hc, 0, p, 0.5, expl_pair=ep)
        assert np.allclose(b, wb, rtol=1e-10, atol=1e-12)
        assert np.allclose(Sig, wSig, rtol=1e-10, atol=1e-13) and np.allclose(Vg, wVg, rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("bfs", ["gen", "sin", "all"])
def test_abfs_are_normal_density_ratios(orc, bfs):
    hc = hybrid_chunk(3)
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    out = orc.assoc_hybrid(hc, bfs, 0.5, gL, gS)
    masks = list(config_masks(bfs, hc.S)) if bfs != "gen" else []
    for g in range(hc.G):
        for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
            b, Sig, Vg = numpy_pair(hc, g, p, 0.5)
            one = np.ones(hc.S)
            want = [mvn_l10abf(b, Sig, Vg, one, ph, om) for ph, om in gL]
            want += [mvn_l10abf(b, Sig, Vg, one, 0.0, ph + om) for ph, om in gL]
            want += [mvn_l10abf(b, Sig, Vg, one, ph + om, 0.0) for ph, om in gL]
            for mk in masks:
                gam = np.array([(mk >> s) & 1 for s in range(hc.S)], float)
                want += [mvn_l10abf(b, Sig, Vg, gam, ph, om) for ph, om in gS]
            assert np.allclose(out.raw[:, p], want, rtol=1e-8, atol=1e-9)
            assert out.w[0, p] == pytest.approx(orc.log10_weighted_sum(out.raw[:len(gL), p]), abs=1e-12)


def test_fiterr0_complete_data_is_mvlr(orc):
    hc = synth.mvlr(seed=21, S=4, N=60, n_genes=3, snps_per_gene=6, Q=2)
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    h = orc.assoc_hybrid(hc, "all", 0.0, gL, gS)
    m = orc.assoc_mvlr(hc, "all", 0.0, gL, gS)
    assert np.allclose(h.raw, m.raw, rtol=1e-9, atol=1e-9)
    assert np.allclose(h.w, m.w, rtol=1e-9, atol=1e-9, equal_nan=True)


def test_sstats_are_the_separate_analysis(orc):
    hc = hybrid_chunk(4, miss=0.3)
    gL = synth.make_grid("general")
    h = orc.assoc_hybrid(hc, "gen", 0.5, gL, None)
    u = orc.assoc_uvlr(hc, "sep", "gen", gL, None)
    assert np.array_equal(h.n, u.n)
    assert np.allclose(h.sstats, u.sstats, rtol=1e-10, atol=1e-12, equal_nan=True)


def test_gene_or_snp_missing_in_a_subgroup_gives_nan(orc):
    hc = hybrid_chunk(7)
    hc.gene_in_sbgrp[1, 2] = 0
    hc.explevels[1, 2] = np.nan
    hc.snp_in_sbgrp[0, 1] = 0
    out = orc.assoc_hybrid(hc, "sin", 0.5, synth.make_grid("general"), synth.make_grid("small"))
    P0, P1 = hc.pair_offset[1], hc.pair_offset[2]
    assert np.isnan(out.w[:, P0:P1]).all() and np.isnan(out.w[:, 0]).all()
    assert np.isfinite(out.w[:3, 1:P0]).all()


def test_no_common_individual_is_an_error(orc):
    hc = hybrid_chunk(8, miss=0.0)
    hc.explevels[0, 0, :35] = np.nan
    hc.explevels[0, 1, 35:] = np.nan
    with pytest.raises(RuntimeError, match="no individuals in common"):
        orc.assoc_hybrid(hc, "gen", 0.5, synth.make_grid("general"), None)


def _l10_mean(v):
    m = np.max(v)
    return m + np.log10(np.mean(10.0 ** (np.asarray(v) - m)))


@pytest.mark.parametrize("use_max", [True, False])
def test_permutations_permute_the_subgroup_fits_only(orc, use_max):
    hc = hybrid_chunk(71, S=3, N=48, n_genes=2, P=5, Q=1, miss=0.15)
    gL, gS = synth.make_grid("general")[::3], synth.make_grid("small")
    rng = np.random.default_rng(5)
    table = np.array([rng.permutation(hc.N) for _ in range(3)] + [np.arange(hc.N)], np.int32)
    one = np.ones(hc.S)
    for g in range(hc.G):
        got = orc.perm_join_stats_hybrid(hc, g, "gen", use_max, 0.5, gL, gS, table)
        assert np.isfinite(got).all()
        for r, perm in enumerate(table):
            v = []
            for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
                b, Sig, Vg = numpy_pair(hc, g, p, 0.5, perm=perm)
                v.append(_l10_mean([mvn_l10abf(b, Sig, Vg, one, ph, om) for ph, om in gL]))
            want = max(v) if use_max else _l10_mean(v)
            assert got[r] == pytest.approx(want, abs=1e-8)
        out = orc.assoc_hybrid(hc, "gen", 0.5, gL, gS)
        w0 = out.w[0, hc.pair_offset[g]:hc.pair_offset[g + 1]]
        assert got[-1] == pytest.approx(max(w0) if use_max else _l10_mean(w0), abs=1e-12)
