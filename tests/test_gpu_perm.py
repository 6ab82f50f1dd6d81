"""GPU permutation statistics against the CPU oracle.

The index tables come from the library's GSL-semantics generator (seeded
like the reference, once per chunk); both sides evaluate the same tables.
Tolerances: plain FP64 path 1e-9 absolute on each permutation statistic
(log10 BF) or 1e-10 relative (p-values); mixed path 1e-4 absolute on the
log10 BF statistic, 1e-6 relative on p-values.  Exceedance counts and
p-values from the reference's counting loop must then agree exactly: on the
FP64 paths directly, on the mixed path after eqbf_perm_stats_fp64_ties has
recomputed in FP64 the true statistics and every permutation statistic
within 5e-5 of them (about 40 times the mixed path's measured error).
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import engine, synth

pytestmark = pytest.mark.gpu

GL25 = synth.make_grid("general")
GS10 = synth.make_grid("small")


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL25, GS10)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_stats(orc, hc, tables, gene_table, analysis, pbf, use_max, sep_sbgrp):
    out = np.full((hc.G, tables.shape[1]), np.nan)
    for g in range(hc.G):
        if gene_table[g] < 0:
            continue
        if analysis == "join":
            out[g] = orc.perm_join_stats(hc, g, pbf, use_max, GL25, GS10, tables[gene_table[g]])
        else:
            out[g] = orc.perm_sep_stats(hc, g, sep_sbgrp, tables[gene_table[g]])
    return out


def _check(gpu, ref, mode, analysis):
    assert np.array_equal(np.isnan(gpu), np.isnan(ref))
    fin = np.isfinite(ref)
    if analysis == "join":
        tol = 1e-9 if mode == "plain" else 1e-4
        assert np.max(np.abs(gpu[fin] - ref[fin]), initial=0) <= tol
    else:
        rtol = 1e-10 if mode == "plain" else 1e-6
        assert np.allclose(gpu[fin], ref[fin], rtol=rtol, atol=1e-300)


def _true_in(eng, orc, hc, gene_table, pbf, use_max, mode):
    """The caller's true statistics, as the host program forms them: from
    the association pass of the same precision."""
    a = eng.assoc_uvlr(hc, "join", "all", precision=mode)
    nc = 2 ** hc.S - 1
    row = {"all": 3 + nc + 1, "gen-sin": 3 + nc, "gen": 0}[pbf]
    return np.array([orc.true_stat(a.w[row, hc.pair_offset[g]:hc.pair_offset[g + 1]], use_max)
                     if gene_table[g] >= 0 else np.nan for g in range(hc.G)])


def _count_parity(eng, orc, hc, tables, gene_table, nperm, pbf, use_max, mode, ref):
    """Counting loop on both sides: the library's statistics and true
    statistics (FP64 wherever a permutation came within the bound) against
    the oracle's FP64 ones; counts, used permutations and p-values
    identical."""
    t_in = _true_in(eng, orc, hc, gene_table, pbf, use_max, mode)
    gpu, true_gpu, n_re = eng.perm_stats_fp64_ties(hc, tables, gene_table, t_in, "join", pbf, use_max,
                                                   precision=mode)
    _check(gpu, ref, mode, "join")
    true = orc.assoc_uvlr(hc, "join", "all", GL25, GS10)
    nc = 2 ** hc.S - 1
    row = {"all": 3 + nc + 1, "gen-sin": 3 + nc, "gen": 0}[pbf]
    for g in range(hc.G):
        if gene_table[g] < 0:
            continue
        w = true.w[row, hc.pair_offset[g]:hc.pair_offset[g + 1]]
        t = orc.true_stat(w, use_max)
        # replaced by the FP64 value, or the caller's within its error
        tol = 1e-9 if (true_gpu[g] != t_in[g] or mode != "mixed") else 1e-4
        assert abs(true_gpu[g] - t) <= tol, (g, true_gpu[g], t)
        a = engine.perm_pvalue(gpu[g], true_gpu[g], nperm)
        b = orc.perm_count(ref[g], t, nperm)
        assert a[0] == b[0] and a[1] == b[1], (g, a, b)
    return n_re, true_gpu, t_in


@pytest.mark.parametrize("mode", ["plain", "fp64_batched", "mixed"])
@pytest.mark.parametrize("pbf,use_max", [("all", True), ("gen-sin", False)])
def test_join_perm_stats(eng, orc, mode, pbf, use_max):
    hc = synth.plain(seed=7, S=3, N=40, n_genes=4, snps_per_gene=12, frac_egenes=0.5)
    nperm = 40
    tables = engine.Rng(7).tables(hc.N, 3, nperm)
    gene_table = np.array([0, 1, -1, 2], np.int32)
    _say("[gpu] perm join mode=%s pbf=%s max=%s" % (mode, pbf, use_max))
    gpu = eng.perm_stats(hc, tables, gene_table, "join", pbf, use_max, precision=mode)
    ref = _oracle_stats(orc, hc, tables, gene_table, "join", pbf, use_max, -1)
    _check(gpu, ref, "mixed" if mode == "mixed" else "plain", "join")
    if mode == "fp64_batched":
        plain = eng.perm_stats(hc, tables, gene_table, "join", pbf, use_max, precision="plain")
        assert np.array_equal(gpu, plain, equal_nan=True)
    n_re, _t, _ti = _count_parity(eng, orc, hc, tables, gene_table, nperm, pbf, use_max, mode, ref)
    _say("[gpu]   re-checked in FP64: %d" % n_re)
    if mode != "mixed":
        assert n_re == 0


def test_exact_ties_count_on_both_sides(eng, orc):
    # identity permutations reproduce the true statistic exactly: the
    # reference counts them (>=); on the mixed path they are near-ties and
    # their FP64 re-check gives the true statistic's own arithmetic
    hc = synth.plain(seed=19, S=3, N=40, n_genes=3, snps_per_gene=10, frac_egenes=0.67)
    nperm = 30
    tables = engine.Rng(9).tables(hc.N, 3, nperm)
    for k in range(3):
        tables[k, 3] = np.arange(hc.N)
        tables[k, 17] = np.arange(hc.N)
    gene_table = np.arange(3, dtype=np.int32)
    _say("[gpu] exact ties (identity permutations)")
    ref = _oracle_stats(orc, hc, tables, gene_table, "join", "all", True, -1)
    n_re, _t, t_in = _count_parity(eng, orc, hc, tables, gene_table, nperm, "all", True, "mixed", ref)
    assert n_re >= 6
    gpu, true_gpu, _ = eng.perm_stats_fp64_ties(hc, tables, gene_table, t_in, "join", "all", True,
                                                precision="mixed")
    for k in range(3):
        assert true_gpu[k] != t_in[k]  # replaced by the FP64 value
        assert gpu[k, 3] == true_gpu[k] and gpu[k, 17] == true_gpu[k]


def test_every_statistic_rechecked_equals_fp64(eng, orc):
    # a bound that covers every statistic: the mixed call returns the FP64
    # path's statistics (contraction genes and general-path genes alike)
    hc = synth.plain(seed=29, S=4, N=48, n_genes=3, snps_per_gene=14, frac_egenes=0.34)
    hc.explevels[2, 1, 5] = np.nan  # gene 2 takes the general path
    nperm = 24
    tables = engine.Rng(4).tables(hc.N, 3, nperm)
    gene_table = np.arange(3, dtype=np.int32)
    _say("[gpu] every statistic re-checked")
    plain = eng.perm_stats(hc, tables, gene_table, "join", "all", True, precision="plain")
    t_in = _true_in(eng, orc, hc, gene_table, "all", True, "mixed")
    t_plain = _true_in(eng, orc, hc, gene_table, "all", True, "plain")
    gpu, true_gpu, n_re = eng.perm_stats_fp64_ties(hc, tables, gene_table, t_in, "join", "all", True,
                                                   precision="mixed", tie_bound=1e6)
    assert n_re == int(np.isfinite(plain).sum())
    assert np.array_equal(gpu, plain, equal_nan=True)
    assert np.allclose(true_gpu, t_plain, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_join_perm_missing_values_and_covariates(eng, orc, mode):
    # membership changes with the permutation (NaN expression values)
    hc = synth.batched(seed=17, S=3, N_all=50, n_genes=3, p_mean=8, p_lo=4, p_hi=12, ns_lo=30, ns_hi=50,
                       Q=2, nan_frac=0.1)
    nperm = 25
    tables = engine.Rng(3).tables(hc.N, 3, nperm)
    gene_table = np.arange(3, dtype=np.int32)
    _say("[gpu] perm join missing mode=%s" % mode)
    gpu = eng.perm_stats(hc, tables, gene_table, "join", "all", True, precision=mode)
    ref = _oracle_stats(orc, hc, tables, gene_table, "join", "all", True, -1)
    _check(gpu, ref, mode, "join")


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("sbgrp", [-1, 1])
def test_sep_perm_stats(eng, orc, mode, sbgrp):
    hc = synth.batched(seed=23, S=3, N_all=45, n_genes=3, p_mean=6, p_lo=3, p_hi=10, ns_lo=25, ns_hi=45,
                       Q=1, nan_frac=0.05, absent_frac=0.2)
    nperm = 30
    tables = engine.Rng(5).tables(hc.N, 3, nperm)
    gene_table = np.arange(3, dtype=np.int32)
    _say("[gpu] perm sep mode=%s sbgrp=%d" % (mode, sbgrp))
    gpu = eng.perm_stats(hc, tables, gene_table, "sep", "gen", True, sep_sbgrp=sbgrp, precision=mode)
    ref = _oracle_stats(orc, hc, tables, gene_table, "sep", "gen", True, sbgrp)
    _check(gpu, ref, mode, "sep")
    # FP64 true statistics (FindMinTruePvalue*) and re-checked near-ties:
    # the counting loop (<=) gives the oracle's counts
    def min_p(a, g):
        t = 1.0
        for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
            for s in range(hc.S):
                if (sbgrp >= 0 and s != sbgrp) or not hc.gene_in_sbgrp[g, s]:
                    continue
                if a.n[s, p] > 0 and a.sstats[4, s, p] < t:
                    t = a.sstats[4, s, p]
        return t

    a_gpu = eng.assoc_uvlr(hc, "sep", "gen", precision=mode)
    t_in = np.array([min_p(a_gpu, g) for g in range(hc.G)])
    gpu2, true_gpu, n_re = eng.perm_stats_fp64_ties(hc, tables, gene_table, t_in, "sep", "gen", True,
                                                    sep_sbgrp=sbgrp, precision=mode, tie_bound=1e-6)
    true = orc.assoc_uvlr(hc, "sep", "gen", GL25, None)
    for g in range(hc.G):
        t = min_p(true, g)
        assert true_gpu[g] == pytest.approx(t, rel=1e-9, abs=0), (g, true_gpu[g], t)
        a = engine.perm_pvalue(gpu2[g], true_gpu[g], nperm, sep=True)
        b = orc.perm_count(ref[g], t, nperm, sep=True)
        assert a[0] == b[0] and a[1] == b[1], (g, a, b)
