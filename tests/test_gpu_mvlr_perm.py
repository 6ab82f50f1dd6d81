"""Permutations with --error mvlr (Gene::MakePermutationsJoin, gene.cpp:
598-717, the mvlr branch :662-672) against the oracle.

Each permutation's statistic comes from the MVLR Bayes factors of the
permuted expression levels (FP64 on both sides); a SNP without genotypes in
every subgroup contributes 0 to the statistic (its initial value in the
reference's loop).  Tolerance: 1e-9 absolute on each statistic (log10 BF);
counts, used permutations and p-values of the counting loop identical.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import engine, synth

pytestmark = pytest.mark.gpu

GL = synth.make_grid("general")[::3]   # 9 points
GS = synth.make_grid("small")[::2]     # 5 points


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL, GS)
    yield e
    e.close()


@pytest.mark.parametrize("pbf,use_max", [("all", True), ("gen", False), ("gen-sin", True)])
@pytest.mark.parametrize("fiterr", [0.0, 0.5])
def test_mvlr_perm_stats_match_oracle(eng, orc, pbf, use_max, fiterr):
    hc = synth.mvlr(seed=61, S=3, N=36, n_genes=3, snps_per_gene=10, frac_egenes=0.67, Q=1)
    hc.snp_in_sbgrp[hc.pair_snp[hc.pair_offset[1] + 2], 1] = 0  # skipped by mvlr: contributes 0
    nperm = 16
    tables = engine.Rng(23).tables(hc.N, 3, nperm)
    gene_table = np.array([0, 1, 2], np.int32)
    _say("[gpu] mvlr perm pbf=%s max=%s fiterr=%g" % (pbf, use_max, fiterr))
    gpu = eng.perm_stats(hc, tables, gene_table, "join", pbf, use_max, precision="mixed", error_model="mvlr",
                         fiterr=fiterr)
    ref = np.array([orc.perm_join_stats_mvlr(hc, g, pbf, use_max, fiterr, GL, GS, tables[g]) for g in range(3)])
    assert np.array_equal(np.isnan(gpu), np.isnan(ref))
    fin = np.isfinite(ref)
    err = float(np.max(np.abs(gpu[fin] - ref[fin])))
    _say("[gpu]   max |stat - oracle| = %.3g" % err)
    assert err <= 1e-9
    # counting loop with each side's true statistic (FP64 association)
    bfs = {"all": "all", "gen": "gen", "gen-sin": "sin"}[pbf]
    a_gpu = eng.assoc_mvlr(hc, bfs, fiterr)
    a_ref = orc.assoc_mvlr(hc, bfs, fiterr, GL, GS)
    nc = {"all": 7, "gen": 0, "gen-sin": 3}[pbf]
    row = {"all": 3 + nc + 1, "gen": 0, "gen-sin": 3 + nc}[pbf]
    for g in range(3):
        p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
        t_gpu = orc.true_stat(a_gpu.w[row, p0:p1], use_max)
        t_ref = orc.true_stat(a_ref.w[row, p0:p1], use_max)
        assert abs(t_gpu - t_ref) <= 1e-9
        a = engine.perm_pvalue(gpu[g], t_gpu, nperm)
        b = orc.perm_count(ref[g], t_ref, nperm)
        assert a[0] == b[0] and a[1] == b[1], (g, a, b)


def test_mvlr_perm_snp_without_all_genotypes_counts_zero(eng, orc):
    # every SNP of gene 0 lacks genotypes in subgroup 2: every permutation's
    # statistic is the maximum (or average) of zeros, i.e. 0
    hc = synth.mvlr(seed=62, S=3, N=30, n_genes=2, snps_per_gene=6, frac_egenes=0.5)
    for p in range(hc.pair_offset[0], hc.pair_offset[1]):
        hc.snp_in_sbgrp[hc.pair_snp[p], 2] = 0
    tables = engine.Rng(5).tables(hc.N, 2, 8)
    gpu = eng.perm_stats(hc, tables, np.array([0, 1], np.int32), "join", "all", True, error_model="mvlr")
    assert np.all(gpu[0] == 0.0)
    ref = orc.perm_join_stats_mvlr(hc, 1, "all", True, 0.5, GL, GS, tables[1])
    assert np.max(np.abs(gpu[1] - ref)) <= 1e-9
