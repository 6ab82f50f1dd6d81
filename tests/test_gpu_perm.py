"""GPU permutation statistics against the CPU oracle.

The index tables come from the library's GSL-semantics generator (seeded
like the reference, once per chunk); both sides evaluate the same tables.
Tolerances: plain FP64 path 1e-9 absolute on each permutation statistic
(log10 BF) or 1e-10 relative (p-values); mixed path 1e-4 absolute on the
log10 BF statistic, 1e-6 relative on p-values.  Exceedance counts and
p-values from the reference's counting loop must then agree exactly for the
plain path; for the mixed path they agree except where a permutation
statistic lies within the tolerance of the true statistic (near-ties,
reported by the test).
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


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("pbf,use_max", [("all", True), ("gen-sin", False)])
def test_join_perm_stats(eng, orc, mode, pbf, use_max):
    hc = synth.plain(seed=7, S=3, N=40, n_genes=4, snps_per_gene=12, frac_egenes=0.5)
    nperm = 40
    tables = engine.Rng(7).tables(hc.N, 3, nperm)
    gene_table = np.array([0, 1, -1, 2], np.int32)
    _say("[gpu] perm join mode=%s pbf=%s max=%s" % (mode, pbf, use_max))
    gpu = eng.perm_stats(hc, tables, gene_table, "join", pbf, use_max, precision=mode)
    ref = _oracle_stats(orc, hc, tables, gene_table, "join", pbf, use_max, -1)
    _check(gpu, ref, mode, "join")
    # the reference's counting loop on both sides
    true = orc.assoc_uvlr(hc, "join", "all", GL25, GS10)
    row = {"all": 3 + 7 + 1, "gen-sin": 3 + 7}[pbf]
    for g in range(hc.G):
        if gene_table[g] < 0:
            continue
        w = true.w[row, hc.pair_offset[g]:hc.pair_offset[g + 1]]
        t = orc.true_stat(w, use_max)
        a = engine.perm_pvalue(gpu[g], t, nperm)
        b = orc.perm_count(ref[g], t, nperm)
        near = np.sum(np.abs(ref[g] - t) <= 1e-4)
        if mode == "plain" or near == 0:
            assert a[0] == b[0] and a[1] == b[1]


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
