"""GPU --error hybrid Bayes factors (eqbf_assoc_hybrid) against the oracle
(FP64 both sides).

The device splits every regression into a per-gene covariate block and
per-SNP sums (Schur complements) where the oracle forms and pseudo-inverts
each design in long double, so values agree to rounding: 1e-8 absolute plus
1e-9 relative on every raw and grid-averaged log10 ABF and on the summary
statistics; sample sizes exactly; NaN where the oracle has NaN.  Cases cover
the register-matrix variants (S <= 4, 8, 12, 16), the covariate-block
variants (Q + 1 <= 4, 8, 17), covariate counts differing per subgroup,
dosages in FP32, genes with several 64-pair tiles, every --bfs, the
pairwise covariates by sample index and the pairwise levels before
normalisation, skipped genes and SNPs, and the error for two subgroups
without a common individual.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import engine, synth

pytestmark = pytest.mark.gpu

GL = synth.make_grid("general")
GS = synth.make_grid("small")


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL, GS)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(a, b, what):
    assert np.array_equal(np.isnan(a), np.isnan(b)), what
    fin = np.isfinite(b)
    err = np.abs(a[fin] - b[fin])
    assert np.all(err <= 1e-8 + 1e-9 * np.abs(b[fin])), (what, float(err.max(initial=0)))
    return float(err.max(initial=0))


def _compare(gpu, ref):
    assert np.array_equal(gpu.n, ref.n)
    _close(gpu.sstats, ref.sstats, "sstats")
    e1 = _close(gpu.raw, ref.raw, "raw")
    e2 = _close(gpu.w, ref.w, "w")
    _say("[gpu]   max abs err raw %.2e  w %.2e" % (e1, e2))


@pytest.mark.parametrize("S,bfs,fiterr,Q,P,extra", [
    (3, "all", 0.5, 2, 4, {}),
    (2, "sin", 0.0, 0, 5, {}),
    (3, "gen", 1.0, 1, 150, {"dosages": True}),
    (5, "all", 0.5, 3, 7, {"q_per_sbgrp": [3, 1, 0, 2, 3]}),
    (6, "sin", 0.25, 5, 6, {}),
    (9, "sin", 0.5, 1, 5, {"miss": 0.15}),
    (11, "gen", 0.5, 9, 4, {"N": 90, "miss": 0.1}),
    (13, "sin", 0.5, 2, 3, {"miss": 0.1}),
])
def test_hybrid_matches_oracle(eng, orc, S, bfs, fiterr, Q, P, extra):
    kw = dict(S=S, Q=Q, P=P, n_genes=3)
    kw.update(extra)
    hc = synth.hybrid(seed=200 + S, **kw)
    _say("[gpu] hybrid S=%d bfs=%s fiterr=%g Q=%d pairs=%d %s" % (S, bfs, fiterr, Q, hc.n_pairs, extra))
    gpu = eng.assoc_hybrid(hc, bfs, fiterr)
    ref = orc.assoc_hybrid(hc, bfs, fiterr, GL, GS)
    _compare(gpu, ref)


def test_hybrid_pairwise_covariates_and_levels(eng, orc):
    hc = synth.hybrid(seed=31, S=4, Q=2, P=6, n_genes=2)
    rng = np.random.default_rng(32)
    cp = rng.normal(size=hc.covars.shape)
    ep = hc.explevels * 1.9 + 0.2
    _say("[gpu] hybrid pairwise covariates by sample index, levels before normalisation")
    gpu = eng.assoc_hybrid(hc, "sin", 0.5, covars_pos=cp, expl_pair=ep)
    ref = orc.assoc_hybrid(hc, "sin", 0.5, GL, GS, covars_pos=cp, expl_pair=ep)
    _compare(gpu, ref)


def test_hybrid_skips_genes_and_snps(eng, orc):
    hc = synth.hybrid(seed=33, S=3, Q=1, P=5, n_genes=3)
    hc.gene_in_sbgrp[1, 2] = 0
    hc.explevels[1, 2] = np.nan
    hc.snp_in_sbgrp[3, 0] = 0
    _say("[gpu] hybrid skipped gene and SNP")
    gpu = eng.assoc_hybrid(hc, "all", 0.5)
    ref = orc.assoc_hybrid(hc, "all", 0.5, GL, GS)
    _compare(gpu, ref)
    assert (gpu.n[:, 5:10] == 0).all() and np.isnan(gpu.w[:, 5:10]).all()


def test_hybrid_complete_data_fiterr0_is_mvlr(eng):
    hc = synth.mvlr(seed=34, S=4, N=60, n_genes=2, snps_per_gene=8, Q=2)
    _say("[gpu] hybrid = mvlr at fiterr 0 on complete data")
    h = eng.assoc_hybrid(hc, "all", 0.0)
    m = eng.assoc_mvlr(hc, "all", 0.0)
    assert np.allclose(h.raw, m.raw, rtol=1e-9, atol=1e-9)
    assert np.allclose(h.w, m.w, rtol=1e-9, atol=1e-9, equal_nan=True)


def test_hybrid_no_common_individual_is_an_error(eng):
    hc = synth.hybrid(seed=35, S=3, miss=0.0)
    hc.explevels[0, 0, :35] = np.nan
    hc.explevels[0, 1, 35:] = np.nan
    _say("[gpu] hybrid subgroups without a common individual")
    with pytest.raises(engine.EngineError, match="no individuals in common"):
        eng.assoc_hybrid(hc, "gen", 0.5)
