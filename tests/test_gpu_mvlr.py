"""GPU --error mvlr Bayes factors against the CPU oracle (FP64 both sides).

Tolerance: 1e-8 absolute (plus 1e-9 relative) on every raw and grid-averaged
log10 ABF; the sample sizes must match exactly and skipped pairs must be NaN
on both sides.  Cases cover the register-matrix variants (S <= 4, 8, 12, 16),
covariates, fiterr 0 (Sigma0) and > 0 (mixed ML covariance with the F to
chi-square correction), every --bfs, skipped genes and SNPs, and the error
for subgroups with different samples.
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


def _compare(gpu, ref):
    assert np.array_equal(gpu.n, ref.n)
    for a, b in ((gpu.raw, ref.raw), (gpu.w, ref.w)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        fin = np.isfinite(b)
        err = np.abs(a[fin] - b[fin])
        assert np.all(err <= 1e-8 + 1e-9 * np.abs(b[fin])), float(err.max())


@pytest.mark.parametrize("S,bfs,fiterr,Q", [
    (3, "all", 0.0, 0), (3, "all", 0.5, 1), (2, "sin", 1.0, 2),
    (6, "sin", 0.5, 2), (5, "all", 0.0, 1),
    (9, "sin", 0.5, 1), (11, "gen", 0.25, 0),
    (13, "sin", 0.5, 2),
])
def test_mvlr_matches_oracle(eng, orc, S, bfs, fiterr, Q):
    hc = synth.mvlr(seed=100 + S, S=S, N=70, n_genes=3, snps_per_gene=9, Q=Q)
    _say("[gpu] mvlr S=%d bfs=%s fiterr=%g Q=%d" % (S, bfs, fiterr, Q))
    gpu = eng.assoc_mvlr(hc, bfs, fiterr)
    ref = orc.assoc_mvlr(hc, bfs, fiterr, GL, GS)
    _compare(gpu, ref)


def test_mvlr_skips_genes_and_snps(eng, orc):
    hc = synth.mvlr(seed=7, S=4, N=60, n_genes=3, snps_per_gene=6, Q=1)
    hc.gene_in_sbgrp[1, 2] = 0          # gene 1 lacks expression in subgroup 2
    hc.explevels[1, 2, :] = np.nan
    hc.snp_in_sbgrp[3, 1] = 0           # SNP 3 lacks genotypes in subgroup 1
    _say("[gpu] mvlr skipped genes/snps")
    gpu = eng.assoc_mvlr(hc, "all", 0.5, want_sstats=True)
    ref = orc.assoc_mvlr(hc, "all", 0.5, GL, GS)
    _compare(gpu, ref)
    assert (gpu.n[:, 6:12] == 0).all() and (gpu.n[:, 3] == 0).all()
    assert np.isnan(gpu.w[:, 6:12]).all()
    assert np.isnan(gpu.sstats).all()  # no summary statistics with --error mvlr


def test_mvlr_rejects_different_samples(eng, orc):
    hc = synth.mvlr(seed=8, S=3, N=40, n_genes=2, snps_per_gene=4)
    hc.explevels[0, 1, 5] = np.nan      # sample 5 missing in subgroup 1 only
    _say("[gpu] mvlr different samples")
    with pytest.raises(engine.EngineError) as ei:
        eng.assoc_mvlr(hc, "gen", 0.0)
    assert "same samples" in str(ei.value)
    with pytest.raises(RuntimeError):
        orc.assoc_mvlr(hc, "gen", 0.0, GL, GS)
