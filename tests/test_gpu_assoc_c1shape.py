"""GPU parity at the batched-ABF workload's real shape.

44 subgroups of U(80, 400) individuals drawn from 500, FP32 dosages, 3
covariates, cis SNPs ~ Poisson(1500) clipped to [200, 6000], 50-point grids,
--bfs sin: the shape the throughput figure is measured on.  At this shape
the batched sums kernel is the LDS-direct instance with two matrix-core
column waves (44 x (1 + 4) = 220 coefficient columns, even, so the 16-byte
staging applies; 44 > 32 subgroups, so two column blocks per wave pair), its
epilogue runs in several passes of subgroups (the default shape takes
floor(80 KiB / (64 x 8 x 6 B)) = 26 of the 44 per pass, the tuned shapes 22
or 11), and with more 64-pair tiles than the 512 workgroups of the
persistent grid some workgroups loop over several tiles.

With Bayes factors as the only outputs the sums are the int8-digit product
(exact integer digit products, FP64 recompute of the ill-conditioned
items); with the FP64 statistics written, or EQBF_I8_SUMS=0, the FP64
product.  The whole chunk runs on the GPU; the oracle (FP64 / long double, Gene::
TestForAssociations, gene.cpp:263-332, GeneSnpPair::CalcAbfsUvlr,
gene_snp_pair.cpp:364-464, 552-570) runs on genes spread across the chunk.
Tolerances are those of the mixed path (north-star precision statement):
beta-hat within 1e-6 of max(|beta-hat|, se), se / sigma-hat / p within 1e-6
relative, log10 ABFs (raw and grid-averaged) within 1e-4 absolute; sample
sizes and NaN patterns exact.
"""
import concurrent.futures as cf
import sys

import numpy as np
import pytest

from eqtlbma_amd import _abi, synth

pytestmark = pytest.mark.gpu

G50 = synth.make_grid("general", n_aes=10)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _subchunk(hc, genes):
    """The listed genes of a chunk as a chunk of their own (same genotype
    matrix, pairs re-based)."""
    sel = [np.arange(hc.pair_offset[g], hc.pair_offset[g + 1]) for g in genes]
    offs = np.concatenate([[0], np.cumsum([len(s) for s in sel])]).astype(np.int64)
    pairs = np.concatenate(sel)
    sub = _abi.HostChunk(hc.explevels[genes], hc.gene_in_sbgrp[genes], hc.geno_in_sbgrp, hc.genotypes,
                         hc.snp_in_sbgrp, offs, hc.pair_snp[pairs], covars=hc.covars, covar_count=hc.covar_count,
                         covar_in_sbgrp=hc.covar_in_sbgrp, geno_set_of_sbgrp=hc.geno_set_of_sbgrp)
    return sub, pairs


@pytest.fixture(scope="module")
def case(orc):
    hc = synth.batched(seed=1001, S=44, N_all=500, n_genes=28, p_mean=1500, p_lo=200, p_hi=6000,
                       ns_lo=80, ns_hi=400, Q=3, frac_egenes=0.15)
    P = np.diff(hc.pair_offset)
    n_tiles = int(np.sum((P + 63) // 64))
    assert n_tiles > 600, n_tiles
    genes = [0, 9, 18, 27]
    # the oracle, one gene per thread (the ctypes call releases the GIL)
    def one(g):
        sub, pairs = _subchunk(hc, [g])
        return orc.assoc_uvlr(sub, "join", "sin", G50, G50), pairs
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        refs = list(ex.map(one, genes))
    _say("[gpu] c1 shape: %d genes, %d pairs, %d tiles, oracle on genes %s" % (hc.G, hc.n_pairs, n_tiles, genes))
    return hc, refs


def _cmp(gpu, refs, with_sstats):
    for ref, pairs in refs:
        assert np.array_equal(gpu.n[:, pairs], ref.n), "sample sizes"
        has = ref.n > 0
        if with_sstats:
            pve, sig, bet, se, pv = ref.sstats
            gpve, gsig, gbet, gse, gpv = (a[:, pairs] for a in gpu.sstats)
            for a, b, name in [(gbet, bet, "betahat"), (gse, se, "se"), (gpv, pv, "pval")]:
                assert np.array_equal(np.isnan(a[has]), np.isnan(b[has])), name
            ok = has & ~np.isnan(bet)
            assert np.all(np.abs(gbet[ok] - bet[ok]) <= 1e-6 * np.maximum(np.abs(bet[ok]), se[ok])), "betahat"
            assert np.allclose(gse[ok], se[ok], rtol=1e-6, atol=0), "se"
            assert np.allclose(gsig[ok], sig[ok], rtol=1e-6, atol=0), "sigmahat"
            m = ok & (pv > 1e-290)
            assert np.allclose(gpv[m], pv[m], rtol=1e-6, atol=0), "pval"
            assert np.allclose(gpve[ok], pve[ok], rtol=0, atol=1e-8), "pve"
        for a, b, name in [(gpu.raw[:, pairs].astype(np.float64), ref.raw, "raw"), (gpu.w[:, pairs], ref.w, "weighted")]:
            assert np.array_equal(np.isnan(a), np.isnan(b)), name + " NaN pattern"
            fin = np.isfinite(b)
            err = float(np.max(np.abs(a[fin] - b[fin]), initial=0))
            _say("[gpu]   %s max abs err %.2e over %d values" % (name, err, int(fin.sum())))
            assert err <= 1e-4, name


@pytest.mark.parametrize("with_sstats,i8", [(False, True), (False, False), (True, True)])
def test_c1_shape_default_launch(case, with_sstats, i8, monkeypatch):
    """Bayes factors only: the int8-digit sums (or, with EQBF_I8_SUMS=0, the
    FP64 product); with the FP64 statistics written: the FP64 product."""
    from eqtlbma_amd.engine import Engine
    hc, refs = case
    if not i8:
        monkeypatch.setenv("EQBF_I8_SUMS", "0")
    eng = Engine(0)
    try:
        eng.set_grids(G50, G50)
        eng.profile(True)
        _say("[gpu] c1 shape, default launch shape, FP64 statistic outputs=%s, int8 sums allowed=%s" %
             (with_sstats, i8))
        gpu = eng.assoc_uvlr(hc, "join", "sin", precision="mixed", raw_tiled=True,
                             want_sstats=with_sstats, want_std=with_sstats)
        prof = eng.profile_read()
        # the batched sums kernel and the separate statistics kernel ran
        assert prof["uvlr_sums_kernel"][1] == 1 and prof["uvlr_stats_kernel"][1] == 1
        assert "assoc_mixed_kernel" not in prof
        assert ("b8_convert_kernel" in prof) == (i8 and not with_sstats)
        _cmp(gpu, refs, with_sstats)
    finally:
        eng.close()


def test_c1_shape_tuned_launch_shapes(case, monkeypatch):
    """The autotuner's key names the instance (44 subgroups, 5 coefficient
    columns each, 7 column blocks per wave, 2 column waves, 500 samples,
    FP32 LDS-direct staging); every shape it times and the one it keeps
    give the same sums in the same order, so the outputs equal those of the
    default shape bit for bit."""
    from eqtlbma_amd.engine import Engine
    hc, refs = case
    monkeypatch.setenv("EQBF_I8_SUMS", "0")  # the tuner times the FP64 product's launch shapes
    a = Engine(0)
    b = Engine(0)
    try:
        a.set_grids(G50, G50)
        b.set_grids(G50, G50)
        b.set_autotune(True)
        _say("[gpu] c1 shape, default vs tuned launch shape")
        ga = a.assoc_uvlr(hc, "join", "sin", precision="mixed", raw_tiled=True, want_sstats=False, want_std=False)
        gb = b.assoc_uvlr(hc, "join", "sin", precision="mixed", raw_tiled=True, want_sstats=False, want_std=False)
        choices = b.autotune_read()
        _say("[gpu]   tuned: %s" % choices)
        assert list(choices) == ["sums/44/5/7/2/500/dma32"]
        ch = choices["sums/44/5/7/2/500/dma32"]
        assert 1 <= ch["spp"] < 44
        assert np.array_equal(ga.n, gb.n)
        assert np.array_equal(ga.raw, gb.raw, equal_nan=True)
        assert np.array_equal(ga.w, gb.w, equal_nan=True)
        _cmp(gb, refs, False)
    finally:
        a.close()
        b.close()
