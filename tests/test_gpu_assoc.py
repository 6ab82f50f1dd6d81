"""GPU parity of the association test (uvlr) against the CPU oracle.

Tolerances (north-star precision statement):
  plain FP64 path: every summary statistic within 1e-10 relative, log10 ABFs
    within 1e-9 absolute (both sides FP64; the oracle factorises in long
    double).
  mixed path: betahat within 1e-6 of max(|betahat|, se) (relative error with
    the standard error as floor, so that betahat ~ 0 is not assessed by a
    meaningless ratio), se / sigmahat / p-value within 1e-6 relative, pve
    within 1e-8 absolute, log10 ABFs (raw and grid-averaged) within 1e-4
    absolute.
Sample sizes n and the NaN pattern must match exactly.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import synth

pytestmark = pytest.mark.gpu

GL25 = synth.make_grid("general")
GS10 = synth.make_grid("small")
G50 = synth.make_grid("general", n_aes=10)


@pytest.fixture(scope="module")
def eng():
    from eqtlbma_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _nan_same(a, b, what):
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), "%s: NaN pattern differs at %s" % (what, np.argwhere(na != nb)[:5])


def compare(gpu, ref, mode, analysis="join"):
    assert np.array_equal(gpu.n, ref.n), "sample sizes differ"
    if mode == "plain":
        rel, beta_rel, abf_abs = 1e-10, 1e-10, 1e-9
    else:
        rel, beta_rel, abf_abs = 1e-6, 1e-6, 1e-4
    has = ref.n > 0
    pve, sig, bet, se, pv = ref.sstats
    gpve, gsig, gbet, gse, gpv = gpu.sstats
    for name, a, b in [("pve", gpve, pve), ("sigmahat", gsig, sig), ("betahat", gbet, bet),
                       ("se", gse, se), ("pval", gpv, pv)]:
        _nan_same(a[has], b[has], name)
    ok = has & ~np.isnan(bet)
    assert np.all(np.abs(gbet[ok] - bet[ok]) <= beta_rel * np.maximum(np.abs(bet[ok]), se[ok]))
    for name, a, b in [("se", gse, se), ("sigmahat", gsig, sig)]:
        m = has & ~np.isnan(b)
        assert np.allclose(a[m], b[m], rtol=rel, atol=0), name
    m = has & ~np.isnan(pv) & (pv > 1e-290)
    assert np.allclose(gpv[m], pv[m], rtol=rel if mode == "plain" else 1e-6, atol=0), "pval"
    m = has & ~np.isnan(pve)
    assert np.allclose(gpve[m], pve[m], rtol=0, atol=1e-8 if mode != "plain" else 1e-11), "pve"
    # standardised statistics
    for k, name in enumerate(("b_std", "se_std", "z")):
        a, b = gpu.stdsstats[k][has], ref.stdsstats[k][has]
        _nan_same(a, b, name)
        fin = np.isfinite(b)
        assert np.allclose(a[fin], b[fin], rtol=rel if mode == "plain" else 1e-6,
                           atol=1e-12 if mode == "plain" else 1e-7), name
    if analysis != "join":
        return
    _nan_same(gpu.raw.astype(np.float64), ref.raw, "raw")
    fin = np.isfinite(ref.raw)
    err = np.abs(gpu.raw.astype(np.float64)[fin] - ref.raw[fin])
    assert err.max(initial=0) <= abf_abs, "raw ABF max abs err %g" % err.max()
    _nan_same(gpu.w, ref.w, "weighted")
    fin = np.isfinite(ref.w)
    err = np.abs(gpu.w[fin] - ref.w[fin])
    assert err.max(initial=0) <= abf_abs, "weighted ABF max abs err %g" % err.max()


def _run(eng, orc, hc, bfs, gridL, gridS, mode, analysis="join", tag=""):
    _say("[gpu] assoc %s mode=%s bfs=%s analysis=%s G=%d S=%d N=%d pairs=%d" %
         (tag, mode, bfs, analysis, hc.G, hc.S, hc.N, hc.n_pairs))
    eng.set_grids(gridL, gridS)
    gpu = eng.assoc_uvlr(hc, analysis=analysis, bfs=bfs, precision=mode)
    ref = orc.assoc_uvlr(hc, analysis, bfs, gridL, gridS if gridS is not None else gridL)
    compare(gpu, ref, mode, analysis)
    return gpu, ref


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_plain_method_small(eng, orc, mode):
    hc = synth.plain(seed=42, S=3, N=100, n_genes=6, snps_per_gene=40, frac_egenes=0.5)
    _run(eng, orc, hc, "all", GL25, GS10, mode, tag="plain-small")


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_batched_covariates_uneven(eng, orc, mode):
    hc = synth.batched(seed=1001, S=7, N_all=120, n_genes=4, p_mean=70, p_lo=20, p_hi=150,
                       ns_lo=30, ns_hi=110, Q=3)
    _run(eng, orc, hc, "sin", G50, G50, mode, tag="batched-cov")


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("n_all,q", [(800, 3), (120, 5)])
def test_prep_wave_path(eng, orc, mode, n_all, q):
    # more than 512 samples or more than 3 covariates: the prep keeps its
    # columns in LDS (prep_wave_kernel) instead of registers
    hc = synth.batched(seed=40 + q, S=3, N_all=n_all, n_genes=3, p_mean=20, p_lo=5, p_hi=40,
                       ns_lo=n_all // 3, ns_hi=n_all - 10, Q=q, nan_frac=0.02)
    _run(eng, orc, hc, "all", GL25, GS10, mode, tag="prep N=%d Q=%d" % (n_all, q))


@pytest.mark.parametrize("n_small", [1, 3, 8, 16])
def test_small_grid_sizes(eng, orc, n_small):
    # the batched Bayes factors sweep the small grid in blocks of 8 points
    # (full blocks, then a remainder): grids of one, part of one, exactly
    # one and exactly two blocks, with absent subgroups and missing values
    hc = synth.batched(seed=60 + n_small, S=5, N_all=100, n_genes=4, p_mean=40, p_lo=10, p_hi=80,
                       ns_lo=20, ns_hi=90, Q=2, nan_frac=0.03, absent_frac=0.2)
    gs = G50[3:3 + n_small]
    _run(eng, orc, hc, "all", GL25, gs, "mixed", tag="small grid of %d" % n_small)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_missing_values_and_absent_genes(eng, orc, mode):
    hc = synth.batched(seed=7, S=5, N_all=90, n_genes=5, p_mean=50, p_lo=10, p_hi=90,
                       ns_lo=20, ns_hi=80, Q=2, nan_frac=0.05, absent_frac=0.25)
    _run(eng, orc, hc, "all", GL25, GS10, mode, tag="missing")


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_too_few_samples_and_sep(eng, orc, mode):
    # some subgroups with fewer than Q+3 members -> NaN statistics
    hc = synth.batched(seed=11, S=6, N_all=40, n_genes=3, p_mean=30, p_lo=5, p_hi=40,
                       ns_lo=3, ns_hi=12, Q=2)
    _run(eng, orc, hc, "sin", GL25, GS10, mode, tag="too-few")
    _run(eng, orc, hc, "gen", GL25, None, mode, analysis="sep", tag="sep")


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_two_genotype_sets_and_bfs_gen(eng, orc, mode):
    hc = synth.batched(seed=5, S=4, N_all=60, n_genes=3, p_mean=40, p_lo=10, p_hi=60,
                       ns_lo=25, ns_hi=60, Q=0)
    g2 = np.clip(hc.genotypes + np.float32(0.5), 0, 2).astype(np.float32)
    hc2 = type(hc)(hc.explevels, hc.gene_in_sbgrp, hc.geno_in_sbgrp,
                   np.stack([hc.genotypes[0], g2[0]]), hc.snp_in_sbgrp, hc.pair_offset, hc.pair_snp,
                   covars=hc.covars, covar_count=hc.covar_count, covar_in_sbgrp=hc.covar_in_sbgrp,
                   geno_set_of_sbgrp=np.array([0, 1, 0, 1], np.int32))
    # SNPs missing in one subgroup
    hc2.snp_in_sbgrp[::7, 2] = 0
    _run(eng, orc, hc2, "gen", GL25, None, mode, tag="two-sets")


def test_ttail_tables_vs_oracle(eng, orc):
    _say("[gpu] t-tail tables")
    ts, nus = [], []
    for nu in [1, 2, 3, 5, 10, 30, 97, 238, 395]:
        for t in [0.0, 1e-5, 0.1, 0.5, 1.0, 1.7, 2.5, 4.0, 7.0, 12.0, 25.0, 40.0, 1e3, 1e6]:
            ts.append(t)
            nus.append(nu)
    w_tab, w_cf, z_tab, z_cf = eng.debug_ttail(np.array(ts), np.array(nus))
    for t, nu, wt, wc, zt, zc in zip(ts, nus, w_tab, w_cf, z_tab, z_cf):
        p_ref = orc.tdist_P(-t, nu)
        z_ref = orc.gauss_Pinv(p_ref)
        if p_ref > 1e-300:
            w_ref = -np.log(p_ref)
            assert wc == pytest.approx(w_ref, rel=1e-12, abs=1e-14), (t, nu)
            assert wt == pytest.approx(w_ref, rel=1e-10, abs=1e-11), (t, nu)
            assert zc == pytest.approx(z_ref, rel=1e-12, abs=1e-13), (t, nu)
            assert zt == pytest.approx(z_ref, rel=1e-9, abs=1e-10), (t, nu)


@pytest.mark.parametrize("mode", ["mixed"])
def test_large_sample_prep_path(eng, orc, mode):
    # N large enough that the per-(gene, subgroup) columns do not fit the
    # wave-level prep kernel's LDS; exercises the workgroup prep kernel
    hc = synth.batched(seed=21, S=3, N_all=2000, n_genes=2, p_mean=12, p_lo=5, p_hi=20,
                       ns_lo=1600, ns_hi=2000, Q=3, nan_frac=0.02)
    _run(eng, orc, hc, "all", GL25, GS10, mode, tag="large-N")


@pytest.mark.parametrize("bfs", ["sin", "all"])
def test_mixed_bayes_factors_only(eng, orc, bfs):
    # without FP64 statistic outputs the Bayes-factor inputs come from the
    # FP32 identities s'^2 = (n - rank) / ((n - P) x~'x~), b' = sign(beta)
    # |z| s' (FP64 where |t| < 1e-3, the residual sum of squares is not
    # positive, or the tail runs past the table): the Bayes factors stay
    # within 1e-4 of the oracle's; includes strong effects (far tails), an
    # exactly null effect and subgroups with too few samples
    hc = synth.batched(seed=31, S=6, N_all=150, n_genes=5, p_mean=60, p_lo=10, p_hi=120,
                       ns_lo=4, ns_hi=140, Q=2, nan_frac=0.02)  # 6 x (1 + 4) coefficient columns: LDS-direct path
    g = hc.genotypes[0]
    y = hc.explevels
    y[0, :, :] += 4.0 * g[:, hc.pair_snp[hc.pair_offset[0]]][None, :]  # |t| in the tens
    y[1, :, :] = 0.5 + 1e-9 * np.arange(hc.N)[None, :]                # near-null effects
    eng.set_grids(GL25, GS10)
    _say("[gpu] assoc mixed without statistic outputs bfs=%s pairs=%d" % (bfs, hc.n_pairs))
    gpu = eng.assoc_uvlr(hc, "join", bfs, precision="mixed", want_sstats=False, want_std=False)
    ref = orc.assoc_uvlr(hc, "join", bfs, GL25, GS10)
    assert np.array_equal(gpu.n, ref.n)
    for a, b, name in [(gpu.raw.astype(np.float64), ref.raw, "raw"), (gpu.w, ref.w, "weighted")]:
        _nan_same(a, b, name)
        fin = np.isfinite(b)
        err = np.abs(a[fin] - b[fin])
        _say("[gpu]   %s max abs err %.2e" % (name, err.max(initial=0)))
        assert err.max(initial=0) <= 1e-4, name


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_raw_tiled_layout(eng, orc, mode):
    # EQBF_F32_TILED: the same FP32 raw ABFs, in tiles of 64 pairs (a last,
    # narrower tile included); every other output unchanged.  Covers the
    # batched path (LDS-direct sums), the register-staged sums (odd column
    # count) and the path for two genotype sets.
    eng.set_grids(GL25, GS10)
    cases = [synth.batched(seed=3, S=4, N_all=80, n_genes=3, p_mean=70, p_lo=20, p_hi=150, ns_lo=30, ns_hi=80, Q=1),
             synth.batched(seed=4, S=3, N_all=70, n_genes=3, p_mean=50, p_lo=10, p_hi=90, ns_lo=25, ns_hi=70, Q=2)]
    hc = cases[0]
    g2 = np.clip(hc.genotypes + np.float32(0.25), 0, 2).astype(np.float32)
    cases.append(type(hc)(hc.explevels, hc.gene_in_sbgrp, hc.geno_in_sbgrp, np.stack([hc.genotypes[0], g2[0]]),
                          hc.snp_in_sbgrp, hc.pair_offset, hc.pair_snp, covars=hc.covars, covar_count=hc.covar_count,
                          covar_in_sbgrp=hc.covar_in_sbgrp, geno_set_of_sbgrp=np.array([0, 1, 0, 1], np.int32)))
    assert any(hc.n_pairs % 64 for hc in cases)
    for k, hc in enumerate(cases):
        _say("[gpu] raw tiled layout mode=%s case=%d pairs=%d" % (mode, k, hc.n_pairs))
        a = eng.assoc_uvlr(hc, "join", "all", precision=mode, raw_dtype=np.float32)
        b = eng.assoc_uvlr(hc, "join", "all", precision=mode, raw_tiled=True)
        assert np.array_equal(a.raw, b.raw, equal_nan=True)
        assert np.array_equal(a.w, b.w, equal_nan=True) and np.array_equal(a.n, b.n)
