"""GPU: the int8-digit sums of the batched path (Bayes factors only) against
the oracle on the cases the digit arithmetic cannot settle by itself, and
against the FP64 product of the same path.

The sums kernel centres each SNP's genotypes on the mean of the tile's first
64 samples, scales them so that |x - c| <= 2^(k+1) fits 30 bits (2^k >
max(max |x - c| over those samples, 1)), and writes a recompute mark for a
SNP with a genotype that is not finite or outside that range, and for a
(pair, subgroup) whose x~'x~ keeps less than 1/8 of x''x'; the statistics
kernel recomputes those, and items with R^2 >= 7/8, in FP64 over the
members.  Each case below lands on one of those branches:
  - a NaN genotype of a sample that is a member of some subgroups only (the
    reference reads the members' genotypes only, gene_snp_pair.cpp:79-170:
    NaN statistics where the sample is a member, finite elsewhere);
  - a SNP monomorphic everywhere, and one constant over one subgroup's
    members (rank-deficient fits: NaN genotype effect, ABF roles of the
    reference);
  - a rare variant whose carriers all come after the first 64 samples
    (zero spread where the scale is chosen);
  - a genotype far outside the first samples' range (x = 40);
  - a SNP within 1% of a covariate over one subgroup (x~'x~ ~ 1e-4 x'x);
  - a very strong effect (R^2 > 0.99).
Tolerances: sample sizes and NaN patterns exact, log10 ABFs within 1e-4 of
the oracle's (the mixed path's bound); between the int8 and the FP64
products of the same path, within 2e-5.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import synth

pytestmark = pytest.mark.gpu

GL25 = synth.make_grid("general", n_aes=5)
GS10 = synth.make_grid("small", n_aes=5)


@pytest.fixture(scope="module")
def eng():
    from eqtlbma_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edge_case():
    hc = synth.batched(seed=77, S=5, N_all=150, n_genes=3, p_mean=30, p_lo=24, p_hi=40,
                       ns_lo=50, ns_hi=140, Q=2, nan_frac=0.02)
    G = hc.genotypes[0]  # [N][n_snps] FP32
    y = hc.explevels
    N = hc.N
    col = [int(hc.pair_snp[hc.pair_offset[0] + k]) for k in range(7)]
    memb = [(hc.geno_in_sbgrp[s] > 0) & ~np.isnan(y[0, s]) for s in range(hc.S)]
    # NaN at a sample that is a member of subgroup 0 and not of subgroup 1
    i0 = int(np.nonzero(memb[0] & ~memb[1])[0][0])
    G[i0, col[0]] = np.nan
    G[:, col[1]] = 1.0
    G[memb[2], col[2]] = 0.5
    G[:, col[3]] = 0.0
    G[[120, 131, 140], col[3]] = 2.0
    G[100, col[4]] = 40.0
    cov = hc.covars[3, 0].astype(np.float32)
    G[:, col[5]] = cov + np.float32(0.01) * np.random.default_rng(5).normal(size=N).astype(np.float32)
    y[0, 4] = np.where(memb[4], 12.0 * G[:, col[6]] + 0.05 * np.random.default_rng(6).normal(size=N), np.nan)
    return hc


def _cmp(gpu, ref, tol):
    assert np.array_equal(gpu.n, ref.n), "sample sizes"
    for a, b, name in [(gpu.raw.astype(np.float64), ref.raw, "raw"), (gpu.w, ref.w, "weighted")]:
        na, nb = np.isnan(a), np.isnan(b)
        assert np.array_equal(na, nb), "%s NaN pattern differs at %s" % (name, np.argwhere(na != nb)[:5])
        fin = np.isfinite(b)
        err = float(np.max(np.abs(a[fin] - b[fin]), initial=0))
        _say("[gpu]   %s max abs err %.2e" % (name, err))
        assert err <= tol, name


@pytest.mark.parametrize("bfs", ["sin", "all"])
def test_i8_sums_edge_cases_vs_oracle(eng, orc, bfs):
    hc = _edge_case()
    eng.set_grids(GL25, GS10)
    eng.profile(True)
    _say("[gpu] int8 sums edge cases bfs=%s pairs=%d" % (bfs, hc.n_pairs))
    gpu = eng.assoc_uvlr(hc, "join", bfs, precision="mixed", want_sstats=False, want_std=False)
    prof = eng.profile_read()
    eng.profile(False)
    assert "b8_convert_kernel" in prof, "the int8-digit sums did not run"
    ref = orc.assoc_uvlr(hc, "join", bfs, GL25, GS10)
    _cmp(gpu, ref, 1e-4)
    # the NaN genotype: NaN where its sample is a member, finite elsewhere
    p0 = int(hc.pair_offset[0])
    assert ref.n[0, p0] > 0 and ref.n[1, p0] > 0


def test_i8_sums_vs_fp64_product(eng, monkeypatch):
    """Same chunk through the int8 and the FP64 products: the Bayes-factor
    inputs differ by ~1e-8 relative, below their FP32 rounding."""
    hc = synth.batched(seed=91, S=12, N_all=300, n_genes=6, p_mean=150, p_lo=60, p_hi=260,
                       ns_lo=40, ns_hi=290, Q=3, nan_frac=0.01)
    eng.set_grids(GL25, GS10)
    _say("[gpu] int8 vs FP64 sums pairs=%d" % hc.n_pairs)
    a = eng.assoc_uvlr(hc, "join", "sin", precision="mixed", want_sstats=False, want_std=False)
    monkeypatch.setenv("EQBF_I8_SUMS", "0")
    b = eng.assoc_uvlr(hc, "join", "sin", precision="mixed", want_sstats=False, want_std=False)
    assert np.array_equal(a.n, b.n)
    for x, y, name in [(a.raw.astype(np.float64), b.raw.astype(np.float64), "raw"), (a.w, b.w, "weighted")]:
        assert np.array_equal(np.isnan(x), np.isnan(y)), name
        fin = np.isfinite(y)
        err = float(np.max(np.abs(x[fin] - y[fin]), initial=0))
        _say("[gpu]   %s int8 vs FP64 product max abs diff %.2e" % (name, err))
        assert err <= 2e-5, name


@pytest.mark.parametrize("n_all", [128, 192, 256, 320])
def test_i8_persistent_half_tiles(eng, monkeypatch, n_all):
    """~30,000 pairs, so each of the persistent workgroups runs about four
    32-pair half tiles one after another (the centre and scale of each, the
    digit buffers and the genotype registers rotating over 2, 3, 4 and 5
    sample steps of 64).  Same chunk through the int8 and the FP64
    products, as above."""
    ns_hi = min(300, n_all - 8)
    hc = synth.batched(seed=93, S=6, N_all=n_all, n_genes=60, p_mean=500, p_lo=400, p_hi=600,
                       ns_lo=40, ns_hi=ns_hi, Q=2, nan_frac=0.01)
    eng.set_grids(GL25, GS10)
    _say("[gpu] int8 persistent half tiles N=%d pairs=%d" % (n_all, hc.n_pairs))
    a = eng.assoc_uvlr(hc, "join", "sin", precision="mixed", want_sstats=False, want_std=False)
    monkeypatch.setenv("EQBF_I8_SUMS", "0")
    b = eng.assoc_uvlr(hc, "join", "sin", precision="mixed", want_sstats=False, want_std=False)
    assert np.array_equal(a.n, b.n)
    for x, y, name in [(a.raw.astype(np.float64), b.raw.astype(np.float64), "raw"), (a.w, b.w, "weighted")]:
        assert np.array_equal(np.isnan(x), np.isnan(y)), name
        fin = np.isfinite(y)
        err = float(np.max(np.abs(x[fin] - y[fin]), initial=0))
        _say("[gpu]   %s int8 vs FP64 product max abs diff %.2e" % (name, err))
        assert err <= 2e-5, name
