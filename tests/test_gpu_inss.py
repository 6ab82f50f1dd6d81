"""--inss on the GPU: Bayes factors from given summary statistics
(eqbf_abf_from_sstats) against the oracle.  Both sides are FP64; the
standardised statistics must agree to 1e-10 relative and the log10 ABFs
(raw and grid-averaged) to 1e-9 absolute, NaN where the oracle has NaN.
Cases: random statistics over a range of sample sizes, n = 0 (no result),
n <= 2 (NaN z), NaN and zero effects, a t ratio past FP64's tail (z = -inf
side), one subgroup, and the round trip of the data path's own summary
statistics (which must reproduce the data path's Bayes factors).  Pairs
with a t-tail probability in FP64's subnormal range (0 < p < 1e-300) are
left out of the comparisons: there FP64 keeps fewer digits and the
reference's own value depends on its rounding (parity unpinned).
"""
import sys

import numpy as np
import pytest
from scipy import special

from eqtlbma_amd import engine, synth

pytestmark = pytest.mark.gpu

GL = synth.make_grid("general")
GS = synth.make_grid("small")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL, GS)
    yield e
    e.close()


def _random_sstats(seed, S, NP):
    rng = np.random.default_rng(seed)
    n = rng.integers(3, 1000, size=(S, NP))
    sig = rng.uniform(0.3, 3.0, size=(S, NP))
    beta = rng.normal(scale=0.3, size=(S, NP)) * rng.choice([0.1, 1.0, 5.0], size=(S, NP))
    se = rng.uniform(0.02, 0.6, size=(S, NP))
    k = NP // 10
    n[rng.integers(S, size=k), rng.integers(NP, size=k)] = 0
    n[0, :3] = [1, 2, 0]
    beta[S - 1, 3] = np.nan
    beta[0, 4] = 0.0
    n[0, 5], beta[0, 5], se[0, 5] = 40, 50.0, 1e-3  # |t| = 5e4, nu = 38: p ~ 1e-180
    n[0, 6], beta[0, 6], se[0, 6] = 500, 50.0, 1e-3  # nu = 498: p underflows, z = -inf
    return n, sig, beta, se


def _comparable(n, sig, beta, se):
    """Pairs whose every subgroup's t-tail probability is 0 or normal."""
    with np.errstate(all="ignore"):
        t = np.abs((beta / sig) / (se / sig))
        p = special.stdtr(np.maximum(n - 2.0, 1.0), -t)
    sub = (n > 2) & (p > 0) & (p < 1e-300)
    return ~np.any(sub, axis=0)


def _check(got, std, raw, w, keep):
    std, raw, w = std[..., keep], raw[:, keep], w[:, keep]
    got_std, got_raw, got_w = got.stdsstats[..., keep], got.raw[:, keep], got.w[:, keep]
    fin = np.isfinite(std)
    assert np.array_equal(fin, np.isfinite(got_std))
    assert np.allclose(got_std[fin], std[fin], rtol=1e-10, atol=0)
    assert np.array_equal(np.isnan(got_raw), np.isnan(raw))
    m = ~np.isnan(raw)
    assert np.max(np.abs(got_raw[m] - raw[m]), initial=0.0) <= 1e-9
    assert np.array_equal(np.isnan(got_w), np.isnan(w))
    m = ~np.isnan(w)
    assert np.max(np.abs(got_w[m] - w[m]), initial=0.0) <= 1e-9


@pytest.mark.parametrize("bfs", ["gen", "sin", "all"])
@pytest.mark.parametrize("S", [1, 4])
def test_inss_matches_oracle(eng, orc, bfs, S):
    _say("[gpu] inss bfs=%s S=%d" % (bfs, S))
    n, sig, beta, se = _random_sstats(10 + S, S, 3000)
    got = eng.abf_from_sstats(n, sig, beta, se, bfs)
    std, raw, w = orc.abf_from_sstats(n, sig, beta, se, bfs, GL, GS)
    keep = _comparable(n, sig, beta, se)
    assert keep.sum() > 0.95 * keep.size
    _check(got, std, raw, w, keep)
    assert np.array_equal(got.n, np.where(n > 0, n, 0).astype(np.int32))


def test_inss_raw_float32(eng, orc):
    _say("[gpu] inss raw float32")
    n, sig, beta, se = _random_sstats(3, 3, 500)
    got = eng.abf_from_sstats(n, sig, beta, se, "all", raw_dtype=np.float32)
    _std, raw, w = orc.abf_from_sstats(n, sig, beta, se, "all", GL, GS)
    keep = _comparable(n, sig, beta, se)
    raw, w, graw, gw = raw[:, keep], w[:, keep], got.raw[:, keep], got.w[:, keep]
    m = ~np.isnan(raw)
    assert np.allclose(graw[m], raw[m].astype(np.float32), rtol=1e-6, atol=1e-6)
    m = ~np.isnan(w)
    assert np.max(np.abs(gw[m] - w[m])) <= 1e-9


def test_inss_round_trip_of_data_path(eng, orc):
    _say("[gpu] inss round trip")
    hc = synth.plain(seed=8, S=3, N=80, n_genes=5, snps_per_gene=40, frac_egenes=0.6)
    ref = eng.assoc_uvlr(hc, "join", "all", precision="fp64_plain")
    got = eng.abf_from_sstats(ref.n, ref.sstats[1], ref.sstats[2], ref.sstats[3], "all")
    m = ~np.isnan(ref.w)
    assert np.array_equal(np.isnan(got.w), ~m)
    assert np.max(np.abs(got.w[m] - ref.w[m])) <= 1e-9
    m = ~np.isnan(ref.raw)
    assert np.max(np.abs(got.raw[m] - ref.raw[m])) <= 1e-9


def test_inss_empty_and_errors(eng):
    _say("[gpu] inss empty input and argument errors")
    z = np.zeros((2, 0))
    got = eng.abf_from_sstats(z.astype(np.int64), z, z, z, "sin")
    assert got.w.shape[1] == 0
    big = np.ones((65, 1))
    with pytest.raises(engine.EngineError, match="at most 64 subgroups"):
        eng.abf_from_sstats(big.astype(np.int64), big, big, big, "sin")
