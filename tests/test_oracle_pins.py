"""Pins of oracle functions the reference's own tests cannot reach here.

BMAlite / BMA weights (GeneSnpPair::CalcBMAlite / CalcBMA,
gene_snp_pair.cpp:552-602) are checked against the independent R statement
of the reference's functional tests (tests/functional_tests.R:1343-1381:
log10WeightedSum, calcAvgAbfsSinAndGenSin, calcAvgAbfAll), restated here in
numpy for S = 3 as written there and for S = 4, 5 with the same rule
(1/S)(1/choose(S, k)); and against the property that the weights sum to
one (equal averages c give gen-sin = all = c).

gsl_rng_uniform_int and gsl_ran_shuffle (GSL is not in this image; call
sites gene.cpp:410,523,639) are checked by properties a plausible mistake
fails: the bucket width is floor((2^32 - 1) / n), not 2^32 / n (at n = 2^31
about half of the raw draws are rejected and the accepted ones are the raw
values); k is monotone in the raw draw with equal-width buckets and no
modulo reduction; the shuffle runs i = n-1 down to 1 (the first draw
decides the last position) and is Fisher-Yates, not Sattolo (a 3-element
shuffle gives the identity one time in six, every permutation equally
often).  What stays unpinned is GSL's exact stream, which only a machine
with GSL can diff.
"""
import itertools
import math

import numpy as np
import pytest

from eqtlbma_amd import synth

GL25 = synth.make_grid("general")
GS10 = synth.make_grid("small")


def _r_log10_weighted_sum(values, weights):
    """functional_tests.R:1343-1349 (no NaN handling: NA configurations are
    dropped by its callers)."""
    values = np.asarray(values, dtype=np.float64)
    mx = values.max()
    return mx + math.log10(np.sum(np.asarray(weights) * 10 ** (values - mx)))


def _r_gen_sin(w_cfg, gen, S):
    """calcAvgAbfsSinAndGenSin (functional_tests.R:1351-1366) with S for 3."""
    v, wt = [], []
    for s in range(S):
        if not np.isnan(w_cfg[s]):
            v.append(w_cfg[s])
            wt.append(0.5 * (1.0 / math.comb(S, 1)))
    v.append(gen)
    wt.append(0.5)
    return _r_log10_weighted_sum(v, wt)


def _r_all(w_cfg, S):
    """calcAvgAbfAll (functional_tests.R:1368-1381) with S for 3."""
    v, wt = [], []
    c = 0
    for k in range(1, S + 1):
        for _ in itertools.combinations(range(S), k):
            if not np.isnan(w_cfg[c]):
                v.append(w_cfg[c])
                wt.append((1.0 / S) * (1.0 / math.comb(S, k)))
            c += 1
    return _r_log10_weighted_sum(v, wt)


@pytest.mark.parametrize("S", [3, 4, 5])
def test_bma_weights_match_the_r_statement(orc, S):
    hc = synth.plain(seed=100 + S, S=S, N=60, n_genes=3, snps_per_gene=12, frac_egenes=0.67)
    out = orc.assoc_uvlr(hc, "join", "all", GL25, GS10)
    nc = 2 ** S - 1
    for p in range(hc.n_pairs):
        w = out.w[:, p]
        w_cfg = w[3:3 + nc]
        assert np.all(np.isfinite(w_cfg)) and np.isfinite(w[0])
        assert w[3 + nc] == pytest.approx(_r_gen_sin(w_cfg, w[0], S), abs=1e-12)
        assert w[3 + nc + 1] == pytest.approx(_r_all(w_cfg, S), abs=1e-12)


@pytest.mark.parametrize("S", [1, 2, 3, 6, 9])
@pytest.mark.parametrize("c", [-3.5, 0.0, 0.7, 12.25])
def test_bma_weights_sum_to_one(orc, S, c):
    nc = 2 ** S - 1
    gs, al = orc.bma(S, "all", np.full(nc, c), c)
    assert gs == pytest.approx(c, abs=1e-13) and al == pytest.approx(c, abs=1e-13)
    gs, al = orc.bma(S, "sin", np.full(S, c), c)
    assert gs == pytest.approx(c, abs=1e-13) and np.isnan(al)


def test_bma_single_dominant_term(orc):
    # one configuration far above the rest: 'all' tends to its log10 weight
    # plus its value; BMAlite to log10(1/(2S)) plus the singleton's value
    S = 4
    nc = 2 ** S - 1
    w = np.full(nc, -50.0)
    w[6] = 40.0  # the 7th configuration, "2-3" (size 2)
    gs, al = orc.bma(S, "all", w, -50.0)
    assert al == pytest.approx(40.0 + math.log10((1.0 / S) / math.comb(S, 2)), abs=1e-12)
    w2 = np.full(nc, -50.0)
    w2[1] = 30.0  # singleton "2"
    gs, _ = orc.bma(S, "sin", w2[:S], -50.0)
    assert gs == pytest.approx(30.0 + math.log10(1.0 / (2 * S)), abs=1e-12)


# ----------------------------------------------------------- GSL RNG ---------
def _raw(seed, n):
    return np.random.RandomState(seed)._bit_generator.random_raw(n).astype(np.uint64)


@pytest.mark.parametrize("n", [2 ** 31, 2 ** 31 + 3])
def test_uniform_int_bucket_width_at_2_pow_31(orc, n):
    # floor((2^32 - 1) / n) = 1: k is the raw draw itself and every draw at
    # or above n is rejected, about half of them (a width of 2^32 / n = 2
    # would halve the draw and reject none)
    seed = 1234
    raw = _raw(seed, 4000)
    mt = orc.MT(seed)
    ks = [mt.uniform_int(n) for _ in range(1000)]
    accepted = [int(r) for r in raw if r < n][:1000]
    assert ks == accepted
    used = int(np.searchsorted(np.cumsum(raw < n), 1000)) + 1
    rejected = used - 1000
    assert 0.4 < rejected / used < 0.6


@pytest.mark.parametrize("n", [2, 3, 7, 300, 65537])
def test_uniform_int_monotone_equal_buckets(orc, n):
    # with a rejection probability below 2^-16 each call consumes one raw
    # draw: k must be a non-decreasing function of the raw value whose
    # buckets all have the same width (floor((2^32 - 1) / n)); a modulo
    # reduction would not be monotone
    seed = 99
    m = 20000
    raw = _raw(seed, m)
    mt = orc.MT(seed)
    ks = np.array([mt.uniform_int(n) for _ in range(m)], dtype=np.int64)
    width = (2 ** 32 - 1) // n
    assert np.all(raw < width * n)  # no rejection happened in this stream
    order = np.argsort(raw, kind="stable")
    assert np.all(np.diff(ks[order]) >= 0)
    # bucket edges: the smallest raw value seen in bucket k is >= k * width,
    # the largest < (k + 1) * width
    for k in np.unique(ks):
        r = raw[ks == k]
        assert r.min() >= k * width and r.max() < (k + 1) * width


def test_uniform_int_rejects_only_past_the_last_bucket(orc):
    # n = 2^32 - 1: width 1, so k is the raw draw and only the raw value
    # 2^32 - 1 is rejected (probability 2^-32): every call returns the next
    # raw draw unchanged
    mt = orc.MT(5)
    raw = _raw(5, 50)
    assert [mt.uniform_int(2 ** 32 - 1) for _ in range(50)] == [int(r) for r in raw]


def test_shuffle_is_backward_fisher_yates(orc):
    # the first draw, uniform_int(n), picks the element that ends in the
    # last position (i = n-1 first)
    for seed in range(1, 40):
        n = 10
        a = orc.MT(seed).shuffle(np.arange(n))
        assert a[n - 1] == orc.MT(seed).uniform_int(n)


def test_shuffle_three_elements_uniform(orc):
    mt = orc.MT(2024)
    counts = {}
    T = 60000
    for _ in range(T):
        b = tuple(mt.shuffle(np.arange(3)))
        counts[b] = counts.get(b, 0) + 1
    assert len(counts) == 6
    # the identity about one time in six (Sattolo's algorithm never gives it)
    assert abs(counts[(0, 1, 2)] / T - 1 / 6) < 0.01
    chi2 = sum((v - T / 6) ** 2 / (T / 6) for v in counts.values())
    assert chi2 < 25  # 5 degrees of freedom, p ~ 1e-4


def test_flat_is_affine_in_the_raw_draw(orc):
    # gsl_ran_flat(a, b) = a (1 - u) + b u with u = raw / 2^32
    raw = _raw(77, 100)
    mt = orc.MT(77)
    for r in raw:
        v = mt.flat(2.0, 5.0)
        u = float(r) / 2 ** 32
        assert v == pytest.approx(2.0 * (1 - u) + 5.0 * u, rel=1e-15)
        assert 2.0 <= v < 5.0
