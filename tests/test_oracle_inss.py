"""The oracle's --inss path (Bayes factors from given summary statistics).

Pinned without retyping its formulas:
- round trip: the summary statistics the data path computes for a chunk
  without covariates, given back as input, reproduce the data path's
  standardised statistics, raw and grid-averaged Bayes factors (the data
  path is pinned on its own in test_oracle_math.py);
- the small-sample correction uses nu = n - 2 whatever the covariates were:
  z against scipy's t and normal distributions;
- a subgroup with n = 0 is the same as no line at all (its Bayes factors are
  those of the other subgroups, its singleton is 0); n <= 2 and NaN inputs.
"""
import numpy as np
import pytest
from scipy import special

from eqtlbma_amd import synth

GL = synth.make_grid("general")
GS = synth.make_grid("small")


def _no_cov_chunk(seed=5, S=3):
    return synth.plain(seed=seed, S=S, N=60, n_genes=4, snps_per_gene=25, frac_egenes=0.5)


@pytest.mark.parametrize("bfs", ["gen", "sin", "all"])
def test_round_trip_through_sumstats(orc, bfs):
    hc = _no_cov_chunk()
    ref = orc.assoc_uvlr(hc, "join", bfs, GL, GS)
    sig, beta, se = ref.sstats[1], ref.sstats[2], ref.sstats[3]
    std, raw, w = orc.abf_from_sstats(ref.n, sig, beta, se, bfs, GL, GS)
    assert np.allclose(std, ref.stdsstats, rtol=1e-13, atol=0, equal_nan=True)
    assert np.allclose(raw, ref.raw, rtol=1e-12, atol=1e-13, equal_nan=True)
    assert np.allclose(w, ref.w, rtol=1e-12, atol=1e-13, equal_nan=True)


def test_nu_is_n_minus_two(orc):
    rng = np.random.default_rng(3)
    S, NP = 2, 200
    n = rng.integers(5, 400, size=(S, NP))
    sig = rng.uniform(0.5, 2.0, size=(S, NP))
    beta = rng.normal(scale=0.4, size=(S, NP))
    se = rng.uniform(0.05, 0.5, size=(S, NP))
    std, _raw, _w = orc.abf_from_sstats(n, sig, beta, se, "sin", GL, GS)
    t = (beta / sig) / (se / sig)
    z = special.ndtri(special.stdtr(n - 2.0, -np.abs(t)))
    assert np.allclose(std[2], z, rtol=1e-10, atol=1e-12)
    # b and se rescaled so that b / se = z and the sign of beta is kept
    assert np.allclose(std[0] / std[1], np.sign(beta) * np.abs(z), rtol=1e-12)
    assert np.all(np.sign(std[0]) == np.sign(beta))


def test_absent_subgroup_equals_missing_line(orc):
    rng = np.random.default_rng(4)
    S, NP = 3, 50
    n = rng.integers(30, 200, size=(S, NP))
    sig = rng.uniform(0.5, 2.0, size=(S, NP))
    beta = rng.normal(scale=0.5, size=(S, NP))
    se = rng.uniform(0.05, 0.5, size=(S, NP))
    n0 = n.copy()
    n0[1] = 0
    _s, raw0, w0 = orc.abf_from_sstats(n0, sig, beta, se, "all", GL, GS)
    # the same pairs with subgroup 2 removed entirely (S = 2)
    keep = [0, 2]
    _s, raw2, w2 = orc.abf_from_sstats(n[keep], sig[keep], beta[keep], se[keep], "all", GL, GS)
    # consistent configuration: identical
    assert np.allclose(w0[:3], w2[:3], rtol=1e-13)
    assert np.allclose(raw0[:3 * len(GL)], raw2[:3 * len(GL)], rtol=1e-13)
    # configurations of S=3 containing subgroup 2 (bit 1) equal those without it
    from eqtlbma_amd._abi import config_masks
    m3, m2 = config_masks("all", 3), config_masks("all", 2)
    for k, m in enumerate(m3):
        rest = [(m >> s) & 1 for s in keep]
        if not any(rest):
            assert np.all(w0[3 + k] == 0.0)  # only the absent subgroup: all zeros over the grid
            continue
        m_2 = rest[0] | (rest[1] << 1)
        assert np.allclose(w0[3 + k], w2[3 + m2.index(m_2)], rtol=1e-13, atol=0)


def test_small_n_and_nan(orc):
    S, NP = 2, 4
    n = np.array([[2, 1, 50, 50], [40, 40, 0, 40]])
    sig = np.ones((S, NP))
    beta = np.array([[0.3, 0.3, np.nan, 0.2], [0.1, 0.2, 0.3, 0.0]])
    se = np.full((S, NP), 0.1)
    std, raw, w = orc.abf_from_sstats(n, sig, beta, se, "sin", GL, GS)
    # n <= 2: NaN z, b = 0, se = inf (the |z| > 1e-8 test fails)
    assert np.isnan(std[2, 0, 0]) and std[0, 0, 0] == 0.0 and np.isinf(std[1, 0, 0])
    assert np.isnan(std[2, 0, 1])
    # NaN beta: the subgroup keeps NaN statistics; a configuration holding it
    # has a NaN variance of b-bar, fails the reference's "< infinity" test and
    # evaluates to 0 (gene_snp_pair.cpp:337-351), as does the absent partner
    assert np.all(np.isnan(std[:, 0, 2]))
    assert np.all(raw[:, 2] == 0.0)
    assert w[0, 2] == 0.0 and w[3, 2] == 0.0 and w[4, 2] == 0.0
    # beta = 0: t = 0 -> z = 0 -> b = 0, se = inf, the subgroup adds 0
    assert std[0, 1, 3] == 0.0 and np.isinf(std[1, 1, 3])
    assert w[4, 3] == 0.0
