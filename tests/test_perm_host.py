"""Host-side permutation pieces (no GPU): the library's GSL-semantics RNG and
counting loop against the oracle and numpy, and the oracle's permutation
semantics pinned by exhaustive enumeration on a tiny gene.

The MT19937 stream is pinned by numpy's legacy seeding (init_genrand, the
same as gsl_rng_set); gsl_rng_uniform_int / gsl_ran_shuffle are only
checked for agreement with the oracle and for their invariants -- their
parity with GSL itself is unpinned (GSL is absent here).
"""
import itertools
import math

import numpy as np
import pytest

from eqtlbma_amd import engine, synth


@pytest.mark.parametrize("seed", [0, 7, 42, 2 ** 32 + 3])
def test_library_rng_matches_oracle_and_numpy(orc, seed):
    r = engine.Rng(seed)
    o = orc.MT(seed)
    a = [r.get() for _ in range(3000)]
    b = [o.next() for _ in range(3000)]
    assert a == b
    ref_seed = 4357 if seed == 0 else seed & 0xffffffff
    ref = np.random.RandomState(ref_seed)._bit_generator.random_raw(3000)
    assert a == list(ref)


def test_library_shuffle_and_tables_match_oracle(orc):
    r = engine.Rng(7)
    t = r.tables(300, 3, 20)
    ref = orc.perm_tables(7, 300, 3, 20)
    assert np.array_equal(t, ref)
    for row in t.reshape(-1, 300):
        assert np.array_equal(np.sort(row), np.arange(300))
    r1, o1 = engine.Rng(11), orc.MT(11)
    assert [r1.uniform_int(n) for n in range(1, 200)] == [o1.uniform_int(n) for n in range(1, 200)]
    assert r1.flat(0.1, 0.2) == o1.flat(0.1, 0.2)


@pytest.mark.parametrize("trick", [0, 1, 2])
@pytest.mark.parametrize("sep", [False, True])
def test_counting_loop_matches_oracle(orc, trick, sep):
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = 200
        stats = rng.normal(size=n)
        stats[rng.random(n) < 0.05] = np.nan
        true = rng.normal(0, 1.5)
        rt1 = engine.Rng(9) if trick else None
        rt2 = orc.MT(9) if trick else None
        a = engine.perm_pvalue(stats, true, n, sep=sep, trick=trick, trick_cutoff=5, rng_trick=rt1)
        b = orc.perm_count(stats, true, n, trick=trick, trick_cutoff=5, rng_trick=rt2, sep=sep)
        assert a[0] == b[0] or (math.isnan(a[0]) and math.isnan(b[0]))
        assert a[1:4] == b[1:4] if not sep else (a[1], a[3], a[4]) == (b[1], b[3], b[4])


def test_counting_loop_known_answers():
    # no trick, no NaN: p = (1 + #{stat >= true}) / (n + 1)
    stats = np.array([0.5, 2.0, 1.0, 3.0, -1.0])
    pv, used, med, nb, drawn = engine.perm_pvalue(stats, 1.0, 5)
    assert pv == pytest.approx(4 / 6) and used == 5 and med == 1.0 and nb == 5 and drawn == 5
    # a NaN statistic decrements the shared counter and is not counted
    stats = np.array([np.nan, 2.0, 1.0, 3.0, -1.0])
    pv, used, med, nb, drawn = engine.perm_pvalue(stats, 1.0, 5)
    assert (used, nb) == (4, 4) and pv == pytest.approx(4 / 5)
    # separate analysis counts p-values <= the true one
    pv, used, _, nb, _ = engine.perm_pvalue(np.array([0.01, 0.2, 0.05]), 0.05, 3, sep=True)
    assert pv == pytest.approx(3 / 4)


def _tiny_gene(seed=5, N=6, S=2, P=3):
    rng = np.random.default_rng(seed)
    g = rng.binomial(2, 0.4, size=(P, N)).astype(float)
    for j in range(P):
        while g[j].std() == 0:
            g[j] = rng.binomial(2, 0.4, size=N)
    y = rng.normal(size=(1, S, N))
    y[0, 0] += 0.8 * g[0]
    from eqtlbma_amd._abi import HostChunk
    return HostChunk(explevels=y, gene_in_sbgrp=np.ones((1, S), np.uint8), geno_in_sbgrp=np.ones((S, N), np.uint8),
                     genotypes=np.ascontiguousarray(g.T), snp_in_sbgrp=np.ones((P, S), np.uint8),
                     pair_offset=np.array([0, P]), pair_snp=np.arange(P, dtype=np.int32))


@pytest.mark.parametrize("pbf,use_max", [("all", True), ("gen", False), ("gen-sin", True)])
def test_permutation_semantics_by_enumeration(orc, pbf, use_max):
    """All N! permutations of a tiny gene: the oracle's permutation statistic
    equals its association test run on explicitly permuted expression levels
    (expression permuted, genotypes fixed), and the exact permutation p-value
    follows."""
    hc = _tiny_gene()
    N = hc.N
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    perms = np.array(list(itertools.permutations(range(N))), dtype=np.int32)
    stats = orc.perm_join_stats(hc, 0, pbf, use_max, gL, gS, perms)
    bfs = {"gen": "gen", "gen-sin": "sin", "all": "all"}[pbf]
    row = {"gen": 0, "gen-sin": 3 + (0 if bfs == "gen" else (hc.S if bfs == "sin" else 2 ** hc.S - 1)),
           "all": 3 + 2 ** hc.S - 1 + 1}[pbf]
    for k in range(0, len(perms), 97):
        yk = hc.explevels[:, :, perms[k]]
        hk = type(hc)(yk, hc.gene_in_sbgrp, hc.geno_in_sbgrp, hc.genotypes, hc.snp_in_sbgrp,
                      hc.pair_offset, hc.pair_snp)
        out = orc.assoc_uvlr(hk, "join", bfs, gL, gS)
        w = out.w[row]
        ref = w.max() if use_max else orc.log10_weighted_sum(w)
        assert stats[k] == pytest.approx(ref, rel=1e-12, abs=1e-12)
    # identity permutation reproduces the true statistic
    true = orc.assoc_uvlr(hc, "join", bfs, gL, gS).w[row]
    t = true.max() if use_max else orc.log10_weighted_sum(true)
    assert stats[0] == pytest.approx(t, rel=1e-12, abs=1e-12)
    p_exact = (1 + np.sum(stats >= t)) / (len(perms) + 1)
    pv = orc.perm_count(stats, t, len(perms))[0]
    assert pv == pytest.approx(p_exact)
