"""GPU: eqtlbma_hm's EM (eqbf_hm_em / eqbf_hm_bfs) against the oracle.

Same raw log10 ABFs, same starting point: the device EM must take the same
number of iterations and end at the same parameters and log10 likelihood as
the oracle's restatement of run_EM_classic / run_EM_square.  The device sums
over SNPs, configurations and genes in a different order (wave and block
reductions), so values agree to rounding: parameters within 1e-9, log10
likelihoods within 1e-8 absolute, per-pair and per-gene log10 BFs within
1e-10.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from eqtlbma_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _case(seed, G, K, L, S=None, null_frac=0.5, pmin=2, pmax=40, nan_frac=0.0):
    rng = np.random.default_rng(seed)
    P = rng.integers(pmin, pmax + 1, G)
    off = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    D = S if S is not None else K
    raw = rng.normal(-0.6, 0.4, (off[-1], D, L))
    for g in range(G):
        if rng.random() > null_frac:
            p = off[g] + rng.integers(P[g])
            raw[p] += rng.uniform(1.0, 3.0, (D, L)) * (rng.random((D, L)) < 0.6)
    if nan_frac:
        raw[rng.random(raw.shape) < nan_frac] = np.nan
    return raw, off


def _em_both(eng, orc, raw, off, prm, **kw):
    ref, it_ref, _ = orc.hm_em(raw, off, prm, **kw)
    types = prm.type is not None
    pi0, grid, dimw, sub, lik, it = eng.hm_em(raw, off, prm.pi0, prm.grid, cfg=prm.cfg, type_=prm.type,
                                              sub=prm.sub, **kw)
    lik_ref = orc.hm_loglik(raw, off, ref)
    _say("[gpu]   iterations %d (oracle %d)  log10 lik %.10f (oracle %.10f)" % (it, it_ref, lik, lik_ref))
    assert it == it_ref
    assert lik == pytest.approx(lik_ref, abs=1e-8)
    assert pi0 == pytest.approx(ref.pi0, abs=1e-9)
    assert np.allclose(grid, ref.grid, rtol=0, atol=1e-9)
    assert np.allclose(dimw, ref.type if types else ref.cfg, rtol=0, atol=1e-9)
    if types:
        assert np.allclose(sub, ref.sub, rtol=0, atol=1e-9)
    return ref


@pytest.mark.parametrize("G,thresh", [(1500, 0.05), (300, 1e-4)])
def test_hm_configs_classic_vs_oracle(eng, orc, G, thresh):
    raw, off = _case(31, G=G, K=7, L=10)
    _say("[gpu] hm configs classic thresh=%g pairs=%d" % (thresh, raw.shape[0]))
    prm = orc.HmParams(0.5, np.full(10, 0.1), cfg=np.full(7, 1 / 7))
    _em_both(eng, orc, raw, off, prm, thresh=thresh)


def test_hm_configs_nan_abfs_vs_oracle(eng, orc):
    """NaN ABFs are skipped by every log10 sum.  The first term of each sum
    (first grid point, first configuration) stays defined: a NaN there makes
    the whole sum NaN, the updates then drop terms the likelihood keeps, and
    the reference's EM can see its likelihood decrease and stop."""
    raw, off = _case(37, G=400, K=3, L=4, nan_frac=0.03)
    raw[:, :, 0] = np.where(np.isnan(raw[:, :, 0]), -0.5, raw[:, :, 0])
    raw[:, 0, :] = np.where(np.isnan(raw[:, 0, :]), -0.5, raw[:, 0, :])
    _say("[gpu] hm configs NaN ABFs pairs=%d" % raw.shape[0])
    prm = orc.HmParams(0.4, np.full(4, 0.25), cfg=np.array([0.5, 0.3, 0.2]))
    _em_both(eng, orc, raw, off, prm, thresh=1e-5)


def test_hm_types_vs_oracle(eng, orc):
    S, K, L = 4, 3, 6
    raw, off = _case(41, G=300, K=K, L=L, S=S)
    rng = np.random.default_rng(42)
    prm = orc.HmParams(0.5, np.full(L, 1 / L), type_=np.full(K, 1 / K), sub=rng.uniform(0.2, 0.8, (K, S)))
    _say("[gpu] hm types pairs=%d" % raw.shape[0])
    _em_both(eng, orc, raw, off, prm, thresh=1e-4)


@pytest.mark.parametrize("model", ["configs", "types"])
def test_hm_squarem_vs_oracle(eng, orc, model):
    if model == "configs":
        raw, off = _case(43, G=800, K=4, L=6, null_frac=0.7)
        prm = orc.HmParams(0.5, np.full(6, 1 / 6), cfg=np.full(4, 0.25))
    else:
        raw, off = _case(47, G=500, K=2, L=5, S=3)
        prm = orc.HmParams(0.5, np.full(5, 0.2), type_=np.full(2, 0.5), sub=np.full((2, 3), 0.5))
    _say("[gpu] hm %s SQUAREM msl=3 pairs=%d" % (model, raw.shape[0]))
    _em_both(eng, orc, raw, off, prm, thresh=1e-6, stepmax=3.0)


def test_hm_fixed_and_maxit_vs_oracle(eng, orc):
    raw, off = _case(53, G=300, K=3, L=3)
    prm = orc.HmParams(0.3, np.array([0.2, 0.3, 0.5]), cfg=np.array([0.6, 0.3, 0.1]))
    _say("[gpu] hm fixed pi0 and configs, maxit 4")
    ref = _em_both(eng, orc, raw, off, prm, thresh=1e-12, maxit=4, fixed=(1, 0, 1, 0, 0))
    assert ref.pi0 == 0.3


def test_hm_nan_likelihood_is_an_error(eng, orc):
    from eqtlbma_amd.engine import EngineError
    raw, off = _case(23, G=40, K=3, L=3)
    raw[off[0]:off[1]] = np.nan
    with pytest.raises(EngineError):
        eng.hm_em(raw, off, 0.5, np.full(3, 1 / 3), cfg=np.full(3, 1 / 3))
    with pytest.raises(RuntimeError):
        orc.hm_em(raw, off, orc.HmParams(0.5, np.full(3, 1 / 3), cfg=np.full(3, 1 / 3)))


@pytest.mark.parametrize("model", ["configs", "types"])
def test_hm_bfs_vs_oracle(eng, orc, model):
    if model == "configs":
        raw, off = _case(59, G=500, K=7, L=10, nan_frac=0.01)
        prm = orc.HmParams(0.6, np.linspace(1, 2, 10) / np.linspace(1, 2, 10).sum(),
                           cfg=np.arange(1, 8.0) / 28)
    else:
        raw, off = _case(61, G=300, K=3, L=5, S=4)
        prm = orc.HmParams(0.6, np.full(5, 0.2), type_=np.array([0.5, 0.3, 0.2]),
                           sub=np.random.default_rng(3).uniform(0.1, 0.9, (3, 4)))
    _say("[gpu] hm bfs %s pairs=%d" % (model, raw.shape[0]))
    got = eng.hm_bfs(raw, off, prm.pi0, prm.grid, cfg=prm.cfg, type_=prm.type, sub=prm.sub)
    ref = orc.hm_bfs(raw, off, prm)
    for a, b, name in zip(got, ref, ["gene posterior", "gene BF", "SNP BF", "per-dim BF"]):
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        fin = np.isfinite(b)
        err = float(np.max(np.abs(a[fin] - b[fin]), initial=0))
        _say("[gpu]   %s max abs err %.2e" % (name, err))
        assert err <= 1e-10, name


@pytest.mark.parametrize("model", ["configs", "types"])
def test_hm_ci_vs_oracle(eng, orc, model):
    """--getci bounds: the same 0.001 steps, so the same values."""
    if model == "configs":
        raw, off = _case(67, G=400, K=3, L=4)
        prm = orc.HmParams(0.5, np.full(4, 0.25), cfg=np.full(3, 1 / 3))
    else:
        raw, off = _case(71, G=200, K=2, L=3, S=3)
        prm = orc.HmParams(0.5, np.full(3, 1 / 3), type_=np.full(2, 0.5), sub=np.full((2, 3), 0.5))
    em, _, _ = orc.hm_em(raw, off, prm, thresh=1e-6)
    _say("[gpu] hm profile CI %s pairs=%d" % (model, raw.shape[0]))
    left, right = eng.hm_ci(raw, off, em.pi0, em.grid, cfg=em.cfg, type_=em.type, sub=em.sub)
    rl, rr = orc.hm_ci(raw, off, em)
    _say("[gpu]   left %s\n[gpu]   right %s" % (np.round(left, 4), np.round(right, 4)))
    assert np.allclose(left, rl, rtol=0, atol=1e-12) and np.allclose(right, rr, rtol=0, atol=1e-12)


def test_hm_many_configurations_vs_oracle(eng, orc):
    """1,023 configurations (10 subgroups) x 6 grid points: one pair's raw
    ABFs exceed the LDS tile, so the per-pair kernel runs instead."""
    raw, off = _case(73, G=30, K=1023, L=6, pmin=2, pmax=6)
    prm = orc.HmParams(0.5, np.full(6, 1 / 6), cfg=np.full(1023, 1 / 1023))
    _say("[gpu] hm 1023 configurations pairs=%d" % raw.shape[0])
    _em_both(eng, orc, raw, off, prm, thresh=1e-3)


def test_hm_wide_ranges_and_tiny_weights_vs_oracle(eng, orc):
    """Pairs whose log10 ABFs span more than ~290 units, and a grid weight of
    1e-200, leave the factored sums (the reference's two-pass sums are kept
    where a product of a weight and 10^(R - max) would underflow)."""
    raw, off = _case(79, G=200, K=7, L=5)
    rng = np.random.default_rng(80)
    wide = rng.choice(raw.shape[0], 40, replace=False)
    raw[wide, 2, 1] += 350.0
    raw[wide[:10], 0, 0] -= 400.0
    prm = orc.HmParams(0.5, np.full(5, 0.2), cfg=np.full(7, 1 / 7))
    _say("[gpu] hm wide ranges pairs=%d" % raw.shape[0])
    _em_both(eng, orc, raw, off, prm, thresh=1e-4)
    prm = orc.HmParams(0.5, np.array([1e-200, 0.25, 0.25, 0.25, 0.25 - 1e-200]), cfg=np.full(7, 1 / 7))
    _em_both(eng, orc, raw, off, prm, thresh=1e-4, fixed=(0, 1, 0, 0, 0))
    got = eng.hm_bfs(raw, off, prm.pi0, prm.grid, cfg=prm.cfg)
    ref = orc.hm_bfs(raw, off, prm)
    for a, b in zip(got, ref):
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin)
        assert np.max(np.abs(a[fin] - b[fin]), initial=0) <= 1e-10
