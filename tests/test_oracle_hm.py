"""Pins the CPU oracle of eqtlbma_hm's EM (oracle/eqbf_oracle_hm.cpp).

The reference's own hm test (tests/test_hm.bash) checks one number on data
simulated by its R script, which is not available here, so the oracle is
pinned by what the model fixes:
  - the log10 likelihood it evaluates equals the model's definition written
    directly with numpy (grid- then configuration- or type-weighted BFs,
    averaged over SNPs with weight 1/P, mixed with the null at pi0);
  - the EM's fixed point is the maximum likelihood estimate: a direct
    numerical maximisation of that likelihood (scipy, over unconstrained
    logit / softmax coordinates) reaches the same likelihood and
    parameters, for the configs and the types model;
  - with one configuration and one grid point, pi0 solves the score
    equation sum_g (1 - BF_g) / (pi0 + (1 - pi0) BF_g) = 0 (scipy brentq);
  - the likelihood never decreases over iterations, and a parameter held
    fixed stays at its value.
"""
import numpy as np
import pytest
from scipy import optimize

from oracle import oracle as orc


def _case(seed, G=40, K=4, L=5, S=None, null_frac=0.5):
    rng = np.random.default_rng(seed)
    P = rng.integers(2, 7, G)
    off = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    D = S if S is not None else K
    raw = rng.normal(-0.6, 0.4, (off[-1], D, L))
    for g in range(G):
        if rng.random() > null_frac:  # eGene: one SNP with large BFs in some configurations / grid points
            p = off[g] + rng.integers(P[g])
            raw[p] += rng.uniform(1.0, 3.0, (D, L)) * (rng.random((D, L)) < 0.6)
    return raw, off


def _lse10(x, w, axis):
    """log10 sum w 10^x along axis (weights broadcast)."""
    m = np.max(x, axis=axis, keepdims=True)
    return np.squeeze(m, axis) + np.log10(np.sum(w * 10.0 ** (x - m), axis=axis))


def _loglik_np(raw, off, pi0, grid, cfg=None, typ=None, sub=None):
    if typ is None:
        per_cfg = _lse10(raw, grid[None, None, :], 2)                 # [p][k]
        snp = _lse10(per_cfg, cfg[None, :], 1)                        # [p]
    else:
        q = sub[None, :, :, None]                                     # [1][K][S][1]
        r = raw[:, None, :, :]                                        # [p][1][S][L]
        t = np.sum(np.log10(q * 10.0 ** r + 1 - q), axis=2)           # [p][K][L]
        per_type = _lse10(t, grid[None, None, :], 2)
        snp = _lse10(per_type, typ[None, :], 1)
    tot = 0.0
    for g in range(off.size - 1):
        s = snp[off[g]:off[g + 1]]
        gb = _lse10(s, np.full(s.size, 1.0 / s.size), 0)
        tot += np.log10(pi0 + (1 - pi0) * 10.0 ** gb)
    return tot


def _softmax(z):
    e = np.exp(z - z.max())
    return e / e.sum()


def test_loglik_matches_definition():
    raw, off = _case(1)
    rng = np.random.default_rng(2)
    for _ in range(3):
        prm = orc.HmParams(rng.uniform(0.1, 0.9), _softmax(rng.normal(size=5)), cfg=_softmax(rng.normal(size=4)))
        a = orc.hm_loglik(raw, off, prm)
        b = _loglik_np(raw, off, prm.pi0, prm.grid, cfg=prm.cfg)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-10)
    raw, off = _case(3, K=2, L=3, S=3)
    prm = orc.HmParams(0.4, _softmax(rng.normal(size=3)), type_=_softmax(rng.normal(size=2)),
                       sub=rng.uniform(0.1, 0.9, (2, 3)))
    a = orc.hm_loglik(raw, off, prm)
    b = _loglik_np(raw, off, prm.pi0, prm.grid, typ=prm.type, sub=prm.sub)
    assert a == pytest.approx(b, rel=1e-12, abs=1e-10)


def test_em_configs_reaches_the_mle():
    raw, off = _case(5, G=60, K=3, L=4)
    K, L = 3, 4
    prm = orc.HmParams(0.5, np.full(L, 1.0 / L), cfg=np.full(K, 1.0 / K))
    em, it, trace = orc.hm_em(raw, off, prm, thresh=1e-11)
    assert np.all(np.diff(trace) >= -1e-12) and it > 3

    def negll(z):
        return -_loglik_np(raw, off, 1 / (1 + np.exp(-z[0])), _softmax(z[1:1 + L]), cfg=_softmax(z[1 + L:]))
    best = None
    for s in range(3):
        z0 = np.random.default_rng(s).normal(0, 0.5, 1 + L + K)
        r = optimize.minimize(negll, z0, method="BFGS", options={"gtol": 1e-9, "maxiter": 5000})
        if best is None or r.fun < best.fun:
            best = r
    ll_em = orc.hm_loglik(raw, off, em)
    assert ll_em >= -best.fun - 1e-6
    assert ll_em == pytest.approx(-best.fun, abs=1e-5)
    pi0_opt = 1 / (1 + np.exp(-best.x[0]))
    assert em.pi0 == pytest.approx(pi0_opt, abs=2e-3)
    assert np.allclose(em.grid, _softmax(best.x[1:1 + L]), atol=5e-3)
    assert np.allclose(em.cfg, _softmax(best.x[1 + L:]), atol=5e-3)


def test_em_types_reaches_the_mle():
    S, K, L = 3, 2, 2
    raw, off = _case(7, G=50, L=L, S=S)
    rng = np.random.default_rng(8)
    prm = orc.HmParams(0.5, np.full(L, 1.0 / L), type_=np.full(K, 1.0 / K), sub=rng.uniform(0.3, 0.7, (K, S)))
    em, it, trace = orc.hm_em(raw, off, prm, thresh=1e-11)
    assert np.all(np.diff(trace) >= -1e-12)

    def unpack(z):
        return (1 / (1 + np.exp(-z[0])), _softmax(z[1:1 + L]), _softmax(z[1 + L:1 + L + K]),
                (1 / (1 + np.exp(-z[1 + L + K:]))).reshape(K, S))

    def negll(z):
        pi0, grid, typ, sub = unpack(z)
        return -_loglik_np(raw, off, pi0, grid, typ=typ, sub=sub)
    z0 = np.concatenate([[0.0], np.zeros(L), np.zeros(K), np.log(prm.sub / (1 - prm.sub)).ravel()])
    r = optimize.minimize(negll, z0, method="BFGS", options={"gtol": 1e-9, "maxiter": 10000})
    ll_em = orc.hm_loglik(raw, off, em)
    # both are local maxima of a mixture likelihood started from the same point
    assert ll_em == pytest.approx(-r.fun, abs=1e-4)


def test_pi0_solves_the_score_equation():
    raw, off = _case(11, G=80, K=1, L=1)
    prm = orc.HmParams(0.5, np.ones(1), cfg=np.ones(1))
    em, it, _ = orc.hm_em(raw, off, prm, thresh=1e-13, maxit=0)
    bf = np.array([np.mean(10.0 ** raw[off[g]:off[g + 1], 0, 0]) for g in range(off.size - 1)])
    root = optimize.brentq(lambda p: np.sum((1 - bf) / (p + (1 - p) * bf)), 1e-9, 1 - 1e-9, xtol=1e-14)
    assert em.pi0 == pytest.approx(root, abs=1e-6)
    assert em.grid[0] == 1.0 and em.cfg[0] == 1.0


def test_fixed_parameters_stay():
    raw, off = _case(13, K=3, L=3)
    prm = orc.HmParams(0.3, np.array([0.2, 0.3, 0.5]), cfg=np.array([0.6, 0.3, 0.1]))
    em, _, trace = orc.hm_em(raw, off, prm, fixed=(1, 0, 1, 0, 0))
    assert em.pi0 == 0.3 and np.array_equal(em.cfg, prm.cfg)
    assert not np.array_equal(em.grid, prm.grid)
    assert np.all(np.diff(trace) >= -1e-12)


def test_bfs_outputs_follow_from_the_likelihood():
    raw, off = _case(17, K=3, L=4)
    prm = orc.HmParams(0.4, _softmax(np.arange(4.0)), cfg=_softmax(np.arange(3.0)))
    gp, gb, sb, db = orc.hm_bfs(raw, off, prm)
    assert np.allclose(db, _lse10(raw, prm.grid[None, None, :], 2), rtol=1e-13, atol=1e-12)
    assert np.allclose(sb, _lse10(db, prm.cfg[None, :], 1), rtol=1e-13, atol=1e-12)
    for g in range(off.size - 1):
        s = sb[off[g]:off[g + 1]]
        assert gb[g] == pytest.approx(_lse10(s, np.full(s.size, 1.0 / s.size), 0), rel=1e-13, abs=1e-12)
        ol = np.log10(prm.pi0 + (1 - prm.pi0) * 10.0 ** gb[g])
        assert gp[g] == pytest.approx(min(1.0, (1 - prm.pi0) * 10.0 ** (gb[g] - ol)), rel=1e-12)


def test_squarem_reaches_the_classic_optimum_in_fewer_iterations():
    """SQUAREM (any --msl other than 1): its first two iterations are plain
    EM iterations, and it ends at the same maximum as classic EM; the
    extrapolation shows as fewer iterations on a slowly converging case."""
    raw, off = _case(19, G=80, K=4, L=6, null_frac=0.7)
    prm = orc.HmParams(0.5, np.full(6, 1 / 6), cfg=np.full(4, 0.25))
    cl, it_c, tr_c = orc.hm_em(raw, off, prm, thresh=1e-9)
    sq, it_s, tr_s = orc.hm_em(raw, off, prm, thresh=1e-9, stepmax=4.0)
    assert tr_s[1] == tr_c[1] and tr_s[2] == tr_c[2]
    assert it_s < it_c
    assert orc.hm_loglik(raw, off, sq) == pytest.approx(orc.hm_loglik(raw, off, cl), abs=1e-6)
    assert sq.pi0 == pytest.approx(cl.pi0, abs=1e-3)
    assert np.allclose(sq.grid, cl.grid, atol=2e-3) and np.allclose(sq.cfg, cl.cfg, atol=2e-3)


def test_nan_likelihood_is_an_error():
    """A first gene whose every log10 ABF is NaN makes each unnormalised
    configuration weight NaN, where the reference stops with an error."""
    raw, off = _case(23, K=3, L=3)
    raw[off[0]:off[1]] = np.nan
    prm = orc.HmParams(0.5, np.full(3, 1 / 3), cfg=np.full(3, 1 / 3))
    with pytest.raises(RuntimeError, match="NaN"):
        orc.hm_em(raw, off, prm)


def test_profile_ci_bounds_are_where_the_likelihood_drops_by_two():
    """--getci: each bound is the last 0.001 step at which the natural-log
    likelihood (one weight moved, the other weights of its simplex rescaled
    in proportion) stays within 2 of the maximum; the next step leaves it."""
    raw, off = _case(29, G=120, K=3, L=3)
    prm = orc.HmParams(0.5, np.full(3, 1 / 3), cfg=np.full(3, 1 / 3))
    em, _, _ = orc.hm_em(raw, off, prm, thresh=1e-8)
    left, right = orc.hm_ci(raw, off, em)
    ln = np.log10(np.e)
    mx = orc.hm_loglik(raw, off, em)

    def inside(q):
        return orc.hm_loglik(raw, off, q) / ln >= mx / ln - 2.0

    def moved(name, i, x):
        q = em.copy()
        if name == "pi0":
            q.pi0 = x
            return q
        w = getattr(q, name)
        mle = getattr(em, name)
        for j in range(w.size):
            w[j] = x if j == i else mle[j] + (mle[i] - x) * mle[j] / (1 - mle[i])
        return q

    items = [("pi0", 0, 0)] + [("cfg", i, 1 + i) for i in range(3)] + [("grid", i, 4 + i) for i in range(3)]
    n_interior = 0
    for name, i, j in items:
        est = em.pi0 if name == "pi0" else getattr(em, name)[i]
        assert left[j] <= est + 1e-12 <= right[j] + 2e-12
        for b, step in [(left[j], -1e-3), (right[j], 1e-3)]:
            if 1e-9 < b < 1 - 1e-9:
                assert inside(moved(name, i, b))
                assert not inside(moved(name, i, b + step))
                n_interior += 1
    assert n_interior >= 4
    # the types model keeps its type and subgroup estimates as both bounds
    raw, off = _case(31, G=60, L=2, S=3)
    prm = orc.HmParams(0.5, np.full(2, 0.5), type_=np.full(2, 0.5), sub=np.full((2, 3), 0.4))
    left, right = orc.hm_ci(raw, off, prm)
    assert np.array_equal(left[1:9], np.concatenate([prm.type, prm.sub.ravel()]))
    assert np.array_equal(right[1:9], left[1:9])
