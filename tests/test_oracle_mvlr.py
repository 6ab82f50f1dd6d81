"""Pins the oracle's MVLR restatement (oracle/eqbf_oracle_mvlr.cpp).

* fiterr = 0 (Sigma = Sigma0, symmetric): the MVLR log10 ABF must equal the
  log10 ratio of two multivariate normal densities of the OLS estimates,
  N(bhat; 0, V + W) / N(bhat; 0, V) with V = Sigma0 / k -- the textbook form
  of the approximate Bayes factor -- evaluated with scipy's
  multivariate_normal on estimates from numpy least squares.
* fiterr > 0: Sigma is built independently from numpy least-squares
  residuals and scipy's F / chi-square distributions (the oracle uses its
  own t tail and normal quantile), then the ABF from numpy slogdet / solve,
  with det(I + k S^-1 W) = det(S + k W) / det(S).
* invariances: a shift of the genotypes (removed by the intercept) changes
  nothing for any fiterr; a scale of the phenotypes changes nothing when
  fiterr = 0.
"""
import numpy as np
import pytest
from scipy import stats

from eqtlbma_amd import synth
from eqtlbma_amd._abi import config_masks

LN10 = np.log(10.0)


def _data(seed, s=4, n=60, q=2, effect=0.4):
    rng = np.random.default_rng(seed)
    x = rng.binomial(2, 0.3, size=n).astype(float)
    Xc = rng.normal(size=(q, n))
    R = np.full((s, s), 0.4) + 0.6 * np.eye(s)
    Y = np.linalg.cholesky(R) @ rng.normal(size=(s, n))
    Y += effect * rng.normal(size=(s, 1)) * x[None, :] + 0.3 * Xc[:1]
    return Y, x, Xc


def _ols(Y, x, Xc):
    """b_hat, k = x~'x~, Sigma0 and the residual matrices of the two models."""
    n = Y.shape[1]
    X0 = np.column_stack([np.ones(n)] + list(Xc))
    X1 = np.column_stack([X0, x])
    c1, *_ = np.linalg.lstsq(X1, Y.T, rcond=None)
    c0, *_ = np.linalg.lstsq(X0, Y.T, rcond=None)
    E1 = Y.T - X1 @ c1
    E0 = Y.T - X0 @ c0
    xr = x - X0 @ np.linalg.lstsq(X0, x, rcond=None)[0]
    k = xr @ xr
    return c1[-1], k, E0, E1


def _mvn_l10abf(bhat, V, W):
    a = stats.multivariate_normal(mean=np.zeros(len(bhat)), cov=V + W).logpdf(bhat)
    b = stats.multivariate_normal(mean=np.zeros(len(bhat)), cov=V).logpdf(bhat)
    return (a - b) / LN10


def _W(t, phi2, oma2):
    return oma2 * np.outer(t, t) + phi2 * np.diag(t * t)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fiterr0_equals_normal_density_ratio(orc, seed):
    Y, x, Xc = _data(seed)
    s, n = Y.shape
    bhat, k, E0, _ = _ols(Y, x, Xc)
    Sig0 = E0.T @ E0 / n
    S1, S0 = orc.mvlr_sigma(Y, x, Xc, 0.0, np.ones(s))
    assert np.allclose(S0, Sig0, rtol=1e-12, atol=1e-14)
    grid = synth.make_grid("general")
    for gamma in (np.ones(s, int), np.eye(s, dtype=int)[1], np.array([1, 0, 1, 1])):
        got = orc.mvlr_abfs(Y, x, Xc, 0.0, gamma, grid)
        t = gamma * np.sqrt(np.diag(Sig0))
        want = np.array([_mvn_l10abf(bhat, Sig0 / k, _W(t, p, o)) for p, o in grid])
        assert np.allclose(got, want, rtol=1e-10, atol=1e-10)


def _sigma_fit(Y, x, Xc, fiterr, gamma):
    """Mixed ML error covariance of MVLR.cpp:201-298, 323-437 for one SNP, built
    from least-squares residuals and scipy distributions."""
    s, n = Y.shape
    q = Xc.shape[0] + 1
    m = q + s + 1
    X0 = np.column_stack([np.ones(n)] + list(Xc))
    X1 = np.column_stack([X0, x])
    E = np.zeros((n, s))
    fac = np.ones(s)
    xr = x - X0 @ np.linalg.lstsq(X0, x, rcond=None)[0]
    for i in range(s):
        X = X1 if gamma[i] else X0
        coef, *_ = np.linalg.lstsq(X, Y[i], rcond=None)
        E[:, i] = Y[i] - X @ coef
        if gamma[i]:
            sigma1 = E[:, i] @ E[:, i] / (n - q - 1)
            T2 = coef[-1] * (xr @ xr) * coef[-1] / sigma1 ** 2
            F = T2  # (v2 - v1 + 1) T2 / (v1 v2) with v1 = 1
            qv = stats.f.sf(F, 1, n - q - 1)
            newF = stats.chi2.isf(qv, 1)
            fac[i] = 1.0 if F < 1e-8 else F / newF
    Sig = (1e-4 * np.eye(s) * m / (m + n) + E.T @ E / (m + n)) * np.sqrt(fac)[None, :]
    E0 = Y.T - X0 @ np.linalg.lstsq(X0, Y.T, rcond=None)[0]
    Sig0 = E0.T @ E0 / n
    return fiterr * Sig + (1 - fiterr) * Sig0


@pytest.mark.parametrize("fiterr", [0.5, 1.0])
@pytest.mark.parametrize("seed", [4, 5])
def test_fitted_sigma_and_abfs(orc, seed, fiterr):
    Y, x, Xc = _data(seed, effect=0.6)
    s, n = Y.shape
    grid = synth.make_grid("small")
    _, k, _, _ = _ols(Y, x, Xc)
    X0 = np.column_stack([np.ones(n)] + list(Xc))
    xr = x - X0 @ np.linalg.lstsq(X0, x, rcond=None)[0]
    Yr = Y - (X0 @ np.linalg.lstsq(X0, Y.T, rcond=None)[0]).T
    b = Yr @ xr
    for gamma in (np.ones(s, int), np.array([0, 1, 0, 0]), np.array([1, 1, 0, 1])):
        Sig = _sigma_fit(Y, x, Xc, fiterr, gamma)
        S1, _ = orc.mvlr_sigma(Y, x, Xc, fiterr, gamma)
        assert np.allclose(S1, Sig, rtol=1e-9, atol=1e-12)
        t = gamma * np.sqrt(np.diag(Sig))
        want = []
        for p, o in grid:
            W = _W(t, p, o)
            # det(I + k S^-1 W) = det(S + k W) / det(S); M^-1 eVb = (S + k W)^-1 b
            ld = np.linalg.slogdet(Sig + k * W)[1] - np.linalg.slogdet(Sig)[1]
            eVb = np.linalg.solve(Sig, b)
            z = np.linalg.solve(Sig + k * W, b)
            want.append((0.5 * eVb @ W @ z - 0.5 * ld) / LN10)
        got = orc.mvlr_abfs(Y, x, Xc, fiterr, gamma, grid)
        assert np.allclose(got, want, rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("fiterr", [0.0, 0.5])
def test_invariances(orc, fiterr):
    Y, x, Xc = _data(7)
    s = Y.shape[0]
    grid = synth.make_grid("general")
    g = np.ones(s, int)
    base = orc.mvlr_abfs(Y, x, Xc, fiterr, g, grid)
    assert np.allclose(orc.mvlr_abfs(Y, x + 1.7, Xc, fiterr, g, grid), base, rtol=1e-9, atol=1e-9)
    scaled = orc.mvlr_abfs(3.0 * Y, x, Xc, fiterr, g, grid)
    if fiterr == 0.0:
        assert np.allclose(scaled, base, rtol=1e-9, atol=1e-9)


def test_chunk_driver_matches_per_pair_calls(orc):
    hc = synth.mvlr(seed=11, S=3, N=50, n_genes=2, snps_per_gene=5, Q=1)
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    out = orc.assoc_mvlr(hc, "all", 0.5, gL, gS)
    masks = list(config_masks("all", hc.S))
    for g in range(hc.G):
        Y = hc.explevels[g]
        for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
            x = hc.genotypes[0][:, hc.pair_snp[p]]
            gen = orc.mvlr_abfs(Y, x, hc.covars[0, :1], 0.5, np.ones(hc.S, int), gL)
            assert np.allclose(out.raw[:len(gL), p], gen, rtol=1e-12, atol=1e-12)
            assert out.w[0, p] == pytest.approx(orc.log10_weighted_sum(gen), rel=1e-12, abs=1e-12)
            for ci, mk in enumerate(masks):
                gam = np.array([(mk >> s) & 1 for s in range(hc.S)])
                l = orc.mvlr_abfs(Y, x, hc.covars[0, :1], 0.5, gam, gS)
                r0 = 3 * len(gL) + ci * len(gS)
                assert np.allclose(out.raw[r0:r0 + len(gS), p], l, rtol=1e-12, atol=1e-12)
            assert (out.n[:, p] == hc.N).all()
