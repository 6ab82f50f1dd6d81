"""Pins the oracle's eqtlbma_avg_bfs restatement (oracle/eqbf_oracle_avg.cpp)
against closed forms computed with numpy:

* the grid average is log10 sum_l w_l 10^x_l over the kept points (NaN
  skipped; a NaN first value makes it NaN; |result| <= DBL_EPSILON is 0);
* the nested grid-then-configuration average equals the flat double sum
  log10 sum_j c_j sum_l w_l 10^x_jl, and the gene BF is the log10 mean of
  the SNP BFs;
* posteriors: gene_post = (1 - pi0) BF / (pi0 + (1 - pi0) BF), post_the
  sums to 1 over a gene's SNPs, post_an = (BF/m) / (1 - 1/m + BF/m),
  post_dim sums to 1 over the configurations, best_dim is its argmax, and
  the subgroup posteriors of the configs model are sums over the
  configurations that contain the subgroup;
* the types model against its likelihood sum_k t_k sum_l w_l prod_s
  (q_ks 10^x_sl + 1 - q_ks), with binary q reducing it to the configs model
  whose configuration BFs are products of subgroup BFs, and its subgroup
  posteriors against the same sums restricted to subgroup s being active.
"""
import numpy as np
import pytest

LN10 = np.log(10.0)


def lse10(x, w):
    """log10 sum w 10^x, in numpy (exact reference for the pins)."""
    x = np.asarray(x, float)
    w = np.asarray(w, float)
    m = np.max(x)
    return m + np.log10(np.sum(w * 10.0 ** (x - m)))


def _raw(seed, NP, R, L, scale=1.0):
    rng = np.random.default_rng(seed)
    return rng.normal(0.2, scale, (NP, R, L))


def test_grid_average(orc):
    raw = _raw(1, 40, 1, 6)[:, 0, :]
    gw = np.random.default_rng(2).uniform(0.05, 0.4, 6)
    got = orc.avg_grid(raw, gw)
    want = [lse10(r, gw) for r in raw]
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    keep = np.array([0, 2, 5], np.int32)
    got = orc.avg_grid(raw, gw, keep)
    assert np.allclose(got, [lse10(r[keep], gw[keep]) for r in raw], rtol=1e-13, atol=1e-13)
    r2 = raw[:3].copy()
    r2[0, 3] = np.nan
    r2[1, 0] = np.nan
    r2[2] = [-0.1, 0.1, -5, -5, -5, -5]
    g2 = orc.avg_grid(r2, np.array([0.5, 0.5, 0, 0, 0, 0.0]))
    assert g2[0] == pytest.approx(lse10(np.delete(r2[0], 3), np.delete([.5, .5, 0, 0, 0, 0], 3)), abs=1e-13)
    assert np.isnan(g2[1])  # NaN first value
    assert g2[2] == pytest.approx(np.log10(0.5 * 10 ** -0.1 + 0.5 * 10 ** 0.1), abs=1e-14)
    tiny = orc.avg_grid(np.array([[1e-17, 1e-17]]), np.array([0.5, 0.5]))
    assert tiny[0] == 0.0


def test_configs_nested_equals_flat_and_gene_bf(orc):
    NP, D, L = 30, 7, 5
    raw = _raw(3, NP, D, L)
    off = np.array([0, 4, 11, 30])
    gw = np.random.default_rng(4).dirichlet(np.ones(L))
    cw = np.random.default_rng(5).dirichlet(np.ones(D))
    o = orc.avg_bfs(raw, off, gw, cfg_w=cw)
    for p in range(NP):
        flat = lse10(raw[p].ravel(), np.outer(cw, gw).ravel())
        assert o["snp_bf"][p] == pytest.approx(flat, abs=1e-12)
        for j in range(D):
            assert o["dim_bf"][p, j] == pytest.approx(lse10(raw[p, j], gw), abs=1e-12)
        assert o["best_dim"][p] == int(np.argmax(o["dim_bf"][p]))
    for g in range(3):
        sb = o["snp_bf"][off[g]:off[g + 1]]
        assert o["gene_bf"][g] == pytest.approx(lse10(sb, np.full(sb.size, 1 / sb.size)), abs=1e-12)


def test_posteriors_closed_forms(orc):
    NP, D, L = 25, 3, 4
    raw = _raw(6, NP, D, L, scale=0.8)
    off = np.array([0, 9, 25])
    gw = np.full(L, 1 / L)
    cw = np.array([0.6, 0.3, 0.1])
    csub = np.array([1, 2, 3], np.uint64)  # configurations "1", "2", "1-2"
    pi0 = 0.35
    o = orc.avg_bfs(raw, off, gw, cfg_w=cw, cfg_sub=csub, pi0=pi0, want=("a", "b", "c", "d", "dims"),
                    n_subgroups=2)
    for g in range(2):
        sl = slice(off[g], off[g + 1])
        m = off[g + 1] - off[g]
        bf = 10.0 ** o["gene_bf"][g]
        assert o["gene_post"][g] == pytest.approx((1 - pi0) * bf / (pi0 + (1 - pi0) * bf), rel=1e-12)
        assert o["post_the"][sl].sum() == pytest.approx(1.0, rel=1e-12)
        sbf = 10.0 ** o["snp_bf"][sl]
        assert np.allclose(o["post_an"][sl], (sbf / m) / (1 - 1 / m + sbf / m), rtol=1e-12)
    assert np.allclose(o["post_dim"].sum(1), 1.0, rtol=1e-12)
    assert np.array_equal(o["best_dim"], np.argmax(o["post_dim"], 1))
    assert np.allclose(o["post_sub"][:, 0], o["post_dim"][:, 0] + o["post_dim"][:, 2], rtol=1e-14)
    assert np.allclose(o["post_sub"][:, 1], o["post_dim"][:, 1] + o["post_dim"][:, 2], rtol=1e-14)


def _types_lik(raw_p, tw, sw, gw):
    """sum_k t_k sum_l w_l prod_s (q_ks 10^x_sl + 1 - q_ks), and the same
    with subgroup s active for the subgroup posteriors."""
    S, L = raw_p.shape
    bf = 10.0 ** raw_p
    tot = 0.0
    act = np.zeros(S)
    for k in range(len(tw)):
        fac = sw[k][:, None] * bf + 1 - sw[k][:, None]  # S x L
        tot += tw[k] * np.sum(gw * np.prod(fac, 0))
        for s in range(S):
            f2 = fac.copy()
            f2[s] = sw[k][s] * bf[s]
            act[s] += tw[k] * np.sum(gw * np.prod(f2, 0))
    return tot, act


def test_types_model_closed_form(orc):
    NP, S, L = 12, 3, 4
    raw = _raw(7, NP, S, L, scale=0.6)
    off = np.array([0, 5, 12])
    gw = np.random.default_rng(8).dirichlet(np.ones(L))
    tw = np.array([0.5, 0.3, 0.2])
    sw = np.random.default_rng(9).uniform(0.1, 0.9, (3, S))
    o = orc.avg_bfs(raw, off, gw, model="types", type_w=tw, sub_w=sw, pi0=0.5, want=("b", "d", "dims"))
    for p in range(NP):
        tot, act = _types_lik(raw[p], tw, sw, gw)
        assert o["snp_bf"][p] == pytest.approx(np.log10(tot), abs=1e-12)
        assert np.allclose(o["post_sub"][p], np.minimum(act / tot, 1.0), rtol=1e-11)
    assert np.allclose(o["post_dim"].sum(1), 1.0, rtol=1e-12)


def test_binary_types_reduce_to_configs(orc):
    NP, S, L = 10, 3, 5
    raw = _raw(10, NP, S, L, scale=0.5)
    off = np.array([0, 10])
    gw = np.full(L, 0.2)
    tw = np.array([0.7, 0.2, 0.1])
    sw = np.array([[1, 1, 1], [1, 0, 0], [0, 1, 1]], float)
    o_t = orc.avg_bfs(raw, off, gw, model="types", type_w=tw, sub_w=sw)
    cfg_raw = np.stack([raw[:, np.nonzero(sw[k])[0], :].sum(1) for k in range(3)], 1)
    o_c = orc.avg_bfs(cfg_raw, off, gw, cfg_w=tw)
    assert np.allclose(o_t["snp_bf"], o_c["snp_bf"], rtol=1e-12, atol=1e-12)
    assert np.allclose(o_t["dim_bf"], o_c["dim_bf"], rtol=1e-12, atol=1e-12)
    assert o_t["gene_bf"][0] == pytest.approx(o_c["gene_bf"][0], abs=1e-12)
