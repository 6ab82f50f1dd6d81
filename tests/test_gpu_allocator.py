"""The learned permutation allocator on the GPU.

- The network's forward pass and gradients (HIP kernels, BF16 weights, FP32
  accumulation) against the FP64 CPU oracle evaluated with the same
  BF16-rounded weights (oracle/mlp_oracle.py, itself pinned by finite
  differences): outputs to 1e-4 of their scale, gradients to 1e-4 in norm.
- LD eigenvalues against numpy's eigvalsh of the correlation matrix, the GEV
  fits against the oracle's probability-weighted moments.
- End to end on a small permutation set: the trained network predicts
  held-out genes' null distributions better than one GEV pooled over the
  training genes (the method without the network), the allocation spends
  fewer permutations than permuting every gene fully, and the eGene calls
  (Benjamini-Hochberg, 5%) agree with full permutation.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import _abi, allocator, engine, synth
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

GL25 = synth.make_grid("general")
GS10 = synth.make_grid("small")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL25, GS10)
    yield e
    e.close()


def test_network_forward_and_gradients_match_oracle(eng):
    _say("[gpu] allocator network vs oracle")
    net = engine.Mlp(eng, 24, 256, 3, 3, seed=5)
    ref = mo.Mlp(24, 256, 3, 3)
    assert net.n_params == ref.n_params
    rng = np.random.default_rng(2)
    params = net.get_params()
    params[ref.n_params - 3:] = rng.normal(size=3)  # non-zero output biases
    net.set_params(params)
    x = rng.normal(size=(96, 24)).astype(np.float32)
    t = rng.normal(size=(96, 3)).astype(np.float32)
    w = np.array([1.0, 0.5, 2.0], np.float32)
    y = net.forward(x)
    y_ref = ref.forward(params, x)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = net.loss_grad(x, t, w)
    loss_ref, g_ref = ref.loss_grad(params, x, t, w)
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    _say("[gpu]   gradient relative error %.2e" % err)
    assert err <= 1e-4
    off = 0
    for i, o in ref.dims:  # every layer, weights and biases
        for n in (i * o, o):
            seg, seg_ref = g[off:off + n], g_ref[off:off + n]
            assert np.max(np.abs(seg - seg_ref)) <= 1e-3 * np.max(np.abs(seg_ref)) + 1e-12
            off += n
    net.close()


def test_network_training_reduces_loss(eng):
    _say("[gpu] allocator network training")
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1024, 24)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(2 * x[:, 3]), x[:, 4] ** 2 - 1], axis=1)
    net = engine.Mlp(eng, 24, 256, 3, 3, seed=9)
    w = np.ones(3, np.float32)
    curve = net.train(x, t.astype(np.float32), w, epochs=60, batch=64, lr=1e-3, seed=3)
    _say("[gpu]   loss %.4f -> %.4f" % (curve[0], curve[-1]))
    assert np.all(np.isfinite(curve))
    assert curve[-1] < 0.2 * curve[0]
    net.close()


def test_ld_eigenvalues_match_numpy(eng):
    _say("[gpu] LD eigenvalues")
    hc = synth.plain(seed=3, S=2, N=120, n_genes=3, snps_per_gene=80)
    # correlated SNPs: each SNP copies its neighbour with probability 0.7
    g = hc.genotypes[0]
    rng = np.random.default_rng(8)
    for j in range(1, g.shape[1]):
        keep = rng.random(g.shape[0]) < 0.7
        g[keep, j] = g[keep, j - 1]
    eig = eng.ld_eigenvalues(hc, 20)
    for gi in range(hc.G):
        X = g[:, hc.pair_snp[hc.pair_offset[gi]:hc.pair_offset[gi + 1]]].astype(np.float64)
        sd = X.std(axis=0)
        Z = (X - X.mean(axis=0))[:, sd > 0] / sd[sd > 0]
        lam = np.sort(np.linalg.eigvalsh(Z.T @ Z / X.shape[0]))[::-1][:20]
        assert np.allclose(eig[gi, :10], lam[:10], rtol=1e-3)
        assert np.allclose(eig[gi, 10:], lam[10:], rtol=2e-2)


def test_gev_fit_matches_oracle(eng):
    _say("[gpu] GEV fits")
    from scipy import stats
    rng = np.random.default_rng(5)
    G, n = 40, 5000
    xs = np.stack([stats.genextreme.rvs(-rng.uniform(-0.2, 0.3), loc=rng.normal(), scale=rng.uniform(0.3, 2),
                                        size=n, random_state=100 + k) for k in range(G)])
    xs[3, ::7] = np.nan  # NaN draws are not counted
    got = eng.gev_fit(xs)
    for k in range(G):
        v = xs[k][np.isfinite(xs[k])]
        want = mo.gev_fit_pwm(v)
        assert np.allclose(got[k], want, rtol=1e-9, atol=1e-12)


def _variable_p_chunk(seed, S, N, n_genes, p_lo, p_hi):
    """Genes with different numbers of cis SNPs (and LD within each gene)."""
    rng = np.random.default_rng(seed)
    hc = synth.plain(seed=seed, S=S, N=N, n_genes=n_genes, snps_per_gene=p_hi, frac_egenes=0.15)
    g = hc.genotypes[0]
    for j in range(1, g.shape[1]):
        keep = rng.random(N) < rng.uniform(0.0, 0.8)
        g[keep, j] = g[keep, j - 1]
    P = rng.integers(p_lo, p_hi + 1, size=n_genes)
    snp = np.concatenate([np.arange(k * p_hi, k * p_hi + P[k]) for k in range(n_genes)]).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    # phenotypes from the final genotypes: one causal cis SNP in 20% of genes
    y = rng.normal(size=(n_genes, S, N))
    for gi in rng.choice(n_genes, size=n_genes // 5, replace=False):
        j = gi * p_hi + rng.integers(P[gi])
        beta = rng.choice([-1.0, 1.0]) * rng.uniform(0.3, 0.8)
        act = rng.random(S) < 0.6
        act[rng.integers(S)] = True
        y[gi, act] += beta * g[:, j]
    return _abi.HostChunk(explevels=y, gene_in_sbgrp=hc.gene_in_sbgrp, geno_in_sbgrp=hc.geno_in_sbgrp,
                          genotypes=hc.genotypes, snp_in_sbgrp=hc.snp_in_sbgrp, pair_offset=off, pair_snp=snp), P


def _ks(x, mu, sigma, xi):
    """Kolmogorov distance between the empirical distribution of x and a GEV."""
    x = np.sort(x[np.isfinite(x)])
    n = x.shape[0]
    F = 1.0 - allocator.gev_tail(x, mu, sigma, xi)
    return float(max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n)))


def test_allocator_end_to_end_and_ablation(eng, orc):
    _say("[gpu] allocator end to end")
    S, N, G, ntrain = 3, 100, 240, 160
    hc, P = _variable_p_chunk(21, S, N, G, 8, 120)
    nfull, nlow = 1000, 100
    tables = engine.Rng(7).tables(N, 10, nfull)
    gene_table = (np.arange(G) % 10).astype(np.int32)
    full = eng.perm_stats(hc, tables, gene_table, "join", "all", True, precision="mixed")
    ass = eng.assoc_uvlr(hc, "join", "all", precision="mixed", raw_dtype=np.float32, want_sstats=False,
                         want_std=False, want_raw=False)
    w_all = ass.w[3 + (2 ** S - 1) + 1]
    true = np.array([np.max(w_all[hc.pair_offset[g]:hc.pair_offset[g + 1]]) for g in range(G)])
    eig = eng.ld_eigenvalues(hc, allocator.N_EIG)
    x = allocator.features(P, N, S, eig, true)
    gev = eng.gev_fit(full)
    tr, te = np.arange(ntrain), np.arange(ntrain, G)
    al = allocator.Allocator(eng, seed=3)
    curve = al.fit(x[tr], gev[tr], epochs=150, batch=32, lr=1e-3)
    _say("[gpu]   training loss %.3f -> %.3f" % (curve[0], curve[-1]))
    pred = al.predict(x[te])
    # ablation: the trained network against one GEV pooled over the training genes
    pooled = mo.gev_fit_pwm(full[tr].ravel())
    ks_net = np.mean([_ks(full[g], *pred[k]) for k, g in enumerate(te)])
    ks_pool = np.mean([_ks(full[g], *pooled) for g in te])
    _say("[gpu]   held-out null distributions, mean Kolmogorov distance: network %.4f, pooled GEV %.4f"
         % (ks_net, ks_pool))
    assert ks_net < ks_pool
    # allocation on the held-out genes: short runs are prefixes of the full tables
    p_hat = al.p_hat(x[te], true[te])
    k_gene = allocator.Allocator.allocate(p_hat, 0.2, nlow, nfull)

    def run_perm(k, genes):
        return full[te[genes], :k]

    pv_alloc, used = allocator.allocated_pvalues(eng, run_perm, true[te], k_gene)
    pv_full = np.array([engine.perm_pvalue(full[g], true[g], nfull)[0] for g in te])
    calls_a = allocator.bh_reject(pv_alloc, 0.05)
    calls_f = allocator.bh_reject(pv_full, 0.05)
    agree = np.mean(calls_a == calls_f)
    frac = used.sum() / (nfull * len(te))
    _say("[gpu]   permutations used %.2f of full; eGene calls %d vs %d, agreement %.3f"
         % (frac, calls_a.sum(), calls_f.sum(), agree))
    assert frac < 0.8
    assert calls_f.sum() > 0
    assert agree >= 0.95
    al.close()
