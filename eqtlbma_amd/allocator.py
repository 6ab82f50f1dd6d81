"""Learned permutation allocator (no counterpart in the reference).

eqtlbma_bf gives every gene the same --nperm.  Here a small network predicts,
from a gene's features, the null distribution of its permutation statistic
as a GEV (location mu, scale sigma, shape xi).  A gene whose observed
statistic is far in the bulk of its predicted null (predicted p above a
threshold, 0.2 by default) gets a short run of permutations (1,000); the
others get the full run (10,000).  The p-values always come from actual
permutations of the engine; the network only decides how many.

- Inputs (24): log cis-SNP count, N, subgroup count, the 20 largest
  eigenvalues of the gene's cis LD matrix (log1p), the observed statistic.
- Network: 24 -> 256 x 3 (GELU) -> 3, BF16 weights, FP32 accumulation,
  trained on the device (eqbf_mlp_* kernels) on per-gene GEV fits of null
  draws made by the permutation engine (eqbf_gev_fit), standardised targets
  (mu, log sigma, xi), squared error.
- Permutations of a gene given k of them are the first k of its index
  table, so a short run is a prefix of the full one.
"""
import numpy as np

from . import engine as _engine

N_EIG = 20


def features(n_snps, n_samples, n_subgroups, eig, observed):
    """[G][24] float32 network inputs."""
    G = len(n_snps)
    x = np.zeros((G, 4 + N_EIG), np.float32)
    x[:, 0] = np.log(np.maximum(np.asarray(n_snps, dtype=np.float64), 1.0))
    x[:, 1] = n_samples
    x[:, 2] = n_subgroups
    e = np.nan_to_num(np.asarray(eig, dtype=np.float64)[:, :N_EIG], nan=0.0)
    x[:, 3:3 + N_EIG] = np.log1p(np.maximum(e, 0.0))
    x[:, 3 + N_EIG] = np.nan_to_num(np.asarray(observed, dtype=np.float64), nan=0.0)
    return x


def gev_tail(x, mu, sigma, xi):
    """P(X >= x) under a GEV(mu, sigma, xi), elementwise."""
    x, mu, sigma, xi = (np.asarray(a, dtype=np.float64) for a in (x, mu, sigma, xi))
    z = (x - mu) / sigma
    small = np.abs(xi) < 1e-9
    t = 1.0 + xi * z
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        gen = -np.expm1(-np.power(np.maximum(t, 0.0), -1.0 / np.where(small, 1.0, xi)))
        gum = -np.expm1(-np.exp(-z))
    out = np.where(t > 0, gen, np.where(xi > 0, 1.0, 0.0))
    return np.where(small, gum, out)


class Allocator:
    """Network + feature/target scaling; see the module docstring."""

    def __init__(self, eng, hidden=256, layers=3, seed=1):
        self.eng = eng
        self.net = _engine.Mlp(eng, 4 + N_EIG, hidden, layers, 3, seed)
        self.x_mu = self.x_sd = self.t_mu = self.t_sd = None

    def close(self):
        self.net.close()

    @staticmethod
    def targets(gev):
        g = np.asarray(gev, dtype=np.float64)
        return np.stack([g[:, 0], np.log(np.maximum(g[:, 1], 1e-12)), g[:, 2]], axis=1)

    def fit(self, x, gev, epochs=300, batch=256, lr=1e-3, seed=7):
        """Trains on inputs x [G][24] and per-gene GEV fits gev [G][3];
        returns the loss of each epoch."""
        t = self.targets(gev)
        ok = np.all(np.isfinite(t), axis=1) & np.all(np.isfinite(x), axis=1)
        x, t = x[ok], t[ok]
        self.x_mu, self.x_sd = x.mean(axis=0), x.std(axis=0)
        self.x_sd[self.x_sd < 1e-6] = 1.0
        self.t_mu, self.t_sd = t.mean(axis=0), t.std(axis=0)
        self.t_sd[self.t_sd < 1e-6] = 1.0
        xs = ((x - self.x_mu) / self.x_sd).astype(np.float32)
        ts = ((t - self.t_mu) / self.t_sd).astype(np.float32)
        w = np.ones(3, np.float32)
        return self.net.train(xs, ts, w, epochs, batch=min(batch, xs.shape[0]), lr=lr, seed=seed)

    def predict(self, x):
        """[G][3] predicted (mu, sigma, xi)."""
        xs = ((np.asarray(x, dtype=np.float64) - self.x_mu) / self.x_sd).astype(np.float32)
        y = self.net.forward(xs).astype(np.float64) * self.t_sd + self.t_mu
        return np.stack([y[:, 0], np.exp(y[:, 1]), y[:, 2]], axis=1)

    def p_hat(self, x, observed):
        g = self.predict(x)
        return gev_tail(observed, g[:, 0], g[:, 1], g[:, 2])

    @staticmethod
    def allocate(p_hat, threshold=0.2, low=1000, high=10000):
        """Permutations per gene: low where the predicted p exceeds the
        threshold, high elsewhere (and where the prediction is not finite)."""
        p = np.asarray(p_hat, dtype=np.float64)
        return np.where(np.isfinite(p) & (p > threshold), low, high).astype(np.int64)


def bh_reject(p, fdr=0.05):
    """Benjamini-Hochberg: which of the p-values are rejected at this FDR."""
    p = np.asarray(p, dtype=np.float64)
    n = p.shape[0]
    order = np.argsort(p, kind="stable")
    ranked = p[order] * n / np.arange(1, n + 1)
    ok = ranked <= fdr
    out = np.zeros(n, bool)
    if ok.any():
        k = np.max(np.nonzero(ok)[0])
        out[order[:k + 1]] = True
    return out


def allocated_pvalues(eng, run_perm, true_stat, nperm_gene, sep=False):
    """p-values with a per-gene number of permutations.

    run_perm(k, genes) returns the [len(genes)][k] statistics of the first k
    permutations of those genes (the engine's eqbf_perm_stats on a table
    prefix); genes are grouped by their k.  The reference's counting loop
    (eqbf_perm_pvalue) turns each gene's statistics into its p-value."""
    G = len(true_stat)
    pv = np.full(G, np.nan)
    used = np.zeros(G, np.int64)
    for k in np.unique(nperm_gene):
        genes = np.nonzero(nperm_gene == k)[0]
        st = run_perm(int(k), genes)
        for r, g in enumerate(genes):
            res = _engine.perm_pvalue(st[r], true_stat[g], int(k), sep=sep)
            pv[g], used[g] = res[0], res[1]
    return pv, used
