"""Synthetic eQTL data in the layout of include/eqbf.h.

Generators for the workloads the engine is measured on (numpy, CPU; the
full-size batched workload is generated on the GPU by bench.py):
  plain   -- S subgroups sharing N individuals, hard-call genotypes under HWE
             with MAF ~ U(0.05, 0.5), y = beta g + N(0,1), a fraction of genes
             with one causal SNP active in a random non-empty subset of
             subgroups.
  batched -- uneven subgroup membership (N_s ~ U(lo, hi) drawn from N_all),
             FP32 dosages E[g] + N(0, 0.05^2) clipped to [0, 2], Gaussian
             covariates regressed out by the model.
  mvlr    -- every individual in every subgroup, residuals correlated across
             subgroups (equicorrelation rho), effects multivariate normal with
             between-subgroup correlation r.
Grids follow makeGrid (scripts/utils_eqtlbma.R:65-95).
"""
import numpy as np

from . import _abi


def make_grid(kind="general", n_aes=5):
    """makeGrid: average effect sizes 0.1,0.2,0.4,0.8,1.6 (squared) times
    homogeneities {0,1/4,1/2,3/4,1} ("general") or {3/4,1} (small).  With
    n_aes=10 the effect sizes are 0.1*sqrt(2)^k, k=0..9 (a 50-point grid)."""
    if n_aes == 5:
        aes = [0.1 ** 2, 0.2 ** 2, 0.4 ** 2, 0.8 ** 2, 1.6 ** 2]
    else:
        aes = [(0.1 * 2 ** (k / 2.0)) ** 2 for k in range(n_aes)]
    hom = [0, 0.25, 0.5, 0.75, 1] if kind == "general" else [0.75, 1]
    g = []
    for a in aes:
        for h in hom:
            g.append((a * (1 - h), a * h))
    return np.array(g, dtype=np.float64)


def _hwe_genotypes(rng, n_snps, n, maf_lo=0.05, maf_hi=0.5):
    maf = rng.uniform(maf_lo, maf_hi, size=n_snps)
    g = rng.binomial(2, maf[:, None], size=(n_snps, n)).astype(np.float64)
    return g, maf


def plain(seed=42, S=3, N=100, n_genes=200, snps_per_gene=201, frac_egenes=0.1, beta_sd=0.5,
          geno_dtype=np.float64):
    """Plain method shape: genes each with their own window of cis SNPs."""
    rng = np.random.default_rng(seed)
    P = snps_per_gene
    n_snps = n_genes * P
    g, maf = _hwe_genotypes(rng, n_snps, N)
    # redraw SNPs monomorphic in the sample (all individuals share genotypes)
    mono = g.std(axis=1) == 0
    while mono.any():
        g[mono] = rng.binomial(2, maf[mono][:, None], size=(mono.sum(), N))
        mono = g.std(axis=1) == 0
    y = rng.normal(size=(n_genes, S, N))
    n_e = int(round(frac_egenes * n_genes))
    egenes = rng.choice(n_genes, size=n_e, replace=False)
    for gi in egenes:
        j = gi * P + rng.integers(P)
        beta = rng.normal(0, beta_sd)
        mask = 0
        while mask == 0:
            mask = rng.integers(1, 1 << S)
        for s in range(S):
            if (mask >> s) & 1:
                y[gi, s] += beta * g[j]
    genotypes = np.ascontiguousarray(g.T).astype(geno_dtype)  # [N][n_snps]
    pair_offset = np.arange(n_genes + 1, dtype=np.int64) * P
    pair_snp = np.arange(n_snps, dtype=np.int32)
    return _abi.HostChunk(
        explevels=y, gene_in_sbgrp=np.ones((n_genes, S), np.uint8),
        geno_in_sbgrp=np.ones((S, N), np.uint8), genotypes=genotypes,
        snp_in_sbgrp=np.ones((n_snps, S), np.uint8), pair_offset=pair_offset,
        pair_snp=pair_snp)


def batched(seed=1001, S=44, N_all=500, n_genes=20, p_mean=1500, p_lo=200, p_hi=6000,
            ns_lo=80, ns_hi=400, Q=3, frac_egenes=0.15, beta_sd=0.5, nan_frac=0.0,
            absent_frac=0.0):
    """Batched-ABF shape: uneven subgroup sizes, FP32 dosages, Q covariates.

    nan_frac: fraction of expression values set to NaN (missing);
    absent_frac: fraction of (gene, subgroup) with no expression at all."""
    rng = np.random.default_rng(seed)
    P = np.clip(rng.poisson(p_mean, size=n_genes), p_lo, p_hi)
    n_snps = int(P.sum())
    maf = rng.uniform(0.05, 0.5, size=n_snps)
    eg = rng.binomial(2, maf[:, None], size=(n_snps, N_all)).astype(np.float32)
    dos = np.clip(eg + rng.normal(0, 0.05, size=eg.shape).astype(np.float32), 0, 2).astype(np.float32)
    geno_in = np.zeros((S, N_all), np.uint8)
    for s in range(S):
        ns = rng.integers(ns_lo, ns_hi + 1)
        geno_in[s, rng.choice(N_all, size=ns, replace=False)] = 1
    covars = rng.normal(size=(S, max(Q, 1), N_all))
    covar_count = np.full(S, Q, np.int32)
    y = rng.normal(size=(n_genes, S, N_all))
    y += 0.3 * np.einsum("sqn->sn", covars[:, :Q, :])[None] if Q > 0 else 0
    offs = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    n_e = int(round(frac_egenes * n_genes))
    for gi in rng.choice(n_genes, size=n_e, replace=False):
        j = offs[gi] + rng.integers(P[gi])
        beta = rng.normal(0, beta_sd)
        act = rng.random(S) < 0.5
        if not act.any():
            act[rng.integers(S)] = True
        y[gi, act] += beta * dos[j][None, :]
    # expression only exists for samples of the subgroup
    y[:, geno_in == 0] = np.nan
    if nan_frac > 0:
        y[rng.random(y.shape) < nan_frac] = np.nan
    gene_in = np.ones((n_genes, S), np.uint8)
    if absent_frac > 0:
        gene_in[rng.random((n_genes, S)) < absent_frac] = 0
        y[gene_in == 0] = np.nan
    return _abi.HostChunk(
        explevels=y, gene_in_sbgrp=gene_in, geno_in_sbgrp=geno_in,
        genotypes=np.ascontiguousarray(dos.T), snp_in_sbgrp=np.ones((n_snps, S), np.uint8),
        pair_offset=offs, pair_snp=np.arange(n_snps, dtype=np.int32), covars=covars,
        covar_count=covar_count, covar_in_sbgrp=np.ones((S, N_all), np.uint8))


def mvlr(seed=1003, S=9, N=150, n_genes=4, snps_per_gene=40, rho=0.5, r_eff=0.8,
         frac_egenes=0.5, beta_sd=0.5, Q=0):
    """MVLR shape: every individual in every subgroup, correlated residuals."""
    rng = np.random.default_rng(seed)
    P = snps_per_gene
    n_snps = n_genes * P
    g, _ = _hwe_genotypes(rng, n_snps, N)
    mono = g.std(axis=1) == 0
    while mono.any():
        g[mono] = rng.binomial(2, 0.3, size=(mono.sum(), N))
        mono = g.std(axis=1) == 0
    Sig = np.full((S, S), rho) + (1 - rho) * np.eye(S)
    L = np.linalg.cholesky(Sig)
    y = np.einsum("st,gtn->gsn", L, rng.normal(size=(n_genes, S, N)))
    Ceff = np.full((S, S), r_eff) + (1 - r_eff) * np.eye(S)
    Le = np.linalg.cholesky(Ceff)
    for gi in rng.choice(n_genes, size=int(round(frac_egenes * n_genes)), replace=False):
        j = gi * P + rng.integers(P)
        b = beta_sd * (Le @ rng.normal(size=S))
        y[gi] += b[:, None] * g[j][None, :]
    covars = rng.normal(size=(S, max(Q, 1), N))
    covars[:] = covars[0:1]  # same covariates in every subgroup
    return _abi.HostChunk(
        explevels=y, gene_in_sbgrp=np.ones((n_genes, S), np.uint8),
        geno_in_sbgrp=np.ones((S, N), np.uint8), genotypes=np.ascontiguousarray(g.T),
        snp_in_sbgrp=np.ones((n_snps, S), np.uint8),
        pair_offset=np.arange(n_genes + 1, dtype=np.int64) * P,
        pair_snp=np.arange(n_snps, dtype=np.int32), covars=covars,
        covar_count=np.full(S, Q, np.int32), covar_in_sbgrp=np.ones((S, N), np.uint8))


def hybrid(seed=1005, S=3, N=70, n_genes=2, P=4, Q=2, miss=0.25, same_covars=False, q_per_sbgrp=None,
           dosages=False):
    """--error hybrid shape: one genotype column per individual in every
    subgroup (one genotype file), expression levels missing at random per
    subgroup, covariates of their own per subgroup (q_per_sbgrp: a count per
    subgroup, at most Q); dosages: FP32 genotypes with noise."""
    rng = np.random.default_rng(seed)
    n_snps = n_genes * P
    g = rng.binomial(2, rng.uniform(0.1, 0.5, n_snps)[:, None], size=(n_snps, N)).astype(float)
    if dosages:
        g = np.clip(g + rng.normal(0, 0.05, g.shape), 0, 2).astype(np.float32).astype(np.float64)
    covars = rng.normal(size=(S, max(Q, 1), N))
    if same_covars:
        covars[:] = covars[0:1]
    y = rng.normal(size=(n_genes, S, N)) + 0.4 * covars[None, :, 0, :]
    for gi in range(n_genes):
        j = gi * P + rng.integers(P)
        y[gi] += rng.normal(0, 0.6, size=(S, 1)) * g[j][None, :]
    y[rng.random(y.shape) < miss] = np.nan
    qc = np.full(S, Q, np.int32) if q_per_sbgrp is None else np.asarray(q_per_sbgrp, np.int32)
    return _abi.HostChunk(
        explevels=y, gene_in_sbgrp=np.ones((n_genes, S), np.uint8), geno_in_sbgrp=np.ones((S, N), np.uint8),
        genotypes=np.ascontiguousarray(g.T).astype(np.float32 if dosages else np.float64),
        snp_in_sbgrp=np.ones((n_snps, S), np.uint8),
        pair_offset=np.arange(n_genes + 1, dtype=np.int64) * P, pair_snp=np.arange(n_snps, dtype=np.int32),
        covars=covars, covar_count=qc, covar_in_sbgrp=np.ones((S, N), np.uint8))
