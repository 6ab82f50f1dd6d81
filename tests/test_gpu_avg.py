"""GPU eqtlbma_avg_bfs (eqbf_avg_bfs / eqbf_avg_grid) against the oracle.

The device sums over dimensions and SNPs with wave reductions where the
oracle sums in order, so log10 values and posteriors agree to rounding:
1e-12 absolute (1e-12 relative for posteriors), NaN where the oracle has
NaN, best dimensions exactly.  Cases: the grid-only average with a subset
of grid points and NaN values, the configs model with 127 configurations
(two passes of the 64 lanes) and every posterior, the types model with every
posterior, a gene of 150 SNPs (lanes over SNPs wrap), NaN raw values
including a NaN first value.
"""
import sys

import numpy as np
import pytest

from eqtlbma_amd import engine
from eqtlbma_amd._abi import config_masks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cmp(got, ref, keys):
    for k in keys:
        a, b = got[k], ref[k]
        if k == "best_dim":
            assert np.array_equal(a, b), k
            continue
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        fin = np.isfinite(b)
        err = float(np.max(np.abs(a[fin] - b[fin]), initial=0))
        _say("[gpu]   %s max abs err %.2e" % (k, err))
        assert np.all(np.abs(a[fin] - b[fin]) <= 1e-12 + 1e-12 * np.abs(b[fin])), k


def test_avg_grid_vs_oracle(eng, orc):
    rng = np.random.default_rng(1)
    raw = rng.normal(0.3, 1.5, (5000, 12))
    raw[rng.random(raw.shape) < 0.02] = np.nan
    gw = rng.uniform(0.01, 0.2, 12)
    keep = np.array([0, 1, 4, 7, 11], np.int32)
    _say("[gpu] avg_grid rows=%d" % raw.shape[0])
    for kp in (None, keep):
        got = eng.avg_grid(raw, gw, kp)
        ref = orc.avg_grid(raw, gw, kp)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        fin = np.isfinite(ref)
        assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-12


def _genes(rng, G, lo, hi):
    m = rng.integers(lo, hi + 1, G)
    return np.concatenate([[0], np.cumsum(m)]).astype(np.int64)


def test_avg_configs_all_posteriors(eng, orc):
    rng = np.random.default_rng(2)
    S = 7
    masks = np.array(list(config_masks("all", S)), np.uint64)
    D, L = masks.shape[0], 10
    off = _genes(rng, 40, 1, 150)
    NP = int(off[-1])
    raw = rng.normal(-0.4, 0.7, (NP, D, L))
    hot = rng.choice(NP, 30, replace=False)
    raw[hot] += rng.uniform(0.5, 3, (30, D, 1))
    raw[rng.random(raw.shape) < 0.01] = np.nan
    raw[5, 0, 0] = np.nan  # NaN first grid value of the first configuration
    gw = rng.dirichlet(np.ones(L))
    cw = rng.dirichlet(np.ones(D))
    keep = np.array([0, 1, 2, 3, 5, 8, 9], np.int32)
    kw = dict(cfg_w=cw, cfg_sub=masks, pi0=0.3, want=("a", "b", "c", "d", "dims"), keep=keep, n_subgroups=S)
    _say("[gpu] avg configs D=%d pairs=%d" % (D, NP))
    got = eng.avg_bfs(raw, off, gw, **kw)
    ref = orc.avg_bfs(raw, off, gw, **kw)
    _cmp(got, ref, ["dim_bf", "snp_bf", "gene_bf", "gene_post", "post_the", "post_an", "post_dim", "post_sub",
                    "best_dim"])


def test_avg_configs_best_dim_without_posteriors(eng, orc):
    rng = np.random.default_rng(3)
    off = _genes(rng, 10, 2, 20)
    NP = int(off[-1])
    raw = rng.normal(0, 1, (NP, 3, 5))
    cw = np.array([0.5, 0.25, 0.25])
    got = eng.avg_bfs(raw, off, np.full(5, 0.2), cfg_w=cw)
    ref = orc.avg_bfs(raw, off, np.full(5, 0.2), cfg_w=cw)
    _cmp(got, ref, ["dim_bf", "snp_bf", "gene_bf", "best_dim"])


def test_avg_types_all_posteriors(eng, orc):
    rng = np.random.default_rng(4)
    S, K, L = 5, 3, 6
    off = _genes(rng, 30, 1, 80)
    NP = int(off[-1])
    raw = rng.normal(-0.3, 0.8, (NP, S, L))
    raw[rng.random(raw.shape) < 0.01] = np.nan
    gw = rng.dirichlet(np.ones(L))
    tw = np.array([0.6, 0.3, 0.1])
    sw = rng.uniform(0.05, 0.95, (K, S))
    kw = dict(model="types", type_w=tw, sub_w=sw, pi0=0.6, want=("a", "b", "c", "d", "dims"))
    _say("[gpu] avg types S=%d K=%d pairs=%d" % (S, K, NP))
    got = eng.avg_bfs(raw, off, gw, **kw)
    ref = orc.avg_bfs(raw, off, gw, **kw)
    _cmp(got, ref, ["dim_bf", "snp_bf", "gene_bf", "gene_post", "post_the", "post_an", "post_dim", "post_sub",
                    "best_dim"])
