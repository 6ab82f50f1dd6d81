"""ctypes access to the CPU oracle (test infrastructure only).

Only the tests, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use this module, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from eqtlbma_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc_eqbf.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        d, i32, i64, u32, u64, vp = (C.c_double, C.c_int32, C.c_int64, C.c_uint32,
                                    C.c_uint64, C.c_void_p)
        L.orc_mt_seed.argtypes = [vp, u64]
        L.orc_mt_next.argtypes = [vp]
        L.orc_mt_next.restype = u32
        L.orc_uniform_int.argtypes = [vp, u32]
        L.orc_uniform_int.restype = u32
        L.orc_shuffle.argtypes = [vp, vp, i64]
        L.orc_flat.argtypes = [vp, d, d]
        L.orc_flat.restype = d
        L.orc_perm_tables.argtypes = [u64, i64, i64, i64, vp]
        for f in ("orc_tdist_P", "orc_tdist_Q", "orc_fdist_Q_1"):
            getattr(L, f).argtypes = [d, d]
            getattr(L, f).restype = d
        for f in ("orc_gauss_Pinv", "orc_chisq_Qinv_1"):
            getattr(L, f).argtypes = [d]
            getattr(L, f).restype = d
        L.orc_fit_ols.argtypes = [i64, i64, vp, vp, vp]
        L.orc_log10_weighted_sum.argtypes = [vp, i64]
        L.orc_log10_weighted_sum.restype = d
        L.orc_log10_weighted_sum_w.argtypes = [vp, vp, i64]
        L.orc_log10_weighted_sum_w.restype = d
        L.orc_l10abf_uvlr.argtypes = [i32, vp, vp, d, d]
        L.orc_l10abf_uvlr.restype = d
        L.orc_standardize.argtypes = [d, d, d, d, d, vp]
        L.orc_config_masks.argtypes = [i32, i32, vp]
        L.orc_config_masks.restype = i64
        L.orc_assoc_uvlr.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, i32,
                                     vp, vp, vp, vp, vp]
        L.orc_assoc_uvlr.restype = C.c_int
        L.orc_perm_join_stats.argtypes = [vp, i64, i32, i32, vp, vp, i32, vp, vp, i32,
                                          vp, i64, vp]
        L.orc_perm_join_stats.restype = C.c_int
        L.orc_perm_join_stats_mvlr.argtypes = [vp, i64, i32, i32, d, vp, vp, i32, vp, vp, i32, vp, i64, vp]
        L.orc_perm_join_stats_mvlr.restype = C.c_int
        L.orc_true_stat.argtypes = [vp, i64, i32]
        L.orc_true_stat.restype = d
        L.orc_perm_count.argtypes = [vp, d, vp, i32, i64, vp, vp, vp, vp]
        L.orc_perm_count.restype = i64
        L.orc_perm_sep_stats.argtypes = [vp, i64, i32, vp, i64, vp]
        L.orc_perm_sep_stats.restype = C.c_int
        L.orc_perm_sep_count.argtypes = [vp, d, vp, i32, i64, vp, vp, vp]
        L.orc_perm_sep_count.restype = i64
        L.orc_mvlr_abfs.argtypes = [i32, i32, i32, vp, vp, vp, d, vp, vp, vp, i32, vp]
        L.orc_mvlr_sigma.argtypes = [i32, i32, i32, vp, vp, vp, d, vp, vp, vp]
        L.orc_assoc_mvlr.argtypes = [vp, i32, d, vp, vp, i32, vp, vp, i32, vp, vp, vp]
        L.orc_assoc_mvlr.restype = C.c_int
        L.orc_hm_em.argtypes = [vp, vp, i64, i32, i32, i32, i32, d, i64, d, vp, vp, vp, vp, vp, vp, vp, i64]
        L.orc_hm_em.restype = i64
        L.orc_hm_loglik.argtypes = [vp, vp, i64, i32, i32, i32, i32, vp, vp, vp, vp, vp]
        L.orc_hm_loglik.restype = d
        L.orc_hm_bfs.argtypes = [vp, vp, i64, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_assoc_hybrid.argtypes = [vp, i32, d, vp, vp, vp, vp, i32, vp, vp, i32, vp, vp, vp, vp]
        L.orc_assoc_hybrid.restype = i32
        L.orc_hybrid_pair_stats.argtypes = [vp, i64, i64, d, vp, vp, vp, vp, vp]
        L.orc_perm_join_stats_hybrid.argtypes = [vp, i64, i32, i32, d, vp, vp, vp, vp, i32, vp, vp, i32, vp, i64,
                                                 vp]
        L.orc_perm_join_stats_hybrid.restype = i32
        L.orc_hybrid_pair_stats.restype = i32
        L.orc_avg_grid.argtypes = [vp, i64, i32, vp, vp, i32, vp]
        L.orc_avg_bfs.argtypes = [vp, vp, i64, i32, i32, i32, i32, vp, i32, vp, vp, vp, vp, vp, d, i32,
                                  vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_hm_ci.argtypes = [vp, vp, i64, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]
        L.orc_hm_ci.restype = i32
        L.orc_abf_from_sstats.argtypes = [i32, i64, vp, vp, vp, vp, i32, vp, vp, i32, vp, vp, i32, vp, vp, vp]
        L.orc_abf_from_sstats.restype = C.c_int
        L.orc_bma.argtypes = [i32, i32, vp, d, vp, vp]
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


class MT:
    """GSL-semantics mt19937 stream (gsl_rng_set, uniform_int, shuffle, flat)."""

    def __init__(self, seed):
        self._st = (C.c_uint32 * 625)()
        lib().orc_mt_seed(C.addressof(self._st), seed)

    def copy(self):
        m = MT.__new__(MT)
        m._st = (C.c_uint32 * 625)()
        C.memmove(m._st, self._st, C.sizeof(self._st))
        return m

    def next(self):
        return lib().orc_mt_next(C.addressof(self._st))

    def uniform_int(self, n):
        return lib().orc_uniform_int(C.addressof(self._st), n)

    def shuffle(self, a):
        a = np.ascontiguousarray(a, dtype=np.int64)
        lib().orc_shuffle(C.addressof(self._st), _p(a), a.size)
        return a

    def flat(self, a, b):
        return lib().orc_flat(C.addressof(self._st), a, b)

    @property
    def addr(self):
        return C.addressof(self._st)


def perm_tables(seed, n_samples, n_genes, nperm):
    t = np.zeros((n_genes, nperm, n_samples), np.int32)
    lib().orc_perm_tables(seed, n_samples, n_genes, nperm, _p(t))
    return t


def tdist_P(t, nu):
    return lib().orc_tdist_P(float(t), float(nu))


def tdist_Q(t, nu):
    return lib().orc_tdist_Q(float(t), float(nu))


def gauss_Pinv(p):
    return lib().orc_gauss_Pinv(float(p))


def fdist_Q_1(F, nu):
    return lib().orc_fdist_Q_1(float(F), float(nu))


def chisq_Qinv_1(q):
    return lib().orc_chisq_Qinv_1(float(q))


def fit_ols(X, y):
    """X: N x P with the genotype in the last column.  Returns
    (pve, sigmahat, betahat, sebetahat, betapval)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_fit_ols(X.shape[0], X.shape[1], _p(X), _p(y), _p(out))
    return out


def log10_weighted_sum(v, w=None):
    v = np.ascontiguousarray(v, dtype=np.float64)
    if w is None:
        return lib().orc_log10_weighted_sum(_p(v), v.size)
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().orc_log10_weighted_sum_w(_p(v), _p(w), v.size)


def l10abf_uvlr(gamma, stdsstats, phi2, oma2):
    g = np.ascontiguousarray(gamma, dtype=np.int32)
    st = np.ascontiguousarray(stdsstats, dtype=np.float64)
    return lib().orc_l10abf_uvlr(g.size, _p(g), _p(st), phi2, oma2)


def bma(S, bfs, w_cfg, gen):
    """CalcBMAlite / CalcBMA of one pair: (gen-sin, all) from the
    configuration averages (gsl_combination order) and gen; all is NaN unless
    bfs == 'all'."""
    w = np.ascontiguousarray(w_cfg, dtype=np.float64)
    gs, al = np.zeros(1), np.full(1, np.nan)
    lib().orc_bma(S, _abi.BFS[bfs], _p(w), float(gen), _p(gs), _p(al))
    return float(gs[0]), float(al[0])


def standardize(N, Q, sigmahat, betahat, sebetahat):
    out = np.zeros(3)
    lib().orc_standardize(N, Q, sigmahat, betahat, sebetahat, _p(out))
    return out


def assoc_uvlr(hc, analysis, bfs, gridL, gridS):
    """Run the oracle's association test on a HostChunk; returns HostAssocOut."""
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS if gridS is not None else gridL, dtype=np.float64)
    out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, gL.shape[0], gS.shape[0])
    st = hc.struct()
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    rc = lib().orc_assoc_uvlr(C.addressof(st), _abi.ANALYSIS[analysis], _abi.BFS[bfs],
                              _p(phiL), _p(omaL), gL.shape[0], _p(phiS), _p(omaS), gS.shape[0],
                              _p(out.n), _p(out.sstats), _p(out.stdsstats), _p(out.raw), _p(out.w))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return out


def abf_from_sstats(n, sigmahat, betahat, sebetahat, bfs, gridL, gridS):
    """--inss: Bayes factors from given summary statistics, arrays [S][NP]
    (n <= 0: no result).  Returns (stdsstats [3][S][NP], raw [n_raw][NP],
    w [n_w][NP]) in FP64."""
    n = np.ascontiguousarray(n, dtype=np.int64)
    S, NP = n.shape
    sig, beta, se = (np.ascontiguousarray(a, dtype=np.float64) for a in (sigmahat, betahat, sebetahat))
    assert sig.shape == beta.shape == se.shape == (S, NP)
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS if gridS is not None else gridL, dtype=np.float64)
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    nc = lib().orc_config_masks(_abi.BFS[bfs], S, None)
    std = np.full((3, S, NP), np.nan)
    raw = np.full((3 * gL.shape[0] + nc * gS.shape[0], NP), np.nan)
    w = np.full((3 + nc + 2, NP), np.nan)
    rc = lib().orc_abf_from_sstats(S, NP, _p(n), _p(sig), _p(beta), _p(se), _abi.BFS[bfs], _p(phiL), _p(omaL),
                                   gL.shape[0], _p(phiS), _p(omaS), gS.shape[0], _p(std), _p(raw), _p(w))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return std, raw, w


def assoc_mvlr(hc, bfs, fiterr, gridL, gridS):
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS if gridS is not None else gridL, dtype=np.float64)
    out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, gL.shape[0], gS.shape[0],
                            want_sstats=False, want_std=False)
    st = hc.struct()
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    rc = lib().orc_assoc_mvlr(C.addressof(st), _abi.BFS[bfs], fiterr, _p(phiL), _p(omaL),
                              gL.shape[0], _p(phiS), _p(omaS), gS.shape[0], _p(out.n),
                              _p(out.raw), _p(out.w))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return out


def assoc_hybrid(hc, bfs, fiterr, gridL, gridS, covars_pos=None, expl_pair=None):
    """--error hybrid on a chunk: n, summary statistics [5][S][NP], raw and
    grid-averaged log10 ABFs.  covars_pos [S][Qmax][N]: covariate values by
    file column (the pairwise designs read them there); None: the chunk's.
    expl_pair [G][S][N]: expression levels of the pairwise step (the levels
    before --qnorm); None: the chunk's."""
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS if gridS is not None else gridL, dtype=np.float64)
    out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, gL.shape[0], gS.shape[0], want_sstats=True, want_std=False)
    st = hc.struct()
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    cp = None if covars_pos is None else np.ascontiguousarray(covars_pos, dtype=np.float64)
    ep = None if expl_pair is None else np.ascontiguousarray(expl_pair, dtype=np.float64)
    rc = lib().orc_assoc_hybrid(C.addressof(st), _abi.BFS[bfs], fiterr, None if cp is None else _p(cp),
                                None if ep is None else _p(ep), _p(phiL), _p(omaL), gL.shape[0], _p(phiS), _p(omaS), gS.shape[0], _p(out.n),
                                _p(out.sstats), _p(out.raw), _p(out.w))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return out


def hybrid_pair_stats(hc, g, p, fiterr, covars_pos=None, expl_pair=None):
    """b [S], Sigma [S][S] and Vg [S][S] of pair p (of gene g) under
    --error hybrid."""
    S = hc.S
    b, Sig, Vg = np.zeros(S), np.zeros((S, S)), np.zeros((S, S))
    cp = None if covars_pos is None else np.ascontiguousarray(covars_pos, dtype=np.float64)
    st = hc.struct()
    ep = None if expl_pair is None else np.ascontiguousarray(expl_pair, dtype=np.float64)
    rc = lib().orc_hybrid_pair_stats(C.addressof(st), g, p, fiterr, None if cp is None else _p(cp),
                                     None if ep is None else _p(ep), _p(b), _p(Sig), _p(Vg))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return b, Sig, Vg


def mvlr_abfs(Y, x, Xc, fiterr, gamma, grid):
    """Y: s x n, x: n, Xc: qcov x n; grid: K x 2 (phi2, oma2)."""
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    Xc = np.ascontiguousarray(Xc if Xc is not None else np.zeros((0, Y.shape[1])), dtype=np.float64)
    g = np.ascontiguousarray(gamma, dtype=np.int32)
    grid = np.asarray(grid, dtype=np.float64)
    phi = np.ascontiguousarray(grid[:, 0])
    oma = np.ascontiguousarray(grid[:, 1])
    out = np.zeros(grid.shape[0])
    lib().orc_mvlr_abfs(Y.shape[1], Y.shape[0], Xc.shape[0], _p(Y), _p(x), _p(Xc), fiterr,
                        _p(g), _p(phi), _p(oma), grid.shape[0], _p(out))
    return out


def mvlr_sigma(Y, x, Xc, fiterr, gamma):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    Xc = np.ascontiguousarray(Xc if Xc is not None else np.zeros((0, Y.shape[1])), dtype=np.float64)
    g = np.ascontiguousarray(gamma, dtype=np.int32)
    s = Y.shape[0]
    S1 = np.zeros((s, s))
    S0 = np.zeros((s, s))
    lib().orc_mvlr_sigma(Y.shape[1], s, Xc.shape[0], _p(Y), _p(x), _p(Xc), fiterr, _p(g),
                         _p(S1), _p(S0))
    return S1, S0


def perm_join_stats(hc, g, pbf, use_max_bf, gridL, gridS, table):
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS, dtype=np.float64)
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    table = np.ascontiguousarray(table, dtype=np.int32)
    nperm = table.shape[0]
    stats = np.zeros(nperm)
    st = hc.struct()
    rc = lib().orc_perm_join_stats(C.addressof(st), g, _abi.PBF[pbf], int(use_max_bf),
                                   _p(phiL), _p(omaL), gL.shape[0], _p(phiS), _p(omaS),
                                   gS.shape[0], _p(table), nperm, _p(stats))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return stats


def perm_join_stats_mvlr(hc, g, pbf, use_max_bf, fiterr, gridL, gridS, table):
    """--error mvlr permutation statistics of gene g for every row of table."""
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS, dtype=np.float64)
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    table = np.ascontiguousarray(table, dtype=np.int32)
    stats = np.zeros(table.shape[0])
    st = hc.struct()
    rc = lib().orc_perm_join_stats_mvlr(C.addressof(st), g, _abi.PBF[pbf], int(use_max_bf), float(fiterr),
                                        _p(phiL), _p(omaL), gL.shape[0], _p(phiS), _p(omaS), gS.shape[0],
                                        _p(table), table.shape[0], _p(stats))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return stats


def perm_join_stats_hybrid(hc, g, pbf, use_max_bf, fiterr, gridL, gridS, table, covars_pos=None, expl_pair=None):
    """--error hybrid permutation statistics of gene g for every row of
    table (pairwise step on the levels as loaded)."""
    gL = np.ascontiguousarray(gridL, dtype=np.float64)
    gS = np.ascontiguousarray(gridS, dtype=np.float64)
    phiL, omaL = np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1])
    phiS, omaS = np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1])
    table = np.ascontiguousarray(table, dtype=np.int32)
    stats = np.zeros(table.shape[0])
    cp = None if covars_pos is None else np.ascontiguousarray(covars_pos, dtype=np.float64)
    ep = None if expl_pair is None else np.ascontiguousarray(expl_pair, dtype=np.float64)
    st = hc.struct()
    rc = lib().orc_perm_join_stats_hybrid(C.addressof(st), g, _abi.PBF[pbf], int(use_max_bf), float(fiterr),
                                          None if cp is None else _p(cp), None if ep is None else _p(ep),
                                          _p(phiL), _p(omaL), gL.shape[0], _p(phiS), _p(omaS), gS.shape[0],
                                          _p(table), table.shape[0], _p(stats))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return stats


def perm_sep_stats(hc, g, sbgrp, table):
    table = np.ascontiguousarray(table, dtype=np.int32)
    stats = np.zeros(table.shape[0])
    st = hc.struct()
    rc = lib().orc_perm_sep_stats(C.addressof(st), g, sbgrp, _p(table), table.shape[0], _p(stats))
    if rc != 0:
        raise RuntimeError("oracle: %s" % lib().orc_last_error().decode())
    return stats


def true_stat(wpbf, use_max_bf):
    w = np.ascontiguousarray(wpbf, dtype=np.float64)
    return lib().orc_true_stat(_p(w), w.size, int(use_max_bf))


def perm_count(stats, true_stat_value, nb_permutations, trick=0, trick_cutoff=10, rng_trick=None,
               sep=False):
    """Replays the reference's counting loop.  Returns (pval, nperm_used,
    median, nb_permutations_after, drawn); median is None for sep."""
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    nb = C.c_int64(nb_permutations)
    pv = C.c_double(0)
    used = C.c_int64(0)
    med = C.c_double(0)
    addr = rng_trick.addr if rng_trick is not None else None
    if sep:
        drawn = lib().orc_perm_sep_count(_p(stats), true_stat_value, C.addressof(nb), trick,
                                         trick_cutoff, addr, C.addressof(pv), C.addressof(used))
        return pv.value, used.value, None, nb.value, drawn
    drawn = lib().orc_perm_count(_p(stats), true_stat_value, C.addressof(nb), trick, trick_cutoff,
                                 addr, C.addressof(pv), C.addressof(used), C.addressof(med))
    return pv.value, used.value, med.value, nb.value, drawn


class HmParams:
    """Parameters of eqtlbma_hm's model: pi0, grid weights, and either the
    configuration weights (configs model) or the type weights and the
    per-type subgroup probabilities [K][S] (types model)."""

    def __init__(self, pi0, grid, cfg=None, type_=None, sub=None):
        self.pi0 = float(pi0)
        self.grid = np.ascontiguousarray(grid, dtype=np.float64).copy()
        self.cfg = None if cfg is None else np.ascontiguousarray(cfg, dtype=np.float64).copy()
        self.type = None if type_ is None else np.ascontiguousarray(type_, dtype=np.float64).copy()
        self.sub = None if sub is None else np.ascontiguousarray(sub, dtype=np.float64).copy()

    def copy(self):
        return HmParams(self.pi0, self.grid, self.cfg, self.type, self.sub)


def _hm_args(raw, off, prm):
    raw = np.ascontiguousarray(raw, dtype=np.float64)
    off = np.ascontiguousarray(off, dtype=np.int64)
    types = prm.type is not None
    L = raw.shape[2]
    if types:
        K, S = prm.sub.shape
        cfg = np.zeros(1)
        typ, sub = prm.type, np.ascontiguousarray(prm.sub)
    else:
        K, S = raw.shape[1], 0
        cfg, typ, sub = prm.cfg, np.zeros(1), np.zeros(1)
    return raw, off, int(types), int(K), int(S), int(L), cfg, typ, sub


def hm_em(raw, off, prm, thresh=0.05, maxit=0, fixed=(0, 0, 0, 0, 0), stepmax=1.0, max_trace=100000):
    """EM of eqtlbma_hm from prm (not modified): classic for stepmax == 1,
    SQUAREM otherwise.  raw: [pairs][D][L] (D = configurations, or subgroups
    for the types model).  Returns (params, iterations, log10-likelihood
    trace)."""
    out = prm.copy()
    raw, off, types, K, S, L, cfg, typ, sub = _hm_args(raw, off, out)
    fx = np.ascontiguousarray(fixed, dtype=np.int32)
    pi0 = np.array([out.pi0])
    trace = np.full(max_trace, np.nan)
    it = lib().orc_hm_em(_p(raw), _p(off), off.size - 1, types, K, S, L, thresh, maxit, float(stepmax), _p(fx),
                         _p(pi0), _p(out.grid), _p(cfg), _p(typ), _p(sub), _p(trace), max_trace)
    if it == -1:
        raise RuntimeError("oracle: the log10 likelihood decreased")
    if it < 0:
        raise RuntimeError("oracle: NaN or infinite log10 likelihood or weight")
    out.pi0 = float(pi0[0])
    if types:
        out.type, out.sub = typ, sub
    else:
        out.cfg = cfg
    return out, int(it), trace[:min(it + 1, max_trace)]


def hm_loglik(raw, off, prm):
    raw, off, types, K, S, L, cfg, typ, sub = _hm_args(raw, off, prm)
    pi0 = np.array([prm.pi0])
    return lib().orc_hm_loglik(_p(raw), _p(off), off.size - 1, types, K, S, L, _p(pi0), _p(prm.grid), _p(cfg),
                               _p(typ), _p(sub))


def hm_bfs(raw, off, prm):
    """(gene posterior, gene log10 BF, SNP log10 BF, [pairs][dim] log10 BFs)."""
    raw, off, types, K, S, L, cfg, typ, sub = _hm_args(raw, off, prm)
    G, NP = off.size - 1, raw.shape[0]
    gp, gb, sb, db = np.zeros(G), np.zeros(G), np.zeros(NP), np.zeros((NP, K))
    pi0 = np.array([prm.pi0])
    lib().orc_hm_bfs(_p(raw), _p(off), G, types, K, S, L, _p(pi0), _p(prm.grid), _p(cfg), _p(typ), _p(sub),
                     _p(gp), _p(gb), _p(sb), _p(db))
    return gp, gb, sb, db


def hm_ci(raw, off, prm):
    """--getci bounds (left, right): pi0, configs or types, subgroups per
    type (types model), grid points."""
    raw, off, types, K, S, L, cfg, typ, sub = _hm_args(raw, off, prm)
    n = 1 + K + (K * S if types else 0) + L
    left, right = np.zeros(n), np.zeros(n)
    pi0 = np.array([prm.pi0])
    rc = lib().orc_hm_ci(_p(raw), _p(off), off.size - 1, types, K, S, L, _p(pi0), _p(prm.grid), _p(cfg), _p(typ),
                         _p(sub), _p(left), _p(right))
    if rc:
        raise RuntimeError("oracle: NaN or infinite log10 likelihood")
    return left, right


def avg_grid(raw, grid_w, keep=None):
    """eqtlbma_avg_bfs without configuration weights: each row of raw
    [n_rows][L] averaged over the kept grid points (default: all)."""
    raw = np.ascontiguousarray(raw, dtype=np.float64)
    L = raw.shape[1]
    gw = np.ascontiguousarray(grid_w, dtype=np.float64)
    kp = np.ascontiguousarray(np.arange(L) if keep is None else keep, dtype=np.int32)
    out = np.zeros(raw.shape[0])
    lib().orc_avg_grid(_p(raw), raw.shape[0], L, _p(gw), _p(kp), kp.shape[0], _p(out))
    return out


AVG_POST = {"a": 1, "b": 2, "c": 4, "d": 8, "dims": 16}


def avg_bfs(raw, off, grid_w, model="configs", cfg_w=None, cfg_sub=None, type_w=None, sub_w=None, pi0=np.nan,
            want=(), keep=None, n_subgroups=None):
    """eqtlbma_avg_bfs averaging over the grid and the configurations or
    types, with the posteriors named in want ('a', 'b', 'c', 'd', 'dims').
    raw [n_pairs][R][L] (R = configurations or subgroups); off [G + 1].
    Returns a dict of arrays."""
    raw = np.ascontiguousarray(raw, dtype=np.float64)
    NP, R, L = raw.shape
    off = np.ascontiguousarray(off, dtype=np.int64)
    G = off.shape[0] - 1
    types = model == "types"
    gw = np.ascontiguousarray(grid_w, dtype=np.float64)
    kp = np.ascontiguousarray(np.arange(L) if keep is None else keep, dtype=np.int32)
    if types:
        tw = np.ascontiguousarray(type_w, dtype=np.float64)
        sw = np.ascontiguousarray(sub_w, dtype=np.float64)
        D, S = tw.shape[0], R
        cw, cs = np.zeros(D), np.zeros(D, np.uint64)
    else:
        cw = np.ascontiguousarray(cfg_w, dtype=np.float64)
        D = R
        S = n_subgroups if n_subgroups is not None else 1
        cs = np.ascontiguousarray(cfg_sub if cfg_sub is not None else np.zeros(D), dtype=np.uint64)
        tw, sw = np.zeros(1), np.zeros(1)
    bits = 0
    for k in want:
        bits |= AVG_POST[k]
    o = dict(gene_bf=np.full(G, np.nan), gene_post=np.full(G, np.nan), snp_bf=np.full(NP, np.nan),
             dim_bf=np.full((NP, D), np.nan), post_the=np.full(NP, np.nan), post_an=np.full(NP, np.nan),
             post_dim=np.full((NP, D), np.nan), post_sub=np.full((NP, S), np.nan),
             best_dim=np.full(NP, -1, np.int32))
    lib().orc_avg_bfs(_p(raw), _p(off), G, int(types), D, S, L, _p(kp), kp.shape[0], _p(gw), _p(cw), _p(cs),
                      _p(tw), _p(sw), float(pi0), bits, _p(o["gene_bf"]), _p(o["gene_post"]), _p(o["snp_bf"]),
                      _p(o["dim_bf"]), _p(o["post_the"]), _p(o["post_an"]), _p(o["post_dim"]), _p(o["post_sub"]),
                      _p(o["best_dim"]))
    return o
