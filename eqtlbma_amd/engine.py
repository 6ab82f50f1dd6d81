"""Python binding over the C ABI of libeqbf_hip.so (include/eqbf.h).

The library must have been built (python -m eqtlbma_amd.build); there is no
fallback: a missing or unloadable library raises immediately.
"""
import ctypes as C
import os

import numpy as np

from . import LIB_PATH, _abi

_lib = None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("eqbf error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libeqbf_hip.so is not built (%s); run python -m eqtlbma_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.eqbf_create.argtypes = [C.POINTER(vp), C.c_int]
        L.eqbf_destroy.argtypes = [vp]
        L.eqbf_last_error.argtypes = [vp]
        L.eqbf_last_error.restype = C.c_char_p
        L.eqbf_sync.argtypes = [vp]
        L.eqbf_set_grids.argtypes = [vp, vp, vp, i32, vp, vp, i32]
        L.eqbf_n_configs.argtypes = [i32, i32]
        L.eqbf_n_configs.restype = i64
        L.eqbf_config_masks.argtypes = [i32, i32, vp]
        L.eqbf_n_raw.argtypes = [i32, i32, i32, i32]
        L.eqbf_n_raw.restype = i64
        L.eqbf_n_weighted.argtypes = [i32, i32]
        L.eqbf_n_weighted.restype = i64
        L.eqbf_assoc_uvlr.argtypes = [vp, vp, vp, vp]
        L.eqbf_assoc_uvlr_device.argtypes = [vp, vp, vp, vp, vp]
        L.eqbf_assoc_uvlr_device_noalloc.argtypes = [vp, vp, vp, vp, vp]
        L.eqbf_assoc_mvlr.argtypes = [vp, vp, vp, vp]
        L.eqbf_assoc_mvlr_device.argtypes = [vp, vp, vp, vp, vp]
        L.eqbf_assoc_hybrid.argtypes = [vp, vp, vp, vp]
        L.eqbf_avg_bfs.argtypes = [vp, vp, i64, vp, i64, vp, vp]
        L.eqbf_avg_bfs_device.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp]
        L.eqbf_avg_grid.argtypes = [vp, vp, i64, i32, vp, vp, i32, vp]
        L.eqbf_assoc_hybrid_device.argtypes = [vp, vp, vp, vp, vp]
        L.eqbf_abf_from_sstats.argtypes = [vp, vp, i32, i32, vp]
        L.eqbf_abf_from_sstats_device.argtypes = [vp, vp, i32, i32, vp, vp]
        L.eqbf_debug_ttail.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp]
        L.eqbf_perm_stats.argtypes = [vp, vp, vp, vp, i32, vp, i64, vp]
        L.eqbf_perm_stats_device.argtypes = [vp, vp, vp, vp, i32, vp, i64, vp, vp]
        L.eqbf_perm_stats_fp64_ties.argtypes = [vp, vp, vp, vp, i32, vp, i64, C.c_double, vp, vp, vp]
        L.eqbf_perm_stats_fp64_ties_device.argtypes = [vp, vp, vp, vp, i32, vp, i64, C.c_double, vp, vp, vp, vp]
        L.eqbf_rng_seed.argtypes = [vp, C.c_uint64]
        L.eqbf_rng_get.argtypes = [vp]
        L.eqbf_rng_get.restype = C.c_uint32
        L.eqbf_rng_uniform_int.argtypes = [vp, C.c_uint32]
        L.eqbf_rng_uniform_int.restype = C.c_uint32
        L.eqbf_rng_flat.argtypes = [vp, C.c_double, C.c_double]
        L.eqbf_rng_flat.restype = C.c_double
        L.eqbf_rng_shuffle.argtypes = [vp, vp, i64]
        L.eqbf_perm_tables.argtypes = [vp, i64, i64, i64, vp]
        L.eqbf_perm_pvalue.argtypes = [vp, C.c_double, i32, vp, i32, i64, vp, vp, vp, vp]
        L.eqbf_profile.argtypes = [vp, C.c_int]
        L.eqbf_profile_read.argtypes = [vp, vp, vp, vp, i32, vp]
        L.eqbf_mlp_create.argtypes = [vp, i32, i32, i32, i32, C.c_uint64, C.POINTER(vp)]
        L.eqbf_mlp_destroy.argtypes = [vp]
        L.eqbf_mlp_n_params.argtypes = [vp]
        L.eqbf_mlp_n_params.restype = i64
        L.eqbf_mlp_set_params.argtypes = [vp, vp, vp]
        L.eqbf_mlp_get_params.argtypes = [vp, vp, vp]
        L.eqbf_mlp_forward.argtypes = [vp, vp, vp, i64, vp]
        L.eqbf_mlp_loss_grad.argtypes = [vp, vp, vp, vp, vp, i64, vp, vp]
        L.eqbf_mlp_train.argtypes = [vp, vp, vp, vp, vp, i64, i32, i32, C.c_float, C.c_uint64, vp]
        L.eqbf_set_graphs.argtypes = [vp, C.c_int]
        L.eqbf_set_autotune.argtypes = [vp, C.c_int]
        L.eqbf_graph_stats.argtypes = [vp, vp, vp]
        L.eqbf_autotune_read.argtypes = [vp, vp, vp, vp, vp, vp, i32, vp]
        L.eqbf_ld_eigenvalues.argtypes = [vp, vp, i32, vp]
        L.eqbf_ld_eigenvalues_device.argtypes = [vp, vp, i32, vp]
        L.eqbf_gev_fit.argtypes = [vp, vp, i64, i64, vp]
        L.eqbf_hm_em.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp, vp]
        L.eqbf_hm_em_device.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp, vp]
        L.eqbf_hm_bfs.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp, vp, vp, vp]
        L.eqbf_hm_ci.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp, vp]
        L.eqbf_perm_pvalue.restype = i64
        _lib = L
    return _lib


class PermOpts(C.Structure):
    _fields_ = [("analysis", C.c_int32), ("pbf", C.c_int32), ("use_max_bf", C.c_int32),
                ("sep_sbgrp", C.c_int32), ("precision", C.c_int32), ("error_model", C.c_int32),
                ("fiterr", C.c_double), ("covars_by_position", C.c_void_p), ("explevels_pairs", C.c_void_p)]


class MvlrOpts(C.Structure):
    _fields_ = [("bfs", C.c_int32), ("fiterr", C.c_double)]


class HybridOpts(C.Structure):
    """eqbf_hybrid_opts: bfs, fiterr, pairwise covariates by sample index
    [S][Qmax][N] and pairwise expression levels [G][S][N] (NULL: the
    chunk's)."""
    _fields_ = [("bfs", C.c_int32), ("fiterr", C.c_double), ("covars_by_position", C.c_void_p),
                ("explevels_pairs", C.c_void_p)]


class AvgOpts(C.Structure):
    """eqbf_avg_opts (eqtlbma_avg_bfs)."""
    _fields_ = [("model", C.c_int32), ("dim", C.c_int32), ("n_subgroups", C.c_int32), ("n_grid", C.c_int32),
                ("n_keep", C.c_int32), ("keep", C.c_void_p), ("grid_weights", C.c_void_p),
                ("dim_weights", C.c_void_p), ("cfg_subgroups", C.c_void_p), ("sub_weights", C.c_void_p),
                ("pi0", C.c_double), ("want", C.c_int32)]


class AvgOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("gene_bf", "gene_post", "snp_bf", "dim_bf", "post_the", "post_an",
                                         "post_dim", "post_sub", "best_dim")]


AVG_POST = {"a": 1, "b": 2, "c": 4, "d": 8, "dims": 16}


class HmOpts(C.Structure):
    """eqbf_hm_opts: model (0 configs, 1 types), its dimension, subgroups,
    grid size, stopping threshold, iteration limit, SQUAREM step bound (1:
    classic EM) and the fixed flags (pi0, grid, configs, types, subgroups)."""
    _fields_ = [("model", C.c_int32), ("dim", C.c_int32), ("n_subgroups", C.c_int32), ("n_grid", C.c_int32),
                ("thresh", C.c_double), ("max_iters", C.c_int64), ("stepmax", C.c_double),
                ("fixed", C.c_int32 * 5)]


class HmParamsC(C.Structure):
    _fields_ = [("pi0", C.c_double), ("grid", C.c_void_p), ("config", C.c_void_p), ("type", C.c_void_p),
                ("subgroup", C.c_void_p)]


class SstatsIn(C.Structure):
    """eqbf_sstats_in: given summary statistics, arrays [S][n_pairs]."""
    _fields_ = [("n_pairs", C.c_int64), ("n_subgroups", C.c_int32), ("n", C.c_void_p), ("sigmahat", C.c_void_p),
                ("betahat", C.c_void_p), ("sebetahat", C.c_void_p)]


class Rng:
    """GSL-semantics mt19937 of the library (host side, no GPU needed)."""

    def __init__(self, seed):
        self._st = (C.c_uint32 * 625)()
        lib().eqbf_rng_seed(C.addressof(self._st), seed)

    @property
    def addr(self):
        return C.addressof(self._st)

    def get(self):
        return lib().eqbf_rng_get(self.addr)

    def uniform_int(self, n):
        return lib().eqbf_rng_uniform_int(self.addr, n)

    def flat(self, a, b):
        return lib().eqbf_rng_flat(self.addr, a, b)

    def shuffle(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        lib().eqbf_rng_shuffle(self.addr, a.ctypes.data, a.size)
        return a

    def tables(self, n_samples, n_genes, nperm):
        t = np.zeros((n_genes, nperm, n_samples), np.int32)
        lib().eqbf_perm_tables(self.addr, n_samples, n_genes, nperm, t.ctypes.data)
        return t


def perm_pvalue(stats, true_stat, nb_permutations, sep=False, trick=0, trick_cutoff=10, rng_trick=None):
    """Replays the reference's counting loop; returns (pval, nperm_used,
    median, nb_permutations_after, drawn)."""
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    nb = C.c_int64(nb_permutations)
    pv, used, med = C.c_double(0), C.c_int64(0), C.c_double(0)
    drawn = lib().eqbf_perm_pvalue(stats.ctypes.data, float(true_stat), int(sep), C.addressof(nb), trick,
                                   trick_cutoff, rng_trick.addr if rng_trick is not None else None,
                                   C.addressof(pv), C.addressof(used), C.addressof(med))
    return pv.value, used.value, (None if sep else med.value), nb.value, drawn


def _p(a):
    return None if a is None else a.ctypes.data


class Engine:
    """One context on one HIP device."""

    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        rc = lib().eqbf_create(C.byref(self._ctx), device)
        if rc != 0:
            raise EngineError(rc, "eqbf_create(device=%d) failed" % device)
        self.GL = self.GS = 0

    def close(self):
        if self._ctx:
            lib().eqbf_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise EngineError(rc, lib().eqbf_last_error(self._ctx).decode())

    def set_grids(self, gridL, gridS=None):
        gL = np.ascontiguousarray(gridL, dtype=np.float64)
        gS = np.ascontiguousarray(gridS if gridS is not None else np.zeros((0, 2)), dtype=np.float64)
        self._gL = (np.ascontiguousarray(gL[:, 0]), np.ascontiguousarray(gL[:, 1]))
        self._gS = (np.ascontiguousarray(gS[:, 0]), np.ascontiguousarray(gS[:, 1]))
        self._check(lib().eqbf_set_grids(self._ctx, _p(self._gL[0]), _p(self._gL[1]), gL.shape[0],
                                         _p(self._gS[0]), _p(self._gS[1]), gS.shape[0]))
        self.GL, self.GS = gL.shape[0], gS.shape[0]

    def assoc_uvlr(self, hc, analysis="join", bfs="gen", precision="mixed", raw_dtype=np.float64,
                   want_sstats=True, want_std=True, want_raw=True, raw_tiled=False):
        """raw_tiled: FP32 raw ABFs written in 64-pair tiles (EQBF_F32_TILED),
        returned rearranged to [n_raw][n_pairs]."""
        if raw_tiled:
            raw_dtype = np.float32
        out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, self.GL, self.GS, raw_dtype=raw_dtype,
                                want_sstats=want_sstats, want_std=want_std, want_raw=want_raw)
        code = _abi.F32_TILED if raw_tiled else (_abi.F64 if raw_dtype == np.float64 else _abi.F32)
        opts = _abi.AssocOpts(_abi.ANALYSIS[analysis], _abi.BFS[bfs], _abi.PREC[precision], code)
        st = hc.struct()
        o = out.struct()
        self._check(lib().eqbf_assoc_uvlr(self._ctx, C.byref(st), C.byref(opts), C.byref(o)))
        if raw_tiled and out.raw is not None:
            out.raw = _abi.untile_raw(out.raw, out.n_raw, hc.n_pairs)
        return out

    def assoc_uvlr_device(self, chunk_struct, opts, out_struct, stream=None):
        """Device-pointer entry: chunk_struct / out_struct are _abi.Chunk /
        _abi.AssocOut whose pointers are device addresses."""
        self._check(lib().eqbf_assoc_uvlr_device(self._ctx, C.byref(chunk_struct), C.byref(opts),
                                                 C.byref(out_struct), stream))

    def assoc_uvlr_device_noalloc(self, chunk_struct, opts, out_struct, stream=None):
        """eqbf_assoc_uvlr_device_noalloc: device-pointer entry, FP64
        precisions, scratch from the library's static array (no device
        allocation)."""
        self._check(lib().eqbf_assoc_uvlr_device_noalloc(self._ctx, C.byref(chunk_struct), C.byref(opts),
                                                         C.byref(out_struct), stream))

    def assoc_mvlr(self, hc, bfs="gen", fiterr=0.0, want_sstats=False):
        """--error mvlr Bayes factors (FP64); raw log10 ABFs are FP64."""
        out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, self.GL, self.GS, raw_dtype=np.float64,
                                want_sstats=want_sstats, want_std=want_sstats, want_raw=True)
        opts = MvlrOpts(_abi.BFS[bfs], float(fiterr))
        st = hc.struct()
        o = out.struct()
        self._check(lib().eqbf_assoc_mvlr(self._ctx, C.byref(st), C.byref(opts), C.byref(o)))
        return out

    def assoc_hybrid(self, hc, bfs="gen", fiterr=0.0, covars_pos=None, expl_pair=None, want_sstats=True):
        """--error hybrid Bayes factors (FP64), with the sample sizes and
        summary statistics of the separate analysis."""
        out = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, self.GL, self.GS, raw_dtype=np.float64,
                                want_sstats=want_sstats, want_std=False, want_raw=True)
        cp = None if covars_pos is None else np.ascontiguousarray(covars_pos, dtype=np.float64)
        ep = None if expl_pair is None else np.ascontiguousarray(expl_pair, dtype=np.float64)
        if cp is not None:
            assert cp.shape == hc.covars.shape
        if ep is not None:
            assert ep.shape == hc.explevels.shape
        opts = HybridOpts(_abi.BFS[bfs], float(fiterr), None if cp is None else cp.ctypes.data,
                          None if ep is None else ep.ctypes.data)
        st = hc.struct()
        o = out.struct()
        self._check(lib().eqbf_assoc_hybrid(self._ctx, C.byref(st), C.byref(opts), C.byref(o)))
        return out

    def avg_grid(self, raw, grid_w, keep=None):
        """eqtlbma_avg_bfs without configuration weights: each raw row
        [n_rows][L] averaged over the kept grid points."""
        raw = np.ascontiguousarray(raw, dtype=np.float64)
        L = raw.shape[1]
        gw = np.ascontiguousarray(grid_w, dtype=np.float64)
        kp = np.ascontiguousarray(np.arange(L) if keep is None else keep, dtype=np.int32)
        out = np.zeros(raw.shape[0])
        self._check(lib().eqbf_avg_grid(self._ctx, raw.ctypes.data, raw.shape[0], L, gw.ctypes.data,
                                        kp.ctypes.data, kp.shape[0], out.ctypes.data))
        return out

    def avg_bfs(self, raw, off, grid_w, model="configs", cfg_w=None, cfg_sub=None, type_w=None, sub_w=None,
                pi0=float("nan"), want=(), keep=None, n_subgroups=None):
        """eqtlbma_avg_bfs averaging over the grid and the configurations
        or types, with the posteriors named in want; raw [n_pairs][R][L].
        Returns a dict of arrays (the layout of the oracle's avg_bfs)."""
        raw = np.ascontiguousarray(raw, dtype=np.float64)
        NP, R, L = raw.shape
        off = np.ascontiguousarray(off, dtype=np.int64)
        G = off.shape[0] - 1
        types = model == "types"
        gw = np.ascontiguousarray(grid_w, dtype=np.float64)
        kp = np.ascontiguousarray(np.arange(L) if keep is None else keep, dtype=np.int32)
        if types:
            dw = np.ascontiguousarray(type_w, dtype=np.float64)
            sw = np.ascontiguousarray(sub_w, dtype=np.float64)
            D, S, cs = dw.shape[0], R, None
        else:
            dw = np.ascontiguousarray(cfg_w, dtype=np.float64)
            D, sw = R, None
            S = n_subgroups if n_subgroups is not None else 1
            cs = np.ascontiguousarray(cfg_sub if cfg_sub is not None else np.zeros(D), dtype=np.uint64)
        bits = 0
        for k in want:
            bits |= AVG_POST[k]
        o = dict(gene_bf=np.full(G, np.nan), gene_post=np.full(G, np.nan), snp_bf=np.full(NP, np.nan),
                 dim_bf=np.full((NP, D), np.nan), post_the=np.full(NP, np.nan), post_an=np.full(NP, np.nan),
                 post_dim=np.full((NP, D), np.nan), post_sub=np.full((NP, S), np.nan),
                 best_dim=np.full(NP, -1, np.int32))
        opts = AvgOpts(int(types), D, S, L, kp.shape[0], kp.ctypes.data, gw.ctypes.data, dw.ctypes.data,
                       None if cs is None else cs.ctypes.data, None if sw is None else sw.ctypes.data, float(pi0),
                       bits)
        ao = AvgOut(*[o[n].ctypes.data for n in ("gene_bf", "gene_post", "snp_bf", "dim_bf", "post_the", "post_an",
                                                  "post_dim", "post_sub", "best_dim")])
        self._check(lib().eqbf_avg_bfs(self._ctx, raw.ctypes.data, NP, off.ctypes.data, G, C.byref(opts),
                                       C.byref(ao)))
        return o

    def avg_bfs_device(self, raw_dev, n_pairs, off_dev, n_genes, opts, out_struct, stream=None):
        self._check(lib().eqbf_avg_bfs_device(self._ctx, raw_dev, n_pairs, off_dev, n_genes, C.byref(opts),
                                              C.byref(out_struct), stream))

    def assoc_hybrid_device(self, chunk_struct, opts, out_struct, stream=None):
        self._check(lib().eqbf_assoc_hybrid_device(self._ctx, C.byref(chunk_struct), C.byref(opts),
                                                   C.byref(out_struct), stream))

    def assoc_mvlr_device(self, chunk_struct, opts, out_struct, stream=None):
        self._check(lib().eqbf_assoc_mvlr_device(self._ctx, C.byref(chunk_struct), C.byref(opts),
                                                 C.byref(out_struct), stream))

    def abf_from_sstats(self, n, sigmahat, betahat, sebetahat, bfs="gen", raw_dtype=np.float64, want_std=True,
                        want_raw=True):
        """--inss: Bayes factors from given summary statistics, arrays
        [S][n_pairs] (n <= 0: no result in that subgroup); FP64 throughout.
        Returns a HostAssocOut (sstats not filled)."""
        n = np.ascontiguousarray(n, dtype=np.int64)
        S, NP = n.shape
        sig, beta, se = (np.ascontiguousarray(a, dtype=np.float64) for a in (sigmahat, betahat, sebetahat))
        if not (sig.shape == beta.shape == se.shape == (S, NP)):
            raise ValueError("summary statistics must all be [S][n_pairs]")
        out = _abi.HostAssocOut(S, NP, bfs, self.GL, self.GS, raw_dtype=raw_dtype, want_sstats=False,
                                want_std=want_std, want_raw=want_raw)
        si = SstatsIn(NP, S, _p(n), _p(sig), _p(beta), _p(se))
        o = out.struct()
        self._check(lib().eqbf_abf_from_sstats(self._ctx, C.byref(si), _abi.BFS[bfs],
                                               _abi.F64 if raw_dtype == np.float64 else _abi.F32, C.byref(o)))
        return out

    def abf_from_sstats_device(self, sstats_struct, bfs, raw_dtype, out_struct, stream=None):
        """Device-pointer entry: sstats_struct (SstatsIn) and out_struct
        (_abi.AssocOut) hold device addresses."""
        self._check(lib().eqbf_abf_from_sstats_device(self._ctx, C.byref(sstats_struct), _abi.BFS[bfs], raw_dtype,
                                                      C.byref(out_struct), stream))

    def perm_stats(self, hc, tables, gene_table, analysis="join", pbf="all", use_max_bf=True, sep_sbgrp=-1,
                   precision="mixed", error_model="uvlr", fiterr=0.5, covars_pos=None, expl_pair=None):
        """Statistic of every permutation: tables [n_tables][nperm][N],
        gene_table [n_genes] (-1 skips a gene); returns [n_genes][nperm].
        error_model 'mvlr' / 'hybrid': MVLR / hybrid Bayes factors (join);
        covars_pos, expl_pair: the hybrid model's pairwise inputs."""
        tables = np.ascontiguousarray(tables, dtype=np.int32)
        gt = np.ascontiguousarray(gene_table, dtype=np.int32)
        nperm = tables.shape[1]
        out = np.full((hc.G, nperm), np.nan)
        cp = None if covars_pos is None else np.ascontiguousarray(covars_pos, dtype=np.float64)
        ep = None if expl_pair is None else np.ascontiguousarray(expl_pair, dtype=np.float64)
        po = PermOpts(_abi.ANALYSIS[analysis], _abi.PBF[pbf], int(use_max_bf), sep_sbgrp, _abi.PREC[precision],
                      {"uvlr": 0, "mvlr": 1, "hybrid": 2}[error_model], float(fiterr),
                      None if cp is None else cp.ctypes.data, None if ep is None else ep.ctypes.data)
        st = hc.struct()
        self._check(lib().eqbf_perm_stats(self._ctx, C.byref(st), C.byref(po), _p(tables), tables.shape[0],
                                          _p(gt), nperm, _p(out)))
        return out

    def perm_stats_fp64_ties(self, hc, tables, gene_table, true_stats, analysis="join", pbf="all",
                             use_max_bf=True, sep_sbgrp=-1, precision="mixed", tie_bound=5e-5):
        """perm_stats, then for every gene with a permutation statistic within
        tie_bound of true_stats[g] (absolute for join, relative for sep): its
        true statistic and those permutation statistics in FP64.  Returns
        (stats [n_genes][nperm], true_stats [n_genes], number re-computed)."""
        tables = np.ascontiguousarray(tables, dtype=np.int32)
        gt = np.ascontiguousarray(gene_table, dtype=np.int32)
        nperm = tables.shape[1]
        out = np.full((hc.G, nperm), np.nan)
        true = np.array(true_stats, dtype=np.float64, copy=True)
        assert true.shape == (hc.G,)
        n_re = C.c_int64(0)
        po = PermOpts(_abi.ANALYSIS[analysis], _abi.PBF[pbf], int(use_max_bf), sep_sbgrp, _abi.PREC[precision])
        st = hc.struct()
        self._check(lib().eqbf_perm_stats_fp64_ties(self._ctx, C.byref(st), C.byref(po), _p(tables), tables.shape[0],
                                                    _p(gt), nperm, float(tie_bound), _p(out), _p(true),
                                                    C.byref(n_re)))
        return out, true, n_re.value

    def perm_stats_fp64_ties_device(self, chunk_struct, opts, tables_ptr, n_tables, gene_table, nperm, tie_bound,
                                    stats_ptr, true_ptr, stream=None):
        """Device-pointer entry of perm_stats_fp64_ties (true_ptr: the
        caller's true statistics, updated in place); returns the number of
        statistics re-computed."""
        gt = np.ascontiguousarray(gene_table, dtype=np.int32)
        n_re = C.c_int64(0)
        self._check(lib().eqbf_perm_stats_fp64_ties_device(self._ctx, C.byref(chunk_struct), C.byref(opts),
                                                           tables_ptr, int(n_tables), _p(gt), int(nperm),
                                                           float(tie_bound), stats_ptr, true_ptr, C.byref(n_re),
                                                           stream))
        return n_re.value

    def perm_stats_device(self, chunk_struct, opts, tables_ptr, n_tables, gene_table, nperm, stats_ptr, stream=None):
        """Device-pointer entry: tables_ptr [n_tables][nperm][N] int32 and
        stats_ptr [n_genes][nperm] float64 are device addresses; gene_table
        is a host int32 array."""
        gt = np.ascontiguousarray(gene_table, dtype=np.int32)
        self._check(lib().eqbf_perm_stats_device(self._ctx, C.byref(chunk_struct), C.byref(opts), tables_ptr,
                                                 int(n_tables), _p(gt), int(nperm), stats_ptr, stream))

    def ld_eigenvalues(self, hc, n_eig=20):
        """[G][n_eig] largest eigenvalues of each gene's cis LD matrix."""
        out = np.zeros((hc.G, n_eig), np.float32)
        st = hc.struct()
        self._check(lib().eqbf_ld_eigenvalues(self._ctx, C.byref(st), n_eig, _p(out)))
        return out

    def ld_eigenvalues_device(self, chunk_struct, n_genes, n_eig=20):
        out = np.zeros((n_genes, n_eig), np.float32)
        self._check(lib().eqbf_ld_eigenvalues_device(self._ctx, C.byref(chunk_struct), n_eig, _p(out)))
        return out

    def gev_fit(self, stats):
        """[G][3] (mu, sigma, xi) of each row of permutation statistics."""
        st = np.ascontiguousarray(stats, dtype=np.float64)
        out = np.zeros((st.shape[0], 3))
        self._check(lib().eqbf_gev_fit(self._ctx, _p(st), st.shape[0], st.shape[1], _p(out)))
        return out

    def _hm_setup(self, raw, off, pi0, grid, cfg, type_, sub, thresh, maxit, stepmax, fixed):
        raw = np.ascontiguousarray(raw, dtype=np.float64)
        off = np.ascontiguousarray(off, dtype=np.int64)
        types = type_ is not None
        o = HmOpts()
        o.model = 1 if types else 0
        o.n_grid = raw.shape[2]
        if types:
            sub = np.ascontiguousarray(sub, dtype=np.float64).copy()
            o.dim, o.n_subgroups = sub.shape
            dimw = np.ascontiguousarray(type_, dtype=np.float64).copy()
        else:
            o.dim, o.n_subgroups = raw.shape[1], 0
            dimw = np.ascontiguousarray(cfg, dtype=np.float64).copy()
        o.thresh, o.max_iters, o.stepmax = float(thresh), int(maxit), float(stepmax)
        for i in range(5):
            o.fixed[i] = int(fixed[i])
        grid = np.ascontiguousarray(grid, dtype=np.float64).copy()
        p = HmParamsC()
        p.pi0 = float(pi0)
        p.grid = grid.ctypes.data
        if types:
            p.type, p.subgroup = dimw.ctypes.data, sub.ctypes.data
        else:
            p.config = dimw.ctypes.data
        return raw, off, o, p, grid, dimw, sub

    def hm_em(self, raw, off, pi0, grid, cfg=None, type_=None, sub=None, thresh=0.05, maxit=0, stepmax=1.0,
              fixed=(0, 0, 0, 0, 0)):
        """EM of eqtlbma_hm's model on raw log10 ABFs [pairs][D][L] grouped by
        gene (off: [G+1]).  Returns (pi0, grid, configuration or type
        weights, subgroup probabilities or None, log10 likelihood,
        iterations)."""
        raw, off, o, p, grid, dimw, sub = self._hm_setup(raw, off, pi0, grid, cfg, type_, sub, thresh, maxit,
                                                         stepmax, fixed)
        lik, it = C.c_double(), C.c_int64()
        self._check(lib().eqbf_hm_em(self._ctx, _p(raw), raw.shape[0], _p(off), off.size - 1, C.byref(o),
                                     C.byref(p), C.byref(lik), C.byref(it)))
        return p.pi0, grid, dimw, (sub if type_ is not None else None), lik.value, it.value

    def hm_em_device(self, raw_ptr, n_pairs, off_ptr, n_genes, opts, params):
        lik, it = C.c_double(), C.c_int64()
        self._check(lib().eqbf_hm_em_device(self._ctx, raw_ptr, int(n_pairs), off_ptr, int(n_genes), C.byref(opts),
                                            C.byref(params), C.byref(lik), C.byref(it)))
        return lik.value, it.value

    def hm_bfs(self, raw, off, pi0, grid, cfg=None, type_=None, sub=None):
        """(gene posterior, gene log10 BF, SNP log10 BF, [pairs][dim] log10
        BFs averaged over the grid) at the given parameters."""
        raw, off, o, p, grid, dimw, sub = self._hm_setup(raw, off, pi0, grid, cfg, type_, sub, 0.05, 0, 1.0,
                                                         (0,) * 5)
        G, NP = off.size - 1, raw.shape[0]
        gp, gb, sb, db = np.zeros(G), np.zeros(G), np.zeros(NP), np.zeros((NP, o.dim))
        self._check(lib().eqbf_hm_bfs(self._ctx, _p(raw), NP, _p(off), G, C.byref(o), C.byref(p), _p(gp), _p(gb),
                                      _p(sb), _p(db)))
        return gp, gb, sb, db

    def hm_ci(self, raw, off, pi0, grid, cfg=None, type_=None, sub=None):
        """--getci bounds (left, right): pi0, configs or types, subgroups per
        type (types model), grid points."""
        raw, off, o, p, grid, dimw, sub = self._hm_setup(raw, off, pi0, grid, cfg, type_, sub, 0.05, 0, 1.0,
                                                         (0,) * 5)
        n = 1 + o.dim + (o.dim * o.n_subgroups if type_ is not None else 0) + o.n_grid
        left, right = np.zeros(n), np.zeros(n)
        self._check(lib().eqbf_hm_ci(self._ctx, _p(raw), raw.shape[0], _p(off), off.size - 1, C.byref(o),
                                     C.byref(p), _p(left), _p(right)))
        return left, right

    def sync(self):
        self._check(lib().eqbf_sync(self._ctx))

    def set_graphs(self, enable=True):
        """hipGraph replay of repeated batched association calls."""
        self._check(lib().eqbf_set_graphs(self._ctx, int(enable)))

    def graph_stats(self):
        """(graphs held, graph launches so far)."""
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(lib().eqbf_graph_stats(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_autotune(self, enable=True):
        """Per-shape timing of the sums kernel's launch shapes."""
        self._check(lib().eqbf_set_autotune(self._ctx, int(enable)))

    def autotune_read(self, n_max=32):
        keys = C.create_string_buffer(64 * n_max)
        kch, spp, gm = (np.zeros(n_max, np.int32) for _ in range(3))
        ms = np.zeros(n_max, np.float32)
        n = C.c_int32(0)
        self._check(lib().eqbf_autotune_read(self._ctx, keys, _p(kch), _p(spp), _p(gm), _p(ms), n_max, C.byref(n)))
        raw = keys.raw
        return {raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode(): {"kch": int(kch[k]), "spp": int(spp[k]),
                                                                    "grid_mult": int(gm[k]), "ms": float(ms[k])}
                for k in range(n.value)}

    def profile(self, enable=True):
        """Turns the live per-kernel timing of this context on or off."""
        self._check(lib().eqbf_profile(self._ctx, int(enable)))

    def profile_read(self, n_max=64):
        """{kernel name: (total ms, launches)} since profiling was turned on."""
        names = C.create_string_buffer(64 * n_max)
        tot = np.zeros(n_max)
        cnt = np.zeros(n_max, np.int64)
        nk = C.c_int32(0)
        self._check(lib().eqbf_profile_read(self._ctx, names, _p(tot), _p(cnt), n_max, C.byref(nk)))
        raw = names.raw
        return {raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode(): (float(tot[k]), int(cnt[k]))
                for k in range(nk.value)}

    def debug_ttail(self, t, nu):
        t = np.ascontiguousarray(t, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.int32)
        n = t.size
        outs = [np.zeros(n) for _ in range(4)]
        self._check(lib().eqbf_debug_ttail(self._ctx, _p(t), _p(nu), n, *[_p(o) for o in outs]))
        return outs


class Mlp:
    """The allocator's network on the device (eqbf_mlp_* in include/eqbf.h):
    n_in -> n_hidden x n_layers (GELU) -> n_out, BF16 weights, FP32 master
    copy and accumulation."""

    def __init__(self, engine, n_in, n_hidden=256, n_layers=3, n_out=3, seed=1):
        self._eng = engine
        self._m = C.c_void_p()
        self.n_in, self.n_out = n_in, n_out
        engine._check(lib().eqbf_mlp_create(engine._ctx, n_in, n_hidden, n_layers, n_out, seed, C.byref(self._m)))
        self.n_params = lib().eqbf_mlp_n_params(self._m)

    def close(self):
        if self._m:
            lib().eqbf_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        self._eng._check(lib().eqbf_mlp_get_params(self._eng._ctx, self._m, _p(p)))
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (self.n_params,)
        self._eng._check(lib().eqbf_mlp_set_params(self._eng._ctx, self._m, _p(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._eng._check(lib().eqbf_mlp_forward(self._eng._ctx, self._m, _p(x), x.shape[0], _p(y)))
        return y

    def loss_grad(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double(0)
        self._eng._check(lib().eqbf_mlp_loss_grad(self._eng._ctx, self._m, _p(x), _p(t), _p(w), x.shape[0],
                                                   C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, t, w, epochs, batch=256, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        curve = np.zeros(max(epochs, 1), np.float32)
        self._eng._check(lib().eqbf_mlp_train(self._eng._ctx, self._m, _p(x), _p(t), _p(w), x.shape[0], epochs,
                                               batch, float(lr), seed, _p(curve)))
        return curve[:epochs]
