"""ctypes mirror of include/eqbf.h (structures, enums and the chunk builder).

The C ABI is the product boundary; this module only lays numpy buffers out
the way the header describes so that Python callers (tests, the benchmark)
can hand them to the shared library.
"""
import ctypes as C

import numpy as np

EQBF_OK = 0
EQBF_ERR_ARG = 1
EQBF_ERR_HIP = 2
EQBF_ERR_DATA = 3
EQBF_ERR_UNSUPPORTED = 4

ANALYSIS = {"sep": 0, "join": 1}
BFS = {"gen": 0, "sin": 1, "all": 2}
PBF = {"gen": 0, "gen-sin": 1, "all": 2}
PREC = {"fp64_plain": 0, "plain": 0, "mixed": 1, "fp64_batched": 2, "fp64": 2}
F32, F64 = 0, 1
F32_TILED = 2  # raw ABFs only: FP32 in tiles of 64 pairs


def untile_raw(flat, n_raw, n_pairs):
    """[n_raw][n_pairs] view of raw ABFs written as EQBF_F32_TILED: element
    (row, p) at (p & ~63) * n_raw + row * w + (p & 63), w = min(64, n_pairs -
    (p & ~63))."""
    flat = np.asarray(flat).ravel()
    out = np.empty((n_raw, n_pairs), dtype=flat.dtype)
    for tb in range(0, n_pairs, 64):
        w = min(64, n_pairs - tb)
        out[:, tb:tb + w] = flat[tb * n_raw:tb * n_raw + n_raw * w].reshape(n_raw, w)
    return out


class Chunk(C.Structure):
    _fields_ = [
        ("n_genes", C.c_int32),
        ("n_subgroups", C.c_int32),
        ("n_samples", C.c_int32),
        ("n_covars_max", C.c_int32),
        ("n_pairs", C.c_int64),
        ("n_snps", C.c_int64),
        ("explevels", C.c_void_p),
        ("gene_in_sbgrp", C.c_void_p),
        ("geno_in_sbgrp", C.c_void_p),
        ("covar_count", C.c_void_p),
        ("covars", C.c_void_p),
        ("covar_in_sbgrp", C.c_void_p),
        ("n_geno_sets", C.c_int32),
        ("geno_set_of_sbgrp", C.c_void_p),
        ("genotypes", C.c_void_p),
        ("geno_dtype", C.c_int32),
        ("geno_ld", C.c_int64),
        ("geno_set_stride", C.c_int64),
        ("snp_in_sbgrp", C.c_void_p),
        ("pair_offset", C.c_void_p),
        ("pair_snp", C.c_void_p),
    ]


class AssocOpts(C.Structure):
    _fields_ = [
        ("analysis", C.c_int32),
        ("bfs", C.c_int32),
        ("precision", C.c_int32),
        ("raw_dtype", C.c_int32),
    ]


class AssocOut(C.Structure):
    _fields_ = [
        ("n", C.c_void_p),
        ("sstats", C.c_void_p),
        ("stdsstats", C.c_void_p),
        ("raw_abf", C.c_void_p),
        ("w_abf", C.c_void_p),
    ]


def n_configs(bfs, S):
    return {"gen": 0, "sin": S, "all": (1 << S) - 1}[bfs]


def config_masks(bfs, S):
    """gsl_combination order: subsets of size 1..S, lexicographic in a size."""
    from itertools import combinations
    if bfs == "gen":
        return []
    if bfs == "sin":
        return [1 << s for s in range(S)]
    out = []
    for k in range(1, S + 1):
        for c in combinations(range(S), k):
            m = 0
            for s in c:
                m |= 1 << s
            out.append(m)
    return out


def config_label(mask):
    return "-".join(str(s + 1) for s in range(64) if (mask >> s) & 1)


def _ptr(a):
    return None if a is None else a.ctypes.data


class HostChunk:
    """Numpy-backed chunk in the layout of include/eqbf.h.

    Arrays (all C-contiguous):
      explevels      float64 [G, S, N]
      gene_in_sbgrp  uint8   [G, S]
      geno_in_sbgrp  uint8   [S, N]
      covar_count    int32   [S]
      covars         float64 [S, Qmax, N]
      covar_in_sbgrp uint8   [S, N]
      geno_set_of_sbgrp int32 [S]
      genotypes      float32|float64 [n_sets, N, n_snps]
      snp_in_sbgrp   uint8   [n_snps, S]
      pair_offset    int64   [G+1]
      pair_snp       int32   [n_pairs]
    """

    def __init__(self, explevels, gene_in_sbgrp, geno_in_sbgrp, genotypes,
                 snp_in_sbgrp, pair_offset, pair_snp, covars=None,
                 covar_count=None, covar_in_sbgrp=None, geno_set_of_sbgrp=None):
        G, S, N = explevels.shape
        self.explevels = np.ascontiguousarray(explevels, dtype=np.float64)
        self.gene_in_sbgrp = np.ascontiguousarray(gene_in_sbgrp, dtype=np.uint8)
        self.geno_in_sbgrp = np.ascontiguousarray(geno_in_sbgrp, dtype=np.uint8)
        if genotypes.ndim == 2:
            genotypes = genotypes[None]
        assert genotypes.dtype in (np.float32, np.float64)
        self.genotypes = np.ascontiguousarray(genotypes)
        self.snp_in_sbgrp = np.ascontiguousarray(snp_in_sbgrp, dtype=np.uint8)
        self.pair_offset = np.ascontiguousarray(pair_offset, dtype=np.int64)
        self.pair_snp = np.ascontiguousarray(pair_snp, dtype=np.int32)
        if covars is None:
            covars = np.zeros((S, 1, N))
            covar_count = np.zeros(S, np.int32)
            covar_in_sbgrp = np.ones((S, N), np.uint8)
        self.covars = np.ascontiguousarray(covars, dtype=np.float64)
        self.covar_count = np.ascontiguousarray(covar_count, dtype=np.int32)
        self.covar_in_sbgrp = np.ascontiguousarray(covar_in_sbgrp, dtype=np.uint8)
        if geno_set_of_sbgrp is None:
            geno_set_of_sbgrp = np.zeros(S, np.int32)
        self.geno_set_of_sbgrp = np.ascontiguousarray(geno_set_of_sbgrp, dtype=np.int32)
        self.G, self.S, self.N = G, S, N
        self.n_pairs = int(self.pair_snp.shape[0])
        self.n_snps = int(self.genotypes.shape[2])
        self.Qmax = int(self.covars.shape[1])

    def struct(self):
        c = Chunk()
        c.n_genes, c.n_subgroups, c.n_samples = self.G, self.S, self.N
        c.n_covars_max = self.Qmax
        c.n_pairs, c.n_snps = self.n_pairs, self.n_snps
        c.explevels = _ptr(self.explevels)
        c.gene_in_sbgrp = _ptr(self.gene_in_sbgrp)
        c.geno_in_sbgrp = _ptr(self.geno_in_sbgrp)
        c.covar_count = _ptr(self.covar_count)
        c.covars = _ptr(self.covars)
        c.covar_in_sbgrp = _ptr(self.covar_in_sbgrp)
        c.n_geno_sets = self.genotypes.shape[0]
        c.geno_set_of_sbgrp = _ptr(self.geno_set_of_sbgrp)
        c.genotypes = _ptr(self.genotypes)
        c.geno_dtype = F32 if self.genotypes.dtype == np.float32 else F64
        c.geno_ld = self.n_snps
        c.geno_set_stride = self.N * self.n_snps
        c.snp_in_sbgrp = _ptr(self.snp_in_sbgrp)
        c.pair_offset = _ptr(self.pair_offset)
        c.pair_snp = _ptr(self.pair_snp)
        return c


class HostAssocOut:
    """Output buffers of one association call (structure of arrays)."""

    def __init__(self, S, n_pairs, bfs, GL, GS, raw_dtype=np.float64,
                 want_sstats=True, want_std=True, want_raw=True):
        nc = n_configs(bfs, S)
        self.n_raw = 3 * GL + nc * GS
        self.n_w = 3 + nc + 2
        self.n = np.zeros((S, n_pairs), np.int32)
        self.sstats = np.full((5, S, n_pairs), np.nan) if want_sstats else None
        self.stdsstats = np.full((3, S, n_pairs), np.nan) if want_std else None
        self.raw = np.full((self.n_raw, n_pairs), np.nan, dtype=raw_dtype) if want_raw else None
        self.w = np.full((self.n_w, n_pairs), np.nan)

    def struct(self):
        o = AssocOut()
        o.n = _ptr(self.n)
        o.sstats = _ptr(self.sstats)
        o.stdsstats = _ptr(self.stdsstats)
        o.raw_abf = _ptr(self.raw)
        o.w_abf = _ptr(self.w)
        return o
