"""The association entry point that allocates nothing
(eqbf_assoc_uvlr_device_noalloc) against the existing device entry point
(eqbf_assoc_uvlr_device) with the FP64 batched precision.

Both run the same kernels on the same inputs, so every output must be
identical bit for bit; two calls in a row, the second on a smaller chunk
(the static scratch keeps the first call's contents), then the size and
precision limits, which must be refused before any kernel runs.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from eqtlbma_amd import _abi, engine, synth

pytestmark = pytest.mark.gpu

GL = synth.make_grid("general")
GS = synth.make_grid("small")
FIELDS = ("explevels", "gene_in_sbgrp", "geno_in_sbgrp", "covar_count", "covars", "covar_in_sbgrp",
          "geno_set_of_sbgrp", "genotypes", "snp_in_sbgrp", "pair_offset", "pair_snp")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Hip:
    """Device buffers through the HIP runtime the library uses."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        self.ptrs.append(p)
        if arr.nbytes:
            assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        assert self.L.hipDeviceSynchronize() == 0
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _outputs(hip, hc, bfs):
    host = _abi.HostAssocOut(hc.S, hc.n_pairs, bfs, GL.shape[0], GS.shape[0], raw_dtype=np.float64)
    host.n[:] = -1
    d = {k: hip.put(getattr(host, k)) for k in ("n", "sstats", "stdsstats", "raw", "w")}
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats, o.raw_abf, o.w_abf = (d[k].value for k in ("n", "sstats", "stdsstats", "raw", "w"))
    return host, d, o


def _read(hip, host, d):
    return {k: hip.get(d[k], getattr(host, k)) for k in ("n", "sstats", "stdsstats", "raw", "w")}


def _device_chunk(hip, hc):
    c = hc.struct()
    for f in FIELDS:
        setattr(c, f, hip.put(getattr(hc, f)).value)
    return c


@pytest.mark.parametrize("bfs", ["all", "sin"])
def test_noalloc_matches_device_entry_twice(bfs):
    cases = [synth.batched(seed=71, S=5, N_all=150, n_genes=6, p_mean=90, p_lo=30, p_hi=160, ns_lo=40, ns_hi=150,
                           Q=2, nan_frac=0.03, absent_frac=0.1),
             synth.batched(seed=72, S=5, N_all=90, n_genes=3, p_mean=40, p_lo=10, p_hi=70, ns_lo=20, ns_hi=90,
                           Q=1, nan_frac=0.02)]
    eng = engine.Engine(0)
    eng.set_grids(GL, GS)
    hip = _Hip()
    try:
        for k, hc in enumerate(cases):
            _say("[gpu] no-alloc call %d bfs=%s G=%d N=%d pairs=%d" % (k, bfs, hc.G, hc.N, hc.n_pairs))
            c = _device_chunk(hip, hc)
            opts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS[bfs], _abi.PREC["fp64_batched"], _abi.F64)
            h1, d1, o1 = _outputs(hip, hc, bfs)
            h2, d2, o2 = _outputs(hip, hc, bfs)
            eng.assoc_uvlr_device(c, opts, o1)
            eng.assoc_uvlr_device_noalloc(c, opts, o2)
            eng.sync()
            a, b = _read(hip, h1, d1), _read(hip, h2, d2)
            for key in a:
                assert np.array_equal(a[key], b[key], equal_nan=True), (k, key)
            assert np.all(b["n"] >= 0) and np.isfinite(b["w"][0]).any()
            # the plain precision runs the same arithmetic
            opts_p = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS[bfs], _abi.PREC["plain"], _abi.F64)
            h3, d3, o3 = _outputs(hip, hc, bfs)
            eng.assoc_uvlr_device_noalloc(c, opts_p, o3)
            eng.sync()
            c3 = _read(hip, h3, d3)
            for key in a:
                assert np.array_equal(a[key], c3[key], equal_nan=True), (k, key, "plain")
    finally:
        hip.free()
        eng.close()


def test_noalloc_refuses_oversize_and_mixed():
    eng = engine.Engine(0)
    eng.set_grids(GL, GS)
    hip = _Hip()
    try:
        hc = synth.batched(seed=73, S=3, N_all=40, n_genes=2, p_mean=10, p_lo=5, p_hi=15, ns_lo=20, ns_hi=40, Q=0)
        c = _device_chunk(hip, hc)
        h, d, o = _outputs(hip, hc, "gen")
        mixed = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["gen"], _abi.PREC["mixed"], _abi.F64)
        with pytest.raises(engine.EngineError) as e:
            eng.assoc_uvlr_device_noalloc(c, mixed, o)
        assert e.value.code == _abi.EQBF_ERR_UNSUPPORTED
        # a chunk whose scratch exceeds the static array (refused from its
        # sizes alone, before any kernel reads the buffers)
        big = _abi.Chunk.from_buffer_copy(c)
        big.n_genes = 1
        big.n_samples = 100000
        big.n_subgroups = 44
        fp64 = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["gen"], _abi.PREC["fp64_batched"], _abi.F64)
        with pytest.raises(engine.EngineError) as e:
            eng.assoc_uvlr_device_noalloc(big, fp64, o)
        assert e.value.code == _abi.EQBF_ERR_ARG and "scratch" in str(e.value)
        # and the entry point still works afterwards
        eng.assoc_uvlr_device_noalloc(c, fp64, o)
        eng.sync()
        assert np.all(hip.get(d["n"], h.n) >= 0)
    finally:
        hip.free()
        eng.close()
