"""hipGraph replay and the launch-shape autotuner of the batched association.

Both are pure scheduling: every replay of a captured graph, and every launch
shape the autotuner may pick, must give bit-identical outputs to the eager
call.  The graph test replays several times with eager work (zeroing the
outputs) and synchronisations between the replays, changes the inputs'
contents between two replays (the graph must follow them), and checks every
replay against an eager evaluation of the same inputs.
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


def _chunk():
    return synth.batched(seed=41, S=6, N_all=120, n_genes=6, p_mean=150, p_lo=50, p_hi=300, ns_lo=60, ns_hi=120,
                         Q=2, nan_frac=0.02)


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def _host_outputs(eng, hc):
    return eng.assoc_uvlr(hc, "join", "sin", precision="mixed", raw_dtype=np.float32)


class _Hip:
    """Device buffers through the HIP runtime the library itself uses (the
    process's libamdhip64.so.7), so no second runtime is involved."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
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

    def write(self, p, arr):
        arr = np.ascontiguousarray(arr)
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0

    def free(self):
        self.L.hipDeviceSynchronize()
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def test_graph_replays_match_eager():
    hc = _chunk()
    eng = engine.Engine(0)
    eng.set_grids(GL, GS)
    ref = _host_outputs(eng, hc)
    hip = _Hip()
    dev = {}
    c = hc.struct()
    for f in FIELDS:
        dev[f] = hip.put(getattr(hc, f))
        setattr(c, f, dev[f].value)
    zeros = {"n": np.zeros_like(ref.n), "sstats": np.zeros_like(ref.sstats), "stdsstats": np.zeros_like(ref.stdsstats),
             "raw": np.zeros_like(ref.raw), "w": np.zeros_like(ref.w)}
    out_d = {k: hip.put(v) for k, v in zeros.items()}
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats = out_d["n"].value, out_d["sstats"].value, out_d["stdsstats"].value
    o.raw_abf, o.w_abf = out_d["raw"].value, out_d["w"].value
    opts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["sin"], _abi.PREC["mixed"], _abi.F32)
    eng.set_graphs(True)
    eng2 = None
    for k in range(7):
        _say("[gpu] graph call %d" % k)
        if k == 4:
            # new contents in the same buffers: the replays must follow them
            hc2 = _chunk()
            hc2.explevels[...] = hc.explevels * 2.0 + 1.0
            hip.write(dev["explevels"], hc2.explevels)
            eng2 = engine.Engine(0)
            eng2.set_grids(GL, GS)
            ref = _host_outputs(eng2, hc2)
        for name, z in zeros.items():  # eager work between the replays
            hip.write(out_d[name], z)
        eng.assoc_uvlr_device(c, opts, o, None)
        eng.sync()
        got = {name: hip.get(out_d[name], z) for name, z in zeros.items()}
        assert _same(got["n"], ref.n)
        assert _same(got["sstats"], ref.sstats)
        assert _same(got["stdsstats"], ref.stdsstats)
        assert _same(got["raw"], ref.raw)
        assert _same(got["w"], ref.w)
    graphs, launches = eng.graph_stats()
    _say("[gpu]   graphs %d, graph launches %d" % (graphs, launches))
    assert graphs >= 1 and launches >= 5
    eng.set_graphs(False)
    hip.free()
    eng.close()
    if eng2 is not None:
        eng2.close()


def test_autotuner_keeps_results_and_records_choice():
    hc = _chunk()
    plain = engine.Engine(0)
    plain.set_grids(GL, GS)
    ref = _host_outputs(plain, hc)
    tuned = engine.Engine(0)
    tuned.set_grids(GL, GS)
    tuned.set_autotune(True)
    _say("[gpu] autotuned association")
    got = _host_outputs(tuned, hc)
    got2 = _host_outputs(tuned, hc)  # second call: the recorded choice
    for g in (got, got2):
        assert _same(g.n, ref.n) and _same(g.sstats, ref.sstats) and _same(g.stdsstats, ref.stdsstats)
        assert _same(g.raw, ref.raw) and _same(g.w, ref.w)
    choices = tuned.autotune_read()
    _say("[gpu]   choices %s" % choices)
    assert len(choices) == 1
    (ch,) = choices.values()
    assert ch["kch"] >= 4 and 1 <= ch["spp"] <= hc.S and ch["grid_mult"] in (1, 2) and ch["ms"] > 0
    plain.close()
    tuned.close()
