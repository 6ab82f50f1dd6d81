#!/usr/bin/env python3
"""Throughput of the eqtlbma_bf Bayes-factor engine on one MI355X.

Workload (the batched-ABF configuration): 20,000 genes, 44 subgroups drawn
from 500 individuals with uneven sizes N_s ~ U(80, 400), cis SNPs per gene
~ Poisson(1500) clipped to [200, 6000], FP32 dosages E[g] + N(0, 0.05^2)
clipped to [0, 2], 3 Gaussian covariates in the model, 15% eGenes,
--analys join --bfs sin (consistent + 44 singleton configurations) on a
50-point grid for both grids.  Synthetic data generated on the device;
inputs are resident in HBM before the timed region.

A step is one pass of the association engine (summary statistics, raw and
grid-averaged Bayes factors for every pair) over the whole workload, issued
as chunks of --batch-genes genes into reused output buffers.

Also reported: the plain FP64 one-gene-per-launch path on a gene subset,
the error of the mixed path against the FP64 CPU oracle on a sample, and the
oracle's CPU throughput on that sample (cpu_baseline).
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "SNP-gene BFs/s + gene perm p-values/s on 1 MI355X vs plain GPU, at stated error"
HBM_PEAK_GBS = 8000.0
FP64_MATRIX_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic.json")


def _load_traffic(kernel, args):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size) for this
    workload, if such a measurement is recorded."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(kernel)
    if not e or e.get("batch_genes") != args.batch_genes:
        return None
    return e.get("bytes_per_launch")


def _load_traffic_hm(kernel, genes):
    """Counter-measured HBM bytes per launch of an eqtlbma_hm kernel on this
    workload (same file, keyed by kernel and gene count), if recorded."""
    try:
        with open(TRAFFIC_FILE) as f:
            e = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not e or e.get("genes") != genes:
        return None
    return e.get("bytes_per_launch")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--genes", type=int, default=20000)
    ap.add_argument("--batch-genes", type=int, default=2000)
    ap.add_argument("--plain-genes", type=int, default=60)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--diag", action="store_true", help="per-phase split of the association kernel")
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--perm-genes", type=int, default=5000)
    ap.add_argument("--perm-nperm", type=int, default=10000)
    ap.add_argument("--no-perm", action="store_true", help="skip the permutation workload")
    ap.add_argument("--no-alloc", action="store_true", help="skip the learned-allocation workload")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape of the sums kernel (profiler counter runs)")
    ap.add_argument("--alloc-train", type=int, default=2000)
    ap.add_argument("--alloc-genes", type=int, default=20000)
    ap.add_argument("--alloc-holdout", type=int, default=1000)
    ap.add_argument("--alloc-epochs", type=int, default=200)
    ap.add_argument("--only-perm", action="store_true", help="run only the permutation workload")
    ap.add_argument("--mvlr-genes", type=int, default=2000)
    ap.add_argument("--no-mvlr", action="store_true", help="skip the MVLR workload")
    ap.add_argument("--only-mvlr", action="store_true", help="run only the MVLR (and --error hybrid) workloads")
    ap.add_argument("--hybrid-genes", type=int, default=2000)
    ap.add_argument("--no-hybrid", action="store_true", help="skip the --error hybrid workload")
    ap.add_argument("--only-hybrid", action="store_true", help="run only the --error hybrid workload")
    ap.add_argument("--no-c0", action="store_true", help="skip the plain-method workload")
    ap.add_argument("--hm-genes", type=int, default=20000)
    ap.add_argument("--hm-iters", type=int, default=6, help="EM iterations timed in the eqtlbma_hm workload")
    ap.add_argument("--no-hm", action="store_true", help="skip the eqtlbma_hm workload")
    ap.add_argument("--only-hm", action="store_true", help="run only the eqtlbma_hm workload")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, to scale the HBM byte counters of a profiler run)")
    return ap.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def host_cpu():
    """(CPU model, threads used for the CPU baselines): the host cores this
    process may run on, at most 16 (one GPU's share of a box's cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, max(1, min(16, n))


def threaded(fn, items, threads):
    """fn over items on `threads` host threads (the oracle's ctypes calls
    release the interpreter lock, so the threads run in parallel)."""
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(fn, items))


def cpu_label(model, threads, what):
    return {"cores": threads, "kind": "port", "label": "proxy, not reference", "cpu_model": model,
            "note": "the reference needs GSL, absent from this image; this is the FP64 oracle restatement of "
                    "its algorithm, %s, on %d host threads" % (what, threads)}


class Workload:
    """Device-resident synthetic data of the batched configuration."""

    def __init__(self, torch, G, seed, S=44, N=500, Q=3):
        dev = torch.device("cuda:0")
        self.torch, self.dev = torch, dev
        self.G, self.S, self.N, self.Q = G, S, N, Q
        rng = np.random.default_rng(seed)
        P = np.clip(rng.poisson(1500, size=G), 200, 6000).astype(np.int64)
        self.P = P
        self.offs = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
        n_snps = int(self.offs[-1])
        self.n_snps = n_snps
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        t0 = time.time()
        self.geno = torch.empty((N, n_snps), dtype=torch.float32, device=dev)
        step = 1 << 20
        for a in range(0, n_snps, step):
            b = min(n_snps, a + step)
            maf = torch.rand(b - a, device=dev, generator=gen) * 0.45 + 0.05
            g = (torch.rand((N, b - a), device=dev, generator=gen) < maf).float()
            g += (torch.rand((N, b - a), device=dev, generator=gen) < maf).float()
            g += torch.randn((N, b - a), device=dev, generator=gen) * 0.05
            self.geno[:, a:b] = g.clamp_(0.0, 2.0)
            del g, maf
        # membership: N_s ~ U(80, 400) individuals per subgroup
        geno_in = np.zeros((S, N), np.uint8)
        for s in range(S):
            ns = rng.integers(80, 401)
            geno_in[s, rng.choice(N, size=ns, replace=False)] = 1
        self.geno_in = torch.from_numpy(geno_in).to(dev)
        self.covars = torch.randn((S, Q, N), device=dev, generator=gen, dtype=torch.float64)
        self.covar_count = torch.full((S,), Q, dtype=torch.int32, device=dev)
        self.covar_in = torch.ones((S, N), dtype=torch.uint8, device=dev)
        self.gene_in = torch.ones((G, S), dtype=torch.uint8, device=dev)
        self.gset = torch.zeros((S,), dtype=torch.int32, device=dev)
        self.snp_in = torch.ones((n_snps, S), dtype=torch.uint8, device=dev)
        self.pair_snp = torch.arange(n_snps, dtype=torch.int32, device=dev)
        # expression: noise + covariate effects + one eQTL in 15% of genes
        y = torch.randn((G, S, N), device=dev, generator=gen, dtype=torch.float64)
        y += 0.3 * self.covars.sum(dim=1)[None]
        eg = rng.choice(G, size=int(round(0.15 * G)), replace=False)
        for g in eg:
            j = int(self.offs[g] + rng.integers(P[g]))
            beta = rng.normal(0, 0.5)
            act = rng.random(S) < 0.5
            if not act.any():
                act[rng.integers(S)] = True
            idx = torch.from_numpy(np.nonzero(act)[0]).to(dev)
            y[g, idx] += beta * self.geno[:, j].double()[None, :]
        y[:, ~self.geno_in.bool()] = float("nan")
        self.y = y
        torch.cuda.synchronize()
        log("generated %d genes, %d pairs, %.1f GB genotypes in %.1f s" %
            (G, n_snps, self.geno.numel() * 4 / 1e9, time.time() - t0))
        # per-chunk pair offsets (relative) on the device
        self.batches = []

    def make_batches(self, bg):
        torch = self.torch
        self.batches = []
        for g0 in range(0, self.G, bg):
            g1 = min(self.G, g0 + bg)
            rel = self.offs[g0:g1 + 1] - self.offs[g0]
            self.batches.append((g0, g1, torch.from_numpy(rel).to(self.dev)))
        return max(int(self.offs[g1] - self.offs[g0]) for g0, g1, _ in self.batches)

    def chunk(self, abi, g0, g1, rel_off):
        c = abi.Chunk()
        S, N = self.S, self.N
        c.n_genes, c.n_subgroups, c.n_samples, c.n_covars_max = g1 - g0, S, N, self.Q
        c.n_pairs = int(self.offs[g1] - self.offs[g0])
        c.n_snps = self.n_snps
        c.explevels = self.y.data_ptr() + g0 * S * N * 8
        c.gene_in_sbgrp = self.gene_in.data_ptr() + g0 * S
        c.geno_in_sbgrp = self.geno_in.data_ptr()
        c.covar_count = self.covar_count.data_ptr()
        c.covars = self.covars.data_ptr()
        c.covar_in_sbgrp = self.covar_in.data_ptr()
        c.n_geno_sets = 1
        c.geno_set_of_sbgrp = self.gset.data_ptr()
        c.genotypes = self.geno.data_ptr()
        c.geno_dtype = abi.F32
        c.geno_ld = self.n_snps
        c.geno_set_stride = N * self.n_snps
        c.snp_in_sbgrp = self.snp_in.data_ptr()
        c.pair_offset = rel_off.data_ptr()
        c.pair_snp = self.pair_snp.data_ptr() + int(self.offs[g0]) * 4
        return c

    def host_chunk(self, abi, g0, g1):
        """Host copy of genes [g0, g1) with their own SNP columns."""
        p0, p1 = int(self.offs[g0]), int(self.offs[g1])
        return abi.HostChunk(
            explevels=self.y[g0:g1].cpu().numpy(),
            gene_in_sbgrp=self.gene_in[g0:g1].cpu().numpy(),
            geno_in_sbgrp=self.geno_in.cpu().numpy(),
            genotypes=np.ascontiguousarray(self.geno[:, p0:p1].cpu().numpy()),
            snp_in_sbgrp=self.snp_in[p0:p1].cpu().numpy(),
            pair_offset=self.offs[g0:g1 + 1] - p0,
            pair_snp=np.arange(p1 - p0, dtype=np.int32),
            covars=self.covars.cpu().numpy(), covar_count=self.covar_count.cpu().numpy(),
            covar_in_sbgrp=self.covar_in.cpu().numpy())


class PermWorkload:
    """Device-resident synthetic data of the permutation configuration: S
    subgroups sharing all N individuals, hard-call genotypes 0/1/2 under HWE
    with MAF~U(0.05, 0.5), y = N(0,1) plus one eQTL (beta~N(0, 0.5^2) in a
    random non-empty set of subgroups) in 10% of genes, each gene with its
    own cis SNPs (P~Poisson(1000) clipped to [500, 1500])."""

    def __init__(self, torch, G, seed, S=5, N=300):
        dev = torch.device("cuda:0")
        self.torch, self.dev, self.G, self.S, self.N = torch, dev, G, S, N
        rng = np.random.default_rng(seed)
        P = np.clip(rng.poisson(1000, size=G), 500, 1500).astype(np.int64)
        self.P = P
        self.offs = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
        n_snps = int(self.offs[-1])
        self.n_snps = n_snps
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        t0 = time.time()
        self.geno = torch.empty((N, n_snps), dtype=torch.float32, device=dev)
        step = 1 << 20
        for a in range(0, n_snps, step):
            b = min(n_snps, a + step)
            maf = torch.rand(b - a, device=dev, generator=gen) * 0.45 + 0.05
            g = (torch.rand((N, b - a), device=dev, generator=gen) < maf).float()
            g += (torch.rand((N, b - a), device=dev, generator=gen) < maf).float()
            self.geno[:, a:b] = g
            del g, maf
        y = torch.randn((G, S, N), device=dev, generator=gen, dtype=torch.float64)
        for gi in rng.choice(G, size=int(round(0.1 * G)), replace=False):
            j = int(self.offs[gi] + rng.integers(P[gi]))
            beta = rng.normal(0, 0.5)
            mask = int(rng.integers(1, 1 << S))
            idx = torch.tensor([s for s in range(S) if (mask >> s) & 1], device=dev)
            y[gi, idx] += beta * self.geno[:, j].double()[None, :]
        self.y = y - y.mean(dim=2, keepdim=True)
        self.gene_in = torch.ones((G, S), dtype=torch.uint8, device=dev)
        self.geno_in = torch.ones((S, N), dtype=torch.uint8, device=dev)
        self.covars = torch.zeros((S, 1, N), dtype=torch.float64, device=dev)
        self.covar_count = torch.zeros((S,), dtype=torch.int32, device=dev)
        self.covar_in = torch.ones((S, N), dtype=torch.uint8, device=dev)
        self.gset = torch.zeros((S,), dtype=torch.int32, device=dev)
        self.snp_in = torch.ones((n_snps, S), dtype=torch.uint8, device=dev)
        self.pair_snp = torch.arange(n_snps, dtype=torch.int32, device=dev)
        self.rel = torch.from_numpy(self.offs.copy()).to(dev)
        self._rels = {}
        torch.cuda.synchronize()
        log("perm workload: %d genes, %d pairs, %.1f GB genotypes in %.1f s" %
            (G, n_snps, self.geno.numel() * 4 / 1e9, time.time() - t0))

    def chunk(self, abi, g0, g1):
        c = abi.Chunk()
        S, N = self.S, self.N
        c.n_genes, c.n_subgroups, c.n_samples, c.n_covars_max = g1 - g0, S, N, 0
        c.n_pairs = int(self.offs[g1] - self.offs[g0])
        c.n_snps = self.n_snps
        c.explevels = self.y.data_ptr() + g0 * S * N * 8
        c.gene_in_sbgrp = self.gene_in.data_ptr() + g0 * S
        c.geno_in_sbgrp = self.geno_in.data_ptr()
        c.covar_count = self.covar_count.data_ptr()
        c.covars = self.covars.data_ptr()
        c.covar_in_sbgrp = self.covar_in.data_ptr()
        c.n_geno_sets = 1
        c.geno_set_of_sbgrp = self.gset.data_ptr()
        c.genotypes = self.geno.data_ptr()
        c.geno_dtype = abi.F32
        c.geno_ld = self.n_snps
        c.geno_set_stride = N * self.n_snps
        c.snp_in_sbgrp = self.snp_in.data_ptr()
        # pair offsets relative to gene g0: a shifted copy, kept alive as long
        # as the workload (the chunk holds only its address)
        key = (g0, g1)
        if key not in self._rels:
            self._rels[key] = self.rel[g0:g1 + 1] - int(self.offs[g0])
        c.pair_offset = self._rels[key].data_ptr()
        c.pair_snp = self.pair_snp.data_ptr() + int(self.offs[g0]) * 4
        return c

    def host_chunk(self, abi, g0, g1):
        p0, p1 = int(self.offs[g0]), int(self.offs[g1])
        return abi.HostChunk(
            explevels=self.y[g0:g1].cpu().numpy(), gene_in_sbgrp=self.gene_in[g0:g1].cpu().numpy(),
            geno_in_sbgrp=self.geno_in.cpu().numpy(),
            genotypes=np.ascontiguousarray(self.geno[:, p0:p1].cpu().numpy()),
            snp_in_sbgrp=self.snp_in[p0:p1].cpu().numpy(), pair_offset=self.offs[g0:g1 + 1] - p0,
            pair_snp=np.arange(p1 - p0, dtype=np.int32))


WRTSIZE = 10  # genes per chunk of the reference's run loop (--wrtsize default); the RNG restarts per chunk
TIE_BOUND = 5e-5  # FP64 re-check of permutation statistics this close to the true one (log10 BF)


def perm_leg(args, torch, result):
    """Gene-level permutation p-values (--analys join --pbf all --maxbf,
    joint shuffle, seed 7): true statistics from the association pass, the
    statistic of every permutation, and the reference's counting loop."""
    from eqtlbma_amd import _abi, synth
    from eqtlbma_amd.engine import Engine, PermOpts, Rng, perm_pvalue
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    G, nperm = args.perm_genes, args.perm_nperm
    wl = PermWorkload(torch, G, 1002)
    S, N = wl.S, wl.N
    dev = wl.dev
    eng = Engine(0)
    eng.set_grids(gL, gS)
    n_cfg = 2 ** S - 1
    n_w = 3 + n_cfg + 2
    n_raw = 3 * gL.shape[0] + n_cfg * gS.shape[0]
    NP = int(wl.offs[-1])
    out_n = torch.empty((S, NP), dtype=torch.int32, device=dev)
    out_w = torch.empty((n_w, NP), dtype=torch.float64, device=dev)
    out_raw = torch.empty((n_raw, 1), dtype=torch.float32, device=dev)
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats, o.raw_abf, o.w_abf = out_n.data_ptr(), None, None, None, out_w.data_ptr()
    aopts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["all"], _abi.PREC["mixed"], _abi.F32)
    popts = PermOpts(_abi.ANALYSIS["join"], _abi.PBF["all"], 1, -1, _abi.PREC["mixed"])
    t = time.perf_counter()
    tables = Rng(7).tables(N, WRTSIZE, nperm)  # [WRTSIZE][nperm][N]: one RNG stream per chunk of genes
    t_tab = time.perf_counter() - t
    tables_d = torch.from_numpy(tables).to(dev)
    stats = torch.empty((G, nperm), dtype=torch.float64, device=dev)
    gene_table = (np.arange(G) % WRTSIZE).astype(np.int32)
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)

    true_d = torch.empty((G,), dtype=torch.float64, device=dev)
    w_all_row = 3 + n_cfg + 1

    def run(g1):
        # association pass (testForAssociations), then every permutation's
        # statistic with the FP64 true statistics and FP64 re-checks of the
        # near-ties, then the reference's counting loop on the host
        c = wl.chunk(_abi, 0, g1)
        eng.assoc_uvlr_device(c, aopts, o, sptr)
        npc = int(wl.offs[g1])  # the outputs' leading dimension is this chunk's pair count
        w = out_w.view(-1)[w_all_row * npc:(w_all_row + 1) * npc].cpu().numpy()
        true_d[:g1].copy_(torch.from_numpy(np.fmax.reduceat(w, wl.offs[:g1].astype(np.int64))))
        n_re = eng.perm_stats_fp64_ties_device(c, popts, C.c_void_p(tables_d.data_ptr()), WRTSIZE, gene_table[:g1],
                                               nperm, TIE_BOUND, C.c_void_p(stats.data_ptr()),
                                               C.c_void_p(true_d.data_ptr()), sptr)
        true = true_d[:g1].cpu().numpy()
        st = stats[:g1].cpu().numpy()
        pv = np.array([perm_pvalue(st[g], true[g], nperm)[0] for g in range(g1)])
        return pv, st, true, n_re

    run(min(G, 2 * WRTSIZE))  # warm-up: buffers and tables
    torch.cuda.synchronize()
    eng.profile(True)
    before = eng.profile_read()
    t0 = time.perf_counter()
    pv, st, true, n_re = run(G)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: (v[0] - before.get(k, (0.0, 0))[0], v[1] - before.get(k, (0.0, 0))[1]) for k, v in after.items()}
    kern = {k: v for k, v in kern.items() if v[1] > 0}
    value = G / dt
    perm = {"metric": "gene permutation p-values/s", "value": value, "unit": "genes/s", "seconds": dt,
            "workload": "permutations: %d genes, %d subgroups sharing N=%d, cis SNPs~Pois(1000) in [500,1500], "
                        "hard calls, %d joint permutations per gene (GSL mt19937 seed 7, one stream per %d-gene "
                        "chunk), --pbf all --maxbf, grids 25/10" % (G, S, N, nperm, WRTSIZE),
            "pairs": NP, "host_table_generation_s": t_tab,
            "fp64_recheck": {"bound_abs_log10_bf": TIE_BOUND, "permutations_rechecked": int(n_re),
                             "of": int(G) * int(nperm),
                             "what": "permutation statistics within the bound of the true statistic, and the true "
                                     "statistic of their genes, recomputed in FP64 before counting"},
            "kernels_ms_total": {k: v[0] for k, v in kern.items()},
            "kernels_launches": {k: v[1] for k, v in kern.items()},
            "egenes_p_below_0.01": int(np.sum(pv < 0.01))}
    # roofline of the contraction: 2 N P S nperm int8 multiply-adds per gene counted once
    if "pf_gemm_kernel" in kern:
        ops = 2.0 * N * NP * S * nperm
        gm = kern["pf_gemm_kernel"][0] / 1e3
        perm["roofline_gemm"] = {"bound": "mfma", "achieved": ops / gm / 1e12, "peak": 5000.0,
                                 "unit": "TOP/s (int8)", "frac": ops / gm / 1e12 / 5000.0, "kernel": "pf_gemm_kernel",
                                 "alg_ops": ops, "total_ms": gm * 1e3}
    if "pf_bma_kernel" in kern:
        # --maxbf: every z of every (SNP, subgroup, permutation) is read once to
        # bound the SNP's statistic; only the SNPs whose bound reaches the
        # running maximum are evaluated (310 exponentials each), so the floor
        # is the read of the z buffer (FP32)
        zbytes = float(NP) * S * nperm * 4
        bm = kern["pf_bma_kernel"][0] / 1e3
        terms = float(NP) * nperm * n_cfg * gS.shape[0]
        perm["roofline_bma"] = {"bound": "hbm", "achieved": zbytes / bm / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": zbytes / bm / 1e9 / HBM_PEAK_GBS, "kernel": "pf_bma_kernel",
                                "alg_bytes": zbytes, "total_ms": bm * 1e3,
                                "terms_if_every_snp_were_evaluated_per_s": terms / bm}
    log("perm: %d genes x %d permutations in %.2f s = %.1f genes/s; kernels %s" %
        (G, nperm, dt, value, {k: round(v[0], 1) for k, v in kern.items()}))

    # plain FP64 one-gene-per-launch path: a few permutations of one gene, per (gene, permutation)
    if not args.no_plain:
        k = 8
        c = wl.chunk(_abi, 0, 1)
        tab1 = torch.from_numpy(np.ascontiguousarray(tables[:1, :k])).to(dev)
        st1 = torch.empty((1, k), dtype=torch.float64, device=dev)
        pp = PermOpts(_abi.ANALYSIS["join"], _abi.PBF["all"], 1, -1, _abi.PREC["fp64_plain"])
        eng.perm_stats_device(c, pp, C.c_void_p(tab1.data_ptr()), 1, np.zeros(1, np.int32), k,
                              C.c_void_p(st1.data_ptr()), sptr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.perm_stats_device(c, pp, C.c_void_p(tab1.data_ptr()), 1, np.zeros(1, np.int32), k,
                              C.c_void_p(st1.data_ptr()), sptr)
        torch.cuda.synchronize()
        dtp = time.perf_counter() - t
        plain_rate = 1.0 / (dtp / k * nperm)
        perm["plain_gpu"] = {"value": plain_rate, "unit": "genes/s",
                             "sample": "gene 0 (%d SNPs), %d permutations, FP64, one launch per gene and "
                                       "permutation; rate per gene at %d permutations" % (wl.P[0], k, nperm),
                             "speedup_vs_plain": value / plain_rate}
        e_plain = float(np.max(np.abs(st1.cpu().numpy()[0] - st[0, :k])))
        perm["plain_vs_mixed_stat_abs"] = e_plain
        log("perm plain %.3e genes/s" % plain_rate)
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        model, T = host_cpu()
        hc = wl.host_chunk(_abi, 0, 1)
        k = 2
        t = time.perf_counter()
        orc.perm_join_stats(hc, 0, "all", True, gL, gS, tables[0, :k])
        d1 = (time.perf_counter() - t) / k
        kt = max(1, min(nperm // T, int(args.cpu_seconds / max(d1, 1e-4))))  # permutations per thread
        parts = [tables[0, r * kt:(r + 1) * kt] for r in range(T)]
        t = time.perf_counter()
        refs = threaded(lambda tb: orc.perm_join_stats(hc, 0, "all", True, gL, gS, tb), parts, T)
        dtc = time.perf_counter() - t
        ref = np.concatenate(refs)
        k = ref.size
        perm["cpu_baseline"] = dict(value=1.0 / (dtc / k * nperm), unit="genes/s",
                                    sample="gene 0 (%d SNPs), its first %d permutations split over the threads, "
                                           "%.1f s wall; rate per gene at %d permutations" % (wl.P[0], k, dtc, nperm),
                                    **cpu_label(model, T, "permutations of one gene sharded over threads"))
        perm["error_vs_fp64_oracle"] = {"perm_stat_abs": float(np.max(np.abs(st[0, :k] - ref))), "sample_perms": k,
                                        "true_stat_abs": float(abs(true[0] - orc.true_stat(
                                            orc.assoc_uvlr(hc, "join", "all", gL, gS).w[3 + n_cfg + 1], True)))}
        log("perm cpu oracle %.3e genes/s; stat error %.3g" % (perm["cpu_baseline"]["value"],
                                                                 perm["error_vs_fp64_oracle"]["perm_stat_abs"]))
    eng.close()
    del wl, stats, tables_d, out_n, out_w
    torch.cuda.empty_cache()
    result["perm"] = perm


def perm_alloc_leg(args, torch, result):
    """Learned permutation allocator at scale: trained on null draws of 2,000
    genes x 10,000 permutations from the permutation engine, then applied to
    20,000 genes (predicted p > 0.2: 1,000 permutations, else 10,000), with
    1,000 of them also permuted fully to check the p-values and eGene calls."""
    from eqtlbma_amd import _abi, allocator, synth
    from eqtlbma_amd.engine import Engine, PermOpts, Rng, perm_pvalue
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    n_train, n_inf, n_hold = args.alloc_train, args.alloc_genes, args.alloc_holdout
    nfull, nlow = args.perm_nperm, max(1, args.perm_nperm // 10)
    G = n_train + n_inf
    wl = PermWorkload(torch, G, 1004)
    S, N, dev = wl.S, wl.N, wl.dev
    eng = Engine(0)
    eng.set_grids(gL, gS)
    n_cfg = 2 ** S - 1
    NP = int(wl.offs[-1])
    out_n = torch.empty((S, NP), dtype=torch.int32, device=dev)
    out_w = torch.empty((3 + n_cfg + 2, NP), dtype=torch.float64, device=dev)
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats, o.raw_abf, o.w_abf = out_n.data_ptr(), None, None, None, out_w.data_ptr()
    aopts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["all"], _abi.PREC["mixed"], _abi.F32)
    popts = PermOpts(_abi.ANALYSIS["join"], _abi.PBF["all"], 1, -1, _abi.PREC["mixed"])
    tables = Rng(7).tables(N, WRTSIZE, nfull)
    tab_full = torch.from_numpy(tables).to(dev)
    tab_low = torch.from_numpy(np.ascontiguousarray(tables[:, :nlow])).to(dev)
    stats = torch.empty((G, nfull), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)
    c_all = wl.chunk(_abi, 0, G)
    slot = (np.arange(G) % WRTSIZE).astype(np.int32)

    def perm(genes, k):
        gt = np.full(G, -1, np.int32)
        gt[genes] = slot[genes]
        tab = tab_full if k == nfull else tab_low
        eng.perm_stats_device(c_all, popts, C.c_void_p(tab.data_ptr()), WRTSIZE, gt, k,
                              C.c_void_p(stats.data_ptr()), sptr)
        return stats[:, :k]  # statistics of gene g in row g (leading dimension k)

    def stats_rows(genes, k):
        flat = stats.view(-1)[:G * k].view(G, k)
        return flat[torch.from_numpy(np.asarray(genes)).to(dev)].cpu().numpy()

    t0 = time.perf_counter()
    eng.assoc_uvlr_device(c_all, aopts, o, sptr)
    w_all = out_w[3 + n_cfg + 1].cpu().numpy()
    true = np.maximum.reduceat(w_all, wl.offs[:G].astype(np.int64))
    t_true = time.perf_counter() - t0
    # ---- training: null draws of the training genes, GEV fits, network ----
    t0 = time.perf_counter()
    tr = np.arange(n_train)
    perm(tr, nfull)
    null = stats_rows(tr, nfull)
    gev = eng.gev_fit(null)
    eig_tr = eng.ld_eigenvalues_device(wl.chunk(_abi, 0, n_train), n_train, allocator.N_EIG)
    x_tr = allocator.features(wl.P[:n_train], N, S, eig_tr, true[:n_train])
    al = allocator.Allocator(eng, seed=3)
    curve = al.fit(x_tr, gev, epochs=args.alloc_epochs, batch=256, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # ---- inference on the other genes (timed end to end) ----
    inf = np.arange(n_train, G)
    eng.profile(True)
    before = eng.profile_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phase = {}
    eng.assoc_uvlr_device(wl.chunk(_abi, n_train, G), aopts, o, sptr)
    npc = int(wl.offs[G] - wl.offs[n_train])  # the outputs' leading dimension is this chunk's pair count
    w_inf = out_w.view(-1)[(3 + n_cfg + 1) * npc:(3 + n_cfg + 2) * npc].cpu().numpy()
    base = int(wl.offs[n_train])
    true_inf = np.maximum.reduceat(w_inf, (wl.offs[n_train:G] - base).astype(np.int64))
    phase["true_statistics"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    eig = eng.ld_eigenvalues_device(wl.chunk(_abi, n_train, G), n_inf, allocator.N_EIG)
    x = allocator.features(wl.P[n_train:], N, S, eig, true_inf)
    p_hat = al.p_hat(x, true_inf)
    k_gene = allocator.Allocator.allocate(p_hat, 0.2, nlow, nfull)
    phase["features_and_network"] = time.perf_counter() - t1
    t1 = time.perf_counter()

    def run_perm(k, local):
        perm(inf[local], k)
        return stats_rows(inf[local], k)

    pv, used = allocator.allocated_pvalues(eng, run_perm, true_inf, k_gene)
    torch.cuda.synchronize()
    phase["permutations_and_pvalues"] = time.perf_counter() - t1
    dt = time.perf_counter() - t0
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: round(v[0] - before.get(k, (0.0, 0))[0], 2) for k, v in after.items() if v[1] > before.get(k, (0, 0))[1]}
    value = n_inf / dt
    # ---- every inference gene fully permuted: BH 5 % eGene calls over all of
    # them, and the p-value error on the first n_hold (held out from nothing
    # the network saw: it trained on other genes) ----
    t1 = time.perf_counter()
    perm(inf, nfull)
    full_all = stats_rows(inf, nfull)
    pv_full_all = np.array([perm_pvalue(full_all[r], true_inf[r], nfull)[0] for r in range(n_inf)])
    t_full = time.perf_counter() - t1
    calls_a_all = allocator.bh_reject(pv, 0.05)
    calls_f_all = allocator.bh_reject(pv_full_all, 0.05)
    hold = np.arange(n_hold)
    pv_full = pv_full_all[hold]
    calls_a = allocator.bh_reject(pv[hold], 0.05)
    calls_f = allocator.bh_reject(pv_full, 0.05)
    lowk = k_gene[hold] == nlow
    res = {"metric": "gene permutation p-values/s, learned allocation", "value": value, "unit": "genes/s",
           "seconds": dt, "phase_seconds": phase, "genes": int(n_inf), "train_genes": int(n_train),
           "train_seconds": t_train,
           "true_stat_seconds_train_set": t_true,
           "workload": "as the permutation workload (5 subgroups sharing N=300, hard calls, cis SNPs~Pois(1000), "
                       "--pbf all --maxbf), %d genes; network 24 -> 256x3 GELU -> 3 (BF16 weights) trained on "
                       "%d genes x %d permutations; predicted p > 0.2: %d permutations, else %d" %
                       (n_inf, n_train, nfull, nlow, nfull),
           "epochs": args.alloc_epochs, "final_train_loss": float(curve[-1]) if len(curve) else None,
           "permutations_used_fraction_of_full": float(used.sum() / (nfull * n_inf)),
           "genes_given_short_run": int(np.sum(k_gene == nlow)),
           "speedup_vs_full_permutation_same_engine": None,
           "heldout": {"genes": int(n_hold), "egenes_bh5_full": int(calls_f.sum()),
                       "egenes_bh5_allocated": int(calls_a.sum()),
                       "call_agreement": float(np.mean(calls_a == calls_f)),
                       "max_abs_p_error_short_runs": float(np.max(np.abs(pv[hold] - pv_full)[lowk])) if lowk.any()
                       else 0.0,
                       "max_abs_p_error_full_runs": float(np.max(np.abs(pv[hold] - pv_full)[~lowk])) if (~lowk).any()
                       else 0.0},
           "all_genes": {"genes": int(n_inf), "egenes_bh5_full": int(calls_f_all.sum()),
                         "egenes_bh5_allocated": int(calls_a_all.sum()),
                         "calls_identical": bool(np.array_equal(calls_a_all, calls_f_all)),
                         "call_disagreements": int(np.sum(calls_a_all != calls_f_all)),
                         "full_permutation_seconds": t_full},
           "kernels_ms_total": kern}
    if "perm" in result and result["perm"].get("value"):
        res["speedup_vs_full_permutation_same_engine"] = value / result["perm"]["value"]
        if "plain_gpu" in result["perm"]:
            res["speedup_vs_plain_gpu"] = value / result["perm"]["plain_gpu"]["value"]
    log("allocated: %d genes in %.2f s = %.1f genes/s; %.2f of full permutations; BH 5%% calls over all genes "
        "%d allocated vs %d full (%d differ)" % (n_inf, dt, value, res["permutations_used_fraction_of_full"],
                                                 calls_a_all.sum(), calls_f_all.sum(),
                                                 int(np.sum(calls_a_all != calls_f_all))))
    al.close()
    eng.close()
    del wl, stats, tab_full, tab_low, out_n, out_w
    torch.cuda.empty_cache()
    result["perm_allocated"] = res


def to_device(torch, hc, abi):
    """A HostChunk's arrays as device tensors and a chunk struct pointing at
    them (the tensors are returned so that they stay alive)."""
    dev = torch.device("cuda:0")
    keep = {}
    c = hc.struct()
    for f in ("explevels", "gene_in_sbgrp", "geno_in_sbgrp", "covar_count", "covars", "covar_in_sbgrp",
              "geno_set_of_sbgrp", "genotypes", "snp_in_sbgrp", "pair_offset", "pair_snp"):
        t = torch.from_numpy(np.ascontiguousarray(getattr(hc, f))).to(dev)
        keep[f] = t
        setattr(c, f, t.data_ptr())
    return c, keep


def c0_leg(args, torch, result):
    """The plain-method configuration (the reference's CPU-runnable case): 3
    subgroups sharing N = 100 individuals, 200 genes with ~201 cis SNPs each,
    hard calls, 10 % eGenes, --analys join --bfs all, grids 25 / 10, FP64,
    seed 42.  Latency of one association call over the whole chunk."""
    from eqtlbma_amd import _abi, synth
    from eqtlbma_amd.engine import Engine
    from oracle import oracle as orc
    gL, gS = synth.make_grid("general"), synth.make_grid("small")
    hc = synth.plain(seed=42, S=3, N=100, n_genes=200, snps_per_gene=201, frac_egenes=0.1)
    S, NP = hc.S, hc.n_pairs
    dev = torch.device("cuda:0")
    c, keep = to_device(torch, hc, _abi)
    eng = Engine(0)
    eng.set_grids(gL, gS)
    n_cfg = 2 ** S - 1
    n_raw, n_w = 3 * gL.shape[0] + n_cfg * gS.shape[0], 3 + n_cfg + 2
    outs = {"n": torch.empty((S, NP), dtype=torch.int32, device=dev),
            "sstats": torch.empty((5, S, NP), dtype=torch.float64, device=dev),
            "stdsstats": torch.empty((3, S, NP), dtype=torch.float64, device=dev),
            "raw": torch.empty((n_raw, NP), dtype=torch.float64, device=dev),
            "w": torch.empty((n_w, NP), dtype=torch.float64, device=dev)}
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats = outs["n"].data_ptr(), outs["sstats"].data_ptr(), outs["stdsstats"].data_ptr()
    o.raw_abf, o.w_abf = outs["raw"].data_ptr(), outs["w"].data_ptr()
    sptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    res = {"workload": "plain method: 200 genes, 3 subgroups sharing N=100, ~201 cis SNPs per gene, hard calls, "
                       "--bfs all, grids 25/10, seed 42", "pairs": NP}
    t = time.perf_counter()
    ref = orc.assoc_uvlr(hc, "join", "all", gL, gS)
    t_orc = time.perf_counter() - t
    for name in ("fp64_batched", "mixed", "fp64_plain"):
        opts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["all"], _abi.PREC[name], _abi.F64)
        eng.assoc_uvlr_device(c, opts, o, sptr)
        torch.cuda.synchronize()
        reps = 20 if name != "fp64_plain" else 3
        t = time.perf_counter()
        for _ in range(reps):
            eng.assoc_uvlr_device(c, opts, o, sptr)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / reps
        w = outs["w"].cpu().numpy()
        raw = outs["raw"].cpu().numpy()
        fw, fr = np.isfinite(ref.w), np.isfinite(ref.raw)
        res[name] = {"ms_per_call": ms, "pairs_per_s": NP / (ms / 1e3),
                     "log10_abf_weighted_abs_err": float(np.max(np.abs(w[fw] - ref.w[fw]))),
                     "log10_abf_raw_abs_err": float(np.max(np.abs(raw[fr] - ref.raw[fr])))}
    model, T = host_cpu()
    # the oracle over the same 200 genes, sharded by gene over the threads
    subs = [_abi.HostChunk(hc.explevels[g:g + 1], hc.gene_in_sbgrp[g:g + 1], hc.geno_in_sbgrp, hc.genotypes,
                           hc.snp_in_sbgrp, hc.pair_offset[g:g + 2] - hc.pair_offset[g],
                           hc.pair_snp[hc.pair_offset[g]:hc.pair_offset[g + 1]]) for g in range(hc.G)]
    t = time.perf_counter()
    threaded(lambda h: orc.assoc_uvlr(h, "join", "all", gL, gS), subs, T)
    t_par = time.perf_counter() - t
    res["cpu_baseline"] = dict(value=NP / t_par, unit="pairs/s", single_thread_pairs_per_s=NP / t_orc,
                               sample="all 200 genes, %.2f s wall (%.2f s on one thread)" % (t_par, t_orc),
                               **cpu_label(model, T, "genes sharded over threads"))
    log("c0: %s" % {k: (round(v["ms_per_call"], 3) if isinstance(v, dict) and "ms_per_call" in v else None)
                    for k, v in res.items()})
    eng.close()
    del keep, outs
    result["plain_method_c0"] = res


class MvlrWorkload:
    """Device-resident synthetic data of the MVLR configuration: S = 9
    subgroups, every one of N = 150 individuals in every subgroup, hard-call
    genotypes under HWE with MAF ~ U(0.05, 0.5), residuals correlated across
    subgroups (equicorrelation rho = 0.5), one eQTL in 10 % of genes with
    multivariate normal effects (between-subgroup correlation 0.8, sd 0.5),
    cis SNPs per gene ~ Poisson(800) clipped to [400, 1200]."""

    def __init__(self, torch, G, seed, S=9, N=150, rho=0.5, r_eff=0.8):
        dev = torch.device("cuda:0")
        self.torch, self.dev, self.G, self.S, self.N = torch, dev, G, S, N
        rng = np.random.default_rng(seed)
        P = np.clip(rng.poisson(800, size=G), 400, 1200).astype(np.int64)
        self.P = P
        self.offs = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
        n_snps = int(self.offs[-1])
        self.n_snps = n_snps
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        maf = torch.rand(n_snps, device=dev, generator=gen) * 0.45 + 0.05
        g = (torch.rand((N, n_snps), device=dev, generator=gen) < maf).float()
        g += (torch.rand((N, n_snps), device=dev, generator=gen) < maf).float()
        self.geno = g
        L = torch.from_numpy(np.linalg.cholesky(np.full((S, S), rho) + (1 - rho) * np.eye(S))).to(dev)
        Le = np.linalg.cholesky(np.full((S, S), r_eff) + (1 - r_eff) * np.eye(S))
        y = torch.einsum("st,gtn->gsn", L, torch.randn((G, S, N), device=dev, generator=gen, dtype=torch.float64))
        for gi in rng.choice(G, size=int(round(0.1 * G)), replace=False):
            j = int(self.offs[gi] + rng.integers(P[gi]))
            b = torch.from_numpy(0.5 * (Le @ rng.normal(size=S))).to(dev)
            y[gi] += b[:, None] * self.geno[:, j].double()[None, :]
        self.y = y.contiguous()
        self.gene_in = torch.ones((G, S), dtype=torch.uint8, device=dev)
        self.geno_in = torch.ones((S, N), dtype=torch.uint8, device=dev)
        self.covars = torch.zeros((S, 1, N), dtype=torch.float64, device=dev)
        self.covar_count = torch.zeros((S,), dtype=torch.int32, device=dev)
        self.covar_in = torch.ones((S, N), dtype=torch.uint8, device=dev)
        self.gset = torch.zeros((S,), dtype=torch.int32, device=dev)
        self.snp_in = torch.ones((n_snps, S), dtype=torch.uint8, device=dev)
        self.pair_snp = torch.arange(n_snps, dtype=torch.int32, device=dev)
        self.rel = torch.from_numpy(self.offs.copy()).to(dev)
        self._rels = {}
        torch.cuda.synchronize()

    chunk = PermWorkload.chunk
    host_chunk = PermWorkload.host_chunk


def mvlr_leg(args, torch, result):
    """MVLR Bayes factors (--error mvlr --fiterr 0.5 --bfs gen, 50-point
    grid): 2,000 genes x ~800 cis SNPs, 9 subgroups sharing 150 individuals.
    Each (pair, grid point, gen / gen-fix / gen-maxh) is one 9 x 9 system:
    M = I + Vg^-1 Wg, partial-pivot LU, log-determinant and a quadratic form
    in FP64 (MVLR.cpp:559-605)."""
    from eqtlbma_amd import _abi, synth
    from eqtlbma_amd.engine import Engine, MvlrOpts
    from oracle import oracle as orc
    grid = synth.make_grid("general", n_aes=10)
    GLn = grid.shape[0]
    G = args.mvlr_genes
    wl = MvlrWorkload(torch, G, 1003)
    S, dev = wl.S, wl.dev
    NP = int(wl.offs[-1])
    eng = Engine(0)
    eng.set_grids(grid, grid)
    out_n = torch.empty((S, NP), dtype=torch.int32, device=dev)
    out_raw = torch.empty((3 * GLn, NP), dtype=torch.float64, device=dev)
    out_w = torch.empty((5, NP), dtype=torch.float64, device=dev)
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats, o.raw_abf, o.w_abf = out_n.data_ptr(), None, None, out_raw.data_ptr(), out_w.data_ptr()
    mo = MvlrOpts(_abi.BFS["gen"], 0.5)
    sptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    c = wl.chunk(_abi, 0, G)
    eng.assoc_mvlr_device(c, mo, o, sptr)  # warm-up
    torch.cuda.synchronize()
    eng.profile(True)
    before = eng.profile_read()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        eng.assoc_mvlr_device(c, mo, o, sptr)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: ((v[0] - before.get(k, (0.0, 0))[0]) / reps, (v[1] - before.get(k, (0.0, 0))[1]) // reps)
            for k, v in after.items() if v[1] > before.get(k, (0, 0))[1]}
    systems = float(NP) * 3 * GLn
    flops = mvlr_rows_flops(NP, 9, GLn)
    gk = kern.get("mvlr_rows_kernel", (float("nan"), 0))
    res = {"metric": "SNP-gene MVLR Bayes factors/s", "value": NP / dt, "unit": "pairs/s", "seconds_per_pass": dt,
           "workload": "MVLR: %d genes x ~800 cis SNPs (%d pairs), 9 subgroups sharing N=150, residual "
                       "equicorrelation 0.5, effects MVN with correlation 0.8, --fiterr 0.5 --bfs gen, 50-point "
                       "grid (150 9x9 systems per pair), FP64" % (G, NP),
           "kernels_ms_per_pass": {k: v[0] for k, v in kern.items()},
           "roofline": {"bound": "fp64", "kernel": "mvlr_rows_kernel", "systems": systems, "alg_flops": flops,
                        "achieved": flops / (gk[0] / 1e3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": flops / (gk[0] / 1e3) / 1e12 / 78.6, "avg_launch_ms": gk[0] / max(gk[1], 1)}}
    # oracle: parity on two genes, CPU baseline over the host threads
    model, T = host_cpu()
    h0 = wl.host_chunk(_abi, 0, 1)
    t = time.perf_counter()
    r0 = orc.assoc_mvlr(h0, "gen", 0.5, grid, grid)
    d1 = time.perf_counter() - t
    raw0 = out_raw[:, :h0.n_pairs].cpu().numpy()
    w0 = out_w[:, :h0.n_pairs].cpu().numpy()
    fr, fw = np.isfinite(r0.raw), np.isfinite(r0.w)
    res["error_vs_fp64_oracle"] = {"sample_pairs": int(h0.n_pairs),
                                   "log10_abf_raw_abs": float(np.max(np.abs(raw0[fr] - r0.raw[fr]))),
                                   "log10_abf_weighted_abs": float(np.max(np.abs(w0[fw] - r0.w[fw])))}
    k = min(G, T * max(1, int(args.cpu_seconds / max(d1, 1e-3))))
    hcs = [wl.host_chunk(_abi, g, g + 1) for g in range(k)]
    t = time.perf_counter()
    threaded(lambda h: orc.assoc_mvlr(h, "gen", 0.5, grid, grid), hcs, T)
    tc = time.perf_counter() - t
    cp = sum(h.n_pairs for h in hcs)
    res["cpu_baseline"] = dict(value=cp / tc, unit="pairs/s",
                               sample="first %d genes (%d pairs), one gene per task, %.1f s wall" % (k, cp, tc),
                               **cpu_label(model, T, "genes sharded over threads"))
    log("mvlr: %d pairs in %.3f s = %.3e pairs/s; kernels %s; cpu %.3e pairs/s" %
        (NP, dt, NP / dt, {k: round(v[0], 2) for k, v in kern.items()}, cp / tc))
    eng.close()
    del wl, out_n, out_raw, out_w
    torch.cuda.empty_cache()
    result["mvlr"] = res


def mvlr_rows_flops(NP, S, GL):
    """Algorithmic FP64 flops of mvlr_rows_kernel for --bfs gen: per pair,
    the gen and gen-maxh groups take GL rows each of: forming I + p2 H over
    the h = S(S+1)/2 + S - 1 entries of the Hessenberg form (2h), the
    pivoted Hessenberg LU with two right-hand sides (S^2 + 4S), two back
    substitutions (2 S^2), the rank-one update and the dot products (20 S);
    the gen-fix rows are scalar (4 flops).  The reduction to Hessenberg form
    (mvlr_hess_kernel, once per pair: 2 S^2 + 10/3 S^3 + 8 S^2) is not
    counted here."""
    h = S * (S + 1) // 2 + S - 1
    row = 2 * h + 3 * S * S + 20 * S
    return float(NP) * (2 * GL * row + 4 * GL)


class HybridWorkload(MvlrWorkload):
    """The MVLR configuration's shape with what --error hybrid is for: each
    subgroup misses the expression levels of 15 % of the individuals (NaN,
    drawn per gene and subgroup) and has 3 Gaussian covariates of its own."""

    Q = 3

    def __init__(self, torch, G, seed, miss=0.15):
        super().__init__(torch, G, seed)
        dev, S, N = self.dev, self.S, self.N
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + 1)
        self.covars = torch.randn((S, self.Q, N), device=dev, generator=gen, dtype=torch.float64)
        self.covar_count = torch.full((S,), self.Q, dtype=torch.int32, device=dev)
        self.y += 0.3 * self.covars[None, :, 0, :]
        drop = torch.rand(self.y.shape, device=dev, generator=gen) < miss
        self.y[drop] = float("nan")
        self.y = self.y.contiguous()
        torch.cuda.synchronize()

    def chunk(self, abi, g0, g1):
        c = PermWorkload.chunk(self, abi, g0, g1)
        c.n_covars_max = self.Q
        return c

    def host_chunk(self, abi, g0, g1):
        h = PermWorkload.host_chunk(self, abi, g0, g1)
        return abi.HostChunk(h.explevels, h.gene_in_sbgrp, h.geno_in_sbgrp, h.genotypes, h.snp_in_sbgrp,
                             h.pair_offset, h.pair_snp, covars=self.covars.cpu().numpy(),
                             covar_count=self.covar_count.cpu().numpy(), covar_in_sbgrp=self.covar_in.cpu().numpy())


def hybrid_flops(n, n12, S, Q):
    """Algorithmic FP64 flops of the per-pair sums of hyb_pair_kernel: per
    subgroup s, over its n[s] members, x^2, x y and (1, covariates) x for the
    subgroup's own and the pairwise covariates (2 + 2 (Q + 1) products); per
    pair of subgroups, over the n[s2] members of the second, x^2 and
    (1, covariates) x (Q + 2), and over the n12 common ones x^2, two x y and
    (1, covariates) x (Q + 4); a multiply-add counts 2."""
    f = 0
    k = 0
    for s1 in range(S):
        f += n[s1] * (2 + 2 * (Q + 1))
        for s2 in range(s1 + 1, S):
            f += n[s2] * (Q + 2) + n12[k] * (Q + 4)
            k += 1
    return 2 * f


def hybrid_leg(args, torch, result):
    """--error hybrid Bayes factors (--fiterr 0.5 --bfs gen, 50-point grid):
    2,000 genes x ~800 cis SNPs, 9 subgroups over 150 individuals, 15 % of
    the levels missing per subgroup, 3 covariates per subgroup
    (gene_snp_pair.cpp:1384-1423)."""
    from eqtlbma_amd import _abi, synth
    from eqtlbma_amd.engine import Engine, HybridOpts
    from oracle import oracle as orc
    grid = synth.make_grid("general", n_aes=10)
    GLn = grid.shape[0]
    G = args.hybrid_genes
    wl = HybridWorkload(torch, G, 1007)
    S, N, dev = wl.S, wl.N, wl.dev
    NP = int(wl.offs[-1])
    eng = Engine(0)
    eng.set_grids(grid, grid)
    out_n = torch.empty((S, NP), dtype=torch.int32, device=dev)
    out_ss = torch.empty((5, S, NP), dtype=torch.float64, device=dev)
    out_raw = torch.empty((3 * GLn, NP), dtype=torch.float64, device=dev)
    out_w = torch.empty((5, NP), dtype=torch.float64, device=dev)
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats = out_n.data_ptr(), out_ss.data_ptr(), None
    o.raw_abf, o.w_abf = out_raw.data_ptr(), out_w.data_ptr()
    ho = HybridOpts(_abi.BFS["gen"], 0.5, None, None)
    sptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    c = wl.chunk(_abi, 0, G)
    eng.assoc_hybrid_device(c, ho, o, sptr)  # warm-up
    torch.cuda.synchronize()
    eng.profile(True)
    before = eng.profile_read()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        eng.assoc_hybrid_device(c, ho, o, sptr)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: ((v[0] - before.get(k, (0.0, 0))[0]) / reps, (v[1] - before.get(k, (0.0, 0))[1]) // reps)
            for k, v in after.items() if v[1] > before.get(k, (0, 0))[1]}
    # members per (gene, subgroup) and common members per (gene, pair of subgroups)
    pres = (~torch.isnan(wl.y)).double()  # [G][S][N]
    n_gs = pres.sum(2).cpu().numpy()
    common = torch.einsum("gsn,gtn->gst", pres, pres).cpu().numpy()
    iu = np.triu_indices(S, 1)
    flops = 0.0
    for g in range(G):
        flops += float(wl.P[g]) * hybrid_flops(n_gs[g], common[g][iu], S, wl.Q)
    pk = kern.get("hyb_pair_kernel", (float("nan"), 0))
    gk = kern.get("mvlr_rows_kernel", (float("nan"), 0))
    systems = float(NP) * 3 * GLn
    rflops = mvlr_rows_flops(NP, S, GLn)
    res = {"metric": "SNP-gene hybrid Bayes factors/s", "value": NP / dt, "unit": "pairs/s",
           "seconds_per_pass": dt,
           "workload": "hybrid: %d genes x ~800 cis SNPs (%d pairs), 9 subgroups over N=150 with 15%% of the "
                       "levels missing per subgroup, 3 covariates per subgroup, --fiterr 0.5 --bfs gen, "
                       "50-point grid, FP64" % (G, NP),
           "kernels_ms_per_pass": {k: v[0] for k, v in kern.items()},
           "roofline": {"bound": "fp64", "kernel": "hyb_pair_kernel", "alg_flops": flops,
                        "achieved": flops / (pk[0] / 1e3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": flops / (pk[0] / 1e3) / 1e12 / 78.6, "avg_launch_ms": pk[0] / max(pk[1], 1)},
           "roofline_rows": {"bound": "fp64", "kernel": "mvlr_rows_kernel", "systems": systems, "alg_flops": rflops,
                             "achieved": rflops / (gk[0] / 1e3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                             "frac": rflops / (gk[0] / 1e3) / 1e12 / 78.6}}
    model, T = host_cpu()
    h0 = wl.host_chunk(_abi, 0, 1)
    t = time.perf_counter()
    r0 = orc.assoc_hybrid(h0, "gen", 0.5, grid, grid)
    d1 = time.perf_counter() - t
    raw0 = out_raw[:, :h0.n_pairs].cpu().numpy()
    w0 = out_w[:, :h0.n_pairs].cpu().numpy()
    fr, fw = np.isfinite(r0.raw), np.isfinite(r0.w)
    res["error_vs_fp64_oracle"] = {"sample_pairs": int(h0.n_pairs),
                                   "log10_abf_raw_abs": float(np.max(np.abs(raw0[fr] - r0.raw[fr]))),
                                   "log10_abf_weighted_abs": float(np.max(np.abs(w0[fw] - r0.w[fw]))),
                                   "n_equal": bool(np.array_equal(out_n[:, :h0.n_pairs].cpu().numpy(), r0.n))}
    if not args.no_cpu_baseline:
        k = min(G, T * max(1, int(args.cpu_seconds / max(d1, 1e-3))))
        hcs = [wl.host_chunk(_abi, g, g + 1) for g in range(k)]
        t = time.perf_counter()
        threaded(lambda h: orc.assoc_hybrid(h, "gen", 0.5, grid, grid), hcs, T)
        tc = time.perf_counter() - t
        cp = sum(h.n_pairs for h in hcs)
        res["cpu_baseline"] = dict(value=cp / tc, unit="pairs/s",
                                   sample="first %d genes (%d pairs), one gene per task, %.1f s wall" % (k, cp, tc),
                                   **cpu_label(model, T, "genes sharded over threads"))
        log("hybrid: cpu %.3e pairs/s" % (cp / tc))
    log("hybrid: %d pairs in %.3f s = %.3e pairs/s; kernels %s; pair-sums %.1f%% of FP64 peak" %
        (NP, dt, NP / dt, {k: round(v[0], 2) for k, v in kern.items()}, 100 * res["roofline"]["frac"]))
    eng.close()
    del wl, out_n, out_ss, out_raw, out_w
    torch.cuda.empty_cache()
    result["hybrid"] = res


def hm_leg(args, torch, result):
    """eqtlbma_hm's EM (configs model) on device-resident raw log10 ABFs: the
    rows eqtlbma_bf --bfs all writes for 5 subgroups (31 configurations) on a
    10-point grid, for 20,000 genes with cis SNPs ~ Poisson(1500) clipped to
    [200, 6000] (3.0e7 pairs, 74 GB of FP64).  Values N(-0.6, 0.4) plus, in
    15 % of genes, one SNP with U(1, 3) added to 60 % of its entries.  Each EM
    iteration is one pass over the raw ABFs; the timed run makes --hm-iters
    iterations (thresh below reach, --maxit iters + 1 as the reference counts
    them) from the default start."""
    from eqtlbma_amd.engine import Engine, HmOpts, HmParamsC
    from oracle import oracle as orc
    K, L, G = 31, 10, args.hm_genes
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1005)
    P = np.clip(rng.poisson(1500, G), 200, 6000).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    NP = int(off[-1])
    gen = torch.Generator(device=dev)
    gen.manual_seed(1005)
    raw = torch.empty((NP, K, L), dtype=torch.float64, device=dev)
    raw.normal_(-0.6, 0.4, generator=gen)
    eg = np.nonzero(rng.random(G) < 0.15)[0]
    pe = torch.from_numpy(off[eg] + rng.integers(0, P[eg])).to(dev)
    bump = torch.empty((eg.size, K, L), dtype=torch.float64, device=dev).uniform_(1.0, 3.0, generator=gen)
    bump *= (torch.rand((eg.size, K, L), device=dev, generator=gen) < 0.6)
    raw[pe] += bump
    off_d = torch.from_numpy(off).to(dev)
    torch.cuda.synchronize()
    eng = Engine(0)

    def params():
        grid, cfg = np.full(L, 1.0 / L), np.full(K, 1.0 / K)
        p = HmParamsC()
        p.pi0, p.grid, p.config = 0.5, grid.ctypes.data, cfg.ctypes.data
        return p, grid, cfg

    o = HmOpts()
    o.model, o.dim, o.n_subgroups, o.n_grid = 0, K, 5, L
    o.thresh, o.max_iters, o.stepmax = 1e-300, args.hm_iters, 1.0
    p, grid, cfg = params()
    o.max_iters = 2
    eng.hm_em_device(raw.data_ptr(), NP, off_d.data_ptr(), G, o, p)  # warm-up
    o.max_iters = args.hm_iters
    p, grid, cfg = params()
    eng.profile(True)
    before = eng.profile_read()
    t = time.perf_counter()
    lik, iters = eng.hm_em_device(raw.data_ptr(), NP, off_d.data_ptr(), G, o, p)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: ((v[0] - before.get(k, (0.0, 0))[0]), (v[1] - before.get(k, (0, 0))[1]))
            for k, v in after.items() if v[1] > before.get(k, (0, 0))[1]}
    passes = kern.get("hm_pair_kernel", (0.0, 1))[1]
    pk = kern.get("hm_pair_kernel", (float("nan"), 1))
    pair_ms = pk[0] / max(pk[1], 1)
    # algorithmic bytes of one pair-kernel launch: the raw ABFs read once, the
    # per-pair configuration, grid-point and SNP log10 sums written once
    alg = NP * (K * L + K + L + 1) * 8.0
    res = {"metric": "EM iterations x SNP-gene pairs per second", "value": iters * NP / dt, "unit": "pairs/s",
           "seconds": dt, "iterations": int(iters), "passes": int(passes), "log10_lik": lik,
           "ms_per_pass": dt * 1e3 / max(passes, 1),
           "workload": "eqtlbma_hm configs model: %d genes, cis SNPs~Pois(1500) in [200,6000] (%d pairs), 31 "
                       "configurations x 10 grid points of FP64 raw log10 ABFs (%.1f GB) resident in HBM, "
                       "default start, %d iterations" % (G, NP, NP * K * L * 8 / 1e9, iters),
           "kernels_ms_per_pass": {k: v[0] / max(v[1], 1) for k, v in kern.items()},
           "roofline": {"bound": "hbm", "kernel": "hm_pair_kernel", "alg_bytes_per_launch": alg,
                        "achieved": alg / (pair_ms / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                        "frac": alg / (pair_ms / 1e3) / 1e9 / 8000.0, "avg_launch_ms": pair_ms,
                        "traffic": _load_traffic_hm("hm_pair_kernel", G)}}
    avg_leg(args, torch, result, eng, raw, off, off_d, K, L, 5, p.pi0, grid, cfg)
    if args.no_cpu_baseline:
        eng.close()
        del raw, bump, off_d
        torch.cuda.empty_cache()
        result["hm"] = res
        log("hm: %d pairs, %d iterations in %.3f s (%.2f ms per pass); kernels %s" %
            (NP, iters, dt, dt * 1e3 / max(passes, 1), {k: round(v[0] / max(v[1], 1), 3) for k, v in kern.items()}))
        return
    # parity on the first genes and the oracle's rate over the host threads
    model, T = host_cpu()
    gs = 40
    raw_h = raw[:int(off[gs])].cpu().numpy()
    off_h = off[:gs + 1].copy()
    prm = orc.HmParams(0.5, np.full(L, 1.0 / L), cfg=np.full(K, 1.0 / K))
    ref, it_ref, _ = orc.hm_em(raw_h, off_h, prm, thresh=1e-300, maxit=3)
    e2 = Engine(0)
    g_pi0, g_grid, g_cfg, _s, g_lik, g_it = e2.hm_em(raw_h, off_h, 0.5, np.full(L, 1.0 / L), cfg=np.full(K, 1.0 / K),
                                                    thresh=1e-300, maxit=3)
    e2.close()
    res["error_vs_fp64_oracle"] = {"sample_genes": gs, "iterations": int(it_ref), "same_iterations": it_ref == g_it,
                                   "pi0_abs": abs(g_pi0 - ref.pi0),
                                   "grid_abs": float(np.max(np.abs(g_grid - ref.grid))),
                                   "config_abs": float(np.max(np.abs(g_cfg - ref.cfg)))}
    t = time.perf_counter()
    orc.hm_loglik(raw_h, off_h, prm)
    d1 = time.perf_counter() - t
    k = min(G, max(gs, int(gs * args.cpu_seconds * T / max(4 * d1, 1e-3))))
    k = max(gs, int(np.searchsorted(off, 4e9 / (K * L * 8.0), side="right")) - 1) if off[k] * K * L * 8.0 > 4e9 else k
    bounds = [(int(c[0]), int(c[-1]) + 1) for c in np.array_split(np.arange(k), min(T, k))]
    shards = [(raw[int(off[a]):int(off[b])].cpu().numpy(), off[a:b + 1] - off[a]) for a, b in bounds]

    def one(i):
        r_, o_ = shards[i]
        orc.hm_em(r_, o_, prm, thresh=1e-300, maxit=2)  # 3 passes: initial + 2 iterations
    t = time.perf_counter()
    threaded(one, list(range(len(shards))), T)
    tc = time.perf_counter() - t
    cp = int(off[k])
    res["cpu_baseline"] = dict(value=2 * cp / tc, unit="pairs/s",
                               sample="first %d genes (%d pairs), 2 EM iterations (3 likelihood passes) per gene "
                                      "shard, %.1f s wall" % (k, cp, tc),
                               **cpu_label(model, T, "gene ranges sharded over threads"))
    log("hm: %d pairs, %d iterations in %.3f s (%.2f ms per pass); kernels %s; cpu %.3e pairs/s" %
        (NP, iters, dt, dt * 1e3 / max(passes, 1), {k: round(v[0] / max(v[1], 1), 3) for k, v in kern.items()},
         2 * cp / tc))
    eng.close()
    del raw, bump, off_d
    torch.cuda.empty_cache()
    result["hm"] = res


def avg_leg(args, torch, result, eng, raw, off, off_d, K, L, S, pi0, grid, cfg):
    """eqtlbma_avg_bfs after the EM on the same device-resident raw log10
    ABFs (configs model, 31 configurations of 5 subgroups, 10 grid points,
    the EM's weights): per-configuration and SNP BFs, gene BFs and every
    posterior (--save bf+post --post a+b+c+d --alldim --bestdim)."""
    from eqtlbma_amd import _abi
    from eqtlbma_amd.engine import AvgOpts, AvgOut
    from oracle import oracle as orc
    G, NP = off.shape[0] - 1, int(off[-1])
    dev = raw.device
    masks = np.array(list(_abi.config_masks("all", S)), np.uint64)
    keep = np.arange(L, dtype=np.int32)
    gw, cw = np.ascontiguousarray(grid, dtype=np.float64), np.ascontiguousarray(cfg, dtype=np.float64)
    want = 1 | 2 | 4 | 8 | 16
    outs = {"gene_bf": torch.empty(G, dtype=torch.float64, device=dev),
            "gene_post": torch.empty(G, dtype=torch.float64, device=dev),
            "snp_bf": torch.empty(NP, dtype=torch.float64, device=dev),
            "dim_bf": torch.empty((NP, K), dtype=torch.float64, device=dev),
            "post_the": torch.empty(NP, dtype=torch.float64, device=dev),
            "post_an": torch.empty(NP, dtype=torch.float64, device=dev),
            "post_dim": torch.empty((NP, K), dtype=torch.float64, device=dev),
            "post_sub": torch.empty((NP, S), dtype=torch.float64, device=dev),
            "best_dim": torch.empty(NP, dtype=torch.int32, device=dev)}
    names = ("gene_bf", "gene_post", "snp_bf", "dim_bf", "post_the", "post_an", "post_dim", "post_sub", "best_dim")
    ao = AvgOut(*[outs[n].data_ptr() for n in names])
    opts = AvgOpts(0, K, S, L, L, keep.ctypes.data, gw.ctypes.data, cw.ctypes.data, masks.ctypes.data, None,
                   float(pi0), want)
    sptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    eng.avg_bfs_device(raw.data_ptr(), NP, off_d.data_ptr(), G, opts, ao, sptr)  # warm-up
    torch.cuda.synchronize()
    eng.profile(True)
    before = eng.profile_read()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        eng.avg_bfs_device(raw.data_ptr(), NP, off_d.data_ptr(), G, opts, ao, sptr)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    after = eng.profile_read()
    eng.profile(False)
    kern = {k: ((v[0] - before.get(k, (0.0, 0))[0]) / reps, (v[1] - before.get(k, (0, 0))[1]) // reps)
            for k, v in after.items() if v[1] > before.get(k, (0, 0))[1]}
    pk = kern.get("avg_pair_kernel", (float("nan"), 1))
    # algorithmic bytes of avg_pair_kernel: the raw ABFs read once, the
    # per-configuration BFs and posteriors, the subgroup posteriors, the SNP
    # BF and the best configuration written once
    alg = NP * (K * L * 8.0 + 2 * K * 8.0 + S * 8.0 + 8.0 + 4.0)
    res = {"metric": "SNP-gene pairs averaged per second (eqtlbma_avg_bfs)", "value": NP / dt, "unit": "pairs/s",
           "seconds_per_pass": dt,
           "workload": "eqtlbma_avg_bfs configs model on the eqtlbma_hm workload's raw ABFs (%d pairs, 31 "
                       "configurations x 10 grid points, FP64, resident in HBM), the EM's weights, every "
                       "posterior" % NP,
           "kernels_ms_per_pass": {k: v[0] for k, v in kern.items()},
           "roofline": {"bound": "hbm", "kernel": "avg_pair_kernel", "alg_bytes_per_launch": alg,
                        "achieved": alg / (pk[0] / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                        "frac": alg / (pk[0] / 1e3) / 1e9 / 8000.0, "avg_launch_ms": pk[0] / max(pk[1], 1),
                        "traffic": _load_traffic_hm("avg_pair_kernel", G)}}
    gs = 40
    raw_h = raw[:int(off[gs])].cpu().numpy()
    off_h = off[:gs + 1].copy()
    kw = dict(cfg_w=cw, cfg_sub=masks, pi0=float(pi0), want=("a", "b", "c", "d", "dims"), n_subgroups=S)
    t = time.perf_counter()
    ref = orc.avg_bfs(raw_h, off_h, gw, **kw)
    d1 = time.perf_counter() - t
    np_h = int(off[gs])
    err = {}
    for n in names[:-1]:
        a = outs[n][:gs].cpu().numpy() if n in ("gene_bf", "gene_post") else outs[n][:np_h].cpu().numpy()
        b = ref[n]
        fin = np.isfinite(b)
        err[n] = float(np.max(np.abs(a[fin] - b[fin]), initial=0))
    err["best_dim_equal"] = bool(np.array_equal(outs["best_dim"][:np_h].cpu().numpy(), ref["best_dim"]))
    res["error_vs_fp64_oracle"] = dict(sample_genes=gs, **err)
    if not args.no_cpu_baseline:
        model, T = host_cpu()
        k = min(G, max(gs, int(gs * args.cpu_seconds * T / max(d1, 1e-3))))
        if off[k] * K * L * 8.0 > 4e9:
            k = max(gs, int(np.searchsorted(off, 4e9 / (K * L * 8.0), side="right")) - 1)
        bounds = [(int(c[0]), int(c[-1]) + 1) for c in np.array_split(np.arange(k), min(T, k))]
        shards = [(raw[int(off[a]):int(off[b])].cpu().numpy(), off[a:b + 1] - off[a]) for a, b in bounds]
        t = time.perf_counter()
        threaded(lambda sh: orc.avg_bfs(sh[0], sh[1], gw, **kw), shards, T)
        tc = time.perf_counter() - t
        cp = int(off[k])
        res["cpu_baseline"] = dict(value=cp / tc, unit="pairs/s",
                                   sample="first %d genes (%d pairs), one averaging pass with every posterior, "
                                          "%.1f s wall" % (k, cp, tc),
                                   **cpu_label(model, T, "gene ranges sharded over threads"))
        del shards
    log("avg_bfs: %d pairs in %.1f ms; kernels %s; avg_pair_kernel %.1f%% of HBM peak" %
        (NP, dt * 1e3, {k: round(v[0], 2) for k, v in kern.items()}, 100 * res["roofline"]["frac"]))
    del outs
    torch.cuda.empty_cache()
    result["avg_bfs"] = res


def alg_bytes(P_g, S, N, GL, GS):
    """Algorithmic HBM bytes of the association pass for one gene: per pair
    the genotype row (4N), the raw Bayes factors in FP32 (3 GL + S GS), the
    grid-averaged ones in FP64 (3 + S + 1), plus the gene's expression
    levels once (8 S N)."""
    per_pair = 4 * N + 4 * (3 * GL + S * GS) + 8 * (3 + S + 1)
    return P_g * per_pair + 8 * S * N


def abf_bytes(P_g, S, GL, GS):
    """Algorithmic HBM bytes of the Bayes-factor kernel for one gene: per
    pair the standardised inputs of every subgroup (3 FP32 + 1 byte), the
    raw Bayes factors in FP32 and the grid-averaged ones in FP64."""
    return P_g * (13 * S + 4 * (3 * GL + S * GS) + 8 * (3 + S))


SUMS_FLOPS_PER_MEMBER = 2 * (4 + 2)  # x.x, x.y~ and x.U for 4 basis columns (3 covariates + 1)
# int8-digit sums: per (pair, member) the y~ and 4 U columns take ten digit
# products each and x'^2 four, two ops per multiply-add
SUMS_I8_OPS_PER_MEMBER = 2 * (10 * (1 + 4) + 4)
INT8_MATRIX_PEAK_TOPS = 5000.0


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine shards genes across GPUs with no "
              "collective on the hot path; --gpus %d is not run." % args.gpus)
        sys.exit(2)
    import torch
    from eqtlbma_amd import _abi, synth
    from eqtlbma_amd.engine import Engine, lib

    if args.only_hm:
        result = {"metric": "eqtlbma_hm EM only"}
        if args.calib_copy_mib > 0:
            src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device="cuda:0")
            dst = torch.empty_like(src)
            torch.neg(src, out=dst)  # reads and writes every byte once: the counters' scale
            torch.cuda.synchronize()
            log("calibration copy: %d MiB read + %d MiB written" % (args.calib_copy_mib, args.calib_copy_mib))
            del src, dst
        hm_leg(args, torch, result)
        print(json.dumps(result), flush=True)
        return
    if args.only_mvlr:
        result = {"metric": "MVLR and hybrid only"}
        mvlr_leg(args, torch, result)
        if not args.no_hybrid:
            hybrid_leg(args, torch, result)
        print(json.dumps(result), flush=True)
        return
    if args.only_hybrid:
        result = {"metric": "--error hybrid only"}
        hybrid_leg(args, torch, result)
        print(json.dumps(result), flush=True)
        return
    if args.only_perm:
        result = {"metric": METRIC, "n_gpus": 1, "data": "synthetic"}
        perm_leg(args, torch, result)
        if not args.no_alloc:
            perm_alloc_leg(args, torch, result)
        result.update({"value": result["perm"]["value"], "unit": "genes/s"})
        print(json.dumps(result), flush=True)
        return

    grid = synth.make_grid("general", n_aes=10)  # 50 points: 10 effect sizes x 5 homogeneities
    GL = GS = grid.shape[0]
    wl = Workload(torch, args.genes, args.seed)
    S, N = wl.S, wl.N
    max_np = wl.make_batches(args.batch_genes)
    n_raw = 3 * GL + S * GS
    n_w = 3 + S + 2
    dev = wl.dev
    out_n = torch.empty((S, max_np), dtype=torch.int32, device=dev)
    out_raw = torch.empty((n_raw, max_np), dtype=torch.float32, device=dev)
    out_w = torch.empty((n_w, max_np), dtype=torch.float64, device=dev)
    eng = Engine(0)
    eng.set_grids(grid, grid)
    eng.set_autotune(not args.no_autotune)  # the first call of the warm-up picks the sums kernel's launch shape
    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)  # an elementwise kernel (the only neg of the run): reads and writes every byte once
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read + %d MiB written" % (args.calib_copy_mib, args.calib_copy_mib))
        del src, dst
    # raw ABFs in FP32, 64-pair tiles (each workgroup writes one contiguous block)
    opts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["sin"], _abi.PREC["mixed"], _abi.F32_TILED)
    o = _abi.AssocOut()
    o.n, o.sstats, o.stdsstats = out_n.data_ptr(), None, None
    o.raw_abf, o.w_abf = out_raw.data_ptr(), out_w.data_ptr()
    chunks = [wl.chunk(_abi, g0, g1, rel) for g0, g1, rel in wl.batches]
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)

    def step():
        for c in chunks:
            eng.assoc_uvlr_device(c, opts, o, sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L = lib()
    L.eqbf_profile.argtypes = [C.c_void_p, C.c_int]
    L.eqbf_profile_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    L.eqbf_profile(eng._ctx, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.eqbf_profile(eng._ctx, 0)
    names = C.create_string_buffer(64 * 16)
    tot = np.zeros(16)
    cnt = np.zeros(16, np.int64)
    nk = C.c_int32(0)
    L.eqbf_profile_read(eng._ctx, names, tot.ctypes.data, cnt.ctypes.data, 16, C.byref(nk))
    kern = {}
    for k in range(nk.value):
        nm = names.raw[64 * k:64 * k + 64].split(b"\0")[0].decode()
        kern[nm] = (tot[k], int(cnt[k]))
    ms_step = (t1 - t0) * 1e3 / args.steps
    pairs = int(wl.offs[-1])
    value = pairs / (ms_step / 1e3)
    launches_per_step = len(chunks)
    members = int(wl.geno_in.sum().item())  # sum over subgroups of N_s (no missing expression levels)
    dominant = max(kern, key=lambda k: kern[k][0])
    dm_ms, dm_n = kern[dominant]
    avg_ms = dm_ms / max(dm_n, 1)
    i8_sums = "b8_convert_kernel" in kern
    if dominant == "uvlr_sums_kernel" and i8_sums:
        # int8-digit sums: the digit products of the member entries (the
        # non-member rows of the dense product, padding and empty column
        # slots are not counted)
        ops_per_launch = pairs * members * SUMS_I8_OPS_PER_MEMBER / launches_per_step
        achieved = ops_per_launch / (avg_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": INT8_MATRIX_PEAK_TOPS, "unit": "TOP/s (int8)",
                "frac": achieved / INT8_MATRIX_PEAK_TOPS, "traffic": None, "kernel": dominant,
                "avg_launch_ms": avg_ms, "alg_ops_per_launch": ops_per_launch,
                "alg_ops_per_pair": members * SUMS_I8_OPS_PER_MEMBER,
                "fp64_equivalent_tflops": pairs * members * SUMS_FLOPS_PER_MEMBER / launches_per_step / (avg_ms / 1e3) / 1e12}
    elif dominant == "uvlr_sums_kernel":
        # FP64 matrix-core sums: algorithmic flops = 12 per (pair, member of a subgroup)
        flops_per_launch = pairs * members * SUMS_FLOPS_PER_MEMBER / launches_per_step
        achieved = flops_per_launch / (avg_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_MATRIX_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_MATRIX_PEAK_TFLOPS, "traffic": None, "kernel": dominant,
                "avg_launch_ms": avg_ms, "alg_flops_per_launch": flops_per_launch,
                "alg_flops_per_pair": members * SUMS_FLOPS_PER_MEMBER}
    else:
        total_bytes = sum(alg_bytes(int(p), S, N, GL, GS) for p in wl.P)
        bytes_per_launch = total_bytes / launches_per_step
        achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant, "avg_launch_ms": avg_ms,
                "alg_bytes_per_launch": bytes_per_launch}
    others = {}
    if "uvlr_abf_kernel" in kern:
        a_ms = kern["uvlr_abf_kernel"][0] / max(kern["uvlr_abf_kernel"][1], 1)
        a_bytes = sum(abf_bytes(int(p), S, GL, GS) for p in wl.P) / launches_per_step
        others["uvlr_abf_kernel"] = {"bound": "hbm", "achieved": a_bytes / (a_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": a_bytes / (a_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                                     "avg_launch_ms": a_ms, "alg_bytes_per_launch": a_bytes,
                                     "traffic": _load_traffic("uvlr_abf_kernel", args)}
    traffic = _load_traffic(dominant, args)
    if traffic is not None:
        roof["traffic"] = traffic
    # the whole association pass against the HBM bytes it must move (genotype
    # rows, raw and grid-averaged Bayes factors, expression levels)
    path_bytes = sum(alg_bytes(int(p), S, N, GL, GS) for p in wl.P)
    path_roof = {"bound": "hbm", "achieved": path_bytes / (ms_step / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": path_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_step": path_bytes}
    log("step %.2f ms, %.3e pairs/s; kernels: %s" % (ms_step, value, {k: (round(v[0] / max(v[1], 1), 3), v[1]) for k, v in kern.items()}))

    # the same steps replayed from hipGraphs (first pass eager, second captured)
    graph_ms = None
    if args.steps > 0:
        eng.set_graphs(True)
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        tg = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - tg) * 1e3 / args.steps
        n_graphs, n_launch = eng.graph_stats()
        eng.set_graphs(False)
        log("graph replay: %.2f ms per step (%d graphs, %d launches)" % (graph_ms, n_graphs, n_launch))

    result = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("exact int8-digit sums (FP64 recompute of ill-conditioned items) + f32 Bayes-factor terms"
                  if i8_sums else "f64 sums + f32 Bayes-factor terms (mixed)"), "data": "synthetic",
        "config": {"workload": "batched ABFs: %d genes, 44 subgroups of N_s~U(80,400) from 500 "
                               "individuals, cis SNPs~Pois(1500) in [200,6000], FP32 dosages, 3 covariates, "
                               "--bfs sin, 50-point grid" % args.genes,
                   "pairs": pairs, "genes": args.genes, "subgroups": S, "samples": N,
                   "grid_points": GL, "chunk_genes": args.batch_genes,
                   "outputs": "sample sizes, FP32 raw ABFs (64-pair tiles), FP64 grid-averaged ABFs"},
        "roofline": roof,
        "roofline_path": path_roof,
        "roofline_other_kernels": others,
        "kernels_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in kern.items()},
        "autotune": eng.autotune_read(),
        "graph_replay_ms_per_step": graph_ms,
    }

    # ---- per-phase split of the dominant kernel on one chunk ----
    if args.diag:
        c0 = chunks[0]
        diag = {}
        for name, an, bfs in [("sumstats_only", "sep", "gen"), ("gen", "join", "gen"), ("sin", "join", "sin")]:
            dopts = _abi.AssocOpts(_abi.ANALYSIS[an], _abi.BFS[bfs], _abi.PREC["mixed"], _abi.F32)
            eng.assoc_uvlr_device(c0, dopts, o, sptr)
            torch.cuda.synchronize()
            L.eqbf_profile(eng._ctx, 1)
            for _ in range(3):
                eng.assoc_uvlr_device(c0, dopts, o, sptr)
            torch.cuda.synchronize()
            L.eqbf_profile(eng._ctx, 0)
            L.eqbf_profile_read(eng._ctx, names, tot.ctypes.data, cnt.ctypes.data, 16, C.byref(nk))
            cur = {}
            for k in range(nk.value):
                nm = names.raw[64 * k:64 * k + 64].split(b"\0")[0].decode()
                cur[nm] = (tot[k], int(cnt[k]))
            diag[name] = {k: (v[0] - kern.get(k, (0, 0))[0]) / max(v[1] - kern.get(k, (0, 0))[1], 1)
                          for k, v in cur.items()}
            kern = cur
        result["diag_ms_per_launch_first_chunk"] = diag
        log("diag %s" % diag)

    # ---- CPU baseline (oracle on the host cores) and parity sample ----
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        model, T = host_cpu()
        hc1 = wl.host_chunk(_abi, 0, 1)
        t = time.perf_counter()
        orc.assoc_uvlr(hc1, "join", "sin", grid, grid)
        d1 = time.perf_counter() - t
        k = min(wl.G, T * max(1, int(args.cpu_seconds / max(d1, 1e-3))))
        hcs = [wl.host_chunk(_abi, g, g + 1) for g in range(k)]
        t = time.perf_counter()
        refs = threaded(lambda h: orc.assoc_uvlr(h, "join", "sin", grid, grid), hcs, T)
        dt = time.perf_counter() - t
        cpu_pairs = sum(h.n_pairs for h in hcs)
        result["cpu_baseline"] = dict(value=cpu_pairs / dt, unit="pairs/s",
                                      sample="first %d genes of the workload (%d pairs), one gene per task, "
                                             "%.1f s wall" % (k, cpu_pairs, dt),
                                      **cpu_label(model, T, "genes sharded over threads"))
        errs = {}
        for h, ref in list(zip(hcs, refs))[:5]:
            gpu = eng.assoc_uvlr(h, analysis="join", bfs="sin", precision="mixed", raw_dtype=np.float32)
            # the timed configuration: Bayes factors without the FP64 statistic outputs
            gbf = eng.assoc_uvlr(h, analysis="join", bfs="sin", precision="mixed", raw_dtype=np.float32,
                                 want_sstats=False, want_std=False)
            has = ref.n > 0
            bet, se, pv = ref.sstats[2], ref.sstats[3], ref.sstats[4]
            ok = has & ~np.isnan(bet)
            okp = ok & (pv > 1e-290)
            fr, fw = np.isfinite(ref.raw), np.isfinite(ref.w)
            e = {"betahat_rel_to_max(|b|,se)": np.max(np.abs(gpu.sstats[2][ok] - bet[ok]) /
                                                      np.maximum(np.abs(bet[ok]), se[ok])),
                 "se_rel": np.max(np.abs(gpu.sstats[3][ok] - se[ok]) / se[ok]),
                 "pval_rel": np.max(np.abs(gpu.sstats[4][okp] - pv[okp]) / pv[okp]),
                 "log10_abf_raw_abs": np.max(np.abs(gbf.raw.astype(np.float64)[fr] - ref.raw[fr])),
                 "log10_abf_weighted_abs": np.max(np.abs(gbf.w[fw] - ref.w[fw])),
                 "log10_abf_raw_abs_with_fp64_statistics": np.max(np.abs(gpu.raw.astype(np.float64)[fr] - ref.raw[fr])),
                 "log10_abf_weighted_abs_with_fp64_statistics": np.max(np.abs(gpu.w[fw] - ref.w[fw]))}
            for kk, v in e.items():
                errs[kk] = max(errs.get(kk, 0.0), float(v))
        errs["sample_pairs"] = int(sum(h.n_pairs for h in hcs[:5]))
        result["error_vs_fp64_oracle"] = errs
        log("cpu oracle %.1f pairs/s on %d threads; errors %s" % (cpu_pairs / dt, T, errs))

    # ---- plain FP64 one-gene-per-launch path on a subset ----
    if not args.no_plain and args.plain_genes > 0:
        gp = min(args.plain_genes, wl.G)
        rel = torch.from_numpy(wl.offs[:gp + 1].copy()).to(dev)
        c = wl.chunk(_abi, 0, gp, rel)
        popts = _abi.AssocOpts(_abi.ANALYSIS["join"], _abi.BFS["sin"], _abi.PREC["fp64_plain"], _abi.F32)
        eng.assoc_uvlr_device(c, popts, o, sptr)  # warm
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.assoc_uvlr_device(c, popts, o, sptr)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        plain_rate = c.n_pairs / dt
        # mixed path on the same subset, same call shape
        eng.assoc_uvlr_device(c, opts, o, sptr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.assoc_uvlr_device(c, opts, o, sptr)
        torch.cuda.synchronize()
        dtm = time.perf_counter() - t
        result["plain_gpu"] = {"value": plain_rate, "unit": "pairs/s",
                               "sample": "first %d genes (%d pairs), FP64, one launch per gene" % (gp, c.n_pairs),
                               "mixed_same_sample": c.n_pairs / dtm,
                               "speedup_mixed_vs_plain_same_sample": (c.n_pairs / dtm) / plain_rate,
                               "speedup_full_run_vs_plain": value / plain_rate}
        log("plain %.3e pairs/s" % plain_rate)

    eng.close()
    del wl, chunks, out_n, out_raw, out_w
    torch.cuda.empty_cache()
    if not args.no_c0:
        c0_leg(args, torch, result)
    if not args.no_mvlr:
        mvlr_leg(args, torch, result)
    if not args.no_hybrid:
        hybrid_leg(args, torch, result)
    if not args.no_hm:
        hm_leg(args, torch, result)
    if not args.no_perm:
        perm_leg(args, torch, result)
        if not args.no_alloc:
            perm_alloc_leg(args, torch, result)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
