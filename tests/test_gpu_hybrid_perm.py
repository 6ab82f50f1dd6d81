"""Permutations with --error hybrid (Gene::MakePermutationsJoin, gene.cpp:
598-717, the hybrid branch :669-672) against the oracle.

Each permutation's statistic comes from the hybrid Bayes factors with the
per-subgroup fits on the permuted expression levels and the pairwise step on
the levels as loaded (CalcOffDiagCovarsFromPairsOfSubgroups takes no
permutation, gene_snp_pair.cpp:1157-1159); a SNP without genotypes in every
subgroup contributes 0.  FP64 on both sides: 1e-8 absolute on each
statistic (log10 BF).  The end-to-end case runs eqtlbma_bf --error hybrid
--nperm --trick 2 and replays the reference's counting loop with the oracle.
"""
import os
import sys

import numpy as np
import pytest

from eqtlbma_amd import engine, synth
import cli_case
from test_gpu_cli import _close, _common_args, _num, _run
from test_gpu_cli_trick import _advance, _stats_from

pytestmark = pytest.mark.gpu

GL = synth.make_grid("general")[::3]   # 9 points
GS = synth.make_grid("small")[::2]     # 5 points


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = engine.Engine(0)
    e.set_grids(GL, GS)
    yield e
    e.close()


@pytest.mark.parametrize("pbf,use_max,pairwise_inputs", [("all", True, False), ("gen", False, True),
                                                         ("gen-sin", True, False)])
def test_hybrid_perm_stats_match_oracle(eng, orc, pbf, use_max, pairwise_inputs):
    hc = synth.hybrid(seed=71, S=3, N=48, n_genes=3, P=8, Q=1, miss=0.15)
    hc.snp_in_sbgrp[hc.pair_snp[hc.pair_offset[1] + 2], 1] = 0  # skipped by hybrid: contributes 0
    nperm = 12
    tables = engine.Rng(29).tables(hc.N, 3, nperm)
    gene_table = np.array([0, 1, 2], np.int32)
    kw = {}
    if pairwise_inputs:
        rng = np.random.default_rng(72)
        kw = dict(covars_pos=rng.normal(size=hc.covars.shape), expl_pair=hc.explevels * 1.3 + 0.1)
    _say("[gpu] hybrid perm pbf=%s max=%s pairwise inputs=%s" % (pbf, use_max, pairwise_inputs))
    gpu = eng.perm_stats(hc, tables, gene_table, "join", pbf, use_max, precision="mixed", error_model="hybrid",
                         fiterr=0.5, **kw)
    ref = np.array([orc.perm_join_stats_hybrid(hc, g, pbf, use_max, 0.5, GL, GS, tables[g], **kw)
                    for g in range(3)])
    assert np.array_equal(np.isnan(gpu), np.isnan(ref))
    fin = np.isfinite(ref)
    err = float(np.max(np.abs(gpu[fin] - ref[fin])))
    _say("[gpu]   max |stat - oracle| = %.3g" % err)
    assert err <= 1e-8


def test_cli_hybrid_permutations_trick2(tmp_path, orc):
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=37, S=3, N=40, one_geno=True)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "hyp")
    nperm, seed, cutoff = 20, 13, 3
    _run(_common_args(root, out, "join") + ["--covar", os.path.join(root, "list_cov.txt"), "--error", "hybrid",
                                            "--fiterr", "0.5", "--bfs", "all", "--nperm", str(nperm), "--seed",
                                            str(seed), "--pbf", "all", "--maxbf", "--trick", "2", "--tricut",
                                            str(cutoff)], root)
    gL, gS = c["gL"], c["gS"]
    nb = nperm
    exp = []
    for names, hc, _psn in cli_case.expected_chunks(inp, 2500, 2, mvlr=True):
        ref = orc.assoc_hybrid(hc, "all", 0.5, gL, gS, covars_pos=inp["covars_pos"]) if hc.n_pairs else None
        mt = orc.MT(seed)
        mtT = orc.MT(seed)
        for g, gname in enumerate(names):
            p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
            if p1 == p0:
                continue
            true = orc.true_stat(ref.w[3 + 7 + 1, p0:p1], True)
            st = _stats_from(orc, mt, nb, hc.N, lambda t: orc.perm_join_stats_hybrid(
                hc, g, "all", True, 0.5, gL, gS, t, covars_pos=inp["covars_pos"]))
            pv, used, _med, nb, drawn = orc.perm_count(st, true, nb, 2, cutoff, mtT)
            _advance(mt, drawn, hc.N)
            exp.append([gname, p1 - p0, pv, used, true])
    rows = cli_case.read_table(out + "_joinPermPvals.txt.gz")
    assert len(rows) - 2 == len(exp)
    for got, want in zip(rows[2:], exp):
        assert got[0] == want[0] and int(got[1]) == want[1]
        assert int(got[3]) == want[3], (got, want)
        assert _num(got[2]) == pytest.approx(want[2], rel=1e-5), (got, want)
        assert _close(_num(got[4]), want[4]), (got, want)
    _say("[gpu]   %d genes' permutation p-values match" % len(exp))
