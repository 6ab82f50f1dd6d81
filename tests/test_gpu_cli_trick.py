"""End to end: permutations through the eqtlbma_bf host program with
--trick 1 / 2 and NaN permutation statistics, every gene's p-value and
number of permutations against the oracle's sequential replay.

The reference draws every gene's shuffles from one generator reseeded per
chunk and shares nb_permutations along the run: a NaN statistic decrements
it, so later genes get fewer permutations, and --trick 1 stops a gene after
tricut + 1 exceedances, so the next gene's shuffles start earlier in the
stream (Gene::MakePermutationsJoin, gene.cpp:598-717; makePermutationsJoin,
eqtlbma_bf.cpp:826-864).  The expected side replays exactly that with the
oracle: for each gene the statistics of nb_permutations shuffles drawn from
a copy of the stream, the counting loop (orc_perm_count, gene.cpp:348-364
and :677-703), then the stream advanced by the shuffles the gene drew.

The inputs: uneven sample sets, NA expression levels, and the first cis SNP
of geneB monomorphic but for one carrier, so that with --analys sep
--permsep 2 every permutation that leaves the carrier without an
expression level in a subgroup has a NaN p-value at the first SNP, hence a
NaN statistic (std::min_element, gene.cpp:429).  (A NaN subgroup statistic
makes a Bayes factor 0, not NaN, in CalcLog10AbfUvlr,
gene_snp_pair.cpp:297-356, so join statistics are not NaN here.)
"""
import os
import sys

import numpy as np
import pytest

import cli_case
from test_gpu_cli import _close, _common_args, _num, _run

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("cli_trick"))
    c = cli_case.make_case(root, seed=31, covariates=False, near_mono=0)
    return c, cli_case.load_expected_inputs(c)


def _stats_from(orc, mt, nb_cur, N, fn):
    """Statistics of nb_cur cumulative shuffles drawn from a copy of mt."""
    if nb_cur <= 0:
        return np.zeros(1)
    tab = np.zeros((nb_cur, N), np.int32)
    m2 = mt.copy()
    perm = np.arange(N)
    for p in range(nb_cur):
        perm = m2.shuffle(perm)
        tab[p] = perm
    return fn(tab)


def _advance(mt, drawn, N):
    perm = np.arange(N)
    for _ in range(drawn):
        perm = mt.shuffle(perm)


def _expected_join(orc, c, inp, seed, nperm, trick, cutoff, wrtsize):
    gL, gS = c["gL"], c["gS"]
    nb = nperm
    out = []
    for names, hc, _psn in cli_case.expected_chunks(inp, 2500, wrtsize):
        ref = orc.assoc_uvlr(hc, "join", "gen", gL, gS)
        mt = orc.MT(seed)
        mtT = orc.MT(seed) if trick else None
        for g, gname in enumerate(names):
            p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
            if p1 == p0:
                continue
            true = orc.true_stat(ref.w[0, p0:p1], True)
            st = _stats_from(orc, mt, nb, hc.N, lambda t: orc.perm_join_stats(hc, g, "gen", True, gL, gS, t))
            pv, used, _med, nb, drawn = orc.perm_count(st, true, nb, trick, cutoff, mtT)
            _advance(mt, drawn, hc.N)
            out.append([gname, p1 - p0, pv, used, true])
    return {0: out}, 0


def _expected_mvlr(orc, c, inp, seed, nperm, trick, cutoff, wrtsize):
    gL, gS = c["gL"], c["gS"]
    nb = nperm
    out = []
    for names, hc, _psn in cli_case.expected_chunks(inp, 2500, wrtsize, mvlr=True):
        ref = orc.assoc_mvlr(hc, "all", 0.5, gL, gS) if hc.n_pairs else None
        mt = orc.MT(seed)
        mtT = orc.MT(seed) if trick else None
        for g, gname in enumerate(names):
            p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
            if p1 == p0:
                continue
            true = orc.true_stat(ref.w[3 + 7 + 1, p0:p1], True)
            st = _stats_from(orc, mt, nb, hc.N,
                             lambda t: orc.perm_join_stats_mvlr(hc, g, "all", True, 0.5, gL, gS, t))
            pv, used, _med, nb, drawn = orc.perm_count(st, true, nb, trick, cutoff, mtT)
            _advance(mt, drawn, hc.N)
            out.append([gname, p1 - p0, pv, used, true])
    return {0: out}, 0


def test_cli_mvlr_permutations_trick2(tmp_path, orc):
    # --error mvlr --nperm --pbf all --maxbf --trick 2 (the manual's recipe)
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=14, S=3, N=30, mvlr=True)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "mvp")
    nperm, seed, cutoff = 30, 9, 3
    _run(_common_args(root, out, "join") + ["--covar", os.path.join(root, "list_cov.txt"), "--error", "mvlr",
                                            "--fiterr", "0.5", "--bfs", "all", "--nperm", str(nperm), "--seed",
                                            str(seed), "--pbf", "all", "--maxbf", "--trick", "2", "--tricut",
                                            str(cutoff)], root)
    exp, _ = _expected_mvlr(orc, c, inp, seed, nperm, 2, cutoff, 2)
    rows = cli_case.read_table(out + "_joinPermPvals.txt.gz")
    assert len(rows) - 2 == len(exp[0])
    for got, want in zip(rows[2:], exp[0]):
        assert got[0] == want[0] and int(got[1]) == want[1]
        assert int(got[3]) == want[3], (got, want)
        assert _num(got[2]) == pytest.approx(want[2], rel=1e-5), (got, want)
        assert _close(_num(got[4]), want[4]), (got, want)


def _expected_sep2(orc, c, inp, seed, nperm, trick, cutoff, wrtsize):
    S = len(inp["sbgrps"])
    nb = nperm
    out = {k: [] for k in range(S)}
    n_nan = 0
    for names, hc, _psn in cli_case.expected_chunks(inp, 2500, wrtsize):
        ref = orc.assoc_uvlr(hc, "sep", "gen", c["gL"], c["gS"])
        qual = [g for g in range(hc.G) if hc.pair_offset[g + 1] > hc.pair_offset[g]]
        for k in range(S):
            mt = orc.MT(seed)
            mtT = orc.MT(seed) if trick else None
            for g in qual:
                p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
                true = 1.0
                for p in range(p0, p1):
                    if ref.n[k, p] > 0 and ref.sstats[4, k, p] < true:
                        true = ref.sstats[4, k, p]
                nb_cur = nb
                st = _stats_from(orc, mt, nb_cur, hc.N, lambda t: orc.perm_sep_stats(hc, g, k, t))
                n_nan += int(np.isnan(st[:nb_cur]).sum()) if nb_cur > 0 else 0
                pv, used, _m, nb, drawn = orc.perm_count(st, true, nb_cur, trick, cutoff, mtT, sep=True)
                _advance(mt, drawn, hc.N)
                out[k].append([names[g], int((ref.n[k, p0:p1] > 0).sum()), pv, used, true])
    return out, n_nan


@pytest.mark.parametrize("analysis,trick,precision", [("join", 0, "mixed"), ("join", 1, "mixed"),
                                                      ("join", 2, "mixed"), ("join", 1, "fp64"),
                                                      ("sep2", 0, "mixed"), ("sep2", 1, "mixed"),
                                                      ("sep2", 2, "fp64")])
def test_cli_trick_and_nan_permutations(case, orc, analysis, trick, precision):
    c, inp = case
    root = c["root"]
    out = os.path.join(root, "%s_t%d_%s" % (analysis, trick, precision))
    nperm, seed, cutoff = 40, 17, 3
    if analysis == "join":
        args = _common_args(root, out, "join") + ["--bfs", "gen", "--pbf", "gen", "--maxbf"]
    else:
        args = _common_args(root, out, "sep") + ["--permsep", "2"]
    args += ["--nperm", str(nperm), "--seed", str(seed)]
    args[args.index("--precision") + 1] = precision
    if trick:
        args += ["--trick", str(trick), "--tricut", str(cutoff)]
    _run(args, root)
    fn = _expected_join if analysis == "join" else _expected_sep2
    exp, n_nan = fn(orc, c, inp, seed, nperm, trick, cutoff, 2)
    _say("[gpu]   NaN permutation statistics on the expected side: %d" % n_nan)
    if analysis == "sep2":
        assert n_nan > 0
    for k, rows_exp in exp.items():
        name = out + ("_joinPermPvals.txt.gz" if analysis == "join" else
                      "_sepPermPvals_%s.txt.gz" % inp["sbgrps"][k])
        rows = cli_case.read_table(name)
        assert len(rows) - 2 == len(rows_exp)
        for got, want in zip(rows[2:], rows_exp):
            assert got[0] == want[0] and int(got[1]) == want[1]
            assert int(got[3]) == want[3], (got, want)
            assert _num(got[2]) == pytest.approx(want[2], rel=1e-5), (got, want)
            assert _close(_num(got[4]), want[4]), (got, want)
