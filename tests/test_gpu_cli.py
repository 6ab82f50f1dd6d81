"""End to end: the eqtlbma_bf host program on files, against the oracle.

The inputs (tests/cli_case.py) exercise uneven sample sets and column
orders per subgroup, NA expression levels, a genotype line with 'NA' (skipped
for that subgroup) and one with 'nan' (SNP dropped for that subgroup),
covariates sorted by name, genes without coordinates, without expression or
without cis SNPs, a gene absent from one subgroup, and chunks of --wrtsize
genes.  The program runs with --precision fp64; the printed values (%.6e,
seven significant digits) must match the oracle's to 2e-6 relative (one unit
in the last printed digit) or 1e-9 absolute, and the permutation counts and
p-values exactly.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from eqtlbma_amd import CLI_PATH, synth
import cli_case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args, cwd):
    _say("[gpu] eqtlbma_bf " + " ".join(args))
    r = subprocess.run([CLI_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _num(s):
    return float(s)


def _close(a, b):
    if math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= 1e-9 + 2e-6 * abs(b)


def _common_args(root, out, analys):
    return ["--geno", os.path.join(root, "list_geno.txt"), "--scoord", os.path.join(root, "snps.bed"),
            "--exp", os.path.join(root, "list_exp.txt"), "--gcoord", os.path.join(root, "genes.bed"),
            "--cis", "2500", "--out", out, "--analys", analys, "--gridL", os.path.join(root, "gridL.txt"),
            "--gridS", os.path.join(root, "gridS.txt"), "--precision", "fp64", "--wrtsize", "2", "-v", "0"]


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("cli"))
    c = cli_case.make_case(root)
    inp = cli_case.load_expected_inputs(c)
    return c, inp


def test_cli_join_all_with_permutations(case, orc):
    c, inp = case
    root = c["root"]
    out = os.path.join(root, "join")
    nperm, seed = 25, 11
    _run(_common_args(root, out, "join") + ["--covar", os.path.join(root, "list_cov.txt"), "--bfs", "all",
                                            "--outss", "--outw", "--nperm", str(nperm), "--seed", str(seed),
                                            "--pbf", "all", "--maxbf"], root)
    S = len(inp["sbgrps"])
    labels = cli_case.config_labels("all", S)
    gL, gS = c["gL"], c["gS"]
    exp_ss = {s: [] for s in range(S)}
    exp_raw, exp_avg, exp_perm = [], [], []
    for names, hc, psn in cli_case.expected_chunks(inp, 2500, 2):
        ref = orc.assoc_uvlr(hc, "join", "all", gL, gS)
        tables = orc.perm_tables(seed, hc.N, sum(hc.pair_offset[g + 1] > hc.pair_offset[g] for g in range(hc.G)),
                                 nperm)
        t = 0
        for g, gname in enumerate(names):
            p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
            for p in range(p0, p1):
                for s in range(S):
                    if ref.n[s, p] > 0:
                        exp_ss[s].append([gname, psn[p], inp["geno_vals"][s][psn[p]][1], ref.n[s, p]] +
                                         list(ref.sstats[:, s, p]))
                for k, typ in enumerate(("gen", "gen-fix", "gen-maxh")):
                    exp_raw.append([gname, psn[p], typ] + list(ref.raw[k * len(gL):(k + 1) * len(gL), p]))
                for ci, lab in enumerate(labels):
                    r0 = 3 * len(gL) + ci * len(gS)
                    exp_raw.append([gname, psn[p], lab] + list(ref.raw[r0:r0 + len(gS), p]) +
                                   [float("nan")] * (len(gL) - len(gS)))
                nc = len(labels)
                exp_avg.append([gname, psn[p], int((ref.n[:, p] > 0).sum()), ref.w[0, p], ref.w[1, p], ref.w[2, p],
                                ref.w[3 + nc, p], ref.w[3 + nc + 1, p]] + [ref.w[3 + k, p] for k in range(nc)])
            if p1 > p0:
                true = orc.true_stat(ref.w[3 + len(labels) + 1, p0:p1], True)
                st = orc.perm_join_stats(hc, g, "all", True, gL, gS, tables[t])
                pv, used, med, _nb, _dr = orc.perm_count(st, true, nperm)
                exp_perm.append([gname, p1 - p0, pv, used, true, med])
                t += 1
    # sumstats
    for s, sg in enumerate(inp["sbgrps"]):
        rows = cli_case.read_table(out + "_sumstats_%s.txt.gz" % sg)
        assert rows[0] == ["gene", "snp", "maf", "n", "pve", "sigmahat", "betahat.geno", "sebetahat.geno",
                           "betapval.geno"]
        assert len(rows) - 1 == len(exp_ss[s])
        for got, want in zip(rows[1:], exp_ss[s]):
            assert got[:2] == want[:2] and int(got[3]) == want[3]
            assert all(_close(_num(a), b) for a, b in zip([got[2]] + got[4:], [want[2]] + want[4:])), (got, want)
    # raw ABFs
    rows = cli_case.read_table(out + "_l10abfs_raw.txt.gz")
    assert rows[0] == ["gene", "snp", "config"] + ["l10abf.grid%d" % (i + 1) for i in range(len(gL))]
    assert len(rows) - 1 == len(exp_raw)
    for got, want in zip(rows[1:], exp_raw):
        assert got[:3] == want[:3]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    # grid averages
    rows = cli_case.read_table(out + "_l10abfs_avg-grids.txt.gz")
    assert rows[0][:8] == ["gene", "snp", "nb.subgroups", "l10abf.gen", "l10abf.gen.fix", "l10abf.gen.maxh",
                           "l10abf.gen.sin", "l10abf.all"]
    assert rows[0][8:] == ["l10abf." + l for l in labels]
    assert len(rows) - 1 == len(exp_avg)
    for got, want in zip(rows[1:], exp_avg):
        assert got[:2] == want[:2] and int(got[2]) == want[2]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    # permutation p-values
    rows = cli_case.read_table(out + "_joinPermPvals.txt.gz")
    assert rows[0] == ["# perm.bf=all seed=%d" % seed]
    assert rows[1] == ["gene", "nb.snps", "join.perm.pval", "nb.permutations", "true.l10abf", "med.perm.l10abf"]
    assert len(rows) - 2 == len(exp_perm)
    for got, want in zip(rows[2:], exp_perm):
        assert got[0] == want[0] and int(got[1]) == want[1] and int(got[3]) == want[3]
        assert _num(got[2]) == pytest.approx(want[2], rel=1e-5)
        assert _close(_num(got[4]), want[4]) and _close(_num(got[5]), want[5])


@pytest.mark.parametrize("permsep", [1, 2])
def test_cli_sep_with_permutations(case, orc, permsep):
    c, inp = case
    root = c["root"]
    out = os.path.join(root, "sep%d" % permsep)
    nperm, seed = 20, 5
    _run(_common_args(root, out, "sep") + ["--covar", os.path.join(root, "list_cov.txt"), "--nperm", str(nperm),
                                           "--seed", str(seed), "--permsep", str(permsep)], root)
    S = len(inp["sbgrps"])
    passes = 1 if permsep == 1 else S
    expected = {k: [] for k in range(passes)}
    for names, hc, psn in cli_case.expected_chunks(inp, 2500, 2):
        ref = orc.assoc_uvlr(hc, "sep", "gen", c["gL"], c["gS"])
        qual = [g for g in range(hc.G) if hc.pair_offset[g + 1] > hc.pair_offset[g]]
        for k in range(passes):
            tables = orc.perm_tables(seed, hc.N, len(qual), nperm)
            for t, g in enumerate(qual):
                p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
                subs = range(S) if permsep == 1 else [k]
                true = 1.0
                for p in range(p0, p1):
                    for s in subs:
                        if ref.n[s, p] > 0 and ref.sstats[4, s, p] < true:
                            true = ref.sstats[4, s, p]
                st = orc.perm_sep_stats(hc, g, -1 if permsep == 1 else k, tables[t])
                pv, used, _m, _nb, _dr = orc.perm_count(st, true, nperm, sep=True)
                nsnp = p1 - p0 if permsep == 1 else int((ref.n[k, p0:p1] > 0).sum())
                expected[k].append([names[g], nsnp, pv, used, true])
    for k in range(passes):
        name = out + ("_sepPermPvals.txt.gz" if permsep == 1 else "_sepPermPvals_%s.txt.gz" % inp["sbgrps"][k])
        rows = cli_case.read_table(name)
        assert rows[0] == ["# seed=%d" % seed]
        assert len(rows) - 2 == len(expected[k])
        for got, want in zip(rows[2:], expected[k]):
            assert got[0] == want[0] and int(got[1]) == want[1] and int(got[3]) == want[3]
            assert _num(got[2]) == pytest.approx(want[2], rel=1e-5)
            assert _close(_num(got[4]), want[4])


def test_cli_mvlr(tmp_path, orc):
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=9, S=3, N=30, mvlr=True)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "mv")
    _run(_common_args(root, out, "join") + ["--covar", os.path.join(root, "list_cov.txt"), "--error", "mvlr",
                                            "--fiterr", "0.5", "--bfs", "sin", "--outw"], root)
    S = len(inp["sbgrps"])
    labels = cli_case.config_labels("sin", S)
    gL, gS = c["gL"], c["gS"]
    exp_raw = []
    for names, hc, psn in cli_case.expected_chunks(inp, 2500, 2, mvlr=True):
        if hc.n_pairs == 0:
            continue
        # the reference's MVLR uses the first subgroup's covariates
        ref = orc.assoc_mvlr(hc, "sin", 0.5, gL, gS)
        for g, gname in enumerate(names):
            for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
                for k, typ in enumerate(("gen", "gen-fix", "gen-maxh")):
                    exp_raw.append([gname, psn[p], typ] + list(ref.raw[k * len(gL):(k + 1) * len(gL), p]))
                for ci, lab in enumerate(labels):
                    r0 = 3 * len(gL) + ci * len(gS)
                    exp_raw.append([gname, psn[p], lab] + list(ref.raw[r0:r0 + len(gS), p]) +
                                   [float("nan")] * (len(gL) - len(gS)))
    rows = cli_case.read_table(out + "_l10abfs_raw.txt.gz")
    assert len(rows) - 1 == len(exp_raw)
    for got, want in zip(rows[1:], exp_raw):
        assert got[:3] == want[:3]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    assert not os.path.exists(out + "_sumstats_s1.txt.gz")  # not written with --error mvlr


def test_cli_join_permutations_contraction_path(tmp_path, orc):
    # every sample in every subgroup, one genotype file of hard calls, no
    # covariates: the permutation statistics take the contraction path
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=12, S=3, N=40, mvlr=True, covariates=False)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "jc")
    nperm, seed = 30, 21
    r = _run(_common_args(root, out, "join") + ["--bfs", "all", "--nperm", str(nperm), "--seed", str(seed),
                                                "--pbf", "all", "--maxbf", "--precision", "mixed", "--ktimes"], root)
    kernels = {l.split("\t")[1] for l in r.stderr.splitlines() if l.startswith("kernel\t")}
    _say("[gpu]   kernels: %s" % sorted(kernels))
    assert "pf_gemm_kernel" in kernels and "pf_bma_kernel" in kernels
    assert "perm_gather_kernel" not in kernels  # no gene fell back to the general path
    S = len(inp["sbgrps"])
    labels = cli_case.config_labels("all", S)
    gL, gS = c["gL"], c["gS"]
    exp_perm = []
    for names, hc, psn in cli_case.expected_chunks(inp, 2500, 2):
        ref = orc.assoc_uvlr(hc, "join", "all", gL, gS)
        tables = orc.perm_tables(seed, hc.N, sum(hc.pair_offset[g + 1] > hc.pair_offset[g] for g in range(hc.G)),
                                 nperm)
        t = 0
        for g, gname in enumerate(names):
            p0, p1 = hc.pair_offset[g], hc.pair_offset[g + 1]
            if p1 > p0:
                true = orc.true_stat(ref.w[3 + len(labels) + 1, p0:p1], True)
                st = orc.perm_join_stats(hc, g, "all", True, gL, gS, tables[t])
                pv, used, med, _nb, _dr = orc.perm_count(st, true, nperm)
                exp_perm.append([gname, p1 - p0, pv, used, true, med])
                t += 1
    rows = cli_case.read_table(out + "_joinPermPvals.txt.gz")
    assert rows[0] == ["# perm.bf=all seed=%d" % seed]
    assert len(rows) - 2 == len(exp_perm)
    for got, want in zip(rows[2:], exp_perm):
        # counts exact: near-ties are re-checked in FP64 by the library; the
        # true statistic is the mixed path's (FP64 for the genes with a
        # near-tie), within 1e-4
        assert got[0] == want[0] and int(got[1]) == want[1]
        assert int(got[3]) == want[3]
        assert _num(got[2]) == pytest.approx(want[2], rel=1e-5)
        assert abs(_num(got[4]) - want[4]) <= 1e-4 and abs(_num(got[5]) - want[5]) <= 1e-4


def test_cli_inss(tmp_path, orc):
    # --inss: Bayes factors from summary-statistics files (no genotypes or
    # expression levels); files read like the reference, values against the
    # oracle's restatement of SetSstats + CalcAbfsUvlr
    root = str(tmp_path)
    lst = cli_case.make_inss_case(root)
    c = cli_case.make_case(root)  # for the grid files only
    gL, gS = c["gL"], c["gS"]
    out = os.path.join(root, "ss")
    r = _run(["--inss", lst, "--out", out, "--analys", "join", "--gridL", os.path.join(root, "gridL.txt"),
              "--gridS", os.path.join(root, "gridS.txt"), "--bfs", "all", "--outw", "--outss", "--wrtsize", "2",
              "-v", "0", "--ktimes"], root)
    assert "inss_abf_kernel" in r.stderr
    sbgrps, pairs, n, sig, beta, se, present = cli_case.expected_inss(lst)
    S = len(sbgrps)
    labels = cli_case.config_labels("all", S)
    _std, raw, w = orc.abf_from_sstats(np.where(present, n, 0), sig, beta, se, "all", gL, gS)
    nc = len(labels)
    exp_raw, exp_avg = [], []
    p = 0
    for gname, snps in pairs:
        for snp in snps:
            for k, typ in enumerate(("gen", "gen-fix", "gen-maxh")):
                exp_raw.append([gname, snp, typ] + list(raw[k * len(gL):(k + 1) * len(gL), p]))
            for ci, lab in enumerate(labels):
                r0 = 3 * len(gL) + ci * len(gS)
                exp_raw.append([gname, snp, lab] + list(raw[r0:r0 + len(gS), p]) + [float("nan")] * (len(gL) - len(gS)))
            exp_avg.append([gname, snp, int(present[:, p].sum()), w[0, p], w[1, p], w[2, p], w[3 + nc, p],
                            w[3 + nc + 1, p]] + [w[3 + k, p] for k in range(nc)])
            p += 1
    rows = cli_case.read_table(out + "_l10abfs_raw.txt.gz")
    assert rows[0] == ["gene", "snp", "config"] + ["l10abf.grid%d" % (i + 1) for i in range(len(gL))]
    assert len(rows) - 1 == len(exp_raw)
    for got, want in zip(rows[1:], exp_raw):
        assert got[:3] == want[:3]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    rows = cli_case.read_table(out + "_l10abfs_avg-grids.txt.gz")
    assert rows[0][8:] == ["l10abf." + l for l in labels]
    assert len(rows) - 1 == len(exp_avg)
    for got, want in zip(rows[1:], exp_avg):
        assert got[:2] == want[:2] and int(got[2]) == want[2]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    # --outss: header only (no gene has expression levels in any subgroup)
    for sg in sbgrps:
        rows = cli_case.read_table(out + "_sumstats_%s.txt.gz" % sg)
        assert rows == [["gene", "snp", "maf", "n", "pve", "sigmahat", "betahat.geno", "sebetahat.geno",
                         "betapval.geno"]]


def test_cli_qnorm(tmp_path, orc):
    # --qnorm: every (gene, subgroup) expression vector over its members (an
    # expression level and a genotype column) replaced by normal quantiles of
    # its ranks, (rank - a) / (n + 1 - 2a) with a = 3/8 for n <= 10 else 1/2
    # (utils_math.cpp:80-96); continuous values, so no ties.  The expected
    # side transforms with scipy's ndtri and runs the oracle's association.
    from scipy import special
    import copy
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=17, S=3, N=36, covariates=False)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "qn")
    _run(_common_args(root, out, "join") + ["--bfs", "all", "--outss", "--outw", "--qnorm"], root)
    q = copy.deepcopy(inp)
    for gname, per in q["explev"].items():
        for s, y in per.items():
            mem = np.nonzero(~np.isnan(y) & (q["geno_in"][s] > 0))[0]
            n = mem.size
            if n == 0:
                continue
            a = 0.375 if n <= 10 else 0.5
            order = np.argsort(y[mem], kind="stable")
            z = np.empty(n)
            z[order] = special.ndtri((np.arange(1, n + 1) - a) / (n + 1 - 2 * a))
            y2 = y.copy()
            y2[mem] = z
            per[s] = y2
    S = len(q["sbgrps"])
    gL, gS = c["gL"], c["gS"]
    labels = cli_case.config_labels("all", S)
    exp_ss = {s: [] for s in range(S)}
    exp_avg = []
    for names, hc, psn in cli_case.expected_chunks(q, 2500, 2):
        ref = orc.assoc_uvlr(hc, "join", "all", gL, gS)
        nc = len(labels)
        for g, gname in enumerate(names):
            for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
                for s in range(S):
                    if ref.n[s, p] > 0:
                        exp_ss[s].append([gname, psn[p], ref.n[s, p]] + list(ref.sstats[:, s, p]))
                exp_avg.append([gname, psn[p], int((ref.n[:, p] > 0).sum()), ref.w[0, p], ref.w[1, p], ref.w[2, p],
                                ref.w[3 + nc, p], ref.w[3 + nc + 1, p]] + [ref.w[3 + k, p] for k in range(nc)])
    for s, sg in enumerate(q["sbgrps"]):
        rows = cli_case.read_table(out + "_sumstats_%s.txt.gz" % sg)
        assert len(rows) - 1 == len(exp_ss[s])
        for got, want in zip(rows[1:], exp_ss[s]):
            assert got[:2] == want[:2] and int(got[3]) == want[2]
            assert all(_close(_num(a), b) for a, b in zip(got[4:], want[3:])), (got, want)
    rows = cli_case.read_table(out + "_l10abfs_avg-grids.txt.gz")
    assert len(rows) - 1 == len(exp_avg)
    for got, want in zip(rows[1:], exp_avg):
        assert got[:2] == want[:2] and int(got[2]) == want[2]
        assert all(_close(_num(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
