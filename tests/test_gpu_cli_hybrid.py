"""End to end: eqtlbma_bf --error hybrid on files, against the oracle.

One genotype file for every subgroup (as the reference requires for
--error mvlr / hybrid), expression files with their own sample subsets,
column orders and NA levels, so the pairs of subgroups share some
individuals and not others; covariate files list the samples in another
order than the union of samples, so the pairwise designs read covariates by
sample index (covars_by_position); a gene absent from one subgroup is
skipped.  --qnorm normalises the per-subgroup levels only: the pairwise step
reads the levels as loaded.  Values printed with seven significant digits
must match the oracle's to 2e-6 relative or 1e-9 absolute.
"""
import copy
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from eqtlbma_amd import CLI_PATH
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


def _close(a, b):
    if math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= 1e-9 + 2e-6 * abs(b)


def _args(root, out):
    return ["--geno", os.path.join(root, "list_geno.txt"), "--scoord", os.path.join(root, "snps.bed"),
            "--exp", os.path.join(root, "list_exp.txt"), "--gcoord", os.path.join(root, "genes.bed"),
            "--cis", "2500", "--out", out, "--analys", "join", "--gridL", os.path.join(root, "gridL.txt"),
            "--gridS", os.path.join(root, "gridS.txt"), "--wrtsize", "2", "-v", "0",
            "--covar", os.path.join(root, "list_cov.txt"), "--error", "hybrid", "--fiterr", "0.5",
            "--bfs", "all", "--outss", "--outw"]


def _qnorm(inp):
    from scipy import special
    q = copy.deepcopy(inp)
    for per in q["explev"].values():
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
    return q


def _check(orc, c, inp, out, qnorm_inp=None):
    S = len(inp["sbgrps"])
    gL, gS = c["gL"], c["gS"]
    labels = cli_case.config_labels("all", S)
    nc = len(labels)
    exp_ss = {s: [] for s in range(S)}
    exp_raw, exp_avg = [], []
    use = qnorm_inp if qnorm_inp is not None else inp
    raw_chunks = cli_case.expected_chunks(inp, 2500, 2, mvlr=True)
    for (names, hc, psn), (_n2, hraw, _p2) in zip(cli_case.expected_chunks(use, 2500, 2, mvlr=True), raw_chunks):
        if hc.n_pairs == 0:
            continue
        ref = orc.assoc_hybrid(hc, "all", 0.5, gL, gS, covars_pos=inp["covars_pos"],
                               expl_pair=hraw.explevels if qnorm_inp is not None else None)
        for g, gname in enumerate(names):
            for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
                for s in range(S):
                    if ref.n[s, p] > 0:
                        exp_ss[s].append([gname, psn[p], ref.n[s, p]] + list(ref.sstats[:, s, p]))
                for k, typ in enumerate(("gen", "gen-fix", "gen-maxh")):
                    exp_raw.append([gname, psn[p], typ] + list(ref.raw[k * len(gL):(k + 1) * len(gL), p]))
                for ci, lab in enumerate(labels):
                    r0 = 3 * len(gL) + ci * len(gS)
                    exp_raw.append([gname, psn[p], lab] + list(ref.raw[r0:r0 + len(gS), p]) +
                                   [float("nan")] * (len(gL) - len(gS)))
                exp_avg.append([gname, psn[p], int((ref.n[:, p] > 0).sum()), ref.w[0, p], ref.w[1, p], ref.w[2, p],
                                ref.w[3 + nc, p], ref.w[3 + nc + 1, p]] + [ref.w[3 + k, p] for k in range(nc)])
    assert exp_raw, "no pair analysed"
    rows = cli_case.read_table(out + "_l10abfs_raw.txt.gz")
    assert len(rows) - 1 == len(exp_raw)
    for got, want in zip(rows[1:], exp_raw):
        assert got[:3] == want[:3]
        assert all(_close(float(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    rows = cli_case.read_table(out + "_l10abfs_avg-grids.txt.gz")
    assert len(rows) - 1 == len(exp_avg)
    for got, want in zip(rows[1:], exp_avg):
        assert got[:2] == want[:2] and int(got[2]) == want[2]
        assert all(_close(float(a), b) for a, b in zip(got[3:], want[3:])), (got, want)
    for s, sg in enumerate(inp["sbgrps"]):
        rows = cli_case.read_table(out + "_sumstats_%s.txt.gz" % sg)
        assert len(rows) - 1 == len(exp_ss[s])
        for got, want in zip(rows[1:], exp_ss[s]):
            assert got[:2] == want[:2] and int(got[3]) == want[2]
            assert all(_close(float(a), b) for a, b in zip(got[4:], want[3:])), (got, want)
    _say("[gpu]   %d raw rows, %d pairs match the oracle" % (len(exp_raw), len(exp_avg)))


def test_cli_hybrid(tmp_path, orc):
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=23, S=3, N=40, one_geno=True)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "hy")
    _run(_args(root, out), root)
    _check(orc, c, inp, out)


def test_cli_hybrid_qnorm(tmp_path, orc):
    root = str(tmp_path)
    c = cli_case.make_case(root, seed=29, S=3, N=40, one_geno=True)
    inp = cli_case.load_expected_inputs(c)
    out = os.path.join(root, "hq")
    _run(_args(root, out) + ["--qnorm"], root)
    _check(orc, c, inp, out, qnorm_inp=_qnorm(inp))
