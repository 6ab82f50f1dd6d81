"""End to end: the eqtlbma_avg_bfs host program on raw ABF files, against
the oracle.

Inputs (tests/avg_case.py): two gzipped raw files in eqtlbma_bf's layout
(gen, gen-fix, gen-maxh and the 7 configurations of 3 subgroups per SNP,
"nan" tokens here and there).  The expected side parses them independently,
runs the oracle's averaging and posteriors, orders each gene's SNPs by
decreasing BF (NaN last) and applies --bestsnp.  Values print with %.6e
(grid only) or %.4e: they must match the oracle's to the printed precision
(relative 1e-6 / 1e-4, i.e. within rounding of the last digit).
"""
import gzip
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from eqtlbma_amd import CLI_AVG_PATH
import avg_case

pytestmark = pytest.mark.gpu

L = 6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args, cwd):
    _say("[gpu] eqtlbma_avg_bfs " + " ".join(args))
    r = subprocess.run([CLI_AVG_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _read(path):
    with gzip.open(path, "rt") as f:
        return [l.rstrip("\n").split("\t") for l in f]


def _num_ok(tok, want, rel):
    if math.isnan(want):
        return tok in ("nan", "-nan")
    got = float(tok)
    return abs(got - want) <= rel * abs(want) + 1e-300


def _w(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("avg"))
    paths = avg_case.write_raw(root, seed=5, S=3, L=L)
    rng = np.random.default_rng(6)
    gw = rng.dirichlet(np.ones(L))
    _w(os.path.join(root, "gw.txt"), ["%.10g" % x for x in gw])
    return root, paths, gw


def test_grid_only(case, orc):
    root, paths, gw = case
    out = os.path.join(root, "grid.txt.gz")
    _run(["--in", os.path.join(root, "out_*_l10abfs_raw.txt.gz"), "--gwts", os.path.join(root, "gw.txt"),
          "--gtk", "0+2+3+5", "--nsubgrp", "3", "--dim", "7", "--out", out, "-v", "0"], root)
    keep = np.array([0, 2, 3, 5], np.int32)
    exp = []
    for p in sorted(paths):
        with gzip.open(p, "rt") as fh:
            lines = [l for l in fh.read().split("\n") if l][1:]
        rows = [l.split("\t") for l in lines]
        raw = np.array([[float(x) for x in r[3:3 + L]] for r in rows])
        avg = orc.avg_grid(raw, gw, keep)
        exp += [(r[0], r[1], r[2], a) for r, a in zip(rows, avg)]
    got = _read(out)
    assert got[0] == ["gene", "snp", "config", "l10abf.grid.avg"]
    assert len(got) - 1 == len(exp)
    for g, e in zip(got[1:], exp):
        assert g[:3] == list(e[:3])
        assert _num_ok(g[3], e[3], 1e-6), (g, e)


def _expected(orc, paths, gw, keep_cfgs, model, kw, post, best_snps, filters=None):
    """Rows (gene, snp, dict of values) in the output order."""
    filters = filters or {}
    out = []
    for gnames, snames, raw, off in avg_case.read_raw(paths, L, keep_cfgs, **filters):
        if not gnames:
            continue
        r = orc.avg_bfs(raw, off, gw, model=model, want=post, **kw)
        for g, gname in enumerate(gnames):
            idx = list(range(off[g], off[g + 1]))
            idx.sort(key=lambda p: (math.isnan(r["snp_bf"][p]), -r["snp_bf"][p] if not math.isnan(r["snp_bf"][p])
                                    else 0))
            last = len(idx) - 1
            if best_snps == 1:
                last = 0
            elif best_snps == 2:
                cum = 0.0
                for q, p in enumerate(idx):
                    cum += r["post_the"][p]
                    if cum >= 0.95:
                        last = q
                        break
            for p in idx[:last + 1]:
                out.append((gname, snames[p], g, p, r))
    return out


def test_configs_all_outputs(case, orc):
    root, paths, gw = case
    names = avg_case.cfg_names(3)
    cw = np.array([0.3, 0.2, 0.1, 0.1, 0.1, 0.05, 0.15])
    _w(os.path.join(root, "cw.txt"), ["%s\t%.10g" % (n, w) for n, w in zip(names, cw)])
    _w(os.path.join(root, "genes.txt"), ["gene000", "gene002", "gene005", "gene007"])
    out = os.path.join(root, "cfg.txt.gz")
    _run(["--in", os.path.join(root, "out_*_l10abfs_raw.txt.gz"), "--gwts", os.path.join(root, "gw.txt"),
          "--model", "configs", "--nsubgrp", "3", "--dim", "7", "--cwts", os.path.join(root, "cw.txt"),
          "--save", "bf+post", "--pi0", "0.4", "--post", "a+b+c+d", "--bestdim", "--alldim",
          "--gene", os.path.join(root, "genes.txt"), "--out", out, "-v", "0"], root)
    masks = [sum(1 << (int(t) - 1) for t in n.split("-")) for n in names]
    exp = _expected(orc, paths, gw, names, "configs",
                    dict(cfg_w=cw, cfg_sub=np.array(masks, np.uint64), pi0=0.4, n_subgroups=3),
                    ("a", "b", "c", "d", "dims"), 0,
                    dict(genes={"gene000", "gene002", "gene005", "gene007"}))
    got = _read(out)
    assert got[0] == ["#pi0=4.0000e-01"]
    hdr = (["gene", "snp", "gene.log10.bf", "snp.log10.bf"] + ["log10.bf." + n for n in names] +
           ["gene.post", "snp.post.the", "snp.post.an", "snp.post.s1", "snp.post.s2", "snp.post.s3"] +
           ["post.config." + n for n in names] + ["best.config", "post.best.config"])
    assert got[1] == hdr
    assert len(got) - 2 == len(exp)
    for row, (gname, sname, g, p, r) in zip(got[2:], exp):
        assert row[:2] == [gname, sname]
        want = ([r["gene_bf"][g], r["snp_bf"][p]] + list(r["dim_bf"][p]) +
                [r["gene_post"][g], r["post_the"][p], r["post_an"][p]] + list(r["post_sub"][p]) +
                list(r["post_dim"][p]))
        for tok, w in zip(row[2:2 + len(want)], want):
            assert _num_ok(tok, w, 1e-4), (row, want)
        b = int(r["best_dim"][p])
        assert row[-2] == names[b]
        assert _num_ok(row[-1], r["post_dim"][p][b], 1e-4)
    _say("[gpu]   %d SNP lines match" % len(exp))


def test_configs_subset_bestsnp2(case, orc):
    root, paths, gw = case
    sub = ["1", "2", "3", "1-2-3"]
    cw = np.array([0.2, 0.2, 0.2, 0.4])
    _w(os.path.join(root, "cw4.txt"), ["%s %.10g" % (n, w) for n, w in zip(sub, cw)])
    out = os.path.join(root, "cfg4.txt.gz")
    _run(["--in", os.path.join(root, "out_*_l10abfs_raw.txt.gz"), "--gwts", os.path.join(root, "gw.txt"),
          "--gtk", "1+2+3+4", "--model", "configs", "--nsubgrp", "3", "--dim", "4", "--cwts",
          os.path.join(root, "cw4.txt"), "--save", "bf+post", "--pi0", "0.7", "--post", "b",
          "--bestsnp", "2", "--out", out, "-v", "0"], root)
    exp = _expected(orc, paths, gw, sub, "configs", dict(cfg_w=cw, pi0=0.7, keep=np.array([1, 2, 3, 4], np.int32)),
                    ("b",), 2)
    got = _read(out)
    assert got[1] == ["gene", "snp", "gene.log10.bf", "snp.log10.bf", "snp.post.the"]
    assert len(got) - 2 == len(exp)
    for row, (gname, sname, g, p, r) in zip(got[2:], exp):
        assert row[:2] == [gname, sname]
        for tok, w in zip(row[2:], [r["gene_bf"][g], r["snp_bf"][p], r["post_the"][p]]):
            assert _num_ok(tok, w, 1e-4), row


def test_types_all_outputs(case, orc):
    root, paths, gw = case
    tw = np.array([0.7, 0.3])
    sw = np.array([[0.9, 0.8, 0.7], [0.2, 0.5, 0.1]])
    # the names eqtlbma_hm writes (#type.k, #subgroup.k-s), types first
    lines = ["type.%d\t%.10g" % (k + 1, tw[k]) for k in range(2)]
    lines += ["subgroup.%d-%d\t%.10g" % (k + 1, s + 1, sw[k, s]) for k in range(2) for s in range(3)]
    _w(os.path.join(root, "tsw.txt"), lines)
    out = os.path.join(root, "types.txt.gz")
    _run(["--in", os.path.join(root, "out_*_l10abfs_raw.txt.gz"), "--gwts", os.path.join(root, "gw.txt"),
          "--model", "types", "--nsubgrp", "3", "--dim", "2", "--tswts", os.path.join(root, "tsw.txt"),
          "--save", "bf+post", "--pi0", "0.5", "--post", "a+b+c+d", "--bestdim", "--alldim", "--out", out,
          "-v", "0"], root)
    exp = _expected(orc, paths, gw, ["1", "2", "3"], "types", dict(type_w=tw, sub_w=sw, pi0=0.5),
                    ("a", "b", "c", "d", "dims"), 0)
    got = _read(out)
    assert got[1][-2:] == ["best.type", "post.best.type"]
    assert len(got) - 2 == len(exp)
    for row, (gname, sname, g, p, r) in zip(got[2:], exp):
        assert row[:2] == [gname, sname]
        want = ([r["gene_bf"][g], r["snp_bf"][p]] + list(r["dim_bf"][p]) +
                [r["gene_post"][g], r["post_the"][p], r["post_an"][p]] + list(r["post_sub"][p]) +
                list(r["post_dim"][p]))
        for tok, w in zip(row[2:2 + len(want)], want):
            assert _num_ok(tok, w, 1e-4), (row, want)
        b = int(r["best_dim"][p])
        assert row[-2] == str(b + 1)
