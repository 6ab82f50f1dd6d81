"""GPU: the eqtlbma_hm host program end to end, against the oracle.

The expected side parses the raw-ABF file independently (the reference's
loader rules: header "gene snp config", consecutive rows of a gene and of a
SNP, "gen*" rows skipped unless --keepgen, the first --ngrid values of a
row), runs the oracle's EM from the reference's starting point
(init_params: pi0 0.5, uniform weights, subgroup probabilities 0.5; an
--init file; or GSL mt19937 draws for --seed), and the oracle's --getbf and
--getci outputs.  The program prints every number as %.4e, so values are
compared to 1e-4 relative; names, order, NaNs and the fixed flags exactly.
"""
import gzip
import itertools
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from eqtlbma_amd import CLI_HM_PATH, CLI_PATH
import cli_case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hm(args):
    _say("[gpu] eqtlbma_hm " + " ".join(args))
    r = subprocess.run([CLI_HM_PATH] + args + ["-v", "0"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _labels(S):
    out = []
    for k in range(1, S + 1):
        for c in itertools.combinations(range(1, S + 1), k):
            out.append("-".join(str(i) for i in c))
    return out


def _write_raw(path, raw, off, labels, L_file, rng):
    """eqtlbma_bf's raw-ABF layout: per pair three gen rows then one row per
    configuration, L_file values each (NaN after the ones given)."""
    with gzip.open(path, "wt") as f:
        f.write("gene\tsnp\tconfig" + "".join("\tl10abf.grid%d" % (i + 1) for i in range(L_file)) + "\n")
        for g in range(off.size - 1):
            for p in range(off[g], off[g + 1]):
                head = "gene%03d\trs%d_%d\t" % (g, g, p - off[g])
                for typ in ("gen", "gen-fix", "gen-maxh"):
                    f.write(head + typ + "".join("\t%.6e" % v for v in rng.normal(0, 1, L_file)) + "\n")
                for k, lab in enumerate(labels):
                    vals = list(raw[p, k]) + [float("nan")] * (L_file - raw.shape[2])
                    f.write(head + lab + "".join("\t%.6e" % v for v in vals) + "\n")


def _parse_raw(path, ngrid, keep):
    """Independent loader: rows whose config passes keep(), grouped by
    consecutive gene and SNP names; returns raw [pairs][D][ngrid], offsets,
    gene names, SNP names and the first SNP's config names."""
    genes, snps, off, rows, names = [], [], [0], [], []
    cur_g = cur_s = None
    per_snp = []

    def flush_snp():
        if per_snp:
            rows.append(list(per_snp))
            snps.append(cur_s)
            per_snp.clear()

    with gzip.open(path, "rt") as f:
        next(f)
        for line in f:
            t = line.split()
            if not keep(t[2]):
                continue
            if t[0] != cur_g:
                flush_snp()
                if cur_g is not None:
                    genes.append(cur_g)
                    off.append(len(snps))
                cur_g = t[0]
            if t[1] != cur_s:
                flush_snp()
                cur_s = t[1]
            vals = [float(x) for x in t[3:3 + ngrid]] + [float("nan")] * max(0, ngrid - len(t[3:]))
            per_snp.append(vals)
            if len(snps) == 0 and len(rows) == 0:
                names.append(t[2])
    flush_snp()
    genes.append(cur_g)
    off.append(len(snps))
    return np.array(rows, dtype=np.float64), np.array(off, dtype=np.int64), genes, snps, names


def _read_out(path):
    params, header, rows = [], None, []
    with gzip.open(path, "rt") as f:
        for line in f:
            t = line.rstrip("\n").split("\t")
            if t[0].startswith("#"):
                params.append(t)
            elif header is None:
                header = t
            else:
                rows.append(t)
    return params, header, rows


def _close(txt, want):
    v = float(txt)
    if math.isnan(want):
        return math.isnan(v)
    return abs(v - want) <= 1.01e-4 * abs(want) + 1e-300


def _check(path, orc, raw, off, genes, snps, names, em, model, fixed=(0, 0, 0, 0, 0), ci=None, getbf=True):
    params, header, rows = _read_out(path)
    assert params[0] == ["#param", "mle", "left.ci", "right.ci", "fixed"]
    K = len(em.cfg) if model == "configs" else len(em.type)
    L = len(em.grid)
    exp = [("#pi0", em.pi0, fixed[0])]
    if model == "configs":
        exp += [("#config." + names[k], em.cfg[k], fixed[2]) for k in range(K)]
    else:
        S = em.sub.shape[1]
        exp += [("#type.%d" % (k + 1), em.type[k], fixed[3]) for k in range(K)]
        exp += [("#subgroup.%d-%d" % (k + 1, s + 1), em.sub[k, s], fixed[4]) for k in range(K) for s in range(S)]
    exp += [("#grid.%d" % (l + 1), em.grid[l], fixed[1]) for l in range(L)]
    assert len(params) - 1 == len(exp)
    for i, (got, (name, v, fx)) in enumerate(zip(params[1:], exp)):
        assert got[0] == name and got[4] == ("true" if fx else "false"), got
        assert _close(got[1], v), (got, v)
        lo, hi = (float("nan"), float("nan")) if ci is None else (ci[0][i], ci[1][i])
        assert _close(got[2], lo) and _close(got[3], hi), (got, lo, hi)
    if not getbf:
        assert header is None
        return
    gp, gb, sb, db = orc.hm_bfs(raw, off, em)
    want_h = ["gene", "gene.posterior.prob", "gene.log10.bf", "snp", "snp.log10.bf"]
    want_h += ["log10.bf." + (names[k] if model == "configs" else str(k + 1)) for k in range(K)]
    assert header == want_h
    assert len(rows) == len(snps)
    for g in range(off.size - 1):
        for p in range(off[g], off[g + 1]):
            r = rows[p]
            assert r[0] == genes[g] and r[3] == snps[p]
            assert _close(r[1], gp[g]) and _close(r[2], gb[g]) and _close(r[4], sb[p]), (r, gp[g], gb[g], sb[p])
            assert all(_close(r[5 + k], db[p, k]) for k in range(K)), (r, db[p])


def _default_prm(orc, model, K, L, S=None):
    if model == "configs":
        return orc.HmParams(0.5, np.full(L, 1.0 / L), cfg=np.full(K, 1.0 / K))
    return orc.HmParams(0.5, np.full(L, 1.0 / L), type_=np.full(K, 1.0 / K), sub=np.full((K, S), 0.5))


@pytest.fixture(scope="module")
def synth_raw(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("hmcli"))
    rng = np.random.default_rng(5)
    G, S, L = 250, 3, 5
    labels = _labels(S)
    P = rng.integers(2, 25, G)
    off = np.concatenate([[0], np.cumsum(P)]).astype(np.int64)
    raw = rng.normal(-0.6, 0.4, (off[-1], len(labels), L))
    for g in range(G):
        if rng.random() < 0.4:
            raw[off[g] + rng.integers(P[g])] += rng.uniform(1, 3, (len(labels), L)) * (rng.random((len(labels), L)) < .6)
    path = os.path.join(root, "obs_bf_l10abfs_raw.txt.gz")
    _write_raw(path, raw, off, labels, L + 2, rng)
    return root, path, S, L, labels


def test_hm_cli_after_eqtlbma_bf(tmp_path, orc):
    """The reference's own sequence (tests/test_hm.bash): eqtlbma_bf --bfs all
    writes the raw ABFs, eqtlbma_hm fits the configs model on them with
    --getbf --getci."""
    root = str(tmp_path)
    c = cli_case.make_case(root)
    out = os.path.join(root, "obs_bf")
    args = ["--geno", os.path.join(root, "list_geno.txt"), "--scoord", os.path.join(root, "snps.bed"),
            "--exp", os.path.join(root, "list_exp.txt"), "--gcoord", os.path.join(root, "genes.bed"),
            "--cis", "2500", "--out", out, "--analys", "join", "--gridL", os.path.join(root, "gridL.txt"),
            "--gridS", os.path.join(root, "gridS.txt"), "--bfs", "all", "--outss", "--outw", "-v", "0"]
    _say("[gpu] eqtlbma_bf " + " ".join(args))
    r = subprocess.run([CLI_PATH] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    S, ng = 3, len(c["gS"])
    rawf = out + "_l10abfs_raw.txt.gz"
    hm_out = os.path.join(root, "obs_hm.txt.gz")
    _hm(["--data", rawf, "--nsubgrp", str(S), "--dim", "7", "--ngrid", str(ng), "--out", hm_out, "--getbf",
         "--getci"])
    raw, off, genes, snps, names = _parse_raw(rawf, ng, lambda c: "gen" not in c)
    assert names == _labels(S)
    em, _, _ = orc.hm_em(raw, off, _default_prm(orc, "configs", 7, ng))
    _check(hm_out, orc, raw, off, genes, snps, names, em, "configs", ci=orc.hm_ci(raw, off, em))


@pytest.mark.parametrize("msl", ["1", "3"])
def test_hm_cli_configs(synth_raw, orc, msl):
    root, path, S, L, labels = synth_raw
    out = os.path.join(root, "hm_cfg_%s.txt.gz" % msl)
    _hm(["--data", path, "--nsubgrp", str(S), "--dim", "7", "--ngrid", str(L), "--out", out, "--getbf",
         "--thresh", "1e-4", "--msl", msl])
    raw, off, genes, snps, names = _parse_raw(path, L, lambda c: "gen" not in c)
    em, _, _ = orc.hm_em(raw, off, _default_prm(orc, "configs", 7, L), thresh=1e-4, stepmax=float(msl))
    _check(out, orc, raw, off, genes, snps, names, em, "configs")


def test_hm_cli_types(synth_raw, orc):
    root, path, S, L, labels = synth_raw
    out = os.path.join(root, "hm_types.txt.gz")
    _hm(["--data", path, "--nsubgrp", str(S), "--model", "types", "--dim", "2", "--ngrid", str(L), "--out", out,
         "--getbf", "--getci", "--configs", "1-2|3"])  # --configs is ignored by the types model
    keep = {"1", "2", "3"}
    raw, off, genes, snps, names = _parse_raw(path, L, lambda c: c in keep)
    em, _, _ = orc.hm_em(raw, off, _default_prm(orc, "types", 2, L, S))
    _check(out, orc, raw, off, genes, snps, names, em, "types", ci=orc.hm_ci(raw, off, em))


def test_hm_cli_init_file_fixed_and_subset(synth_raw, orc):
    """--init with fixed grid weights, --pi0, --configs subset with --keepgen
    (keeps "gen", still drops gen-fix and gen-maxh), --maxit."""
    root, path, S, L, labels = synth_raw
    init = os.path.join(root, "init.txt")
    grid0 = np.array([0.1, 0.2, 0.3, 0.2, 0.2])
    with open(init, "w") as f:
        f.write("param\tvalue\tfixed\n# a comment\n")
        for l in range(L):
            f.write("grid.%d\t%g\tTRUE\n" % (l + 1, grid0[l]))
        for k, w in enumerate([0.4, 0.3, 0.2, 0.1]):
            f.write("config.%d\t%g\tFALSE\n" % (k + 1, w))
    out = os.path.join(root, "hm_init.txt.gz")
    _hm(["--data", path, "--nsubgrp", str(S), "--dim", "4", "--ngrid", str(L), "--out", out, "--init", init,
         "--pi0", "0.7", "--configs", "gen|1|2|1-2-3", "--keepgen", "--maxit", "5"])
    keep = {"gen", "1", "2", "1-2-3"}
    raw, off, genes, snps, names = _parse_raw(path, L, lambda c: c in keep)
    assert names == ["gen", "1", "2", "1-2-3"]
    prm = orc.HmParams(0.7, grid0, cfg=np.array([0.4, 0.3, 0.2, 0.1]))
    em, it, _ = orc.hm_em(raw, off, prm, maxit=5, fixed=(1, 1, 0, 0, 0))
    _check(out, orc, raw, off, genes, snps, names, em, "configs", fixed=(1, 1, 0, 0, 0), getbf=False)


def test_hm_cli_random_start(synth_raw, orc):
    """--seed: pi0 = gsl_rng_uniform, grid and configuration weights from
    gsl_ran_exponential(1) = -log1p(-u), normalised (randomize_parameter_sp)."""
    root, path, S, L, labels = synth_raw
    out = os.path.join(root, "hm_seed.txt.gz")
    _hm(["--data", path, "--nsubgrp", str(S), "--dim", "7", "--ngrid", str(L), "--out", out, "--seed", "1234",
         "--maxit", "4"])
    rng = orc.MT(1234)
    pi0 = rng.flat(0.0, 1.0)
    grid = np.array([-math.log1p(-rng.flat(0.0, 1.0)) for _ in range(L)])
    cfg = np.array([-math.log1p(-rng.flat(0.0, 1.0)) for _ in range(7)])
    prm = orc.HmParams(pi0, grid / grid.sum(), cfg=cfg / cfg.sum())
    raw, off, genes, snps, names = _parse_raw(path, L, lambda c: "gen" not in c)
    em, _, _ = orc.hm_em(raw, off, prm, maxit=4)
    _check(out, orc, raw, off, genes, snps, names, em, "configs", getbf=False)
