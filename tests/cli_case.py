"""A small eqtlbma_bf input set written to disk, and the results the
reference's semantics give on it, computed with the oracle.

The expected side restates the reference's data handling independently of
the host program: sorted union of sample names, subgroups in name order,
genes in name order processed in chunks of --wrtsize, BED start + 1 as the
TSS, cis SNPs by position, the custom dose format's missing-value rules, the
MAF of each genotype file, covariates sorted by name, and the permutation
generator reseeded at every chunk (eqtlbma_bf.cpp:1546-1574, 826-917).
"""
import gzip
import os

import numpy as np

from eqtlbma_amd._abi import HostChunk, config_masks


def _w(path, lines):
    with open(path, "w") as f:
        for l in lines:
            f.write(l + "\n")


def make_case(root, seed=3, S=3, N=36, mvlr=False, covariates=True, near_mono=None, one_geno=False):
    """Writes the input files under root; returns a dict describing them.
    one_geno: one genotype file of every sample for all subgroups (as
    --error mvlr / hybrid require) while the expression files keep their own
    sample subsets and NA levels.
    near_mono: index of a SNP made monomorphic but for one carrier (its
    statistics are NaN whenever that carrier is not among the members)."""
    rng = np.random.default_rng(seed)
    samples = ["ind%02d" % (i + 1) for i in range(N)]
    sbgrps = ["s%d" % (s + 1) for s in range(S)]
    # SNPs on two chromosomes, every 400 bp
    snps = []
    for chrom in ("chr1", "chr2"):
        for k in range(40):
            snps.append(("rs%s_%d" % (chrom[-1], k), chrom, 1000 + 400 * k))
    maf = rng.uniform(0.1, 0.5, size=len(snps))
    G = rng.binomial(2, maf[:, None], size=(len(snps), N)).astype(float)
    if near_mono is not None:
        G[near_mono] = 1.0
        G[near_mono, 0] = 2.0
    genes = [("geneB", "chr1", 3000, 4000), ("geneA", "chr1", 9000, 9500), ("geneD", "chr2", 5000, 5200),
             ("geneC", "chr2", 12000, 13000), ("geneE", "chr2", 90000, 91000),  # no cis SNP
             ("geneF", "chr1", 6000, 7000)]
    # membership: uvlr subgroups see different sample subsets; mvlr all samples
    exp_in, geno_in = [], []
    for s in range(S):
        if mvlr:
            e = np.ones(N, bool)
            g = np.ones(N, bool)
        elif one_geno:
            e = rng.random(N) < 0.85
            g = np.ones(N, bool)
        else:
            e = rng.random(N) < 0.85
            g = rng.random(N) < 0.85
        if near_mono is not None:
            e[0] = g[0] = True  # the carrier is in every file
        exp_in.append(e)
        geno_in.append(g)
    Y = rng.normal(size=(len(genes), S, N))
    Y[0] += 0.8 * G[5][None, :]
    Y[3, 1] += 0.7 * G[40 + 27]
    cov = rng.normal(size=(2, N))
    Y += 0.3 * cov[0]
    paths = {}
    exp_lines, geno_lines, cov_lines = [], [], []
    for s, name in enumerate(sbgrps):
        order = [i for i in rng.permutation(N) if exp_in[s][i]]
        hdr = ("Id " if s == 0 else "") + " ".join(samples[i] for i in order)
        rows = [hdr]
        for gi, (gname, *_rest) in enumerate(genes):
            if gname == "geneC" and s == 2 and not mvlr:
                continue  # no expression of geneC in s3
            vals = []
            for i in order:
                v = Y[gi, s, i]
                vals.append("NA" if (not mvlr and rng.random() < 0.03) else "%.10g" % v)
            rows.append(gname + " " + " ".join(vals))
        rows.append("geneZ " + " ".join("0.5" for _ in order))  # no coordinates: skipped
        p = os.path.join(root, "exp_%s.txt" % name)
        _w(p, rows)
        exp_lines.append("%s %s" % (name, p))
        if mvlr or one_geno:
            gorder = list(range(N))
            gpath = os.path.join(root, "geno.txt")
        else:
            gorder = [i for i in rng.permutation(N) if geno_in[s][i]]
            gpath = os.path.join(root, "geno_%s.txt" % name)
        if not (mvlr or one_geno) or s == 0:
            rows = ["Id\t" + "\t".join(samples[i] for i in gorder)]
            for j, (sname, _c, _p) in enumerate(snps):
                vals = ["%g" % G[j, i] for i in gorder]
                if not mvlr and s == 1 and j == 7:
                    vals[0] = "NA"      # the line is skipped for this subgroup
                if not mvlr and s == 2 and j == 9:
                    vals[1] = "nan"     # the SNP is dropped for this subgroup
                rows.append(sname + "\t" + "\t".join(vals))
            _w(gpath, rows)
        geno_lines.append("%s\t%s" % (name, gpath))
        if covariates and s != 1:
            corder = list(rng.permutation(N))
            rows = ["Id " + " ".join(samples[i] for i in corder)]
            rows.append("pc2 " + " ".join("%.8g" % (cov[1, i] * (s + 1)) for i in corder))
            rows.append("pc1 " + " ".join("%.8g" % cov[0, i] for i in corder))
            p = os.path.join(root, "cov_%s.txt" % name)
            _w(p, rows)
            cov_lines.append("%s %s" % (name, p))
    _w(os.path.join(root, "list_exp.txt"), exp_lines)
    _w(os.path.join(root, "list_geno.txt"), geno_lines)
    if cov_lines:
        _w(os.path.join(root, "list_cov.txt"), cov_lines)
    _w(os.path.join(root, "genes.bed"), ["%s\t%d\t%d\t%s\t0\t+" % (c, a, b, n) for n, c, a, b in genes] +
       ["chr1\t50000\t50100\tgeneY\t0\t+"])  # no expression: dropped
    _w(os.path.join(root, "snps.bed"), ["%s\t%d\t%d\t%s" % (c, p - 1, p, n) for n, c, p in snps])
    gL = np.array([[0.04, 0.0], [0.01, 0.03], [0.16, 0.04], [0.0, 0.25]])
    gS = np.array([[0.02, 0.01], [0.2, 0.05], [0.5, 0.1]])
    _w(os.path.join(root, "gridL.txt"), ["%g\t%g" % tuple(r) for r in gL])
    _w(os.path.join(root, "gridS.txt"), ["%g\t%g" % tuple(r) for r in gS])
    return dict(root=root, samples=samples, sbgrps=sbgrps, snps=snps, G=G, genes=genes, Y=Y, cov=cov,
                gL=gL, gS=gS, mvlr=mvlr, covariates=covariates, exp_lines=exp_lines)


def read_table(path):
    with gzip.open(path, "rt") as f:
        return [l.rstrip("\n").split("\t") for l in f]


def load_expected_inputs(case):
    """Parses the written files the way the reference does and returns the
    per-gene data needed to build oracle chunks."""
    root = case["root"]
    sbgrps = sorted(case["sbgrps"])
    S = len(sbgrps)

    def rd(p):
        with open(p) as f:
            return [l.rstrip("\n") for l in f]

    def cols(line):
        t = line.split()
        return t[1:] if t[0] in ("Id", "id", "ID") else t

    exp, geno, covf = {}, {}, {}
    for l in rd(os.path.join(root, "list_exp.txt")):
        a, b = l.split()
        exp[a] = b
    for l in rd(os.path.join(root, "list_geno.txt")):
        a, b = l.split()
        geno[a] = b
    if os.path.exists(os.path.join(root, "list_cov.txt")):
        for l in rd(os.path.join(root, "list_cov.txt")):
            a, b = l.split()
            covf[a] = b
    names = set()
    for s in sbgrps:
        names |= set(cols(rd(exp[s])[0]))
        names |= set(cols(rd(geno[s])[0]))
    samples = sorted(names)
    N = len(samples)
    idx = {n: i for i, n in enumerate(samples)}
    gcoord = {}
    for l in rd(os.path.join(root, "genes.bed")):
        t = l.split()
        gcoord.setdefault(t[3], (t[0], int(t[1]) + 1, int(t[2])))
    explev = {}
    for s, sg in enumerate(sbgrps):
        lines = rd(exp[sg])
        c = cols(lines[0])
        for l in lines[1:]:
            t = l.split()
            if t[0] not in gcoord:
                continue
            y = np.full(N, np.nan)
            for k, v in enumerate(t[1:]):
                y[idx[c[k]]] = np.nan if v in ("NA", "na", "NaN", "nan") else float(v)
            explev.setdefault(t[0], {}).setdefault(s, y)
    scoord = {}
    for l in rd(os.path.join(root, "snps.bed")):
        t = l.split()
        scoord.setdefault(t[3], (t[0], int(t[2])))
    # genotype files (custom dose): per subgroup the values over the union
    files = {}
    geno_vals = []     # per subgroup: dict snp -> (values over union, maf)
    geno_in = np.zeros((S, N), np.uint8)
    for s, sg in enumerate(sbgrps):
        lines = rd(geno[sg])
        c = cols(lines[0])
        for n in c:
            geno_in[s, idx[n]] = 1
        if geno[sg] not in files:
            d = {}
            for l in lines[1:]:
                t = l.split()
                if t[0] not in scoord or "NA" in t:
                    continue
                v = t[1:]
                if any(x in ("na", "NaN", "nan") for x in v):
                    continue
                x = np.array([float(a) for a in v])
                f = x.sum() / (2 * len(x))
                f = f if f <= 0.5 else 1 - f
                full = np.zeros(N)
                for k, a in enumerate(x):
                    full[idx[c[k]]] = a
                d[t[0]] = (full, f)
            files[geno[sg]] = d
        geno_vals.append(files[geno[sg]])
    covars = np.zeros((S, 2, N))
    covars_pos = np.full((S, 2, N), np.nan)  # covariate value at file column i (hybrid's pairwise designs)
    covar_count = np.zeros(S, np.int32)
    covar_in = np.zeros((S, N), np.uint8)
    for s, sg in enumerate(sbgrps):
        if sg not in covf:
            continue
        lines = rd(covf[sg])
        c = cols(lines[0])
        rows = {}
        for l in lines[1:]:
            t = l.split()
            rows.setdefault(t[0], [float(a) for a in t[1:]])
        for q, name in enumerate(sorted(rows)):
            for k, v in enumerate(rows[name]):
                covars[s, q, idx[c[k]]] = v
                covar_in[s, idx[c[k]]] = 1
                if k < N:
                    covars_pos[s, q, k] = v
        covar_count[s] = len(rows)
    return dict(sbgrps=sbgrps, samples=samples, gcoord=gcoord, explev=explev, scoord=scoord,
                geno_vals=geno_vals, geno_in=geno_in, covars=covars, covar_count=covar_count,
                covar_in=covar_in, covars_pos=covars_pos)


def expected_chunks(inp, radius, wrtsize, mvlr=False, anchor_tes=False):
    """Oracle chunks in the reference's order: [(gene names, HostChunk, snp names)].
    anchor_tes: cis window [start - radius, end + radius] (--anchor TSS+TES)."""
    S = len(inp["sbgrps"])
    N = len(inp["samples"])
    genes = sorted(g for g in inp["gcoord"] if g in inp["explev"])
    snps_by_chr = {}
    for n, (c, p) in inp["scoord"].items():
        if any(n in gv for gv in inp["geno_vals"]):
            snps_by_chr.setdefault(c, []).append((p, n))
    for c in snps_by_chr:
        snps_by_chr[c].sort()
    out = []
    for k in range(0, len(genes), wrtsize):
        names = genes[k:k + wrtsize]
        Gc = len(names)
        expl = np.full((Gc, S, N), np.nan)
        gene_in = np.zeros((Gc, S), np.uint8)
        pair_offset = [0]
        pair_snps = []
        for gi, gname in enumerate(names):
            for s, y in inp["explev"][gname].items():
                expl[gi, s] = y
                gene_in[gi, s] = 1
            chrom, start, end = inp["gcoord"][gname]
            lo = start - radius if start >= radius else -1
            hi = (end if anchor_tes else start) + radius
            cis = [n for p, n in snps_by_chr.get(chrom, []) if lo <= p <= hi]
            has = any(gene_in[gi, s] and n in inp["geno_vals"][s] for n in cis for s in range(S))
            if not has or (mvlr and not gene_in[gi].all()):
                cis = []
            pair_snps += cis
            pair_offset.append(len(pair_snps))
        uniq = sorted(set(pair_snps), key=pair_snps.index)
        col = {n: j for j, n in enumerate(uniq)}
        P = max(len(uniq), 1)
        geno = np.zeros((S, N, P))
        snp_in = np.zeros((P, S), np.uint8)
        for j, n in enumerate(uniq):
            for s in range(S):
                if n in inp["geno_vals"][s]:
                    geno[s, :, j] = inp["geno_vals"][s][n][0]
                    snp_in[j, s] = 1
        hc = HostChunk(explevels=expl, gene_in_sbgrp=gene_in, geno_in_sbgrp=inp["geno_in"], genotypes=geno,
                       snp_in_sbgrp=snp_in, pair_offset=np.array(pair_offset),
                       pair_snp=np.array([col[n] for n in pair_snps], np.int32), covars=inp["covars"],
                       covar_count=inp["covar_count"], covar_in_sbgrp=inp["covar_in"],
                       geno_set_of_sbgrp=np.arange(S, dtype=np.int32))
        out.append((names, hc, pair_snps))
    return out


def config_labels(bfs, S):
    return ["-".join(str(s + 1) for s in range(S) if (m >> s) & 1) for m in config_masks(bfs, S)]


def make_inss_case(root, seed=7):
    """Summary-statistics inputs for --inss: a list file naming one file per
    subgroup (one gzipped, one with extra columns in another order), pairs
    that appear in only some subgroups, a repeated line (the first is kept),
    n = 0, n = 2, a 'nan' effect and a zero effect.  Returns the list path."""
    rng = np.random.default_rng(seed)
    sbgrps = ["tissueB", "tissueA", "tissueC"]  # sorted: tissueA, tissueB, tissueC
    genes = {"gZ": ["rs%d" % k for k in range(1, 6)], "gA": ["rs%d" % k for k in range(3, 11)],
             "gM": ["rs%d" % k for k in range(20, 23)]}
    paths = {}
    for si, sg in enumerate(sbgrps):
        lines = []
        for gname, snps in genes.items():
            for j, snp in enumerate(snps):
                if (j + si) % 4 == 3:
                    continue  # this pair has no line in this subgroup
                n = int(rng.integers(20, 300))
                sig = rng.uniform(0.5, 2.0)
                beta = rng.normal(scale=0.3) * (3.0 if j == 1 else 1.0)
                se = rng.uniform(0.03, 0.3)
                lines.append((gname, snp, n, sig, beta, se))
        rng.shuffle(lines)
        if si == 0:
            g, sn, _n, sig, beta, se = lines[0]
            lines.append((g, sn, 999, sig * 2, -beta, se))  # repeated pair: ignored
            lines[1] = lines[1][:2] + (0,) + lines[1][3:]
            lines[2] = lines[2][:2] + (2,) + lines[2][3:]
        if si == 1:
            lines[0] = lines[0][:4] + ("nan",) + lines[0][5:]
            lines[1] = lines[1][:4] + (0.0,) + lines[1][5:]
        if si == 2:
            hdr = ["maf", "sebetahat.geno", "snp", "gene", "betahat.geno", "pve", "n", "sigmahat", "betapval.geno"]
            rows = ["\t".join(hdr)]
            for g, sn, n, sig, beta, se in lines:
                d = {"maf": "0.2", "sebetahat.geno": "%.10g" % se, "snp": sn, "gene": g,
                     "betahat.geno": beta if isinstance(beta, str) else "%.10g" % beta, "pve": "0.01",
                     "n": str(n), "sigmahat": "%.10g" % sig, "betapval.geno": "0.5"}
                rows.append("\t".join(d[h] for h in hdr))
        else:
            rows = ["gene\tsnp\tn\tsigmahat\tbetahat.geno\tsebetahat.geno"]
            for g, sn, n, sig, beta, se in lines:
                rows.append("%s\t%s\t%d\t%.10g\t%s\t%.10g" % (g, sn, n, sig, beta if isinstance(beta, str)
                                                                else "%.10g" % beta, se))
        p = os.path.join(root, "ss_%s.txt" % sg) + (".gz" if si == 1 else "")
        if si == 1:
            with gzip.open(p, "wt") as f:
                f.write("\n".join(rows) + "\n")
        else:
            _w(p, rows)
        paths[sg] = p
    lst = os.path.join(root, "list_ss.txt")
    _w(lst, ["#subgroup path"] + ["%s %s" % (sg, paths[sg]) for sg in sbgrps])
    return lst


def expected_inss(list_path):
    """The pairs --inss builds, in the reference's order: genes by name, pairs
    by first appearance over the subgroups' files (subgroups by name), the
    first line of a (gene, SNP, subgroup) kept.  Returns (subgroups,
    [(gene, [snp, ...])], n, sigmahat, betahat, sebetahat [S][NP], present
    [S][NP])."""
    files = {}
    with open(list_path) as f:
        for l in f:
            t = l.split()
            if t and not t[0].startswith("#"):
                files.setdefault(t[0], t[1])
    sbgrps = sorted(files)
    S = len(sbgrps)
    genes, order = {}, []
    for s, sg in enumerate(sbgrps):
        opener = gzip.open if files[sg].endswith(".gz") else open
        with opener(files[sg], "rt") as f:
            lines = [l.rstrip("\n") for l in f if l.strip()]
        hdr = lines[0].split("\t")
        c = {k: hdr.index(k) for k in ("gene", "snp", "n", "sigmahat", "betahat.geno", "sebetahat.geno")}
        for l in lines[1:]:
            t = l.split("\t")
            g = genes.setdefault(t[c["gene"]], {})
            if t[c["snp"]] not in g:
                g[t[c["snp"]]] = [None] * S
            if g[t[c["snp"]]][s] is None:
                g[t[c["snp"]]][s] = (int(t[c["n"]]), float(t[c["sigmahat"]]), float(t[c["betahat.geno"]]),
                                     float(t[c["sebetahat.geno"]]))
    pairs = [(gname, list(genes[gname].keys())) for gname in sorted(genes)]
    NP = sum(len(p) for _g, p in pairs)
    n = np.zeros((S, NP), np.int64)
    sig, beta, se = (np.full((S, NP), np.nan) for _ in range(3))
    present = np.zeros((S, NP), bool)
    p = 0
    for gname, snps in pairs:
        for snp in snps:
            for s in range(S):
                e = genes[gname][snp][s]
                if e is not None:
                    present[s, p] = True
                    n[s, p], sig[s, p], beta[s, p], se[s, p] = e
            p += 1
    return sbgrps, pairs, n, sig, beta, se, present
