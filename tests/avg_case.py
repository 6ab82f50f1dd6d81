"""Raw log10 ABF files in eqtlbma_bf's --outraw layout for eqtlbma_avg_bfs,
and the reference program's reading of them restated independently of the
host program (src/eqtlbma_avg_bfs.cpp:1103-1193): header "gene snp
config", rows of a gene and of a SNP consecutive, only the configurations
named in the weights file (or "1".."S" for the types model) kept, the first
|grid weights| values of a row, --gene / --snp / --gene-snp filters.
"""
import gzip
import os

import numpy as np

from eqtlbma_amd._abi import config_masks


def cfg_names(S):
    return ["-".join(str(s + 1) for s in range(S) if (m >> s) & 1) for m in config_masks("all", S)]


def write_raw(root, seed, S=3, L=6, n_files=2, genes_per_file=4, snps=(3, 12)):
    """Writes n_files gzipped raw files; returns their paths.  Rows per SNP:
    gen, gen-fix, gen-maxh, then every configuration (values with NaN
    here and there and "nan" tokens as eqtlbma_bf prints them)."""
    rng = np.random.default_rng(seed)
    names = cfg_names(S)
    paths = []
    g = 0
    for f in range(n_files):
        lines = ["gene\tsnp\tconfig" + "".join("\tl10abf.grid%d" % (l + 1) for l in range(L))]
        for _ in range(genes_per_file):
            gname = "gene%03d" % g
            g += 1
            for j in range(rng.integers(snps[0], snps[1] + 1)):
                sname = "rs%d_%d" % (g, j)
                hot = rng.random() < 0.15
                for c in ["gen", "gen-fix", "gen-maxh"] + names:
                    v = rng.normal(-0.5, 0.6, L) + (rng.uniform(1, 4) if hot else 0)
                    toks = ["nan" if rng.random() < 0.01 else "%.6e" % x for x in v]
                    lines.append("\t".join([gname, sname, c] + toks))
        p = os.path.join(root, "out_%d_l10abfs_raw.txt.gz" % f)
        with gzip.open(p, "wt") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(p)
    return paths


def read_raw(paths, L, keep_cfgs, genes=None, snps=None, pairs=None):
    """[(gene names, snp names, raw [pairs][R][L], off)] per file."""
    out = []
    for p in sorted(paths):
        with gzip.open(p, "rt") as fh:
            lines = fh.read().split("\n")
        lines = [l for l in lines if l]
        gnames, snames, rows, off = [], [], [], [0]
        cur_g, cur_s = None, None
        for l in lines[1:]:
            t = l.split("\t")
            if t[2] not in keep_cfgs:
                continue
            if genes is not None and t[0] not in genes:
                continue
            if snps is not None and t[1] not in snps:
                continue
            if pairs is not None and (t[0] not in pairs or t[1] not in pairs[t[0]]):
                continue
            if t[0] != cur_g:
                if cur_g is not None:
                    off.append(len(snames))
                cur_g = t[0]
                gnames.append(cur_g)
                cur_s = None
            if t[1] != cur_s:
                cur_s = t[1]
                snames.append(cur_s)
                rows.append([])
            rows[-1].append([float(x) for x in t[3:3 + L]])
        if cur_g is not None:
            off.append(len(snames))
        R = len(keep_cfgs)
        raw = np.full((len(snames), R, L), np.nan)
        for i, r in enumerate(rows):
            raw[i, :len(r)] = np.array(r)
        out.append((gnames, snames, raw, np.array(off, np.int64)))
    return out
