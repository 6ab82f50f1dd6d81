"""End to end: genotype file formats and the SNP / subgroup filters of the
eqtlbma_bf host program.

Formats (Snp::AddSubgroupFrom{Vcf,Impute,Dose}Line, snp.cpp:88-185; the
loaders, data_loader.cpp:240-320, 570-720): the same genotypes written as
VCF (GT at the second FORMAT index, '.' alleles for a missing call, SNP
coordinates from the file) and as IMPUTE (genotype-probability triples, an
all-zero triple for a missing call, fractional probabilities giving a
dosage AB + 2 BB, sample names from the 'name_a1a1' header) must give the
same output files, byte for byte, as the custom dose format with 'NA' at
the same calls and --scoord (whose outputs test_gpu_cli.py checks against
the oracle).

Filters, against the oracle with the reference's rules restated here:
--maf (a SNP is dropped in a subgroup whose genotype file gives it a minor
allele frequency below the cutoff, Snp::EraseIfLowMafPerSubgroup,
snp.cpp:245-254), --snp (only the listed SNPs are read, loadSnpsToKeep,
data_loader.cpp:529-556), --anchor TSS+TES (cis window [start - r, end + r],
Snp::IsInCis, snp.cpp:274-297) and --sbgrp (only the listed subgroups are
loaded, eqtlbma_bf.cpp:438): the program's results must equal the oracle's
on the inputs those rules leave.
"""
import math
import os
import shutil

import numpy as np
import pytest

import cli_case
from test_gpu_cli import _close, _common_args, _num, _run

pytestmark = pytest.mark.gpu

FILES = ["_sumstats_s1.txt.gz", "_sumstats_s2.txt.gz", "_sumstats_s3.txt.gz", "_l10abfs_raw.txt.gz",
         "_l10abfs_avg-grids.txt.gz"]


def _read_lines(path):
    with open(path) as f:
        return [l.rstrip("\n").split() for l in f]


def _custom_geno(root):
    """{subgroup: (file, header samples, {snp: values as written})}."""
    out = {}
    for sg, path in (l for l in _read_lines(os.path.join(root, "list_geno.txt"))):
        rows = _read_lines(path)
        out[sg] = (path, rows[0][1:], {r[0]: r[1:] for r in rows[1:]})
    return out


def _snp_pos(root):
    return {r[3]: (r[0], int(r[2])) for r in _read_lines(os.path.join(root, "snps.bed"))}


def _write_vcf(root, geno, pos):
    lst = []
    for sg, (path, samples, vals) in geno.items():
        p = os.path.join(root, "geno_%s.vcf" % sg)
        with open(p, "w") as f:
            f.write("##fileformat=VCFv4.1\n##source=test\n")
            f.write("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(samples) + "\n")
            for snp, v in vals.items():
                chrom, ps = pos[snp]
                gts = []
                for k, x in enumerate(v):
                    if x in ("NA", "na", "nan", "NaN"):
                        gt = "./." if k % 2 == 0 else ".|0"
                    else:
                        d = int(float(x))
                        gt = {0: "0/0", 1: "0|1" if k % 2 else "1|0", 2: "1/1"}[d]
                    gts.append("%s:%s" % (x if x[0].isdigit() else "0", gt))  # DS:GT, GT second
                f.write("%s\t%d\t%s\tA\tG\t.\tPASS\t.\tDS:GT\t%s\n" % (chrom, ps, snp, "\t".join(gts)))
        lst.append("%s %s" % (sg, p))
    with open(os.path.join(root, "list_geno_vcf.txt"), "w") as f:
        f.write("\n".join(lst) + "\n")


def _write_impute(root, geno, pos):
    lst = []
    for sg, (path, samples, vals) in geno.items():
        p = os.path.join(root, "geno_%s.impute" % sg)
        with open(p, "w") as f:
            hdr = ["chr", "name", "coord", "a1", "a2"]
            for smp in samples:
                hdr += [smp + "_a1a1", smp + "_a1a2", smp + "_a2a2"]
            f.write(" ".join(hdr) + "\n")
            for snp, v in vals.items():
                chrom, ps = pos[snp]
                trip = []
                for x in v:
                    if x in ("NA", "na", "nan", "NaN"):
                        trip += ["0", "0", "0"]
                    elif float(x) == 1.375:
                        trip += ["0.125", "0.375", "0.5"]  # AB + 2 BB = 1.375 exactly
                    else:
                        d = int(float(x))
                        trip += {0: ["1", "0", "0"], 1: ["0", "1", "0"], 2: ["0", "0", "1"]}[d]
                f.write("%s %s %d A G %s\n" % (chrom, snp, ps, " ".join(trip)))
        lst.append("%s %s" % (sg, p))
    with open(os.path.join(root, "list_geno_impute.txt"), "w") as f:
        f.write("\n".join(lst) + "\n")


def _set_custom_value(root, geno, sg, snp, k, value):
    path, samples, vals = geno[sg]
    vals[snp][k] = value
    rows = ["Id\t" + "\t".join(samples)] + [s + "\t" + "\t".join(v) for s, v in vals.items()]
    with open(path, "w") as f:
        f.write("\n".join(rows) + "\n")


def _args_without_scoord(args):
    i = args.index("--scoord")
    return args[:i] + args[i + 2:]


def _same_outputs(a, b):
    import gzip
    for suf in FILES:
        with gzip.open(a + suf, "rb") as fa, gzip.open(b + suf, "rb") as fb:
            assert fa.read() == fb.read(), suf


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("cli_formats"))
    c = cli_case.make_case(root, seed=23, covariates=False)
    return c


@pytest.mark.parametrize("fmt", ["vcf", "impute"])
def test_cli_genotype_formats_equal_custom(case, fmt):
    root = case["root"]
    geno = _custom_geno(root)
    pos = _snp_pos(root)
    if fmt == "impute":
        # fractional dosages, representable exactly both ways
        _set_custom_value(root, geno, "s1", "rs1_3", 2, "1.375")
        _set_custom_value(root, geno, "s3", "rs2_5", 0, "1.375")
        _write_impute(root, geno, pos)
    else:
        _write_vcf(root, geno, pos)
    ref_out = os.path.join(root, "custom_" + fmt)
    _run(_common_args(root, ref_out, "join") + ["--bfs", "all", "--outss", "--outw"], root)
    out = os.path.join(root, fmt)
    args = _common_args(root, out, "join") + ["--bfs", "all", "--outss", "--outw"]
    args[args.index("--geno") + 1] = os.path.join(root, "list_geno_%s.txt" % fmt)
    _run(_args_without_scoord(args), root)
    _same_outputs(ref_out, out)


def _oracle_filtered(orc, case, inp, radius, wrtsize, anchor_tes):
    """Expected rows of the sumstats files with the filters applied to the
    inputs (the oracle's association test on what the reference keeps)."""
    gL, gS = case["gL"], case["gS"]
    exp_ss = {}
    chunks = cli_case.expected_chunks(inp, radius, wrtsize, anchor_tes=anchor_tes)
    S = len(inp["sbgrps"])
    for names, hc, psn in chunks:
        ref = orc.assoc_uvlr(hc, "join", "all", gL, gS)
        for g, gname in enumerate(names):
            for p in range(hc.pair_offset[g], hc.pair_offset[g + 1]):
                for s in range(S):
                    if ref.n[s, p] > 0:
                        exp_ss.setdefault(s, []).append([gname, psn[p], inp["geno_vals"][s][psn[p]][1], ref.n[s, p]] +
                                                        list(ref.sstats[:, s, p]))
    return exp_ss


def test_cli_maf_snp_anchor_sbgrp_filters(case, orc, tmp_path):
    root = case["root"]
    # a low-frequency SNP: one carrier (heterozygous) in every genotype file
    geno = _custom_geno(root)
    for sg, (path, samples, vals) in geno.items():
        v = ["0"] * len(samples)
        v[0] = "1"
        vals["rs1_4"] = v
        rows = ["Id\t" + "\t".join(samples)] + [s + "\t" + "\t".join(x) for s, x in vals.items()]
        with open(path, "w") as f:
            f.write("\n".join(rows) + "\n")
    keep = [n for n, _c, _p in case["snps"] if n not in ("rs1_2", "rs2_6", "rs1_11")]
    snp_file = os.path.join(root, "snps_keep.txt")
    with open(snp_file, "w") as f:
        f.write("#names\n" + "\n".join(keep) + "\n")
    min_maf = 0.05
    out = os.path.join(root, "filt")
    args = _common_args(root, out, "join") + ["--bfs", "all", "--outss", "--maf", str(min_maf), "--snp", snp_file,
                                              "--anchor", "TSS+TES", "--sbgrp", "s1+s3"]
    _run(args, root)
    # the expected side: list files with s1 and s3 only, then the rules
    sub = str(tmp_path)
    for lst in ("list_exp.txt", "list_geno.txt"):
        rows = [r for r in _read_lines(os.path.join(root, lst)) if r[0] in ("s1", "s3")]
        with open(os.path.join(sub, lst), "w") as f:
            f.write("\n".join(" ".join(r) for r in rows) + "\n")
    for fname in ("genes.bed", "snps.bed"):
        shutil.copy(os.path.join(root, fname), os.path.join(sub, fname))
    inp = cli_case.load_expected_inputs({"root": sub, "sbgrps": ["s1", "s3"]})
    for s in range(len(inp["sbgrps"])):
        # the reference holds the cutoff as a float (eqtlbma_bf.cpp:1590)
        inp["geno_vals"][s] = {k: v for k, v in inp["geno_vals"][s].items()
                               if k in keep and not v[1] < float(np.float32(min_maf))}
    assert "rs1_4" not in inp["geno_vals"][0] and "rs1_4" not in inp["geno_vals"][1]
    exp_ss = _oracle_filtered(orc, case, inp, 2500, 2, True)
    for s, sg in enumerate(inp["sbgrps"]):
        rows = cli_case.read_table(out + "_sumstats_%s.txt.gz" % sg)
        assert len(rows) - 1 == len(exp_ss.get(s, [])), sg
        for got, want in zip(rows[1:], exp_ss[s]):
            assert got[:2] == want[:2] and int(got[3]) == want[3]
            assert all(_close(_num(a), b) for a, b in zip([got[2]] + got[4:], [want[2]] + want[4:])), (got, want)
    assert not os.path.exists(out + "_sumstats_s2.txt.gz")
