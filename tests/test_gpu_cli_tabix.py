"""GPU: eqtlbma_bf with a tabix-indexed --scoord (loadSnpInfo with a .tbi,
data_loader.cpp:1345-1400) against the reference's own tabix.

The BED file is compressed and indexed with the reference's bgzip and tabix
(built from its sources into oracle/_ref).  Expected SNP set: the records
the reference's tabix returns for each gene's region
(Gene::GetRegionInTabixFormat, gene.cpp:94-104, genes in name order), the
first record of each SNP name kept.  The run on the indexed file must write
the same files as a run on a plain BED holding exactly that set.  The cases:
a radius larger than a gene's start (the region's start wraps in size_t and
the query starts at 0), --anchor TSS+TES, and a SNP name given twice at two
positions, where the index's per-gene order decides which one is kept (the
plain path keeps the first in file order).
"""
import gzip
import os
import shutil
import subprocess
import sys

import pytest

from eqtlbma_amd import CLI_PATH, REPO_DIR
import cli_case

pytestmark = pytest.mark.gpu

REF = os.path.join(REPO_DIR, "oracle", "_ref")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args):
    _say("[gpu] eqtlbma_bf " + " ".join(args))
    r = subprocess.run([CLI_PATH] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


def _index(root, lines):
    """bgzip + tabix -p bed of the (sorted) lines, with the reference's tools."""
    plain = os.path.join(root, "snps_idx.bed")
    with open(plain, "w") as f:
        f.write("".join(l + "\n" for l in lines))
    subprocess.run([os.path.join(REF, "bgzip"), "-f", plain], check=True)
    subprocess.run([os.path.join(REF, "tabix"), "-f", "-p", "bed", plain + ".gz"], check=True)
    return plain + ".gz"


def _expected_lines(gz, genes_bed, anchor, radius):
    """Records the reference's tabix returns per gene region, first of each
    name kept (genes in name order)."""
    genes = []
    for l in open(genes_bed):
        t = l.split()
        genes.append((t[3], t[0], int(t[1]) + 1, int(t[2])))
    kept, order = {}, []
    for name, chrom, start, end in sorted(genes):
        lo = (start - radius) % (1 << 64)  # size_t arithmetic
        hi = (end if anchor == "TSS+TES" else start) + radius
        region = "%s:%d-%d" % (chrom, lo, hi)
        r = subprocess.run([os.path.join(REF, "tabix"), gz, region], stdout=subprocess.PIPE, text=True, check=True)
        for rec in r.stdout.splitlines():
            t = rec.split("\t")
            if t[3] not in kept:
                kept[t[3]] = rec
                order.append(t[3])
    return [kept[n] for n in order]


def _outputs(prefix):
    d = os.path.dirname(prefix)
    base = os.path.basename(prefix)
    out = {}
    for f in sorted(os.listdir(d)):
        if f.startswith(base + "_"):
            with gzip.open(os.path.join(d, f), "rt") as h:
                out[f[len(base):]] = h.read()
    return out


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "tabix")), reason="oracle/_ref has no tabix build")
@pytest.mark.parametrize("anchor,radius", [("TSS", 5000), ("TSS+TES", 2500)])
def test_tabix_scoord_matches_reference_tabix(tmp_path, anchor, radius):
    root = str(tmp_path)
    cli_case.make_case(root)
    lines = [l.rstrip("\n") for l in open(os.path.join(root, "snps.bed"))]
    # rs1_5 (chr1:3000) given a second time near geneA (chr1:9100)
    lines.append("chr1\t9099\t9100\trs1_5")
    lines.sort(key=lambda l: (l.split("\t")[0], int(l.split("\t")[1])))
    gz = _index(root, lines)
    exp = _expected_lines(gz, os.path.join(root, "genes.bed"), anchor, radius)
    assert "chr1\t9099\t9100\trs1_5" in exp and "chr1\t2999\t3000\trs1_5" not in exp
    plain = os.path.join(root, "snps_expected.bed")
    with open(plain, "w") as f:
        f.write("".join(l + "\n" for l in exp))
    common = ["--geno", os.path.join(root, "list_geno.txt"), "--exp", os.path.join(root, "list_exp.txt"),
              "--gcoord", os.path.join(root, "genes.bed"), "--cis", str(radius), "--anchor", anchor,
              "--analys", "join", "--gridL", os.path.join(root, "gridL.txt"), "--gridS",
              os.path.join(root, "gridS.txt"), "--bfs", "all", "--outss", "--outw", "--precision", "fp64",
              "-v", "0"]
    os.makedirs(os.path.join(root, "a"))
    os.makedirs(os.path.join(root, "b"))
    _run(common + ["--scoord", gz, "--out", os.path.join(root, "a", "obs")])
    _run(common + ["--scoord", plain, "--out", os.path.join(root, "b", "obs")])
    got, want = _outputs(os.path.join(root, "a", "obs")), _outputs(os.path.join(root, "b", "obs"))
    assert sorted(got) == sorted(want) and len(got) >= 4
    for k in want:
        assert got[k] == want[k], k
    # the duplicated name is tested against geneA, not geneB
    raw = got["_l10abfs_raw.txt.gz"]
    assert "geneA\trs1_5\t" in raw
