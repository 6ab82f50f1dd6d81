"""The eqtlbma_bf host program without a GPU: help, the reference's option
checks (eqtlbma_bf.cpp:432-693) and the early exit when nothing is left to
analyse (run() returns before any output, :1503-1504)."""
import os
import subprocess

import pytest

from eqtlbma_amd import CLI_PATH
import cli_case


def _run(args, cwd):
    return subprocess.run([CLI_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


@pytest.fixture(scope="module")
def root(tmp_path_factory):
    r = str(tmp_path_factory.mktemp("clihost"))
    cli_case.make_case(r)
    return r


def _base(root, out):
    return ["--geno", os.path.join(root, "list_geno.txt"), "--scoord", os.path.join(root, "snps.bed"),
            "--exp", os.path.join(root, "list_exp.txt"), "--gcoord", os.path.join(root, "genes.bed"),
            "--out", out, "-v", "0"]


def test_help_and_version():
    r = _run(["--help"], "/tmp")
    assert r.returncode == 0 and "--gridL" in r.stdout and "--pbf" in r.stdout
    assert _run(["-V"], "/tmp").returncode == 0


@pytest.mark.parametrize("extra,msg", [
    ([], "missing compulsory option --analys"),
    (["--analys", "both"], "--analys both is not valid"),
    (["--analys", "join"], "missing compulsory option --gridL with --analys join"),
    (["--analys", "join", "--gridL", "GL", "--bfs", "sin"], "--gridS is required with --analys join and --bfs sin"),
    (["--analys", "join", "--gridL", "GL", "--bfs", "one"], "--bfs one is not valid"),
    (["--analys", "join", "--gridL", "GL", "--error", "xx"], "--error xx is not valid"),
    (["--analys", "sep", "--nperm", "10"], "--permsep should be '1' or '2'"),
    (["--analys", "join", "--gridL", "GL", "--nperm", "10"], "--pbf should be different than 'none'"),
    (["--analys", "join", "--gridL", "GL", "--nperm", "10", "--pbf", "all"], "--pbf should be 'gen'"),
    (["--analys", "join", "--gridL", "GL", "--gridS", "GS", "--bfs", "sin", "--nperm", "10", "--pbf", "all"],
     "--pbf should be 'gen' or 'gen-sin'"),
    (["--analys", "sep", "--nperm", "5", "--permsep", "1", "--trick", "1", "--tricut", "10"],
     "--tricut 10 is larger than --nperm 5"),
    (["--analys", "sep", "--wrtsize", "0"], "--wrtsize should be greater than 1"),
    (["--analys", "sep", "--anchor", "TES"], "--anchor should be TSS or TSS+TES"),
])
def test_option_checks(root, extra, msg):
    extra = [os.path.join(root, "gridL.txt") if a == "GL" else os.path.join(root, "gridS.txt") if a == "GS" else a
             for a in extra]
    r = _run(_base(root, os.path.join(root, "o")) + extra, root)
    assert r.returncode == 1
    assert msg in r.stderr


def test_no_subgroup_left_writes_nothing(root):
    out = os.path.join(root, "none")
    r = _run(_base(root, out) + ["--analys", "sep", "--sbgrp", "nosuch"], root)
    assert r.returncode == 0, r.stderr
    assert not [f for f in os.listdir(root) if f.startswith("none_")]


@pytest.mark.parametrize("extra,msg", [
    (["--analys", "sep"], "--inss requires --analys join"),
    (["--analys", "join", "--gridL", "GL", "--error", "mvlr"], "--inss requires --error uvlr"),
    (["--analys", "join", "--gridL", "GL", "--nperm", "5", "--pbf", "gen"],
     "--nperm with --inss is not supported by this build"),
])
def test_inss_option_checks(root, extra, msg):
    lst = cli_case.make_inss_case(root)
    extra = [os.path.join(root, "gridL.txt") if a == "GL" else a for a in extra]
    r = _run(["--inss", lst, "--out", os.path.join(root, "i"), "-v", "0"] + extra, root)
    assert r.returncode == 1
    assert msg in r.stderr


def test_inss_missing_column_and_empty(root):
    bad = os.path.join(root, "ss_bad.txt")
    with open(bad, "w") as f:
        f.write("gene\tsnp\tn\tsigmahat\tbetahat.geno\nga\trs1\t10\t1\t0.1\n")
    lst = os.path.join(root, "list_bad.txt")
    with open(lst, "w") as f:
        f.write("s1 %s\n" % bad)
    args = ["--analys", "join", "--gridL", os.path.join(root, "gridL.txt"), "-v", "0"]
    r = _run(["--inss", lst, "--out", os.path.join(root, "b")] + args, root)
    assert r.returncode == 1
    # the reference prints the missing column's index, std::string::npos
    assert "missing 18446744073709551615 in header of " + bad in r.stderr
    # a file with its header only: nothing to analyse, nothing written
    with open(bad, "w") as f:
        f.write("gene\tsnp\tn\tsigmahat\tbetahat.geno\tsebetahat.geno\n")
    r = _run(["--inss", lst, "--out", os.path.join(root, "empty")] + args, root)
    assert r.returncode == 0, r.stderr
    assert not [f for f in os.listdir(root) if f.startswith("empty_")]


def test_tabix_index_older_than_data(root, tmp_path):
    """A .tbi older than its BED file is refused (data_loader.cpp:1358-1363)."""
    import shutil
    import time
    from eqtlbma_amd import REPO_DIR
    ref = os.path.join(REPO_DIR, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "tabix")):
        pytest.skip("oracle/_ref has no tabix build")
    bed = str(tmp_path / "snps.bed")
    lines = sorted(open(os.path.join(root, "snps.bed")), key=lambda l: (l.split()[0], int(l.split()[1])))
    open(bed, "w").write("".join(lines))
    subprocess.run([os.path.join(ref, "bgzip"), "-f", bed], check=True)
    subprocess.run([os.path.join(ref, "tabix"), "-f", "-p", "bed", bed + ".gz"], check=True)
    t = os.path.getmtime(bed + ".gz.tbi")
    os.utime(bed + ".gz", (t + 10, t + 10))
    args = _base(root, str(tmp_path / "out"))
    args[args.index("--scoord") + 1] = bed + ".gz"
    r = _run(args + ["--analys", "sep"], str(tmp_path))
    assert r.returncode == 1
    assert "is older than data file" in r.stderr

