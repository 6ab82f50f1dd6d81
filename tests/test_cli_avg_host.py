"""The eqtlbma_avg_bfs host program without a GPU: help, the reference's
option checks in its order and words (eqtlbma_avg_bfs.cpp:281-365) and the
weight loaders' errors, all of which end before any GPU call."""
import os
import subprocess

import pytest

from eqtlbma_amd import CLI_AVG_PATH


def _run(args, cwd):
    return subprocess.run([CLI_AVG_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


@pytest.fixture
def root(tmp_path):
    (tmp_path / "gw.txt").write_text("0.5\n0.5\n")
    (tmp_path / "cw.txt").write_text("1\t0.5\n2\t0.3\n1-2\t0.2\n")
    return str(tmp_path)


def test_help(root):
    r = _run(["--help"], root)
    assert r.returncode == 0 and "--gwts" in r.stdout and "--bestsnp" in r.stdout


@pytest.mark.parametrize("extra,msg", [
    ([], "missing compulsory option --bf-out"),
    (["--in", "x"], "missing compulsory option --gwts"),
    (["--in", "x", "--gwts", "nope.txt"], "can't find file nope.txt"),
    (["--in", "x", "--gwts", "gw.txt", "--model", "both"], "--model both is unvalid"),
    (["--in", "x", "--gwts", "gw.txt"], "missing compulsory option --nsubgrp"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2"], "missing compulsory option --dim"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3"], "missing compulsory option --hm-in"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--save", "post"],
     "missing compulsory option --post with --save post"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--pi0", "0.5"],
     "missing compulsory option --post with --pi0"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--post", "a"],
     "missing compulsory option --pi0 with --post"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--thread", "0"],
     "--thread 0 is invalid"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--bestsnp", "3"],
     "--bestsnp 3 is invalid"),
    (["--in", "x", "--gwts", "gw.txt", "--nsubgrp", "2", "--dim", "3", "--out", "o.gz", "--bestsnp", "2"],
     "--bestsnp 2 can only be used with --save post"),
])
def test_option_checks(root, extra, msg):
    r = _run(extra, root)
    assert r.returncode == 1
    assert msg in r.stderr


@pytest.mark.parametrize("extra,msg", [
    (["--gtk", "0+2"], "--gtk doesn't correspond to --gwts"),
    (["--cwts", "cw.txt", "--dim", "4"], "should contain 4 config weights"),
])
def test_weight_file_errors(root, extra, msg):
    args = ["--in", "none_*", "--gwts", "gw.txt", "--nsubgrp", "2", "--out", "o.gz", "-v", "0"]
    if "--dim" not in extra:
        args += ["--dim", "3"]
    r = _run(args + extra, root)
    assert r.returncode == 1
    assert msg in r.stderr
