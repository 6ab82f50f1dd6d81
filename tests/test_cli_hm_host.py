"""The eqtlbma_hm host program without a GPU: help, the reference's option
checks (eqtlbma_hm.cpp:1956-2065) and the loader's errors (wrong header, no
input file), all of which end before any GPU call."""
import gzip
import os
import subprocess

import pytest

from eqtlbma_amd import CLI_HM_PATH


def _run(args, cwd="/tmp"):
    return subprocess.run([CLI_HM_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=60)


def test_help_and_version():
    r = _run(["--help"])
    assert r.returncode == 0 and "--nsubgrp" in r.stdout and "--msl" in r.stdout
    assert _run(["-V"]).returncode == 0


@pytest.mark.parametrize("extra,msg", [
    ([], "missing compulsory option --data"),
    (["--data", "x", "--model", "mixed"], "--model mixed is invalid"),
    (["--data", "x"], "missing compulsory option --nsubgrp"),
    (["--data", "x", "--nsubgrp", "3"], "missing compulsory option --dim"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "8"], "--model configs --nsubgrp 3 --dim 8 is invalid"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "7"], "missing compulsory option --ngrid"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "7", "--ngrid", "10"], "missing compulsory option --out"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "7", "--ngrid", "10", "--out", "o.gz", "--init", "/nonexistent"],
     "can't find /nonexistent"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "7", "--ngrid", "10", "--out", "o.gz", "--thresh", "0"],
     "--thresh 0 is invalid"),
    (["--data", "x", "--nsubgrp", "3", "--dim", "7", "--ngrid", "10", "--out", "o.gz", "--pi0", "1.5"],
     "--pi0 1.5 is invalid"),
])
def test_option_checks(extra, msg):
    r = _run(extra)
    assert r.returncode == 1
    assert "ERROR: " + msg in r.stderr


def test_ci_needs_getci(tmp_path):
    ci = tmp_path / "ci.txt"
    ci.write_text("pi0\t0.5\n")
    r = _run(["--data", "x", "--nsubgrp", "3", "--dim", "7", "--ngrid", "10", "--out", "o.gz", "--ci", str(ci)])
    assert r.returncode == 1 and "ERROR: --getci should be given with --ci" in r.stderr


def test_loader_errors(tmp_path):
    args = ["--nsubgrp", "2", "--dim", "3", "--ngrid", "2", "--out", str(tmp_path / "o.gz"), "-v", "0"]
    r = _run(["--data", str(tmp_path / "none_*.gz")] + args)
    assert r.returncode == 1 and "ERROR: no input file was found from pattern" in r.stderr
    bad = tmp_path / "bad.txt.gz"
    with gzip.open(bad, "wt") as f:
        f.write("gene\tsnp\tcfg\tl10abf.grid1\n")
    r = _run(["--data", str(bad)] + args)
    assert r.returncode == 1 and "has wrong header line" in r.stderr
