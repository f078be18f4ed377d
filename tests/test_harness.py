"""The ffgemm_bench command-line harness (the TensileLite.out drop-in):
argument parsing on the CPU; a verified run on the GPU whose
"]Performance = X TFLOPs" line parses the way test/perf_test.py:31-41 reads it."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tensilelite_amd", "bin", "ffgemm_bench")


def _need_bin():
    if not os.path.isfile(BIN):
        pytest.skip("ffgemm_bench not built")


def test_help_lists_reference_flags():
    _need_bin()
    out = subprocess.run([BIN, "-h"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    for flag in ("-d,", "-m,", "-n,", "-k,", "-p,", "-r,", "-s,", "-x,", "-y,", "-z,", "-u,", "-v,", "-l,", "-f,", "-t,"):
        assert flag in out.stdout, flag


def test_unknown_option_is_an_error():
    _need_bin()
    out = subprocess.run([BIN, "--bogus", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2


def _perf(stdout: str) -> float:
    # test/perf_test.py:35-40 reads the text between "]Performance = " and "TFLOPs"
    for line in stdout.splitlines():
        i, j = line.rfind("]Performance = "), line.rfind("TFLOPs")
        if i != -1:
            return float(line[i + len("]Performance = "):j - 1])
    raise AssertionError("no Performance line")


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["-d", "1", "-m", "480", "-n", "512", "-k", "512", "-p", "32"],          # the reference's CPU-runnable case
    ["-d", "2", "-m", "960", "-n", "1024", "-k", "1024", "-p", "32"],
    ["-d", "3", "-m", "1024", "-n", "1024", "-k", "1024", "--layout", "NN", "--init", "normal", "--seed", "1"],
    ["-d", "3", "-m", "512", "-n", "512", "-k", "512", "--round", "truncate", "-r", "2", "-s", "2", "-x", "2", "-y", "2"],
    ["-d", "1", "-m", "512", "-n", "512", "-k", "512", "--compute", "bf16x3", "--init", "normal"],
    ["-d", "4", "-m", "512", "-n", "512", "-k", "512", "--init", "normal"],
    # the reference's tile flags asking for 256 x 256 with 2 x 4 waves (the eight-wave engine)
    ["-d", "3", "-m", "1024", "-n", "1024", "-k", "1024", "-r", "4", "-s", "2", "-x", "2", "-y", "4", "-z", "32",
     "--init", "normal"],
    # odd padding (ld = rows + pad off the 16-byte grid) and K off the vector:
    # the general-stride kernel, with the record's tile or, for 256 x 256, the
    # library's choice (a warning)
    ["-d", "1", "-m", "480", "-n", "512", "-k", "512", "-p", "1"],
    ["-d", "2", "-m", "768", "-n", "640", "-k", "256", "-p", "3", "-r", "2", "-s", "2", "-x", "2", "-y", "2"],
    ["-d", "3", "-m", "480", "-n", "512", "-k", "100", "-p", "1", "--init", "normal"],
    ["-d", "3", "-m", "512", "-n", "512", "-k", "512", "-p", "1", "-r", "4", "-s", "2", "-x", "2", "-y", "4",
     "-z", "32", "--init", "normal"],
])
def test_cli_run_verifies_and_reports(cuda, args):
    _need_bin()
    print("case", " ".join(args), flush=True)
    out = subprocess.run([BIN, "-l", "5"] + args, capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-1000:], flush=True)
    assert out.returncode == 0, out.returncode
    assert "Verify Success." in out.stdout
    assert _perf(out.stdout) > 0
    assert re.search(r"Efficiency = [0-9.]+%", out.stdout)
