"""The `hunyuangraph` command on the CPU: argument handling and file errors
(the reference's program takes `<graph> <nparts> <pool>`, hunyuangraph.cu:5-13).
Partitioning through it is covered by the GPU tests."""
import os
import subprocess

import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "hunyuansc25_amd", "hunyuangraph")
GOLD = os.path.join(ROOT, "tests", "golden", "io")


def run(*args, cwd=None):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_built():
    assert os.access(CLI, os.X_OK), "build the package first (__graft_entry__.build())"


@pytest.mark.parametrize("args,reason", [
    ((), "expected <graph> <nparts> [pool]"),
    (("g.graph",), "expected <graph> <nparts> [pool]"),
    (("g.graph", "0"), "<nparts> must be a positive integer"),
    (("g.graph", "x"), "<nparts> must be a positive integer"),
    (("g.graph", "4", "2"), "[pool] must be 0 or 1"),
    (("g.graph", "4", "1", "extra"), "expected <graph> <nparts> [pool]"),
    (("g.graph", "4", "--runs", "0"), "--runs must be"),
    (("g.graph", "4", "--runs"), "--runs needs a value"),
    (("g.graph", "4", "--ub", "0.9"), "--ub must be"),
    (("g.graph", "4", "--study", "nothing"), "unknown --study mode"),
    (("g.graph", "4", "--study", "coarsen-sum=f"), "bad --study value"),
    (("g.graph", "4", "--frobnicate"), "unknown option"),
])
def test_usage_errors(args, reason):
    r = run(*args)
    assert r.returncode == 2
    assert reason in r.stderr and "usage: hunyuangraph" in r.stderr


def test_help():
    r = run("--help")
    assert r.returncode == 0 and "--study" in r.stdout


def test_file_errors(tmp_path):
    r = run(tmp_path / "absent.graph", 4)
    assert r.returncode == 1 and "cannot open" in r.stderr
    bad = tmp_path / "bad.graph"
    bad.write_text("3 1\n4\n1\n\n")
    r = run(bad, 2)
    assert r.returncode == 1 and "out of bounds" in r.stderr
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 4 1\n1 2 1\n")
    r = run(bad, 2)
    assert r.returncode == 1 and "square" in r.stderr


def test_no_device_is_an_error(hgp):
    if hgp.device_count() > 0:
        pytest.skip("a HIP device is present")
    r = run(os.path.join(GOLD, "metis_manual.graph"), 2, 1)
    # the graph line is printed (as the reference does after reading), then the
    # partition fails loudly: there is no CPU fallback
    assert r.stdout.startswith("graph:") and " 7 22 2 1" in r.stdout
    assert r.returncode == 1 and "no HIP device" in r.stderr
