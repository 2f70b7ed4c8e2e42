"""GPU: the coarsening phase alone (hgp_coarsen) against the oracle, the
per-level statistics, and the `hunyuangraph` command end to end: best of two
runs with a host recount and the partition file (hunyuangraph.cu:47-96), and
the study modes of the reference's -DFIGURE9/10/14 builds."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, csr_from_edges, random_graph

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "hunyuansc25_amd", "hunyuangraph")


def small_graphs(hgp):
    out = {"grid48": hgp.gen_grid2d(48, 48)}
    out["er400w"] = random_graph(400, 0.02, seed=4, wmax=9, vmax=3)
    out["rmat12"] = hgp.gen_rmat(12, 1 << 14, seed=5)
    out["rgg3kw"] = hgp.gen_rgg3d(3000, hgp.rgg_radius(3000, 12), seed=7, vw=(1, 4), ew=(1, 16))
    edges = [(0, i) for i in range(1, 301)] + [(301, 302)]
    x, a, w = csr_from_edges(303, edges)
    out["star300"] = (x, a, np.ones(303, np.int32), w)
    return out


@pytest.mark.parametrize("k", [2, 32])
def test_coarsen_bit_exact(hgp, orc, k):
    for name, (x, a, vw, aw) in small_graphs(hgp).items():
        got, st = hgp.coarsen(x, a, k, vw, aw, seed=3, level_stats=1)
        want, so = orc.dga_coarsen(x, a, k, vw, aw, seed=3)
        for g, w, key in zip(got, want, ["xadj", "adjncy", "vwgt", "adjwgt"]):
            np.testing.assert_array_equal(g, w, err_msg=f"{name} k={k} {key}")
        for key in ["levels", "coarsest_n", "coarsest_m", "level_n", "level_m", "level_adjwgt"]:
            assert st[key] == so[key], (name, k, key)
        L = st["levels"]
        assert all(t >= 0 for t in st["level_ms"][:L]) and st["level_ms"][L] == 0
        assert sum(st["level_ms"]) <= st["ms_coarsen"] + 0.05


def test_level_stats_leave_the_partition_alone(hgp, orc):
    x, a, vw, aw = hgp.gen_rmat(13, 1 << 15, seed=2)
    p1, s1 = hgp.partition(x, a, 16, vw, aw, seed=2)
    p2, s2 = hgp.partition(x, a, 16, vw, aw, seed=2, level_stats=1)
    np.testing.assert_array_equal(p1, p2)
    assert s1["cut"] == s2["cut"]
    _, so = orc.dga_partition(x, a, 16, vw, aw, seed=2)
    assert s2["level_adjwgt"] == so["level_adjwgt"]
    assert not any(s1["level_adjwgt"])


def cli(*args):
    cmd = [CLI, *map(str, args)]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cli_best_of_two_and_partition_file(hgp, orc, tmp_path):
    n = 20000
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12), seed=11, vw=(1, 3), ew=(1, 9))
    path = str(tmp_path / "rgg20k.graph")
    hgp.write_graph(path, x, a, vw, aw)
    out = cli(path, 8, 1, "--write")
    assert out.startswith(f"graph:{path} {n} {len(a)} 8 1")
    cuts = [int(c) for c in re.findall(r"edge-cut=\s+(\d+)", out)]
    imbs = [float(c) for c in re.findall(r"imbalance=\s+([\d.]+)", out)]
    runs = [hgp.partition(x, a, 8, vw, aw, seed=s) for s in (1, 2)]
    assert cuts == [st["cut"] for _, st in runs]
    for (p, st), imb in zip(runs, imbs):
        assert abs(imb - st["max_pwgt"] * 8 / st["tvwgt"]) < 6e-4
    best = min(range(2), key=lambda i: cuts[i])
    part = np.loadtxt(f"{path}.part.8", dtype=np.int32)
    np.testing.assert_array_equal(part, runs[best][0])
    np.testing.assert_array_equal(part, orc.dga_partition(x, a, 8, vw, aw, seed=best + 1)[0])
    for key in ["best_alltime", "best_coarsentime", "best_inittime", "best_uncoarsentime"]:
        assert re.search(key + r"=\s+[\d.]+", out), key


def test_cli_reads_matrix_market(hgp, orc, tmp_path):
    import scipy.io
    import scipy.sparse as sp
    A = sp.random(3000, 3000, density=0.002, random_state=np.random.default_rng(3), format="coo")
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, A)
    out = cli(path, 4, 0, "--runs", "1", "--seed", "5", "--write")
    x, a, vw, aw = hgp.read_mtx(path)
    p, st = hgp.partition(x, a, 4, vw, aw, seed=5)
    assert int(re.search(r"edge-cut=\s+(\d+)", out).group(1)) == st["cut"]
    np.testing.assert_array_equal(np.loadtxt(f"{path}.part.4", dtype=np.int32), p)


def test_cli_studies(hgp, orc, tmp_path):
    n = 30000
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 10), seed=13, vw=(1, 2), ew=(1, 5))
    path = str(tmp_path / "rgg30k.graph")
    hgp.write_graph(path, x, a, vw, aw)
    (cx, ca, cv, cw), st = hgp.coarsen(x, a, 16, vw, aw, seed=1, level_stats=1)
    L = st["levels"]

    # -DFIGURE9_SUM: one line per coarse level
    out = cli(path, 16, "--study", "coarsen-sum")
    rows = re.findall(r"level\s+(\d+): nvtxs\s+(\d+) nedges\s+(\d+) nedges/nvtxs=\s*[\d.]+ adjwgtsum\s+(\d+)", out)
    assert [tuple(map(int, r)) for r in rows] == [
        (l, st["level_n"][l], st["level_m"][l], st["level_adjwgt"][l]) for l in range(1, L + 1)]

    # -DFIGURE9_TIME: one time per level built
    out = cli(path, 16, "--study", "coarsen-time")
    times = re.findall(r"level=\s*(\d+) time=\s*([\d.]+) ms", out)
    assert [int(l) for l, _ in times] == list(range(1, L + 1))

    # -DFIGURE10_CGRAPH: the coarsest graph as a weighted .graph file
    cg = str(tmp_path / "coarsest.graph")
    cli(path, 16, "--study", f"coarsest-graph={cg}")
    assert open(cg).readline().split()[2] == "11"
    for got, want in zip(hgp.read_graph(cg), (cx, ca, cv, cw)):
        np.testing.assert_array_equal(got, want)

    # -DFIGURE10_SAMPLING / _EXHAUSTIVE: initial partition of the dumped graph
    out = cli(cg, 16, "--study", "init-sampling")
    w = orc.dga_initpart(cx, ca, cv, cw, 16, seed=1)
    assert int(re.search(r"edgecut=\s*(\d+)", out).group(1)) == orc.edge_cut(cx, ca, cw, w)
    out = cli(cg, 2, "--study", "init-exhaustive")
    nc = len(cx) - 1
    w = orc.dga_initpart(cx, ca, cv, cw, 2, seed=1, trials=nc)
    assert re.search(r"trials=(\d+)", out).group(1) == str(nc)
    assert int(re.search(r"edgecut=\s*(\d+)", out).group(1)) == orc.edge_cut(cx, ca, cw, w)

    # -DFIGURE14_EDGECUT: refinement patience 12, best cut of the runs
    out = cli(path, 16, "--study", "edgecut")
    best = int(re.search(r"best_edgecut=\s+(\d+)", out).group(1))
    assert best == min(hgp.partition(x, a, 16, vw, aw, seed=s, patience=12)[1]["cut"] for s in (1, 2))
