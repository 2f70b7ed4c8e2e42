"""The library's CPU partitioner (hgp_partition_cpu, the reference's
hunyuangraph_rbbisection, hunyuangraph_CPU_initialpartition.h:361-374)
against the oracle's restatement of the same path (oracle_cpu_rb_partition):
the same seed gives the same partition, bit for bit, on meshes, geometric
and power-law graphs, weighted graphs and every k shape of the recursion
(powers of two, odd k, k = 3).  Host code only: runs without a GPU."""
import numpy as np
import pytest

from conftest import random_graph


def _graphs(hgp):
    return {
        "grid-40x30": hgp.gen_grid2d(40, 30),
        "er-700-w": random_graph(700, 0.01, seed=4, wmax=6, vmax=4),
        "rgg-3000-w": hgp.gen_rgg3d(3000, hgp.rgg_radius(3000, 10.0), seed=3, vw=(1, 4), ew=(1, 9)),
        "rmat-s12": hgp.gen_rmat(12, 1 << 15, seed=6),
        "banded-2000": hgp.gen_banded(2000, 16, 0.1, seed=2),
    }


@pytest.mark.parametrize("k", [2, 3, 7, 16, 64])
def test_cpu_path_equals_oracle(hgp, orc, k):
    for name, (x, a, vw, aw) in _graphs(hgp).items():
        for seed in (1, 9):
            p, cut = hgp.partition_cpu(x, a, k, vw, aw, seed=seed)
            q, cut_o = orc.cpu_rb_partition(x, a, k, vw, aw, seed=seed)
            np.testing.assert_array_equal(p, q, err_msg=f"{name} k={k} seed={seed}")
            assert cut == cut_o == orc.edge_cut(x, a, aw, p)
            assert p.min() >= 0 and p.max() < k


def test_cpu_path_edge_cases(hgp):
    p, cut = hgp.partition_cpu(np.array([0], np.int64), np.zeros(0, np.int32), 4)
    assert len(p) == 0 and cut == 0
    x, a, vw, aw = random_graph(50, 0.1, seed=1)
    p, cut = hgp.partition_cpu(x, a, 1, vw, aw)
    assert not p.any() and cut == 0
    with pytest.raises(hgp.HgpError) as e:
        hgp.partition_cpu(x, a, 0, vw, aw)
    assert e.value.code == -1
    bad = a.copy()
    bad[0] = 50
    with pytest.raises(hgp.HgpError) as e:
        hgp.partition_cpu(x, bad, 4, vw, aw)
    assert e.value.code == -1
