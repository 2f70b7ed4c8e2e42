"""Batched partitioning: every graph of a batch gets exactly the partition it
gets alone (same options), whatever the lane count and the mix."""
import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu


def mix(hgp):
    gs, ks = [], []
    gs.append(hgp.gen_grid2d(20, 30)); ks.append(4)
    gs.append(hgp.gen_rmat(11, 1 << 13, seed=3)); ks.append(8)
    gs.append(hgp.gen_rgg3d(1500, hgp.rgg_radius(1500, 12), seed=4, vw=(1, 4), ew=(1, 16))); ks.append(16)
    gs.append(hgp.gen_banded(800, 32, 0.1, seed=5)); ks.append(5)
    x, a, vw, aw = random_graph(300, 0.03, seed=6, wmax=9, vmax=4)
    gs.append((x, a, vw, aw)); ks.append(3)
    edges = [(0, i) for i in range(1, 301)] + [(301, 302)]
    x, a, w = csr_from_edges(303, edges)
    gs.append((x, a, None, None)); ks.append(4)
    gs.append((np.zeros(51, np.int64), np.zeros(0, np.int32), None, None)); ks.append(2)   # edgeless
    gs.append((np.zeros(2, np.int64), np.zeros(0, np.int32), None, None)); ks.append(3)    # one vertex
    gs.append(hgp.gen_grid2d(64, 64)); ks.append(1)
    return gs, ks


@pytest.mark.parametrize("lanes", [1, 3, 8])
def test_batch_equals_alone(hgp, lanes):
    gs, ks = mix(hgp)
    parts, stats = hgp.partition_batch(gs, ks, lanes=lanes, seed=3)
    for i, ((x, a, vw, aw), k) in enumerate(zip(gs, ks)):
        p, st = hgp.partition(x, a, k, vw, aw, seed=3)
        np.testing.assert_array_equal(parts[i], p, err_msg=f"graph {i}")
        assert stats[i]["cut"] == st["cut"], i


def test_batch_repeat_and_reorder(hgp):
    gs, ks = mix(hgp)
    p1, _ = hgp.partition_batch(gs, ks, lanes=4, seed=9)
    rev = list(range(len(gs)))[::-1]
    p2, _ = hgp.partition_batch([gs[i] for i in rev], [ks[i] for i in rev], lanes=2, seed=9)
    for j, i in enumerate(rev):
        np.testing.assert_array_equal(p1[i], p2[j])


def test_batch_reports_bad_graph(hgp):
    gs, ks = mix(hgp)
    x, a, _, _ = gs[0]
    bad = a.copy()
    bad[0] = len(x) + 5          # neighbour id out of range
    gs[2] = (x, bad, None, None)
    with pytest.raises(hgp.HgpError) as e:
        hgp.partition_batch(gs, ks, lanes=4)
    assert e.value.code == -1 and "graph 2" in str(e.value)


def test_batch_fused_small_levels_match_alone(hgp):
    """The batch's fused launch over the small levels (one block per graph,
    levels of up to 16384 vertices: the block's scans and compactions then
    run in several chunks) gives every graph its partition alone."""
    x, a, vw, aw = random_graph(3000, 0.004, seed=8, wmax=7, vmax=5)
    gs = [hgp.gen_rgg3d(12000, hgp.rgg_radius(12000, 12), seed=2), hgp.gen_grid2d(110, 110),
          hgp.gen_banded(9000, 32, 0.1, seed=3), (x, a, vw, aw), hgp.gen_rmat(13, 1 << 16, seed=4)]
    ks = [8, 16, 32, 5, 64]
    alone = [hgp.partition(g[0], g[1], k, g[2], g[3], seed=4)[0] for g, k in zip(gs, ks)]
    t0 = hgp.get_tuning()
    try:
        for fuse in (16384, 4096, 0):
            print(f"fuse {fuse}", flush=True)
            hgp.set_tuning(batch_fuse_max_vertices=fuse)
            parts, stats = hgp.partition_batch(gs, ks, lanes=3, seed=4)
            for i, (p, q) in enumerate(zip(parts, alone)):
                np.testing.assert_array_equal(p, q, err_msg=f"fuse {fuse} graph {i}")
                assert stats[i]["cut"] > 0 and stats[i]["max_pwgt"] > 0
    finally:
        hgp.set_tuning(**t0)
