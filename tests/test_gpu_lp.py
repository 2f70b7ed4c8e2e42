"""GPU parity of the older label-propagation refiners (hunyuangraph_GPU_LP.h):
the k-way rounds of hunyuangraph_k_refinement_me as a level refiner
(refiner=1) and the 2-way jetLP pass, called through the C ABI and compared
with the oracle bit for bit."""
import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu


def lp_graphs(hgp):
    out = {}
    out["grid32"] = hgp.gen_grid2d(32, 32)
    out["er300w"] = random_graph(300, 0.03, seed=3, wmax=9, vmax=4)
    out["rmat12"] = hgp.gen_rmat(12, 1 << 14, seed=5)
    out["rgg2k"] = hgp.gen_rgg3d(2000, hgp.rgg_radius(2000, 12), seed=7, vw=(1, 4), ew=(1, 16))
    # a 3000-entry hub and a 200-entry row (rows of > 64 entries keep their
    # edge weight per part in HBM), leaves, a path and isolated vertices
    edges = [(0, i) for i in range(1, 3001)] + [(3001, i) for i in range(1, 201)]
    edges += [(i, i + 1) for i in range(3002, 3100)]
    x, a, w = csr_from_edges(3200, edges, weights=[1 + (i % 5) for i in range(len(edges))])
    out["hubs"] = (x, a, np.ones(3200, np.int32), w)
    return out


@pytest.mark.parametrize("k", [2, 7, 64, 256])
def test_lp_level_matches_oracle(hgp, orc, k):
    for name, (x, a, vw, aw) in lp_graphs(hgp).items():
        n = len(x) - 1
        rng = np.random.default_rng(k * 31 + n)
        # a random start and a blocky start (few boundary vertices)
        starts = [rng.integers(0, k, size=n).astype(np.int32),
                  (np.arange(n, dtype=np.int64) * k // max(n, 1)).astype(np.int32)]
        for si, where in enumerate(starts):
            print(f"{name} k={k} start={si}", flush=True)
            tv = int(vw.sum())
            ow, ocut, (i1, i2) = orc.lp_refine(x, a, vw, aw, k, tv, where)
            gw, gcut, git = hgp.refine(x, a, vw, aw, k, tv, 0, where, 0, refiner=1)
            np.testing.assert_array_equal(gw, ow, err_msg=f"{name} k={k} start={si}")
            assert gcut == ocut == orc.edge_cut(x, a, aw, ow)
            assert git == i1 + i2


def test_lp_level_rmat_hub_rows(hgp, orc):
    """R-MAT 2^15 (hub rows of thousands of entries) at k = 256."""
    x, a, vw, aw = hgp.gen_rmat(15, 1 << 18, seed=11)
    n = len(x) - 1
    where = np.random.default_rng(2).integers(0, 256, size=n).astype(np.int32)
    tv = int(vw.sum())
    ow, ocut, (i1, i2) = orc.lp_refine(x, a, vw, aw, 256, tv, where)
    gw, gcut, git = hgp.refine(x, a, vw, aw, 256, tv, 1, where, 0, refiner=1)
    np.testing.assert_array_equal(gw, ow)
    assert gcut == ocut and git == i1 + i2


@pytest.mark.parametrize("k", [2, 16, 64])
def test_pipeline_with_older_lp(hgp, orc, k):
    for name, (x, a, vw, aw) in lp_graphs(hgp).items():
        print(f"{name} k={k}", flush=True)
        op, ost = orc.dga_partition(x, a, k, vw, aw, seed=5, refiner=1)
        gp, gst = hgp.partition(x, a, k, vw, aw, seed=5, refiner=1)
        np.testing.assert_array_equal(gp, op, err_msg=f"{name} k={k}")
        assert gst["cut"] == ost["cut"]
        assert gst["refine_iters"] == ost["refine_iters"]


def test_batch_with_older_lp(hgp):
    """hgp_partition_batch runs the older LP graph by graph (no fused small
    levels): each graph gets its partition alone."""
    gs = [hgp.gen_grid2d(64, 64), hgp.gen_rmat(12, 1 << 14, seed=5), random_graph(800, 0.01, seed=4, wmax=3)]
    ks = [8, 32, 5]
    alone = [hgp.partition(g[0], g[1], k, g[2], g[3], seed=3, refiner=1)[0] for g, k in zip(gs, ks)]
    parts, _ = hgp.partition_batch(gs, ks, lanes=2, seed=3, refiner=1)
    for p, q in zip(parts, alone):
        np.testing.assert_array_equal(p, q)


def test_lp2_matches_oracle(hgp, orc):
    for name, (x, a, vw, aw) in lp_graphs(hgp).items():
        n = len(x) - 1
        rng = np.random.default_rng(n)
        starts = {
            "random": rng.integers(0, 2, size=n).astype(np.int32),
            "halves": (np.arange(n) >= n // 2).astype(np.int32),
            "skewed": (rng.random(n) < 0.1).astype(np.int32),
            "all0": np.zeros(n, np.int32),
        }
        for sname, where in starts.items():
            print(f"{name} {sname}", flush=True)
            ow, oran = orc.lp2_refine(x, a, vw, aw, where)
            gw, gran = hgp.refine_2way_lp(x, a, vw, aw, where)
            np.testing.assert_array_equal(gw, ow, err_msg=f"{name} {sname}")
            assert gran == oran


def test_lp2_large_rmat(hgp, orc):
    x, a, vw, aw = hgp.gen_rmat(16, 1 << 19, seed=3)
    n = len(x) - 1
    where = (np.random.default_rng(5).random(n) < 0.3).astype(np.int32)
    ow, oran = orc.lp2_refine(x, a, vw, aw, where)
    gw, gran = hgp.refine_2way_lp(x, a, vw, aw, where)
    np.testing.assert_array_equal(gw, ow)
    assert gran == oran



@pytest.mark.parametrize("k", [1024, 4096])
def test_lp_level_many_parts(hgp, orc, k):
    """k of 1024 and 4096 (LDS rows of k ints per wave, k 8-byte keys per
    block), on a 6000-vertex weighted RGG and on graphs with fewer vertices
    than parts."""
    cases = [hgp.gen_rgg3d(6000, hgp.rgg_radius(6000, 12), seed=k, vw=(1, 3), ew=(1, 9))]
    x, a, w = csr_from_edges(5, [(0, 1), (1, 2), (3, 4)])
    cases.append((x, a, np.ones(5, np.int32), w))
    for ci, (x, a, vw, aw) in enumerate(cases):
        n = len(x) - 1
        where = np.random.default_rng(k + ci).integers(0, k, size=n).astype(np.int32)
        tv = int(vw.sum())
        print(f"case {ci} n={n} k={k}", flush=True)
        ow, ocut, (i1, i2) = orc.lp_refine(x, a, vw, aw, k, tv, where)
        gw, gcut, git = hgp.refine(x, a, vw, aw, k, tv, 1, where, 0, refiner=1)
        np.testing.assert_array_equal(gw, ow)
        assert gcut == ocut and git == i1 + i2
