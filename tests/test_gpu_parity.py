"""GPU parity: every stage of the HIP path, called through the C ABI, must equal
the oracle bit for bit (integer arithmetic throughout)."""
import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu


def graphs(hgp):
    """Small graphs covering the shapes the pipeline meets."""
    out = {}
    out["grid32"] = hgp.gen_grid2d(32, 32)
    x, a, vw, aw = random_graph(300, 0.03, seed=3, wmax=9, vmax=4)
    out["er300w"] = (x, a, vw, aw)
    out["rmat12"] = hgp.gen_rmat(12, 1 << 14, seed=5)
    out["rgg2k"] = hgp.gen_rgg3d(2000, hgp.rgg_radius(2000, 12), seed=7, vw=(1, 4), ew=(1, 16))
    out["band1k"] = hgp.gen_banded(1000, 8, 0.1, seed=9)
    # star with leaves plus isolated vertices: exercises leaf/isolate/forced
    edges = [(0, i) for i in range(1, 40)] + [(40, 41), (41, 42)]
    x, a, w = csr_from_edges(60, edges)
    out["star_iso"] = (x, a, np.ones(60, np.int32), w)
    # hub with 300 leaves: leaf pairing over several 64-entry chunks
    edges = [(0, i) for i in range(1, 301)] + [(301, 302)]
    x, a, w = csr_from_edges(303, edges)
    out["star300"] = (x, a, np.ones(303, np.int32), w)
    # comb: hub 0 -- teeth 1..200, tooth i -- heavy private vertex 200+i; with
    # k=64 the heavy vertices cannot be matched, so the relative pass pairs teeth
    edges = [(0, i) for i in range(1, 201)] + [(i, 200 + i) for i in range(1, 201)]
    x, a, w = csr_from_edges(401, edges)
    vw = np.ones(401, np.int32)
    vw[201:] = 50
    out["comb"] = (x, a, vw, w)
    return out


@pytest.mark.parametrize("n", [1, 2, 255, 2047, 2048, 2049, 100003, (1 << 20) + 7])
def test_scan_i32(hgp, orc, n):
    rng = np.random.default_rng(n)
    a = rng.integers(-1000, 1000, size=n).astype(np.int32)
    np.testing.assert_array_equal(hgp.scan_i32(a), orc.scan_i32(a))
    np.testing.assert_array_equal(hgp.scan_i32(a), np.cumsum(a, dtype=np.int64).astype(np.int32))


def test_scan_i32_wraps(hgp, orc):
    a = np.full(5000, 2**30, dtype=np.int32)
    np.testing.assert_array_equal(hgp.scan_i32(a), orc.scan_i32(a))


@pytest.mark.parametrize("n", [1, 3000, 4194305])
def test_scan_i64(hgp, orc, n):
    rng = np.random.default_rng(n + 1)
    a = rng.integers(-(2**40), 2**40, size=n).astype(np.int64)
    np.testing.assert_array_equal(hgp.scan_i64(a), orc.scan_i64(a))
    np.testing.assert_array_equal(hgp.scan_i64(a), np.cumsum(a))


def test_part_weights_and_cut(hgp, orc):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        n = len(x) - 1
        for k in (2, 7, 64):
            where = np.random.default_rng(k).integers(0, k, size=n).astype(np.int32)
            np.testing.assert_array_equal(hgp.part_weights(vw, where, k), orc.part_weights(vw, where, k), err_msg=name)
            assert hgp.edge_cut(x, a, aw, where) == orc.edge_cut(x, a, aw, where), name


@pytest.mark.parametrize("level", [0, 3])
def test_match_bit_exact(hgp, orc, level):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        tv = int(vw.sum())
        for k in (2, 64):
            mv = orc.maxvwgt(tv, orc.coarsen_threshold(k))
            m_g, c_g = hgp.match(x, a, vw, aw, mv, 11, level, 8)
            m_o, c_o = orc.dga_match(x, a, vw, aw, mv, 11, level, 8)
            np.testing.assert_array_equal(m_g, m_o, err_msg=f"{name} k={k}")
            assert c_g == c_o


def test_contract_bit_exact(hgp, orc):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        tv = int(vw.sum())
        mv = orc.maxvwgt(tv, orc.coarsen_threshold(4))
        mt, _ = orc.dga_match(x, a, vw, aw, mv, 1, 0, 8)
        got = hgp.contract(x, a, aw, vw, mt)
        exp = orc.contract(x, a, aw, vw, mt)
        for g_arr, e_arr, what in zip(got, exp, ["match", "cmap", "cxadj", "cadjncy", "cadjwgt", "cvwgt"]):
            np.testing.assert_array_equal(g_arr, e_arr, err_msg=f"{name}: {what}")


def test_contract_random_matchings(hgp, orc):
    """Contraction of arbitrary (non-adjacent too) fixed matchings."""
    x, a, vw, aw = random_graph(500, 0.02, seed=21, wmax=5, vmax=3)
    n = len(x) - 1
    rng = np.random.default_rng(4)
    for trial in range(3):
        perm = rng.permutation(n)
        mt = -np.ones(n, np.int32)
        npair = n // 3
        for i in range(npair):
            u, v = perm[2 * i], perm[2 * i + 1]
            mt[u], mt[v] = v, u
        got = hgp.contract(x, a, aw, vw, mt)
        exp = orc.contract(x, a, aw, vw, mt)
        for g_arr, e_arr in zip(got, exp):
            np.testing.assert_array_equal(g_arr, e_arr)


@pytest.mark.parametrize("k", [2, 3, 5, 8, 16])
def test_initpart_bit_exact(hgp, orc, k):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        if len(x) - 1 > 1200:
            continue
        w_g = hgp.initpart(x, a, vw, aw, k, seed=3, trials=64)
        w_o = orc.dga_initpart(x, a, vw, aw, k, seed=3, trials=64)
        np.testing.assert_array_equal(w_g, w_o, err_msg=f"{name} k={k}")


def test_initpart_exhaustive_and_sampled(hgp, orc):
    x, a, vw, aw = random_graph(200, 0.05, seed=8, wmax=4, vmax=3)
    for trials in (1024, 37):     # n <= trials: exhaustive; n > trials: sampled starts
        w_g = hgp.initpart(x, a, vw, aw, 6, seed=9, trials=trials)
        w_o = orc.dga_initpart(x, a, vw, aw, 6, seed=9, trials=trials)
        np.testing.assert_array_equal(w_g, w_o)


@pytest.mark.parametrize("k", [4, 8, 64])
@pytest.mark.parametrize("finest", [0, 1])
def test_refine_bit_exact(hgp, orc, k, finest):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        n = len(x) - 1
        tv = int(vw.sum())
        where = np.random.default_rng(k).integers(0, k, size=n).astype(np.int32)
        cut = orc.edge_cut(x, a, aw, where)
        w_g, c_g, i_g = hgp.refine(x, a, vw, aw, k, tv, finest, where, cut)
        w_o, c_o, i_o = orc.dga_refine(x, a, vw, aw, k, tv, finest, where, cut)
        np.testing.assert_array_equal(w_g, w_o, err_msg=f"{name} k={k}")
        assert (c_g, i_g) == (c_o, i_o), name
        assert c_g == orc.edge_cut(x, a, aw, w_g)


def _bucket_fifty_case(n, npairs, seed):
    """Random graph plus isolated pairs joined by an edge of weight 20000,
    placed with ids spread over the graph, in three overloaded parts: under
    weak rebalancing the pairs' vertices (no target part, gain -20000) fall
    in gain bucket 50, beside the next part's first-bucket candidates."""
    rng = np.random.default_rng(seed)
    pairs = rng.choice(n, 2 * npairs, replace=False)
    normal = np.setdiff1d(np.arange(n), pairs)
    edges = {}
    for _ in range(int(2.5 * n)):
        u, v = rng.choice(normal, 2, replace=False)
        edges.setdefault((int(min(u, v)), int(max(u, v))), int(rng.integers(1, 4)))
    for i in range(0, 2 * npairs, 2):
        edges[(int(min(pairs[i], pairs[i + 1])), int(max(pairs[i], pairs[i + 1])))] = 20000
    x, a, aw = csr_from_edges(n, list(edges), list(edges.values()))
    k = 8
    where = rng.integers(0, k, size=n).astype(np.int32)
    over = rng.random(n) < 0.45
    where[over] = rng.integers(0, 3, size=int(over.sum()))
    for i in range(0, 2 * npairs, 2):
        where[pairs[i]] = where[pairs[i + 1]] = rng.integers(0, 3)
    return x, a, np.ones(n, np.int32), aw, k, where


@pytest.mark.parametrize("n,npairs", [(1600, 60), (400, 20)])
def test_refine_weak_rebalance_bucket_fifty(hgp, orc, n, npairs):
    """Jet refinement from a partition whose first iterations are weak
    rebalancing with gain-bucket-50 candidates, on the row-class kernels
    (1600 vertices) and the single-block loop (400), against the oracle."""
    assert orc.gain_bucket(-20000, 1) == 50
    x, a, vw, aw, k, where = _bucket_fifty_case(n, npairs, 5)
    tv = int(vw.sum())
    cut = orc.edge_cut(x, a, aw, where)
    for finest in (0, 1):
        print(f"bucket-50 case n={n} finest={finest}", flush=True)
        w_g, c_g, i_g = hgp.refine(x, a, vw, aw, k, tv, finest, where, cut)
        w_o, c_o, i_o = orc.dga_refine(x, a, vw, aw, k, tv, finest, where, cut)
        np.testing.assert_array_equal(w_g, w_o)
        assert (c_g, i_g) == (c_o, i_o)
        assert c_g == orc.edge_cut(x, a, aw, w_g)


@pytest.mark.parametrize("k", [2, 8, 32])
def test_pipeline_bit_exact(hgp, orc, k):
    for name, (x, a, vw, aw) in graphs(hgp).items():
        p_g, st_g = hgp.partition(x, a, k, vw, aw, seed=5)
        p_o, st_o = orc.dga_partition(x, a, k, vw, aw, seed=5)
        np.testing.assert_array_equal(p_g, p_o, err_msg=f"{name} k={k}")
        assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)
        assert st_g["levels"] == st_o["levels"]
        assert st_g["level_iters"] == st_o["level_iters"]


def test_pipeline_grid256_k8(hgp, orc):
    """The plain small config: 256x256 5-point grid, k=8, ub 1.03, seed 1."""
    x, a, vw, aw = hgp.gen_grid2d(256, 256)
    p_g, st = hgp.partition(x, a, 8, vw, aw, seed=1)
    p_o, st_o = orc.dga_partition(x, a, 8, vw, aw, seed=1)
    np.testing.assert_array_equal(p_g, p_o)
    assert st["cut"] == orc.edge_cut(x, a, aw, p_g)
    assert orc.balanced(vw, p_g, 8, 103, 100)


def test_dropin_matches_versioned(hgp):
    x, a, vw, aw = hgp.gen_grid2d(40, 50)
    n = len(x) - 1
    p1 = hgp.PartitionGraph(n, x.astype(np.int32), a, vw, aw, 6)
    p2, _ = hgp.partition(x, a, 6, vw, aw, seed=1)
    np.testing.assert_array_equal(p1, p2)


def test_deterministic(hgp):
    x, a, vw, aw = hgp.gen_rmat(13, 1 << 15, seed=2)
    p1, s1 = hgp.partition(x, a, 16, seed=7)
    p2, s2 = hgp.partition(x, a, 16, seed=7)
    np.testing.assert_array_equal(p1, p2)
    assert s1["cut"] == s2["cut"]


def test_edge_cases(hgp, orc):
    # single vertex, no edges
    x = np.array([0, 0], np.int64)
    a = np.zeros(0, np.int32)
    p, st = hgp.partition(x, a, 4)
    assert p.tolist() == [0] or p[0] in range(4)
    # k > n
    x, a, w = csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    p, _ = hgp.partition(x, a, 8)
    q, _ = orc.dga_partition(x, a, 8)
    np.testing.assert_array_equal(p, q)
    # k = 1
    p, _ = hgp.partition(x, a, 1)
    assert (p == 0).all()
    # edgeless graph
    x = np.zeros(101, np.int64)
    p, _ = hgp.partition(x, np.zeros(0, np.int32), 4)
    q, _ = orc.dga_partition(x, np.zeros(0, np.int32), 4)
    np.testing.assert_array_equal(p, q)


def test_invalid_inputs_raise(hgp):
    x, a, w = csr_from_edges(4, [(0, 1), (1, 2)])
    bad = a.copy()
    bad[0] = 9
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, bad, 2)
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, a, 0)
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, a, 2, vwgt=np.array([1, 0, 1, 1], np.int32))


def _hub_graph(hgp):
    """R-MAT with hubs of > 8k entries: its coarse rows reach the register,
    wave-register and both LDS contraction classes (<= 4096 and <= 16384
    entries), and its hub rows are cut into several segments."""
    x, a, vw, aw = hgp.gen_rmat(16, 1 << 20, seed=12)
    deg = np.diff(x)
    assert deg.max() > 2 * 4096
    return x, a, vw, aw


def test_contract_hub_rows(hgp, orc):
    x, a, vw, aw = _hub_graph(hgp)
    tv = int(vw.sum())
    for lvl, k in ((0, 64), (2, 8)):
        mv = orc.maxvwgt(tv, orc.coarsen_threshold(k))
        m_g, _ = hgp.match(x, a, vw, aw, mv, 3, lvl, 8)
        m_o, _ = orc.dga_match(x, a, vw, aw, mv, 3, lvl, 8)
        np.testing.assert_array_equal(m_g, m_o)
        got = hgp.contract(x, a, aw, vw, m_o)
        exp = orc.contract(x, a, aw, vw, m_o)
        for g_arr, e_arr, what in zip(got, exp, ["match", "cmap", "cxadj", "cadjncy", "cadjwgt", "cvwgt"]):
            np.testing.assert_array_equal(g_arr, e_arr, err_msg=what)
    # the same rows with weights: merged weights are sums
    rng = np.random.default_rng(1)
    aw2 = np.zeros_like(aw)
    n = len(x) - 1
    rows = np.repeat(np.arange(n), np.diff(x))
    key = np.minimum(rows, a).astype(np.int64) * n + np.maximum(rows, a)
    aw2 = (key % 13 + 1).astype(np.int32)     # symmetric per undirected edge
    mt, _ = orc.dga_match(x, a, vw, aw2, orc.maxvwgt(tv, 512), 4, 0, 8)
    for g_arr, e_arr in zip(hgp.contract(x, a, aw2, vw, mt), orc.contract(x, a, aw2, vw, mt)):
        np.testing.assert_array_equal(g_arr, e_arr)


def test_contract_wide_coarse_ids(hgp, orc):
    """More than 2^16 coarse vertices: the block-sorted rows (513..16384
    entries) take three 8-bit digit passes; a row of > 16384 entries takes
    the coarse-id bins."""
    x, a, vw, aw = hgp.gen_rmat(18, 1 << 22, seed=21)
    deg = np.diff(x)
    assert ((deg > 1024) & (deg <= 4096)).sum() > 20
    assert ((deg > 4096) & (deg <= 16384)).sum() > 20 and (deg > 16384).sum() >= 1
    tv = int(vw.sum())
    mt, _ = orc.dga_match(x, a, vw, aw, orc.maxvwgt(tv, orc.coarsen_threshold(64)), 5, 0, 8)
    got = hgp.contract(x, a, aw, vw, mt)
    exp = orc.contract(x, a, aw, vw, mt)
    assert len(exp[2]) - 1 > (1 << 16)
    for g_arr, e_arr, what in zip(got, exp, ["match", "cmap", "cxadj", "cadjncy", "cadjwgt", "cvwgt"]):
        np.testing.assert_array_equal(g_arr, e_arr, err_msg=what)


def test_pipeline_quarter_rows(hgp, orc):
    """About 330k rows of 33..128 entries: the wave-per-row kernels take 41
    rows per wave chunk and walk them four at a time, one per 16-lane quarter,
    so the last round of a chunk often has fewer rows than quarters and its
    rows sit in lanes of the idle quarters."""
    n = 400000
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 40), seed=33, vw=(1, 3), ew=(1, 7))
    deg = np.diff(x)
    assert ((deg > 32) & (deg <= 128)).sum() > 16 * 8192
    p_g, st_g = hgp.partition(x, a, 32, vw, aw, seed=2)
    p_o, st_o = orc.dga_partition(x, a, 32, vw, aw, seed=2)
    np.testing.assert_array_equal(p_g, p_o)
    assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)


@pytest.mark.parametrize("hub_track", [0, 1])
def test_pipeline_hub_graph(hgp, orc, hub_track):
    """Long rows rebuild their connectivity in every selection (hub_track 0)
    or keep it current through the move accounting (1): the same partition
    as the oracle either way."""
    x, a, vw, aw = _hub_graph(hgp)
    t0 = hgp.get_tuning()
    try:
        hgp.set_tuning(hub_track=hub_track)
        p_g, st_g = hgp.partition(x, a, 64, vw, aw, seed=2)
    finally:
        hgp.set_tuning(**t0)
    p_o, st_o = orc.dga_partition(x, a, 64, vw, aw, seed=2)
    np.testing.assert_array_equal(p_g, p_o)
    assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)
    assert orc.balanced(vw, p_g, 64, 103, 100)


def test_golden_fixtures_on_gpu(hgp):
    """The committed fixtures (tests/golden, derived from the definitions)
    hold for the GPU path too."""
    import json
    import os
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for g in json.load(open(os.path.join(gdir, "contract.json"))):
        got = hgp.contract(np.array(g["xadj"], np.int64), np.array(g["adjncy"], np.int32),
                           np.array(g["adjwgt"], np.int32), np.array(g["vwgt"], np.int32),
                           np.array(g["match"], np.int32))
        for arr, key in zip(got[1:], ["cmap", "cxadj", "cadjncy", "cadjwgt", "cvwgt"]):
            assert arr.tolist() == g[key], (g["name"], key)
    for g in json.load(open(os.path.join(gdir, "known_cut.json"))):
        p, st = hgp.partition(np.array(g["xadj"], np.int64), np.array(g["adjncy"], np.int32), g["k"])
        assert st["cut"] == g["optimal_cut"]


def test_pipeline_weighted_rgg_k64(hgp, orc):
    """The 3-D RGG config at 2^18 points (vertex weights 1..4, edge weights 1..16, k = 64)."""
    n = 1 << 18
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12.0), seed=1, vw=(1, 4), ew=(1, 16))
    p_g, st_g = hgp.partition(x, a, 64, vw, aw, seed=1)
    p_o, st_o = orc.dga_partition(x, a, 64, vw, aw, seed=1)
    np.testing.assert_array_equal(p_g, p_o)
    assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)
    assert orc.balanced(vw, p_g, 64, 103, 100)


def test_large_rmat_properties(hgp, orc):
    """R-MAT 2^20 / 2^24 at k = 256 (too large for the oracle in a test): the
    part ids, the exact balance bound, the cut against an exact recount and a
    repeat run that must be identical."""
    x, a, vw, aw = hgp.gen_rmat(20, 1 << 24, seed=3)
    p1, s1 = hgp.partition(x, a, 256, vw, aw, seed=4)
    assert p1.min() >= 0 and p1.max() < 256
    assert orc.balanced(vw, p1, 256, 103, 100)
    assert s1["cut"] == orc.edge_cut(x, a, aw, p1)
    p2, s2 = hgp.partition(x, a, 256, vw, aw, seed=4)
    np.testing.assert_array_equal(p1, p2)


def test_files_end_to_end(hgp, orc, tmp_path):
    """Graph file -> partition -> partition file, as the reference's command
    line does it (read, partition, write <graph>.part.<k>)."""
    import os
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
    cases = [("metis_manual.graph", 2), ("weighted_011.graph", 3), ("symmetric_pattern.mtx", 2)]
    x, a, vw, aw = hgp.gen_rgg3d(5000, hgp.rgg_radius(5000, 12), seed=21, vw=(1, 3), ew=(1, 9))
    big = str(tmp_path / "rgg5k.graph")
    hgp.write_graph(big, x, a, vw, aw)
    cases.append((big, 16))
    for name, k in cases:
        path = os.path.join(gdir, name)
        x, a, vw, aw = hgp.read_mtx(path) if name.endswith(".mtx") else hgp.read_graph(path)
        p_g, st = hgp.partition(x, a, k, vw, aw, seed=1)
        p_o, _ = orc.dga_partition(x, a, k, vw, aw, seed=1)
        np.testing.assert_array_equal(p_g, p_o, err_msg=name)
        out = hgp.write_partition(str(tmp_path / os.path.basename(name)), p_g, k)
        assert np.loadtxt(out, dtype=np.int32, ndmin=1).tolist() == p_g.tolist()


@pytest.mark.parametrize("n,k", [(400, 200), (400, 300), (20000, 300), (3000, 64)])
def test_pipeline_level_paths(hgp, orc, n, k):
    """Levels of <= 512 vertices run the whole Jet loop in one block when
    k <= 256 and the multi-kernel loop otherwise; both equal the oracle."""
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 10.0), seed=n + k, vw=(1, 3), ew=(1, 9))
    p_g, st_g = hgp.partition(x, a, k, vw, aw, seed=2)
    p_o, st_o = orc.dga_partition(x, a, k, vw, aw, seed=2)
    np.testing.assert_array_equal(p_g, p_o, err_msg=f"n={n} k={k}")
    assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)
    assert st_g["level_iters"] == st_o["level_iters"]


def test_mostly_isolated_graph(hgp, orc):
    """40,000 vertices, 97.5 % of them isolated, beside a 1000-vertex random
    graph: after one level the coarse graph has fewer than n/2 entries, so
    coarsening stops with ~20,000 vertices and the initial partition's trials
    run on that graph from global scratch sized by it.  The result equals the
    oracle (16 trials keep the oracle's run to seconds)."""
    xs, as_, vs, ws = random_graph(1000, 0.008, seed=31, wmax=5, vmax=3)
    n = 40000
    x = np.concatenate([xs, np.full(n - 1000, xs[-1], dtype=np.int64)])
    vw = np.concatenate([vs, np.ones(n - 1000, np.int32)])
    for k in (4, 37):
        p_g, st_g = hgp.partition(x, as_, k, vw, ws, seed=3, trials=16)
        p_o, st_o = orc.dga_partition(x, as_, k, vw, ws, seed=3, trials=16)
        np.testing.assert_array_equal(p_g, p_o, err_msg=f"k={k}")
        assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, as_, ws, p_g)
        assert orc.balanced(vw, p_g, k)
        assert st_g["levels"] == st_o["levels"] == 1 and st_g["coarsest_n"] > 20000
