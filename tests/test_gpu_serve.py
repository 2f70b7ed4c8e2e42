"""Serving loop: jobs submitted while others run give the same partition as
alone, the slots' arenas stay within the device-memory budget, and errors
come back per job."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def jobs(hgp):
    out = []
    for i in range(10):
        if i % 2 == 0:
            g = hgp.gen_rmat(10 + i % 4, 1 << (12 + i % 4), seed=20 + i)
        else:
            n = 1500 + 300 * i
            g = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12), seed=30 + i)
        out.append((g, [8, 16, 32, 64, 128][i % 5], 100 + i))
    return out


def test_serve_equals_alone(hgp):
    js = jobs(hgp)
    with hgp.Server(slots=4, budget_bytes=8 << 30) as srv:
        tickets = [srv.submit(g[0], g[1], k, g[2], g[3], seed=s) for g, k, s in js]
        res = [srv.wait(t) for t in tickets]
        info = srv.info()
    for (g, k, s), (part, st, lat) in zip(js, res):
        p, st1 = hgp.partition(g[0], g[1], k, g[2], g[3], seed=s)
        np.testing.assert_array_equal(part, p)
        assert st["cut"] == st1["cut"]
        assert lat > 0
    assert info["completed"] == len(js)


def test_serve_budget_and_concurrent_submitters(hgp):
    js = jobs(hgp)
    budget = 150 << 20                # each job's arena is ~66-70 MB: at most two run at once
    out = {}
    with hgp.Server(slots=4, budget_bytes=budget) as srv:
        def client(sub):
            for idx in sub:
                g, k, s = js[idx]
                t = srv.submit(g[0], g[1], k, g[2], g[3], seed=s)
                out[idx] = srv.wait(t)[0]
        th = [threading.Thread(target=client, args=(list(range(c, len(js), 3)),)) for c in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        info = srv.info()
    assert info["peak_reserved"] <= budget
    for idx, (g, k, s) in enumerate(js):
        p, _ = hgp.partition(g[0], g[1], k, g[2], g[3], seed=s)
        np.testing.assert_array_equal(out[idx], p)


def test_serve_rejects(hgp):
    g = hgp.gen_grid2d(30, 30)
    with hgp.Server(slots=2, budget_bytes=1 << 20) as srv:
        with pytest.raises(hgp.HgpError) as e:
            srv.submit(g[0], g[1], 4)            # arena alone exceeds the budget
        assert e.value.code == -3
    with hgp.Server(slots=2, budget_bytes=1 << 30) as srv:
        bad = g[1].copy()
        bad[0] = 10 ** 6
        with pytest.raises(hgp.HgpError) as e:
            srv.submit(g[0], bad, 4)
        assert e.value.code == -1
        t = srv.submit(g[0], g[1], 4)
        part, st, _ = srv.wait(t)
        assert st["cut"] > 0 and part.max() == 3


def test_serve_ticket_waited_once(hgp):
    """Two threads waiting on one ticket: one gets the result, the other an
    error; a wait after the result was taken is refused too."""
    g = hgp.gen_rmat(12, 1 << 14, seed=4)
    res, errs = [], []
    with hgp.Server(slots=2, budget_bytes=1 << 30) as srv:
        t = srv.submit(g[0], g[1], 8, seed=2)

        def waiter():
            try:
                res.append(srv.wait(t))
            except hgp.HgpError as e:
                errs.append(e.code)

        th = [threading.Thread(target=waiter) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        with pytest.raises(hgp.HgpError):
            srv.wait(t)
    assert len(res) == 1 and errs == [-1]
    p, _ = hgp.partition(g[0], g[1], 8, seed=2)
    np.testing.assert_array_equal(res[0][0], p)


def _resident_cases(hgp):
    g0 = hgp.gen_rmat(13, 1 << 17, seed=41)
    n1 = 20000
    g1 = hgp.gen_rgg3d(n1, hgp.rgg_radius(n1, 12), seed=42, vw=(1, 3), ew=(1, 7))
    # (graph, k, seed): k 128, 256 and 512 share the coarsening threshold
    # 1024, so the second and third requests of graph 0 reuse the first's
    # hierarchy; another seed or a k with another threshold builds its own
    reqs = [(0, 128, 1), (0, 256, 1), (0, 512, 1), (0, 64, 1), (0, 256, 2), (1, 8, 3), (1, 8, 3), (1, 16, 3),
            (0, 300, 1)]
    return [g0, g1], reqs


def test_resident_graphs_reuse_hierarchies(hgp, orc):
    graphs, reqs = _resident_cases(hgp)
    with hgp.Server(slots=3, budget_bytes=8 << 30) as srv:
        ids = [srv.register(*g) for g in graphs]
        res = []
        for gi, k, s in reqs:           # one at a time: the hits are deterministic
            print(f"resident case graph {gi} k={k} seed={s}", flush=True)
            res.append(srv.wait(srv.submit_graph(ids[gi], k, seed=s)))
        kept = srv.kept()
        before = srv.info()["reserved"]
        srv.unregister(ids[0])
        after = srv.info()["reserved"]
        # the same requests again, all at once, on the graph left
        tickets = [srv.submit_graph(ids[1], k, seed=s) for gi, k, s in reqs if gi == 1]
        again = [srv.wait(t)[0] for t in tickets]
    assert kept["hits"] == 4 and kept["misses"] == 5, kept   # 256/1, 512/1, 8/3, 300/1 reuse
    assert kept["kept"] == 5 and kept["kept_bytes"] > 0     # graph 0: 1024/1, 512/1, 1024/2; graph 1: 64/3, 128/3
    assert after < before
    for (gi, k, s), (part, st, _) in zip(reqs, res):
        x, a, vw, aw = graphs[gi]
        p, st1 = hgp.partition(x, a, k, vw, aw, seed=s)
        np.testing.assert_array_equal(part, p, err_msg=f"graph {gi} k={k} seed={s}")
        assert st["cut"] == st1["cut"] == orc.edge_cut(x, a, aw, part)
    mine = [r[0] for (gi, _, _), r in zip(reqs, res) if gi == 1]
    for p0, p1 in zip(mine, again):
        np.testing.assert_array_equal(p0, p1)


def test_resident_hierarchy_not_kept_without_room(hgp):
    graphs, reqs = _resident_cases(hgp)
    g = graphs[0]
    # room for the graph and one job's arena, not for a kept hierarchy as well
    with hgp.Server(slots=1, budget_bytes=120 << 20) as srv:
        gid = srv.register(*g)
        parts = [srv.wait(srv.submit_graph(gid, k, seed=1))[0] for k in (128, 256)]
        kept = srv.kept()
        assert srv.info()["peak_reserved"] <= 120 << 20
    for k, part in zip((128, 256), parts):
        p, _ = hgp.partition(g[0], g[1], k, g[2], g[3], seed=1)
        np.testing.assert_array_equal(part, p)
    assert kept["kept"] <= 1


def test_resident_rejects(hgp):
    with hgp.Server(slots=1, budget_bytes=1 << 20) as srv:
        g = hgp.gen_rmat(12, 1 << 16, seed=3)
        with pytest.raises(hgp.HgpError) as e:
            srv.register(*g)                    # 1 MB cannot hold the graph
        assert e.value.code == -3
        with pytest.raises(hgp.HgpError) as e:
            srv.submit_graph(7, 8)              # unknown graph id
        assert e.value.code == -1
