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
