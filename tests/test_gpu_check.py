"""Device-side CSR checks (hgp_check_graph_device, also run by
hgp_partition_device) and the evaluation of a device partition
(hgp_evaluate_device), through the C ABI on device buffers.

The malformed graphs below are rejected by the row pass or the entry pass;
in every case a kernel only compares ids with n and never gathers through
a bad one (reference: check_graph, hunyuangraph_GPU_match.h:1718-1800)."""
import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu


def _dev(torch, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrs]


def _ptr(t):
    return t.data_ptr() if t is not None and t.numel() else 0


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _base():
    x, a, w = csr_from_edges(6, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (1, 4)])
    vw = np.array([1, 2, 1, 3, 1, 1], np.int32)
    return x, a, vw, w


def _expect_check(hgp, torch, x, a, vw, aw, symmetric, msg):
    d = _dev(torch, x, a, vw, aw)
    n = len(x) - 1
    print(f"check case: {msg}", flush=True)
    if msg is None:
        hgp.check_graph_device(n, *(t.data_ptr() for t in d[:2]), _ptr(d[2]), _ptr(d[3]), symmetric=symmetric)
        return
    with pytest.raises(hgp.HgpError) as e:
        hgp.check_graph_device(n, *(t.data_ptr() for t in d[:2]), _ptr(d[2]), _ptr(d[3]), symmetric=symmetric)
    assert e.value.code == -1
    assert msg in str(e.value), str(e.value)
    # the partition entry refuses the same graph before any pipeline kernel runs
    part = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    with pytest.raises(hgp.HgpError) as e2:
        hgp.partition_device(n, d[0].data_ptr(), d[1].data_ptr(), _ptr(d[2]), _ptr(d[3]), 2, part.data_ptr(),
                             validate=1 if symmetric else 0)
    assert e2.value.code == -1 and msg in str(e2.value)


def test_check_accepts_valid(hgp, torch_mod):
    x, a, vw, aw = _base()
    _expect_check(hgp, torch_mod, x, a, vw, aw, True, None)
    x, a, vw, aw = hgp.gen_rmat(14, 1 << 17, seed=3)
    _expect_check(hgp, torch_mod, x, a, vw, aw, True, None)


def test_check_row_faults(hgp, torch_mod):
    x, a, vw, aw = _base()
    bad = x.copy()
    bad[0] = 1
    _expect_check(hgp, torch_mod, bad, a, vw, aw, False, "xadj[0] != 0")
    bad = x.copy()
    bad[3] = bad[4] + 1          # row 3 ends before it starts
    _expect_check(hgp, torch_mod, bad, a, vw, aw, False, "not monotone at vertex 3")
    v0 = vw.copy()
    v0[4] = 0                    # zero vertex weights are refused on every entry point
    _expect_check(hgp, torch_mod, x, a, v0, aw, False, "vertex weight < 1 at vertex 4")


def test_check_entry_faults(hgp, torch_mod):
    x, a, vw, aw = _base()
    bad = a.copy()
    bad[x[2]] = 6                # neighbour id == n
    _expect_check(hgp, torch_mod, x, bad, vw, aw, False, "neighbour id out of range in row 2")
    bad = a.copy()
    bad[x[5]] = -3
    _expect_check(hgp, torch_mod, x, bad, vw, aw, False, "neighbour id out of range in row 5")
    bad = a.copy()
    bad[x[1]] = 1                # self-loop
    _expect_check(hgp, torch_mod, x, bad, vw, aw, False, "self-loop at vertex 1")
    w0 = aw.copy()
    w0[x[3] + 1] = 0
    _expect_check(hgp, torch_mod, x, a, vw, w0, False, "edge weight < 1 in row 3")


def test_check_symmetry(hgp, torch_mod):
    x, a, vw, aw = _base()
    w1 = aw.copy()
    w1[0] = 7                    # (0,1) has weight 7 in row 0 and 1 in row 1
    _expect_check(hgp, torch_mod, x, a, vw, w1, False, None)   # not asked: accepted
    _expect_check(hgp, torch_mod, x, a, vw, w1, True, "graph is not symmetric")
    # one-directional edge: row 0 lists 3 instead of 1
    a2 = a.copy()
    a2[x[0]] = 3
    _expect_check(hgp, torch_mod, x, a2, vw, aw, True, "graph is not symmetric")


def test_check_hub_rows_and_empty_rows(hgp, torch_mod):
    """Rows far longer than a thread's 16 entries and runs of empty rows: a
    fault deep in a hub row and one right after empty rows are found at
    their rows."""
    n = 5000
    edges = [(0, i) for i in range(1, 3000)] + [(4000, 4001), (4998, 4999)]
    x, a, w = csr_from_edges(n, edges)
    vw = np.ones(n, np.int32)
    _expect_check(hgp, torch_mod, x, a, vw, w, True, None)
    bad = a.copy()
    bad[x[0] + 2500] = n + 11
    _expect_check(hgp, torch_mod, x, bad, vw, w, False, "neighbour id out of range in row 0")
    bad = a.copy()
    bad[x[4998]] = 4998
    _expect_check(hgp, torch_mod, x, bad, vw, w, False, "self-loop at vertex 4998")


def test_evaluate_matches_oracle(hgp, orc, torch_mod):
    torch = torch_mod
    cases = [random_graph(400, 0.02, seed=4, wmax=9, vmax=5), hgp.gen_rmat(15, 1 << 18, seed=8),
             hgp.gen_grid2d(50, 70)]
    for ci, (x, a, vw, aw) in enumerate(cases):
        n = len(x) - 1
        for k in (1, 7, 256, 4096):
            where = np.random.default_rng(k + ci).integers(0, k, size=n).astype(np.int32)
            d = _dev(torch, x, a, aw, vw, where)
            out = torch.full((4,), -1, dtype=torch.int64, device="cuda:0")
            hgp.evaluate_device(n, *(t.data_ptr() for t in d), k, out.data_ptr())
            cut, mx, bp, bn = out.cpu().tolist()
            assert cut == orc.edge_cut(x, a, aw, where), (ci, k)
            assert mx == int(orc.part_weights(vw, where, k).max()), (ci, k)
            assert bp == 0 and bn == 0
    # ids outside the ranges are counted, not gathered
    x, a, vw, aw = cases[0]
    n = len(x) - 1
    where = np.zeros(n, np.int32)
    where[[3, 9]] = [5, -1]
    bad = a.copy()
    bad[[0, 10, 20]] = [n, -5, n + 100]
    d = _dev(torch, x, bad, aw, vw, where)
    out = torch.zeros(4, dtype=torch.int64, device="cuda:0")
    hgp.evaluate_device(n, *(t.data_ptr() for t in d), 4, out.data_ptr())
    assert out.cpu().tolist()[2:] == [2, 3]


def test_evaluate_rejects(hgp, torch_mod):
    x, a, vw, aw = _base()
    d = _dev(torch_mod, x, a, aw, vw, np.zeros(6, np.int32))
    out = torch_mod.zeros(4, dtype=torch_mod.int64, device="cuda:0")
    with pytest.raises(hgp.HgpError) as e:
        hgp.evaluate_device(6, *(t.data_ptr() for t in d), 4097, out.data_ptr())
    assert e.value.code == -4
    with pytest.raises(hgp.HgpError) as e:
        hgp.evaluate_device(6, *(t.data_ptr() for t in d), 0, out.data_ptr())
    assert e.value.code == -1
