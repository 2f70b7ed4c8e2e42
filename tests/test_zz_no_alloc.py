"""The evaluation entry point that allocates nothing
(hgp_evaluate_device_noalloc) against the arena one (hgp_evaluate_device),
through the C ABI on device buffers: two calls in a row, the second with a
smaller graph and fewer parts, must give the same four numbers as the arena
entry, and those must equal the exact host counts."""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _host(x, a, aw, vw, where, k):
    n = len(x) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(x))
    ok = (a >= 0) & (a < n)
    r, c, w = rows[ok], a[ok], aw[ok]
    cut = int(w[where[r] != where[c]].astype(np.int64).sum()) // 2
    good = (where >= 0) & (where < k)
    pw = np.bincount(where[good], weights=vw[good].astype(np.float64), minlength=k)
    return [cut, int(pw.max()), int((~good).sum()), int((~ok).sum())]


def _both(hgp, torch, x, a, aw, vw, where, k):
    d = [torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for v in (x, a, aw, vw, where)]
    n = len(x) - 1
    out_a = torch.full((4,), -7, dtype=torch.int64, device="cuda:0")
    out_f = torch.full((4,), -7, dtype=torch.int64, device="cuda:0")
    print(f"no-alloc case: n={n} m={len(a)} k={k}", flush=True)
    hgp.evaluate_device(n, *(t.data_ptr() for t in d), k, out_a.data_ptr())
    hgp.evaluate_device_noalloc(n, *(t.data_ptr() for t in d), k, out_f.data_ptr())
    return out_a.cpu().tolist(), out_f.cpu().tolist()


def test_noalloc_matches_arena_entry(hgp):
    import torch
    assert torch.cuda.is_available()
    # first call: an R-MAT with hub rows and the largest k
    x, a, vw, aw = hgp.gen_rmat(15, 1 << 18, seed=8)
    where = np.random.default_rng(1).integers(0, 4096, size=len(x) - 1).astype(np.int32)
    ra, rf = _both(hgp, torch, x, a, aw, vw, where, 4096)
    assert rf == ra == _host(x, a, aw, vw, where, 4096)
    # second call, smaller graph and fewer parts, with out-of-range ids that
    # are counted: the scratch keeps the first call's sums unless cleared
    x, a, vw, aw = random_graph(300, 0.03, seed=4, wmax=9, vmax=5)
    where = np.random.default_rng(2).integers(0, 7, size=len(x) - 1).astype(np.int32)
    where[[5, 17]] = [7, -2]
    a = a.copy()
    a[[0, 9]] = [len(x) - 1, -4]
    ra, rf = _both(hgp, torch, x, a, aw, vw, where, 7)
    assert rf == ra == _host(x, a, aw, vw, where, 7)
    assert rf[2:] == [2, 2]


def test_noalloc_rejects_before_launch(hgp):
    import torch
    x, a, vw, aw = random_graph(50, 0.1, seed=3)
    d = [torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for v in (x, a, aw, vw, np.zeros(50, np.int32))]
    out = torch.full((4,), -7, dtype=torch.int64, device="cuda:0")
    with pytest.raises(hgp.HgpError) as e:
        hgp.evaluate_device_noalloc(50, *(t.data_ptr() for t in d), 4097, out.data_ptr())
    assert e.value.code == -4
    with pytest.raises(hgp.HgpError) as e:
        hgp.evaluate_device_noalloc(50, *(t.data_ptr() for t in d), 0, out.data_ptr())
    assert e.value.code == -1
    with pytest.raises(hgp.HgpError) as e:
        hgp.evaluate_device_noalloc(50, *(t.data_ptr() for t in d), 4, out.data_ptr(), device=99)
    assert e.value.code == -1
    assert out.cpu().tolist() == [-7] * 4
