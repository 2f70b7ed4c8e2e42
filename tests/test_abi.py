"""CPU checks of the C ABI library: it loads, exports every symbol that
include/hunyuangraph.h declares, rejects bad input on the host, and its
(host-only) graph generators produce valid symmetric CSR."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, csr_from_edges

HEADER = os.path.join(ROOT, "include", "hunyuangraph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if (n.startswith("hgp_") or n.startswith("hunyuangraph_")) and n not in skip})


def test_header_declares_dropin():
    assert "hunyuangraph_PartitionGraph" in declared_functions()


def test_library_exports_every_declared_symbol(hgp):
    lib = ctypes.CDLL(hgp.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding knows every declared symbol's signature
    assert set(declared_functions()) <= set(hgp.SIGNATURES)


def test_version_and_no_compute_without_device(hgp):
    assert b"gfx950" in hgp.lib().hgp_version()
    assert hgp.device_count() >= 0


def check_csr(x, a, vw, aw):
    n = len(x) - 1
    assert x[0] == 0 and np.all(np.diff(x) >= 0) and x[-1] == len(a)
    assert a.min(initial=0) >= 0 and (len(a) == 0 or a.max() < n)
    rows = np.repeat(np.arange(n), np.diff(x))
    assert not np.any(rows == a), "self-loop"
    fwd = np.stack([rows, a, aw], 1)
    bwd = np.stack([a, rows, aw], 1)
    fwd = fwd[np.lexsort(fwd.T[::-1])]
    bwd = bwd[np.lexsort(bwd.T[::-1])]
    np.testing.assert_array_equal(fwd, bwd)   # symmetric with equal weights
    for v in range(min(n, 2000)):
        r = a[x[v]:x[v + 1]]
        assert np.all(np.diff(r) > 0), "rows sorted, no duplicates"
    assert np.all(vw > 0) and np.all(aw > 0)


def test_grid_generator(hgp):
    x, a, vw, aw = hgp.gen_grid2d(256, 256)
    assert len(x) - 1 == 65536 and len(a) == 2 * 130560
    check_csr(x, a, vw, aw)


def test_rgg_generator(hgp):
    n = 20000
    x, a, vw, aw = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12), seed=1, vw=(1, 4), ew=(1, 16))
    check_csr(x, a, vw, aw)
    assert 10.0 < len(a) / n < 12.5          # mean degree ~12 minus boundary loss
    assert vw.min() == 1 and vw.max() == 4 and aw.min() == 1 and aw.max() == 16
    x2, a2, vw2, aw2 = hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12), seed=1, vw=(1, 4), ew=(1, 16))
    np.testing.assert_array_equal(a, a2)
    np.testing.assert_array_equal(aw, aw2)


def test_rmat_generator(hgp):
    x, a, vw, aw = hgp.gen_rmat(12, 20000, seed=3)
    assert len(x) - 1 == 4096 and len(a) == 40000
    check_csr(x, a, vw, aw)
    deg = np.diff(x)
    assert deg.max() > 20 * deg.mean()       # power law hubs


def test_banded_generator(hgp):
    x, a, vw, aw = hgp.gen_banded(2000, 32, 0.1, seed=4)
    check_csr(x, a, vw, aw)
    rows = np.repeat(np.arange(2000), np.diff(x))
    off = np.abs(rows - a) > 32
    band = (~off).sum() // 2
    assert abs(off.sum() // 2 - round(0.1 * band)) == 0


def test_host_validation_rejects_bad_csr(hgp):
    x, a, w = csr_from_edges(4, [(0, 1), (1, 2)])
    with pytest.raises(hgp.HgpError) as e:
        hgp.partition(x, np.array([1, 0, 9, 1], np.int32), 2)
    assert e.value.code == -1
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, a, 0)
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, a, 2, vwgt=np.array([1, 0, 1, 1], np.int32))
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, np.array([0, 0, 2, 1], np.int32), 2)        # self-loop at vertex 0
    with pytest.raises(hgp.HgpError):
        hgp.partition(x, a, 5000)                                  # k beyond the supported range
    vw = np.ones(4, np.int32)
    with pytest.raises(hgp.HgpError) as e:
        hgp.partition(x, a, 2, refiner=2)                          # no such refiner
    assert e.value.code == -1
    with pytest.raises(hgp.HgpError) as e:
        hgp.refine(x, a, vw, w, 2, 4, 1, np.zeros(4, np.int32), 0, refiner=-1)
    assert e.value.code == -1
    with pytest.raises(hgp.HgpError) as e:
        hgp.refine_2way_lp(x, a, vw, w, np.array([0, 2, 1, 0], np.int32))   # side 2
    assert e.value.code == -1


def test_struct_layouts_match_header(hgp, tmp_path):
    """The ctypes mirrors of hgp_options / hgp_stats have the C layout."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "hunyuangraph.h"\n'
        'int main(void) { printf("%zu %zu %zu %zu %zu\\n", sizeof(hgp_options), sizeof(hgp_stats),'
        ' offsetof(hgp_stats, level_iters), offsetof(hgp_stats, arena_peak), sizeof(hgp_graph)); return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(hgp.HgpOptions), C.sizeof(hgp.HgpStats), hgp.HgpStats.level_iters.offset,
            hgp.HgpStats.arena_peak.offset, C.sizeof(hgp.HgpGraph)]
    assert got == want


def test_tuning_round_trip_and_range(hgp):
    t0 = hgp.get_tuning()
    try:
        hgp.set_tuning(tiny_max_vertices=1024, batch_lanes=3, hub_track=1, graph_max_vertices=77)
        t = hgp.get_tuning()
        assert t["tiny_max_vertices"] == 1024 and t["batch_lanes"] == 3 and t["hub_track"] == 1
        assert t["graph_max_vertices"] == 77
        for bad in (dict(tiny_max_vertices=5000), dict(fill_waves=0), dict(batch_lanes=0), dict(seg_grid=0),
                    dict(hub_track=2), dict(graph_max_vertices=-1), dict(batch_fuse_max_vertices=16385)):
            with pytest.raises(hgp.HgpError) as e:
                hgp.set_tuning(**bad)
            assert e.value.code == -1
        assert hgp.get_tuning()["batch_lanes"] == 3       # a refused setting changes nothing
    finally:
        hgp.set_tuning(**t0)
