"""Scheduling knobs (hgp_tuning) change no result: partitions under several
non-default settings equal those under the defaults, on the small-level
paths (one block for a level's whole refinement loop, one launch for a
level's short rows) and the long-row segment kernels, for single calls and
batches."""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu

SETTINGS = [
    dict(tiny_max_vertices=0, few_rows=0, fill_waves=1, seg_grid=1, batch_lanes=1, graph_max_vertices=1 << 30,
         batch_fuse_max_vertices=0),
    dict(tiny_max_vertices=4096, few_rows=1 << 20, fill_waves=65536, seg_grid=4, batch_lanes=3, hub_track=1,
         batch_fuse_max_vertices=300),
    dict(tiny_max_vertices=1024, few_rows=2048, fill_waves=4096, seg_grid=8192, batch_lanes=12, hub_track=1,
         graph_max_vertices=0),
]


def test_tuning_changes_no_result(hgp):
    graphs = [hgp.gen_grid2d(60, 70), random_graph(900, 0.01, seed=3, wmax=5, vmax=3), hgp.gen_rmat(14, 1 << 17, seed=9),
              hgp.gen_banded(5000, 32, 0.1, seed=4)]
    ks = [8, 5, 64, 16]
    default = hgp.get_tuning()
    try:
        hgp.set_tuning(**{k: v for k, v in default.items()})
        ref = [hgp.partition(g[0], g[1], k, g[2], g[3], seed=2)[0] for g, k in zip(graphs, ks)]
        for s in SETTINGS:
            print(f"tuning {s}", flush=True)
            hgp.set_tuning(**s)
            for g, k, r in zip(graphs, ks, ref):
                p, _ = hgp.partition(g[0], g[1], k, g[2], g[3], seed=2)
                np.testing.assert_array_equal(p, r, err_msg=str(s))
            parts, _ = hgp.partition_batch(graphs, ks, lanes=0, seed=2)
            for p, r in zip(parts, ref):
                np.testing.assert_array_equal(p, r, err_msg=f"batch {s}")
    finally:
        hgp.set_tuning(**default)
