"""The bisection trials' FM driven by device priority queues (fm_kind 1: the
same moves as the wave arg-max scan; fm_kind 2: the reference's queue-driven
FM, FM_2way_cut_refinement with priority_queue_t,
hunyuangraph_GPU_initialpartition.h:160-300 and
hunyuangraph_GPU_priorityqueue.h:27-225) on the GPU against the oracle, bit
for bit, with queues in LDS (small node graphs) and in global scratch (node
graphs above the LDS capacity)."""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu


def _cases(hgp):
    return [
        ("er-300-w", random_graph(300, 0.03, seed=21, wmax=7, vmax=3)),
        ("grid-40x50", hgp.gen_grid2d(40, 50)),
        ("rmat-s12", hgp.gen_rmat(12, 1 << 15, seed=4)),           # 4096 vertices: queues in global scratch
        ("rgg-2500", hgp.gen_rgg3d(2500, hgp.rgg_radius(2500, 10.0), seed=6, ew=(1, 5))),
    ]


@pytest.mark.parametrize("fm_kind", [1, 2])
def test_initpart_heap_fm_bit_exact(hgp, orc, fm_kind):
    for name, (x, a, vw, aw) in _cases(hgp):
        for k in (2, 6, 16):
            print(f"heap fm case {name} k={k} fm_kind={fm_kind}", flush=True)
            w_g = hgp.initpart(x, a, vw, aw, k, seed=3, trials=48, fm_kind=fm_kind)
            w_o = orc.dga_initpart(x, a, vw, aw, k, seed=3, trials=48, fm_kind=fm_kind)
            np.testing.assert_array_equal(w_g, w_o, err_msg=f"{name} k={k}")
            if fm_kind == 1:
                w_0 = hgp.initpart(x, a, vw, aw, k, seed=3, trials=48, fm_kind=0)
                np.testing.assert_array_equal(w_g, w_0, err_msg=f"{name} k={k} scan")


def test_pipeline_reference_heap_fm(hgp, orc):
    for name, g, k in [("er-2000", random_graph(2000, 0.004, seed=5, wmax=4, vmax=2), 8),
                       ("rmat-s14", hgp.gen_rmat(14, 1 << 17, seed=2), 32)]:
        x, a, vw, aw = g
        print(f"heap fm pipeline {name} k={k}", flush=True)
        p_g, st_g = hgp.partition(x, a, k, vw, aw, seed=4, fm_kind=2)
        p_o, st_o = orc.dga_partition(x, a, k, vw, aw, seed=4, fm_kind=2)
        np.testing.assert_array_equal(p_g, p_o, err_msg=name)
        assert st_g["cut"] == st_o["cut"] == orc.edge_cut(x, a, aw, p_g)
        assert orc.balanced(vw, p_g, k)
        p_1, _ = hgp.partition(x, a, k, vw, aw, seed=4, fm_kind=1)
        p_0, _ = hgp.partition(x, a, k, vw, aw, seed=4, fm_kind=0)
        np.testing.assert_array_equal(p_1, p_0, err_msg=name)
