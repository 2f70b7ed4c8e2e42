"""bench.py's reading of the committed counter summaries (CPU only): the
traffic figure of the bench line's roofline comes from
profiles/<workload>_pmc_traffic.md, found by the table's header."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_measured_traffic_of_the_dominant_kernel():
    import bench
    path = os.path.join(ROOT, "profiles", "rmat_s22_e26_k256_pmc_traffic.md")
    if not os.path.exists(path):
        pytest.skip("no committed counter summary")
    t = bench.measured_traffic("rmat_s22_e26_k256", "k_jet_select_seg")
    assert isinstance(t, int) and t > 0
    # the row's algorithmic bytes, for a sanity bound on the figure read
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "k_jet_select_seg":
            alg = float(cells[-2])
            assert 0.5 * alg < t < 2.0 * alg
            break
    else:
        pytest.fail("k_jet_select_seg row missing")
    assert bench.measured_traffic("rmat_s22_e26_k256", "no_such_kernel") is None
    assert bench.measured_traffic("no_such_workload", "k_jet_select_seg") is None
