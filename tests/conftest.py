"""Test configuration: the ``gpu`` marker and shared graph fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# PyTorch (the device buffers of some GPU tests) ships its own HIP runtime and
# loads it as torch/lib/libamdhip64.so, a name the dynamic loader does not
# match against the libamdhip64.so.7 that libhgp.so links.  If libhgp.so is
# loaded first, torch then loads a second HIP runtime, and whichever of the
# two initialises the GPU second finds no device.  Importing torch before the
# library makes libhgp.so bind to torch's runtime: one runtime per process.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR (int64 xadj) from undirected (u, v) pairs; rows sorted."""
    rows = [[] for _ in range(n)]
    for i, (u, v) in enumerate(edges):
        w = 1 if weights is None else int(weights[i])
        rows[u].append((v, w))
        rows[v].append((u, w))
    xadj = np.zeros(n + 1, dtype=np.int64)
    adj, wts = [], []
    for u in range(n):
        rows[u].sort()
        for v, w in rows[u]:
            adj.append(v)
            wts.append(w)
        xadj[u + 1] = len(adj)
    return xadj, np.array(adj, dtype=np.int32), np.array(wts, dtype=np.int32)


def random_graph(n, p, seed, wmax=1, vmax=1):
    """Erdos-Renyi-like graph with optional random weights (rows sorted)."""
    rng = np.random.default_rng(seed)
    edges, ws = [], []
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < p:
                edges.append((u, v))
                ws.append(int(rng.integers(1, wmax + 1)))
    xadj, adj, aw = csr_from_edges(n, edges, ws)
    vw = rng.integers(1, vmax + 1, size=n).astype(np.int32)
    return xadj, adj, vw, aw


@pytest.fixture(scope="session")
def hgp():
    import hunyuansc25_amd as H
    H.lib()
    return H


@pytest.fixture(scope="session")
def orc():
    import oracle as O
    O.lib()
    return O
