"""Synthetic workloads of the batched and serving modes (host-side generators
over the library's own graph generators; seeded, so every run and test sees
the same graphs)."""
from __future__ import annotations

import os

import numpy as np


def batch_graphs(hgp, count, seed, n_lo=1000, n_hi=65536):
    """The batched workload: vertex counts log-uniform in [n_lo, n_hi]
    (default 1k..64k), cycling 2-D grids (about sqrt(n) x sqrt(n)), 3-D random
    geometric graphs (mean degree 12) and banded FEM-like patterns (bandwidth
    32, 10% random off-band entries); unit weights; k drawn from
    {2, 4, 8, 16, 32, 64}.  Returns (graphs, ks); a graph is (xadj, adjncy,
    vwgt, adjwgt)."""
    rng = np.random.default_rng(seed)
    specs, ks = [], []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        specs.append((i, n))
        ks.append(int(rng.choice([2, 4, 8, 16, 32, 64])))

    def make(spec):
        i, n = spec
        kind = i % 3
        if kind == 0:
            side = max(2, int(round(np.sqrt(n))))
            return hgp.gen_grid2d(side, max(2, n // side))
        if kind == 1:
            return hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12.0), seed=seed + i)
        return hgp.gen_banded(n, 32, 0.1, seed=seed + i)

    return _generate(make, specs), ks


def _generate(make, specs):
    # the generators are C code called through ctypes, which releases the
    # GIL: threads build the graphs in parallel, in order
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(make, specs))


def serve_requests(hgp, count, seed, n_lo=1e4, n_hi=4e6):
    """Serving requests: vertex counts log-uniform in [n_lo, n_hi] (default
    10k..4M), alternating R-MAT (edge factor 16, power-of-two vertex counts)
    and 3-D random geometric graphs (mean degree 12); k log-uniform over
    {8, 16, ..., 512}.  Returns [((xadj, adjncy, vwgt, adjwgt), k), ...]."""
    rng = np.random.default_rng(seed)
    specs, ks = [], []
    for i in range(count):
        n = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        ks.append(int(2 ** rng.integers(3, 10)))
        specs.append((i, n))

    def make(spec):
        i, n = spec
        if i % 2 == 0:
            scale = int(round(np.log2(n)))
            return hgp.gen_rmat(scale, 16 << scale, seed=seed + i)
        n = int(n)
        return hgp.gen_rgg3d(n, hgp.rgg_radius(n, 12.0), seed=seed + i)

    return list(zip(_generate(make, specs), ks))
