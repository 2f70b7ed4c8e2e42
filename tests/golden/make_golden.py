"""Writes the golden fixtures of this folder.

The reference ships no known-answer vectors for its partitioning path, so the
fixtures here are derived independently of the oracle, from the definitions:

* contract.json    -- coarse CSR of small graphs under fixed matchings, by a
                      dictionary contraction (rename by cmap, drop self
                      entries, sum duplicates, sort rows).
* bisection.json   -- minimum edge cut of tiny graphs over all 2-way splits
                      with part-0 weight in [lo, hi], by enumeration.
* gain_bucket.json -- Jet gain buckets by exact rational arithmetic: bucket 0
                      for gain > 0, 1 for gain = 0, else 25 shifted by the
                      number of factors 3/2 between |gain|/vwgt and 1.
* edge_hash.json   -- the SplitMix64-based edge priorities of the handshake
                      matching, by an independent Python SplitMix64.
* known_cut.json   -- graphs whose optimal k-way cut is known by
                      construction (k cliques on a ring of single edges).

Run: python tests/golden/make_golden.py  (writes next to this file).
"""
import itertools
import json
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def csr(n, edges, weights=None):
    rows = [[] for _ in range(n)]
    for i, (u, v) in enumerate(edges):
        w = 1 if weights is None else int(weights[i])
        rows[u].append((v, w))
        rows[v].append((u, w))
    xadj, adj, wts = [0], [], []
    for u in range(n):
        for v, w in sorted(rows[u]):
            adj.append(v)
            wts.append(w)
        xadj.append(len(adj))
    return xadj, adj, wts


def contract(n, xadj, adj, wts, vw, match):
    rep = [min(v, match[v]) if match[v] >= 0 else v for v in range(n)]
    reps = sorted(set(rep))
    cid = {r: i for i, r in enumerate(reps)}
    cmap = [cid[rep[v]] for v in range(n)]
    cn = len(reps)
    rows = [dict() for _ in range(cn)]
    cvw = [0] * cn
    for v in range(n):
        cvw[cmap[v]] += vw[v]
        for j in range(xadj[v], xadj[v + 1]):
            c, cc = cmap[v], cmap[adj[j]]
            if c != cc:
                rows[c][cc] = rows[c].get(cc, 0) + wts[j]
    cx, ca, cw = [0], [], []
    for c in range(cn):
        for cc in sorted(rows[c]):
            ca.append(cc)
            cw.append(rows[c][cc])
        cx.append(len(ca))
    return cmap, cx, ca, cw, cvw


def splitmix_mix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def edge_hash(seed, level, lo, hi):
    z = splitmix_mix(seed ^ ((1 << 32) + level))
    z = splitmix_mix(z ^ lo)
    z = splitmix_mix(z ^ hi)
    return z >> 32


def gain_bucket(gain, vwgt):
    if gain > 0:
        return 0
    if gain == 0:
        return 1
    r = Fraction(-gain, vwgt)
    t = 25
    if r < 1:
        while r < 1 and t >= 2:
            r *= Fraction(3, 2)
            t -= 1
        return max(t, 2)
    while r > 1 and t <= 50:
        r *= Fraction(2, 3)
        t += 1
    return 49 if t > 50 else t


def main():
    rng = np.random.default_rng(2026)
    # contraction
    cases = []
    graphs = [
        ("path6", 6, [(i, i + 1) for i in range(5)], None),
        ("triangle_pairs", 6, [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (2, 3)], [1, 2, 3, 4, 5, 6, 7]),
        ("star7", 7, [(0, i) for i in range(1, 7)], [1, 2, 3, 4, 5, 6]),
        ("k5", 5, list(itertools.combinations(range(5), 2)), list(range(1, 11))),
    ]
    for i in range(4):
        n = 10
        e = [(u, v) for u, v in itertools.combinations(range(n), 2) if rng.random() < 0.35]
        graphs.append((f"er10_{i}", n, e, [int(x) for x in rng.integers(1, 9, size=len(e))]))
    for name, n, e, w in graphs:
        xadj, adj, wts = csr(n, e, w)
        vw = [int(x) for x in rng.integers(1, 5, size=n)]
        for t in range(3):
            perm = [int(x) for x in rng.permutation(n)]
            match = [-1] * n
            for p in range(int(rng.integers(0, n // 2 + 1))):
                u, v = perm[2 * p], perm[2 * p + 1]
                match[u], match[v] = v, u
            cmap, cx, ca, cw, cvw = contract(n, xadj, adj, wts, vw, match)
            cases.append({"name": f"{name}_{t}", "n": n, "xadj": xadj, "adjncy": adj, "adjwgt": wts, "vwgt": vw,
                          "match": match, "cmap": cmap, "cxadj": cx, "cadjncy": ca, "cadjwgt": cw, "cvwgt": cvw})
    json.dump(cases, open(os.path.join(HERE, "contract.json"), "w"))

    # bisection optimum by enumeration
    bis = []
    for i in range(8):
        n = int(rng.integers(6, 13))
        e = [(u, v) for u, v in itertools.combinations(range(n), 2) if rng.random() < 0.4]
        w = [int(x) for x in rng.integers(1, 6, size=len(e))]
        xadj, adj, wts = csr(n, e, w)
        vw = [int(x) for x in rng.integers(1, 4, size=n)]
        tv = sum(vw)
        lo, hi = tv // 2 - 1, (tv + 1) // 2 + 1
        best = None
        for mask in range(1 << n):
            w0 = sum(vw[v] for v in range(n) if mask >> v & 1)
            if not (lo <= w0 <= hi):
                continue
            cut = sum(wts[j] for v in range(n) for j in range(xadj[v], xadj[v + 1])
                      if (mask >> v & 1) != (mask >> adj[j] & 1)) // 2
            best = cut if best is None else min(best, cut)
        bis.append({"n": n, "xadj": xadj, "adjncy": adj, "adjwgt": wts, "vwgt": vw, "lo": lo, "hi": hi,
                    "min_cut": best})
    json.dump(bis, open(os.path.join(HERE, "bisection.json"), "w"))

    # gain buckets
    gb = []
    for g in list(range(-40, 6)) + [-(1 << 20), -(1 << 30), -999999937]:
        for v in (1, 2, 3, 7, 64, 1000, 1 << 20):
            gb.append([g, v, gain_bucket(g, v)])
    json.dump(gb, open(os.path.join(HERE, "gain_bucket.json"), "w"))

    # edge hashes
    eh = []
    for _ in range(64):
        seed = int(rng.integers(0, 1 << 63))
        level = int(rng.integers(0, 40))
        lo = int(rng.integers(0, 1 << 31))
        hi = int(rng.integers(lo, 1 << 31))
        eh.append([seed, level, lo, hi, edge_hash(seed, level, lo, hi)])
    json.dump(eh, open(os.path.join(HERE, "edge_hash.json"), "w"))

    # known optimal k-way cuts: k cliques of size s joined in a ring by one edge each
    kc = []
    for k, s in ((4, 12), (8, 10), (3, 20)):
        n = k * s
        e = []
        for c in range(k):
            base = c * s
            e += [(base + i, base + j) for i, j in itertools.combinations(range(s), 2)]
            e.append((base + s - 1, ((c + 1) % k) * s))
        xadj, adj, wts = csr(n, e)
        kc.append({"k": k, "n": n, "xadj": xadj, "adjncy": adj, "optimal_cut": k})
    json.dump(kc, open(os.path.join(HERE, "known_cut.json"), "w"))


if __name__ == "__main__":
    main()
