"""Pins of the oracle's heuristic steps on constructed inputs whose outcome is
forced, each expectation worked out by hand from the reference's rules (the
comments give the derivation), plus brute-force and semantic checks:

* matching fallbacks (hunyuangraph_GPU_match.h:1374-1716): leaf, isolate,
  twin, relative and forced pairing, their id order and the maxvwgt gate;
* weak rebalancing, jetrw (hunyuangraph_GPU_krefine.h:3136-3366): the
  destination rules, bucket order and the eviction cut-off at the excess;
* strong rebalancing, jetrs (:2515-2766): candidate order, eviction and the
  cookie cutter (:2710-2745) against a brute-force closest-prefix search;
* the Jet LP afterburner (:2137-2178): each kept move has a non-negative
  gain once the higher-priority moves are actually applied;
* the acceptance rule of k_refine (:4805-4839) on a constructed sequence;
* the CPU recursive bisection (hunyuangraph_CPU_initialpartition.h:361-374)
  on graphs whose optimal partition is known in closed form: a wrong gain
  sign, a dropped FM pass or a bad split leaves cut edges that the optimum
  does not have.
"""
import numpy as np
import pytest

from conftest import csr_from_edges


def _graph(n, edges, weights=None):
    x, a, w = csr_from_edges(n, edges, weights)
    return x, a, w


# ------------------------------------------------------------------ matching fallbacks

def test_leaf_pairs_in_id_order_per_hub(orc):
    # hub 0 with leaves 1..5, hub 6 with leaves 7, 8; the hubs are joined.
    # Leaves of one hub pair as they come in id order: (1,2), then (3,4) is
    # refused (3 weighs 9: 9 + 1 >= maxvwgt 5) and 5 is left alone; (7,8).
    edges = [(0, i) for i in range(1, 6)] + [(6, 7), (6, 8), (0, 6)]
    x, a, w = _graph(9, edges)
    vw = np.ones(9, np.int32)
    vw[3] = 9
    got = orc.fb_apply("leaf", x, a, w, vw, 5, -np.ones(9, np.int32))
    assert got.tolist() == [-1, 2, 1, -1, -1, -1, -1, 8, 7]


def test_isolate_pairs_consecutive_isolated_vertices(orc):
    # isolated 0, 2, 3, 5, 6 (1-4 is an edge): list pairs (0,2), (3,5), 6;
    # (3,5) is refused by weight (3 weighs 9); 6 has no partner
    x, a, w = _graph(7, [(1, 4)])
    vw = np.ones(7, np.int32)
    vw[3] = 9
    got = orc.fb_apply("isolate", x, a, w, vw, 5, -np.ones(7, np.int32))
    assert got.tolist() == [2, -1, 0, -1, -1, -1, -1]


def test_twin_groups_by_degree_and_row_sum(orc):
    # hubs 0, 1, 2 (weight 10).  Rows: 3, 4, 5 = {0, 1}; 6 = {1} (row sum 1
    # like 3..5, degree 1); 7, 8, 9 = {0, 2}.  4 is already matched (to 0).
    # Twin groups (degree, sum): {3, 5} -> (3,5); {6} alone; {7, 8, 9}:
    # (7,8) refused (8 weighs 9), 9 left alone; hubs have distinct keys.
    edges = [(3, 0), (3, 1), (4, 0), (4, 1), (5, 0), (5, 1), (6, 1), (7, 0), (7, 2), (8, 0), (8, 2),
             (9, 0), (9, 2)]
    x, a, w = _graph(10, edges)
    vw = np.ones(10, np.int32)
    vw[:3] = 10
    vw[8] = 9
    m0 = -np.ones(10, np.int32)
    m0[0], m0[4] = 4, 0
    got = orc.fb_apply("twin", x, a, w, vw, 5, m0)
    assert got.tolist() == [4, -1, -1, 5, 0, 3, -1, -1, -1, -1]


def test_twin_equal_degree_and_sum_pairs_different_rows(orc):
    # the reference compares only the degree and the hashed row sum, so rows
    # {0, 3} and {1, 2} (both sum 3, degree 2) pair although they differ
    x, a, w = _graph(6, [(4, 0), (4, 3), (5, 1), (5, 2)])
    vw = np.ones(6, np.int32)
    vw[:4] = 10
    got = orc.fb_apply("twin", x, a, w, vw, 5, -np.ones(6, np.int32))
    assert got[4] == 5 and got[5] == 4


def test_relative_prefers_degree_then_edge_weight(orc):
    # hubs P=0, Q=1, R=2 (weight 10), light 3..8.  Degrees: P 3, Q 4, R 4.
    #   3 {P, R}      -> R (degree 4 > 3)
    #   4 {R}         -> R
    #   5 {P, Q w5}   -> Q
    #   6 {P w2, Q w2}-> Q (degree)
    #   7 {Q, R}      -> Q (tie in degree and weight: first in row)
    #   8 {Q w1, R w3}-> R (tie in degree, heavier edge)
    # Hubs pick P->6, Q->5, R->8 (ties in degree to the heavier edge).
    # In id order: P, Q, R claim 6, 5, 8 alone; 3 and 4 share R -> (3,4);
    # 5 and 6 share Q -> (5,6); 7 (Q) and 8 (R) find no partner.
    edges = [(3, 0), (3, 2), (4, 2), (5, 0), (5, 1), (6, 0), (6, 1), (7, 1), (7, 2), (8, 1), (8, 2)]
    wts = [1, 1, 1, 1, 5, 2, 2, 1, 1, 1, 3]
    x, a, w = _graph(9, edges, wts)
    vw = np.ones(9, np.int32)
    vw[:3] = 10
    got = orc.fb_apply("relative", x, a, w, vw, 5, -np.ones(9, np.int32))
    assert got.tolist() == [-1, -1, -1, 4, 3, 6, 5, -1, -1]


def test_forced_pairs_consecutive_free_vertices(orc):
    # free in id order: 0, 2, 3, 4, 5, 7 (1-6 matched): (0,2) refused
    # (1 + 4 >= 5), then (3,4), (5,7)
    x, a, w = _graph(8, [(1, 6)])
    vw = np.array([1, 1, 4, 1, 1, 1, 1, 1], np.int32)
    m0 = -np.ones(8, np.int32)
    m0[1], m0[6] = 6, 1
    got = orc.fb_apply("forced", x, a, w, vw, 5, m0)
    assert got.tolist() == [-1, 6, -1, 4, 3, 7, 1, 5]


def test_fallback_chain_in_match(orc):
    # The same rules in sequence inside dga_match (all handshakes refused by
    # maxvwgt): hubs 0, 1; 2 = {1}; 3..7 = {0, 1}.  Twin pairs (3,4), (5,6);
    # 7 is left.  Still < 75 % matched, so relative runs: 2 and 7 both prefer
    # hub 1 (degree 6 > 5) -> (2,7).  6 of 8 matched: forced does not run.
    edges = [(2, 1)] + [(v, h) for v in range(3, 8) for h in (0, 1)]
    x, a, w = _graph(8, edges)
    vw = np.ones(8, np.int32)
    vw[:2] = 10
    m, cnt = orc.dga_match(x, a, vw, w, 5, 1, 0, 8)
    assert m.tolist() == [-1, -1, 7, 4, 3, 6, 5, 2] and cnt == 6


# ------------------------------------------------------------------ weak rebalancing

def test_rw_bucket_order_and_cutoff(orc):
    # part 0 = {0, 1, 2, 4}, part 1 = {3}; pwgts [4, 1], maxwgt 2, opt 2.
    # Every vertex of part 0 qualifies (2*1 < 3*(4-2)).  Part 1 is the only
    # part under maxwgt:
    #   0: conn to 1 = 2, own 1 -> gain +1, bucket 0
    #   1: conn to 1 = 1, own 1 -> gain  0, bucket 1
    #   2: conn to 1 = 1, own 2 -> gain -1, bucket 25
    #   4: no edge to part 1, and no part is below maxdest 1 -> no candidate
    # Excess 4 - 2 = 2: evict while the weight ahead is < 2 -> 0 and 1 only.
    edges = [(0, 3), (0, 4), (1, 3), (1, 4), (2, 3), (2, 4)]
    wts = [2, 1, 1, 1, 1, 2]
    x, a, w = _graph(5, edges, wts)
    vw = np.ones(5, np.int32)
    where = np.array([0, 0, 0, 1, 0], np.int32)
    dest, moves = orc.jet_rw_step(x, a, w, vw, 2, where, [4, 1], 2, 2)
    assert dest.tolist() == [1, 1, 0, 1, 0] and moves == 2


def test_rw_undersized_parts_by_vertex_id(orc):
    # part 0 = path 0-1-2-3-4-5 (pwgts 6), parts 1 = {6}, 2 = {7}, edge 6-7;
    # maxwgt 3, opt 2, maxdest max(99*3/100, 3-100) = 2: parts 1 and 2 are
    # undersized.  No vertex of part 0 touches them, so dest = under[v % 2]
    # and gain = -own: ends 0, 5 -> -1 (bucket 25), inner -> -2 (bucket 27).
    # Order 0, 5, 1, 2, 3, 4; excess 3 -> 0, 5, 1 evicted: 0 -> 1, 5 -> 2, 1 -> 2.
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (6, 7)]
    x, a, w = _graph(8, edges)
    vw = np.ones(8, np.int32)
    where = np.array([0, 0, 0, 0, 0, 0, 1, 2], np.int32)
    dest, moves = orc.jet_rw_step(x, a, w, vw, 3, where, [6, 1, 1], 3, 2)
    assert dest.tolist() == [1, 2, 0, 0, 0, 2, 1, 2] and moves == 3


def test_rw_skips_heavy_vertices(orc):
    # a vertex qualifies only if 2*vwgt < 3*(pwgts[p] - opt): with pwgts 5,
    # opt 3 the bound is 2*vwgt < 6, so the weight-3 vertex 0 stays and the
    # weight-2 vertex 1 (gain +1 into part 1) covers the excess 5 - 4 alone
    edges = [(0, 2), (1, 2)]
    x, a, w = _graph(3, edges)
    vw = np.array([3, 2, 1], np.int32)
    where = np.array([0, 0, 1], np.int32)
    dest, moves = orc.jet_rw_step(x, a, w, vw, 2, where, [5, 1], 4, 3)
    assert dest.tolist() == [0, 1, 1] and moves == 1


def test_rw_bucket_fifty_stays_in_its_part(orc):
    # gain_bucket returns 50 for -gain/vwgt in (1.5^24, 1.5^25]: vertices 1
    # and 2 of part 0 (joined by an edge of weight 20000, no edge to the only
    # target part 2, dest = undersized part 2, gain -20000).  They sort into
    # their own part's last bucket, not into part 1's first bucket beside
    # vertex 0 (part 1, gain +1 into part 2, bucket 0).
    #   part 0 (pwgts 302, excess 102): 3 (bucket 1, weight 300), then 1, 2
    #     -> only 3 is evicted;
    #   part 1 (pwgts 301, excess 101): 0 (bucket 0), then 4 (bucket 1)
    #     -> both evicted (weight ahead of 4 is 1).
    # Counting 1 and 2 in part 1's run would evict them as well.
    assert orc.gain_bucket(-20000, 1) == 50
    x, a, w = _graph(6, [(1, 2), (0, 5)], [20000, 1])
    vw = np.array([1, 1, 1, 300, 300, 1], np.int32)
    where = np.array([1, 0, 0, 0, 1, 2], np.int32)
    dest, moves = orc.jet_rw_step(x, a, w, vw, 3, where, [302, 301, 1], 200, 100)
    assert dest.tolist() == [2, 0, 0, 2, 2, 2] and moves == 3


# ------------------------------------------------------------------ strong rebalancing

def test_rs_eviction_and_cookie_cutter(orc):
    # part 0 = {0, 1, 2, 3}, part 1 = {4}; pwgts [4, 1], maxwgt 3, opt 2.
    # Loss = own connectivity: 0 -> 4 (0-1 w3, 0-3), 1 -> 4 (0-1 w3, 1-2),
    # 2 -> 1, 3 -> 1; buckets 29, 29, 25, 25.  Order 2, 3, 0, 1; excess 1
    # -> only 2 is evicted.  Cookie cutter: part 0 capacity 0 (empty slice),
    # part 1 capacity 2: prefix [0, 1] closest to 2 ends at 1 -> 2 goes to 1.
    edges = [(0, 1), (1, 2), (2, 4), (3, 4), (0, 3)]
    wts = [3, 1, 1, 2, 1]
    x, a, w = _graph(5, edges, wts)
    vw = np.ones(5, np.int32)
    where = np.array([0, 0, 0, 0, 1], np.int32)
    dest, moves = orc.jet_rs_step(x, a, w, vw, 2, where, [4, 1], 3, 2)
    assert dest.tolist() == [0, 0, 1, 0, 1] and moves == 1


def _cookie_brute(bs, maxwgt, pwgts):
    """Consecutive slices; each ends at the prefix index e >= start whose
    bs[e] is closest to bs[start] + max(maxwgt - pwgts[p], 0), ties to the
    lower index (linear search over every candidate end)."""
    ne = len(bs) - 1
    es, ee, start = [], [], 0
    for pw in pwgts:
        find = bs[start] + max(maxwgt - pw, 0)
        end = min(range(start, ne + 1), key=lambda e: (abs(bs[e] - find), e))
        es.append(start)
        ee.append(end)
        start = end
    return es, ee


def test_cookie_cutter_brute_force(orc):
    rng = np.random.default_rng(17)
    for _ in range(400):
        k = int(rng.integers(1, 7))
        ne = int(rng.integers(0, 14))
        wts = rng.integers(1, 6, size=ne)
        bs = np.concatenate([[0], np.cumsum(wts)]).astype(np.int64)
        maxwgt = int(rng.integers(1, 20))
        pwgts = rng.integers(0, 25, size=k).astype(np.int64)
        es, ee = orc.cookie_cutter(bs, maxwgt, pwgts)
        bes, bee = _cookie_brute(bs.tolist(), maxwgt, pwgts.tolist())
        assert es.tolist() == bes and ee.tolist() == bee, (bs.tolist(), maxwgt, pwgts.tolist())


def test_cookie_cutter_hand_case(orc):
    # weights 2, 1, 3, 1 -> prefix 0, 2, 3, 6, 7; capacities 3, 0, 4:
    # part 0 wants 3 -> ends at 2 (bs 3); part 1 wants 0 -> empty at 2;
    # part 2 wants 3 + 4 = 7 -> ends at 4
    es, ee = orc.cookie_cutter([0, 2, 3, 6, 7], 10, [7, 12, 6])
    assert es.tolist() == [0, 2, 2] and ee.tolist() == [2, 2, 4]
    # a tie (wants 4 between prefixes 3 and 5) goes to the lower end
    es, ee = orc.cookie_cutter([0, 3, 5], 4, [0])
    assert ee.tolist() == [1]


# ------------------------------------------------------------------ LP afterburner

def _conn(x, a, aw, v, where):
    c = {}
    for j in range(x[v], x[v + 1]):
        q = int(where[a[j]])
        c[q] = c.get(q, 0) + int(aw[j])
    return c


def test_lp_afterburner_applies_higher_priority_moves(orc):
    """For every candidate of an LP step: it is kept iff its gain is >= 0
    after the moves of its higher-priority candidate neighbours (larger
    gain, ties to the smaller id) have been applied to the partition."""
    from conftest import random_graph
    for seed in range(6):
        n, k = 70, 4
        x, a, vw, aw = random_graph(n, 0.09, seed=40 + seed, wmax=5)
        where = np.random.default_rng(seed).integers(0, k, size=n).astype(np.int32)
        lock = (np.random.default_rng(seed + 3).random(n) < 0.15).astype(np.uint8)
        fnum, fden = (1, 4) if seed % 2 else (3, 4)
        dest_o, lock_o, moves = orc.jet_lp_step(x, a, aw, k, where, fnum, fden, lock)
        # candidates from first principles
        dest, gain = {}, {}
        for v in range(n):
            c = _conn(x, a, aw, v, where)
            p = int(where[v])
            pc = c.get(p, 0)
            others = [(cq, -q) for q, cq in c.items() if q != p and cq > 0]
            if not others or lock[v]:
                continue
            bc, nq = max(others)
            if bc >= pc or pc - bc < (fnum * pc) // fden:
                dest[v], gain[v] = -nq, bc - pc
        kept = set()
        for v in dest:
            ahead = [int(u) for u in a[x[v]:x[v + 1]] if int(u) in dest and
                     (gain[int(u)] > gain[v] or (gain[int(u)] == gain[v] and int(u) < v))]
            w2 = where.copy()
            for u in ahead:
                w2[u] = dest[u]
            c2 = _conn(x, a, aw, v, w2)
            if c2.get(dest[v], 0) - c2.get(int(where[v]), 0) >= 0:
                kept.add(v)
        moved = {v for v in range(n) if dest_o[v] != where[v]}
        assert moved == kept, seed
        assert all(dest_o[v] == dest[v] for v in moved)
        assert moves == len(kept) and lock_o.tolist() == [1 if v in kept else 0 for v in range(n)]


def test_lp_hand_case(orc):
    # path 0-1-2-3 plus 1-3, where [0, 0, 1, 1], k = 2, filter 3/4, no locks.
    # conn: 0 {0:1} none; 1 {0:1, 1:2} -> dest 1, gain +1; 2 {0:1, 1:1}: 1 - 1
    # = 0 >= 0 -> dest 0, gain 0; 3 {0:1, 1:1} -> dest 0, gain 0.
    # Afterburner: 1 (gain 1) goes first and is kept.  2: neighbour 1 is
    # ahead (gain 1 > 0) and moves into 2's part: conn becomes {1:2} -> 0-2 < 0,
    # dropped.  3: 1 is ahead (moves to 1) and 2 is ahead (same gain, smaller
    # id, dest 0): conn {1: 1 (from 1), 0: 1 (from 2)} -> 0, kept.
    x, a, w = _graph(4, [(0, 1), (1, 2), (2, 3), (1, 3)])
    where = np.array([0, 0, 1, 1], np.int32)
    dest, lock, moves = orc.jet_lp_step(x, a, w, 2, where, 3, 4, np.zeros(4, np.uint8))
    assert dest.tolist() == [0, 1, 1, 0] and moves == 2 and lock.tolist() == [0, 1, 0, 1]


# ------------------------------------------------------------------ acceptance rule

def test_acceptance_sequence(orc):
    maxwgt = 8
    st = (10000, 10, 5)                     # best cut, best heaviest part (unbalanced), patience count
    steps = [
        # (cur cut, cur heaviest, count before) -> (accepted, best cut, best heaviest, count)
        ((12000, 9, 5), (True, 12000, 9, 0)),     # unbalanced best: lighter heaviest part wins, cut ignored
        ((11000, 9, 1), (True, 11000, 9, 0)),     # smaller cut, no worse balance, > 0.1 % better: reset
        ((10000, 12, 1), (False, 11000, 9, 1)),   # smaller cut but heavier than the best and over maxwgt
        ((10990, 8, 2), (True, 10990, 8, 0)),     # balanced now: lighter heaviest part of an unbalanced best
        ((10980, 8, 2), (True, 10980, 8, 2)),     # smaller cut by < 0.1 %: kept, count not reset
        ((10960, 8, 2), (True, 10960, 8, 0)),     # smaller by > 0.1 %: reset
        ((10950, 9, 1), (False, 10960, 8, 1)),    # over maxwgt and heavier than the balanced best
        ((10960, 7, 1), (False, 10960, 8, 1)),    # equal cut: not kept, whatever the balance
        ((10959, 7, 1), (True, 10959, 7, 1)),     # one less cut edge: kept, no reset
    ]
    best_cut, best_max, _ = st
    for (cut, mx, count), exp in steps:
        got = orc.jet_accept(cut, mx, maxwgt, best_cut, best_max, count)
        assert got == exp, ((cut, mx, count), got, exp)
        _, best_cut, best_max, _ = got


# ------------------------------------------------------------------ CPU recursive bisection

def _clique(off, s):
    return [(off + i, off + j) for i in range(s) for j in range(i + 1, s)]


def _unit(n, edges, weights=None):
    x, a, w = csr_from_edges(n, edges, weights)
    return x, a, np.ones(n, np.int32), w


@pytest.mark.parametrize("s", [8, 20, 50])
def test_cpu_rb_barbell_cuts_the_bridge(orc, s):
    # two K_s joined by one edge: the only balanced bisection of cut 1 puts
    # each clique in its own part; any other split cuts >= s - 1 clique edges
    x, a, vw, aw = _unit(2 * s, _clique(0, s) + _clique(s, s) + [(s - 1, s)])
    for seed in (1, 2, 3):
        p, cut = orc.cpu_rb_partition(x, a, 2, vw, aw, seed=seed)
        assert cut == 1
        assert len(set(p[:s])) == 1 and len(set(p[s:])) == 1 and p[0] != p[s]


@pytest.mark.parametrize("n,k", [(64, 2), (64, 4), (200, 8)])
def test_cpu_rb_path_is_cut_into_runs(orc, n, k):
    # a path into k balanced parts: the optimum is k - 1 cut edges, each part
    # one run of consecutive vertices
    x, a, vw, aw = _unit(n, [(i, i + 1) for i in range(n - 1)])
    p, cut = orc.cpu_rb_partition(x, a, k, vw, aw, seed=1)
    assert cut == k - 1
    assert int(np.count_nonzero(np.diff(p))) == k - 1
    assert np.bincount(p, minlength=k).max() <= 1.03 ** 3 * n / k


def test_cpu_rb_separates_components(orc):
    # four disjoint K_8 into four parts: cut 0, one clique per part
    edges = [e for c in range(4) for e in _clique(8 * c, 8)]
    x, a, vw, aw = _unit(32, edges)
    for seed in (1, 5):
        p, cut = orc.cpu_rb_partition(x, a, 4, vw, aw, seed=seed)
        assert cut == 0
        assert sorted(p.reshape(4, 8)[:, 0].tolist()) == [0, 1, 2, 3]
        assert all(len(set(r)) == 1 for r in p.reshape(4, 8))


def test_cpu_rb_ladder_and_weighted_bridge(orc):
    # a 2 x 32 ladder: the optimal bisection cuts the two rails once each
    n = 32
    e = [(i, i + 1) for i in range(n - 1)] + [(n + i, n + i + 1) for i in range(n - 1)] + [(i, n + i) for i in range(n)]
    x, a, vw, aw = _unit(2 * n, e)
    assert orc.cpu_rb_partition(x, a, 2, vw, aw, seed=1)[1] == 2
    # two K_12 joined by three parallel bridges of weights 1, 1, 5: the
    # bisection cuts the bridges (weight 7), not the unit clique edges
    s = 12
    edges = _clique(0, s) + _clique(s, s) + [(0, s), (1, s + 1), (2, s + 2)]
    wts = [1] * (len(edges) - 3) + [1, 1, 5]
    x, a, vw, aw = _unit(2 * s, edges, wts)
    p, cut = orc.cpu_rb_partition(x, a, 2, vw, aw, seed=2)
    assert cut == 7 and len(set(p[:s])) == 1 and len(set(p[s:])) == 1
