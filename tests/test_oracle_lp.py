"""Pins of the oracle's older label-propagation refiners (CPU only).

The k-way rounds restate hunyuangraph_k_refinement_me
(hunyuangraph_GPU_LP.h:1245-1374) and the 2-way pass restates jetLP
(:673-781).  The cases below are worked out by hand from those lines; each
one is chosen so that a plausible slip fails it: moves applied one by one
instead of together, ties broken by the highest id, moved vertices allowed to
move again, the rebalancing loop ending early, float bounds computed in
exact arithmetic, the flip pass reading sides already flipped, the sort
order reversed.
"""
import numpy as np
import pytest

from conftest import csr_from_edges, random_graph

import oracle as O


def lp(n, edges, where, k, vw=None, w=None):
    x, a, aw = csr_from_edges(n, edges, w)
    vw = np.ones(n, np.int32) if vw is None else np.asarray(vw, np.int32)
    return O.lp_refine(x, a, vw, aw, k, int(vw.sum()), np.asarray(where, np.int32))


def lp2(n, edges, where, vw=None, w=None):
    x, a, aw = csr_from_edges(n, edges, w)
    vw = np.ones(n, np.int32) if vw is None else np.asarray(vw, np.int32)
    return O.lp2_refine(x, a, vw, aw, np.asarray(where, np.int32))


PATH6 = [(i, i + 1) for i in range(5)]


def test_threshold_closed_form():
    # max(n / (6000 * floor(log2 k)), 8 k), hunyuangraph_GPU_LP.h:1251
    assert O.lp_threshold(100000, 2) == 16          # 100000 / 6000 = 16 = 8 * 2
    assert O.lp_threshold(100000, 4) == 32          # 100000 / 12000 = 8 < 32
    assert O.lp_threshold(6, 2) == 16
    assert O.lp_threshold(10**7, 256) == 2048       # 10^7 / 48000 = 208
    assert O.lp_threshold(10**8, 8) == 5555         # 10^8 / 18000
    assert O.lp_threshold(10**8, 7) == 8333         # floor(log2 7) = 2
    assert O.lp_threshold(1000, 1) == 0             # one part: no rounds
    for k in (2, 3, 5, 8, 9, 255, 256, 257, 4096):
        for n in (0, 1, 5999, 6000, 10**6, 2**31 - 1):
            lg = k.bit_length() - 1
            assert O.lp_threshold(n, k) == max(n // (6000 * lg), 8 * k)


def test_part_bounds_are_floats():
    lo, hi = O.lp_part_bounds(2, 1000)
    # 0.5 / 1.03 * 1000 = 485.43..., 0.5 * 1.03 * 1000 = 515 (as doubles, then float)
    assert lo == np.float32(485.43689) and hi == np.float32(515.0)
    # 2^30 at k = 3: floats there are 32 apart; the exact bound
    # 2^30 / 3 * 1.03 = 368,651,359.57 is stored as 368,651,360
    lo, hi = O.lp_part_bounds(3, 2**30)
    assert hi == np.float32(368651360.0)
    assert lo == np.float32(347489280.0)


@pytest.mark.parametrize("p0,over", [(368651370, False), (368651400, True)])
def test_overweight_compared_as_float(p0, over):
    # Parts of 2^30 total weight, k = 3.  Part 0 weighs p0: 368,651,370 is
    # above the exact bound 368,651,359.57 but converts to the float
    # 368,651,360, which is not above the float bound, so the rebalancing loop
    # does not start; 368,651,400 converts to 368,651,392 and does.  Part 1
    # (347,489,000) is under its bound either way.  Heavy edges inside each
    # part keep the first loop still; vertex 0 borders part 1 by one edge.
    # When the loop starts, its first pass moves vertex 0 into part 1, which
    # makes part 1 over and part 0 under; the next moves it back (the
    # rebalancer does not check the destination's bound), and so on for all
    # 8k = 24 rounds,
    # an even number: the partition ends where it began.
    tot = 2**30
    p1 = 347489000
    vw = [p0 - 1, 1, p1 - 1, 1, tot - p0 - p1 - 1, 1]
    edges = [(0, 1), (2, 3), (4, 5), (0, 2)]
    w, cut, it = lp(6, edges, [0, 0, 1, 1, 2, 2], 3, vw=vw, w=[10, 10, 10, 1])
    assert w.tolist() == [0, 0, 1, 1, 2, 2]
    assert it == ((0, 24) if over else (0, 0))


def test_round_moves_together():
    # path 0-1-2-3-4-5, sides 0 0 1 0 1 1: vertex 2 wants side 0 (gain 2),
    # vertex 3 wants side 1 (gain 2); 1 and 4 have equal weight both ways.
    # Both move in the same round -> 0 0 0 1 1 1, cut 1; the next round finds
    # no candidate (0 and 5 are interior, 2 and 3 moved).
    w, cut, it = lp(6, PATH6, [0, 0, 1, 0, 1, 1], 2)
    assert w.tolist() == [0, 0, 0, 1, 1, 1] and cut == 1 and it == (1, 0)


def test_one_winner_per_part():
    # path 0-1-2-3, sides 0 1 0 1: candidates for side 0 are 1 (gain 2) and
    # 3 (gain 1), for side 1 are 0 (gain 1) and 2 (gain 2).  Only the best
    # of each part moves: 1 -> 0, 2 -> 1.
    w, cut, it = lp(4, [(0, 1), (1, 2), (2, 3)], [0, 1, 0, 1], 2)
    assert w.tolist() == [0, 0, 1, 1] and cut == 1 and it == (1, 0)


def test_ties_lowest_id_and_moved_lock():
    # two separate edges 0-1 and 2-3, sides 1 0 1 0: every vertex has gain 1.
    # First pass: side 0 takes 0 (not 2), side 1 takes 1 (not 3) -> 0 1 1 0.
    # Second: 0 and 1 are locked; 2 -> 0 and 3 -> 1 -> 0 1 0 1.
    # Third: every vertex has moved; the loop ends with cut 2.
    w, cut, it = lp(4, [(0, 1), (2, 3)], [1, 0, 1, 0], 2)
    assert w.tolist() == [0, 1, 0, 1] and cut == 2 and it == (2, 0)


def test_destination_is_lowest_part_of_max():
    # star 0-1 (w 2), 0-2 (w 2), 0-3 (w 1), parts 2 1 0 2, k = 3.
    # First pass: vertex 0 sees 2 in parts 0 and 1, 1 in its own: destination
    # part 0 (the lower id), gain 1.  Vertices 1 and 2 both want part 2 with
    # gain 2: vertex 1 wins.  -> 0 2 0 2.  Second: vertex 3 (own part 2,
    # neighbour 0 now in part 0) moves to 0 -> 0 2 0 0; the third has no
    # candidate.  Rebalancing (bounds 1.29 / 1.37 of total 4): part 0 (3) is
    # over, parts 1 (0) and 2 (1) under; vertex 0 borders part 2 through
    # vertex 1 and moves there -> 2 2 0 0.  Then parts 0 and 2 are over and
    # part 1 under, but nothing borders the empty part 1: the remaining 23 of
    # the 24 rounds move nothing.
    edges = [(0, 1), (0, 2), (0, 3)]
    w, cut, it = lp(4, edges, [2, 1, 0, 2], 3, w=[2, 2, 1])
    assert w.tolist() == [2, 2, 0, 0] and cut == 3 and it == (2, 24)


def test_rebalance_rounds_without_moves_run_to_threshold():
    # all of a path in part 0 but vertex 5: the first pass moves 5 into part 0
    # (gain 1), after which part 1 is empty and nothing borders it.  The
    # rebalancing loop (part 0 over, part 1 under) finds no boundary vertex
    # and runs its full threshold of 16 rounds.
    w, cut, it = lp(6, PATH6, [0, 0, 0, 0, 0, 1], 2)
    assert w.tolist() == [0] * 6 and cut == 0 and it == (1, 16)


def test_rebalance_moves_into_thin_part():
    # sides 0 0 0 0 1 1: no first-loop move (3 and 4 are tied), part 0 weighs
    # 4 > 3.09, part 1 weighs 2 < 2.91; vertex 3 borders part 1 and moves.
    w, cut, it = lp(6, PATH6, [0, 0, 0, 0, 1, 1], 2)
    assert w.tolist() == [0, 0, 0, 1, 1, 1] and cut == 1 and it == (0, 1)


def test_lp_refine_invariants():
    x, a, vw, aw = random_graph(400, 0.02, seed=11, wmax=5, vmax=3)
    n = len(x) - 1
    rng = np.random.default_rng(3)
    for k in (2, 5, 16):
        where = rng.integers(0, k, size=n).astype(np.int32)
        w, cut, (i1, i2) = O.lp_refine(x, a, vw, aw, k, int(vw.sum()), where)
        assert cut == O.edge_cut(x, a, aw, w)
        assert w.min() >= 0 and w.max() < k
        thr = O.lp_threshold(n, k)
        assert 0 <= i1 <= thr and 0 <= i2 <= thr
        # the first loop moves each vertex at most once: fewer changed vertices
        # than rounds times parts
        assert int((w != where).sum()) <= (i1 + i2) * k
        if i2 < thr:
            lo, hi = O.lp_part_bounds(k, int(vw.sum()))
            pw = np.bincount(w, weights=vw, minlength=k).astype(np.float32)
            assert not ((pw > hi).any() and (pw < lo).any())


def test_lp_refine_improves_random_start():
    x, a, vw, aw = random_graph(300, 0.03, seed=5)
    n = len(x) - 1
    where = np.random.default_rng(1).integers(0, 4, size=n).astype(np.int32)
    c0 = O.edge_cut(x, a, aw, where)
    _, cut, _ = O.lp_refine(x, a, vw, aw, 4, int(vw.sum()), where)
    assert cut < c0


def test_lp2_synchronous_flip():
    # path 0-1-2-3, sides 0 1 0 1: every vertex has all its edge weight on
    # the other side, so all four flip together (a pass that read flipped
    # sides would stop flipping after vertex 0).  Sides then weigh 2 and 2.
    w, ran = lp2(4, [(0, 1), (1, 2), (2, 3)], [0, 1, 0, 1])
    assert w.tolist() == [1, 0, 1, 0] and not ran


def test_lp2_isolated_vertex_flips():
    # 0 <= 0: an isolated vertex changes side; the transfer then has nothing
    # it may move (num = (1 - 0) / 2 = 0).
    w, ran = lp2(1, [], [0])
    assert w.tolist() == [1] and ran


def test_lp2_transfer_prefix():
    # two triangles {0,1,2}, {3,4,5} joined by 2-3, sides 0 0 0 0 0 1.
    # Flip: 4 (1 vs 1) -> 1, 5 (0 vs 2) -> 0, others stay: 0 0 0 0 1 0.
    # Sides weigh 5 and 1: from = 0, num = (5 - 1) / 2 = 2.  Keys on side 0:
    # 5: 1-1 = 0, 3: 1-2 = -1, 0: -2, 1: -2, 2: -3; vertex 4 gets -m = -14.
    # Sorted: 5, 3, 0, 1, 2, 4; prefix weights 1, 2, 3 ... -> 5 and 3 move.
    tri = [(0, 1), (0, 2), (1, 2), (3, 4), (3, 5), (4, 5), (2, 3)]
    w, ran = lp2(6, tri, [0, 0, 0, 0, 0, 1])
    assert w.tolist() == [0, 0, 0, 1, 1, 1] and ran


def test_lp2_transfer_weights_and_ties():
    # star centre 0 on side 0 with leaves 1..6; leaves 1..5 on side 0, leaf 6
    # on side 1, leaf weights 3.  Flip: the centre (5 vs 1) stays; leaves 1..5
    # (1 vs 0) stay; leaf 6 (0 vs 1) flips to 0.  Side 0 then holds all 19 of
    # the weight: from = 0, num = 19 / 2 = 9.  Keys: leaves -1 each (ties, id
    # order), centre -6.  Prefix 3, 6, 9, 12: leaves 1, 2, 3 move to side 1.
    edges = [(0, i) for i in range(1, 7)]
    vw = [1] + [3] * 6
    w, ran = lp2(7, edges, [0, 0, 0, 0, 0, 0, 1], vw=vw)
    assert w.tolist() == [0, 1, 1, 1, 0, 0, 0] and ran


def test_partition_with_older_lp_is_valid():
    x, a, vw, aw = random_graph(500, 0.02, seed=9, wmax=4)
    for k in (2, 8):
        part, st = O.dga_partition(x, a, k, vw, aw, refiner=1)
        assert part.min() >= 0 and part.max() < k
        assert st["cut"] == O.edge_cut(x, a, aw, part)
        jet, _ = O.dga_partition(x, a, k, vw, aw)
        # same coarsening and initial partition: only the refinement differs
        _, st0 = O.dga_partition(x, a, k, vw, aw)
        assert st["init_cut"] == st0["init_cut"] and st["levels"] == st0["levels"]
