/*
 * oracle_dga.c -- serial emulation of the deterministic GPU algorithm (DGA).
 * TEST INFRASTRUCTURE ONLY (see hgp_oracle.h).
 *
 * Every decision of the reference GPU pipeline that depends on atomics order
 * or CAS races is replaced by a deterministic rule; each function says which
 * reference lines it follows and what it changes.  Float control decisions
 * are restated in exact integer arithmetic, except the per-bisection bounds of
 * the initial partition, which are computed in IEEE double with a fixed
 * operation order (the GPU computes the identical expression).
 */
#include "hgp_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <limits.h>

/* ------------------------------------------------------------------ rng */

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rng3(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return mix64(mix64(mix64(seed ^ a) ^ b) ^ c);
}

uint32_t oracle_edge_hash(uint64_t seed, int32_t level, int32_t lo, int32_t hi)
{
    return (uint32_t)(rng3(seed, 0x100000000ull + (uint64_t)(uint32_t)level,
                           (uint64_t)(uint32_t)lo, (uint64_t)(uint32_t)hi) >> 32);
}

static inline uint64_t trial_rng(uint64_t seed, uint32_t node, int32_t trial, int32_t draw)
{
    return rng3(seed, 0x200000000ull + node, (uint64_t)(uint32_t)trial, (uint64_t)(uint32_t)draw);
}

void oracle_default_opts(oracle_opts *o)
{
    o->seed = 1;
    o->ub_num = 103;
    o->ub_den = 100;
    o->match_rounds = 8;
    o->trials = 1024;
    o->fm_passes = 10;
    o->fm_kind = 0;
    o->patience = 6;
    o->refiner = 0;
}

/* hunyuangraph_partitiongraph.h:298-302 */
int32_t oracle_coarsen_threshold(int32_t k)
{
    int32_t t = k * 8;
    if (t > 1024) {
        t = k * 2;
        if (t < 1024) t = 1024;
    }
    return t;
}

/* hunyuangraph_GPU_coarsen.h:165, 1.5*tvwgt/threshold restated as
 * floor(3*tvwgt / (2*threshold)). */
int32_t oracle_maxvwgt(int64_t tvwgt, int32_t threshold)
{
    int64_t v = (3 * tvwgt) / (2 * (int64_t)threshold);
    if (v > INT32_MAX) v = INT32_MAX;
    return (int32_t)v;
}

/* ------------------------------------------------------------- matching */

static inline int64_t deg_of(const int64_t *xadj, int32_t v) { return xadj[v + 1] - xadj[v]; }

static int64_t count_matched(int32_t n, const int32_t *match)
{
    int64_t c = 0;
    for (int32_t v = 0; v < n; v++) c += match[v] >= 0;
    return c;
}

static inline int need_more(int64_t matched, int32_t n)
{
    /* sum/nvtxs < 0.75  (hunyuangraph_GPU_match.h:2577) */
    return 4 * matched < 3 * (int64_t)n;
}

/* Leaf matches (hunyuangraph_GPU_match.h:1374-1423): the unmatched degree-1
 * vertices hanging off the same hub are paired consecutively in increasing id
 * order (the reference pairs them in atomics order); a pair whose weight is
 * not below maxvwgt stays unmatched. */
static void fb_leaf(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                    int32_t maxvwgt, int32_t *match)
{
    int32_t *flag = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t v = 0; v < n; v++) flag[v] = -1;
    for (int32_t v = 0; v < n; v++) {
        if (match[v] != -1 || deg_of(xadj, v) != 1) continue;
        int32_t h = adjncy[xadj[v]];
        if (flag[h] == -1) { flag[h] = v; continue; }
        int32_t u = flag[h];
        if ((int64_t)vwgt[u] + vwgt[v] < maxvwgt) { match[u] = v; match[v] = u; }
        flag[h] = -1;
    }
    free(flag);
}

/* Isolate matches (hunyuangraph_GPU_match.h:1577-1594): degree-0 vertices in
 * increasing id order (the reference uses its atomics-ordered bin list) are
 * paired (list[2i], list[2i+1]). */
static void fb_isolate(int32_t n, const int64_t *xadj, const int32_t *vwgt, int32_t maxvwgt,
                       int32_t *match)
{
    int32_t prev = -1, idx = 0;
    for (int32_t v = 0; v < n; v++) {
        if (deg_of(xadj, v) != 0) continue;
        if (idx & 1) {
            if (match[v] == -1 && (int64_t)vwgt[v] + vwgt[prev] < maxvwgt) {
                match[v] = prev; match[prev] = v;
            }
        }
        prev = v;
        idx++;
    }
}

typedef struct { int64_t deg; int64_t sum; int32_t v; } twin_key;

static int cmp_twin(const void *a, const void *b)
{
    const twin_key *x = (const twin_key *)a, *y = (const twin_key *)b;
    if (x->deg != y->deg) return x->deg < y->deg ? -1 : 1;
    if (x->sum != y->sum) return x->sum < y->sum ? -1 : 1;
    return (x->v > y->v) - (x->v < y->v);
}

/* Twin matches (hunyuangraph_GPU_match.h:1519-1574): the reference hashes the
 * row sum (int64) of each unmatched vertex and pairs first-come vertices of
 * equal degree that probe the same slot.  Restated deterministically: group
 * unmatched vertices by exact (degree, row sum) and pair consecutive members
 * of each group in increasing id order. */
static void fb_twin(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                    int32_t maxvwgt, int32_t *match)
{
    twin_key *keys = (twin_key *)malloc(sizeof(twin_key) * (size_t)(n > 0 ? n : 1));
    int32_t nc = 0;
    for (int32_t v = 0; v < n; v++) {
        if (match[v] != -1) continue;
        int64_t s = 0;
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) s += adjncy[j];
        keys[nc].deg = deg_of(xadj, v);
        keys[nc].sum = s;
        keys[nc].v = v;
        nc++;
    }
    qsort(keys, (size_t)nc, sizeof(twin_key), cmp_twin);
    for (int32_t i = 0; i < nc;) {
        int32_t j = i;
        while (j < nc && keys[j].deg == keys[i].deg && keys[j].sum == keys[i].sum) j++;
        for (int32_t a = i; a + 1 < j; a += 2) {
            int32_t x = keys[a].v, y = keys[a + 1].v;
            if ((int64_t)vwgt[x] + vwgt[y] < maxvwgt) { match[x] = y; match[y] = x; }
        }
        i = j;
    }
    free(keys);
}

/* Relative matches (hunyuangraph_GPU_match.h:1597-1680): every unmatched
 * vertex with neighbours picks dist = the neighbour of largest degree, ties to
 * the larger edge weight, then to the first in row order.  The vertices that
 * picked the same hub are paired consecutively in increasing id order.  The
 * reference pairs every marked neighbour of the hub and settles double claims
 * with reset_match; here a hub pairs only the vertices whose dist is that
 * hub. */
static void fb_relative(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *adjwgt, const int32_t *vwgt, int32_t maxvwgt,
                        int32_t *match)
{
    int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *flag = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t v = 0; v < n; v++) {
        dist[v] = -1;
        flag[v] = -1;
        if (match[v] != -1) continue;
        int64_t bl = -1;
        int32_t bw = 0;
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            int32_t u = adjncy[j];
            int64_t l = deg_of(xadj, u);
            if (dist[v] == -1 || l > bl || (l == bl && adjwgt[j] > bw)) {
                dist[v] = u; bl = l; bw = adjwgt[j];
            }
        }
    }
    for (int32_t v = 0; v < n; v++) {
        int32_t h = dist[v];
        if (h < 0) continue;
        if (flag[h] == -1) { flag[h] = v; continue; }
        int32_t u = flag[h];
        if ((int64_t)vwgt[u] + vwgt[v] < maxvwgt) { match[u] = v; match[v] = u; }
        flag[h] = -1;
    }
    free(dist);
    free(flag);
}

/* Forced matches (hunyuangraph_GPU_match.h:1682-1716): unmatched vertices in
 * increasing id order (reference: atomics order) are paired consecutively. */
static void fb_forced(int32_t n, const int32_t *vwgt, int32_t maxvwgt, int32_t *match)
{
    int32_t prev = -1, idx = 0;
    for (int32_t v = 0; v < n; v++) {
        if (match[v] != -1) continue;
        if (idx & 1) {
            if ((int64_t)vwgt[v] + vwgt[prev] < maxvwgt) { match[v] = prev; match[prev] = v; }
        }
        prev = v;
        idx++;
    }
}

/* One fallback pass on its own, for the tests: kind 0 leaf, 1 isolate,
 * 2 twin, 3 relative, 4 forced; match (n) is updated in place (-1 = free). */
void oracle_fb_apply(int32_t kind, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                     const int32_t *adjwgt, const int32_t *vwgt, int32_t maxvwgt, int32_t *match)
{
    switch (kind) {
    case 0: fb_leaf(n, xadj, adjncy, vwgt, maxvwgt, match); break;
    case 1: fb_isolate(n, xadj, vwgt, maxvwgt, match); break;
    case 2: fb_twin(n, xadj, adjncy, vwgt, maxvwgt, match); break;
    case 3: fb_relative(n, xadj, adjncy, adjwgt, vwgt, maxvwgt, match); break;
    case 4: fb_forced(n, vwgt, maxvwgt, match); break;
    default: break;
    }
}

/* Handshake heavy-edge matching.  Replaces the level-0 CAS random matching
 * (hunyuangraph_GPU_match.h:454-510) and the level>=1 weight-sorted top-k
 * handshake (:800-861, loop :2373-2507): in each synchronous round every free
 * vertex proposes to its best free neighbour u with vwgt[u]+vwgt[v] < maxvwgt
 * (the strict bound of :823), ranking edges by (weight desc, seeded edge hash
 * desc, lo asc, hi asc) -- a total order on edges, so every locally heaviest
 * edge is mutual -- and mutual proposals are matched.  Then the fallback
 * passes of :2540-2991 run while fewer than 75% of vertices are matched. */
int64_t oracle_dga_match(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t maxvwgt,
                         uint64_t seed, int32_t level, int32_t rounds, int32_t *match)
{
    int32_t *best = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t v = 0; v < n; v++) match[v] = -1;
    for (int32_t r = 0; r < rounds; r++) {
        for (int32_t v = 0; v < n; v++) {
            best[v] = -1;
            if (match[v] != -1) continue;
            int32_t bw = -1, blo = 0, bhi = 0, bu = -1;
            uint32_t bh = 0;
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t u = adjncy[j];
                if (u == v || match[u] != -1) continue;
                if ((int64_t)vwgt[u] + vwgt[v] >= maxvwgt) continue;
                int32_t w = adjwgt[j];
                int32_t lo = u < v ? u : v, hi = u < v ? v : u;
                uint32_t h = oracle_edge_hash(seed, level, lo, hi);
                int better = bu < 0 || w > bw ||
                             (w == bw && (h > bh || (h == bh && (lo < blo || (lo == blo && hi < bhi)))));
                if (better) { bu = u; bw = w; bh = h; blo = lo; bhi = hi; }
            }
            best[v] = bu;
        }
        int64_t got = 0;
        for (int32_t v = 0; v < n; v++) {
            int32_t u = best[v];
            if (u >= 0 && best[u] == v) { match[v] = u; got++; }
        }
        if (got == 0) break;
    }
    free(best);
    int64_t matched = count_matched(n, match);
    if (need_more(matched, n)) {
        fb_leaf(n, xadj, adjncy, vwgt, maxvwgt, match);
        matched = count_matched(n, match);
    }
    if (need_more(matched, n)) {
        fb_isolate(n, xadj, vwgt, maxvwgt, match);
        matched = count_matched(n, match);
    }
    if (need_more(matched, n)) {
        fb_twin(n, xadj, adjncy, vwgt, maxvwgt, match);
        matched = count_matched(n, match);
    }
    if (need_more(matched, n)) {
        fb_relative(n, xadj, adjncy, adjwgt, vwgt, maxvwgt, match);
        matched = count_matched(n, match);
    }
    if (need_more(matched, n)) {
        fb_forced(n, vwgt, maxvwgt, match);
        matched = count_matched(n, match);
    }
    return matched;
}

/* ------------------------------------------------------ level storage */

typedef struct {
    int32_t n;
    int64_t m;
    int64_t *xadj;
    int32_t *adjncy, *adjwgt, *vwgt;
    int32_t *cmap;      /* fine -> coarse map (set when a coarser level exists) */
    int32_t *where;
} olevel;

static void level_free(olevel *L)
{
    free(L->xadj); free(L->adjncy); free(L->adjwgt); free(L->vwgt);
    free(L->cmap); free(L->where);
    memset(L, 0, sizeof(*L));
}

/* ------------------------------------------------------ initial partition */

typedef struct {
    int32_t n;
    int64_t *xadj;
    int32_t *adjncy, *adjwgt, *vwgt, *label;
    int32_t k, fpart;
    uint32_t id;
} rbnode;

static int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
static int64_t iabs64(int64_t a) { return a < 0 ? -a : a; }

/* ------------------------------------------------ device priority queue
 * Binary max-heap of vertices with a position index, restating
 * priority_queue_Insert/_Delete/_Update/_GetTop
 * (hunyuangraph_GPU_priorityqueue.h:27-225): a key moves above another when
 * it is strictly greater; a removal refills the hole from the last slot and
 * sifts that element up or down.  With `ties` set, equal keys are ordered by
 * the smaller vertex id, so the top is the unique (gain desc, id asc)
 * maximum; without it equal keys keep the reference's heap order. */
typedef struct {
    int32_t cnt;
    int64_t *key;
    int32_t *val;
} oheap;

static int heap_above(int64_t ka, int32_t va, int64_t kb, int32_t vb, int ties)
{
    return ka > kb || (ties && ka == kb && va < vb);
}

static void heap_place(oheap *h, int32_t *loc, int32_t i, int64_t key, int32_t v)
{
    h->key[i] = key;
    h->val[i] = v;
    loc[v] = i;
}

static void heap_up(oheap *h, int32_t *loc, int32_t i, int64_t key, int32_t v, int ties)
{
    while (i > 0) {
        int32_t j = (i - 1) >> 1;
        if (!heap_above(key, v, h->key[j], h->val[j], ties)) break;
        heap_place(h, loc, i, h->key[j], h->val[j]);
        i = j;
    }
    heap_place(h, loc, i, key, v);
}

static void heap_down(oheap *h, int32_t *loc, int32_t i, int64_t key, int32_t v, int ties)
{
    int32_t j;
    while ((j = 2 * i + 1) < h->cnt) {
        if (heap_above(h->key[j], h->val[j], key, v, ties)) {
            if (j + 1 < h->cnt && heap_above(h->key[j + 1], h->val[j + 1], h->key[j], h->val[j], ties)) j++;
        } else if (j + 1 < h->cnt && heap_above(h->key[j + 1], h->val[j + 1], key, v, ties)) {
            j++;
        } else {
            break;
        }
        heap_place(h, loc, i, h->key[j], h->val[j]);
        i = j;
    }
    heap_place(h, loc, i, key, v);
}

static void heap_insert(oheap *h, int32_t *loc, int32_t v, int64_t key, int ties)
{
    heap_up(h, loc, h->cnt++, key, v, ties);
}

static int32_t heap_top(oheap *h, int32_t *loc, int ties)
{
    if (h->cnt == 0) return -1;
    int32_t top = h->val[0];
    loc[top] = -1;
    if (--h->cnt > 0) heap_down(h, loc, 0, h->key[h->cnt], h->val[h->cnt], ties);
    return top;
}

static void heap_delete(oheap *h, int32_t *loc, int32_t v, int ties)
{
    int32_t i = loc[v];
    loc[v] = -1;
    if (--h->cnt > 0 && h->val[h->cnt] != v) {
        int64_t nk = h->key[h->cnt], ok = h->key[i];
        int32_t nv = h->val[h->cnt];
        if (heap_above(nk, nv, ok, v, ties)) heap_up(h, loc, i, nk, nv, ties);
        else heap_down(h, loc, i, nk, nv, ties);
    }
}

static void heap_update(oheap *h, int32_t *loc, int32_t v, int64_t key, int ties)
{
    int32_t i = loc[v];
    if (key > h->key[i]) heap_up(h, loc, i, key, v, ties);
    else heap_down(h, loc, i, key, v, ties);
}

/* Runs a sequence of operations on one queue over vertices [0, nv):
 * op 0 insert (v, key), 1 top, 2 delete v, 3 update (v, key).  tops[i]
 * receives the vertex a top returned (-1 when empty; unchanged for other
 * ops).  The final slots (vertex, key) and the position index are written
 * out; returns the final count.  Exposed for tests of the queue alone. */
int32_t oracle_heap_run(int32_t nv, int32_t nops, const int32_t *op, const int32_t *v, const int64_t *key,
                        int32_t ties, int32_t *tops, int32_t *final_val, int64_t *final_key, int32_t *final_loc)
{
    oheap h;
    h.cnt = 0;
    h.key = final_key;
    h.val = final_val;
    for (int32_t i = 0; i < nv; i++) final_loc[i] = -1;
    for (int32_t i = 0; i < nops; i++) {
        switch (op[i]) {
        case 0: heap_insert(&h, final_loc, v[i], key[i], ties); break;
        case 1: tops[i] = heap_top(&h, final_loc, ties); break;
        case 2: heap_delete(&h, final_loc, v[i], ties); break;
        default: heap_update(&h, final_loc, v[i], key[i], ties); break;
        }
    }
    return h.cnt;
}

/* FM passes driven by two device priority queues, one per side
 * (FM_2way_cut_refinement, hunyuangraph_GPU_initialpartition.h:160-300):
 * every pass inserts the boundary vertices in id order with key ed - id,
 * takes the top of the heavier side's queue, and keeps the queues current
 * as the neighbours' gains change (delete when a neighbour leaves the
 * boundary, update, insert when one joins it; moved vertices are never
 * queued again in the pass), then rolls back to the best prefix.
 *   fm_kind 1: keys ordered (gain desc, id asc) and the acceptance rule of
 *     the live wave FM (:3244-3505, the part bounds): the move sequence is
 *     exactly that of the arg-max scan, found in O(log n) per step.
 *   fm_kind 2: the reference's queue semantics (equal gains in heap order)
 *     and its balance rule (|tpwgts0 - pwgts0| <= origdiff + avgvwgt,
 *     avgvwgt = min(tvwgt/20, 2 tvwgt/n)), with the move's weights applied
 *     before the test as in hunyuangraph_CPU_2wayrefine.h:133-145 (the
 *     device version tests the weights before the move, :196-203).
 * ed/id/bnd/where/pw are the trial's state after growing; returns the cut. */
static int64_t fm_heap_passes(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                              const int32_t *adjwgt, int64_t tvwgt, int64_t onemax, int64_t ideal0,
                              int32_t fm_passes, int fm_kind, int64_t cut, uint8_t *where, int64_t *ed,
                              int64_t *id, uint8_t *bnd, uint8_t *moved, int32_t *swaps, int64_t *pw)
{
    const int ties = fm_kind == 1;
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * n + 2));
    int32_t *vals = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 2));
    int32_t *loc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    oheap q[2];
    q[0].key = keys; q[0].val = vals;
    q[1].key = keys + n + 1; q[1].val = vals + n + 1;
    for (int32_t v = 0; v < n; v++) loc[v] = -1;
    int64_t ideal[2] = {ideal0, tvwgt - ideal0};
    int64_t limit = imin64(imax64(n / 100, 15), 100);
    int64_t avgvwgt = imin64((pw[0] + pw[1]) / 20, 2 * (pw[0] + pw[1]) / n);
    int64_t origdiff = iabs64(ideal[0] - pw[0]);
    for (int32_t pass = 0; pass < fm_passes; pass++) {
        int64_t mincutorder = -1;
        int64_t newcut = cut, mincut = cut, initcut = cut;
        int64_t mindiff = iabs64(ideal[0] - pw[0]);
        q[0].cnt = q[1].cnt = 0;
        for (int32_t v = 0; v < n; v++)
            if (bnd[v]) heap_insert(&q[where[v]], loc, v, ed[v] - id[v], ties);
        int32_t nswaps;
        for (nswaps = 0; nswaps < n; nswaps++) {
            int from = (ideal[0] - pw[0] < ideal[1] - pw[1]) ? 0 : 1;
            int to = from ^ 1;
            int32_t v = heap_top(&q[from], loc, ties);
            if (v == -1) break;
            int64_t g = ed[v] - id[v];
            newcut -= g;
            int accept;
            if (fm_kind == 1) {
                int64_t dd = iabs64(ideal[0] - pw[0]);
                accept = (newcut < mincut && pw[to] + vwgt[v] <= onemax && pw[from] - vwgt[v] <= onemax) ||
                         (newcut == mincut && dd < mindiff);
                if (accept) mindiff = dd;
                pw[to] += vwgt[v];
                pw[from] -= vwgt[v];
            } else {
                pw[to] += vwgt[v];
                pw[from] -= vwgt[v];
                int64_t dd = iabs64(ideal[0] - pw[0]);
                accept = (newcut < mincut && dd <= origdiff + avgvwgt) || (newcut == mincut && dd < mindiff);
                if (accept) mindiff = dd;
            }
            if (accept) {
                mincut = newcut;
                mincutorder = nswaps;
            } else if (nswaps - mincutorder > limit) {
                newcut += g;
                pw[from] += vwgt[v];
                pw[to] -= vwgt[v];
                break;
            }
            where[v] = (uint8_t)to;
            moved[v] = 1;
            swaps[nswaps] = v;
            { int64_t t = ed[v]; ed[v] = id[v]; id[v] = t; }
            if (ed[v] == 0 && xadj[v] < xadj[v + 1]) bnd[v] = 0;
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t x = adjncy[j];
                int64_t kw = (where[x] == to) ? (int64_t)adjwgt[j] : -(int64_t)adjwgt[j];
                id[x] += kw;
                ed[x] -= kw;
                if (bnd[x]) {
                    if (ed[x] == 0) {
                        bnd[x] = 0;
                        if (!moved[x]) heap_delete(&q[where[x]], loc, x, ties);
                    } else if (!moved[x]) {
                        heap_update(&q[where[x]], loc, x, ed[x] - id[x], ties);
                    }
                } else if (ed[x] > 0) {
                    bnd[x] = 1;
                    if (!moved[x]) heap_insert(&q[where[x]], loc, x, ed[x] - id[x], ties);
                }
            }
        }
        for (int32_t i = 0; i < nswaps; i++) moved[swaps[i]] = 0;
        for (int32_t i = 0; i < q[0].cnt; i++) loc[q[0].val[i]] = -1;
        for (int32_t i = 0; i < q[1].cnt; i++) loc[q[1].val[i]] = -1;
        for (nswaps--; nswaps > mincutorder; nswaps--) {
            int32_t v = swaps[nswaps];
            int from = where[v];
            int to = from ^ 1;
            where[v] = (uint8_t)to;
            { int64_t t = ed[v]; ed[v] = id[v]; id[v] = t; }
            if (ed[v] == 0 && bnd[v] && xadj[v] < xadj[v + 1]) bnd[v] = 0;
            else if (ed[v] > 0 && !bnd[v]) bnd[v] = 1;
            pw[to] += vwgt[v];
            pw[from] -= vwgt[v];
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t x = adjncy[j];
                int64_t kw = (where[x] == from) ? -(int64_t)adjwgt[j] : (int64_t)adjwgt[j];
                id[x] += kw;
                ed[x] -= kw;
                if (bnd[x] && ed[x] == 0) bnd[x] = 0;
                else if (!bnd[x] && ed[x] > 0) bnd[x] = 1;
            }
        }
        cut = mincut;
        if (mincutorder <= 0 || mincut == initcut) break;
    }
    free(keys); free(vals); free(loc);
    return cut;
}

/* One GGGP + FM trial, one wave per trial on the GPU
 * (hunyuangraph_gpu_BFS_warp_2wayrefine_noedit[_sampling],
 * hunyuangraph_GPU_initialpartition.h:2426-3659).  Start vertices come from a
 * counter-based generator instead of curand XORWOW; the FM arg-max breaks ties
 * to the smallest vertex id instead of the lowest lane. */
int64_t oracle_dga_trial(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt,
                         int64_t tvwgt, int64_t onemin, int64_t onemax, int64_t ideal0,
                         int32_t ntrials, int32_t trial, uint64_t seed, uint32_t node_id,
                         int32_t fm_passes, int32_t fm_kind, uint8_t *where)
{
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int64_t *ed = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *id = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    uint8_t *moved = (uint8_t *)calloc((size_t)n, 1);
    uint8_t *bnd = (uint8_t *)calloc((size_t)n, 1);
    int64_t pw[2];
    int32_t draw = 0;

    for (int32_t v = 0; v < n; v++) where[v] = 1;
    pw[0] = 0;
    pw[1] = tvwgt;

    /* start vertex (:3050-3108) */
    int32_t start;
    if (ntrials >= n) {
        start = trial;
    } else {
        int32_t len = n / ntrials, rem = n % ntrials, left;
        if (trial < rem) { len++; left = trial * len; }
        else left = rem * (len + 1) + (trial - rem) * len;
        start = left + (int32_t)(trial_rng(seed, node_id, trial, draw++) % (uint64_t)len);
    }

    /* BFS growing (:3109-3170) */
    int32_t first = 0, last = 1, nleft = n - 1, drain = 0;
    queue[0] = start;
    moved[start] = 1;
    for (;;) {
        if (first == last) {
            if (nleft == 0 || drain) break;
            int32_t r = (int32_t)(trial_rng(seed, node_id, trial, draw++) % (uint64_t)nleft);
            int32_t v;
            for (v = 0; v < n; v++) {
                if (moved[v] == 0) {
                    if (r == 0) break;
                    r--;
                }
            }
            queue[0] = v;
            moved[v] = 1;
            first = 0;
            last = 1;
            nleft--;
        }
        int32_t v = queue[first++];
        if (pw[0] > 0 && pw[1] - vwgt[v] < onemin) { drain = 1; continue; }
        where[v] = 0;
        pw[0] += vwgt[v];
        pw[1] -= vwgt[v];
        if (pw[1] <= onemax) break;
        drain = 0;
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            int32_t x = adjncy[j];
            if (moved[x] == 0) { queue[last++] = x; moved[x] = 1; nleft--; }
        }
    }

    /* ed/id/bnd (compute_v_ed_id_bnd, :23-62) and cut */
    int64_t cut = 0;
    for (int32_t v = 0; v < n; v++) {
        int64_t e = 0, i = 0;
        if (xadj[v] == xadj[v + 1]) { bnd[v] = 1; ed[v] = 0; id[v] = 0; continue; }
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            if (where[adjncy[j]] != where[v]) e += adjwgt[j]; else i += adjwgt[j];
        }
        ed[v] = e; id[v] = i; bnd[v] = e > 0;
        cut += e;
    }
    cut /= 2;
    memset(moved, 0, (size_t)n);

    if (fm_kind != 0) {
        int64_t pwh[2] = {pw[0], pw[1]};
        cut = fm_heap_passes(n, xadj, adjncy, vwgt, adjwgt, tvwgt, onemax, ideal0, fm_passes, fm_kind, cut,
                             where, ed, id, bnd, moved, queue, pwh);
        free(queue); free(ed); free(id); free(moved); free(bnd);
        return cut;
    }

    /* FM passes (:3244-3505) */
    int32_t *swaps = queue;
    int64_t ideal[2] = {ideal0, tvwgt - ideal0};
    int64_t limit = imin64(imax64(n / 100, 15), 100);
    for (int32_t pass = 0; pass < fm_passes; pass++) {
        int64_t mincutorder = -1;
        int64_t newcut = cut, mincut = cut, initcut = cut;
        int64_t mindiff = iabs64(ideal[0] - pw[0]);
        int32_t nswaps;
        for (nswaps = 0; nswaps < n; nswaps++) {
            int from = (ideal[0] - pw[0] < ideal[1] - pw[1]) ? 0 : 1;
            int to = from ^ 1;
            int32_t v = -1;
            int64_t bg = 0;
            for (int32_t u = 0; u < n; u++) {
                if (where[u] != from || !bnd[u] || moved[u]) continue;
                int64_t g = ed[u] - id[u];
                if (v == -1 || g > bg) { v = u; bg = g; }
            }
            if (v == -1) break;
            newcut -= ed[v] - id[v];
            if ((newcut < mincut && pw[to] + vwgt[v] <= onemax && pw[from] - vwgt[v] <= onemax) ||
                (newcut == mincut && iabs64(ideal[0] - pw[0]) < mindiff)) {
                mincut = newcut;
                mindiff = iabs64(ideal[0] - pw[0]);
                mincutorder = nswaps;
            } else if (nswaps - mincutorder > limit) {
                newcut += ed[v] - id[v];
                break;
            }
            pw[to] += vwgt[v];
            pw[from] -= vwgt[v];
            where[v] = (uint8_t)to;
            moved[v] = 1;
            swaps[nswaps] = v;
            { int64_t t = ed[v]; ed[v] = id[v]; id[v] = t; }
            if (ed[v] == 0 && xadj[v] < xadj[v + 1]) bnd[v] = 0;
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t x = adjncy[j];
                int64_t kw = (where[x] == from) ? -(int64_t)adjwgt[j] : (int64_t)adjwgt[j];
                id[x] += kw;
                ed[x] -= kw;
                if (bnd[x]) { if (ed[x] == 0) bnd[x] = 0; }
                else if (ed[x] > 0) bnd[x] = 1;
            }
        }
        for (int32_t i = 0; i < nswaps; i++) moved[swaps[i]] = 0;
        for (nswaps--; nswaps > mincutorder; nswaps--) {
            int32_t v = swaps[nswaps];
            int from = where[v];
            int to = from ^ 1;
            where[v] = (uint8_t)to;
            { int64_t t = ed[v]; ed[v] = id[v]; id[v] = t; }
            if (ed[v] == 0 && bnd[v] && xadj[v] < xadj[v + 1]) bnd[v] = 0;
            else if (ed[v] > 0 && !bnd[v]) bnd[v] = 1;
            pw[to] += vwgt[v];
            pw[from] -= vwgt[v];
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t x = adjncy[j];
                int64_t kw = (where[x] == from) ? -(int64_t)adjwgt[j] : (int64_t)adjwgt[j];
                id[x] += kw;
                ed[x] -= kw;
                if (bnd[x] && ed[x] == 0) bnd[x] = 0;
                else if (!bnd[x] && ed[x] > 0) bnd[x] = 1;
            }
        }
        cut = mincut;
        if (mincutorder <= 0 || mincut == initcut) break;
    }

    free(queue); free(ed); free(id); free(moved); free(bnd);
    return cut;
}

/* Per-bisection bounds: ub = 1.03^(1/ln k) (hunyuangraph_GPU_initialpartition.h:4476),
 * onemax = ub*tvwgt*tp1, onemin = tvwgt*tp1/ub (:3986-3987) with tp1 = kr/k
 * for uniform target weights.  Same double expression on the GPU. */
static void bisect_bounds(int64_t tv, int32_t k, double ub, int64_t *onemin, int64_t *onemax,
                          int64_t *ideal0)
{
    int32_t kl = k >> 1, kr = k - kl;
    double t1 = ((double)tv * (double)kr) / (double)k;
    *onemax = (int64_t)floor(ub * t1);
    *onemin = (int64_t)floor(t1 / ub);
    *ideal0 = (tv * kl) / k;
}

/* One bisection node of the recursion (hunyuangraph_gpu_RecursiveBisection,
 * :3913-4385): trials, arg-min cut (first minimum wins, :3713-3812), write
 * answer through label (:3896-3902), split (hunyuangraph_GPU_splitgraph.h:802-944)
 * and recurse right then left (:4367-4382; the order does not affect results). */
static void rb_node(rbnode *nd, const oracle_opts *o, double ub, int32_t *answer)
{
    int32_t n = nd->n, k = nd->k;
    if (n == 0) return;
    int64_t tv = 0;
    for (int32_t v = 0; v < n; v++) tv += nd->vwgt[v];
    int64_t onemin, onemax, ideal0;
    bisect_bounds(tv, k, ub, &onemin, &onemax, &ideal0);
    int32_t T = o->trials < n ? o->trials : n;
    uint8_t *w = (uint8_t *)malloc((size_t)n);
    uint8_t *bw = (uint8_t *)malloc((size_t)n);
    int64_t bcut = -1;
    for (int32_t t = 0; t < T; t++) {
        int64_t c = oracle_dga_trial(n, nd->xadj, nd->adjncy, nd->vwgt, nd->adjwgt, tv, onemin,
                                     onemax, ideal0, T, t, o->seed, nd->id, o->fm_passes, o->fm_kind, w);
        if (bcut < 0 || c < bcut) { bcut = c; memcpy(bw, w, (size_t)n); }
    }
    for (int32_t v = 0; v < n; v++) answer[nd->label[v]] = nd->fpart + bw[v];
    if (k > 2) {
        /* order-preserving split into the two induced subgraphs */
        int32_t *ren = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        int32_t cnt[2] = {0, 0};
        int64_t ecnt[2] = {0, 0};
        for (int32_t v = 0; v < n; v++) {
            int s = bw[v];
            ren[v] = cnt[s]++;
            for (int64_t j = nd->xadj[v]; j < nd->xadj[v + 1]; j++) ecnt[s] += bw[nd->adjncy[j]] == s;
        }
        rbnode ch[2];
        for (int s = 0; s < 2; s++) {
            ch[s].n = cnt[s];
            ch[s].xadj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cnt[s] + 1));
            ch[s].adjncy = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ecnt[s] + 1));
            ch[s].adjwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ecnt[s] + 1));
            ch[s].vwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt[s] + 1));
            ch[s].label = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt[s] + 1));
            ch[s].xadj[0] = 0;
            cnt[s] = 0;
            ecnt[s] = 0;
        }
        ch[0].k = k >> 1; ch[0].fpart = nd->fpart; ch[0].id = 2 * nd->id;
        ch[1].k = k - (k >> 1); ch[1].fpart = nd->fpart + (k >> 1); ch[1].id = 2 * nd->id + 1;
        for (int32_t v = 0; v < n; v++) {
            int s = bw[v];
            rbnode *c = &ch[s];
            for (int64_t j = nd->xadj[v]; j < nd->xadj[v + 1]; j++) {
                int32_t x = nd->adjncy[j];
                if (bw[x] != s) continue;
                c->adjncy[ecnt[s]] = ren[x];
                c->adjwgt[ecnt[s]] = nd->adjwgt[j];
                ecnt[s]++;
            }
            c->vwgt[cnt[s]] = nd->vwgt[v];
            c->label[cnt[s]] = nd->label[v];
            cnt[s]++;
            c->xadj[cnt[s]] = ecnt[s];
        }
        free(ren);
        for (int s = 1; s >= 0; s--) {
            if (ch[s].k >= 2) rb_node(&ch[s], o, ub, answer);
            free(ch[s].xadj); free(ch[s].adjncy); free(ch[s].adjwgt); free(ch[s].vwgt); free(ch[s].label);
        }
    }
    free(w);
    free(bw);
}

void oracle_dga_initpart(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                         const oracle_opts *opts, int32_t *where)
{
    for (int32_t v = 0; v < n; v++) where[v] = 0;
    if (k <= 1 || n == 0) return;
    double ub = pow((double)opts->ub_num / (double)opts->ub_den, 1.0 / log((double)k));
    rbnode root;
    root.n = n;
    root.xadj = (int64_t *)xadj;
    root.adjncy = (int32_t *)adjncy;
    root.adjwgt = (int32_t *)adjwgt;
    root.vwgt = (int32_t *)vwgt;
    root.label = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; v++) root.label[v] = v;
    root.k = k;
    root.fpart = 0;
    root.id = 1;
    rb_node(&root, opts, ub, where);
    free(root.label);
}

/* ------------------------------------------------------ Jet refinement */

static const int32_t kMaxBuckets = 50, kMidBucket = 25;

/* gain_bucket (hunyuangraph_GPU_krefine.h:2477-2513) restated in integers:
 * gain/vwgt is multiplied (or divided) by 1.5 until it crosses 1, counting
 * the steps; x*1.5^j < 1  <=>  |g|*3^j < w*2^j. */
int32_t oracle_gain_bucket(int64_t gain, int64_t vwgt)
{
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    int64_t a = -gain, b = vwgt;
    int32_t t = kMidBucket;
    if (a < b) {
        /* while x < 1: x *= 1.5, t-- ; clamp t >= 2 (t stops mattering below 2) */
        while (a < b) {
            a *= 3; b *= 2; t--;
            if (t < 2) break;
        }
        if (t < 2) t = 2;
    } else {
        /* while x > 1: x /= 1.5, t++ ; t > 50 -> 49 */
        while (a > b) {
            a *= 2; b *= 3; t++;
            if (t > kMaxBuckets) break;
        }
        if (t > kMaxBuckets) t = kMaxBuckets - 1;
    }
    return t;
}

typedef struct {
    int32_t n, k;
    const int64_t *xadj;
    const int32_t *adjncy, *adjwgt, *vwgt;
    int64_t *conn;      /* dense k scratch */
    int32_t *touched;
} connctx;

/* Accumulate conn[q] = sum of edge weights from v into part q. Returns the
 * number of touched parts (listed in ctx->touched).  Caller clears. */
static int32_t conn_build(connctx *c, int32_t v, const int32_t *where)
{
    int32_t nt = 0;
    for (int64_t j = c->xadj[v]; j < c->xadj[v + 1]; j++) {
        int32_t q = where[c->adjncy[j]];
        if (c->conn[q] == 0) c->touched[nt++] = q;
        c->conn[q] += c->adjwgt[j];
    }
    return nt;
}

static void conn_clear(connctx *c, int32_t nt)
{
    for (int32_t i = 0; i < nt; i++) c->conn[c->touched[i]] = 0;
}

/* jetlp (hunyuangraph_GPU_krefine.h:2250-2471): select_dest_part (:1995-2046)
 * without the dest cache (every vertex is evaluated each step),
 * filter_potential_vertex (:2114-2135), afterburner_heuristic (:2137-2178),
 * lock handling (:2213-2221).  filter ratio r = num/den; floor(r*p_conn) is
 * computed exactly as (num*p_conn)/den. */
int64_t oracle_jet_lp_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, int32_t k, const int32_t *where,
                           int32_t filter_num, int32_t filter_den,
                           uint8_t *lock, int32_t *dest)
{
    connctx c = {n, k, xadj, adjncy, adjwgt, NULL, NULL, NULL};
    c.conn = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    c.touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int64_t *pregain = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t ncand = 0;
    for (int32_t v = 0; v < n; v++) {
        int32_t p = where[v];
        int32_t nt = conn_build(&c, v, where);
        int64_t pc = c.conn[p], bc = 0;
        int32_t best = p;
        for (int32_t q = 0; q < k; q++) {
            if (q == p) continue;
            if (c.conn[q] > bc) { bc = c.conn[q]; best = q; }
        }
        int64_t gain = 0;
        if (best != p) {
            if (bc >= pc || (pc - bc) < (filter_num * pc) / filter_den) gain = bc - pc;
            else best = p;
        }
        conn_clear(&c, nt);
        dest[v] = best;
        if (best != p && lock[v] == 0) { pregain[v] = gain; ncand++; }
        else pregain[v] = INT64_MIN;
    }
    if (ncand == 0) {
        for (int32_t v = 0; v < n; v++) dest[v] = where[v];
        free(c.conn); free(c.touched); free(pregain);
        return 0;
    }
    uint8_t *sel = (uint8_t *)calloc((size_t)n, 1);
    for (int32_t v = 0; v < n; v++) {
        if (pregain[v] == INT64_MIN) continue;
        int32_t p = where[v], best = dest[v];
        int64_t ig = pregain[v], change = 0;
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            int32_t u = adjncy[j];
            int64_t ng = pregain[u];
            if (ng > ig || (ng == ig && u < v)) {
                int32_t w = adjwgt[j];
                int32_t nd = dest[u];
                if (nd == p) change -= w; else if (nd == best) change += w;
                int32_t nw = where[u];
                if (nw == p) change += w; else if (nw == best) change -= w;
            }
        }
        if (ig + change >= 0) sel[v] = 1;
    }
    int64_t moves = 0;
    for (int32_t v = 0; v < n; v++) lock[v] = 0;
    for (int32_t v = 0; v < n; v++) {
        if (sel[v]) { lock[v] = 1; moves++; }
        else dest[v] = where[v];
    }
    free(sel);
    free(c.conn); free(c.touched); free(pregain);
    return moves;
}

typedef struct { int64_t key; int32_t v; } ckey;

static int cmp_ckey(const void *a, const void *b)
{
    const ckey *x = (const ckey *)a, *y = (const ckey *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->v > y->v) - (x->v < y->v);
}

/* jetrw (hunyuangraph_GPU_krefine.h:3381-3562, kernels :3136-3366).  The
 * reference orders a part's candidates by (gain bucket, section = v mod
 * sections, atomics order); here by (gain bucket, vertex id).  A vertex is
 * evicted iff the weight of the candidates ahead of it in its part is below
 * pwgts[p]-maxwgt (filter_scores_below_cutoff, :3344-3366). */
static int64_t jet_rw_step(connctx *c, const int32_t *where, const int64_t *pwgts,
                           int64_t maxwgt, int64_t opt, int32_t *dest)
{
    int32_t n = c->n, k = c->k;
    int64_t maxdest = (99 * maxwgt) / 100;
    if (maxdest < maxwgt - 100) maxdest = maxwgt - 100;
    int32_t *under = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int32_t nu = 0;
    for (int32_t q = 0; q < k; q++) if (pwgts[q] < maxdest) under[nu++] = q;
    ckey *cand = (ckey *)malloc(sizeof(ckey) * (size_t)(n > 0 ? n : 1));
    int32_t nc = 0;
    for (int32_t v = 0; v < n; v++) {
        dest[v] = where[v];
        int32_t p = where[v];
        if (!(pwgts[p] > maxwgt && 2 * (int64_t)c->vwgt[v] < 3 * (pwgts[p] - opt))) continue;
        int32_t nt = conn_build(c, v, where);
        int32_t best = -1;
        int64_t bc = 0;
        for (int32_t q = 0; q < k; q++) {
            if (pwgts[q] < maxwgt && c->conn[q] > bc) { bc = c->conn[q]; best = q; }
        }
        int64_t pc = c->conn[p], gain;
        conn_clear(c, nt);
        if (best >= 0) gain = bc - pc;
        else if (nu > 0) { best = under[v % nu]; gain = -pc; }
        else continue;
        if (best == p) continue;
        dest[v] = best;
        /* gain_bucket may return kMaxBuckets itself: the reference's bucket id
         * (:3243-3266) then lies in the next part's first bucket and counts
         * against that part's cut-off; here the vertex stays in its own
         * part's last bucket (jetrs drops such vertices, :2549). */
        int32_t b = oracle_gain_bucket(gain, c->vwgt[v]);
        if (b >= kMaxBuckets) b = kMaxBuckets - 1;
        cand[nc].key = (int64_t)p * kMaxBuckets + b;
        cand[nc].v = v;
        nc++;
    }
    qsort(cand, (size_t)nc, sizeof(ckey), cmp_ckey);
    int64_t moves = 0;
    int64_t run = 0;
    int32_t curp = -1;
    for (int32_t i = 0; i < nc; i++) {
        int32_t v = cand[i].v, p = where[v];
        if (p != curp) { curp = p; run = 0; }
        if (run < pwgts[p] - maxwgt) moves++;
        else dest[v] = where[v];
        run += c->vwgt[v];
    }
    free(under);
    free(cand);
    return moves;
}

/* cookie_cutter (hunyuangraph_GPU_krefine.h:2710-2745): the evicted list,
 * whose vertex weights prefix to bs[0..ne], is cut into consecutive slices
 * [es[p], ee[p]), one per part in part order; each slice starts where the
 * previous ended and ends at the prefix index closest to
 * bs[start] + max(maxwgt - pwgts[p], 0), found by the reference's binary
 * search (ties to the lower index). */
void oracle_cookie_cutter(int32_t k, int32_t ne, const int64_t *bs, int64_t maxwgt,
                          const int64_t *pwgts, int32_t *es, int32_t *ee)
{
    int32_t start = 0;
    for (int32_t p = 0; p < k; p++) {
        es[p] = start;
        int64_t cap = maxwgt > pwgts[p] ? maxwgt - pwgts[p] : 0;
        int64_t find = bs[start] + cap;
        int32_t lo = start, hi = ne;
        while (lo + 1 < hi) {
            int32_t mid = (lo + hi) / 2;
            if (bs[mid] >= find) hi = mid; else lo = mid;
        }
        int32_t end = (iabs64(bs[hi] - find) < iabs64(bs[lo] - find)) ? hi : lo;
        ee[p] = end;
        start = end;
    }
}

/* jetrs (hunyuangraph_GPU_krefine.h:2774-3134, kernels :2515-2766).  The
 * reference's loss term averages connectivity over parts gated by
 * poverload[p] of the vertex's own (overloaded) part, so it reduces to
 * -conn[p]; kept as written.  Candidates are ordered by (part, bucket, id);
 * each overloaded part evicts the prefix whose preceding weight is below its
 * excess (find_score_cutoffs/filter_below_cutoffs); the evicted list is then
 * cut into consecutive slices, one per part, by the serial cookie cutter
 * (:2710-2745) and select_dest_parts_rs (:2747-2766). */
static int64_t jet_rs_step(connctx *c, const int32_t *where, const int64_t *pwgts,
                           int64_t maxwgt, int64_t opt, int32_t *dest)
{
    int32_t n = c->n, k = c->k;
    ckey *cand = (ckey *)malloc(sizeof(ckey) * (size_t)(n > 0 ? n : 1));
    int32_t nc = 0;
    for (int32_t v = 0; v < n; v++) {
        dest[v] = where[v];
        int32_t p = where[v];
        if (!(pwgts[p] > maxwgt)) continue;
        int64_t pc = 0;
        for (int64_t j = c->xadj[v]; j < c->xadj[v + 1]; j++)
            if (where[c->adjncy[j]] == p) pc += c->adjwgt[j];
        int32_t b = oracle_gain_bucket(-pc, c->vwgt[v]);
        if (b < kMaxBuckets && (int64_t)c->vwgt[v] < 2 * (pwgts[p] - opt)) {
            cand[nc].key = (int64_t)p * kMaxBuckets + b;
            cand[nc].v = v;
            nc++;
        }
    }
    if (nc == 0) { free(cand); return 0; }
    qsort(cand, (size_t)nc, sizeof(ckey), cmp_ckey);
    int32_t *ev = (int32_t *)malloc(sizeof(int32_t) * (size_t)nc);
    int32_t ne = 0;
    int64_t run = 0;
    int32_t curp = -1;
    for (int32_t i = 0; i < nc; i++) {
        int32_t v = cand[i].v, p = where[v];
        if (p != curp) { curp = p; run = 0; }
        if (run < pwgts[p] - maxwgt) ev[ne++] = v;
        run += c->vwgt[v];
    }
    if (ne == 0) { free(cand); free(ev); return 0; }
    int64_t *bs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ne + 1));
    bs[0] = 0;
    for (int32_t i = 0; i < ne; i++) bs[i + 1] = bs[i] + c->vwgt[ev[i]];
    int32_t *es = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int32_t *ee = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    oracle_cookie_cutter(k, ne, bs, maxwgt, pwgts, es, ee);
    int64_t moves = 0;
    int32_t p = 0;
    for (int32_t i = 0; i < ne; i++) {
        while (p + 1 < k && es[p + 1] <= i) p++;
        int32_t v = ev[i];
        if (i < ee[p] && p != where[v]) { dest[v] = p; moves++; }
    }
    free(cand); free(ev); free(bs); free(es); free(ee);
    return moves;
}

/* One weak (jetrw) or strong (jetrs) rebalancing step on its own, for the
 * tests: pwgts, maxwgt and opt as the refinement loop holds them; dest[v]
 * receives the chosen part (where[v] if v stays).  Returns the moves. */
int64_t oracle_jet_rw_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, const int32_t *vwgt, int32_t k,
                           const int32_t *where, const int64_t *pwgts, int64_t maxwgt,
                           int64_t opt, int32_t *dest)
{
    connctx c = {n, k, xadj, adjncy, adjwgt, vwgt, NULL, NULL};
    c.conn = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    c.touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int64_t m = jet_rw_step(&c, where, pwgts, maxwgt, opt, dest);
    free(c.conn); free(c.touched);
    return m;
}

int64_t oracle_jet_rs_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, const int32_t *vwgt, int32_t k,
                           const int32_t *where, const int64_t *pwgts, int64_t maxwgt,
                           int64_t opt, int32_t *dest)
{
    connctx c = {n, k, xadj, adjncy, adjwgt, vwgt, NULL, NULL};
    c.conn = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    c.touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int64_t m = jet_rs_step(&c, where, pwgts, maxwgt, opt, dest);
    free(c.conn); free(c.touched);
    return m;
}

/* perform_moves (hunyuangraph_GPU_krefine.h:4034-4155): apply dest[], update
 * pwgts and the cut.  The reference keeps the cut incrementally from
 * connectivity lists; here the exact change is summed over the rows of moved
 * vertices (edges between two moved vertices are seen from both ends). */
static int64_t apply_moves(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, const int32_t *vwgt, int32_t *where,
                           const int32_t *dest, int64_t *pwgts)
{
    int64_t d2 = 0;
    for (int32_t v = 0; v < n; v++) {
        if (dest[v] == where[v]) continue;
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            int32_t u = adjncy[j];
            int um = dest[u] != where[u];
            int32_t nu = um ? dest[u] : where[u];
            int64_t before = where[v] != where[u], after = dest[v] != nu;
            d2 += (after - before) * (int64_t)adjwgt[j] * (um ? 1 : 2);
        }
    }
    for (int32_t v = 0; v < n; v++) {
        if (dest[v] == where[v]) continue;
        pwgts[where[v]] -= vwgt[v];
        pwgts[dest[v]] += vwgt[v];
    }
    for (int32_t v = 0; v < n; v++) where[v] = dest[v];
    return d2 / 2;
}

static int64_t max_of(const int64_t *a, int32_t k)
{
    int64_t m = a[0];
    for (int32_t i = 1; i < k; i++) if (a[i] > m) m = a[i];
    return m;
}

/* The acceptance rule of k_refine (hunyuangraph_GPU_krefine.h:4805-4839)
 * after an iteration that left cut cur_cut and heaviest part curmax: an
 * unbalanced best (best_max > maxwgt) is replaced by any state with a lighter
 * heaviest part, and the patience counter restarts; otherwise a smaller cut
 * is kept if the state is balanced or no less balanced than the best, and
 * the counter restarts only if the cut fell by more than 0.1 % (the
 * reference's tol 0.999, as 1000 cur < 999 best).  The reference compares
 * max(pwgts)/opt; the heaviest part weight orders identically.  Returns 1
 * when the state becomes the best. */
int oracle_jet_accept(int64_t cur_cut, int64_t curmax, int64_t maxwgt, int64_t *best_cut,
                      int64_t *best_max, int32_t *count)
{
    if (*best_max > maxwgt && curmax < *best_max) {
        *best_max = curmax;
        *best_cut = cur_cut;
        *count = 0;
        return 1;
    }
    if (cur_cut < *best_cut && (curmax <= maxwgt || curmax <= *best_max)) {
        if (1000 * cur_cut < 999 * *best_cut) *count = 0;
        *best_max = curmax;
        *best_cut = cur_cut;
        return 1;
    }
    return 0;
}

/* k_refine (hunyuangraph_GPU_krefine.h:4434-4882) with the driver limits of
 * hunyuangraph_GPU_uncoarsen.h:428-437: opt = tvwgt/k, maxwgt = the exact
 * ub bound floor(ub_num*tvwgt/(ub_den*k)) (the reference uses (int)(opt*1.04)).
 * Imbalance comparisons of the acceptance rule (:4805-4839) are made on the
 * maximum part weight, which orders identically to max(pwgts)/opt. */
/* Trace of the refinement loops, in the records the GPU library's trace
 * uses (level vertices, iterations done, mode of the coming iteration, cut
 * so far, loop ended); levels of <= 512 vertices record nothing, as there. */
static int64_t *g_trace_buf, g_trace_cap, g_trace_len;

void oracle_jet_trace(int64_t *buf, int64_t cap)
{
    g_trace_buf = buf;
    g_trace_cap = buf ? cap : 0;
    g_trace_len = 0;
}

int64_t oracle_jet_trace_len(void) { return g_trace_len; }

static int64_t g_dump_n = -1;
static int32_t g_dump_iter = -1, *g_dump_where, *g_dump_dest, g_dumped;

void oracle_jet_dump(int64_t n, int32_t iter, int32_t *where, int32_t *dest)
{
    g_dump_n = n;
    g_dump_iter = iter;
    g_dump_where = where;
    g_dump_dest = dest;
    g_dumped = 0;
}

int32_t oracle_jet_dumped(void) { return g_dumped; }

static void trace_rec(int32_t n, int32_t iters, int32_t mode, int64_t cut, int32_t done)
{
    if (!g_trace_buf || n <= 512) return;
    int64_t r[5] = {n, iters, mode, cut, done};
    for (int i = 0; i < 5; i++) {
        if (g_trace_len < g_trace_cap) g_trace_buf[g_trace_len] = r[i];
        g_trace_len++;
    }
}

int32_t oracle_dga_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                          const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                          int64_t tvwgt, int32_t is_finest, const oracle_opts *opts,
                          int32_t *where, int64_t *cut)
{
    int64_t opt = tvwgt / k;
    int64_t maxwgt = ((int64_t)opts->ub_num * tvwgt) / ((int64_t)opts->ub_den * k);
    int64_t *pwgts = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    for (int32_t v = 0; v < n; v++) pwgts[where[v]] += vwgt[v];
    int32_t *best_where = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *dest = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint8_t *lock = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    memcpy(best_where, where, sizeof(int32_t) * (size_t)n);
    int64_t best_cut = *cut, best_max = max_of(pwgts, k), cur_cut = *cut;
    connctx c = {n, k, xadj, adjncy, adjwgt, vwgt, NULL, NULL};
    c.conn = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    c.touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int32_t fnum = is_finest ? 1 : 3, fden = 4;     /* filter_ratio 0.25 / 0.75 (:2256-2258) */
    int32_t count = 0, balance_counter = 0, iters = 0;
    while (count++ < opts->patience) {
        int64_t curmax = max_of(pwgts, k);
        trace_rec(n, iters, curmax <= maxwgt ? 0 : (balance_counter < 2 ? 1 : 2), cur_cut, 0);
        if (curmax <= maxwgt) {
            oracle_jet_lp_step(n, xadj, adjncy, adjwgt, k, where, fnum, fden, lock, dest);
            balance_counter = 0;
        } else {
            if (balance_counter < 2) jet_rw_step(&c, where, pwgts, maxwgt, opt, dest);
            else jet_rs_step(&c, where, pwgts, maxwgt, opt, dest);
            balance_counter++;
        }
        if (g_dump_where && n == g_dump_n && iters == g_dump_iter) {
            memcpy(g_dump_where, where, sizeof(int32_t) * (size_t)n);
            memcpy(g_dump_dest, dest, sizeof(int32_t) * (size_t)n);
            g_dumped = 1;
        }
        cur_cut += apply_moves(n, xadj, adjncy, adjwgt, vwgt, where, dest, pwgts);
        iters++;
        curmax = max_of(pwgts, k);
        if (oracle_jet_accept(cur_cut, curmax, maxwgt, &best_cut, &best_max, &count))
            memcpy(best_where, where, sizeof(int32_t) * (size_t)n);
    }
    trace_rec(n, iters, 0, cur_cut, 1);
    memcpy(where, best_where, sizeof(int32_t) * (size_t)n);
    *cut = best_cut;
    free(pwgts); free(best_where); free(dest); free(lock); free(c.conn); free(c.touched);
    return iters;
}

/* ---------------------------------------- older label propagation (k-way, 2-way) */

/* hunyuangraph_compute_log2 (hunyuangraph_common.h:26-31): floor(log2 a), 0 for a <= 1. */
static int32_t ref_log2(int32_t a)
{
    int32_t i;
    for (i = 1; a > 1; i++, a >>= 1) {}
    return i - 1;
}

/* hunyuangraph_GPU_LP.h:1251: max(nvtxs / (6000 * log2(nparts)), 8 * nparts).
 * The reference divides by zero for one part; no rounds then. */
int32_t oracle_lp_threshold(int32_t n, int32_t k)
{
    if (k < 2) return 0;
    int32_t a = n / (6000 * ref_log2(k)), b = 8 * k;
    return a > b ? a : b;
}

/* calculate_overweight (hunyuangraph_GPU_LP.h:1008-1026) with tpwgts[0] =
 * (float)(1.0 / nparts) (hunyuangraph_admin.h:21-23): each bound is a double
 * expression stored in a float; a part weight is compared as a float. */
void oracle_lp_part_bounds(int32_t k, int64_t tvwgt, float *fmin, float *fmax)
{
    float temp = (float)(1.0 / k);
    int32_t tv = (int32_t)tvwgt;
    *fmin = (float)(temp / 1.03 * tv);
    *fmax = (float)(temp * 1.03 * tv);
}

static void lp_flags(const int64_t *pw, int32_t k, float fmin, float fmax, uint8_t *over, uint8_t *thin,
                     int32_t *nover, int32_t *nthin)
{
    *nover = *nthin = 0;
    for (int32_t p = 0; p < k; p++) {
        float val = (float)(int32_t)pw[p];
        over[p] = val > fmax;
        thin[p] = val < fmin;
        *nover += over[p];
        *nthin += thin[p];
    }
}

/* Arg-max of conn over all k parts, the lowest part id among equal values;
 * only the parts in touched[0..nt) can be non-zero. */
static int32_t lp_argmax(const int64_t *conn, const int32_t *touched, int32_t nt)
{
    int32_t best = -1;
    int64_t bc = 0;
    for (int32_t i = 0; i < nt; i++) {
        int32_t q = touched[i];
        if (conn[q] > bc || (conn[q] == bc && conn[q] > 0 && q < best)) {
            bc = conn[q];
            best = q;
        }
    }
    return best < 0 ? 0 : best;   /* all zero: part 0 */
}

/* calculate_move (hunyuangraph_GPU_LP.h:990-1006): the chosen vertex of each
 * destination part moves there; returns the number of moves. */
static int32_t lp_apply(int32_t k, const int32_t *bv, const int32_t *vwgt, int64_t *pw, int32_t *where,
                        uint8_t *moved)
{
    int32_t moves = 0;
    for (int32_t p = 0; p < k; p++) {
        int32_t v = bv[p];
        if (v < 0) continue;
        pw[where[v]] -= vwgt[v];
        pw[p] += vwgt[v];
        where[v] = p;
        moved[v] = 1;
        moves++;
    }
    return moves;
}

/* hunyuangraph_k_refinement_me (hunyuangraph_GPU_LP.h:1245-1374).
 * Deterministic where the reference races:
 *  - calculate_to / calculate_to_reba (:822-917, :1059-1162) reduce the
 *    connectivity row in shared memory with unsynchronised steps (and, past
 *    64 parts, pair a value with the index of another slot, :858-867); here
 *    the destination is the part of largest connectivity, the lowest part id
 *    among equal values, over all parts including the own one.
 *  - calculate_Max (:930-988) and calculate_Max_reba (:1164-1179) pick, per
 *    destination, whichever vertex of the largest gain wrote last; here the
 *    lowest vertex id among equal gains.
 *  - set_bnd_reba (:1028-1057) reads an uninitialised flag when no neighbour
 *    qualifies; here such a vertex is not a boundary vertex.
 *  - The first loop stops after a round without moves: nothing changed, so
 *    every later round would repeat it. */
void oracle_lp_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                      const int32_t *adjwgt, int32_t k, int64_t tvwgt, int32_t *where, int64_t *cut,
                      int32_t *iters)
{
    iters[0] = iters[1] = 0;
    if (n == 0 || k < 2) {
        *cut = oracle_edge_cut(n, xadj, adjncy, adjwgt, where);
        return;
    }
    int64_t *pw = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    int64_t *conn = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    int64_t *bg = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    int32_t *bv = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    uint8_t *seen = (uint8_t *)calloc((size_t)k, 1);
    uint8_t *over = (uint8_t *)calloc((size_t)k, 1);
    uint8_t *thin = (uint8_t *)calloc((size_t)k, 1);
    uint8_t *moved = (uint8_t *)calloc((size_t)n, 1);
    for (int32_t v = 0; v < n; v++) pw[where[v]] += vwgt[v];
    const int32_t thr = oracle_lp_threshold(n, k);

    /* first loop (:1283-1308): set_bnd, calculate_to, init_max, calculate_Max, calculate_move */
    for (int32_t it = 0; it < thr; it++) {
        for (int32_t p = 0; p < k; p++) { bv[p] = -1; bg[p] = 0; }
        for (int32_t v = 0; v < n; v++) {
            if (moved[v]) continue;                       /* set_bnd (:791-820) */
            const int32_t me = where[v];
            int bnd = 0;
            for (int64_t e = xadj[v]; e < xadj[v + 1]; e++)
                if (where[adjncy[e]] != me) { bnd = 1; break; }
            if (!bnd) continue;
            int32_t nt = 0;
            for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
                int32_t q = where[adjncy[e]];
                if (!seen[q]) { seen[q] = 1; touched[nt++] = q; }
                conn[q] += adjwgt[e];
            }
            const int32_t best = lp_argmax(conn, touched, nt);
            const int64_t gain = conn[best] - conn[me];
            for (int32_t i = 0; i < nt; i++) { conn[touched[i]] = 0; seen[touched[i]] = 0; }
            if (best == me || gain <= 0) continue;        /* moveto -1, or calculate_Max's val > 0 */
            if (bv[best] < 0 || gain > bg[best]) { bv[best] = v; bg[best] = gain; }
        }
        if (!lp_apply(k, bv, vwgt, pw, where, moved)) break;
        iters[0]++;
    }

    /* rebalancing loop (:1313-1364) */
    float fmin, fmax;
    oracle_lp_part_bounds(k, tvwgt, &fmin, &fmax);
    int32_t nover, nthin;
    lp_flags(pw, k, fmin, fmax, over, thin, &nover, &nthin);
    int32_t it2 = 0;
    while (nover > 0 && nthin > 0) {
        if (it2 >= thr) break;
        it2++;
        for (int32_t p = 0; p < k; p++) { bv[p] = -1; bg[p] = 0; }
        for (int32_t v = 0; v < n; v++) {
            const int32_t me = where[v];
            if (!over[me]) continue;                      /* set_bnd_reba (:1028-1057) */
            int bnd = 0;
            for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
                int32_t q = where[adjncy[e]];
                if (q != me && thin[q]) { bnd = 1; break; }
            }
            if (!bnd) continue;
            int32_t nt = 0;                                /* calculate_to_reba (:1059-1162) */
            for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
                int32_t q = where[adjncy[e]];
                if (q == me || !thin[q]) continue;
                if (!seen[q]) { seen[q] = 1; touched[nt++] = q; }
                conn[q] += adjwgt[e];
            }
            const int32_t best = lp_argmax(conn, touched, nt);
            const int64_t gain = conn[best];
            for (int32_t i = 0; i < nt; i++) { conn[touched[i]] = 0; seen[touched[i]] = 0; }
            if (!thin[best]) continue;
            if (bv[best] < 0 || gain > bg[best]) { bv[best] = v; bg[best] = gain; }   /* calculate_Max_reba */
        }
        lp_apply(k, bv, vwgt, pw, where, moved);
        lp_flags(pw, k, fmin, fmax, over, thin, &nover, &nthin);
    }
    iters[1] = it2;
    *cut = oracle_edge_cut(n, xadj, adjncy, adjwgt, where);
    free(pw); free(conn); free(bg); free(touched); free(bv); free(seen); free(over); free(thin); free(moved);
}

typedef struct { int32_t key, ptr; } lpkp;

/* compRule (hunyuangraph_GPU_LP.h:499-508): key descending, then id ascending. */
static int cmp_lpkp(const void *a, const void *b)
{
    const lpkp *x = (const lpkp *)a, *y = (const lpkp *)b;
    if (x->key != y->key) return x->key > y->key ? -1 : 1;
    return x->ptr < y->ptr ? -1 : (x->ptr > y->ptr);
}

/* jetLP (hunyuangraph_GPU_LP.h:673-781).  filter_atomic (:253-296) lets a
 * thread read sides its neighbours have already flipped; here every vertex
 * decides on the sides before the pass (a synchronous pass). */
int32_t oracle_lp2_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                          const int32_t *adjwgt, int32_t *where)
{
    if (n == 0) return 0;
    int32_t *nw = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; v++) {
        int64_t same = 0, other = 0;
        for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
            if (where[adjncy[e]] == where[v]) same += adjwgt[e];
            else other += adjwgt[e];
        }
        nw[v] = same <= other ? 1 - where[v] : where[v];
    }
    memcpy(where, nw, sizeof(int32_t) * (size_t)n);
    free(nw);
    /* compute_pwgts + reduction6 (:410-466, :714-722) */
    int64_t tv = 0, pw[2] = {0, 0};
    for (int32_t v = 0; v < n; v++) {
        tv += vwgt[v];
        if (where[v] == 0) pw[0] += vwgt[v];
    }
    pw[1] = tv - pw[0];
    const int32_t tvi = (int32_t)tv;
    const double lo = tvi * 0.5 / 1.03, hi = tvi * 0.5 * 1.03;
    if (pw[0] >= lo && pw[0] <= hi && pw[1] >= lo && pw[1] <= hi) return 0;
    const int32_t from = pw[0] > pw[1] ? 0 : 1, to = 1 - from;
    /* compute_gainkp (:468-497); nedges = the number of adjacency entries */
    const int32_t nedges = (int32_t)xadj[n];
    lpkp *kp = (lpkp *)malloc(sizeof(lpkp) * (size_t)n);
    for (int32_t v = 0; v < n; v++) {
        int32_t id = 0, ed = 0;
        for (int64_t e = xadj[v]; e < xadj[v + 1]; e++) {
            if (where[adjncy[e]] == where[v]) id += adjwgt[e];
            else ed += adjwgt[e];
        }
        kp[v].key = where[v] == from ? ed - id : -nedges;
        kp[v].ptr = v;
    }
    qsort(kp, (size_t)n, sizeof(lpkp), cmp_lpkp);
    /* compute_gainvkp, inclusive_scan, rebalancekp (:510-532, :751-773) */
    const int32_t num = (int32_t)((pw[from] - pw[to]) / 2);
    int64_t prefix = 0;
    for (int32_t i = 0; i < n; i++) {
        prefix += vwgt[kp[i].ptr];
        if (prefix > num) break;
        where[kp[i].ptr] = to;
    }
    free(kp);
    return 1;
}

/* ------------------------------------------------------ full pipeline */

/* hunyuangraph_PartitionGraph (hunyuangraph_partitiongraph.h:284-347) ->
 * hunyuangraph_kway_partition (:13-149): coarsen (hunyuangraph_GPU_coarsen.h:163-253),
 * initial partition (hunyuangraph_GPU_initialpartition.h:4405-4496), uncoarsen
 * (hunyuangraph_GPU_uncoarsen.h:555-587). */
/* Level 0 as a private copy of the input graph. */
static void level0_copy(olevel *l0, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *vwgt, const int32_t *adjwgt)
{
    int64_t m0 = xadj[n];
    l0->n = n; l0->m = m0;
    l0->xadj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    l0->adjncy = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m0 + 1));
    l0->adjwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m0 + 1));
    l0->vwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    memcpy(l0->xadj, xadj, sizeof(int64_t) * (size_t)(n + 1));
    memcpy(l0->adjncy, adjncy, sizeof(int32_t) * (size_t)m0);
    memcpy(l0->adjwgt, adjwgt, sizeof(int32_t) * (size_t)m0);
    memcpy(l0->vwgt, vwgt, sizeof(int32_t) * (size_t)n);
}

/* Coarsening loop of hunyuangarph_coarsen (hunyuangraph_GPU_coarsen.h:163-253):
 * a do-while over levels with the stop rule n > thr && n < 0.85 n_finer &&
 * m > n/2 (hunyuangraph_GPU_coarsen.h:244-247), in exact integers.  lv[0]
 * holds the input; returns the index L of the coarsest level. */
static int32_t coarsen_levels(olevel *lv, int32_t k, int64_t tvwgt, const oracle_opts *o)
{
    int32_t thr = oracle_coarsen_threshold(k);
    int32_t maxvwgt = oracle_maxvwgt(tvwgt, thr);
    int32_t L = 0;
    for (;;) {
        olevel *f = &lv[L];
        int32_t *match = (int32_t *)malloc(sizeof(int32_t) * (size_t)f->n);
        oracle_dga_match(f->n, f->xadj, f->adjncy, f->vwgt, f->adjwgt, maxvwgt, o->seed, L,
                         o->match_rounds, match);
        f->cmap = (int32_t *)malloc(sizeof(int32_t) * (size_t)f->n);
        int32_t cn = oracle_cmap(f->n, match, f->cmap);
        olevel *c = &lv[L + 1];
        c->n = cn;
        c->xadj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cn + 1));
        c->adjncy = (int32_t *)malloc(sizeof(int32_t) * (size_t)(f->m + 1));
        c->adjwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(f->m + 1));
        c->vwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)cn);
        c->m = oracle_contract(f->n, f->xadj, f->adjncy, f->adjwgt, f->vwgt, match, f->cmap, cn,
                               c->xadj, c->adjncy, c->adjwgt, c->vwgt);
        free(match);
        L++;
        if (!(c->n > thr && 100 * (int64_t)c->n < 85 * (int64_t)f->n && c->m > c->n / 2) || L >= 62)
            break;
    }
    return L;
}

/* Level sizes and total edge weights (the per-level figures of the
 * reference's -DFIGURE9_SUM build, hunyuangraph_GPU_coarsen.h:240-242). */
static void level_stats(const olevel *lv, int32_t L, oracle_stats *st)
{
    st->levels = L;
    st->coarsest_n = lv[L].n;
    st->coarsest_m = lv[L].m;
    for (int32_t l = 0; l <= L && l < 64; l++) {
        st->level_n[l] = lv[l].n;
        st->level_m[l] = lv[l].m;
        int64_t w = 0;
        for (int64_t j = 0; j < lv[l].m; j++) w += lv[l].adjwgt[j];
        st->level_adjwgt[l] = w;
    }
}

int oracle_dga_coarsen(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                       const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                       const oracle_opts *opts, int32_t *cn, int64_t *cm, int64_t *cxadj,
                       int32_t *cadjncy, int32_t *cadjwgt, int32_t *cvwgt, oracle_stats *st)
{
    oracle_opts o;
    if (opts) o = *opts; else oracle_default_opts(&o);
    if (st) memset(st, 0, sizeof(*st));
    if (n < 1 || k < 1) return -1;
    int64_t tvwgt = 0;
    for (int32_t v = 0; v < n; v++) tvwgt += vwgt[v];
    olevel lv[64];
    memset(lv, 0, sizeof(lv));
    level0_copy(&lv[0], n, xadj, adjncy, vwgt, adjwgt);
    int32_t L = coarsen_levels(lv, k, tvwgt, &o);
    if (st) level_stats(lv, L, st);
    const olevel *c = &lv[L];
    *cn = c->n;
    *cm = c->m;
    memcpy(cxadj, c->xadj, sizeof(int64_t) * (size_t)(c->n + 1));
    memcpy(cadjncy, c->adjncy, sizeof(int32_t) * (size_t)c->m);
    memcpy(cadjwgt, c->adjwgt, sizeof(int32_t) * (size_t)c->m);
    memcpy(cvwgt, c->vwgt, sizeof(int32_t) * (size_t)c->n);
    for (int32_t l = 0; l <= L; l++) level_free(&lv[l]);
    return 0;
}

int oracle_dga_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                         const oracle_opts *opts, int32_t *part, oracle_stats *st)
{
    oracle_opts o;
    if (opts) o = *opts; else oracle_default_opts(&o);
    if (st) memset(st, 0, sizeof(*st));
    if (n < 0 || k < 1) return -1;
    if (n == 0) return 0;
    if (k == 1) {
        for (int32_t v = 0; v < n; v++) part[v] = 0;
        return 0;
    }
    int64_t tvwgt = 0;
    for (int32_t v = 0; v < n; v++) tvwgt += vwgt[v];

    olevel lv[64];
    memset(lv, 0, sizeof(lv));
    level0_copy(&lv[0], n, xadj, adjncy, vwgt, adjwgt);
    int32_t L = coarsen_levels(lv, k, tvwgt, &o);

    olevel *cg = &lv[L];
    cg->where = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cg->n > 0 ? cg->n : 1));
    oracle_dga_initpart(cg->n, cg->xadj, cg->adjncy, cg->vwgt, cg->adjwgt, k, &o, cg->where);
    int64_t cut = oracle_edge_cut(cg->n, cg->xadj, cg->adjncy, cg->adjwgt, cg->where);
    if (st) {
        level_stats(lv, L, st);
        st->init_cut = cut;
    }
    for (int32_t l = L; l >= 0; l--) {
        olevel *g = &lv[l];
        int32_t it;
        if (o.refiner == 1) {
            int32_t two[2];
            oracle_lp_refine(g->n, g->xadj, g->adjncy, g->vwgt, g->adjwgt, k, tvwgt, g->where, &cut, two);
            it = two[0] + two[1];
        } else {
            it = oracle_dga_refine(g->n, g->xadj, g->adjncy, g->vwgt, g->adjwgt, k, tvwgt, l == 0, &o, g->where,
                                   &cut);
        }
        if (st) { st->level_iters[l] = it; st->refine_iters += it; }
        if (l == 0) break;
        olevel *fg = &lv[l - 1];
        fg->where = (int32_t *)malloc(sizeof(int32_t) * (size_t)fg->n);
        for (int32_t v = 0; v < fg->n; v++) fg->where[v] = g->where[fg->cmap[v]];
        level_free(g);
    }
    memcpy(part, lv[0].where, sizeof(int32_t) * (size_t)n);
    if (st) {
        st->cut = cut;
        int64_t *pw = (int64_t *)calloc((size_t)k, sizeof(int64_t));
        oracle_part_weights(n, vwgt, part, k, pw);
        st->max_pwgt = max_of(pw, k);
        free(pw);
    }
    level_free(&lv[0]);
    return 0;
}
