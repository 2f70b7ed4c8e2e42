// hgp_cpurb.cpp -- the reference's CPU partitioner as a library entry
// (hgp_partition_cpu): serial METIS-style multilevel recursive bisection,
// hunyuangraph_rbbisection (hunyuangraph_CPU_initialpartition.h:361-374).
//
// A host-only path the caller chooses explicitly; the GPU entry points never
// fall back to it.  It runs with the parameters of the reference's standalone
// admin (hunyuangraph_admin.h:7-31): coarsening threshold 200, one cut, ten
// FM passes, ubfactor 1.03 per bisection, uniform target weights.  Random
// choices (matching visit order, growing starts, FM insertion order) come
// from a seeded xorshift64* stream where the reference calls glibc rand()
// unseeded (hunyuangraph_common.h:34-47).  Arithmetic is int32 with the
// reference's float balance factors.
#include "hunyuangraph.h"

#include "hgp_hostcheck.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

namespace hgp {

void clear_error();

namespace {

class Rng {
  public:
    explicit Rng(uint64_t seed) : s_(seed ? seed * 0x9E3779B97F4A7C15ull : 0x1234567ull) {}
    int next()
    {
        s_ ^= s_ >> 12;
        s_ ^= s_ << 25;
        s_ ^= s_ >> 27;
        return (int)((s_ * 0x2545F4914F6CDD1Dull) >> 33);
    }
    int below(int max) { return next() % max; }

    // hunyuangraph_int_randarrayofp (hunyuangraph_common.h:105-137): p is
    // set to 0..n-1 when fresh, then shuffled by `rounds` swaps of 4-blocks
    // (single swaps for n < 10).
    void shuffle(std::vector<int> &p, int n, int rounds, bool fresh)
    {
        if (fresh)
            for (int i = 0; i < n; i++) p[i] = i;
        if (n < 10) {
            for (int i = 0; i < n; i++) {
                int a = below(n), b = below(n);
                std::swap(p[a], p[b]);
            }
            return;
        }
        for (int i = 0; i < rounds; i++) {
            int a = below(n - 3), b = below(n - 3);
            std::swap(p[a + 0], p[b + 2]);
            std::swap(p[a + 1], p[b + 3]);
            std::swap(p[a + 2], p[b + 0]);
            std::swap(p[a + 3], p[b + 1]);
        }
    }

  private:
    uint64_t s_;
};

// Max-priority queue of vertices keyed by gain with a position index
// (hunyuangraph_priorityqueue.h:22-262): a key moves above its parent when
// strictly greater; a removal refills the hole from the last slot.
class GainQueue {
  public:
    explicit GainQueue(int cap) : key_(cap > 0 ? cap : 1), val_(cap > 0 ? cap : 1), pos_(cap > 0 ? cap : 1, -1) {}
    void clear()
    {
        for (int i = 0; i < cnt_; i++) pos_[val_[i]] = -1;
        cnt_ = 0;
    }
    void push(int v, int key) { up(cnt_++, v, key); }
    int pop()
    {
        if (cnt_ == 0) return -1;
        int top = val_[0];
        pos_[top] = -1;
        if (--cnt_ > 0) down(0, val_[cnt_], key_[cnt_]);
        return top;
    }
    void change(int v, int key)
    {
        int i = pos_[v];
        if (key > key_[i]) up(i, v, key);
        else down(i, v, key);
    }
    void erase(int v)
    {
        int i = pos_[v];
        pos_[v] = -1;
        if (--cnt_ > 0 && val_[cnt_] != v) {
            int nv = val_[cnt_], nk = key_[cnt_];
            if (nk > key_[i]) up(i, nv, nk);
            else down(i, nv, nk);
        }
    }

  private:
    void put(int i, int v, int key)
    {
        key_[i] = key;
        val_[i] = v;
        pos_[v] = i;
    }
    void up(int i, int v, int key)
    {
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (!(key > key_[p])) break;
            put(i, val_[p], key_[p]);
            i = p;
        }
        put(i, v, key);
    }
    void down(int i, int v, int key)
    {
        for (int c; (c = 2 * i + 1) < cnt_;) {
            if (key_[c] > key) {
                if (c + 1 < cnt_ && key_[c + 1] > key_[c]) c++;
            } else if (c + 1 < cnt_ && key_[c + 1] > key) {
                c++;
            } else {
                break;
            }
            put(i, val_[c], key_[c]);
            i = c;
        }
        put(i, v, key);
    }
    std::vector<int> key_, val_, pos_;
    int cnt_ = 0;
};

// A graph of the recursion or of a coarsening hierarchy, with its bisection
// state (where, part weights, internal/external degrees, boundary list).
struct Graph {
    int n = 0, m = 0, tvwgt = 0;
    float tvwgt_rev = 0;
    std::vector<int> xadj, vwgt, adjncy, adjwgt, label, cmap;
    std::vector<int> where, bndptr, bndlist, id, ed;
    int pwgts[2] = {0, 0};
    int nbnd = 0, mincut = 0;
    std::unique_ptr<Graph> coarser;
    Graph *finer = nullptr;

    void set_total()
    {
        long s = 0;
        for (int i = 0; i < n; i++) s += vwgt[i];
        tvwgt = (int)s;
        tvwgt_rev = 1.0f / (float)(tvwgt > 0 ? tvwgt : 1);
    }
    void alloc_bisection()
    {
        where.assign(n, 0);
        bndptr.assign(n, -1);
        bndlist.assign(n, 0);
        id.assign(n, 0);
        ed.assign(n, 0);
    }
    int deg(int v) const { return xadj[v + 1] - xadj[v]; }
    void bnd_insert(int v)
    {
        bndlist[nbnd] = v;
        bndptr[v] = nbnd++;
    }
    void bnd_delete(int v)
    {
        bndlist[bndptr[v]] = bndlist[--nbnd];
        bndptr[bndlist[nbnd]] = bndptr[v];
        bndptr[v] = -1;
    }
    // hunyuangraph_compute_cpu_2wayparam (hunyuangraph_CPU_2wayrefine.h:22-75)
    void bisection_params()
    {
        pwgts[0] = pwgts[1] = 0;
        nbnd = 0;
        int cut = 0;
        for (int i = 0; i < n; i++) {
            bndptr[i] = -1;
            pwgts[where[i]] += vwgt[i];
        }
        for (int i = 0; i < n; i++) {
            int in = 0, ex = 0;
            for (int j = xadj[i]; j < xadj[i + 1]; j++) (where[adjncy[j]] == where[i] ? in : ex) += adjwgt[j];
            id[i] = in;
            ed[i] = ex;
            if (ex > 0 || deg(i) == 0) {
                bnd_insert(i);
                cut += ex;
            }
        }
        mincut = cut / 2;
    }
};

class Bisector {
  public:
    explicit Bisector(uint64_t seed) : rng_(seed) {}

    // hunyuangraph_mlevel_rbbisection (hunyuangraph_CPU_initialpartition.h:284-358)
    int recurse(std::unique_ptr<Graph> g, int nparts, int *part, float *tpwgts, int fpart, int level)
    {
        if (g->n == 0) return 0;
        const int n = g->n;
        float t0 = 0;
        for (int i = 0; i < (nparts >> 1); i++) t0 += tpwgts[i];
        const float tp[2] = {t0, 1.0f - t0};
        int obj = bisect(*g, tp);
        level++;
        for (int i = 0; i < n; i++) part[level == 1 ? i : g->label[i]] = g->where[i] + fpart;
        std::unique_ptr<Graph> left, right;
        if (nparts > 2) split(*g, level == 1, left, right);
        g.reset();
        float ws = 0;
        for (int i = 0; i < (nparts >> 1); i++) ws += tpwgts[i];
        for (int i = 0; i < (nparts >> 1); i++) tpwgts[i] *= 1.0f / ws;
        for (int i = 0; i < nparts - (nparts >> 1); i++) tpwgts[(nparts >> 1) + i] *= 1.0f / (1.0f - ws);
        if (nparts > 3) {
            obj += recurse(std::move(left), nparts >> 1, part, tpwgts, fpart, level);
            obj += recurse(std::move(right), nparts - (nparts >> 1), part, tpwgts + (nparts >> 1),
                           fpart + (nparts >> 1), level);
        } else if (nparts == 3) {
            obj += recurse(std::move(right), nparts - (nparts >> 1), part, tpwgts + (nparts >> 1),
                           fpart + (nparts >> 1), level);
        }
        return obj;
    }

  private:
    static constexpr int kThreshold = 200;   // standalone admin's coarsening threshold
    static constexpr int kPasses = 10;       // FM passes
    static constexpr float kUb = 1.03f;

    // hunyuangraph_cpu_mlevelbisect (hunyuangraph_CPU_initialpartition.h:148-281), one cut
    int bisect(Graph &g, const float *tp)
    {
        balance_[0] = g.tvwgt_rev / tp[0];
        balance_[1] = g.tvwgt_rev / tp[1];
        Graph *c = coarsen(g);
        grow(*c, tp, c->n <= kThreshold ? 5 : 7);
        // hunyuangraph_cpu_refinement (hunyuangraph_CPU_2wayrefine.h:335-386)
        c->bisection_params();
        for (Graph *h = c;;) {
            balance(*h, tp);
            fm(*h, tp);
            if (h == &g) break;
            h = h->finer;
            project(*h);
        }
        return g.mincut;
    }

    // hunyuangraph_cpu_coarsen (hunyuangraph_CPU_coarsen.h:11-46)
    Graph *coarsen(Graph &g)
    {
        bool equal = true;
        for (int i = 1; i < g.m && equal; i++) equal = g.adjwgt[i] == g.adjwgt[0];
        maxvwgt_ = (int)(1.5 * g.tvwgt / kThreshold);
        Graph *h = &g;
        do {
            if ((int)h->cmap.size() < h->n) h->cmap.assign(h->n, 0);
            if (equal || h->m == 0) match_random(*h);
            else match_heavy(*h);
            h = h->coarser.get();
        } while (h->n > kThreshold && h->n < 0.85 * h->finer->n && h->m > h->n / 2);
        return h;
    }

    // hunyuangraph_cpu_match_RM (hunyuangraph_CPU_match.h:194-282)
    void match_random(Graph &g)
    {
        const int n = g.n;
        std::vector<int> match(n, -1), perm(n);
        rng_.shuffle(perm, n, n / 8, true);
        int cn = 0, last = 0, unmatched = 0;
        for (int pi = 0; pi < n; pi++) {
            const int i = perm[pi];
            if (match[i] != -1) continue;
            int pick = i;
            if (g.vwgt[i] < maxvwgt_) {
                if (g.deg(i) == 0) {
                    pick = next_free(perm, match, pi, last);
                } else {
                    for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) {
                        int u = g.adjncy[j];
                        if (match[u] == -1 && g.vwgt[i] + g.vwgt[u] <= maxvwgt_) {
                            pick = u;
                            break;
                        }
                    }
                    if (pick == i && 3 * g.vwgt[i] < maxvwgt_) {
                        unmatched++;
                        pick = -1;
                    }
                }
            }
            if (pick != -1) pair_up(g, match, i, pick, cn);
        }
        if (unmatched > 0.1 * n) cn = match_two_hop(g, perm, match, cn, unmatched);
        contract(g, finish(g, match), match);
    }

    // hunyuangraph_cpu_match_HEM (hunyuangraph_CPU_match.h:285-402), with the
    // visit order of hunyuangraph_matching_sort (:10-31): a random
    // permutation bucket-sorted by degree capped at 0.7 x the mean degree
    void match_heavy(Graph &g)
    {
        const int n = g.n;
        std::vector<int> match(n, -1), perm(n), rperm(n), d(n);
        rng_.shuffle(rperm, n, n / 8, true);
        const int cap = (int)(0.7 * (g.xadj[n] / n));
        for (int i = 0; i < n; i++) d[i] = std::min(g.deg(i), cap);
        std::vector<int> start(cap + 2, 0);
        for (int i = 0; i < n; i++) start[d[i] + 1]++;
        for (int b = 1; b < cap + 2; b++) start[b] += start[b - 1];
        for (int ii = 0; ii < n; ii++) {
            int i = rperm[ii];
            perm[start[d[i]]++] = i;
        }
        int cn = 0, last = 0;
        for (int pi = 0; pi < n; pi++) {
            const int i = perm[pi];
            if (match[i] != -1) continue;
            int pick = i, best = -1;
            if (g.vwgt[i] < maxvwgt_) {
                if (g.deg(i) == 0) {
                    pick = next_free(perm, match, pi, last);
                } else {
                    for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) {
                        int u = g.adjncy[j];
                        if (match[u] == -1 && best < g.adjwgt[j] && g.vwgt[i] + g.vwgt[u] <= maxvwgt_) {
                            pick = u;
                            best = g.adjwgt[j];
                        }
                    }
                    if (pick == i && 3 * g.vwgt[i] < maxvwgt_) pick = -1;
                }
            }
            if (pick != -1) pair_up(g, match, i, pick, cn);
        }
        contract(g, finish(g, match), match);
    }

    // the first free vertex after position max(pi, last) of the visit order
    static int next_free(const std::vector<int> &perm, const std::vector<int> &match, int pi, int &last)
    {
        last = std::max(pi, last) + 1;
        for (; last < (int)perm.size(); last++)
            if (match[perm[last]] == -1) return perm[last];
        return perm[pi];
    }

    static void pair_up(Graph &g, std::vector<int> &match, int a, int b, int &cn)
    {
        g.cmap[a] = g.cmap[b] = cn++;
        match[a] = b;
        match[b] = a;
    }

    // Match_2Hop (hunyuangraph_CPU_match.h:162-172) over Match_2HopAny
    // (:37-98): unmatched vertices sharing a neighbour are paired.
    // Match_2HopAll (:100-160) changes nothing in the reference (its match
    // pointer is swapped for a zeroed array) and is not repeated here.
    int match_two_hop(Graph &g, const std::vector<int> &perm, std::vector<int> &match, int cn, int unmatched)
    {
        cn = two_hop_any(g, perm, match, cn, unmatched, 2);
        if (unmatched > 1.5 * 0.1 * g.n) cn = two_hop_any(g, perm, match, cn, unmatched, 3);
        if (unmatched > 2.0 * 0.1 * g.n) cn = two_hop_any(g, perm, match, cn, unmatched, g.n);
        return cn;
    }

    static int two_hop_any(Graph &g, const std::vector<int> &perm, std::vector<int> &match, int cn, int &unmatched,
                           int maxdeg)
    {
        const int n = g.n;
        // for each vertex, the unmatched low-degree vertices next to it (in visit order)
        std::vector<int> off(n + 1, 0);
        auto small = [&](int i) { return match[i] == -1 && g.deg(i) < maxdeg; };
        for (int i = 0; i < n; i++)
            if (small(i))
                for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) off[g.adjncy[j] + 1]++;
        for (int i = 0; i < n; i++) off[i + 1] += off[i];
        std::vector<int> fill(off.begin(), off.end() - 1), who(std::max(off[n], 1));
        for (int pi = 0; pi < n; pi++) {
            int i = perm[pi];
            if (small(i))
                for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) who[fill[g.adjncy[j]]++] = i;
        }
        for (int pi = 0; pi < n; pi++) {
            const int i = perm[pi];
            if (off[i + 1] - off[i] < 2) continue;
            int hi = off[i + 1];
            for (int j = off[i]; j < hi; j++) {
                if (match[who[j]] != -1) continue;
                for (hi--; hi > j; hi--) {
                    if (match[who[hi]] == -1) {
                        pair_up(g, match, who[j], who[hi], cn);
                        unmatched -= 2;
                        break;
                    }
                }
            }
        }
        return cn;
    }

    // coarse ids in vertex order: unmatched vertices become singletons
    static int finish(Graph &g, std::vector<int> &match)
    {
        int cn = 0;
        for (int i = 0; i < g.n; i++) {
            if (match[i] == -1) {
                match[i] = i;
                g.cmap[i] = cn++;
            } else if (i <= match[i]) {
                g.cmap[i] = g.cmap[match[i]] = cn++;
            }
        }
        return cn;
    }

    // hunyuangraph_cpu_create_cgraph (hunyuangraph_CPU_contraction.h:9-98):
    // each coarse row in first-seen order of its neighbours (a table from
    // coarse id to row slot), the self entry dropped by moving the row's
    // last entry into its slot
    static void contract(Graph &g, int cn, const std::vector<int> &match)
    {
        auto c = std::make_unique<Graph>();
        c->n = cn;
        c->xadj.assign(cn + 1, 0);
        c->vwgt.assign(std::max(cn, 1), 0);
        c->adjncy.assign(std::max(g.m, 1), 0);
        c->adjwgt.assign(std::max(g.m, 1), 0);
        std::vector<int> slot(std::max(cn, 1), -1);
        int row = 0, total = 0;
        for (int v = 0; v < g.n; v++) {
            const int u = match[v];
            if (u < v) continue;
            int *adj = c->adjncy.data() + total, *wgt = c->adjwgt.data() + total;
            int len = 0;
            auto add_row = [&](int x) {
                for (int j = g.xadj[x]; j < g.xadj[x + 1]; j++) {
                    int t = g.cmap[g.adjncy[j]];
                    if (slot[t] == -1) {
                        adj[len] = t;
                        wgt[len] = g.adjwgt[j];
                        slot[t] = len++;
                    } else {
                        wgt[slot[t]] += g.adjwgt[j];
                    }
                }
            };
            c->vwgt[row] = g.vwgt[v];
            add_row(v);
            if (u != v) {
                c->vwgt[row] += g.vwgt[u];
                add_row(u);
                if (int s = slot[row]; s != -1) {
                    adj[s] = adj[--len];
                    wgt[s] = wgt[len];
                    slot[row] = -1;
                }
            }
            for (int j = 0; j < len; j++) slot[adj[j]] = -1;
            total += len;
            c->xadj[++row] = total;
        }
        c->m = total;
        c->set_total();
        c->finer = &g;
        g.coarser = std::move(c);
    }

    // huyuangraph_cpu_growbisection (hunyuangraph_CPU_initialpartition.h:36-145)
    void grow(Graph &g, const float *tp, int tries)
    {
        const int n = g.n;
        g.alloc_bisection();
        std::vector<int> best(n), queue(n), seen(n);
        const int onemax = (int)(kUb * g.tvwgt * tp[1]);
        const int onemin = (int)((1.0 / kUb) * g.tvwgt * tp[1]);
        int bestcut = 0;
        for (int t = 0; t < tries; t++) {
            std::fill(g.where.begin(), g.where.end(), 1);
            std::fill(seen.begin(), seen.end(), 0);
            int pw0 = 0, pw1 = g.tvwgt;
            queue[0] = rng_.below(n);
            seen[queue[0]] = 1;
            int first = 0, last = 1, left = n - 1;
            bool drain = false;
            for (;;) {
                if (first == last) {
                    if (left == 0 || drain) break;
                    int r = rng_.below(left), i = 0;
                    for (; i < n; i++)
                        if (!seen[i] && r-- == 0) break;
                    queue[0] = i;
                    seen[i] = 1;
                    first = 0;
                    last = 1;
                    left--;
                }
                const int i = queue[first++];
                if (pw0 > 0 && pw1 - g.vwgt[i] < onemin) {
                    drain = true;
                    continue;
                }
                g.where[i] = 0;
                pw0 += g.vwgt[i];
                pw1 -= g.vwgt[i];
                if (pw1 <= onemax) break;
                drain = false;
                for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) {
                    int u = g.adjncy[j];
                    if (!seen[u]) {
                        queue[last++] = u;
                        seen[u] = 1;
                        left--;
                    }
                }
            }
            g.bisection_params();
            balance(g, tp);
            fm(g, tp);
            if (t == 0 || bestcut > g.mincut) {
                bestcut = g.mincut;
                best = g.where;
                if (bestcut == 0) break;
            }
        }
        g.mincut = bestcut;
        g.where = best;
    }

    // hunyuangraph_2way_bal (hunyuangraph_balance.h:128-140)
    void balance(Graph &g, const float *tp)
    {
        float worst = -1.0f;
        for (int s = 0; s < 2; s++) worst = std::max(worst, g.pwgts[s] * balance_[s] - kUb);
        if (worst <= 0) return;
        if (std::fabs(tp[0] * g.tvwgt - g.pwgts[0]) < 3 * g.tvwgt / g.n) return;
        boundary_balance(g, tp);
    }

    // hunyuangraph_bndvertex_2way_bal (hunyuangraph_balance.h:27-126): move
    // boundary vertices of the heavy side, highest gain first, while the
    // light side stays within its target
    void boundary_balance(Graph &g, const float *tp)
    {
        const int n = g.n;
        const int t0 = (int)(g.tvwgt * tp[0]), target[2] = {t0, g.tvwgt - t0};
        const int mindiff = std::abs(target[0] - g.pwgts[0]);
        const int from = g.pwgts[0] < target[0] ? 1 : 0, to = from ^ 1;
        GainQueue q(n);
        std::vector<int> moved(n, -1), perm(std::max(g.nbnd, 1));
        rng_.shuffle(perm, g.nbnd, g.nbnd / 5, true);
        for (int ii = 0; ii < g.nbnd; ii++) {
            int i = g.bndlist[perm[ii]];
            if (g.where[i] == from && g.vwgt[i] <= mindiff) q.push(i, g.ed[i] - g.id[i]);
        }
        int cut = g.mincut;
        for (int step = 0; step < n; step++) {
            const int v = q.pop();
            if (v == -1 || g.pwgts[to] + g.vwgt[v] > target[to]) break;
            cut -= g.ed[v] - g.id[v];
            g.pwgts[to] += g.vwgt[v];
            g.pwgts[from] -= g.vwgt[v];
            g.where[v] = to;
            moved[v] = step;
            std::swap(g.id[v], g.ed[v]);
            if (g.ed[v] == 0 && g.deg(v) > 0) g.bnd_delete(v);
            for (int j = g.xadj[v]; j < g.xadj[v + 1]; j++) {
                const int u = g.adjncy[j], w = g.where[u] == to ? g.adjwgt[j] : -g.adjwgt[j];
                g.id[u] += w;
                g.ed[u] -= w;
                const bool queued = moved[u] == -1 && g.where[u] == from && g.vwgt[u] <= mindiff;
                if (g.bndptr[u] != -1) {
                    if (g.ed[u] == 0) {
                        g.bnd_delete(u);
                        if (queued) q.erase(u);
                    } else if (queued) {
                        q.change(u, g.ed[u] - g.id[u]);
                    }
                } else if (g.ed[u] > 0) {
                    g.bnd_insert(u);
                    if (queued) q.push(u, g.ed[u] - g.id[u]);
                }
            }
        }
        g.mincut = cut;
    }

    // hunyuangraph_cpu_2way_refine (hunyuangraph_CPU_2wayrefine.h:77-242):
    // FM passes with one queue per side, rolled back to the best prefix
    void fm(Graph &g, const float *tp)
    {
        const int n = g.n;
        const int t0 = (int)(g.tvwgt * tp[0]), target[2] = {t0, g.tvwgt - t0};
        const int limit = std::min(std::max((int)(0.01 * n), 15), 100);
        int avgvwgt = (g.pwgts[0] + g.pwgts[1]) / 20;
        if (n > 0) avgvwgt = std::min(avgvwgt, 2 * (g.pwgts[0] + g.pwgts[1]) / n);
        const int origdiff = std::abs(target[0] - g.pwgts[0]);
        GainQueue q[2] = {GainQueue(n), GainQueue(n)};
        std::vector<int> moved(std::max(n, 1), -1), swaps(std::max(n, 1)), perm(std::max(n, 1));
        for (int pass = 0; pass < kPasses; pass++) {
            q[0].clear();
            q[1].clear();
            int bestorder = -1, newcut = g.mincut, mincut = g.mincut;
            const int initcut = g.mincut;
            int mindiff = std::abs(target[0] - g.pwgts[0]);
            rng_.shuffle(perm, g.nbnd, g.nbnd, true);
            for (int ii = 0; ii < g.nbnd; ii++) {
                int i = g.bndlist[perm[ii]];
                q[g.where[i]].push(i, g.ed[i] - g.id[i]);
            }
            int steps = 0;
            for (; steps < n; steps++) {
                const int from = target[0] - g.pwgts[0] < target[1] - g.pwgts[1] ? 0 : 1, to = from ^ 1;
                const int v = q[from].pop();
                if (v == -1) break;
                newcut -= g.ed[v] - g.id[v];
                g.pwgts[to] += g.vwgt[v];
                g.pwgts[from] -= g.vwgt[v];
                const int diff = std::abs(target[0] - g.pwgts[0]);
                if ((newcut < mincut && diff <= origdiff + avgvwgt) || (newcut == mincut && diff < mindiff)) {
                    mincut = newcut;
                    mindiff = diff;
                    bestorder = steps;
                } else if (steps - bestorder > limit) {
                    newcut += g.ed[v] - g.id[v];
                    g.pwgts[from] += g.vwgt[v];
                    g.pwgts[to] -= g.vwgt[v];
                    break;
                }
                g.where[v] = to;
                moved[v] = steps;
                swaps[steps] = v;
                std::swap(g.id[v], g.ed[v]);
                if (g.ed[v] == 0 && g.deg(v) > 0) g.bnd_delete(v);
                for (int j = g.xadj[v]; j < g.xadj[v + 1]; j++) {
                    const int u = g.adjncy[j], w = g.where[u] == to ? g.adjwgt[j] : -g.adjwgt[j];
                    g.id[u] += w;
                    g.ed[u] -= w;
                    if (g.bndptr[u] != -1) {
                        if (g.ed[u] == 0) {
                            g.bnd_delete(u);
                            if (moved[u] == -1) q[g.where[u]].erase(u);
                        } else if (moved[u] == -1) {
                            q[g.where[u]].change(u, g.ed[u] - g.id[u]);
                        }
                    } else if (g.ed[u] > 0) {
                        g.bnd_insert(u);
                        if (moved[u] == -1) q[g.where[u]].push(u, g.ed[u] - g.id[u]);
                    }
                }
            }
            for (int i = 0; i < steps; i++) moved[swaps[i]] = -1;
            for (steps--; steps > bestorder; steps--) {   // roll back past the best prefix
                const int v = swaps[steps];
                const int to = g.where[v] ^ 1;
                g.where[v] = to;
                std::swap(g.id[v], g.ed[v]);
                if (g.ed[v] == 0 && g.bndptr[v] != -1 && g.deg(v) > 0) g.bnd_delete(v);
                else if (g.ed[v] > 0 && g.bndptr[v] == -1) g.bnd_insert(v);
                g.pwgts[to] += g.vwgt[v];
                g.pwgts[to ^ 1] -= g.vwgt[v];
                for (int j = g.xadj[v]; j < g.xadj[v + 1]; j++) {
                    const int u = g.adjncy[j], w = g.where[u] == to ? g.adjwgt[j] : -g.adjwgt[j];
                    g.id[u] += w;
                    g.ed[u] -= w;
                    if (g.bndptr[u] != -1 && g.ed[u] == 0) g.bnd_delete(u);
                    if (g.bndptr[u] == -1 && g.ed[u] > 0) g.bnd_insert(u);
                }
            }
            g.mincut = mincut;
            if (bestorder <= 0 || mincut == initcut) break;
        }
    }

    // hunyuangraph_2way_project (hunyuangraph_CPU_2wayrefine.h:245-321):
    // the coarse sides carried to g; vertices whose coarse vertex was not on
    // the boundary count all their weight as internal
    static void project(Graph &g)
    {
        Graph &c = *g.coarser;
        g.alloc_bisection();
        std::vector<int> cb(g.n);
        for (int i = 0; i < g.n; i++) {
            g.where[i] = c.where[g.cmap[i]];
            cb[i] = c.bndptr[g.cmap[i]];
        }
        g.nbnd = 0;
        for (int i = 0; i < g.n; i++) {
            int in = 0, ex = 0;
            for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) {
                if (cb[i] == -1 || g.where[g.adjncy[j]] == g.where[i]) in += g.adjwgt[j];
                else ex += g.adjwgt[j];
            }
            g.id[i] = in;
            g.ed[i] = ex;
            g.bndptr[i] = -1;
            if (ex > 0 || g.deg(i) == 0) g.bnd_insert(i);
        }
        g.mincut = c.mincut;
        g.pwgts[0] = c.pwgts[0];
        g.pwgts[1] = c.pwgts[1];
        g.coarser.reset();
    }

    // hunyuangraph_splitgraph / _first (hunyuangraph_CPU_splitgraph.h:8-346):
    // the two induced subgraphs in vertex order; labels map to the input
    static void split(const Graph &g, bool first, std::unique_ptr<Graph> &left, std::unique_ptr<Graph> &right)
    {
        std::unique_ptr<Graph> s[2] = {std::make_unique<Graph>(), std::make_unique<Graph>()};
        std::vector<int> rename(g.n);
        int cnt[2] = {0, 0}, ecnt[2] = {0, 0};
        for (int i = 0; i < g.n; i++) {
            rename[i] = cnt[g.where[i]]++;
            ecnt[g.where[i]] += g.deg(i);
        }
        for (int p = 0; p < 2; p++) {
            s[p]->xadj.assign(cnt[p] + 1, 0);
            s[p]->vwgt.reserve(cnt[p]);
            s[p]->label.reserve(cnt[p]);
            s[p]->adjncy.reserve(ecnt[p]);
            s[p]->adjwgt.reserve(ecnt[p]);
        }
        for (int i = 0; i < g.n; i++) {
            Graph &h = *s[g.where[i]];
            const bool interior = g.bndptr[i] == -1;
            for (int j = g.xadj[i]; j < g.xadj[i + 1]; j++) {
                if (interior || g.where[g.adjncy[j]] == g.where[i]) {
                    h.adjncy.push_back(rename[g.adjncy[j]]);
                    h.adjwgt.push_back(g.adjwgt[j]);
                }
            }
            h.vwgt.push_back(g.vwgt[i]);
            h.label.push_back(first ? i : g.label[i]);
            h.xadj[h.vwgt.size()] = (int)h.adjncy.size();
        }
        for (int p = 0; p < 2; p++) {
            s[p]->n = cnt[p];
            s[p]->m = (int)s[p]->adjncy.size();
            s[p]->set_total();
        }
        left = std::move(s[0]);
        right = std::move(s[1]);
    }

    Rng rng_;
    int maxvwgt_ = 0;
    float balance_[2] = {0, 0};
};

}  // namespace

}  // namespace hgp

extern "C" int hgp_partition_cpu(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                                 const int32_t *adjwgt, int32_t k, uint64_t seed, int32_t *part, int64_t *cut)
{
    using namespace hgp;
    clear_error();
    try {
        if (k < 1) throw Error{-1, "nparts must be >= 1"};
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        if (n > 0 && !part) throw Error{-1, "part is NULL"};
        const int64_t m = n > 0 ? xadj[n] : 0;
        if (m > INT32_MAX) throw Error{-4, "the CPU path holds offsets in int32: more than 2^31-1 entries"};
        int64_t total = 0;
        for (int32_t i = 0; i < n; i++) total += vwgt ? vwgt[i] : 1;
        if (total > INT32_MAX) throw Error{-4, "the CPU path needs a total vertex weight below 2^31"};
        if (cut) *cut = 0;
        if (n == 0) return HGP_OK;
        if (k == 1) {
            for (int32_t i = 0; i < n; i++) part[i] = 0;
            return HGP_OK;
        }
        auto g = std::make_unique<Graph>();
        g->n = n;
        g->m = (int)m;
        g->xadj.resize((size_t)n + 1);
        for (int32_t i = 0; i <= n; i++) g->xadj[i] = (int)xadj[i];
        g->adjncy.assign(adjncy, adjncy + m);
        g->adjwgt.resize((size_t)m);
        for (int64_t j = 0; j < m; j++) g->adjwgt[j] = adjwgt ? adjwgt[j] : 1;
        g->vwgt.resize((size_t)n);
        for (int32_t i = 0; i < n; i++) g->vwgt[i] = vwgt ? vwgt[i] : 1;
        g->set_total();
        std::vector<float> tp((size_t)k, 1.0f / k);
        Bisector(seed).recurse(std::move(g), k, part, tp.data(), 0, 0);
        if (cut) {
            int64_t c = 0;
            for (int32_t i = 0; i < n; i++)
                for (int64_t j = xadj[i]; j < xadj[i + 1]; j++)
                    if (part[i] != part[adjncy[j]]) c += adjwgt ? adjwgt[j] : 1;
            *cut = c / 2;
        }
        return HGP_OK;
    } catch (const Error &e) {
        set_error(e.code, e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error(-3, "host out of memory");
        return HGP_ERR_OOM;
    }
}
