/*
 * oracle_cpurb.c -- restatement of the reference CPU path
 * hunyuangraph_rbbisection (hunyuangraph_CPU_initialpartition.h:361-374):
 * serial METIS-style multilevel recursive bisection.
 * TEST INFRASTRUCTURE ONLY (see hgp_oracle.h); it is the timed host baseline.
 *
 * Parameters are those of the standalone admin (hunyuangraph_admin.h:7-31):
 * coarsening threshold 200, ncuts 1, FM passes 10, ubfactor 1.03 per
 * bisection, uniform target weights.  glibc rand() (unseeded in the reference,
 * hunyuangraph_common.h:34-47) is replaced by a seeded xorshift64* stream, so
 * random choices are parity unpinned; everything else follows the reference.
 * Int32 arithmetic as in the reference (the caller keeps totals in range).
 */
#include "hgp_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { uint64_t s; } rbrng;

static int rb_rand(rbrng *r)
{
    r->s ^= r->s >> 12;
    r->s ^= r->s << 25;
    r->s ^= r->s >> 27;
    return (int)((r->s * 0x2545F4914F6CDD1Dull) >> 33);
}

static int rb_randinrange(rbrng *r, int max) { return rb_rand(r) % max; }

/* hunyuangraph_int_randarrayofp (hunyuangraph_common.h:105-137) */
static void rb_randarrayofp(rbrng *r, int n, int *p, int m, int flag)
{
    int i, u, v, t;
    if (flag == 1) for (i = 0; i < n; i++) p[i] = i;
    if (n < 10) {
        for (i = 0; i < n; i++) {
            v = rb_randinrange(r, n);
            u = rb_randinrange(r, n);
            t = p[v]; p[v] = p[u]; p[u] = t;
        }
    } else {
        for (i = 0; i < m; i++) {
            v = rb_randinrange(r, n - 3);
            u = rb_randinrange(r, n - 3);
            t = p[v + 0]; p[v + 0] = p[u + 2]; p[u + 2] = t;
            t = p[v + 1]; p[v + 1] = p[u + 3]; p[u + 3] = t;
            t = p[v + 2]; p[v + 2] = p[u + 0]; p[u + 0] = t;
            t = p[v + 3]; p[v + 3] = p[u + 1]; p[u + 1] = t;
        }
    }
}

/* ------------------------------------------------ max-heap priority queue
 * hunyuangraph_priorityqueue.h:22-262 (binary max-heap with locator). */
typedef struct { int key, val; } rkv;
typedef struct { int nnodes; rkv *heap; int *locator; } pqueue;

static pqueue *pq_create(int maxnodes)
{
    pqueue *q = (pqueue *)malloc(sizeof(pqueue));
    q->nnodes = 0;
    q->heap = (rkv *)malloc(sizeof(rkv) * (size_t)(maxnodes > 0 ? maxnodes : 1));
    q->locator = (int *)malloc(sizeof(int) * (size_t)(maxnodes > 0 ? maxnodes : 1));
    for (int i = 0; i < maxnodes; i++) q->locator[i] = -1;
    return q;
}

static void pq_free(pqueue *q) { free(q->heap); free(q->locator); free(q); }

static void pq_reset(pqueue *q)
{
    for (int i = q->nnodes - 1; i >= 0; i--) q->locator[q->heap[i].val] = -1;
    q->nnodes = 0;
}

static void pq_insert(pqueue *q, int node, int key)
{
    int i = q->nnodes++, j;
    while (i > 0) {
        j = (i - 1) >> 1;
        if (key > q->heap[j].key) { q->heap[i] = q->heap[j]; q->locator[q->heap[i].val] = i; i = j; }
        else break;
    }
    q->heap[i].key = key; q->heap[i].val = node; q->locator[node] = i;
}

static void pq_sift(pqueue *q, int i, int node, int key)
{
    int j, nn = q->nnodes;
    while ((j = 2 * i + 1) < nn) {
        if (q->heap[j].key > key) {
            if (j + 1 < nn && q->heap[j + 1].key > q->heap[j].key) j++;
            q->heap[i] = q->heap[j]; q->locator[q->heap[i].val] = i; i = j;
        } else if (j + 1 < nn && q->heap[j + 1].key > key) {
            j++;
            q->heap[i] = q->heap[j]; q->locator[q->heap[i].val] = i; i = j;
        } else break;
    }
    q->heap[i].key = key; q->heap[i].val = node; q->locator[node] = i;
}

static int pq_top(pqueue *q)
{
    if (q->nnodes == 0) return -1;
    q->nnodes--;
    int vtx = q->heap[0].val;
    q->locator[vtx] = -1;
    if (q->nnodes > 0) {
        int key = q->heap[q->nnodes].key, node = q->heap[q->nnodes].val;
        pq_sift(q, 0, node, key);
    }
    return vtx;
}

static void pq_update(pqueue *q, int node, int newkey)
{
    int i = q->locator[node], j, oldkey = q->heap[i].key;
    if (newkey > oldkey) {
        while (i > 0) {
            j = (i - 1) >> 1;
            if (newkey > q->heap[j].key) { q->heap[i] = q->heap[j]; q->locator[q->heap[i].val] = i; i = j; }
            else break;
        }
        q->heap[i].key = newkey; q->heap[i].val = node; q->locator[node] = i;
    } else {
        pq_sift(q, i, node, newkey);
    }
}

static void pq_delete(pqueue *q, int node)
{
    int i = q->locator[node];
    q->locator[node] = -1;
    if (--q->nnodes > 0 && q->heap[q->nnodes].val != node) {
        int nd = q->heap[q->nnodes].val, newkey = q->heap[q->nnodes].key, oldkey = q->heap[i].key, j;
        if (newkey > oldkey) {
            while (i > 0) {
                j = (i - 1) >> 1;
                if (newkey > q->heap[j].key) { q->heap[i] = q->heap[j]; q->locator[q->heap[i].val] = i; i = j; }
                else break;
            }
            q->heap[i].key = newkey; q->heap[i].val = nd; q->locator[nd] = i;
        } else {
            pq_sift(q, i, nd, newkey);
        }
    }
}

/* ------------------------------------------------ graph record */
typedef struct rbgraph {
    int nvtxs, nedges, tvwgt;
    float tvwgt_rev;
    int *xadj, *vwgt, *adjncy, *adjwgt, *label, *cmap;
    int *where, *pwgts, *bndptr, *bndlist, *id, *ed;
    int nbnd, mincut;
    int owns;                 /* 0 for the caller's level-0 arrays */
    struct rbgraph *coarser, *finer;
} rbgraph;

typedef struct {
    rbrng rng;
    int threshold, niter, maxvwgt;
    float ub;
    float part_balance[2];
} rbadmin;

#define LISTINSERT(n, list, lptr, i) do { list[n] = i; lptr[i] = (n)++; } while (0)
#define LISTDELETE(n, list, lptr, i) do { list[lptr[i]] = list[--(n)]; lptr[list[n]] = lptr[i]; lptr[i] = -1; } while (0)

static rbgraph *g_new(void) { return (rbgraph *)calloc(1, sizeof(rbgraph)); }

static void g_free(rbgraph *g)
{
    if (!g) return;
    if (g->owns) { free(g->xadj); free(g->vwgt); free(g->adjncy); free(g->adjwgt); }
    free(g->label); free(g->cmap); free(g->where); free(g->pwgts); free(g->bndptr);
    free(g->bndlist); free(g->id); free(g->ed);
    free(g);
}

static void g_set_tvwgt(rbgraph *g)
{
    int s = 0;
    for (int i = 0; i < g->nvtxs; i++) s += g->vwgt[i];
    g->tvwgt = s;
    g->tvwgt_rev = 1.0f / (float)(s > 0 ? s : 1);
}

static void alloc_2way(rbgraph *g)
{
    int n = g->nvtxs > 0 ? g->nvtxs : 1;
    free(g->pwgts); free(g->where); free(g->bndptr); free(g->bndlist); free(g->id); free(g->ed);
    g->pwgts = (int *)malloc(2 * sizeof(int));
    g->where = (int *)malloc((size_t)n * sizeof(int));
    g->bndptr = (int *)malloc((size_t)n * sizeof(int));
    g->bndlist = (int *)malloc((size_t)n * sizeof(int));
    g->id = (int *)malloc((size_t)n * sizeof(int));
    g->ed = (int *)malloc((size_t)n * sizeof(int));
}

/* hunyuangraph_compute_cpu_2wayparam (hunyuangraph_CPU_2wayrefine.h:22-75) */
static void compute_2wayparam(rbgraph *g)
{
    int nbnd = 0, mincut = 0;
    g->pwgts[0] = g->pwgts[1] = 0;
    for (int i = 0; i < g->nvtxs; i++) { g->bndptr[i] = -1; g->pwgts[g->where[i]] += g->vwgt[i]; }
    for (int i = 0; i < g->nvtxs; i++) {
        int me = g->where[i], tid = 0, ted = 0;
        for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
            if (me == g->where[g->adjncy[j]]) tid += g->adjwgt[j]; else ted += g->adjwgt[j];
        }
        g->id[i] = tid; g->ed[i] = ted;
        if (ted > 0 || g->xadj[i] == g->xadj[i + 1]) { LISTINSERT(nbnd, g->bndlist, g->bndptr, i); mincut += ted; }
    }
    g->mincut = mincut / 2;
    g->nbnd = nbnd;
}

/* ------------------------------------------------ coarsening */

/* hunyuangraph_cpu_create_cgraph (hunyuangraph_CPU_contraction.h:9-98) */
static rbgraph *create_cgraph(rbgraph *g, int cnvtxs, const int *match)
{
    rbgraph *c = g_new();
    c->owns = 1;
    c->nvtxs = cnvtxs;
    c->xadj = (int *)malloc(sizeof(int) * (size_t)(cnvtxs + 1));
    c->adjncy = (int *)malloc(sizeof(int) * (size_t)(g->nedges > 0 ? g->nedges : 1));
    c->adjwgt = (int *)malloc(sizeof(int) * (size_t)(g->nedges > 0 ? g->nedges : 1));
    c->vwgt = (int *)malloc(sizeof(int) * (size_t)(cnvtxs > 0 ? cnvtxs : 1));
    c->finer = g;
    g->coarser = c;
    int *htable = (int *)malloc(sizeof(int) * (size_t)(cnvtxs > 0 ? cnvtxs : 1));
    for (int i = 0; i < cnvtxs; i++) htable[i] = -1;
    int *cmap = g->cmap, cn = 0, cne = 0;
    int *cadj = c->adjncy, *cw = c->adjwgt;
    c->xadj[0] = 0;
    for (int v = 0; v < g->nvtxs; v++) {
        int u = match[v];
        if (u < v) continue;
        c->vwgt[cn] = g->vwgt[v];
        int ne = 0;
        for (int j = g->xadj[v]; j < g->xadj[v + 1]; j++) {
            int k = cmap[g->adjncy[j]], m;
            if ((m = htable[k]) == -1) { cadj[ne] = k; cw[ne] = g->adjwgt[j]; htable[k] = ne++; }
            else cw[m] += g->adjwgt[j];
        }
        if (v != u) {
            c->vwgt[cn] += g->vwgt[u];
            for (int j = g->xadj[u]; j < g->xadj[u + 1]; j++) {
                int k = cmap[g->adjncy[j]], m;
                if ((m = htable[k]) == -1) { cadj[ne] = k; cw[ne] = g->adjwgt[j]; htable[k] = ne++; }
                else cw[m] += g->adjwgt[j];
            }
            int j;
            if ((j = htable[cn]) != -1) {
                cadj[j] = cadj[--ne];
                cw[j] = cw[ne];
                htable[cn] = -1;
            }
        }
        for (int j = 0; j < ne; j++) htable[cadj[j]] = -1;
        cne += ne;
        c->xadj[++cn] = cne;
        cadj += ne;
        cw += ne;
    }
    c->nedges = cne;
    g_set_tvwgt(c);
    free(htable);
    return c;
}

/* Match_2HopAny (hunyuangraph_CPU_match.h:37-98).  Match_2HopAll (:100-160)
 * is a no-op in the reference (match is re-pointed at a zeroed array, so every
 * candidate is skipped) and is omitted. */
static int match_2hop_any(rbgraph *g, const int *perm, int *match, int cnvtxs, int *r_nun, int maxdegree)
{
    int n = g->nvtxs, nun = *r_nun;
    int *colptr = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; i++)
        if (match[i] == -1 && g->xadj[i + 1] - g->xadj[i] < maxdegree)
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) colptr[g->adjncy[j]]++;
    for (int i = 1; i < n; i++) colptr[i] += colptr[i - 1];
    for (int i = n; i > 0; i--) colptr[i] = colptr[i - 1];
    colptr[0] = 0;
    int *rowind = (int *)malloc(sizeof(int) * (size_t)(colptr[n] > 0 ? colptr[n] : 1));
    for (int pi = 0; pi < n; pi++) {
        int i = perm[pi];
        if (match[i] == -1 && g->xadj[i + 1] - g->xadj[i] < maxdegree)
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) rowind[colptr[g->adjncy[j]]++] = i;
    }
    for (int i = n; i > 0; i--) colptr[i] = colptr[i - 1];
    colptr[0] = 0;
    for (int pi = 0; pi < n; pi++) {
        int i = perm[pi];
        if (colptr[i + 1] - colptr[i] < 2) continue;
        for (int jj = colptr[i + 1], j = colptr[i]; j < jj; j++) {
            if (match[rowind[j]] == -1) {
                for (jj--; jj > j; jj--) {
                    if (match[rowind[jj]] == -1) {
                        g->cmap[rowind[j]] = g->cmap[rowind[jj]] = cnvtxs++;
                        match[rowind[j]] = rowind[jj];
                        match[rowind[jj]] = rowind[j];
                        nun -= 2;
                        break;
                    }
                }
            }
        }
    }
    free(colptr);
    free(rowind);
    *r_nun = nun;
    return cnvtxs;
}

/* Match_2Hop (hunyuangraph_CPU_match.h:162-172) */
static int match_2hop(rbgraph *g, const int *perm, int *match, int cnvtxs, int nun)
{
    cnvtxs = match_2hop_any(g, perm, match, cnvtxs, &nun, 2);
    if (nun > 1.5 * 0.1 * g->nvtxs) cnvtxs = match_2hop_any(g, perm, match, cnvtxs, &nun, 3);
    if (nun > 2.0 * 0.1 * g->nvtxs) cnvtxs = match_2hop_any(g, perm, match, cnvtxs, &nun, g->nvtxs);
    return cnvtxs;
}

static int finish_match(rbgraph *g, int *match)
{
    int cn = 0;
    for (int i = 0; i < g->nvtxs; i++) {
        if (match[i] == -1) { match[i] = i; g->cmap[i] = cn++; }
        else if (i <= match[i]) g->cmap[i] = g->cmap[match[i]] = cn++;
    }
    return cn;
}

/* hunyuangraph_cpu_match_RM (hunyuangraph_CPU_match.h:194-282) */
static void match_rm(rbadmin *a, rbgraph *g)
{
    int n = g->nvtxs, cn = 0, last = 0, nun = 0;
    int *match = (int *)malloc(sizeof(int) * (size_t)n);
    int *perm = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) match[i] = -1;
    rb_randarrayofp(&a->rng, n, perm, n / 8, 1);
    for (int pi = 0; pi < n; pi++) {
        int i = perm[pi];
        if (match[i] != -1) continue;
        int maxidx = i;
        if (g->vwgt[i] < a->maxvwgt) {
            if (g->xadj[i] == g->xadj[i + 1]) {
                last = (pi > last ? pi : last) + 1;
                for (; last < n; last++) {
                    int j = perm[last];
                    if (match[j] == -1) { maxidx = j; break; }
                }
            } else {
                for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                    int k = g->adjncy[j];
                    if (match[k] == -1 && g->vwgt[i] + g->vwgt[k] <= a->maxvwgt) { maxidx = k; break; }
                }
                if (maxidx == i && 3 * g->vwgt[i] < a->maxvwgt) { nun++; maxidx = -1; }
            }
        }
        if (maxidx != -1) { g->cmap[i] = g->cmap[maxidx] = cn++; match[i] = maxidx; match[maxidx] = i; }
    }
    if (nun > 0.1 * n) cn = match_2hop(g, perm, match, cn, nun);
    cn = finish_match(g, match);
    create_cgraph(g, cn, match);
    free(match);
    free(perm);
}

/* hunyuangraph_cpu_match_HEM (hunyuangraph_CPU_match.h:285-402) */
static void match_hem(rbadmin *a, rbgraph *g)
{
    int n = g->nvtxs, cn = 0, last = 0;
    int *match = (int *)malloc(sizeof(int) * (size_t)n);
    int *perm = (int *)malloc(sizeof(int) * (size_t)n);
    int *tperm = (int *)malloc(sizeof(int) * (size_t)n);
    int *d = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) match[i] = -1;
    rb_randarrayofp(&a->rng, n, tperm, n / 8, 1);
    int aved = (int)(0.7 * (g->xadj[n] / n));
    for (int i = 0; i < n; i++) {
        int dg = g->xadj[i + 1] - g->xadj[i];
        d[i] = dg > aved ? aved : dg;
    }
    /* hunyuangraph_matching_sort (:10-31) */
    int *counts = (int *)calloc((size_t)aved + 2, sizeof(int));
    for (int i = 0; i < n; i++) counts[d[i]]++;
    for (int i = 1; i < aved + 1; i++) counts[i] += counts[i - 1];
    for (int i = aved + 1; i > 0; i--) counts[i] = counts[i - 1];
    counts[0] = 0;
    for (int ii = 0; ii < n; ii++) { int i = tperm[ii]; perm[counts[d[i]]++] = i; }
    free(counts);
    for (int pi = 0; pi < n; pi++) {
        int i = perm[pi];
        if (match[i] != -1) continue;
        int maxidx = i, maxwgt = -1;
        if (g->vwgt[i] < a->maxvwgt) {
            if (g->xadj[i] == g->xadj[i + 1]) {
                last = (pi > last ? pi : last) + 1;
                for (; last < n; last++) {
                    int j = perm[last];
                    if (match[j] == -1) { maxidx = j; break; }
                }
            } else {
                for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                    int k = g->adjncy[j];
                    if (match[k] == -1 && maxwgt < g->adjwgt[j] && g->vwgt[i] + g->vwgt[k] <= a->maxvwgt) {
                        maxidx = k; maxwgt = g->adjwgt[j];
                    }
                }
                if (maxidx == i && 3 * g->vwgt[i] < a->maxvwgt) maxidx = -1;
            }
        }
        if (maxidx != -1) { g->cmap[i] = g->cmap[maxidx] = cn++; match[i] = maxidx; match[maxidx] = i; }
    }
    cn = finish_match(g, match);
    create_cgraph(g, cn, match);
    free(match); free(perm); free(tperm); free(d);
}

/* hunyuangraph_cpu_coarsen (hunyuangraph_CPU_coarsen.h:11-46) */
static rbgraph *cpu_coarsen(rbadmin *a, rbgraph *g)
{
    int eq = 1;
    for (int i = 1; i < g->nedges; i++) if (g->adjwgt[0] != g->adjwgt[i]) { eq = 0; break; }
    a->maxvwgt = (int)(1.5 * g->tvwgt / a->threshold);
    do {
        if (!g->cmap) g->cmap = (int *)malloc(sizeof(int) * (size_t)(g->nvtxs > 0 ? g->nvtxs : 1));
        if (eq || g->nedges == 0) match_rm(a, g); else match_hem(a, g);
        g = g->coarser;
    } while (g->nvtxs > a->threshold && g->nvtxs < 0.85 * g->finer->nvtxs && g->nedges > g->nvtxs / 2);
    return g;
}

/* ------------------------------------------------ balance and FM */

/* hunyuangraph_bndvertex_2way_bal (hunyuangraph_balance.h:27-126) */
static void bndvertex_2way_bal(rbadmin *a, rbgraph *g, const float *ntpwgts)
{
    int n = g->nvtxs, tp[2], from, to, nbnd = g->nbnd;
    int *pw = g->pwgts, *where = g->where, *id = g->id, *ed = g->ed, *bndptr = g->bndptr, *bndlist = g->bndlist;
    int *moved = (int *)malloc(sizeof(int) * (size_t)n), *perm = (int *)malloc(sizeof(int) * (size_t)n);
    tp[0] = (int)(g->tvwgt * ntpwgts[0]);
    tp[1] = g->tvwgt - tp[0];
    int mindiff = abs(tp[0] - pw[0]);
    from = pw[0] < tp[0] ? 1 : 0;
    to = (from + 1) % 2;
    pqueue *q = pq_create(n);
    for (int i = 0; i < n; i++) moved[i] = -1;
    rb_randarrayofp(&a->rng, nbnd, perm, nbnd / 5, 1);
    for (int ii = 0; ii < nbnd; ii++) {
        int i = bndlist[perm[ii]];
        if (where[i] == from && g->vwgt[i] <= mindiff) pq_insert(q, i, ed[i] - id[i]);
    }
    int mincut = g->mincut;
    for (int nswaps = 0; nswaps < n; nswaps++) {
        int hg = pq_top(q);
        if (hg == -1) break;
        if (pw[to] + g->vwgt[hg] > tp[to]) break;
        mincut -= ed[hg] - id[hg];
        pw[to] += g->vwgt[hg]; pw[from] -= g->vwgt[hg];
        where[hg] = to;
        moved[hg] = nswaps;
        { int t = id[hg]; id[hg] = ed[hg]; ed[hg] = t; }
        if (ed[hg] == 0 && g->xadj[hg] < g->xadj[hg + 1]) LISTDELETE(nbnd, bndlist, bndptr, hg);
        for (int j = g->xadj[hg]; j < g->xadj[hg + 1]; j++) {
            int k = g->adjncy[j];
            int kw = to == where[k] ? g->adjwgt[j] : -g->adjwgt[j];
            id[k] += kw; ed[k] -= kw;
            int inq = moved[k] == -1 && where[k] == from && g->vwgt[k] <= mindiff;
            if (bndptr[k] != -1) {
                if (ed[k] == 0) { LISTDELETE(nbnd, bndlist, bndptr, k); if (inq) pq_delete(q, k); }
                else if (inq) pq_update(q, k, ed[k] - id[k]);
            } else if (ed[k] > 0) {
                LISTINSERT(nbnd, bndlist, bndptr, k);
                if (inq) pq_insert(q, k, ed[k] - id[k]);
            }
        }
    }
    g->mincut = mincut;
    g->nbnd = nbnd;
    pq_free(q);
    free(moved); free(perm);
}

/* hunyuangraph_2way_bal (hunyuangraph_balance.h:128-140) */
static void two_way_bal(rbadmin *a, rbgraph *g, const float *ntpwgts)
{
    float mx = -1.0f;
    for (int j = 0; j < 2; j++) {
        float cur = g->pwgts[j] * a->part_balance[j] - a->ub;
        if (cur > mx) mx = cur;
    }
    if (mx <= 0) return;
    if (fabsf(ntpwgts[0] * g->tvwgt - g->pwgts[0]) < 3 * g->tvwgt / g->nvtxs) return;
    bndvertex_2way_bal(a, g, ntpwgts);
}

/* hunyuangraph_cpu_2way_refine (hunyuangraph_CPU_2wayrefine.h:77-242) */
static void two_way_refine(rbadmin *a, rbgraph *g, const float *ntpwgts, int niter)
{
    int n = g->nvtxs, tp[2];
    int *pw = g->pwgts, *where = g->where, *id = g->id, *ed = g->ed, *bndptr = g->bndptr, *bndlist = g->bndlist;
    int *moved = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *swaps = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *perm = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    tp[0] = (int)(g->tvwgt * ntpwgts[0]);
    tp[1] = g->tvwgt - tp[0];
    int limit = (int)(0.01 * n);
    if (limit < 15) limit = 15;
    if (limit > 100) limit = 100;
    int avgvwgt = (pw[0] + pw[1]) / 20;
    if (n > 0 && 2 * (pw[0] + pw[1]) / n < avgvwgt) avgvwgt = 2 * (pw[0] + pw[1]) / n;
    pqueue *qs[2] = {pq_create(n), pq_create(n)};
    int origdiff = abs(tp[0] - pw[0]);
    for (int i = 0; i < n; i++) moved[i] = -1;
    for (int pass = 0; pass < niter; pass++) {
        pq_reset(qs[0]); pq_reset(qs[1]);
        int mincutorder = -1, newcut, mincut, initcut, nswaps;
        newcut = mincut = initcut = g->mincut;
        int mindiff = abs(tp[0] - pw[0]);
        int nbnd = g->nbnd;
        rb_randarrayofp(&a->rng, nbnd, perm, nbnd, 1);
        for (int ii = 0; ii < nbnd; ii++) {
            int i = bndlist[perm[ii]];
            pq_insert(qs[where[i]], i, ed[i] - id[i]);
        }
        for (nswaps = 0; nswaps < n; nswaps++) {
            int from = tp[0] - pw[0] < tp[1] - pw[1] ? 0 : 1, to = (from + 1) % 2;
            int hg = pq_top(qs[from]);
            if (hg == -1) break;
            newcut -= ed[hg] - id[hg];
            pw[to] += g->vwgt[hg]; pw[from] -= g->vwgt[hg];
            if ((newcut < mincut && abs(tp[0] - pw[0]) <= origdiff + avgvwgt) ||
                (newcut == mincut && abs(tp[0] - pw[0]) < mindiff)) {
                mincut = newcut; mindiff = abs(tp[0] - pw[0]); mincutorder = nswaps;
            } else if (nswaps - mincutorder > limit) {
                newcut += ed[hg] - id[hg];
                pw[from] += g->vwgt[hg]; pw[to] -= g->vwgt[hg];
                break;
            }
            where[hg] = to;
            moved[hg] = nswaps;
            swaps[nswaps] = hg;
            { int t = id[hg]; id[hg] = ed[hg]; ed[hg] = t; }
            if (ed[hg] == 0 && g->xadj[hg] < g->xadj[hg + 1]) LISTDELETE(nbnd, bndlist, bndptr, hg);
            for (int j = g->xadj[hg]; j < g->xadj[hg + 1]; j++) {
                int k = g->adjncy[j];
                int kw = to == where[k] ? g->adjwgt[j] : -g->adjwgt[j];
                id[k] += kw; ed[k] -= kw;
                if (bndptr[k] != -1) {
                    if (ed[k] == 0) {
                        LISTDELETE(nbnd, bndlist, bndptr, k);
                        if (moved[k] == -1) pq_delete(qs[where[k]], k);
                    } else if (moved[k] == -1) pq_update(qs[where[k]], k, ed[k] - id[k]);
                } else if (ed[k] > 0) {
                    LISTINSERT(nbnd, bndlist, bndptr, k);
                    if (moved[k] == -1) pq_insert(qs[where[k]], k, ed[k] - id[k]);
                }
            }
        }
        for (int i = 0; i < nswaps; i++) moved[swaps[i]] = -1;
        for (nswaps--; nswaps > mincutorder; nswaps--) {
            int hg = swaps[nswaps];
            int to = where[hg] = (where[hg] + 1) % 2;
            { int t = id[hg]; id[hg] = ed[hg]; ed[hg] = t; }
            if (ed[hg] == 0 && bndptr[hg] != -1 && g->xadj[hg] < g->xadj[hg + 1]) LISTDELETE(nbnd, bndlist, bndptr, hg);
            else if (ed[hg] > 0 && bndptr[hg] == -1) LISTINSERT(nbnd, bndlist, bndptr, hg);
            pw[to] += g->vwgt[hg]; pw[(to + 1) % 2] -= g->vwgt[hg];
            for (int j = g->xadj[hg]; j < g->xadj[hg + 1]; j++) {
                int k = g->adjncy[j];
                int kw = to == where[k] ? g->adjwgt[j] : -g->adjwgt[j];
                id[k] += kw; ed[k] -= kw;
                if (bndptr[k] != -1 && ed[k] == 0) LISTDELETE(nbnd, bndlist, bndptr, k);
                if (bndptr[k] == -1 && ed[k] > 0) LISTINSERT(nbnd, bndlist, bndptr, k);
            }
        }
        g->mincut = mincut;
        g->nbnd = nbnd;
        if (mincutorder <= 0 || mincut == initcut) break;
    }
    pq_free(qs[0]); pq_free(qs[1]);
    free(moved); free(swaps); free(perm);
}

/* hunyuangraph_2way_project (hunyuangraph_CPU_2wayrefine.h:245-321) */
static void two_way_project(rbgraph *g)
{
    rbgraph *c = g->coarser;
    alloc_2way(g);
    int nbnd = 0;
    for (int i = 0; i < g->nvtxs; i++) {
        int j = g->cmap[i];
        g->where[i] = c->where[j];
        g->cmap[i] = c->bndptr[j];
    }
    for (int i = 0; i < g->nvtxs; i++) {
        int tid = 0, ted = 0;
        g->bndptr[i] = -1;
        if (g->cmap[i] == -1) {
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) tid += g->adjwgt[j];
        } else {
            int me = g->where[i];
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                if (me == g->where[g->adjncy[j]]) tid += g->adjwgt[j]; else ted += g->adjwgt[j];
            }
        }
        g->id[i] = tid; g->ed[i] = ted;
        if (ted > 0 || g->xadj[i] == g->xadj[i + 1]) LISTINSERT(nbnd, g->bndlist, g->bndptr, i);
    }
    g->mincut = c->mincut;
    g->nbnd = nbnd;
    g->pwgts[0] = c->pwgts[0];
    g->pwgts[1] = c->pwgts[1];
    g_free(c);
    g->coarser = NULL;
}

/* huyuangraph_cpu_growbisection (hunyuangraph_CPU_initialpartition.h:36-145) */
static void grow_bisection(rbadmin *a, rbgraph *g, const float *ntpwgts, int niparts)
{
    int n = g->nvtxs, bestcut = 0;
    alloc_2way(g);
    int *bestwhere = (int *)malloc(sizeof(int) * (size_t)n);
    int *queue = (int *)malloc(sizeof(int) * (size_t)n);
    int *tra = (int *)malloc(sizeof(int) * (size_t)n);
    int onemax = (int)(a->ub * g->tvwgt * ntpwgts[1]);
    int onemin = (int)((1.0 / a->ub) * g->tvwgt * ntpwgts[1]);
    for (int iter = 0; iter < niparts; iter++) {
        int pw[2], first, last, nleft, dd, i;
        for (i = 0; i < n; i++) { g->where[i] = 1; tra[i] = 0; }
        pw[1] = g->tvwgt; pw[0] = 0;
        queue[0] = rb_randinrange(&a->rng, n);
        tra[queue[0]] = 1;
        first = 0; last = 1; nleft = n - 1; dd = 0;
        for (;;) {
            if (first == last) {
                if (nleft == 0 || dd) break;
                int k = rb_randinrange(&a->rng, nleft);
                for (i = 0; i < n; i++) {
                    if (tra[i] == 0) { if (k == 0) break; else k--; }
                }
                queue[0] = i; tra[i] = 1; first = 0; last = 1; nleft--;
            }
            i = queue[first++];
            if (pw[0] > 0 && pw[1] - g->vwgt[i] < onemin) { dd = 1; continue; }
            g->where[i] = 0;
            pw[0] += g->vwgt[i]; pw[1] -= g->vwgt[i];
            if (pw[1] <= onemax) break;
            dd = 0;
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                int k = g->adjncy[j];
                if (tra[k] == 0) { queue[last++] = k; tra[k] = 1; nleft--; }
            }
        }
        compute_2wayparam(g);
        two_way_bal(a, g, ntpwgts);
        two_way_refine(a, g, ntpwgts, a->niter);
        if (iter == 0 || bestcut > g->mincut) {
            bestcut = g->mincut;
            memcpy(bestwhere, g->where, sizeof(int) * (size_t)n);
            if (bestcut == 0) break;
        }
    }
    g->mincut = bestcut;
    memcpy(g->where, bestwhere, sizeof(int) * (size_t)n);
    free(bestwhere); free(queue); free(tra);
}

/* hunyuangraph_cpu_refinement (hunyuangraph_CPU_2wayrefine.h:335-386) */
static void cpu_refinement(rbadmin *a, rbgraph *org, rbgraph *g, const float *tp)
{
    compute_2wayparam(g);
    for (;;) {
        two_way_bal(a, g, tp);
        two_way_refine(a, g, tp, a->niter);
        if (g == org) break;
        g = g->finer;
        two_way_project(g);
    }
}

/* hunyuangraph_cpu_mlevelbisect (hunyuangraph_CPU_initialpartition.h:148-281), ncuts = 1 */
static int mlevel_bisect(rbadmin *a, rbgraph *g, const float *tp)
{
    a->part_balance[0] = g->tvwgt_rev / tp[0];
    a->part_balance[1] = g->tvwgt_rev / tp[1];
    rbgraph *c = cpu_coarsen(a, g);
    int niparts = c->nvtxs <= a->threshold ? 5 : 7;
    grow_bisection(a, c, tp, niparts);
    cpu_refinement(a, g, c, tp);
    return g->mincut;
}

/* hunyuangraph_splitgraph / _first (hunyuangraph_CPU_splitgraph.h:8-346) */
static void split_graph(rbgraph *g, int first, rbgraph **rl, rbgraph **rr)
{
    int n = g->nvtxs, sn[2] = {0, 0}, se[2] = {0, 0};
    int *ren = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        int k = g->where[i];
        ren[i] = sn[k]++;
        se[k] += g->xadj[i + 1] - g->xadj[i];
    }
    rbgraph *s[2];
    for (int p = 0; p < 2; p++) {
        s[p] = g_new();
        s[p]->owns = 1;
        s[p]->nvtxs = sn[p];
        s[p]->xadj = (int *)malloc(sizeof(int) * (size_t)(sn[p] + 1));
        s[p]->vwgt = (int *)malloc(sizeof(int) * (size_t)(sn[p] + 1));
        s[p]->adjncy = (int *)malloc(sizeof(int) * (size_t)(se[p] > 0 ? se[p] : 1));
        s[p]->adjwgt = (int *)malloc(sizeof(int) * (size_t)(se[p] > 0 ? se[p] : 1));
        s[p]->label = (int *)malloc(sizeof(int) * (size_t)(sn[p] > 0 ? sn[p] : 1));
        s[p]->xadj[0] = 0;
        sn[p] = se[p] = 0;
    }
    for (int i = 0; i < n; i++) {
        int mp = g->where[i];
        if (g->bndptr[i] == -1) {
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                s[mp]->adjncy[se[mp]] = g->adjncy[j];
                s[mp]->adjwgt[se[mp]++] = g->adjwgt[j];
            }
        } else {
            for (int j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
                int k = g->adjncy[j];
                if (g->where[k] == mp) { s[mp]->adjncy[se[mp]] = k; s[mp]->adjwgt[se[mp]++] = g->adjwgt[j]; }
            }
        }
        s[mp]->vwgt[sn[mp]] = g->vwgt[i];
        s[mp]->label[sn[mp]] = first ? i : g->label[i];
        s[mp]->xadj[++sn[mp]] = se[mp];
    }
    for (int p = 0; p < 2; p++) {
        for (int i = 0; i < s[p]->xadj[sn[p]]; i++) s[p]->adjncy[i] = ren[s[p]->adjncy[i]];
        s[p]->nedges = se[p];
        g_set_tvwgt(s[p]);
    }
    free(ren);
    *rl = s[0];
    *rr = s[1];
}

/* hunyuangraph_mlevel_rbbisection (hunyuangraph_CPU_initialpartition.h:284-358) */
static int mlevel_rb(rbadmin *a, rbgraph *g, int nparts, int *part, float *tpwgts, int fpart, int level)
{
    if (g->nvtxs == 0) { g_free(g); return 0; }
    int n = g->nvtxs;
    float tp2[2];
    float ws = 0;
    for (int i = 0; i < (nparts >> 1); i++) ws += tpwgts[i];
    tp2[0] = ws;
    tp2[1] = 1.0f - tp2[0];
    int obj = mlevel_bisect(a, g, tp2);
    level++;
    if (level == 1) for (int i = 0; i < n; i++) part[i] = g->where[i] + fpart;
    else for (int i = 0; i < n; i++) part[g->label[i]] = g->where[i] + fpart;
    rbgraph *lg = NULL, *rg = NULL;
    if (nparts > 2) split_graph(g, level == 1, &lg, &rg);
    g_free(g);
    float wsum = 0;
    for (int i = 0; i < (nparts >> 1); i++) wsum += tpwgts[i];
    for (int i = 0; i < (nparts >> 1); i++) tpwgts[i] *= 1.0f / wsum;
    for (int i = 0; i < nparts - (nparts >> 1); i++) tpwgts[(nparts >> 1) + i] *= 1.0f / (1.0f - wsum);
    if (nparts > 3) {
        obj += mlevel_rb(a, lg, nparts >> 1, part, tpwgts, fpart, level);
        obj += mlevel_rb(a, rg, nparts - (nparts >> 1), part, tpwgts + (nparts >> 1), fpart + (nparts >> 1), level);
    } else if (nparts == 3) {
        g_free(lg);
        obj += mlevel_rb(a, rg, nparts - (nparts >> 1), part, tpwgts + (nparts >> 1), fpart + (nparts >> 1), level);
    }
    return obj;
}

int64_t oracle_cpu_rb_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                                const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                                uint64_t seed, int32_t *part)
{
    if (n <= 0) return 0;
    if (k <= 1) { for (int i = 0; i < n; i++) part[i] = 0; return 0; }
    rbadmin a;
    a.rng.s = seed ? seed * 0x9E3779B97F4A7C15ull : 0x1234567ull;
    a.threshold = 200;
    a.niter = 10;
    a.ub = 1.03f;
    a.maxvwgt = 0;
    rbgraph *g = g_new();
    g->owns = 1;
    g->nvtxs = n;
    g->nedges = (int)xadj[n];
    g->xadj = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i <= n; i++) g->xadj[i] = (int)xadj[i];
    g->adjncy = (int *)malloc(sizeof(int) * (size_t)(g->nedges > 0 ? g->nedges : 1));
    g->adjwgt = (int *)malloc(sizeof(int) * (size_t)(g->nedges > 0 ? g->nedges : 1));
    g->vwgt = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(g->adjncy, adjncy, sizeof(int) * (size_t)g->nedges);
    memcpy(g->adjwgt, adjwgt, sizeof(int) * (size_t)g->nedges);
    memcpy(g->vwgt, vwgt, sizeof(int) * (size_t)n);
    g_set_tvwgt(g);
    float *tpw = (float *)malloc(sizeof(float) * (size_t)k);
    for (int i = 0; i < k; i++) tpw[i] = 1.0f / k;
    mlevel_rb(&a, g, k, part, tpw, 0, 0);
    free(tpw);
    return oracle_edge_cut(n, xadj, adjncy, adjwgt, part);
}
