/*
 * oracle_prims.c -- exact integer primitives of the partitioning path.
 * TEST INFRASTRUCTURE ONLY (see hgp_oracle.h).
 */
#include "hgp_oracle.h"
#include <stdlib.h>
#include <string.h>

/* hunyuangraph_GPU_prefixsum.h:141-188 (inclusive, in place). */
void oracle_scan_i32(int32_t *a, int64_t n)
{
    uint32_t acc = 0;               /* two's complement wrap, like int32 hardware adds */
    for (int64_t i = 0; i < n; i++) {
        acc += (uint32_t)a[i];
        a[i] = (int32_t)acc;
    }
}

void oracle_scan_i64(int64_t *a, int64_t n)
{
    uint64_t acc = 0;
    for (int64_t i = 0; i < n; i++) {
        acc += (uint64_t)a[i];
        a[i] = (int64_t)acc;
    }
}

/* hunyuangraph_GPU_uncoarsen.h:19-45 (calculateSum). */
void oracle_part_weights(int32_t n, const int32_t *vwgt, const int32_t *where,
                         int32_t k, int64_t *pwgts)
{
    for (int32_t p = 0; p < k; p++) pwgts[p] = 0;
    for (int32_t v = 0; v < n; v++) pwgts[where[v]] += vwgt[v];
}

/* hunyuangraph_graph.h:144-155 (hunyuangraph_computecut_cpu). */
int64_t oracle_edge_cut(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *adjwgt, const int32_t *where)
{
    int64_t cut = 0;
    for (int32_t v = 0; v < n; v++)
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++)
            if (where[v] != where[adjncy[j]]) cut += adjwgt[j];
    return cut / 2;
}

/* Exact form of the balance bound (replaces the float check of
 * hunyuangraph_graph.h:157-171). */
int oracle_balanced(int32_t n, const int32_t *vwgt, const int32_t *where, int32_t k,
                    int64_t ub_num, int64_t ub_den)
{
    int64_t *pw = (int64_t *)calloc((size_t)k, sizeof(int64_t));
    int64_t tv = 0, mx = 0;
    for (int32_t v = 0; v < n; v++) { pw[where[v]] += vwgt[v]; tv += vwgt[v]; }
    for (int32_t p = 0; p < k; p++) if (pw[p] > mx) mx = pw[p];
    free(pw);
    return ub_den * (int64_t)k * mx <= ub_num * tv;
}

/* resolve_conflict_12 (hunyuangraph_GPU_match.h:1012-1024) followed by the
 * inclusive scan of cmap (:3088-3095) and resolve_conflict_4's copy of the
 * representative's id to the partner (:1027-1040). */
int32_t oracle_cmap(int32_t n, int32_t *match, int32_t *cmap)
{
    if (n <= 0) return 0;
    for (int32_t v = 0; v < n; v++) {
        int32_t u = match[v];
        if (u < 0 || u >= n || match[u] != v) match[v] = v;
    }
    /* second pass: a vertex whose partner was rewritten to a singleton in the
     * first pass must become a singleton too (pairs are symmetric after it). */
    for (int32_t v = 0; v < n; v++) {
        int32_t u = match[v];
        if (match[u] != v) match[v] = v;
    }
    int32_t acc = 0;
    for (int32_t v = 0; v < n; v++) {
        if (v > 0 && v <= match[v]) acc++;
        cmap[v] = acc;
    }
    for (int32_t v = 0; v < n; v++)
        if (v > match[v]) cmap[v] = cmap[match[v]];
    return acc + 1;
}

static int cmp_pair(const void *a, const void *b)
{
    const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

/* hunyuangraph_cpu_create_cgraph (hunyuangraph_CPU_contraction.h:9-98) with the
 * GPU path's canonical row order (rows sorted by coarse id,
 * hunyuangraph_GPU_contraction.h:886-1203). */
int64_t oracle_contract(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *adjwgt, const int32_t *vwgt,
                        const int32_t *match, const int32_t *cmap, int32_t cn,
                        int64_t *cxadj, int32_t *cadjncy, int32_t *cadjwgt,
                        int32_t *cvwgt)
{
    int32_t maxdeg = 0;
    for (int32_t v = 0; v < n; v++) {
        int64_t d = xadj[v + 1] - xadj[v];
        int32_t u = match[v];
        if (u != v) d += xadj[u + 1] - xadj[u];
        if (d > maxdeg) maxdeg = (int32_t)d;
    }
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(maxdeg + 1));
    int64_t cm = 0;
    cxadj[0] = 0;
    for (int32_t v = 0; v < n; v++) {
        int32_t u = match[v];
        if (u < v) continue;                       /* partner handled it */
        int32_t c = cmap[v];
        cvwgt[c] = vwgt[v] + (u != v ? vwgt[u] : 0);
        int32_t len = 0;
        for (int e = 0; e < (u != v ? 2 : 1); e++) {
            int32_t x = e ? u : v;
            for (int64_t j = xadj[x]; j < xadj[x + 1]; j++) {
                int32_t cc = cmap[adjncy[j]];
                if (cc == c) continue;             /* internal edge of the pair */
                buf[2 * len] = cc;
                buf[2 * len + 1] = adjwgt[j];
                len++;
            }
        }
        qsort(buf, (size_t)len, 2 * sizeof(int32_t), cmp_pair);
        for (int32_t i = 0; i < len;) {
            int32_t id = buf[2 * i];
            int32_t w = 0;
            while (i < len && buf[2 * i] == id) { w += buf[2 * i + 1]; i++; }
            cadjncy[cm] = id;
            cadjwgt[cm] = w;
            cm++;
        }
        cxadj[c + 1] = cm;
    }
    (void)cn;
    free(buf);
    return cm;
}
