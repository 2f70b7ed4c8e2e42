/*
 * hgp_oracle.h -- CPU oracle for the multilevel k-way partitioning path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may use it, and only as the checker (or, for the CPU baseline,
 * as the timed host path), never as the thing measured on the GPU.
 *
 * Three groups of functions:
 *
 *  (1) Exact primitives -- closed-form integer arithmetic that the GPU must
 *      reproduce bit for bit: inclusive scans, part weights, edge cut, the
 *      exact 1.03 balance test, cmap of a fixed matching and contraction of a
 *      fixed matching into the canonical coarse CSR.
 *
 *  (2) "DGA" -- a serial emulation of the deterministic GPU algorithm: the
 *      reference GPU pipeline (hunyuangraph_PartitionGraph,
 *      hunyuangraph_partitiongraph.h:284) with every atomics-ordered or
 *      CAS-raced decision replaced by a deterministic rule.  The GPU library
 *      implements the same rules, so its part[] must equal oracle_dga_partition
 *      bit for bit on the same inputs and seed.
 *
 *  (3) "CPU RB" -- a restatement of the reference CPU path
 *      hunyuangraph_rbbisection (hunyuangraph_CPU_initialpartition.h:361):
 *      serial METIS-style multilevel recursive bisection.  It is the timed
 *      host-CPU baseline.  The reference draws from glibc rand() (never
 *      seeded, hunyuangraph_common.h:34-47); this restatement draws from a
 *      seeded xorshift generator instead, so its random sequences (visit
 *      order, GGGP start vertices, FM permutations) are "parity unpinned".
 *
 * Index conventions: vertex ids int32, edge offsets (xadj) int64, weights
 * int32, sums int64.  CSR is symmetric, 0-based, no self-loops.
 */
#ifndef HGP_ORACLE_H
#define HGP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- (1) */

/* Inclusive in-place scan (hunyuangraph_GPU_prefixsum.h:141-188 computes an
 * inclusive in-place scan; int32 wraps exactly as two's complement). */
void oracle_scan_i32(int32_t *a, int64_t n);
void oracle_scan_i64(int64_t *a, int64_t n);

/* pwgts[p] = sum of vwgt[v] over where[v]==p  (hunyuangraph_GPU_uncoarsen.h:19-45). */
void oracle_part_weights(int32_t n, const int32_t *vwgt, const int32_t *where,
                         int32_t k, int64_t *pwgts);

/* Edge cut = sum over directed entries with where[u]!=where[v] of adjwgt, /2
 * (hunyuangraph_graph.h:144-155). */
int64_t oracle_edge_cut(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *adjwgt, const int32_t *where);

/* Exact balance test: 100*k*max_p pwgts[p] <= 103*tvwgt, generalised to
 * ub_num/ub_den.  Returns 1 if balanced. */
int oracle_balanced(int32_t n, const int32_t *vwgt, const int32_t *where, int32_t k,
                    int64_t ub_num, int64_t ub_den);

/* cmap of a fixed matching (hunyuangraph_GPU_match.h:1012-1024 + scan :3088):
 * match[v]==-1 or an unreciprocated entry is treated as a singleton and
 * rewritten to v; cmap[v] = number of representatives (v<=match[v]) in
 * 1..v (vertex 0 maps to 0); the partner's cmap copies the representative's.
 * Returns cnvtxs.  match is normalised in place. */
int32_t oracle_cmap(int32_t n, int32_t *match, int32_t *cmap);

/* Contraction of a fixed (normalised) matching into the canonical coarse CSR
 * (hunyuangraph_CPU_contraction.h:9-98, hunyuangraph_GPU_contraction.h:715-1247):
 * rows = concatenated rows of both endpoints renamed by cmap, self entries
 * dropped, duplicates merged by summing weights, sorted by coarse id.
 * Output arrays must hold cn+1, m (fine m is an upper bound), cn entries.
 * Returns the coarse directed entry count. */
int64_t oracle_contract(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                        const int32_t *adjwgt, const int32_t *vwgt,
                        const int32_t *match, const int32_t *cmap, int32_t cn,
                        int64_t *cxadj, int32_t *cadjncy, int32_t *cadjwgt,
                        int32_t *cvwgt);

/* ---------------------------------------------------------------- (2) */

typedef struct oracle_opts {
    uint64_t seed;
    int32_t ub_num, ub_den;     /* final balance bound, default 103/100          */
    int32_t match_rounds;       /* handshake rounds per level, default 8         */
    int32_t trials;             /* initial-bisection trials, default 1024        */
    int32_t fm_passes;          /* FM passes per trial, default 10               */
    int32_t patience;           /* Jet non-improving rounds, default 6           */
    int32_t fm_kind;            /* 0 wave arg-max FM, 1 heap FM with the same    */
                                /* moves, 2 the reference's heap FM (default 0)  */
    int32_t refiner;            /* 0 Jet (default), 1 the older label            */
                                /* propagation (oracle_lp_refine)                */
} oracle_opts;

typedef struct oracle_stats {
    int32_t levels;             /* number of coarse levels built                 */
    int32_t coarsest_n;
    int64_t coarsest_m;
    int64_t init_cut;           /* cut of the initial partition (coarsest level) */
    int64_t cut;                /* final cut (exact)                             */
    int64_t max_pwgt;
    int64_t refine_iters;       /* total Jet iterations over all levels          */
    int64_t level_n[64];
    int64_t level_m[64];
    int32_t level_iters[64];    /* Jet iterations per level (index = level)      */
    int64_t level_adjwgt[64];   /* total edge weight of each level               */
} oracle_stats;

void oracle_default_opts(oracle_opts *o);

/* Full deterministic pipeline.  part[] out.  Returns 0 on success, negative on
 * invalid input. */
int oracle_dga_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                         const oracle_opts *opts, int32_t *part, oracle_stats *st);

/* The coarsening phase of oracle_dga_partition alone: the coarsest graph
 * (cxadj n+1, cadjncy/cadjwgt xadj[n], cvwgt n entries, caller-allocated) and
 * the level statistics.  Returns 0, or -1 for n < 1 or k < 1. */
int oracle_dga_coarsen(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                       const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                       const oracle_opts *opts, int32_t *cn, int64_t *cm, int64_t *cxadj,
                       int32_t *cadjncy, int32_t *cadjwgt, int32_t *cvwgt, oracle_stats *st);

/* Stage entry points of the DGA, exposed for stage-by-stage parity tests. */

/* Coarsening parameters (hunyuangraph_partitiongraph.h:298-302,
 * hunyuangraph_GPU_coarsen.h:165). */
int32_t oracle_coarsen_threshold(int32_t k);
int32_t oracle_maxvwgt(int64_t tvwgt, int32_t threshold);

/* One level of matching, before normalisation: match[v] = partner or -1.
 * Returns the number of matched vertices. */
int64_t oracle_dga_match(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t maxvwgt,
                         uint64_t seed, int32_t level, int32_t rounds, int32_t *match);

/* Edge priority hash of the handshake matching. */
/* One matching fallback pass (0 leaf, 1 isolate, 2 twin, 3 relative,
 * 4 forced) on its own, match updated in place. */
void oracle_fb_apply(int32_t kind, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                     const int32_t *adjwgt, const int32_t *vwgt, int32_t maxvwgt, int32_t *match);
uint32_t oracle_edge_hash(uint64_t seed, int32_t level, int32_t lo, int32_t hi);

/* Recursive-bisection initial partition of a (coarsest) graph into k parts. */
void oracle_dga_initpart(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                         const oracle_opts *opts, int32_t *where);

/* One GGGP+FM trial on a graph (exposed for tests): returns the cut, fills
 * where (0/1).  node_id and trial select the random stream. */
int64_t oracle_dga_trial(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, const int32_t *adjwgt,
                         int64_t tvwgt, int64_t onemin, int64_t onemax, int64_t ideal0,
                         int32_t ntrials, int32_t trial, uint64_t seed, uint32_t node_id,
                         int32_t fm_passes, int32_t fm_kind, uint8_t *where);

/* The device priority queue alone (hunyuangraph_GPU_priorityqueue.h:27-225):
 * applies nops operations (op 0 insert (v, key), 1 top, 2 delete v,
 * 3 update (v, key)) to one queue over vertices [0, nv); tops[i] receives
 * what a top returned; final_val/final_key (nv slots) and final_loc (nv) the
 * final layout.  ties != 0 orders equal keys by the smaller vertex.  Returns
 * the final count. */
int32_t oracle_heap_run(int32_t nv, int32_t nops, const int32_t *op, const int32_t *v, const int64_t *key,
                        int32_t ties, int32_t *tops, int32_t *final_val, int64_t *final_key, int32_t *final_loc);

/* Trace of the Jet loops of the following calls into buf (cap int64; NULL
 * stops): records of 5 int64 (level vertices, iterations done, mode of the
 * coming iteration, cut so far, 1 on the loop's last record), as the GPU
 * library's hgp_jet_trace; levels of <= 512 vertices record nothing.
 * oracle_jet_trace_len returns the values recorded (also past cap). */
void oracle_jet_trace(int64_t *buf, int64_t cap);
int64_t oracle_jet_trace_len(void);
/* Keeps where (before the moves) and dest of iteration `iter` of the level
 * with n vertices in the caller's arrays (NULL stops); oracle_jet_dumped
 * returns 1 once that iteration ran. */
void oracle_jet_dump(int64_t n, int32_t iter, int32_t *where, int32_t *dest);
int32_t oracle_jet_dumped(void);

/* Jet refinement of one level in place.  Returns the number of iterations.
 * cut is in/out. */
int32_t oracle_dga_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                          const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                          int64_t tvwgt, int32_t is_finest, const oracle_opts *opts,
                          int32_t *where, int64_t *cut);

/* One Jet label-propagation step (select + afterburner) without applying the
 * moves: dest[v] = new part for selected vertices, else where[v].
 * lock is in/out with the reference's lock semantics.  Returns #moves. */
int64_t oracle_jet_lp_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, int32_t k, const int32_t *where,
                           int32_t filter_num, int32_t filter_den,
                           uint8_t *lock, int32_t *dest);

/* Integer restatement of gain_bucket (hunyuangraph_GPU_krefine.h:2477-2513). */
int32_t oracle_gain_bucket(int64_t gain, int64_t vwgt);

/* Jet rebalancing steps on their own (weak: jetrw, strong: jetrs), the
 * cookie cutter of jetrs and the acceptance rule of k_refine, for tests. */
int64_t oracle_jet_rw_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, const int32_t *vwgt, int32_t k,
                           const int32_t *where, const int64_t *pwgts, int64_t maxwgt,
                           int64_t opt, int32_t *dest);
int64_t oracle_jet_rs_step(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                           const int32_t *adjwgt, const int32_t *vwgt, int32_t k,
                           const int32_t *where, const int64_t *pwgts, int64_t maxwgt,
                           int64_t opt, int32_t *dest);
void oracle_cookie_cutter(int32_t k, int32_t ne, const int64_t *bs, int64_t maxwgt,
                          const int64_t *pwgts, int32_t *es, int32_t *ee);
int oracle_jet_accept(int64_t cur_cut, int64_t curmax, int64_t maxwgt, int64_t *best_cut,
                      int64_t *best_max, int32_t *count);

/* The older label propagation the reference keeps beside Jet
 * (hunyuangraph_k_refinement_me, hunyuangraph_GPU_LP.h:1245-1374), one level
 * in place: up to `threshold` rounds that move, into each part, the unmoved
 * vertex of highest positive gain whose best part it is, then up to
 * `threshold` rebalancing rounds between overweight and underweight parts.
 * iters[0] = rounds that moved a vertex, iters[1] = rebalancing rounds;
 * *cut receives the exact cut afterwards. */
int32_t oracle_lp_threshold(int32_t n, int32_t k);
void oracle_lp_part_bounds(int32_t k, int64_t tvwgt, float *fmin, float *fmax);
void oracle_lp_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                      const int32_t *vwgt, const int32_t *adjwgt, int32_t k, int64_t tvwgt,
                      int32_t *where, int64_t *cut, int32_t *iters);

/* The reference's 2-way label propagation (jetLP, hunyuangraph_GPU_LP.h:
 * 673-781): one synchronous flip pass, then, when a side is outside
 * [tvwgt/2/1.03, tvwgt/2*1.03], one sort-based transfer from the heavier
 * side.  where (0/1) in place; returns 1 if the transfer ran. */
int32_t oracle_lp2_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                          const int32_t *vwgt, const int32_t *adjwgt, int32_t *where);

/* ---------------------------------------------------------------- (3) */

/* Reference CPU path restated: multilevel recursive bisection with the
 * standalone parameters of hunyuangraph_rbbisection (threshold 200, 1 cut,
 * GGGP x5/7, FM x10, ubfactor 1.03 per bisection).  Returns the cut. */
int64_t oracle_cpu_rb_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy,
                                const int32_t *vwgt, const int32_t *adjwgt, int32_t k,
                                uint64_t seed, int32_t *part);

#ifdef __cplusplus
}
#endif
#endif
