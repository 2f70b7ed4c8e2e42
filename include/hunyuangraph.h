/*
 * hunyuangraph.h -- C ABI of the MI355X multilevel k-way graph partitioner.
 *
 * The first symbol is the drop-in replacement of the reference's entry point;
 * everything else is the versioned API (64-bit edge offsets, honoured weights,
 * a seed, status codes) plus stage entry points used by the parity tests.
 * All pointers are plain host pointers unless a function says "device".
 * No function exits the process; errors are returned as negative status codes
 * and described by hgp_last_error().
 */
#ifndef HUNYUANGRAPH_H
#define HUNYUANGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ drop-in */

/* Replaces hunyuangraph_PartitionGraph (hunyuangraph_partitiongraph.h:284).
 * Same signature and CSR contract: int32 xadj/adjncy, nvtxs vertices,
 * *nparts parts, part[] out.  Unlike the reference, vwgt and adjwgt are
 * honoured (NULL means unit weights), *ubvec (if non-NULL and > 1) is the
 * balance bound (default 1.03, enforced exactly), tpwgts must be NULL or
 * uniform.  On error part[] is left unspecified and hgp_last_error() says why
 * (the reference exits the process instead). */
void hunyuangraph_PartitionGraph(int *nvtxs, int *xadj, int *adjncy, int *vwgt, int *adjwgt,
                                 int *nparts, float *tpwgts, float *ubvec, int *part);

/* ------------------------------------------------------------ status */

enum {
    HGP_OK = 0,
    HGP_ERR_INVALID = -1,     /* malformed CSR, k < 1, weights <= 0, ...          */
    HGP_ERR_DEVICE = -2,      /* HIP runtime error                                 */
    HGP_ERR_OOM = -3,         /* device memory or HBM budget exhausted             */
    HGP_ERR_UNSUPPORTED = -4, /* outside the supported range (e.g. k too large)    */
    HGP_ERR_NODEVICE = -5,    /* no HIP device present                              */
    HGP_ERR_IO = -6           /* a file could not be opened, read or written        */
};

/* Text of the last error on the calling thread ("" if none). */
const char *hgp_last_error(void);
/* Library version string. */
const char *hgp_version(void);
/* Number of HIP devices (0 if none; never errors). */
int hgp_device_count(void);

/* ------------------------------------------------------------ options */

typedef struct hgp_options {
    uint64_t seed;          /* random stream of matching ties and trial starts   */
    int32_t ub_num;         /* balance bound ub_num/ub_den (default 103/100)     */
    int32_t ub_den;
    int32_t match_rounds;   /* handshake rounds per level (default 8)            */
    int32_t trials;         /* initial bisection trials per node (default 1024)  */
    int32_t fm_passes;      /* FM passes per trial (default 10)                  */
    int32_t patience;       /* Jet non-improving rounds (default 6)              */
    int32_t validate;       /* 1: also check CSR symmetry before running         */
    int32_t device;         /* HIP device ordinal (default 0)                    */
    int32_t level_stats;    /* 1: fill level_adjwgt and level_ms of hgp_stats    */
    int32_t fm_kind;        /* FM of the bisection trials: 0 wave arg-max scan,  */
                            /* 1 device priority queues with the same moves,     */
                            /* 2 the reference's priority-queue FM               */
                            /* (FM_2way_cut_refinement); -1 (default) picks 0 or */
                            /* 1 by the coarsest graph's size (same result)      */
    int32_t refiner;        /* refinement of each level: 0 Jet (default), 1 the  */
                            /* older label propagation the reference keeps beside*/
                            /* it (hunyuangraph_k_refinement_me,                 */
                            /* hunyuangraph_GPU_LP.h:1245)                       */
} hgp_options;

#define HGP_MAX_LEVELS 64

typedef struct hgp_stats {
    int64_t cut;            /* final edge cut (exact)                            */
    int64_t max_pwgt;       /* heaviest part                                     */
    int64_t tvwgt;          /* total vertex weight                               */
    int32_t levels;         /* coarse levels built                               */
    int32_t coarsest_n;
    int64_t coarsest_m;
    int64_t init_cut;       /* cut of the initial partition                      */
    int64_t refine_iters;   /* Jet iterations over all levels                    */
    double ms_total;        /* device time, CSR resident -> part resident        */
    double ms_coarsen, ms_init, ms_uncoarsen;
    double alg_bytes;       /* algorithmic HBM bytes of the run                  */
    int64_t level_n[HGP_MAX_LEVELS];
    int64_t level_m[HGP_MAX_LEVELS];
    int32_t level_iters[HGP_MAX_LEVELS];
    int64_t arena_bytes;    /* device arena the run used                         */
    int64_t arena_peak;     /* most of it in use at once                         */
    /* with level_stats: total edge weight of each level (sum of adjwgt over
     * the directed entries) and the device time of building level l+1 from
     * level l (level_ms[levels] = 0) */
    int64_t level_adjwgt[HGP_MAX_LEVELS];
    double level_ms[HGP_MAX_LEVELS];
} hgp_stats;

void hgp_default_options(hgp_options *opt);

/* ------------------------------------------------------------ partition */

/* Full pipeline on host arrays (copied to the device and back).
 * xadj has n+1 int64 offsets; vwgt/adjwgt may be NULL (unit weights). */
int hgp_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                  const int32_t *adjwgt, int32_t k, const hgp_options *opt, int32_t *part,
                  hgp_stats *stats);

/* Full pipeline on device-resident arrays (device pointers on opt->device);
 * d_part (device, n entries) receives the partition.  The caller's arrays are
 * not modified.  ms_total in stats is the device time of this call. */
int hgp_partition_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy,
                         const int32_t *d_vwgt, const int32_t *d_adjwgt, int32_t k,
                         const hgp_options *opt, int32_t *d_part, hgp_stats *stats);

/* Checks a device-resident CSR (device pointers on `device`) with the
 * invariants of check_graph (hunyuangraph_GPU_match.h:1718-1800): xadj[0] = 0,
 * monotone offsets, vertex weights >= 1, neighbour ids in [0, n), no
 * self-loops, edge weights >= 1, and a total edge weight below 2^31.  With
 * symmetric != 0, also that every edge appears in both rows with the same
 * weight (compared through two 64-bit hash sums of the edge multiset, so an
 * asymmetric graph passes with probability about 2^-128).  vwgt/adjwgt may be
 * NULL (unit weights).  Returns HGP_OK or HGP_ERR_INVALID, with the first
 * fault (the smallest vertex showing it) in hgp_last_error().  The offsets
 * are proved monotone before any row is read, so a malformed xadj never
 * sends a kernel outside adjncy.  hgp_partition_device runs this check on
 * every call (symmetric when opt->validate is set). */
int hgp_check_graph_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_vwgt,
                           const int32_t *d_adjwgt, int32_t symmetric, int32_t device);

/* Quality of a device-resident partition (the device forms of
 * compute_edgecut_gpu, hunyuangraph_graph.h:184-268, and of the part-weight
 * sums of calculateSum, hunyuangraph_GPU_uncoarsen.h:19-45).  d_out (device,
 * 4 int64) receives {edge cut, heaviest part weight, number of vertices whose
 * part id is outside [0, k), number of entries whose neighbour id is outside
 * [0, n)}; such vertices and entries are counted, never dereferenced.  The
 * offsets must be valid (as hgp_check_graph_device accepts them); vwgt/adjwgt
 * may be NULL (unit weights); k <= 4096.  Runs on the library's stream of
 * `device` and waits for it. */
int hgp_evaluate_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_adjwgt,
                        const int32_t *d_vwgt, const int32_t *d_part, int32_t k, int32_t device, int64_t *d_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * evaluation and arguments as hgp_evaluate_device, with its scratch in a
 * fixed per-device array of the library (4 + 4096 int64 = 32.8 KB, held
 * until the process ends) instead of the library's arena, so it calls no
 * hipMalloc or hipFree and creates no stream.  k <= 4096 (larger k returns
 * HGP_ERR_UNSUPPORTED before any launch).  Runs on the null stream of
 * `device` and waits for it.  Every call on a device shares that device's
 * array: calls are serialised by a per-device lock, and each call clears the
 * array by a kernel before use. */
int hgp_evaluate_device_noalloc(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_adjwgt,
                                const int32_t *d_vwgt, const int32_t *d_part, int32_t k, int32_t device,
                                int64_t *d_out);


/* ------------------------------------------------------------ CPU path
 * The reference's CPU partitioner, hunyuangraph_rbbisection
 * (hunyuangraph_CPU_initialpartition.h:361-374): serial multilevel recursive
 * bisection on the host (random or heavy-edge matching with two-hop
 * matching, greedy graph growing, boundary balancing and FM with priority
 * queues, k-1 bisections), with the standalone admin's parameters
 * (hunyuangraph_admin.h:7-31: coarsening threshold 200, one cut, ten FM
 * passes, ubfactor 1.03 per bisection).  A path the caller chooses; the GPU
 * entry points never fall back to it.  Its random choices come from `seed`
 * (the reference calls rand() unseeded).  vwgt/adjwgt may be NULL (unit
 * weights); m and the total vertex weight must fit int32, as in the
 * reference.  *cut (optional) receives the partition's edge cut. */
int hgp_partition_cpu(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                      const int32_t *adjwgt, int32_t k, uint64_t seed, int32_t *part, int64_t *cut);

/* ------------------------------------------------------------ tuning
 * No counterpart in the reference (its launch shapes are compile-time).
 * Scheduling knobs that change how fast a partition is computed, never its
 * result: every path they choose between computes the same exact partition.
 * hunyuansc25_amd/autotune.py measures them on the device in use and writes
 * them to hunyuansc25_amd/tuning.json, which the Python binding applies on
 * load.  Set them while no partition runs. */
typedef struct hgp_tuning {
    int32_t tiny_max_vertices;  /* levels of <= this many vertices (k <= 256) run the
                                   whole refinement loop in one block (0..4096, default 512) */
    int32_t few_rows;           /* levels with <= this many rows of <= 512 entries run
                                   those rows in one launch (default 8192) */
    int32_t fill_waves;         /* waves the wave-per-row kernels spread a level over (default 8192) */
    int32_t seg_grid;           /* grid cap of the long-row segment kernels (default 32768) */
    int32_t batch_lanes;        /* lanes of hgp_partition_batch when it is passed <= 0 (1..64, default 8) */
    int32_t hub_track;          /* 1: rows of > 512 entries keep their part-connectivity rows
                                   current through a level's iterations, updated by every move
                                   (k <= 512); 0: rebuilt by each selection (default 0) */
    int32_t graph_max_vertices; /* levels of <= this many vertices record their refinement
                                   launch sequences once as HIP graphs and replay them in
                                   every iteration (0: never, the default) */
    int32_t batch_fuse_max_vertices; /* hgp_partition_batch refines the levels of <= this many
                                   vertices (the coarse end of each hierarchy, k <= 256) of
                                   every graph of a group in one launch, a block per graph
                                   (0: each graph on its own; 0..16384, default 8192) */
} hgp_tuning;
int hgp_get_tuning(hgp_tuning *t);
int hgp_set_tuning(const hgp_tuning *t);

/* Trace of the Jet refinement loops (the reference prints per-iteration
 * state only under debug macros, hunyuangraph_GPU_krefine.h:4434-4882).
 * hgp_jet_trace_enable(1) clears and starts recording, 0 stops; while on,
 * every read of a level's control block appends 5 int64: the level's vertex
 * count, the iterations done, the mode of the coming iteration (0 label
 * propagation, 1 weak rebalance, 2 strong rebalance), the cut so far, and 1
 * on the read that ends the loop (else 0).  Levels of <= 512 vertices run
 * their loop in one kernel and record nothing.  hgp_jet_trace copies up to
 * cap values into buf and returns how many are recorded. */
int hgp_jet_trace_enable(int on);
int64_t hgp_jet_trace(int64_t *buf, int64_t cap);
/* While tracing, also keep one iteration's state: at iteration `iter` of the
 * level with n vertices, where[] before the iteration's moves and dest[] (the
 * part each vertex moves to, its own part if it stays).  hgp_jet_trace_dumped
 * copies up to cap entries of each and returns the level's n (0 if the
 * iteration was not reached). */
int hgp_jet_trace_dump(int64_t n, int32_t iter);
int64_t hgp_jet_trace_dumped(int32_t *where, int32_t *dest, int64_t cap);

/* ------------------------------------------------------------ batches
 * No counterpart in the reference, which partitions one graph per call.
 * Partitions `count` independent graphs: graph i is the CSR n[i], xadj[i]
 * (int64, n[i]+1 entries), adjncy[i], vwgt[i], adjwgt[i] with k[i] parts, and
 * part[i] (n[i] entries) receives its partition.  vwgt/adjwgt may be NULL, or
 * hold NULL entries, for unit weights.  The graphs run concurrently on
 * `lanes` streams of opt->device (<= 0: the tuning's batch_lanes), each lane with its own arena and
 * staging buffers; part[i] equals what hgp_partition returns for graph i
 * alone with the same options.  status[i] (optional) receives each graph's
 * status and stats[i] (optional) its statistics.  Returns HGP_OK when every
 * graph succeeded, else the status of the lowest-index failing graph, which
 * hgp_last_error() describes. */
int hgp_partition_batch(int32_t count, const int32_t *n, const int64_t *const *xadj,
                        const int32_t *const *adjncy, const int32_t *const *vwgt,
                        const int32_t *const *adjwgt, const int32_t *k, const hgp_options *opt,
                        int32_t lanes, int32_t *const *part, int32_t *status, hgp_stats *stats);

/* ------------------------------------------------------------ serving
 * No counterpart in the reference.  A request loop on one device: up to
 * `slots` jobs run at once, each on its own stream and arena.  Jobs are
 * admitted first come, first served, when a slot is free and the arenas of
 * all slots fit in `budget_bytes` of device memory (arenas of idle slots are
 * released when a new job needs the room).  Jobs that arrive while others run
 * overlap with them on the device, and each job's part[] equals what
 * hgp_partition returns for the same graph and options. */
typedef struct hgp_server hgp_server;

/* opt (may be NULL) gives the device and the defaults of every job. */
int hgp_server_create(const hgp_options *opt, int32_t slots, int64_t budget_bytes, hgp_server **out);
/* Queues a job (validated now).  The caller keeps every array alive until
 * hgp_server_wait returns for it.  Returns a ticket >= 0 or a negative status;
 * a job whose arena alone exceeds the budget is refused with HGP_ERR_OOM. */
int64_t hgp_server_submit(hgp_server *s, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                          const int32_t *vwgt, const int32_t *adjwgt, int32_t k, uint64_t seed,
                          int32_t *part);
/* Blocks until the job is done and returns its status; stats and latency_ms
 * (submit to result, host clock) are optional outputs.  A ticket is waited
 * for once: a second wait on it, concurrent or later, returns
 * HGP_ERR_INVALID and leaves the job to the first. */
int hgp_server_wait(hgp_server *s, int64_t ticket, hgp_stats *stats, double *latency_ms);
/* Resident graphs.  hgp_server_register uploads a graph once (validated,
 * NULL weights = unit) and returns a graph id >= 0, or a negative status
 * (HGP_ERR_OOM when it does not fit the budget); the host arrays may be
 * freed afterwards.  hgp_server_submit_graph queues a partition of a
 * resident graph into k parts with `seed` (part[] is written before the
 * wait returns; returns a ticket or a negative status).  Partitions of a
 * resident graph keep their coarsening hierarchy on the device: a later
 * request with the same seed and coarsening threshold (hunyuangraph_partitiongraph.h:298-302:
 * every k in 128..512 shares one) skips the coarsening.  Kept hierarchies
 * count against the budget and are dropped least recently used first when
 * a job, graph or hierarchy needs the room; results equal hgp_partition of
 * the same graph in every case.  hgp_server_unregister waits for the
 * graph's queued jobs, then frees it and its hierarchies. */
int64_t hgp_server_register(hgp_server *s, int32_t n, const int64_t *xadj, const int32_t *adjncy,
                            const int32_t *vwgt, const int32_t *adjwgt);
int64_t hgp_server_submit_graph(hgp_server *s, int64_t graph, int32_t k, uint64_t seed, int32_t *part);
int hgp_server_unregister(hgp_server *s, int64_t graph);
/* Hierarchy reuse so far: requests that found their hierarchy kept (hits)
 * or built it (misses), and the hierarchies and bytes kept now. */
int hgp_server_kept(hgp_server *s, int64_t *hits, int64_t *misses, int64_t *kept, int64_t *kept_bytes);
/* Device bytes held by the slots' arenas, the resident graphs and the kept
 * hierarchies, now and at most so far, and the number of finished jobs (any
 * output may be NULL). */
int hgp_server_info(hgp_server *s, int64_t *reserved, int64_t *peak_reserved, int64_t *completed);
/* Finishes queued jobs, stops the workers, waits until every
 * hgp_server_wait in flight has returned, and frees the device memory.  No
 * call on the server may start once destroy has begun. */
int hgp_server_destroy(hgp_server *s);

/* ------------------------------------------------------------ kernel timer
 * No counterpart entry point in the reference (its phase timers are process
 * globals, hunyuangraph_timer.h:9-319).  Per-kernel timing with HIP events on
 * the library's stream, for measurement: when enabled, the main kernels record an event pair per launch
 * and the algorithmic HBM bytes the launch must move.  enable(0/1) clears the
 * records.  report() writes JSON {"kernel": [launches, total_ms, bytes], ...}
 * into buf (NUL-terminated, truncated to len) and returns the length needed. */
int hgp_ktimer_enable(int on);
int64_t hgp_ktimer_report(char *buf, int64_t len);

/* ------------------------------------------------------------ stages
 * Host-array entry points of single stages, for parity tests.  Each copies
 * its inputs to the device, runs the GPU stage and copies the result back. */

/* Inclusive scans (hunyuangraph_GPU_prefixsum.h:141-188). in and out may alias. */
int hgp_scan_i32(const int32_t *in, int32_t *out, int64_t n);
int hgp_scan_i64(const int64_t *in, int64_t *out, int64_t n);

/* Part weights (hunyuangraph_GPU_uncoarsen.h:19-45) and edge cut
 * (hunyuangraph_graph.h:184-268). */
int hgp_part_weights(int32_t n, const int32_t *vwgt, const int32_t *where, int32_t k,
                     int64_t *pwgts);
int hgp_edge_cut(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                 const int32_t *where, int64_t *cut);

/* One level of matching (hunyuangraph_gpu_match, hunyuangraph_GPU_match.h:1846):
 * match[v] = partner or -1 (unmatched), before normalisation. */
int hgp_match(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
              const int32_t *adjwgt, int32_t maxvwgt, uint64_t seed, int32_t level,
              int32_t rounds, int32_t *match, int64_t *matched);

/* cmap of a matching plus contraction into the canonical coarse CSR
 * (hunyuangraph_gpu_create_cgraph, hunyuangraph_GPU_contraction.h:715).
 * match (n) is normalised in place (unmatched -> self).  Output buffers:
 * cmap n, cxadj n+1, cadjncy/cadjwgt xadj[n], cvwgt n.  *cn, *cm out. */
int hgp_contract(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                 const int32_t *vwgt, int32_t *match, int32_t *cmap, int32_t *cn,
                 int64_t *cxadj, int32_t *cadjncy, int32_t *cadjwgt, int32_t *cvwgt,
                 int64_t *cm);

/* Recursive-bisection initial partition (hunyuangraph_gpu_initialpartition,
 * hunyuangraph_GPU_initialpartition.h:4405) of a graph into k parts. */
int hgp_initpart(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                 const int32_t *adjwgt, int32_t k, const hgp_options *opt, int32_t *where);

/* Jet refinement of one level (k_refine, hunyuangraph_GPU_krefine.h:4434);
 * where and *cut in/out, *iters out.  With opt->refiner = 1 the level runs
 * the older label propagation instead (hunyuangraph_k_refinement_me,
 * hunyuangraph_GPU_LP.h:1245); *iters then counts its rounds that moved a
 * vertex plus its rebalancing rounds. */
int hgp_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
               const int32_t *adjwgt, int32_t k, int64_t tvwgt, int32_t is_finest,
               const hgp_options *opt, int32_t *where, int64_t *cut, int32_t *iters);

/* The 2-way label propagation of jetLP (hunyuangraph_GPU_LP.h:673), which
 * the reference keeps beside its 2-way FM (FM_2WayCutRefine_GPU,
 * hunyuangraph_GPU_2wayrefine.h:653-674): one synchronous pass in which a
 * vertex whose edges into its own side weigh no more than those to the other
 * side changes side, then, when a side lies outside [tvwgt/2/1.03,
 * tvwgt/2*1.03], a transfer of the best-gain prefix of the heavier side.
 * where (0/1) in/out; *ran = 1 if the transfer ran.  Uses opt->device only. */
int hgp_refine_2way_lp(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                       const int32_t *adjwgt, const hgp_options *opt, int32_t *where, int32_t *ran);

/* ------------------------------------------------------------ graphs
 * No counterpart in the reference, which reads graphs from files
 * (hunyuangraph_io.h:25-186).  Seeded synthetic generators (host code).  The library allocates the
 * arrays; release them with hgp_graph_free. */
typedef struct hgp_graph {
    int32_t n;
    int64_t m;              /* directed entries = 2 x undirected edges */
    int64_t *xadj;
    int32_t *adjncy;
    int32_t *adjwgt;
    int32_t *vwgt;
} hgp_graph;

void hgp_graph_free(hgp_graph *g);
/* 5-point grid rows x cols, rows sorted, unit weights. */
int hgp_gen_grid2d(int32_t rows, int32_t cols, hgp_graph *g);
/* 3-D random geometric graph: n points in [0,1)^3, radius r; optional
 * weights vwgt U{vw_lo..vw_hi}, adjwgt U{ew_lo..ew_hi} (0,0 = unit). */
int hgp_gen_rgg3d(int32_t n, double radius, uint64_t seed, int32_t vw_lo, int32_t vw_hi,
                  int32_t ew_lo, int32_t ew_hi, hgp_graph *g);
/* R-MAT with 2^scale vertices and exactly nedges unique undirected edges,
 * self-loops and duplicates removed, seeded vertex relabelling. */
int hgp_gen_rmat(int32_t scale, int64_t nedges, double a, double b, double c, uint64_t seed,
                 hgp_graph *g);
/* Banded FEM-like pattern: |i-j| <= bw plus frac*n*bw random symmetric
 * off-band entries. */
int hgp_gen_banded(int32_t n, int32_t bw, double frac, uint64_t seed, hgp_graph *g);

/* ------------------------------------------------------------ files
 * Host code; the library allocates the graph arrays (release with
 * hgp_graph_free; on error the graph is left empty). */

/* Replaces hunyuangraph_readgraph (hunyuangraph_io.h:25-186).  METIS text
 * format: lines whose first character is '%' are skipped; header "n m [fmt
 * [ncon]]" with m counting each undirected edge once; fmt digits (zero-padded
 * to three) flag vertex sizes, vertex weights and edge weights; then one line
 * per vertex with 1-based neighbours.  Vertex weights must be >= 0, edge
 * weights > 0, neighbours in 1..n, and the rows must hold exactly 2m entries.
 * Differences from the reference: errors return HGP_ERR_INVALID with the
 * reference's reason in hgp_last_error() (it prints and exits), a vertex size
 * is read and ignored (the reference takes it for a neighbour), negative fmt
 * and ncon > 1 are refused. */
int hgp_read_graph(const char *path, hgp_graph *g);

/* Replaces hunyuangraph_writetofile (hunyuangraph_io.h:189-205): writes
 * part[0..n) one value per line to "<fname>.part.<nparts>". */
int hgp_write_partition(const char *fname, const int32_t *part, int32_t n, int32_t nparts);

/* Replaces the converter matrix_to_graph (matrix_to_graph/matrix_to_graph.c:
 * 117-281 with mmio_highlevel.h): a square Matrix Market coordinate matrix A
 * (pattern, real, integer or complex; general, symmetric or hermitian)
 * becomes the graph of pattern(A + A^T) without the diagonal, neighbours
 * ascending, unit vertex and edge weights.  Unlike the reference, duplicate
 * entries are merged whatever the entry order, and a non-square matrix is
 * refused. */
int hgp_read_mtx(const char *path, hgp_graph *g);

/* Writes g as a METIS .graph file readable by hgp_read_graph and by the
 * reference.  weights: 0 none, 1 edge weights, 10 vertex weights, 11 both
 * (the header's fmt field). */
int hgp_write_graph(const char *path, const hgp_graph *g, int32_t weights);

/* ------------------------------------------------------------ coarsening only */

/* The coarsening phase of hgp_partition alone (the same hierarchy for the
 * same graph, k and options), for the reference's study builds that stop
 * after coarsening: -DFIGURE9_SUM / -DFIGURE9_TIME print per-level sizes,
 * edge-weight sums and times, -DFIGURE10_CGRAPH dumps the coarsest graph
 * (hunyuangraph_partitiongraph.h:46-80, hunyuangraph_GPU_coarsen.h:198-242).
 * *coarsest receives the coarsest graph (library-allocated; release with
 * hgp_graph_free) and *stats (optional) levels, level_n, level_m,
 * ms_coarsen and, with opt->level_stats, level_adjwgt and level_ms. */
int hgp_coarsen(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                const int32_t *adjwgt, int32_t k, const hgp_options *opt, hgp_graph *coarsest,
                hgp_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
