// hgp_graph.hpp -- device graph record and the phase entry points of the
// multilevel pipeline.
#pragma once

#include "hgp_common.hpp"
#include "hgp_group.hpp"

#include <vector>

namespace hgp {

// One level of the hierarchy in HBM: CSR with 64-bit offsets, int32 ids and
// weights (xadj[n+1], adjncy[m], adjwgt[m], vwgt[n]).
struct DevGraph {
    int32_t n = 0;
    int64_t m = 0;
    eid_t *xadj = nullptr;
    vid_t *adjncy = nullptr;
    wgt_t *adjwgt = nullptr;
    wgt_t *vwgt = nullptr;
    RowBins bins;             // degree classes of the rows (set by attach_bins)
    RowSegs segs;             // segments of the class-2 rows (set by attach_bins)
};

// Computes g.bins and g.segs (on the left end of the arena: they live with the level).
void attach_bins(Ctx &c, DevGraph &g);

struct Params {
    uint64_t seed = 1;
    int32_t ub_num = 103, ub_den = 100;
    int32_t match_rounds = 8;
    int32_t trials = 1024;
    int32_t fm_passes = 10;
    int32_t patience = 6;
    int32_t fm_kind = -1;       // bisection-trial FM: 0 scan, 1 heap (same moves), 2 reference heap FM, -1 auto
    int32_t refiner = 0;        // level refinement: 0 Jet, 1 the older label propagation
    bool level_stats = false;   // per-level edge-weight sums and build times
};

// ---------------------------------------------------------------- coarsening

// hunyuangraph_partitiongraph.h:298-302 and hunyuangraph_GPU_coarsen.h:165
int32_t coarsen_threshold(int32_t k);
int32_t coarsen_maxvwgt(int64_t tvwgt, int32_t thr);

// One level of matching (device match[n] out: partner or -1).  Returns the
// number of matched vertices.
int64_t match_level(Ctx &c, const DevGraph &g, int32_t maxvwgt, uint64_t seed, int32_t level,
                    int32_t rounds, int32_t *match);

// Normalises match (unmatched -> self) and builds cmap; returns cnvtxs.
int32_t build_cmap(Ctx &c, int32_t n, int32_t *match, int32_t *cmap);

// Contracts g by a normalised matching; the coarse arrays are allocated on
// the left end of the arena.
DevGraph contract(Ctx &c, const DevGraph &g, const int32_t *match, const int32_t *cmap, int32_t cn);

// ---------------------------------------------------------------- initial partition

// Recursive bisection of g into k parts; where (device, n) out.
void initial_partition(Ctx &c, const DevGraph &g, int32_t k, const Params &p, int32_t *where);

// ---------------------------------------------------------------- refinement

struct RefineStats {
    int32_t iters = 0;
};

// Jet refinement of one level in place.  *d_cut (device int64) in/out.
RefineStats refine_level(Ctx &c, const DevGraph &g, int32_t k, int64_t tvwgt, bool finest,
                         const Params &p, int32_t *where, int64_t *d_cut);

// The older k-way label propagation of one level in place
// (hunyuangraph_k_refinement_me, hunyuangraph_GPU_LP.h:1245-1374); *d_cut
// (device int64) receives the exact cut.  iters = rounds that moved a vertex
// plus rebalancing rounds.
RefineStats lp_refine_level(Ctx &c, const DevGraph &g, int32_t k, int64_t tvwgt, int32_t *where, int64_t *d_cut);
int32_t lp_threshold(int32_t n, int32_t k);
void lp_part_bounds(int32_t k, int64_t tvwgt, float *fmin, float *fmax);

// The 2-way label propagation of jetLP (hunyuangraph_GPU_LP.h:673-781) on a
// 0/1 partition in place; returns whether the transfer step ran.
bool lp2_refine(Ctx &c, const DevGraph &g, int32_t *where);

// where_f[v] = where_c[cmap[v]]
void project(Ctx &c, int32_t n, const int32_t *cmap, const int32_t *where_c, int32_t *where_f);

// Exact cut into a device int64 (overwritten), part weights into pwgts[k].
void edge_cut(Ctx &c, const DevGraph &g, const int32_t *where, int64_t *d_cut);
void part_weights(Ctx &c, int32_t n, const wgt_t *vwgt, const int32_t *where, int32_t k, int64_t *pwgts);

// ---------------------------------------------------------------- small levels of many graphs

// Largest level (vertices) and k that the one-block Jet loop takes.
constexpr int32_t kTinyMaxN = 16384;
constexpr int32_t kTinyMaxK = 256;

// Device scratch of the one-block Jet loop for levels of up to n vertices
// and k parts (alloc_tiny_scratch: left end of the arena), plus the run's
// cut (int64; in: the coarsest partition's cut, out: the cut after the
// last level) and the Jet iterations of each level (HGP_MAX_LEVELS ints).
struct TinyScratch {
    int32_t *dest = nullptr, *pregain = nullptr, *best = nullptr, *elist = nullptr;
    int32_t *under = nullptr, *es = nullptr, *ee = nullptr, *pstart = nullptr, *iters = nullptr;
    uint8_t *lock = nullptr, *sel = nullptr, *w8 = nullptr;
    uint16_t *w16 = nullptr, *pd16 = nullptr;
    uint32_t *pd32 = nullptr, *key = nullptr, *val = nullptr, *k2 = nullptr, *v2 = nullptr;
    uint64_t *ab = nullptr;
    int64_t *wsc = nullptr, *bs = nullptr, *pwgts = nullptr, *cut = nullptr;
    void *ctl = nullptr;   // the Jet control block
};
TinyScratch alloc_tiny_scratch(Ctx &c, int32_t n, int32_t k);

// One graph's run of small levels: levels hi (the coarsest) down to lo, each
// refined by the one-block Jet loop and projected to the next finer level;
// the last projection fills where[lo - 1] when lo > 0.  lv, where and cmap
// are indexed by level; cmap[l] maps level l to level l + 1.
struct TinyRun {
    int32_t k = 0, hi = 0, lo = 0;
    int64_t tvwgt = 0;
    Params p;
    std::vector<DevGraph> lv;
    std::vector<int32_t *> where;
    std::vector<const int32_t *> cmap;
    TinyScratch s;
};

// Every run in one launch, one block per run (k <= kTinyMaxK, levels of <=
// kTinyMaxN vertices, checked by the caller); the same partitions as
// refine_level and project level by level.  Waits for the device.
void tiny_batch(Ctx &c, const std::vector<TinyRun> &runs);

}  // namespace hgp
