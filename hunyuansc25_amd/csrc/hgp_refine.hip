// hgp_refine.hip -- Jet k-way refinement of one level (k_refine,
// hunyuangraph_GPU_krefine.h:4434-4882), projection, edge cut, part weights.
//
// Connectivity is recomputed from the CSR each iteration: kG lanes per
// vertex accumulate edge weights per part in a dense LDS row of k ints, then
// reduce the arg-max with xor shuffles.  Every decision reads the state at
// the start of the iteration, and the candidate orders of the rebalancers are
// (part, gain bucket, vertex id) via a stable radix sort, so the result does
// not depend on scheduling.
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"

#include "hunyuangraph.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace hgp {

namespace {

constexpr int kBS = 256;
constexpr int kMaxWaves = 1024 / kWave;   // waves of the largest block here
constexpr int32_t kNoGain = INT32_MIN;

enum Mode : int32_t { kModeLP = 0, kModeRW = 1, kModeRS = 2 };

struct JetCtl {
    int64_t cut, best_cut, best_max, maxwgt, opt, maxdest;
    unsigned long long delta2;      // 2 x cut change (two's complement)
    int32_t count, balance_counter, mode, done, iters, patience;
    int32_t ncand, nunder, copy_best, filt_num;   // ncand: 1 if an LP candidate exists
    int64_t nsorted;                // candidates of RW/RS
    int64_t nevict;
};

// Part ids of neighbours are gathered from a narrow copy of where/dest (uint8
// for k <= 256, else uint16): 1-2 bytes per vertex keep the gathered array
// cache-resident where the int32 original would not be.
struct NarrowParts {
    const uint8_t *p8;
    const uint16_t *p16;
    __device__ int32_t operator()(int32_t u) const { return p8 ? (int32_t)p8[u] : (int32_t)p16[u]; }
    // q[t] = part of u[t] for the live slots, -1 for the others
    template <int T>
    __device__ void gather(const int32_t (&u)[T], int32_t (&q)[T]) const
    {
        if (p8) gather_slots<T>(p8, u, q, -1);
        else gather_slots<T>(p16, u, q, -1);
    }
};

__global__ void k_narrow(int32_t n, const int32_t *__restrict__ src, uint8_t *__restrict__ d8,
                         uint16_t *__restrict__ d16)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    if (d8) d8[v] = (uint8_t)src[v];
    else d16[v] = (uint16_t)src[v];
}

// where and dest of a vertex packed into one word (8 bits each for k <= 256,
// else 16), so that the move accounting gathers one cache line per
// neighbour instead of two.
struct PairParts {
    const uint16_t *p16;
    const uint32_t *p32;
    __device__ void operator()(int32_t u, int32_t &w, int32_t &d) const
    {
        if (p16) {
            uint32_t x = p16[u];
            w = (int32_t)(x & 0xFFu);
            d = (int32_t)(x >> 8);
        } else {
            uint32_t x = p32[u];
            w = (int32_t)(x & 0xFFFFu);
            d = (int32_t)(x >> 16);
        }
    }
    // (where, dest) of u[t] for the live slots, (-1, -1) for the others
    template <int T>
    __device__ void gather(const int32_t (&u)[T], int32_t (&w)[T], int32_t (&d)[T]) const
    {
        uint32_t x[T];
        if (p16) {
#pragma unroll
            for (int t = 0; t < T; t++) x[t] = u[t] >= 0 ? (uint32_t)p16[u[t]] : 0u;
#pragma unroll
            for (int t = 0; t < T; t++) {
                w[t] = u[t] >= 0 ? (int32_t)(x[t] & 0xFFu) : -1;
                d[t] = u[t] >= 0 ? (int32_t)(x[t] >> 8) : -1;
            }
        } else {
#pragma unroll
            for (int t = 0; t < T; t++) x[t] = u[t] >= 0 ? p32[u[t]] : 0u;
#pragma unroll
            for (int t = 0; t < T; t++) {
                w[t] = u[t] >= 0 ? (int32_t)(x[t] & 0xFFFFu) : -1;
                d[t] = u[t] >= 0 ? (int32_t)(x[t] >> 16) : -1;
            }
        }
    }
};

// Connectivity rows of the long rows (class kClsLong, > 512 entries), kept
// current through the Jet iterations of a level, as the reference keeps its
// connectivity lists (set_gain_val_warp / update_large,
// hunyuangraph_GPU_krefine.h:4332-4432, 3728-3871), here only for the rows
// whose k-int row is far shorter than the row itself: tconn[h * k + q] =
// total edge weight from long row h to part q.  Built once per level from
// where; the move accounting then moves each moved vertex's weight between
// the rows of its long neighbours.  Integer sums, so the order of the
// updates does not matter.
struct HubConn {
    const int32_t *hidx = nullptr;   // long-row index of each vertex, -1 for the others; null: not tracked
    int32_t *tconn = nullptr;
    int32_t k = 0;
    // long-row index of u[t] (-1 for dead slots and other rows)
    template <int T>
    __device__ void index(const int32_t (&u)[T], int32_t (&h)[T]) const
    {
#pragma unroll
        for (int t = 0; t < T; t++) h[t] = (hidx && u[t] >= 0) ? hidx[u[t]] : -1;
    }
    // moves weight w of a neighbour from part pv to dv in long row h
    __device__ void move(int32_t h, int32_t w, int32_t pv, int32_t dv) const
    {
        if (h < 0) return;
        int32_t *row = tconn + (size_t)h * k;
        atomicAdd(&row[pv], -w);
        atomicAdd(&row[dv], w);
    }
    template <int T>
    __device__ void moved(const int32_t (&u)[T], const int32_t (&w)[T], int32_t pv, int32_t dv) const
    {
        if (!hidx) return;
        int32_t h[T];
        index<T>(u, h);
#pragma unroll
        for (int t = 0; t < T; t++) move(h[t], w[t], pv, dv);
    }
};

// Writes the (where, dest) word of v.
__device__ inline void store_pair(uint16_t *p16, uint32_t *p32, int32_t v, int32_t w, int32_t d)
{
    if (p16) p16[v] = (uint16_t)((uint32_t)w | ((uint32_t)d << 8));
    else p32[v] = (uint32_t)w | ((uint32_t)d << 16);
}

// The afterburner's view of a neighbour: pregain (low 32 bits), where and
// dest (16 bits each above it), one 8-byte gather per neighbour.
__device__ inline void ab_unpack(uint64_t x, int32_t &g, int32_t &w, int32_t &d)
{
    g = (int32_t)(uint32_t)x;
    w = (int32_t)((x >> 32) & 0xFFFFu);
    d = (int32_t)(x >> 48);
}

__device__ inline uint64_t ab_pack(int32_t g, int32_t w, int32_t d)
{
    return (uint64_t)(uint32_t)g | ((uint64_t)(uint32_t)w << 32) | ((uint64_t)(uint32_t)d << 48);
}

// Integer gain_bucket (hunyuangraph_GPU_krefine.h:2477-2513); identical to the
// oracle's restatement.
__device__ inline int32_t gain_bucket(int64_t gain, int64_t vwgt)
{
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    int64_t a = -gain, b = vwgt;
    int32_t t = kMidBucket;
    if (a < b) {
        while (a < b) {
            a *= 3; b *= 2; t--;
            if (t < 2) break;
        }
        if (t < 2) t = 2;
    } else {
        while (a > b) {
            a *= 2; b *= 3; t++;
            if (t > kMaxBuckets) break;
        }
        if (t > kMaxBuckets) t = kMaxBuckets - 1;
    }
    return t;
}

// ---------------------------------------------------------------- control

// The end of one Jet iteration and the start of the next in one launch
// (k_refine's loop, hunyuangraph_GPU_krefine.h:4658-4842).  End (not on the
// first call): apply the cut change, count the iteration and keep the best
// partition by the acceptance rule (:4805-4839) on the heaviest part weight.
// Start: stop after `patience` iterations without a > 0.1 % improvement,
// else pick the mode (LP when balanced, RW twice, then RS) and list the parts
// below maxdest for RW.
struct NextScratch {
    int64_t wmax[kMaxWaves];
    int32_t sh[kMaxWaves + 1];
    int32_t count, bc;
};

// Block-wide body (any block of whole waves, up to kMaxWaves): every thread
// of the block calls it.
__device__ inline void jet_next(JetCtl &C, int32_t k, const int64_t *__restrict__ pwgts, int32_t *__restrict__ under,
                                int32_t first, NextScratch &S)
{
    const int nw = blockDim.x / kWave;
    int64_t mx = INT64_MIN;
    for (int q = threadIdx.x; q < k; q += blockDim.x) mx = pwgts[q] > mx ? pwgts[q] : mx;
    mx = wave_max(mx);
    if (lane_id() == 0) S.wmax[threadIdx.x / kWave] = mx;
    __syncthreads();
    int64_t curmax = S.wmax[0];
    for (int w = 1; w < nw; w++) curmax = S.wmax[w] > curmax ? S.wmax[w] : curmax;
    if (threadIdx.x == 0) {
        if (!first) {
            C.cut += (int64_t)C.delta2 / 2;
            C.iters++;
            C.copy_best = 0;
            if (C.best_max > C.maxwgt && curmax < C.best_max) {
                C.best_max = curmax;
                C.best_cut = C.cut;
                C.copy_best = 1;
                C.count = 0;
            } else if (C.cut < C.best_cut && (curmax <= C.maxwgt || curmax <= C.best_max)) {
                if (1000 * C.cut < 999 * C.best_cut) C.count = 0;
                C.best_max = curmax;
                C.best_cut = C.cut;
                C.copy_best = 1;
            }
        }
        S.count = C.count;
        S.bc = C.balance_counter;
    }
    __syncthreads();
    bool stop = S.count >= C.patience;
    bool balanced = curmax <= C.maxwgt;
    int32_t mode = balanced ? kModeLP : (S.bc < 2 ? kModeRW : kModeRS);
    if (!stop && mode == kModeRW) {
        int32_t base = 0;
        for (int q0 = 0; q0 < k; q0 += blockDim.x) {
            int q = q0 + threadIdx.x;
            int32_t f = (q < k && pwgts[q] < C.maxdest) ? 1 : 0;
            int32_t tot;
            int32_t ex = block_exclusive_scan(f, S.sh, &tot);
            if (f) under[base + ex] = q;
            base += tot;
        }
        if (threadIdx.x == 0) C.nunder = base;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (stop) {
            C.done = 1;
        } else {
            C.count++;
            C.mode = mode;
            if (balanced) C.balance_counter = 0; else C.balance_counter++;
            C.ncand = 0;
            C.delta2 = 0;
            C.nsorted = 0;
            C.nevict = 0;
        }
    }
    __syncthreads();
}

// mirror (host memory mapped for the device, may be null) receives a copy of
// the control block for the host to read once the kernel has completed, in
// place of a copy launch.
__global__ void __launch_bounds__(kBS) k_jet_next(JetCtl *ctl, int32_t k, const int64_t *__restrict__ pwgts,
                                                  int32_t *__restrict__ under, int32_t first, JetCtl *mirror)
{
    __shared__ NextScratch S;
    __shared__ int32_t was_done;
    if (threadIdx.x == 0) was_done = ctl->done;
    __syncthreads();
    if (!was_done) jet_next(*ctl, k, pwgts, under, first, S);
    __syncthreads();
    if (mirror && threadIdx.x == 0) *mirror = *ctl;
}

// ---------------------------------------------------------------- selection

struct PartKey {
    int32_t c, q;      // connectivity, part (q = -1: none)
};

// Larger connectivity first, ties to the smaller part id, only c > 0 counts.
struct BetterPart {
    __device__ bool operator()(const PartKey &a, const PartKey &b) const
    {
        return a.c > b.c || (a.c == b.c && a.c > 0 && a.q < b.q);
    }
};

// Per-vertex decision of the iteration's mode.
//  LP : select_dest_part (:1995-2046) + filter_potential_vertex (:2114-2135)
//  RW : select_dest_parts_rw (:3184-3241), bucket of assign_move_scores (:3243-3266)
//  RS : assign_move_scores_part1 (:2515-2556) with its effective loss -conn[p]
// The three kernels below differ only in how they find, for a row, the
// connectivity pc to its own part and the best other part (c, q) under
// BetterPart; every part adjacent to the row is a candidate, so the result
// is the same whichever kernel serves a row.

struct SelArgs {
    int32_t k;
    const eid_t *xadj;
    const vid_t *adjncy;
    const wgt_t *adjwgt;
    const wgt_t *vwgt;
    const int32_t *where;
    const int64_t *pwgts;
    const uint8_t *lock;
    const int32_t *under;
    JetCtl *ctl;
    int32_t *dest;
    int32_t *pregain;
    uint32_t *key;
    unsigned long long *kb;   // kernel-timer byte counter (may be null)
    NarrowParts wq;           // where of the neighbours
    uint64_t *ab;             // LP: the afterburner's (pregain, where, dest) records
    uint16_t *pd16;           // RW/RS: (where, dest) words of the move accounting
    uint32_t *pd32;
};

// The per-vertex inputs of a decision besides its connectivity: part, vertex
// weight, lock (LP) and the weight of the own part (RW, RS).  Loaded once
// per vertex when its row is screened.
struct SelRow {
    int32_t p = 0, vw = 0, lk = 0;
    int64_t pwp = 0;
};

__device__ inline SelRow sel_row(const SelArgs &A, int32_t mode, int32_t v)
{
    SelRow r;
    r.p = A.where[v];
    if (mode == kModeLP) {
        r.lk = A.lock[v];
    } else {
        r.vw = A.vwgt[v];
        r.pwp = A.pwgts[r.p];
    }
    return r;
}

__device__ inline bool sel_eligible(int32_t mode, int64_t maxwgt, int64_t opt, const SelRow &r)
{
    if (mode == kModeLP) return true;
    if (mode == kModeRW) return r.pwp > maxwgt && 2 * (int64_t)r.vw < 3 * (r.pwp - opt);
    return r.pwp > maxwgt;
}

// Whether part q may receive v in this mode (q != p).
__device__ inline bool sel_target(const SelArgs &A, int32_t mode, int64_t maxwgt, int32_t q)
{
    if (mode == kModeRS) return false;
    return mode != kModeRW || A.pwgts[q] < maxwgt;
}

// A decision: destination, pregain (LP) or rebalancing key (RW, RS, as its
// bit pattern), and whether v is an LP candidate.
struct SelOut {
    int32_t d, val, cand;
};

__device__ inline SelOut sel_compute(const SelArgs &A, const JetCtl &C, int32_t mode, int32_t v, const SelRow &r,
                                     bool eligible, int32_t pc, PartKey best)
{
    const uint32_t nokey = (uint32_t)A.k * kMaxBuckets;
    const int32_t p = r.p;
    int32_t bc = best.c, bq = best.q;
    SelOut o{p, 0, 0};
    if (mode == kModeLP) {
        int32_t gain = kNoGain;
        if (bq >= 0 && bc > 0) {
            int32_t fn = C.filt_num;
            if (bc >= pc || (int64_t)(pc - bc) < ((int64_t)fn * pc) / 4) {
                o.d = bq;
                if (!r.lk) { gain = bc - pc; o.cand = 1; }
            }
        }
        o.val = gain;
    } else if (mode == kModeRW) {
        uint32_t kk = nokey;
        if (eligible) {
            int64_t gain = 0;
            bool mv = false;
            int32_t d = p;
            if (bq >= 0 && bc > 0) { d = bq; gain = (int64_t)bc - pc; mv = true; }
            else if (C.nunder > 0) { d = A.under[v % C.nunder]; gain = -(int64_t)pc; mv = true; }
            if (mv && d != p) {
                // gain_bucket may return kMaxBuckets itself; the reference's
                // bucket id (:3243-3266) then falls into the next part's
                // first bucket and counts against that part's cut-off.  The
                // vertex stays in its own part's last bucket here.
                int32_t bkt = gain_bucket(gain, r.vw);
                kk = (uint32_t)p * kMaxBuckets + (uint32_t)(bkt < kMaxBuckets ? bkt : kMaxBuckets - 1);
                o.d = d;
            }
        }
        o.val = (int32_t)kk;
    } else {
        uint32_t kk = nokey;
        if (eligible) {
            int32_t bkt = gain_bucket(-(int64_t)pc, r.vw);
            if (bkt < kMaxBuckets && (int64_t)r.vw < 2 * (r.pwp - C.opt)) kk = (uint32_t)p * kMaxBuckets + (uint32_t)bkt;
        }
        o.val = (int32_t)kk;
    }
    return o;
}

// Stores the decision, and the packed copies the next kernels gather: the
// afterburner record in LP (the final LP dest is settled by k_lp_finish), the
// (where, dest) word in RW/RS (changed later only for the vertices the
// eviction keeps or the cookie cutter moves).
__device__ inline void sel_store(const SelArgs &A, int32_t mode, int32_t v, int32_t p, const SelOut &o)
{
    A.dest[v] = o.d;
    if (mode == kModeLP) {
        A.pregain[v] = o.val;
        A.ab[v] = ab_pack(o.val, p, o.d);
    } else {
        A.key[v] = (uint32_t)o.val;
        store_pair(A.pd16, A.pd32, v, p, o.d);
    }
}

// Writes dest/pregain/key of v from (pc, best); returns 1 for an LP candidate.
__device__ inline int32_t sel_decide(const SelArgs &A, const JetCtl &C, int32_t mode, int32_t v, const SelRow &r,
                                     bool eligible, int32_t pc, PartKey best)
{
    SelOut o = sel_compute(A, C, mode, v, r, eligible, pc, best);
    sel_store(A, mode, v, r.p, o);
    return o.cand;
}

__device__ inline SelRow shfl_row(const SelRow &r, int src)
{
    SelRow o;
    o.p = __shfl(r.p, src, kWave);
    o.vw = __shfl(r.vw, src, kWave);
    o.lk = __shfl(r.lk, src, kWave);
    o.pwp = __shfl(r.pwp, src, kWave);
    return o;
}

// LP only needs to know whether any candidate exists: one store per block
// that has some (every thread of the block calls this).
__device__ inline void sel_count(const SelArgs &A, int32_t mode, int32_t ncand)
{
    if (mode != kModeLP) return;
    if (__syncthreads_or(ncand > 0) && threadIdx.x == 0) A.ctl->ncand = 1;
}

// Part flags of the iteration for the block: tg[q] = 1 if q may receive a
// vertex (sel_target), read once per block instead of once per entry.
__device__ inline void load_targets(const SelArgs &A, int32_t mode, int64_t maxwgt, uint8_t *tg)
{
    for (int q = threadIdx.x; q < A.k; q += blockDim.x) tg[q] = sel_target(A, mode, maxwgt, q) ? 1 : 0;
    __syncthreads();
}

// Rows of <= G entries (G = 8): each wave screens 64 rows (one per lane),
// settles the ineligible ones lane by lane, then takes the eligible ones 64/G
// at a time, G lanes and one entry per lane each.  The entry loads of all
// passes are issued before the first pass is reduced.  A lane's connectivity
// to its entry's part is summed over the group with shuffles (no LDS).
template <int G>
__global__ void __launch_bounds__(kRowBS) k_jet_select_g(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A)
{
    extern __shared__ uint8_t tg[];     // k part flags
    __shared__ int32_t slots[kRowBS];   // per wave: lanes holding eligible rows, in rank order
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    load_targets(A, mode, maxwgt, tg);
    constexpr int RP = kWave / G;   // rows per pass
    constexpr int NP = kWave / RP;  // passes that cover 64 rows
    static_assert(NP <= 8, "the passes' entries are held in registers at once");
    const int lane = lane_id(), grp = lane / G, gl = lane % G;
    int32_t *sl = slots + (threadIdx.x / kWave) * kWave;
    const int64_t wid = ((int64_t)blockIdx.x * kRowBS + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * kRowBS) / kWave;
    int32_t ncand = 0;
    uint64_t kbytes = 0;
    for (int64_t base = wid * kWave; base < cnt; base += nwaves * kWave) {
        // one row per lane: screen, settle the ineligible ones, fetch the bounds of the others
        int64_t i = base + lane;
        int32_t mv = i < cnt ? ids[i] : -1;
        SelRow mr;
        if (mv >= 0) mr = sel_row(A, mode, mv);
        bool mel = mv >= 0 && sel_eligible(mode, maxwgt, opt, mr);
        if (mv >= 0 && !mel) sel_decide(A, C, mode, mv, mr, false, 0, PartKey{0, -1});
        eid_t mb = 0, me = 0;
        if (mel) {
            mb = A.xadj[mv];
            me = A.xadj[mv + 1];
        }
        uint64_t bal = __ballot(mel);
        const int ne = __popcll(bal);
        if (mel) sl[__popcll(bal & lanemask_lt())] = lane;
        wave_lds_sync();
        const int np = (ne + RP - 1) / RP;
        int32_t src[NP], rl[NP], u[NP], w[NP], q[NP];
#pragma unroll
        for (int t = 0; t < NP; t++) {
            u[t] = -1;
            w[t] = 0;
            src[t] = -1;
            rl[t] = 0;
            if (t < np) {   // wave-uniform
                int r = t * RP + grp;
                int s = r < ne ? sl[r] : 0;
                eid_t b = __shfl(mb, s, kWave), e = __shfl(me, s, kWave);
                if (r < ne) {
                    src[t] = s;
                    rl[t] = (int32_t)(e - b);
                }
                if (r < ne && b + gl < e) {
                    u[t] = A.adjncy[b + gl];
                    w[t] = A.adjwgt[b + gl];
                }
            }
        }
        A.wq.gather(u, q);
#pragma unroll
        for (int t = 0; t < NP; t++) {
            if (t < np) {
                int s = src[t] >= 0 ? src[t] : 0;
                int32_t v = __shfl(mv, s, kWave);
                const SelRow rr = shfl_row(mr, s);
                const int32_t p = rr.p;
                int32_t conn = w[t];
#pragma unroll
                for (int o = 1; o < G; o++) {
                    int32_t qq = __shfl_xor(q[t], o, G), ww = __shfl_xor(w[t], o, G);
                    if (qq == q[t]) conn += ww;
                }
                int32_t pc = group_sum<G>(q[t] == p ? w[t] : 0);
                PartKey best{0, -1};
                if (q[t] >= 0 && q[t] != p && tg[q[t]]) best = PartKey{conn, q[t]};
                best = group_best<G>(best, BetterPart());
                if (src[t] >= 0 && gl == 0) {
                    ncand += sel_decide(A, C, mode, v, rr, true, pc, best);
                    kbytes += 16 + 12 * (uint64_t)rl[t];
                }
            }
        }
        wave_lds_sync();
    }
    sel_count(A, mode, ncand);
    kb_flush(A.kb, kbytes);
}

// Rows of 9..32 entries (class m): 16-lane groups, four rows per wave at a
// time, each group with its own k-int LDS row.  The groups of a wave run in
// lockstep, so a wave-level fence orders each group's accumulation, reads
// and clearing.  Used when k <= kQuarterK.
constexpr int kQuarterG = 16;
constexpr int32_t kQuarterK = 512;
// The long rows (> 512 entries) keep connectivity rows while k is at most
// this: a k-int row is then at most the size of the shortest long row's
// entries, and the selection reads it instead of the row.
constexpr int32_t kTrackMaxK = 512;

__global__ void __launch_bounds__(kRowBS) k_jet_select_q(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A)
{
    extern __shared__ int32_t acc[];   // k ints per group, then k part flags
    constexpr int G = kQuarterG, NG = kRowBS / G, E = (int)(kBinHalf / G);
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k, gid = threadIdx.x / G, gl = threadIdx.x % G;
    int32_t *my = acc + (size_t)gid * k;
    uint8_t *tg = (uint8_t *)(acc + (size_t)NG * k);
    for (int i = gl; i < k; i += G) my[i] = 0;
    load_targets(A, mode, maxwgt, tg);
    int32_t ncand = 0;
    uint64_t kbytes = 0;
    for (int64_t row = (int64_t)blockIdx.x * NG + gid; row < cnt; row += (int64_t)gridDim.x * NG) {
        const int32_t v = ids[row];
        const SelRow r = sel_row(A, mode, v);
        const int32_t p = r.p;
        if (!sel_eligible(mode, maxwgt, opt, r)) {
            if (gl == 0) sel_decide(A, C, mode, v, r, false, 0, PartKey{0, -1});
            continue;
        }
        const eid_t b = A.xadj[v], e = A.xadj[v + 1];
        int32_t u[E], w[E], q[E];
        load_entries<E, G>(A.adjncy, A.adjwgt, b, e, gl, u, w);
        A.wq.gather(u, q);
        if (gl == 0) kbytes += 16 + 12 * (uint64_t)(e - b);
        if (mode == kModeRS) {
            int32_t own = 0;
#pragma unroll
            for (int t = 0; t < E; t++) own += q[t] == p ? w[t] : 0;
            own = group_sum<G>(own);
            if (gl == 0) sel_decide(A, C, mode, v, r, true, own, PartKey{0, -1});
            continue;
        }
#pragma unroll
        for (int t = 0; t < E; t++)
            if (q[t] >= 0) atomicAdd(&my[q[t]], w[t]);
        wave_lds_sync();
        const int32_t pc = my[p];
        PartKey best{0, -1};
#pragma unroll
        for (int t = 0; t < E; t++) {
            if (q[t] >= 0 && q[t] != p && tg[q[t]]) {
                PartKey cand{my[q[t]], q[t]};
                if (BetterPart()(cand, best)) best = cand;
            }
        }
        best = group_best<G>(best, BetterPart());
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < E; t++)
            if (q[t] >= 0) my[q[t]] = 0;
        wave_lds_sync();
        if (gl == 0) ncand += sel_decide(A, C, mode, v, r, true, pc, best);
    }
    sel_count(A, mode, ncand);
    kb_flush(A.kb, kbytes);
}


// Rows of 9..512 entries (classes 1 and 2): a wave takes 64 rows of the list
// at a time, decides the ineligible ones lane by lane and then walks the
// eligible ones.  With Q (k <= kQuarterK) the rows of <= kQuarterRow entries
// go four at a time, one per 16-lane quarter with 8 entries per lane and its
// own k-int LDS row, so that one chain of loads and gathers serves four rows;
// the longer rows (and all rows without Q) go one at a time with their
// entries in registers (8 per lane, loaded in phases) and a k-int LDS row per
// wave.  Waves loop on their own, without block barriers.
constexpr int64_t kQuarterRow = kQuarterG * 8;

template <bool Q>
__global__ void __launch_bounds__(kRowBS) k_jet_select_w(const int32_t *__restrict__ ids, int64_t cnt, int32_t rpw,
                                                         SelArgs A)
{
    extern __shared__ int32_t acc[];   // k ints per wave (4k with Q), then k part flags
    constexpr int S = (int)(kBinMid / kWave), NQ = Q ? kWave / kQuarterG : 1;
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k, lane = lane_id();
    const int wv = threadIdx.x / kWave, nw = kRowBS / kWave;
    const int quarter = lane / kQuarterG, gl = lane % kQuarterG;
    int32_t *my = acc + (size_t)wv * NQ * k;
    uint8_t *tg = (uint8_t *)(acc + (size_t)nw * NQ * k);
    for (int i = lane; i < NQ * k; i += kWave) my[i] = 0;
    load_targets(A, mode, maxwgt, tg);
    int32_t ncand = 0;
    uint64_t kbytes = 0;
    for (int64_t base = ((int64_t)blockIdx.x * nw + wv) * rpw; base < cnt; base += (int64_t)gridDim.x * nw * rpw) {
        int64_t i = base + lane;
        int32_t mv = lane < rpw && i < cnt ? ids[i] : -1;
        SelRow mr;
        if (mv >= 0) mr = sel_row(A, mode, mv);
        bool mel = mv >= 0 && sel_eligible(mode, maxwgt, opt, mr);
        if (mv >= 0 && !mel) sel_decide(A, C, mode, mv, mr, false, 0, PartKey{0, -1});
        eid_t mb = 0, me = 0;
        if (mel) {
            mb = A.xadj[mv];
            me = A.xadj[mv + 1];
        }
        uint64_t todo = __ballot(mel);
        if constexpr (Q) {
            uint64_t four = __ballot(mel && me - mb <= kQuarterRow);
            todo &= ~four;
            while (four) {
                int ql = -1;   // the lane (row) of this lane's quarter, -1 when the rows ran out
#pragma unroll
                for (int g = 0; g < NQ; g++) {
                    const int l = four ? __ffsll((long long)four) - 1 : -1;
                    if (l >= 0) four &= four - 1;
                    if (g == quarter) ql = l;
                }
                const int src = ql >= 0 ? ql : 0;
                const int32_t v = __shfl(mv, src, kWave);
                const SelRow rr = shfl_row(mr, src);
                const int32_t p = rr.p;
                const eid_t b = __shfl(mb, src, kWave), e1 = __shfl(me, src, kWave);   // every lane shuffles
                const eid_t e = ql >= 0 ? e1 : b;
                int32_t u[8], w[8], q[8];
                load_entries<8, kQuarterG>(A.adjncy, A.adjwgt, b, e, gl, u, w);
                A.wq.gather(u, q);
                const bool lead = gl == 0 && ql >= 0;
                if (lead) kbytes += 16 + 12 * (uint64_t)(e - b);
                if (mode == kModeRS) {
                    int32_t own = 0;
#pragma unroll
                    for (int t = 0; t < 8; t++) own += q[t] == p ? w[t] : 0;
                    own = group_sum<kQuarterG>(own);
                    if (lead) sel_decide(A, C, mode, v, rr, true, own, PartKey{0, -1});
                    continue;
                }
                int32_t *mq = my + (size_t)quarter * k;
#pragma unroll
                for (int t = 0; t < 8; t++)
                    if (q[t] >= 0) atomicAdd(&mq[q[t]], w[t]);
                wave_lds_sync();
                const int32_t pc = ql >= 0 ? mq[p] : 0;
                PartKey best{0, -1};
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    if (q[t] >= 0 && q[t] != p && tg[q[t]]) {
                        PartKey cand{mq[q[t]], q[t]};
                        if (BetterPart()(cand, best)) best = cand;
                    }
                }
                best = group_best<kQuarterG>(best, BetterPart());
                wave_lds_sync();
#pragma unroll
                for (int t = 0; t < 8; t++)
                    if (q[t] >= 0) mq[q[t]] = 0;
                wave_lds_sync();
                if (lead) ncand += sel_decide(A, C, mode, v, rr, true, pc, best);
            }
        }
        while (todo) {
            int l = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            int32_t v = __shfl(mv, l, kWave);
            const SelRow rr = shfl_row(mr, l);
            const int32_t p = rr.p;
            eid_t b = __shfl(mb, l, kWave), e = __shfl(me, l, kWave);
            int32_t u[S], w[S], q[S];
            load_entries<S, kWave>(A.adjncy, A.adjwgt, b, e, lane, u, w);
            A.wq.gather(u, q);
            if (mode == kModeRS) {
                // the strong rebalancer needs only the connectivity to the own part
                int32_t own = 0;
#pragma unroll
                for (int t = 0; t < S; t++) own += q[t] == p ? w[t] : 0;
                own = wave_sum(own);
                if (lane == 0) {
                    sel_decide(A, C, mode, v, rr, true, own, PartKey{0, -1});
                    kbytes += 16 + 12 * (uint64_t)(e - b);
                }
                continue;
            }
#pragma unroll
            for (int t = 0; t < S; t++)
                if (q[t] >= 0) atomicAdd(&my[q[t]], w[t]);
            wave_lds_sync();
            int32_t pc = my[p];
            PartKey best{0, -1};
#pragma unroll
            for (int t = 0; t < S; t++) {
                if (q[t] >= 0 && q[t] != p && tg[q[t]]) {
                    PartKey cand{my[q[t]], q[t]};
                    if (BetterPart()(cand, best)) best = cand;
                }
            }
            best = group_best<kWave>(best, BetterPart());
            wave_lds_sync();
#pragma unroll
            for (int t = 0; t < S; t++)
                if (q[t] >= 0) my[q[t]] = 0;
            wave_lds_sync();
            if (lane == 0) {
                ncand += sel_decide(A, C, mode, v, rr, true, pc, best);
                kbytes += 16 + 12 * (uint64_t)(e - b);
            }
        }
    }
    sel_count(A, mode, ncand);
    kb_flush(A.kb, kbytes);
}

// Rows of > 512 entries by segments (RowSegs), a k-int LDS accumulator per
// block.  The segment's entries stay in registers (kSegLen / kSegBlk per
// thread, loaded in phases) for the accumulation, the arg-max and the
// clearing.  A single-segment row is decided here; a multi row's segments add
// their sums to its k-int row of conn, which k_jet_select_fin reads.
__global__ void __launch_bounds__(kSegBlk) k_jet_select_seg(RowSegs S, SelArgs A, int32_t *__restrict__ conn)
{
    extern __shared__ int32_t acc[];   // k ints, then k part flags
    constexpr int T = (int)(kSegLen / kSegBlk) / 2;   // a segment in two phases of T entries per thread
    constexpr eid_t kHalf = (eid_t)T * kSegBlk;
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k;
    uint8_t *tg = (uint8_t *)(acc + k);
    for (int i = threadIdx.x; i < k; i += kSegBlk) acc[i] = 0;
    load_targets(A, mode, maxwgt, tg);
    int32_t ncand = 0;
    uint64_t kbytes = 0;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t mi = S.mi[s];
        const SelRow rr = sel_row(A, mode, v);
        const int32_t p = rr.p;
        bool eligible = sel_eligible(mode, maxwgt, opt, rr);
        eid_t b = S.j0[s], e = min(b + (eid_t)kSegLen, A.xadj[v + 1]);
        if (threadIdx.x == 0 && eligible) kbytes += 8 + 12 * (uint64_t)(e - b);
        const eid_t ee = eligible ? e : b;
        const bool two = ee - b > kHalf;
        int32_t u[T], w[T], q[T];
        load_entries<T, kSegBlk>(A.adjncy, A.adjwgt, b, min(ee, b + kHalf), threadIdx.x, u, w);
        A.wq.gather(u, q);
        if (mode == kModeRS) {
            // only the connectivity to the own part: a block sum, no k-int row
            int32_t own = 0;
#pragma unroll
            for (int t = 0; t < T; t++) own += q[t] == p ? w[t] : 0;
            if (two) {
                load_entries<T, kSegBlk>(A.adjncy, A.adjwgt, b + kHalf, ee, threadIdx.x, u, w);
                A.wq.gather(u, q);
#pragma unroll
                for (int t = 0; t < T; t++) own += q[t] == p ? w[t] : 0;
            }
            own = group_sum<kSegBlk>(own);
            if (threadIdx.x == 0) {
                if (mi < 0) sel_decide(A, C, mode, v, rr, eligible, own, PartKey{0, -1});
                else if (eligible && own) atomicAdd(&conn[(size_t)mi * k + p], own);
            }
            continue;
        }
        // a two-phase segment keeps only its second half in registers: walk the parts
        bool walk_parts = two || e - b > (eid_t)k;
#pragma unroll
        for (int t = 0; t < T; t++)
            if (q[t] >= 0) atomicAdd(&acc[q[t]], w[t]);
        if (two) {
            load_entries<T, kSegBlk>(A.adjncy, A.adjwgt, b + kHalf, ee, threadIdx.x, u, w);
            A.wq.gather(u, q);
#pragma unroll
            for (int t = 0; t < T; t++)
                if (q[t] >= 0) atomicAdd(&acc[q[t]], w[t]);
        }
        __syncthreads();
        if (mi >= 0) {
            // partial sums of a multi row: move them to conn and clear
            if (eligible) {
                int32_t *row = conn + (size_t)mi * k;
                if (walk_parts) {
                    for (int r = threadIdx.x; r < k; r += kSegBlk) {
                        int32_t x = acc[r];
                        if (x) {
                            atomicAdd(&row[r], x);
                            acc[r] = 0;
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < T; t++) {
                        if (q[t] < 0) continue;
                        int32_t x = atomicExch(&acc[q[t]], 0);
                        if (x) atomicAdd(&row[q[t]], x);
                    }
                }
            }
            __syncthreads();
            continue;
        }
        int32_t pc = eligible ? acc[p] : 0;
        PartKey best{0, -1};
        if (eligible) {
            if (walk_parts) {
                for (int r = threadIdx.x; r < k; r += kSegBlk) {
                    if (r == p || !tg[r]) continue;
                    PartKey cand{acc[r], r};
                    if (BetterPart()(cand, best)) best = cand;
                }
            } else {
#pragma unroll
                for (int t = 0; t < T; t++) {
                    if (q[t] < 0 || q[t] == p || !tg[q[t]]) continue;
                    PartKey cand{acc[q[t]], q[t]};
                    if (BetterPart()(cand, best)) best = cand;
                }
            }
        }
        best = group_best<kSegBlk>(best, BetterPart());
        if (eligible) {
            if (walk_parts) {
                for (int r = threadIdx.x; r < k; r += kSegBlk) acc[r] = 0;
            } else {
#pragma unroll
                for (int t = 0; t < T; t++)
                    if (q[t] >= 0) acc[q[t]] = 0;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) ncand += sel_decide(A, C, mode, v, rr, eligible, pc, best);
    }
    sel_count(A, mode, ncand);
    kb_flush(A.kb, kbytes);
}

// Multi rows: one wave per row reads the summed k-int row (k / 64 parts per
// lane), decides, clears it.
__global__ void __launch_bounds__(kRowBS) k_jet_select_fin(RowSegs S, SelArgs A, int32_t *__restrict__ conn)
{
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k, lane = lane_id();
    const int64_t nwv = (int64_t)gridDim.x * (kRowBS / kWave);
    int32_t ncand = 0;
    for (int64_t t = ((int64_t)blockIdx.x * kRowBS + threadIdx.x) / kWave; t < S.nmulti; t += nwv) {
        int32_t v = S.mrow[t];
        const SelRow rr = sel_row(A, mode, v);
        const int32_t p = rr.p;
        bool eligible = sel_eligible(mode, maxwgt, opt, rr);
        int32_t *row = conn + (size_t)t * k;
        int32_t pc = eligible ? row[p] : 0;
        PartKey best{0, -1};
        if (eligible)
            for (int q = lane; q < k; q += kWave) {
                if (q == p || !sel_target(A, mode, maxwgt, q)) continue;
                PartKey cand{row[q], q};
                if (BetterPart()(cand, best)) best = cand;
            }
        best = group_best<kWave>(best, BetterPart());
        if (eligible)
            for (int q = lane; q < k; q += kWave) row[q] = 0;
        if (lane == 0) ncand += sel_decide(A, C, mode, v, rr, eligible, pc, best);
    }
    sel_count(A, mode, ncand);
}

// Long-row index of every vertex: position in the long rows' id list.
__global__ void k_hub_index(const int32_t *__restrict__ ids, int64_t cnt, int32_t *__restrict__ hidx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) hidx[ids[i]] = (int32_t)i;
}

// Connectivity rows of the long rows from the current partition, one block
// per segment: the segment's entries are summed per part in a k-int LDS row
// (two load phases of 8 entries per thread), then added to the long row's
// connectivity row (tconn zeroed beforehand).
__global__ void __launch_bounds__(kSegBlk) k_hub_build(RowSegs S, const eid_t *__restrict__ xadj,
                                                       const vid_t *__restrict__ adjncy,
                                                       const wgt_t *__restrict__ adjwgt, NarrowParts wq, HubConn hc)
{
    extern __shared__ int32_t acc[];   // k ints
    constexpr int T = (int)(kSegLen / kSegBlk) / 2;
    constexpr eid_t kHalf = (eid_t)T * kSegBlk;
    const int k = hc.k;
    for (int i = threadIdx.x; i < k; i += kSegBlk) acc[i] = 0;
    __syncthreads();
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        const int32_t v = S.row[s];
        const eid_t b = S.j0[s], e = min(b + (eid_t)kSegLen, xadj[v + 1]);
        for (eid_t c0 = b; c0 < e; c0 += kHalf) {
            int32_t u[T], w[T], q[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, min(e, c0 + kHalf), threadIdx.x, u, w);
            wq.gather(u, q);
#pragma unroll
            for (int t = 0; t < T; t++)
                if (q[t] >= 0) atomicAdd(&acc[q[t]], w[t]);
        }
        __syncthreads();
        int32_t *row = hc.tconn + (size_t)hc.hidx[v] * k;
        for (int r = threadIdx.x; r < k; r += kSegBlk) {
            int32_t x = acc[r];
            if (x) {
                atomicAdd(&row[r], x);
                acc[r] = 0;
            }
        }
        __syncthreads();
    }
}

// Decisions of the long rows from their connectivity rows, one wave per row
// (row i of the long rows' list is connectivity row i): the own part's
// connectivity and the best target part (k / 64 parts per lane), then the
// decision of the mode as for any row.
__global__ void __launch_bounds__(kRowBS) k_jet_select_hub(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A,
                                                           HubConn hc)
{
    extern __shared__ uint8_t tg[];   // k part flags
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k, lane = lane_id();
    load_targets(A, mode, maxwgt, tg);
    int32_t ncand = 0;
    uint64_t kbytes = 0;
    const int64_t nwv = (int64_t)gridDim.x * (kRowBS / kWave);
    for (int64_t i = ((int64_t)blockIdx.x * kRowBS + threadIdx.x) / kWave; i < cnt; i += nwv) {
        const int32_t v = ids[i];
        const SelRow rr = sel_row(A, mode, v);
        const int32_t p = rr.p;
        const bool eligible = sel_eligible(mode, maxwgt, opt, rr);
        const int32_t *row = hc.tconn + (size_t)i * k;
        int32_t pc = eligible ? row[p] : 0;
        PartKey best{0, -1};
        if (eligible && mode != kModeRS)
            for (int q = lane; q < k; q += kWave) {
                if (q == p || !tg[q]) continue;
                PartKey cand{row[q], q};
                if (BetterPart()(cand, best)) best = cand;
            }
        best = group_best<kWave>(best, BetterPart());
        if (lane == 0) {
            ncand += sel_decide(A, C, mode, v, rr, eligible, pc, best);
            if (eligible) kbytes += 4 * (uint64_t)k;
        }
    }
    sel_count(A, mode, ncand);
    kb_flush(A.kb, kbytes);
}

// Afterburner (afterburner_heuristic, :2137-2178): recompute a candidate's
// gain assuming every higher-priority neighbour (greater gain, ties to the
// smaller id) has already moved.  A neighbour's pregain, where and dest come
// from one 8-byte record (ab).
__device__ inline int64_t ab_change(int32_t v, int32_t ig, int32_t p, int32_t best, int32_t u, int32_t w, uint64_t x)
{
    int32_t ng, nw, nd;
    ab_unpack(x, ng, nw, nd);
    int64_t ch = 0;
    if (ng > ig || (ng == ig && u < v)) {
        if (nd == p) ch -= w; else if (nd == best) ch += w;
        if (nw == p) ch += w; else if (nw == best) ch -= w;
    }
    return ch;
}

template <int G>
__global__ void __launch_bounds__(kRowBS) k_afterburner(const int32_t *__restrict__ ids, int64_t cnt,
                                                        const eid_t *__restrict__ xadj,
                                                        const vid_t *__restrict__ adjncy,
                                                        const wgt_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ where,
                                                        const int32_t *__restrict__ dest, const uint64_t *__restrict__ ab,
                                                        const int32_t *__restrict__ pregain, const JetCtl *ctl,
                                                        uint8_t *__restrict__ sel, unsigned long long *kb)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    uint64_t kbytes = 0;
    if (gi < cnt) {   // whole groups take the same branches
        int32_t v = ids[gi];
        int32_t ig = pregain[v];
        if (ig == kNoGain) {
            if (lane == 0) sel[v] = 0;
        } else {
            int32_t p = where[v], best = dest[v];
            int64_t change = 0;
            for (eid_t j = xadj[v] + lane; j < xadj[v + 1]; j += G) {
                int32_t u = adjncy[j], w = adjwgt[j];
                change += ab_change(v, ig, p, best, u, w, ab[u]);
            }
            change = group_sum<G>(change);
            if (lane == 0) {
                sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
                kbytes = 24 + 16 * (uint64_t)(xadj[v + 1] - xadj[v]);
            }
        }
    }
    kb_flush(kb, kbytes);
}

// Rows of 33..512 entries: 64 rows per wave at a time; non-candidates are
// settled lane by lane, candidates of <= kQuarterRow entries four at a time
// (a 16-lane quarter and 8 entries per lane each), longer ones one row at a
// time by the wave with the row's entries loaded in phases (8 per lane).
__global__ void __launch_bounds__(kRowBS) k_afterburner_w(const int32_t *__restrict__ ids, int64_t cnt,
                                                          const eid_t *__restrict__ xadj,
                                                          const vid_t *__restrict__ adjncy,
                                                          const wgt_t *__restrict__ adjwgt,
                                                          const int32_t *__restrict__ where,
                                                          const int32_t *__restrict__ dest, const uint64_t *__restrict__ ab,
                                                          const int32_t *__restrict__ pregain, const JetCtl *ctl,
                                                          uint8_t *__restrict__ sel, int32_t rpw,
                                                          unsigned long long *kb)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    constexpr int S = (int)(kBinMid / kWave);
    const int lane = lane_id();
    uint64_t kbytes = 0;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    for (int64_t base = wid * rpw; base < cnt; base += nwaves * rpw) {
        int64_t i = base + lane;
        int32_t mv = lane < rpw && i < cnt ? ids[i] : -1;
        int32_t mg = mv >= 0 ? pregain[mv] : kNoGain;
        bool live = mv >= 0 && mg != kNoGain;
        if (mv >= 0 && !live) sel[mv] = 0;
        eid_t mb = 0, me = 0;
        int32_t mp = 0, md = 0;
        if (live) {
            mb = xadj[mv];
            me = xadj[mv + 1];
            mp = where[mv];
            md = dest[mv];
        }
        uint64_t todo = __ballot(live);
        // rows of <= kQuarterRow entries four at a time, a 16-lane quarter each
        uint64_t four = __ballot(live && me - mb <= kQuarterRow);
        todo &= ~four;
        while (four) {
            int ql = -1;
#pragma unroll
            for (int g = 0; g < kWave / kQuarterG; g++) {
                const int l = four ? __ffsll((long long)four) - 1 : -1;
                if (l >= 0) four &= four - 1;
                if (g == lane / kQuarterG) ql = l;
            }
            const int src = ql >= 0 ? ql : 0;
            const int32_t v = __shfl(mv, src, kWave), ig = __shfl(mg, src, kWave);
            const int32_t p = __shfl(mp, src, kWave), best = __shfl(md, src, kWave);
            const eid_t b = __shfl(mb, src, kWave), e1 = __shfl(me, src, kWave);   // every lane shuffles
            const eid_t e = ql >= 0 ? e1 : b;
            int32_t u[8], w[8];
            uint64_t x[8];
            load_entries<8, kQuarterG>(adjncy, adjwgt, b, e, lane % kQuarterG, u, w);
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = u[t] >= 0 ? ab[u[t]] : 0ull;
            int64_t change = 0;
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (u[t] >= 0) change += ab_change(v, ig, p, best, u[t], w[t], x[t]);
            change = group_sum<kQuarterG>(change);
            if (lane % kQuarterG == 0 && ql >= 0) {
                sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
                kbytes += 24 + 16 * (uint64_t)(e - b);
            }
        }
        while (todo) {
            int l = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            int32_t v = __shfl(mv, l, kWave), ig = __shfl(mg, l, kWave);
            int32_t p = __shfl(mp, l, kWave), best = __shfl(md, l, kWave);
            eid_t b = __shfl(mb, l, kWave), e = __shfl(me, l, kWave);
            int32_t u[S], w[S];
            uint64_t x[S];
            load_entries<S, kWave>(adjncy, adjwgt, b, e, lane, u, w);
#pragma unroll
            for (int t = 0; t < S; t++) x[t] = u[t] >= 0 ? ab[u[t]] : 0ull;
            int64_t change = 0;
#pragma unroll
            for (int t = 0; t < S; t++)
                if (u[t] >= 0) change += ab_change(v, ig, p, best, u[t], w[t], x[t]);
            change = wave_sum(change);
            if (lane == 0) {
                sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
                kbytes += 24 + 16 * (uint64_t)(e - b);
            }
        }
    }
    kb_flush(kb, kbytes);
}

// Long rows by segments: a single-segment row decides here, a multi row adds
// its partial change to macc for k_afterburner_fin.  Entries in phases of 8
// per thread.
__global__ void __launch_bounds__(kSegBlk) k_afterburner_seg(RowSegs S, const eid_t *__restrict__ xadj,
                                                            const vid_t *__restrict__ adjncy,
                                                            const wgt_t *__restrict__ adjwgt,
                                                            const int32_t *__restrict__ where,
                                                            const int32_t *__restrict__ dest,
                                                            const uint64_t *__restrict__ ab,
                                                            const int32_t *__restrict__ pregain, const JetCtl *ctl,
                                                            uint8_t *__restrict__ sel,
                                                            unsigned long long *__restrict__ macc,
                                                            unsigned long long *kb)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    constexpr int T = 8;
    uint64_t kbytes = 0;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t mi = S.mi[s];
        int32_t ig = pregain[v];
        if (ig == kNoGain) {
            if (threadIdx.x == 0 && mi < 0) sel[v] = 0;
            continue;
        }
        int32_t p = where[v], best = dest[v];
        int64_t change = 0;
        eid_t j0 = S.j0[s], j1 = min(j0 + (eid_t)kSegLen, xadj[v + 1]);
        for (eid_t c0 = j0; c0 < j1; c0 += (eid_t)T * kSegBlk) {
            int32_t u[T], w[T];
            uint64_t x[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, j1, threadIdx.x, u, w);
#pragma unroll
            for (int t = 0; t < T; t++) x[t] = u[t] >= 0 ? ab[u[t]] : 0ull;
#pragma unroll
            for (int t = 0; t < T; t++)
                if (u[t] >= 0) change += ab_change(v, ig, p, best, u[t], w[t], x[t]);
        }
        change = group_sum<kSegBlk>(change);
        if (threadIdx.x == 0) {
            if (mi < 0) sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
            else if (change) atomicAdd(&macc[mi], (unsigned long long)change);
            kbytes += 16 + 16 * (uint64_t)(j1 - j0);
        }
    }
    kb_flush(kb, kbytes);
}

__global__ void k_afterburner_fin(RowSegs S, const int32_t *__restrict__ pregain, const JetCtl *ctl,
                                  uint8_t *__restrict__ sel, int64_t *__restrict__ macc)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    int32_t t = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= S.nmulti) return;
    int32_t v = S.mrow[t];
    int32_t ig = pregain[v];
    sel[v] = (ig != kNoGain && (int64_t)ig + macc[t] >= 0) ? 1 : 0;
    macc[t] = 0;
}

// filter_beneficial_moves + set_lock (:2180-2221): with candidates present,
// locks become exactly the selected set; without any, nothing changes.
__global__ void k_lp_finish(int32_t n, const JetCtl *ctl, const uint8_t *__restrict__ sel,
                            const int32_t *__restrict__ where, int32_t *__restrict__ dest,
                            uint8_t *__restrict__ lock, uint16_t *__restrict__ pd16, uint32_t *__restrict__ pd32)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    const int32_t w = where[v];
    int32_t d = w;
    if (ctl->ncand != 0) {
        uint8_t s = sel[v];
        lock[v] = s;
        if (s) d = dest[v];
        else dest[v] = w;
    } else {
        dest[v] = w;
    }
    store_pair(pd16, pd32, v, w, d);
}

// ---------------------------------------------------------------- rebalancers

// After the sort by (key, id): each candidate's weight (scanned next into
// the prefix weights of its part's candidates) and the first position of
// each part in the sorted list.
__global__ void k_rb_weights(int32_t n, int32_t k, const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                             const wgt_t *__restrict__ vwgt, int64_t *__restrict__ w, int32_t *__restrict__ pstart)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const uint32_t kk = key[i];
    const bool cand = kk < (uint32_t)k * kMaxBuckets;
    w[i] = cand ? vwgt[val[i]] : 0;
    if (cand) {
        uint32_t p = kk / kMaxBuckets;
        if (i == 0 || key[i - 1] / kMaxBuckets != p) pstart[p] = i;
    }
}

// Evict a candidate iff the candidates ahead of it in its part weigh less
// than the part's excess (filter_scores_below_cutoff :3344-3366 for RW,
// find_score_cutoffs/filter_below_cutoffs :2641-2691 for RS).  RW keeps the
// destination chosen in selection; RS only flags the evicted ones.
__global__ void k_rb_evict(int32_t n, int32_t k, const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                           const int64_t *__restrict__ wsc, const wgt_t *__restrict__ vwgt,
                           const int32_t *__restrict__ pstart, const int64_t *__restrict__ pwgts, const JetCtl *ctl,
                           const int32_t *__restrict__ where, int32_t *__restrict__ dest, uint8_t *__restrict__ ev,
                           uint16_t *__restrict__ pd16, uint32_t *__restrict__ pd32)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    uint32_t kk = key[i];
    int32_t v = (int32_t)val[i];
    if (kk >= (uint32_t)k * kMaxBuckets) {
        ev[i] = 0;
        return;
    }
    int32_t p = (int32_t)(kk / kMaxBuckets);
    int32_t s = pstart[p];
    int64_t before = wsc[i] - vwgt[v] - (s > 0 ? wsc[s - 1] : 0);
    bool evict = before < pwgts[p] - ctl->maxwgt;
    ev[i] = evict ? 1 : 0;
    if (ctl->mode != kModeRW || !evict) {
        int32_t w = where[v];
        dest[v] = w;
        store_pair(pd16, pd32, v, w, w);
    }
}

// Cookie cutter (:2710-2745): slice the evicted list (in sorted order) into
// consecutive runs, part by part, each filling maxwgt - pwgts.  The parts
// stay sequential (a slice starts where the last one ended); one wave finds
// each slice end.  Vertex weights are >= 1, so bs is strictly increasing: a
// slice of capacity 0 (or an exhausted list) ends where it starts, and the
// first prefix reaching bs[start] + cap lies within cap entries of start.
// Over (start, limit], limit = min(ne, start + cap), the serial binary search
// ends on hi = the first index whose prefix reaches find (limit if none) and
// lo = hi - 1; the wave finds that hi by 64-way narrowing (each round one
// load per lane and a ballot), then keeps the nearer of lo and hi.
__global__ void __launch_bounds__(kWave) k_rs_cookie(int32_t k, const int64_t *__restrict__ bs,
                                                     const int64_t *__restrict__ ne_ptr,
                                                     const int64_t *__restrict__ pwgts, const JetCtl *ctl,
                                                     int32_t *__restrict__ es, int32_t *__restrict__ ee)
{
    extern __shared__ int64_t room[];   // maxwgt - pwgts[p], clamped at 0
    if (blockIdx.x != 0) return;
    const int lane = lane_id();
    const int32_t ne = (int32_t)*ne_ptr;
    const int64_t maxwgt = ctl->maxwgt;
    for (int32_t p = lane; p < k; p += kWave) room[p] = maxwgt > pwgts[p] ? maxwgt - pwgts[p] : 0;
    __syncthreads();
    int32_t start = 0;
    int64_t bstart = bs[0];   // bs[start], carried from slice to slice
    for (int32_t p = 0; p < k; p++) {
        if (start >= ne) {   // the list is used up: every later slice is empty
            for (int32_t q = p + lane; q < k; q += kWave) es[q] = ee[q] = ne;
            break;
        }
        if (lane == 0) es[p] = start;
        const int64_t cap = room[p];
        int32_t end = start;
        if (cap > 0 && start < ne) {
            const int64_t find = bstart + cap;
            const int32_t limit = (int32_t)min((int64_t)ne, (int64_t)start + cap);
            int32_t L = start + 1, H = limit;   // the answer lies in [L, H]; H stands for "none below H"
            while (L < H) {
                const int32_t span = H - L;
                const int32_t step = span <= kWave ? 1 : (span + kWave - 1) / kWave;
                const int32_t pos = L + lane * step;
                const bool ge = pos < H && bs[pos] >= find;
                const uint64_t m = __ballot(ge);
                if (step == 1) {   // every index of [L, H) was tested
                    if (m) H = L + (__ffsll((long long)m) - 1);
                    break;
                }
                if (m == 0) {   // every sample below find: past the last one that lies below H
                    const int32_t nsamp = (span + step - 1) / step;
                    L = L + (nsamp - 1) * step + 1;
                    continue;
                }
                const int f = __ffsll((long long)m) - 1;
                if (f == 0) {
                    H = L;
                    break;
                }
                H = L + f * step;
                L = L + (f - 1) * step + 1;
            }
            const int32_t hi = H, lo = H - 1;
            const int64_t bh = bs[hi], bl = bs[lo];
            int64_t dh = bh - find, dl = bl - find;
            dh = dh < 0 ? -dh : dh;
            dl = dl < 0 ? -dl : dl;
            end = dh < dl ? hi : lo;
            bstart = dh < dl ? bh : bl;
        }
        if (lane == 0) ee[p] = end;
        start = end;
    }
}

// select_dest_parts_rs (:2747-2766) over the compacted evicted list.
__global__ void k_rs_assign(int32_t k, const int64_t *__restrict__ ne_ptr, const int32_t *__restrict__ elist,
                            const uint32_t *__restrict__ val, const int32_t *__restrict__ es,
                            const int32_t *__restrict__ ee, const int32_t *__restrict__ where,
                            int32_t *__restrict__ dest, uint16_t *__restrict__ pd16, uint32_t *__restrict__ pd32)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= *ne_ptr) return;
    int32_t lo = 0, hi = k - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (es[mid] <= i) lo = mid; else hi = mid - 1;
    }
    int32_t v = (int32_t)val[elist[i]];
    int32_t w = where[v];
    if (i < ee[lo] && lo != w) {
        dest[v] = lo;
        store_pair(pd16, pd32, v, w, lo);
    }
}

__global__ void k_gather_w(const int64_t *__restrict__ ne_ptr, const int32_t *__restrict__ elist,
                           const uint32_t *__restrict__ val, const wgt_t *__restrict__ vwgt, int64_t *__restrict__ bs)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    int64_t ne = *ne_ptr;
    if (i == 0) bs[0] = 0;
    if (i < ne) bs[i + 1] = vwgt[val[elist[i]]];
}

// ---------------------------------------------------------------- moves

// Move accounting (perform_moves, :4034-4155): the exact cut change of
// applying dest[] (edges between two moved vertices are seen from both ends,
// so each side counts once and one-sided edges twice) and the part-weight
// updates.  Both are summed in LDS and flushed with one atomic per block and
// touched part.  Rows of <= G entries (G = 8 or 32): each wave screens 64
// rows, then walks the moved ones 64/G at a time (G lanes, one entry per lane
// each), the entry loads of all passes issued before the first is used.
__device__ inline int64_t move_d2(int32_t pv, int32_t dv, int32_t pu, int32_t du, int32_t w)
{
    bool um = pu != du;
    int32_t before = pv != pu, after = dv != (um ? du : pu);
    return (int64_t)(after - before) * w * (um ? 1 : 2);
}

template <int G>
__global__ void __launch_bounds__(kRowBS) k_move_delta_g(const int32_t *__restrict__ ids, int64_t cnt, int32_t k,
                                                         const eid_t *__restrict__ xadj,
                                                         const vid_t *__restrict__ adjncy,
                                                         const wgt_t *__restrict__ adjwgt,
                                                         const wgt_t *__restrict__ vwgt,
                                                         const int32_t *__restrict__ where,
                                                         const int32_t *__restrict__ dest, PairParts pq, HubConn hc,
                                                         JetCtl *ctl, int64_t *__restrict__ pwgts,
                                                         unsigned long long *kb)
{
    extern __shared__ unsigned long long pdelta[];
    __shared__ unsigned long long bdelta;
    __shared__ int32_t slots[kRowBS];
    if (ctl->done) return;
    for (int i = threadIdx.x; i < k; i += kRowBS) pdelta[i] = 0;
    if (threadIdx.x == 0) bdelta = 0;
    __syncthreads();
    constexpr int RP = kWave / G;                       // rows per pass
    constexpr int PB = kWave / RP < 8 ? kWave / RP : 8;   // passes whose loads go out together
    const int lane = lane_id(), grp = lane / G, gl = lane % G;
    int32_t *sl = slots + (threadIdx.x / kWave) * kWave;
    const int64_t wid = ((int64_t)blockIdx.x * kRowBS + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * kRowBS) / kWave;
    uint64_t kbytes = 0;
    for (int64_t base = wid * kWave; base < cnt; base += nwaves * kWave) {
        int64_t i = base + lane;
        int32_t mv = i < cnt ? ids[i] : -1;
        int32_t mpv = mv >= 0 ? where[mv] : 0, mdv = mv >= 0 ? dest[mv] : 0;
        bool moved = mv >= 0 && mpv != mdv;
        eid_t mb = 0, me = 0;
        if (moved) {
            mb = xadj[mv];
            me = xadj[mv + 1];
            int32_t wv = vwgt[mv];
            atomicAdd(&pdelta[mpv], (unsigned long long)(-(int64_t)wv));
            atomicAdd(&pdelta[mdv], (unsigned long long)(int64_t)wv);
            kbytes += 20 + 16 * (uint64_t)(me - mb);
        }
        uint64_t bal = __ballot(moved);
        const int nm = __popcll(bal);
        if (moved) sl[__popcll(bal & lanemask_lt())] = lane;
        wave_lds_sync();
        const int np = (nm + RP - 1) / RP;
        int64_t d2 = 0;
        for (int t0 = 0; t0 < np; t0 += PB) {
            int32_t src[PB], u[PB], w[PB], pu[PB], du[PB];
#pragma unroll
            for (int t = 0; t < PB; t++) {
                u[t] = -1;
                w[t] = 0;
                src[t] = 0;
                if (t0 + t < np) {   // wave-uniform
                    int r = (t0 + t) * RP + grp;
                    int s = r < nm ? sl[r] : 0;
                    eid_t b = __shfl(mb, s, kWave), e = __shfl(me, s, kWave);
                    src[t] = s;
                    if (r < nm && b + gl < e) {
                        u[t] = adjncy[b + gl];
                        w[t] = adjwgt[b + gl];
                    }
                }
            }
            pq.gather(u, pu, du);
            int32_t hh[PB];
            hc.index<PB>(u, hh);
#pragma unroll
            for (int t = 0; t < PB; t++) {
                if (t0 + t < np) {
                    int32_t pv = __shfl(mpv, src[t], kWave), dv = __shfl(mdv, src[t], kWave);
                    if (u[t] >= 0) {
                        d2 += move_d2(pv, dv, pu[t], du[t], w[t]);
                        hc.move(hh[t], w[t], pv, dv);
                    }
                }
            }
        }
        d2 = wave_sum(d2);
        if (lane == 0 && d2) atomicAdd(&bdelta, (unsigned long long)d2);
        wave_lds_sync();
    }
    kb_flush(kb, kbytes);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += kRowBS)
        if (pdelta[i]) atomicAdd((unsigned long long *)&pwgts[i], pdelta[i]);
    if (threadIdx.x == 0 && bdelta) atomicAdd(&ctl->delta2, bdelta);
}

// Rows of 33..512 entries: 64 rows per wave at a time, only the moved ones
// walked: rows of <= kQuarterRow entries four at a time (a 16-lane quarter
// and 8 entries per lane each), longer ones one at a time with their entries
// loaded in phases (8 per lane).
__global__ void __launch_bounds__(kRowBS) k_move_delta_w(const int32_t *__restrict__ ids, int64_t cnt, int32_t k,
                                                         const eid_t *__restrict__ xadj,
                                                         const vid_t *__restrict__ adjncy,
                                                         const wgt_t *__restrict__ adjwgt,
                                                         const wgt_t *__restrict__ vwgt,
                                                         const int32_t *__restrict__ where,
                                                         const int32_t *__restrict__ dest, PairParts pq, HubConn hc,
                                                         JetCtl *ctl, int64_t *__restrict__ pwgts, int32_t rpw,
                                                         unsigned long long *kb)
{
    extern __shared__ unsigned long long pdelta[];
    __shared__ unsigned long long bdelta;
    if (ctl->done) return;
    constexpr int S = (int)(kBinMid / kWave);
    for (int i = threadIdx.x; i < k; i += kRowBS) pdelta[i] = 0;
    if (threadIdx.x == 0) bdelta = 0;
    __syncthreads();
    uint64_t kbytes = 0;
    const int lane = lane_id();
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    int64_t d2 = 0;
    for (int64_t base = wid * rpw; base < cnt; base += nwaves * rpw) {
        int64_t i = base + lane;
        int32_t mv = lane < rpw && i < cnt ? ids[i] : -1;
        int32_t mpv = mv >= 0 ? where[mv] : 0, mdv = mv >= 0 ? dest[mv] : 0;
        bool moved = mv >= 0 && mpv != mdv;
        eid_t mb = 0, me = 0;
        if (moved) {
            mb = xadj[mv];
            me = xadj[mv + 1];
            int32_t wv = vwgt[mv];
            atomicAdd(&pdelta[mpv], (unsigned long long)(-(int64_t)wv));
            atomicAdd(&pdelta[mdv], (unsigned long long)(int64_t)wv);
            kbytes += 20 + 16 * (uint64_t)(me - mb);
        }
        uint64_t todo = __ballot(moved);
        // rows of <= kQuarterRow entries four at a time, a 16-lane quarter each
        uint64_t four = __ballot(moved && me - mb <= kQuarterRow);
        todo &= ~four;
        while (four) {
            int ql = -1;
#pragma unroll
            for (int g = 0; g < kWave / kQuarterG; g++) {
                const int l = four ? __ffsll((long long)four) - 1 : -1;
                if (l >= 0) four &= four - 1;
                if (g == lane / kQuarterG) ql = l;
            }
            const int src = ql >= 0 ? ql : 0;
            const int32_t pv = __shfl(mpv, src, kWave), dv = __shfl(mdv, src, kWave);
            const eid_t b = __shfl(mb, src, kWave), e1 = __shfl(me, src, kWave);   // every lane shuffles
            const eid_t e = ql >= 0 ? e1 : b;
            int32_t u[8], w[8], pu[8], du[8];
            load_entries<8, kQuarterG>(adjncy, adjwgt, b, e, lane % kQuarterG, u, w);
            pq.gather(u, pu, du);
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (u[t] >= 0) d2 += move_d2(pv, dv, pu[t], du[t], w[t]);
            hc.moved<8>(u, w, pv, dv);
        }
        while (todo) {
            int l = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            int32_t pv = __shfl(mpv, l, kWave), dv = __shfl(mdv, l, kWave);
            eid_t b = __shfl(mb, l, kWave), e = __shfl(me, l, kWave);
            int32_t u[S], w[S], pu[S], du[S];
            load_entries<S, kWave>(adjncy, adjwgt, b, e, lane, u, w);
            pq.gather(u, pu, du);
#pragma unroll
            for (int t = 0; t < S; t++)
                if (u[t] >= 0) d2 += move_d2(pv, dv, pu[t], du[t], w[t]);
            hc.moved<S>(u, w, pv, dv);
        }
    }
    d2 = wave_sum(d2);
    if (lane_id() == 0 && d2) atomicAdd(&bdelta, (unsigned long long)d2);
    kb_flush(kb, kbytes);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += kRowBS)
        if (pdelta[i]) atomicAdd((unsigned long long *)&pwgts[i], pdelta[i]);
    if (threadIdx.x == 0 && bdelta) atomicAdd(&ctl->delta2, bdelta);
}

// Long rows by segments; the part weights move with the row's first segment.
// Entries in phases of 8 per thread.
__global__ void __launch_bounds__(kSegBlk) k_move_delta_seg(RowSegs S, int32_t k, const eid_t *__restrict__ xadj,
                                                           const vid_t *__restrict__ adjncy,
                                                           const wgt_t *__restrict__ adjwgt,
                                                           const wgt_t *__restrict__ vwgt,
                                                           const int32_t *__restrict__ where,
                                                           const int32_t *__restrict__ dest, PairParts pq, HubConn hc,
                                                           JetCtl *ctl, int64_t *__restrict__ pwgts,
                                                           unsigned long long *kb)
{
    extern __shared__ unsigned long long pdelta[];
    __shared__ unsigned long long bdelta;
    if (ctl->done) return;
    constexpr int T = 8;
    for (int i = threadIdx.x; i < k; i += kSegBlk) pdelta[i] = 0;
    if (threadIdx.x == 0) bdelta = 0;
    __syncthreads();
    uint64_t kbytes = 0;
    int64_t d2 = 0;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t pv = where[v], dv = dest[v];
        if (pv == dv) continue;
        eid_t j0 = S.j0[s], j1 = min(j0 + (eid_t)kSegLen, xadj[v + 1]);
        for (eid_t c0 = j0; c0 < j1; c0 += (eid_t)T * kSegBlk) {
            int32_t u[T], w[T], pu[T], du[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, j1, threadIdx.x, u, w);
            pq.gather(u, pu, du);
#pragma unroll
            for (int t = 0; t < T; t++)
                if (u[t] >= 0) d2 += move_d2(pv, dv, pu[t], du[t], w[t]);
            hc.moved<T>(u, w, pv, dv);
        }
        if (threadIdx.x == 0) {
            kbytes += 16 + 16 * (uint64_t)(j1 - j0);
            if (j0 == xadj[v]) {
                atomicAdd(&pdelta[pv], (unsigned long long)(-(int64_t)vwgt[v]));
                atomicAdd(&pdelta[dv], (unsigned long long)(int64_t)vwgt[v]);
            }
        }
    }
    d2 = group_sum<kSegBlk>(d2);
    if (threadIdx.x == 0 && d2) atomicAdd(&bdelta, (unsigned long long)d2);
    kb_flush(kb, kbytes);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += kSegBlk)
        if (pdelta[i]) atomicAdd((unsigned long long *)&pwgts[i], pdelta[i]);
    if (threadIdx.x == 0 && bdelta) atomicAdd(&ctl->delta2, bdelta);
}

// Applies dest and refreshes the narrow copy of where.  First, when the
// previous iteration's k_jet_next kept its partition (copy_best, still set
// until this iteration's k_jet_next), copies that partition to best: the
// copy costs no launch of its own (after the last iteration k_copy_best
// makes it).
__global__ void k_move_commit(int32_t n, const JetCtl *ctl, const int32_t *__restrict__ dest, int32_t *__restrict__ where,
                              uint8_t *__restrict__ w8, uint16_t *__restrict__ w16, int32_t *__restrict__ best)
{
    if (ctl->done) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    if (ctl->copy_best) best[v] = where[v];
    int32_t d = dest[v];
    where[v] = d;
    if (w8) w8[v] = (uint8_t)d;
    else w16[v] = (uint16_t)d;
}

// Keeps the iteration's partition when k_jet_next found it better (also when
// that kernel ended the level: the last iteration's result still counts).
__global__ void k_copy_best(int32_t n, const JetCtl *ctl, const int32_t *__restrict__ where, int32_t *__restrict__ best)
{
    if (!ctl->copy_best) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) best[v] = where[v];
}

__global__ void __launch_bounds__(kBS) k_jet_init(JetCtl *ctl, int32_t k, const int64_t *__restrict__ pwgts,
                                                  const int64_t *__restrict__ d_cut, int64_t maxwgt, int64_t opt,
                                                  int32_t patience, int32_t filt_num)
{
    __shared__ int64_t smax[kBS];
    int64_t mx = INT64_MIN;
    for (int q = threadIdx.x; q < k; q += kBS) mx = pwgts[q] > mx ? pwgts[q] : mx;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && smax[threadIdx.x + s] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    JetCtl &C = *ctl;
    C.cut = C.best_cut = *d_cut;
    C.best_max = smax[0];
    C.maxwgt = maxwgt;
    C.opt = opt;
    int64_t md = (99 * maxwgt) / 100;
    C.maxdest = md < maxwgt - 100 ? maxwgt - 100 : md;
    C.delta2 = 0;
    C.count = 0;
    C.balance_counter = 0;
    C.mode = 0;
    C.done = 0;
    C.iters = 0;
    C.patience = patience;
    C.ncand = 0;
    C.nunder = 0;
    C.copy_best = 0;
    C.filt_num = filt_num;
}

__global__ void k_jet_fini(const JetCtl *ctl, int64_t *d_cut) { *d_cut = ctl->best_cut; }

// ---------------------------------------------------------------- utilities

template <int G>
__global__ void __launch_bounds__(kRowBS) k_cut(const int32_t *__restrict__ ids, int64_t cnt,
                                                const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                const wgt_t *__restrict__ adjwgt, const int32_t *__restrict__ where,
                                                unsigned long long *out)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t v = ids[gi];
    int32_t p = where[v];
    int64_t s = 0;
    for (eid_t j = xadj[v] + lane; j < xadj[v + 1]; j += G) s += where[adjncy[j]] != p ? adjwgt[j] : 0;
    s = group_sum<G>(s);
    if (lane == 0 && s) atomicAdd(out, (unsigned long long)s);
}

__global__ void __launch_bounds__(kSegBlk) k_cut_seg(RowSegs S, const eid_t *__restrict__ xadj,
                                                    const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ where, unsigned long long *out)
{
    int64_t tot = 0;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t p = where[v];
        eid_t j1 = min(S.j0[s] + (eid_t)kSegLen, xadj[v + 1]);
        for (eid_t j = S.j0[s] + threadIdx.x; j < j1; j += kSegBlk) tot += where[adjncy[j]] != p ? adjwgt[j] : 0;
    }
    tot = group_sum<kSegBlk>(tot);
    if (threadIdx.x == 0 && tot) atomicAdd(out, (unsigned long long)tot);
}

// Classes 0-2 (the long rows go by segments).
template <class F>
void for_small_bins(const RowBins &b, F f)
{
    if (b.count(0)) f(std::integral_constant<int, 8>{}, b.list(0), b.count(0));
    if (b.count(1)) f(std::integral_constant<int, 32>{}, b.list(1), b.count(1));
    if (b.count(2)) f(std::integral_constant<int, 64>{}, b.list(2), b.count(2));
}

inline unsigned seg_grid(const RowSegs &S) { return (unsigned)std::min<int64_t>(S.nseg, tuning().seg_grid); }

__global__ void k_half(int64_t *d) { *d /= 2; }

// Persistent blocks sum their vertices' weights per part in LDS and flush one
// atomic per block and touched part.
__global__ void __launch_bounds__(kBS) k_pwgts(int32_t n, int32_t k, const wgt_t *__restrict__ vwgt,
                                               const int32_t *__restrict__ where, unsigned long long *pw)
{
    extern __shared__ unsigned long long hist[];
    for (int i = threadIdx.x; i < k; i += kBS) hist[i] = 0;
    __syncthreads();
    for (int64_t v = (int64_t)blockIdx.x * kBS + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBS)
        atomicAdd(&hist[where[v]], (unsigned long long)(int64_t)vwgt[v]);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += kBS)
        if (hist[i]) atomicAdd(&pw[i], hist[i]);
}

__global__ void k_project(int32_t n, const int32_t *__restrict__ cmap, const int32_t *__restrict__ wc,
                          int32_t *__restrict__ wf)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) wf[v] = wc[cmap[v]];
}

// Rows per wave chunk of the wave-per-row kernels.  A wave screens a chunk
// of rows and walks the live ones one after another, so on small levels a
// 64-row chunk is a long serial chain on few waves; the chunk shrinks until
// about Tuning::fill_waves waves share the rows (64 rows once the level is large).
inline int32_t rows_per_wave(int64_t rows)
{
    const int64_t fill = tuning().fill_waves;
    int64_t r = (rows + fill - 1) / fill;
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(r, kWave));
}

// Blocks of kRowBS threads that give each chunk of rpw rows its own wave, at most cap.
inline unsigned wave_blocks(int64_t rows, int32_t rpw, int64_t cap)
{
    int64_t waves = (rows + rpw - 1) / rpw;
    int64_t b = (waves + kRowBS / kWave - 1) / (kRowBS / kWave);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

// LDS of k_jet_select_w: a k-int row per wave, or per quarter wave with Q.
inline size_t sel_w_lds(int32_t k, bool q)
{
    return sizeof(int32_t) * (kRowBS / kWave) * (q ? kWave / kQuarterG : 1) * (size_t)k;
}

void launch_select(Ctx &c, const DevGraph &g, int32_t k, const int32_t *where, const int64_t *pwgts,
                   const uint8_t *lock, const int32_t *under, JetCtl *ctl, int32_t *dest, int32_t *pregain,
                   uint32_t *key, int32_t *conn, NarrowParts wq, uint64_t *ab, uint16_t *pd16, uint32_t *pd32,
                   const HubConn &hc)
{
    const RowBins &b = g.bins;
    SelArgs A{k,   g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, pwgts, lock, under, ctl,
              dest, pregain, key, nullptr, wq, ab, pd16, pd32};
    const size_t tg_bytes = ((size_t)k + 3) & ~(size_t)3;   // part flags of the selection kernels
    // fixed bytes: ids, where, pwgts[p], vwgt, the decision (24 per row); eligible rows add xadj and
    // 12 B per entry (adjncy, adjwgt, where gather) in the kernel
    if (b.few()) {
        if (int64_t nw = b.short_rows()) {
            KScope ks(c, "k_jet_select_few", 24.0 * (double)nw, true);
            A.kb = ks.dev();
            int32_t rpw = rows_per_wave(nw);
            if (k <= kQuarterK)
                k_jet_select_w<true><<<wave_blocks(nw, rpw, 4096), kRowBS, sel_w_lds(k, true) + tg_bytes, c.stream>>>(
                    b.list(0), nw, rpw, A);
            else
                k_jet_select_w<false><<<wave_blocks(nw, rpw, 4096), kRowBS, sel_w_lds(k, false) + tg_bytes,
                                        c.stream>>>(b.list(0), nw, rpw, A);
            HGP_LAUNCH_CHECK();
        }
    }
    if (b.count(0) && !b.few()) {
        KScope ks(c, "k_jet_select_s", 24.0 * (double)b.count(0), true);
        A.kb = ks.dev();
        unsigned blocks = (unsigned)std::min<int64_t>((b.count(0) + kRowBS - 1) / kRowBS, 16384);
        k_jet_select_g<8><<<blocks, kRowBS, tg_bytes, c.stream>>>(b.list(0), b.count(0), A);
        HGP_LAUNCH_CHECK();
    }
    // class m: 16-lane groups while k is small enough for their LDS rows;
    // otherwise classes m and w (consecutive in the id list) share one
    // wave-per-row launch
    const bool quarter = !b.few() && k <= kQuarterK && b.count(1) > 0;
    if (quarter) {
        KScope ks(c, "k_jet_select_q", 24.0 * (double)b.count(1), true);
        A.kb = ks.dev();
        constexpr int NG = kRowBS / kQuarterG;
        unsigned blocks = (unsigned)std::min<int64_t>((b.count(1) + NG - 1) / NG, 32768);
        k_jet_select_q<<<blocks, kRowBS, sizeof(int32_t) * NG * (size_t)k + tg_bytes, c.stream>>>(b.list(1),
                                                                                               b.count(1), A);
        HGP_LAUNCH_CHECK();
    }
    if (int64_t nw = b.few() ? 0 : (quarter ? 0 : b.count(1)) + b.count(2)) {
        KScope ks(c, "k_jet_select_w", 24.0 * (double)nw, true);
        A.kb = ks.dev();
        int32_t rpw = rows_per_wave(nw);
        const int32_t *list = quarter ? b.list(2) : b.list(1);
        if (k <= kQuarterK)
            k_jet_select_w<true><<<wave_blocks(nw, rpw, 4096), kRowBS, sel_w_lds(k, true) + tg_bytes, c.stream>>>(
                list, nw, rpw, A);
        else
            k_jet_select_w<false><<<wave_blocks(nw, rpw, 4096), kRowBS, sel_w_lds(k, false) + tg_bytes, c.stream>>>(
                list, nw, rpw, A);
        HGP_LAUNCH_CHECK();
    }
    if (hc.hidx) {
        // the long rows decide from their connectivity rows
        const int64_t nh = b.count(kClsLong);
        KScope ks(c, "k_jet_select_hub", 24.0 * (double)nh, true);
        A.kb = ks.dev();
        const int32_t rows_per_block = kRowBS / kWave;
        k_jet_select_hub<<<(unsigned)std::min<int64_t>((nh + rows_per_block - 1) / rows_per_block, 8192), kRowBS,
                           tg_bytes, c.stream>>>(b.list(kClsLong), nh, A, hc);
        HGP_LAUNCH_CHECK();
    } else if (g.segs.nseg) {
        {
            KScope ks(c, "k_jet_select_seg", 24.0 * (double)g.segs.nseg, true);
            A.kb = ks.dev();
            k_jet_select_seg<<<seg_grid(g.segs), kSegBlk, sizeof(int32_t) * (size_t)k + tg_bytes, c.stream>>>(
                g.segs, A, conn);
            HGP_LAUNCH_CHECK();
        }
        A.kb = nullptr;
        if (g.segs.nmulti) {
            KScope ks(c, "k_jet_select_fin", (8.0 * k + 24.0) * (double)g.segs.nmulti);
            const int32_t rows_per_block = kRowBS / kWave;
            k_jet_select_fin<<<(unsigned)std::min<int32_t>((g.segs.nmulti + rows_per_block - 1) / rows_per_block, 4096),
                               kRowBS, 0, c.stream>>>(g.segs, A, conn);
            HGP_LAUNCH_CHECK();
        }
    }
}

}  // namespace

void edge_cut(Ctx &c, const DevGraph &g, const int32_t *where, int64_t *d_cut)
{
    fill<int64_t>(c, d_cut, 1, 0);
    if (g.n <= 0) return;
    for_small_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
        constexpr int G = decltype(gc)::value;
        k_cut<G><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt, where,
                                                             (unsigned long long *)d_cut);
        HGP_LAUNCH_CHECK();
    });
    if (g.segs.nseg) {
        k_cut_seg<<<seg_grid(g.segs), kSegBlk, 0, c.stream>>>(g.segs, g.xadj, g.adjncy, g.adjwgt, where,
                                                            (unsigned long long *)d_cut);
        HGP_LAUNCH_CHECK();
    }
    k_half<<<1, 1, 0, c.stream>>>(d_cut);
    HGP_LAUNCH_CHECK();
}

void part_weights(Ctx &c, int32_t n, const wgt_t *vwgt, const int32_t *where, int32_t k, int64_t *pwgts)
{
    fill<int64_t>(c, pwgts, k, 0);
    if (n <= 0) return;
    unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)n + kBS - 1) / kBS, 1024);
    k_pwgts<<<blocks, kBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(n, k, vwgt, where,
                                                                              (unsigned long long *)pwgts);
    HGP_LAUNCH_CHECK();
}

void project(Ctx &c, int32_t n, const int32_t *cmap, const int32_t *where_c, int32_t *where_f)
{
    if (n <= 0) return;
    k_project<<<grid_for(n, kBS), kBS, 0, c.stream>>>(n, cmap, where_c, where_f);
    HGP_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- tiny levels

// A level of at most Tuning::tiny_max vertices (k <= kTinyK) runs its whole Jet loop
// in one block of kTinyBS threads: the same phases as the kernels above
// (selection, afterburner, LP filter, RW/RS eviction with its sort and cookie
// cutter, move accounting, commit, bookkeeping) separated by block barriers
// instead of launches, and no host round trip per iteration.  Rows are
// walked one per wave in chunks of 256 entries; per-vertex phases stride the
// block.  The block's 16 waves walk rows one after another, so the path
// pays off only while a level's rows are few: measured on grids, 3-D RGGs
// and banded graphs, 512 vertices beat both 256 and 1024.
constexpr int kTinyBS = 1024;
constexpr int kTinyWaves = kTinyBS / kWave;
constexpr int32_t kTinyN = 4096;   // chunk of the block-local scans and compactions
constexpr int32_t kTinyK = 256;
constexpr int kTinyChunk = 4;                         // entries per lane and chunk
constexpr eid_t kTinyRow = (eid_t)kTinyChunk * kWave;  // row entries per chunk

struct TinyArgs {
    SelArgs A;                   // selection inputs and outputs (dest, pregain, key, ab, pairs, under, ctl)
    int32_t n, sort_bits;
    int32_t *where, *under, *best, *pstart, *elist, *es, *ee;
    uint8_t *sel, *lock, *w8;
    uint16_t *w16;
    uint32_t *val, *k2, *v2;
    int64_t *pwgts, *wsc, *bs;
};

inline size_t tiny_lds(int32_t k)
{
    return sizeof(int32_t) * (size_t)kTinyWaves * k + (((size_t)k + 7) / 8) * 8 + sizeof(int64_t) * (size_t)k;
}

struct TinyShared {
    NextScratch nx;
    uint32_t cnt[256], base_d[256];
    uint32_t wcnt[kTinyWaves][256];
    uint32_t ssh[kTinyWaves + 1];
    int64_t lsh[kTinyWaves + 1];
    unsigned long long d2[kTinyWaves];
    int32_t done, mode;
};

// Inclusive scan of a[0..n) in place (n <= kTinyBS * 4), one block.
__device__ inline void tiny_scan_i64(int64_t *a, int32_t n, TinyShared &S)
{
    constexpr int E = kTinyN / kTinyBS;
    int64_t carry = 0;
    for (int32_t base = 0; base < n; base += kTinyN) {   // chunks of kTinyN, the running sum carried
        const int32_t m = min(n - base, kTinyN);
        const int32_t b = threadIdx.x * E;
        int64_t x[E], run = 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            x[e] = b + e < m ? a[base + b + e] : 0;
            run += x[e];
        }
        int64_t tot;
        int64_t off = carry + block_exclusive_scan<int64_t>(run, S.lsh, &tot);
#pragma unroll
        for (int e = 0; e < E; e++) {
            off += x[e];
            if (b + e < m) a[base + b + e] = off;
        }
        carry += tot;
        __syncthreads();
    }
}

// Stable LSD radix sort of (key, val)[0..n) over bits [0, hi) in 8-bit
// digits, one block; the result ends in key/val.
__device__ inline void tiny_sort(uint32_t *key, uint32_t *val, uint32_t *k2, uint32_t *v2, int32_t n, int hi,
                                 TinyShared &S)
{
    const int t = threadIdx.x, w = t / kWave;
    uint32_t *sk = key, *sv = val, *dk = k2, *dv = v2;
    for (int shift = 0; shift < hi; shift += 8) {
        if (t < 256) S.cnt[t] = 0;
        __syncthreads();
        for (int32_t i = t; i < n; i += kTinyBS) atomicAdd(&S.cnt[(sk[i] >> shift) & 255u], 1u);
        __syncthreads();
        uint32_t ex = block_exclusive_scan<uint32_t>(t < 256 ? S.cnt[t] : 0u, S.ssh, nullptr);
        if (t < 256) S.base_d[t] = ex;
        __syncthreads();
        for (int32_t first = 0; first < n; first += kTinyBS) {
            for (int i = t; i < kTinyWaves * 256; i += kTinyBS) (&S.wcnt[0][0])[i] = 0;
            __syncthreads();
            int32_t i = first + t;
            bool valid = i < n;
            uint32_t kk = valid ? sk[i] : 0u, vv = valid ? sv[i] : 0u;
            uint32_t d = (kk >> shift) & 255u;
            uint64_t m = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; bit++) {
                bool on = (d >> bit) & 1u;
                uint64_t bb = __ballot(on);
                m &= on ? bb : ~bb;
            }
            uint64_t lower = m & lanemask_lt();
            uint32_t rank = (uint32_t)__popcll(lower);
            if (valid && lower == 0) S.wcnt[w][d] = (uint32_t)__popcll(m);
            __syncthreads();
            if (t < 256) {
                uint32_t acc = S.base_d[t];
                for (int ww = 0; ww < kTinyWaves; ww++) {
                    uint32_t c = S.wcnt[ww][t];
                    S.wcnt[ww][t] = acc;
                    acc += c;
                }
                S.base_d[t] = acc;
            }
            __syncthreads();
            if (valid) {
                uint32_t pos = S.wcnt[w][d] + rank;
                dk[pos] = kk;
                dv[pos] = vv;
            }
            __syncthreads();
        }
        uint32_t *tk = sk; sk = dk; dk = tk;
        uint32_t *tv = sv; sv = dv; dv = tv;
    }
    if (sk != key) {
        for (int32_t i = t; i < n; i += kTinyBS) {
            key[i] = sk[i];
            val[i] = sv[i];
        }
    }
    __syncthreads();
}

// The whole Jet loop of one level in the calling block (kTinyBS threads).
// dyn: per-wave k-int conn rows | k part flags (8-byte padded) | k int64
// deltas (tiny_lds(k) bytes of dynamic LDS).
__device__ inline void jet_tiny_loop(const TinyArgs &T, int32_t *dyn, TinyShared &S)
{
    const SelArgs &A = T.A;
    JetCtl &C = *A.ctl;
    const int32_t n = T.n, k = A.k;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid / kWave;
    int32_t *my = dyn + (size_t)wv * k;
    uint8_t *tg = (uint8_t *)(dyn + (size_t)kTinyWaves * k);
    unsigned long long *pdelta = (unsigned long long *)(dyn + (size_t)kTinyWaves * k + (((size_t)k + 7) / 8) * 2);
    const PairParts pq{A.pd16, A.pd32};
    const uint32_t nokey = (uint32_t)k * kMaxBuckets;
    for (int i = tid; i < kTinyWaves * k; i += kTinyBS) dyn[i] = 0;
    __syncthreads();
    jet_next(C, k, T.pwgts, T.under, 1, S.nx);
    if (tid == 0) { S.done = C.done; S.mode = C.mode; }
    __syncthreads();
    while (!S.done) {
        const int32_t mode = S.mode;
        const int64_t maxwgt = C.maxwgt, opt = C.opt;

        // selection (k_jet_select_*): one row per wave
        for (int q = tid; q < k; q += kTinyBS) tg[q] = sel_target(A, mode, maxwgt, q) ? 1 : 0;
        __syncthreads();
        int32_t ncand = 0;
        for (int32_t v = wv; v < n; v += kTinyWaves) {
            const SelRow r = sel_row(A, mode, v);
            const int32_t p = r.p;
            if (!sel_eligible(mode, maxwgt, opt, r)) {
                if (lane == 0) sel_decide(A, C, mode, v, r, false, 0, PartKey{0, -1});
                continue;
            }
            const eid_t b = A.xadj[v], e = A.xadj[v + 1];
            int32_t u[kTinyChunk], w[kTinyChunk], q[kTinyChunk];
#pragma unroll
            for (int t = 0; t < kTinyChunk; t++) q[t] = -1;
            if (mode == kModeRS) {
                int32_t own = 0;
                for (eid_t c0 = b; c0 < e; c0 += kTinyRow) {
                    load_entries<kTinyChunk, kWave>(A.adjncy, A.adjwgt, c0, min(c0 + kTinyRow, e), lane, u, w);
                    A.wq.gather(u, q);
#pragma unroll
                    for (int t = 0; t < kTinyChunk; t++) own += q[t] == p ? w[t] : 0;
                }
                own = wave_sum(own);
                if (lane == 0) sel_decide(A, C, mode, v, r, true, own, PartKey{0, -1});
                continue;
            }
            const bool one = e - b <= kTinyRow;
            for (eid_t c0 = b; c0 < e; c0 += kTinyRow) {
                load_entries<kTinyChunk, kWave>(A.adjncy, A.adjwgt, c0, min(c0 + kTinyRow, e), lane, u, w);
                A.wq.gather(u, q);
#pragma unroll
                for (int t = 0; t < kTinyChunk; t++)
                    if (q[t] >= 0) atomicAdd(&my[q[t]], w[t]);
            }
            wave_lds_sync();
            const int32_t pc = my[p];
            PartKey best{0, -1};
            if (one) {
#pragma unroll
                for (int t = 0; t < kTinyChunk; t++) {
                    if (q[t] >= 0 && q[t] != p && tg[q[t]]) {
                        PartKey cand{my[q[t]], q[t]};
                        if (BetterPart()(cand, best)) best = cand;
                    }
                }
            } else {
                for (int qq = lane; qq < k; qq += kWave) {
                    if (qq == p || !tg[qq]) continue;
                    PartKey cand{my[qq], qq};
                    if (BetterPart()(cand, best)) best = cand;
                }
            }
            best = group_best<kWave>(best, BetterPart());
            wave_lds_sync();
            if (one) {
#pragma unroll
                for (int t = 0; t < kTinyChunk; t++)
                    if (q[t] >= 0) my[q[t]] = 0;
            } else {
                for (int qq = lane; qq < k; qq += kWave) my[qq] = 0;
            }
            wave_lds_sync();
            if (lane == 0) ncand += sel_decide(A, C, mode, v, r, true, pc, best);
        }
        const int any = __syncthreads_or(ncand > 0);
        if (mode == kModeLP && any && tid == 0) C.ncand = 1;

        if (mode == kModeLP) {
            // afterburner (k_afterburner*)
            for (int32_t v = wv; v < n; v += kTinyWaves) {
                const int32_t ig = A.pregain[v];
                if (ig == kNoGain) {
                    if (lane == 0) T.sel[v] = 0;
                    continue;
                }
                const int32_t p = T.where[v], bd = A.dest[v];
                const eid_t b = A.xadj[v], e = A.xadj[v + 1];
                int64_t change = 0;
                for (eid_t c0 = b; c0 < e; c0 += kTinyRow) {
                    int32_t u[kTinyChunk], w[kTinyChunk];
                    uint64_t x[kTinyChunk];
                    load_entries<kTinyChunk, kWave>(A.adjncy, A.adjwgt, c0, min(c0 + kTinyRow, e), lane, u, w);
#pragma unroll
                    for (int t = 0; t < kTinyChunk; t++) x[t] = u[t] >= 0 ? A.ab[u[t]] : 0ull;
#pragma unroll
                    for (int t = 0; t < kTinyChunk; t++)
                        if (u[t] >= 0) change += ab_change(v, ig, p, bd, u[t], w[t], x[t]);
                }
                change = wave_sum(change);
                if (lane == 0) T.sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
            }
            __syncthreads();
            // filter and locks (k_lp_finish)
            for (int32_t v = tid; v < n; v += kTinyBS) {
                const int32_t w = T.where[v];
                int32_t d = w;
                if (any) {
                    const uint8_t s8 = T.sel[v];
                    T.lock[v] = s8;
                    if (s8) d = A.dest[v];
                    else A.dest[v] = w;
                } else {
                    A.dest[v] = w;
                }
                store_pair(A.pd16, A.pd32, v, w, d);
            }
        } else {
            // RW / RS eviction (k_rb_*, k_rs_*)
            for (int32_t v = tid; v < n; v += kTinyBS) T.val[v] = (uint32_t)v;
            __syncthreads();
            tiny_sort(A.key, T.val, T.k2, T.v2, n, T.sort_bits, S);
            for (int32_t i = tid; i < n; i += kTinyBS) T.wsc[i] = A.key[i] < nokey ? A.vwgt[T.val[i]] : 0;
            __syncthreads();
            tiny_scan_i64(T.wsc, n, S);
            for (int32_t i = tid; i < n; i += kTinyBS) {
                uint32_t kk = A.key[i];
                if (kk >= nokey) continue;
                uint32_t p = kk / kMaxBuckets;
                if (i == 0 || A.key[i - 1] / kMaxBuckets != p) T.pstart[p] = i;
            }
            __syncthreads();
            for (int32_t i = tid; i < n; i += kTinyBS) {
                uint32_t kk = A.key[i];
                int32_t v = (int32_t)T.val[i];
                if (kk >= nokey) {
                    T.sel[i] = 0;
                    continue;
                }
                int32_t p = (int32_t)(kk / kMaxBuckets);
                int32_t s0 = T.pstart[p];
                int64_t before = T.wsc[i] - A.vwgt[v] - (s0 > 0 ? T.wsc[s0 - 1] : 0);
                bool evict = before < T.pwgts[p] - maxwgt;
                T.sel[i] = evict ? 1 : 0;
                if (mode != kModeRW || !evict) {
                    int32_t w = T.where[v];
                    A.dest[v] = w;
                    store_pair(A.pd16, A.pd32, v, w, w);
                }
            }
            __syncthreads();
            if (mode == kModeRS) {
                // the evicted list in sorted order, its weight prefix, the cookie cutter
                constexpr int E = kTinyN / kTinyBS;
                int32_t ne = 0;
                for (int32_t base = 0; base < n; base += kTinyN) {   // chunks of kTinyN, in order
                    const int32_t b0 = base + tid * E;
                    int32_t f[E], cnt = 0;
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        f[e] = (b0 + e < n && T.sel[b0 + e]) ? 1 : 0;
                        cnt += f[e];
                    }
                    uint32_t tot;
                    uint32_t off = (uint32_t)ne + block_exclusive_scan<uint32_t>((uint32_t)cnt, S.ssh, &tot);
#pragma unroll
                    for (int e = 0; e < E; e++)
                        if (f[e]) T.elist[off++] = b0 + e;
                    ne += (int32_t)tot;
                }
                __syncthreads();
                if (tid == 0) T.bs[0] = 0;
                for (int32_t i = tid; i < ne; i += kTinyBS) T.bs[i + 1] = A.vwgt[T.val[T.elist[i]]];
                __syncthreads();
                tiny_scan_i64(T.bs + 1, ne, S);
                if (tid == 0) {
                    int32_t start = 0;
                    for (int32_t p = 0; p < k; p++) {
                        T.es[p] = start;
                        int64_t cap = maxwgt > T.pwgts[p] ? maxwgt - T.pwgts[p] : 0;
                        int32_t end = start;
                        if (cap > 0 && start < ne) {
                            int64_t find = T.bs[start] + cap;
                            int32_t lo = start, hi = (int32_t)min((int64_t)ne, (int64_t)start + cap);
                            while (lo + 1 < hi) {
                                int32_t mid = (lo + hi) / 2;
                                if (T.bs[mid] >= find) hi = mid; else lo = mid;
                            }
                            int64_t dh = T.bs[hi] - find, dl = T.bs[lo] - find;
                            dh = dh < 0 ? -dh : dh;
                            dl = dl < 0 ? -dl : dl;
                            end = dh < dl ? hi : lo;
                        }
                        T.ee[p] = end;
                        start = end;
                    }
                }
                __syncthreads();
                for (int32_t i = tid; i < ne; i += kTinyBS) {
                    int32_t lo = 0, hi = k - 1;
                    while (lo < hi) {
                        int32_t mid = (lo + hi + 1) >> 1;
                        if (T.es[mid] <= i) lo = mid; else hi = mid - 1;
                    }
                    int32_t v = (int32_t)T.val[T.elist[i]];
                    int32_t w = T.where[v];
                    if (i < T.ee[lo] && lo != w) {
                        A.dest[v] = lo;
                        store_pair(A.pd16, A.pd32, v, w, lo);
                    }
                }
            }
        }
        __syncthreads();

        // move accounting (k_move_delta_*): part weights and twice the cut change
        for (int q = tid; q < k; q += kTinyBS) pdelta[q] = 0;
        __syncthreads();
        int64_t d2 = 0;
        for (int32_t v = wv; v < n; v += kTinyWaves) {
            const int32_t pv = T.where[v], dv = A.dest[v];
            if (pv == dv) continue;
            if (lane == 0) {
                const int32_t wvv = A.vwgt[v];
                atomicAdd(&pdelta[pv], (unsigned long long)(-(int64_t)wvv));
                atomicAdd(&pdelta[dv], (unsigned long long)(int64_t)wvv);
            }
            const eid_t b = A.xadj[v], e = A.xadj[v + 1];
            for (eid_t c0 = b; c0 < e; c0 += kTinyRow) {
                int32_t u[kTinyChunk], w[kTinyChunk], pu[kTinyChunk], du[kTinyChunk];
                load_entries<kTinyChunk, kWave>(A.adjncy, A.adjwgt, c0, min(c0 + kTinyRow, e), lane, u, w);
                pq.gather(u, pu, du);
#pragma unroll
                for (int t = 0; t < kTinyChunk; t++)
                    if (u[t] >= 0) d2 += move_d2(pv, dv, pu[t], du[t], w[t]);
            }
        }
        d2 = wave_sum(d2);
        if (lane == 0) S.d2[wv] = (unsigned long long)d2;
        __syncthreads();
        if (tid == 0) {
            unsigned long long tot = 0;
            for (int ww = 0; ww < kTinyWaves; ww++) tot += S.d2[ww];
            C.delta2 += tot;
        }
        for (int q = tid; q < k; q += kTinyBS)
            if (pdelta[q]) T.pwgts[q] += (int64_t)pdelta[q];
        __syncthreads();

        // commit (k_move_commit), bookkeeping and the next mode (k_jet_next), best copy (k_copy_best)
        for (int32_t v = tid; v < n; v += kTinyBS) {
            const int32_t d = A.dest[v];
            T.where[v] = d;
            if (T.w8) T.w8[v] = (uint8_t)d;
            else T.w16[v] = (uint16_t)d;
        }
        __syncthreads();
        jet_next(C, k, T.pwgts, T.under, 0, S.nx);
        if (tid == 0) { S.done = C.done; S.mode = C.mode; }
        const int32_t copy = C.copy_best;
        __syncthreads();
        if (copy)
            for (int32_t v = tid; v < n; v += kTinyBS) T.best[v] = T.where[v];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kTinyBS) k_jet_tiny(TinyArgs T)
{
    extern __shared__ int32_t dyn[];
    __shared__ TinyShared S;
    jet_tiny_loop(T, dyn, S);
}

// ---------------------------------------------------------------- small levels of many graphs

// One level of a run for k_jet_tiny_batch: its loop arguments (where, n and
// the graph set per level, scratch shared by the run's levels), the map of
// the next finer level onto it and that level's partition and size (the
// projection target; nfine = 0 after level 0).
struct TinyStep {
    TinyArgs T;
    const int32_t *cmap_fine;
    int32_t *where_fine;
    int32_t nfine;
    int32_t finest;
};

struct TinyRunDev {
    int32_t first, count;    // steps [first, first + count), coarsest level first
    int32_t patience, lo;
    int64_t maxwgt, opt;
    int64_t *cut;
    int32_t *iters;          // indexed by level
};

// refine_level's small-level path and project, level after level, in one
// block per run: lock cleared, best = where, part weights, the control block
// (k_jet_init), the narrow part copy, the Jet loop, where = best, the cut
// carried to the next level, the projection.
__global__ void __launch_bounds__(kTinyBS) k_jet_tiny_batch(const TinyRunDev *__restrict__ runs,
                                                           const TinyStep *__restrict__ steps)
{
    extern __shared__ int32_t dyn[];
    __shared__ TinyShared S;
    __shared__ int64_t cut;
    const TinyRunDev &R = runs[blockIdx.x];
    const int tid = threadIdx.x, lane = lane_id(), wv = tid / kWave;
    if (tid == 0) cut = *R.cut;
    __syncthreads();
    for (int32_t i = 0; i < R.count; i++) {
        const TinyStep &st = steps[R.first + i];
        const TinyArgs &T = st.T;
        const int32_t n = T.n, k = T.A.k;
        JetCtl &C = *T.A.ctl;
        for (int32_t v = tid; v < n; v += kTinyBS) {
            const int32_t w = T.where[v];
            T.lock[v] = 0;
            T.best[v] = w;
            if (T.w8) T.w8[v] = (uint8_t)w;
            else T.w16[v] = (uint16_t)w;
        }
        for (int32_t q = tid; q < k; q += kTinyBS) T.pwgts[q] = 0;
        __syncthreads();
        for (int32_t v = tid; v < n; v += kTinyBS)
            atomicAdd((unsigned long long *)&T.pwgts[T.where[v]], (unsigned long long)(int64_t)T.A.vwgt[v]);
        __syncthreads();
        int64_t mx = INT64_MIN;
        for (int32_t q = tid; q < k; q += kTinyBS) mx = T.pwgts[q] > mx ? T.pwgts[q] : mx;
        mx = wave_max(mx);
        if (lane == 0) S.lsh[wv] = mx;
        __syncthreads();
        if (tid == 0) {
            int64_t m = S.lsh[0];
            for (int w = 1; w < kTinyWaves; w++) m = S.lsh[w] > m ? S.lsh[w] : m;
            C.cut = C.best_cut = cut;
            C.best_max = m;
            C.maxwgt = R.maxwgt;
            C.opt = R.opt;
            const int64_t md = (99 * R.maxwgt) / 100;
            C.maxdest = md < R.maxwgt - 100 ? R.maxwgt - 100 : md;
            C.delta2 = 0;
            C.count = 0;
            C.balance_counter = 0;
            C.mode = 0;
            C.done = 0;
            C.iters = 0;
            C.patience = R.patience;
            C.ncand = 0;
            C.nunder = 0;
            C.copy_best = 0;
            C.filt_num = st.finest ? 1 : 3;
        }
        __syncthreads();
        jet_tiny_loop(T, dyn, S);
        for (int32_t v = tid; v < n; v += kTinyBS) T.where[v] = T.best[v];
        if (tid == 0) {
            cut = C.best_cut;
            R.iters[R.lo + R.count - 1 - i] = C.iters;
        }
        __syncthreads();
        for (int32_t v = tid; v < st.nfine; v += kTinyBS) st.where_fine[v] = T.where[st.cmap_fine[v]];
        __syncthreads();
    }
    if (tid == 0) *R.cut = cut;
}

// hunyuangraph_malloc_krefine (hunyuangraph_GPU_uncoarsen.h:428-508) + k_refine.
// Launch sequences of one level's Jet loop recorded as HIP graphs: slot 0
// the selection, slots 1-3 the rest of an iteration in modes LP, RW and RS.
// Each is recorded on its first use (stream capture, this thread only) and
// replayed by every later iteration of the level; the kernels read the mode
// and the stop flag from the control block themselves, and every temporary
// they use comes from the same arena mark in each iteration.
struct LevelGraphs {
    bool on = false;
    hipGraphExec_t ex[4] = {};
    LevelGraphs() = default;
    LevelGraphs(const LevelGraphs &) = delete;
    LevelGraphs &operator=(const LevelGraphs &) = delete;
    ~LevelGraphs()
    {
        for (hipGraphExec_t &e : ex)
            if (e) (void)hipGraphExecDestroy(e);
    }
    template <class F>
    void run(Ctx &c, int slot, F &&issue)
    {
        if (!ex[slot]) {
            hipGraph_t gr = nullptr;
            HGP_CHECK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            try {
                issue();
            } catch (...) {
                (void)hipStreamEndCapture(c.stream, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            HGP_CHECK_HIP(hipStreamEndCapture(c.stream, &gr));
            hipError_t e = hipGraphInstantiate(&ex[slot], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            HGP_CHECK_HIP(e);
        }
        HGP_CHECK_HIP(hipGraphLaunch(ex[slot], c.stream));
    }
};

RefineStats refine_level(Ctx &c, const DevGraph &g, int32_t k, int64_t tvwgt, bool finest, const Params &p,
                         int32_t *where, int64_t *d_cut)
{
    RefineStats rs;
    int32_t n = g.n;
    if (n <= 0 || k <= 1) return rs;
    if (k > 4096) throw Error{-4, "refine: k > 4096 is not supported"};
    if (p.refiner == 1) return lp_refine_level(c, g, k, tvwgt, where, d_cut);
    size_t mark = c.arena.right_mark();
    int64_t opt = tvwgt / k;
    int64_t maxwgt = ((int64_t)p.ub_num * tvwgt) / ((int64_t)p.ub_den * k);
    int64_t *pwgts = c.arena.ralloc<int64_t>((size_t)k);
    int32_t *under = c.arena.ralloc<int32_t>((size_t)k);
    int32_t *es = c.arena.ralloc<int32_t>((size_t)k);
    int32_t *ee = c.arena.ralloc<int32_t>((size_t)k);
    int32_t *pstart = c.arena.ralloc<int32_t>((size_t)k);
    JetCtl *ctl = c.arena.ralloc<JetCtl>(1);
    int32_t *dest = c.arena.ralloc<int32_t>((size_t)n);
    int32_t *pregain = c.arena.ralloc<int32_t>((size_t)n);
    int32_t *best = c.arena.ralloc<int32_t>((size_t)n);
    uint8_t *lock = c.arena.ralloc<uint8_t>((size_t)n);
    uint8_t *sel = c.arena.ralloc<uint8_t>((size_t)n);
    uint32_t *key = c.arena.ralloc<uint32_t>((size_t)n);
    uint32_t *val = c.arena.ralloc<uint32_t>((size_t)n);
    int64_t *wsc = c.arena.ralloc<int64_t>((size_t)n);
    int32_t *elist = c.arena.ralloc<int32_t>((size_t)n);
    int64_t *bs = c.arena.ralloc<int64_t>((size_t)n + 1);
    // compact copies for the neighbour gathers: where alone (selection),
    // where and dest in one word (move accounting), pregain, where and dest
    // in 8 bytes (afterburner)
    const bool byte_parts = k <= 256;
    uint8_t *w8 = byte_parts ? c.arena.ralloc<uint8_t>((size_t)n) : nullptr;
    uint16_t *w16 = byte_parts ? nullptr : c.arena.ralloc<uint16_t>((size_t)n);
    uint16_t *pd16 = byte_parts ? c.arena.ralloc<uint16_t>((size_t)n) : nullptr;
    uint32_t *pd32 = byte_parts ? nullptr : c.arena.ralloc<uint32_t>((size_t)n);
    uint64_t *ab = c.arena.ralloc<uint64_t>((size_t)n);
    const NarrowParts wq{w8, w16};
    const PairParts pq{pd16, pd32};
    int32_t *conn = c.arena.ralloc<int32_t>((size_t)g.segs.nmulti * (size_t)k);
    int64_t *macc = c.arena.ralloc<int64_t>((size_t)g.segs.nmulti);
    if (g.segs.nmulti) {
        fill<int32_t>(c, conn, (int64_t)g.segs.nmulti * k, 0);
        fill<int64_t>(c, macc, g.segs.nmulti, 0);
    }
    int64_t *d_ne = c.arena.ralloc<int64_t>(1);
    fill<uint8_t>(c, lock, n, 0);
    HGP_CHECK_HIP(hipMemcpyAsync(best, where, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    part_weights(c, n, g.vwgt, where, k, pwgts);
    k_jet_init<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts, d_cut, maxwgt, opt, p.patience, finest ? 1 : 3);
    k_narrow<<<grid_for(n, kBS), kBS, 0, c.stream>>>(n, where, w8, w16);   // refreshed by each commit
    HGP_LAUNCH_CHECK();
    // connectivity rows of the long rows, when asked for, while k-int rows
    // are short beside them (k <= kTrackMaxK) and the level runs the
    // multi-launch loop.  Off by default: on R-MAT 2^22 / 2^26 at k = 256 the
    // moved vertices' atomic updates of the rows of the heaviest hubs (the
    // same few addresses from every wave) cost 211 ms per partition in the
    // move accounting against 19 ms, and the selection saves only 17 ms.
    HubConn hc;
    hc.k = k;
    const int64_t nhub = g.bins.count(kClsLong);
    const bool tiny = n <= tuning().tiny_max && k <= kTinyK;
    if (nhub > 0 && k <= kTrackMaxK && !tiny && tuning().hub_track) {
        int32_t *hidx = c.arena.ralloc<int32_t>((size_t)n);
        hc.tconn = c.arena.ralloc<int32_t>((size_t)nhub * (size_t)k);
        fill<int32_t>(c, hidx, n, -1);
        fill<int32_t>(c, hc.tconn, nhub * (int64_t)k, 0);
        k_hub_index<<<grid_for(nhub, kBS), kBS, 0, c.stream>>>(g.bins.list(kClsLong), nhub, hidx);
        HGP_LAUNCH_CHECK();
        hc.hidx = hidx;
        KScope ks(c, "k_hub_build", 24.0 * (double)g.segs.nseg + 9.0 * (double)g.bins.ent[kClsLong] +
                                        4.0 * (double)nhub * k);
        k_hub_build<<<seg_grid(g.segs), kSegBlk, sizeof(int32_t) * (size_t)k, c.stream>>>(g.segs, g.xadj, g.adjncy,
                                                                                         g.adjwgt, wq, hc);
        HGP_LAUNCH_CHECK();
    }
    unsigned gv = grid_for(n, kBS);
    if (tiny) {
        // the whole loop in one block
        uint32_t *k2 = c.arena.ralloc<uint32_t>((size_t)n);
        uint32_t *v2 = c.arena.ralloc<uint32_t>((size_t)n);
        TinyArgs T;
        T.A = SelArgs{k,    g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, pwgts, lock, under, ctl,
                      dest, pregain, key, nullptr, wq, ab, pd16, pd32};
        T.n = n;
        T.sort_bits = bits_for((uint64_t)k * kMaxBuckets);
        T.where = where;
        T.under = under;
        T.best = best;
        T.pstart = pstart;
        T.elist = elist;
        T.es = es;
        T.ee = ee;
        T.sel = sel;
        T.lock = lock;
        T.w8 = w8;
        T.w16 = w16;
        T.val = val;
        T.k2 = k2;
        T.v2 = v2;
        T.pwgts = pwgts;
        T.wsc = wsc;
        T.bs = bs;
        const size_t lds = tiny_lds(k);
        {
            KScope ks(c, "k_jet_tiny", 0.0);
            k_jet_tiny<<<1, kTinyBS, lds, c.stream>>>(T);
            HGP_LAUNCH_CHECK();
        }
        read_scalars(c, ctl, sizeof(JetCtl));
    }
    // k_jet_next leaves a copy of the control block in the pinned buffer
    JetCtl *mirror = nullptr;
    if (!tiny) HGP_CHECK_HIP(hipHostGetDevicePointer((void **)&mirror, c.pinned, 0));
    if (!tiny) {
        k_jet_next<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts, under, 1, mirror);
        HGP_LAUNCH_CHECK();
    }
    // The rest of an iteration once its mode is known: the mode's moves, the
    // move accounting, the commit and the bookkeeping.
    auto issue_rest = [&](int32_t mode, int32_t iters) {
        if (mode == kModeLP) {
            {
                // fixed: ids, pregain, sel (9 B per row); candidates add xadj, where, dest and 16 B per entry
                if (int64_t nb = g.bins.few() ? g.bins.short_rows() : 0) {
                    KScope ks(c, "k_afterburner_few", 9.0 * (double)nb, true);
                    int32_t rpw = rows_per_wave(nb);
                    k_afterburner_w<<<wave_blocks(nb, rpw, 4096), kRowBS, 0, c.stream>>>(
                        g.bins.list(0), nb, g.xadj, g.adjncy, g.adjwgt, where, dest, ab, pregain, ctl, sel, rpw,
                        ks.dev());
                    HGP_LAUNCH_CHECK();
                }
                if (int64_t nb = g.bins.few() ? 0 : g.bins.count(0)) {
                    KScope ks(c, "k_afterburner_s", 9.0 * (double)nb, true);
                    k_afterburner<8><<<grid_groups(nb, 8), kRowBS, 0, c.stream>>>(
                        g.bins.list(0), nb, g.xadj, g.adjncy, g.adjwgt, where, dest, ab, pregain, ctl, sel, ks.dev());
                    HGP_LAUNCH_CHECK();
                }
                if (int64_t nb = g.bins.few() ? 0 : g.bins.count(1)) {
                    KScope ks(c, "k_afterburner_m", 9.0 * (double)nb, true);
                    k_afterburner<32><<<grid_groups(nb, 32), kRowBS, 0, c.stream>>>(
                        g.bins.list(1), nb, g.xadj, g.adjncy, g.adjwgt, where, dest, ab, pregain, ctl, sel, ks.dev());
                    HGP_LAUNCH_CHECK();
                }
                if (int64_t nb = g.bins.few() ? 0 : g.bins.count(2)) {
                    KScope ks(c, "k_afterburner_w", 9.0 * (double)nb, true);
                    int32_t rpw = rows_per_wave(nb);
                    k_afterburner_w<<<wave_blocks(nb, rpw, 4096), kRowBS, 0, c.stream>>>(
                        g.bins.list(2), nb, g.xadj, g.adjncy, g.adjwgt, where, dest, ab, pregain, ctl, sel, rpw,
                        ks.dev());
                    HGP_LAUNCH_CHECK();
                }
                if (g.segs.nseg) {
                    {
                        KScope ks(c, "k_afterburner_seg", 24.0 * (double)g.segs.nseg, true);
                        k_afterburner_seg<<<seg_grid(g.segs), kSegBlk, 0, c.stream>>>(
                            g.segs, g.xadj, g.adjncy, g.adjwgt, where, dest, ab, pregain, ctl, sel,
                            (unsigned long long *)macc, ks.dev());
                        HGP_LAUNCH_CHECK();
                    }
                    if (g.segs.nmulti) {
                        k_afterburner_fin<<<grid_for(g.segs.nmulti, kBS), kBS, 0, c.stream>>>(g.segs, pregain, ctl,
                                                                                             sel, macc);
                        HGP_LAUNCH_CHECK();
                    }
                }
            }
            k_lp_finish<<<gv, kBS, 0, c.stream>>>(n, ctl, sel, where, dest, lock, pd16, pd32);
            HGP_LAUNCH_CHECK();
        } else {
            radix_sort_index(c, key, val, n, 0, bits_for((uint64_t)k * kMaxBuckets));
            k_rb_weights<<<gv, kBS, 0, c.stream>>>(n, k, key, val, g.vwgt, wsc, pstart);
            HGP_LAUNCH_CHECK();
            scan_inclusive(c, wsc, wsc, n);
            k_rb_evict<<<gv, kBS, 0, c.stream>>>(n, k, key, val, wsc, g.vwgt, pstart, pwgts, ctl, where, dest, sel,
                                                 pd16, pd32);
            HGP_LAUNCH_CHECK();
            if (mode == kModeRS) {
                compact_indices(c, sel, n, elist, d_ne);
                k_gather_w<<<gv, kBS, 0, c.stream>>>(d_ne, elist, val, g.vwgt, bs);
                HGP_LAUNCH_CHECK();
                scan_inclusive(c, bs + 1, bs + 1, n);
                k_rs_cookie<<<1, kWave, sizeof(int64_t) * (size_t)k, c.stream>>>(k, bs, d_ne, pwgts, ctl, es, ee);
                HGP_LAUNCH_CHECK();
                k_rs_assign<<<gv, kBS, 0, c.stream>>>(k, d_ne, elist, val, es, ee, where, dest, pd16, pd32);
                HGP_LAUNCH_CHECK();
            }
        }
        // fixed: ids, where, dest (12 B per row); moved rows add xadj, vwgt and 16 B per entry
        if (int64_t nb = g.bins.few() ? g.bins.short_rows() : 0) {
            KScope ks(c, "k_move_delta_few", 12.0 * (double)nb, true);
            int32_t rpw = rows_per_wave(nb);
            k_move_delta_w<<<wave_blocks(nb, rpw, 2048), kRowBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                g.bins.list(0), nb, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, pq, hc, ctl, pwgts, rpw,
                ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (int64_t nb = g.bins.few() ? 0 : g.bins.count(0)) {
            KScope ks(c, "k_move_delta_s", 12.0 * (double)nb, true);
            int64_t blocks = std::min<int64_t>((nb + kRowBS - 1) / kRowBS, 2048);
            k_move_delta_g<8><<<(unsigned)blocks, kRowBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                g.bins.list(0), nb, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, pq, hc, ctl, pwgts, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (int64_t nb = g.bins.few() ? 0 : g.bins.count(1)) {
            KScope ks(c, "k_move_delta_m", 12.0 * (double)nb, true);
            int64_t blocks = std::min<int64_t>((nb + kRowBS - 1) / kRowBS, 2048);
            k_move_delta_g<32><<<(unsigned)blocks, kRowBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                g.bins.list(1), nb, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, pq, hc, ctl, pwgts, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (int64_t nb = g.bins.few() ? 0 : g.bins.count(2)) {
            KScope ks(c, "k_move_delta_w", 12.0 * (double)nb, true);
            int32_t rpw = rows_per_wave(nb);
            k_move_delta_w<<<wave_blocks(nb, rpw, 2048), kRowBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                g.bins.list(2), nb, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, pq, hc, ctl, pwgts, rpw,
                ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (g.segs.nseg) {
            KScope ks(c, "k_move_delta_seg", 24.0 * (double)g.segs.nseg, true);
            unsigned blocks = (unsigned)std::min<int64_t>(g.segs.nseg, 4096);
            k_move_delta_seg<<<blocks, kSegBlk, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                g.segs, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, pq, hc, ctl, pwgts, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        {
            JetTrace &jt = jet_trace();
            std::lock_guard<std::mutex> lk(jt.mu);
            if (jt.on && jt.dump_n == n && jt.dump_iter == iters) {
                jt.where.resize((size_t)n);
                jt.dest.resize((size_t)n);
                HGP_CHECK_HIP(hipMemcpyAsync(jt.where.data(), where, sizeof(int32_t) * (size_t)n,
                                             hipMemcpyDeviceToHost, c.stream));
                HGP_CHECK_HIP(hipMemcpyAsync(jt.dest.data(), dest, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                                             c.stream));
                HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
            }
        }
        k_move_commit<<<gv, kBS, 0, c.stream>>>(n, ctl, dest, where, w8, w16, best);
        HGP_LAUNCH_CHECK();
        k_jet_next<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts, under, 0, mirror);
        HGP_LAUNCH_CHECK();
    };
    auto issue_select = [&] {
        launch_select(c, g, k, where, pwgts, lock, under, ctl, dest, pregain, key, conn, wq, ab, pd16, pd32, hc);
    };
    // Levels of at most Tuning::graph_max_vertices vertices record each
    // launch sequence once as a HIP graph and replay it (not while kernels are
    // timed or iterations traced, which act between the launches).
    LevelGraphs graphs;
    graphs.on = !tiny && n <= tuning().graph_max_vertices && !ktimer().on && !jet_trace().on;
    int32_t rounds = 0;   // iterations issued: the control block's count once they are done
    for (; !tiny;) {
        // the selection does not depend on the mode on the host: it runs while
        // the host waits for the last k_jet_next (and does nothing once the
        // loop is done)
        HGP_CHECK_HIP(hipEventRecord(c.ev[7], c.stream));
        if (graphs.on) graphs.run(c, 0, issue_select);
        else issue_select();
        read_scalars_wait(c);
        JetCtl h;
        std::memcpy(&h, c.pinned, sizeof(JetCtl));
        if (h.iters != rounds) {
            // the copy did not show this iteration's count: read the block itself
            read_scalars(c, ctl, sizeof(JetCtl));
            h = *(JetCtl *)c.pinned;
        }
        if (jet_trace().on) {
            std::lock_guard<std::mutex> lk(jet_trace().mu);
            jet_trace().recs.insert(jet_trace().recs.end(),
                                    {(int64_t)n, (int64_t)h.iters, (int64_t)h.mode, h.cut, (int64_t)h.done});
        }
        if (h.done) break;
        if (graphs.on) graphs.run(c, 1 + h.mode, [&] { issue_rest(h.mode, h.iters); });
        else issue_rest(h.mode, h.iters);
        rounds++;
    }
    rs.iters = ((JetCtl *)c.pinned)->iters;
    if (!tiny) {
        // the last iteration's keep decision (the next commit would have made it)
        k_copy_best<<<gv, kBS, 0, c.stream>>>(n, ctl, where, best);
        HGP_LAUNCH_CHECK();
    }
    HGP_CHECK_HIP(hipMemcpyAsync(where, best, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    k_jet_fini<<<1, 1, 0, c.stream>>>(ctl, d_cut);
    HGP_LAUNCH_CHECK();
    c.arena.right_release(mark);
    return rs;
}


TinyScratch alloc_tiny_scratch(Ctx &c, int32_t n, int32_t k)
{
    TinyScratch s;
    const size_t N = (size_t)(n > 0 ? n : 1), K = (size_t)k;
    s.dest = c.arena.lalloc<int32_t>(N);
    s.pregain = c.arena.lalloc<int32_t>(N);
    s.best = c.arena.lalloc<int32_t>(N);
    s.elist = c.arena.lalloc<int32_t>(N);
    s.under = c.arena.lalloc<int32_t>(K);
    s.es = c.arena.lalloc<int32_t>(K);
    s.ee = c.arena.lalloc<int32_t>(K);
    s.pstart = c.arena.lalloc<int32_t>(K);
    s.iters = c.arena.lalloc<int32_t>(HGP_MAX_LEVELS);
    s.lock = c.arena.lalloc<uint8_t>(N);
    s.sel = c.arena.lalloc<uint8_t>(N);
    if (k <= 256) {
        s.w8 = c.arena.lalloc<uint8_t>(N);
        s.pd16 = c.arena.lalloc<uint16_t>(N);
    } else {
        s.w16 = c.arena.lalloc<uint16_t>(N);
        s.pd32 = c.arena.lalloc<uint32_t>(N);
    }
    s.key = c.arena.lalloc<uint32_t>(N);
    s.val = c.arena.lalloc<uint32_t>(N);
    s.k2 = c.arena.lalloc<uint32_t>(N);
    s.v2 = c.arena.lalloc<uint32_t>(N);
    s.ab = c.arena.lalloc<uint64_t>(N);
    s.wsc = c.arena.lalloc<int64_t>(N);
    s.bs = c.arena.lalloc<int64_t>(N + 1);
    s.pwgts = c.arena.lalloc<int64_t>(K);
    s.cut = c.arena.lalloc<int64_t>(1);
    s.ctl = c.arena.lalloc<JetCtl>(1);
    fill<int32_t>(c, s.iters, HGP_MAX_LEVELS, 0);
    return s;
}

void tiny_batch(Ctx &c, const std::vector<TinyRun> &runs)
{
    if (runs.empty()) return;
    std::vector<TinyRunDev> rd;
    std::vector<TinyStep> sd;
    int32_t kmax = 1;
    for (const TinyRun &r : runs) {
        if (r.k > kTinyMaxK || r.k < 2) throw Error{-4, "tiny_batch: k out of range"};
        kmax = std::max(kmax, r.k);
        TinyRunDev d;
        d.first = (int32_t)sd.size();
        d.count = r.hi - r.lo + 1;
        d.patience = r.p.patience;
        d.lo = r.lo;
        d.opt = r.tvwgt / r.k;
        d.maxwgt = ((int64_t)r.p.ub_num * r.tvwgt) / ((int64_t)r.p.ub_den * r.k);
        d.cut = r.s.cut;
        d.iters = r.s.iters;
        rd.push_back(d);
        const TinyScratch &s = r.s;
        for (int32_t l = r.hi; l >= r.lo; l--) {
            const DevGraph &g = r.lv[(size_t)l];
            if (g.n > kTinyMaxN) throw Error{-4, "tiny_batch: level too large"};
            TinyStep st;
            TinyArgs &T = st.T;
            int32_t *where = r.where[(size_t)l];
            T.A = SelArgs{r.k,      g.xadj,   g.adjncy, g.adjwgt, g.vwgt, where, s.pwgts, s.lock,
                          s.under,  (JetCtl *)s.ctl, s.dest, s.pregain, s.key, nullptr, NarrowParts{s.w8, s.w16},
                          s.ab, s.pd16, s.pd32};
            T.n = g.n;
            T.sort_bits = bits_for((uint64_t)r.k * kMaxBuckets);
            T.where = where;
            T.under = s.under;
            T.best = s.best;
            T.pstart = s.pstart;
            T.elist = s.elist;
            T.es = s.es;
            T.ee = s.ee;
            T.sel = s.sel;
            T.lock = s.lock;
            T.w8 = s.w8;
            T.w16 = s.w16;
            T.val = s.val;
            T.k2 = s.k2;
            T.v2 = s.v2;
            T.pwgts = s.pwgts;
            T.wsc = s.wsc;
            T.bs = s.bs;
            st.finest = l == 0;
            if (l > 0 && (l > r.lo || r.lo > 0)) {
                st.cmap_fine = r.cmap[(size_t)l - 1];
                st.where_fine = r.where[(size_t)l - 1];
                st.nfine = r.lv[(size_t)l - 1].n;
            } else {
                st.cmap_fine = nullptr;
                st.where_fine = nullptr;
                st.nfine = 0;
            }
            sd.push_back(st);
        }
    }
    const size_t rb = sizeof(TinyRunDev) * rd.size(), sb = sizeof(TinyStep) * sd.size();
    const size_t mark = c.arena.right_mark();
    char *dev = c.arena.ralloc<char>(rb + sb + 256);
    TinyRunDev *drun = (TinyRunDev *)dev;
    TinyStep *dstep = (TinyStep *)(dev + ((rb + 255) & ~(size_t)255));
    std::vector<char> host(((rb + 255) & ~(size_t)255) + sb);
    std::memcpy(host.data(), rd.data(), rb);
    std::memcpy(host.data() + ((rb + 255) & ~(size_t)255), sd.data(), sb);
    HGP_CHECK_HIP(hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, c.stream));
    {
        KScope ks(c, "k_jet_tiny_batch", 0.0);
        k_jet_tiny_batch<<<(unsigned)rd.size(), kTinyBS, tiny_lds(kmax), c.stream>>>(drun, dstep);
        HGP_LAUNCH_CHECK();
    }
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.arena.right_release(mark);
}

}  // namespace hgp
