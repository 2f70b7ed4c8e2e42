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

#include <algorithm>
#include <type_traits>

namespace hgp {

namespace {

constexpr int kBS = 256;
constexpr int32_t kNoGain = INT32_MIN;

enum Mode : int32_t { kModeLP = 0, kModeRW = 1, kModeRS = 2 };

struct JetCtl {
    int64_t cut, best_cut, best_max, maxwgt, opt, maxdest;
    unsigned long long delta2;      // 2 x cut change (two's complement)
    int32_t count, balance_counter, mode, done, iters, patience;
    int32_t ncand, nunder, copy_best, filt_num;
    int64_t nsorted;                // candidates of RW/RS
    int64_t nevict;
};

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

// Start of an iteration: patience check, mode choice (:4661-4700) and, for
// the weak rebalancer, the undersized list in part order (:3136-3163).
__global__ void __launch_bounds__(kBS) k_jet_begin(JetCtl *ctl, int32_t k, const int64_t *__restrict__ pwgts,
                                                   int32_t *__restrict__ under)
{
    __shared__ int64_t smax[kBS];
    __shared__ int32_t sh[kBS / kWave + 1];
    JetCtl &C = *ctl;
    if (C.done) return;
    int64_t mx = INT64_MIN;
    for (int q = threadIdx.x; q < k; q += kBS) mx = pwgts[q] > mx ? pwgts[q] : mx;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && smax[threadIdx.x + s] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + s];
        __syncthreads();
    }
    bool stop = C.count >= C.patience;
    bool balanced = smax[0] <= C.maxwgt;
    int32_t mode = balanced ? kModeLP : (C.balance_counter < 2 ? kModeRW : kModeRS);
    if (!stop && mode == kModeRW) {
        int32_t base = 0;
        for (int q0 = 0; q0 < k; q0 += kBS) {
            int q = q0 + threadIdx.x;
            int32_t f = (q < k && pwgts[q] < C.maxdest) ? 1 : 0;
            int32_t tot;
            int32_t ex = block_exclusive_scan(f, sh, &tot);
            if (f) under[base + ex] = q;
            base += tot;
        }
        if (threadIdx.x == 0) C.nunder = base;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (stop) { C.done = 1; return; }
        C.count++;
        C.mode = mode;
        if (balanced) C.balance_counter = 0; else C.balance_counter++;
        C.ncand = 0;
        C.delta2 = 0;
        C.copy_best = 0;
        C.nsorted = 0;
        C.nevict = 0;
    }
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
};

__device__ inline bool sel_eligible(const SelArgs &A, int32_t mode, int64_t maxwgt, int64_t opt, int32_t v,
                                    int32_t p)
{
    if (mode == kModeLP) return true;
    int64_t pw = A.pwgts[p];
    if (mode == kModeRW) return pw > maxwgt && 2 * (int64_t)A.vwgt[v] < 3 * (pw - opt);
    return pw > maxwgt;
}

// Whether part q may receive v in this mode (q != p).
__device__ inline bool sel_target(const SelArgs &A, int32_t mode, int64_t maxwgt, int32_t q)
{
    if (mode == kModeRS) return false;
    return mode != kModeRW || A.pwgts[q] < maxwgt;
}

// Writes dest/pregain/key of v from (pc, best); returns 1 for an LP candidate.
__device__ inline int32_t sel_decide(const SelArgs &A, const JetCtl &C, int32_t mode, int32_t v, int32_t p,
                                     bool eligible, int32_t pc, PartKey best)
{
    const uint32_t nokey = (uint32_t)A.k * kMaxBuckets;
    int32_t bc = best.c, bq = best.q;
    int32_t d = p, cand = 0;
    if (mode == kModeLP) {
        int32_t gain = kNoGain;
        if (bq >= 0 && bc > 0) {
            int32_t fn = C.filt_num;
            if (bc >= pc || (int64_t)(pc - bc) < ((int64_t)fn * pc) / 4) {
                d = bq;
                if (!A.lock[v]) { gain = bc - pc; cand = 1; }
            }
        }
        A.pregain[v] = gain;
    } else if (mode == kModeRW) {
        uint32_t kk = nokey;
        if (eligible) {
            int64_t gain = 0;
            bool mv = false;
            if (bq >= 0 && bc > 0) { d = bq; gain = (int64_t)bc - pc; mv = true; }
            else if (C.nunder > 0) { d = A.under[v % C.nunder]; gain = -(int64_t)pc; mv = true; }
            if (mv && d != p) kk = (uint32_t)p * kMaxBuckets + (uint32_t)gain_bucket(gain, A.vwgt[v]);
            else d = p;
        }
        A.key[v] = kk;
    } else {
        uint32_t kk = nokey;
        if (eligible) {
            int32_t bkt = gain_bucket(-(int64_t)pc, A.vwgt[v]);
            if (bkt < kMaxBuckets && (int64_t)A.vwgt[v] < 2 * (A.pwgts[p] - C.opt))
                kk = (uint32_t)p * kMaxBuckets + (uint32_t)bkt;
        }
        A.key[v] = kk;
    }
    A.dest[v] = d;
    return cand;
}

__device__ inline void sel_count(const SelArgs &A, int32_t mode, int32_t ncand)
{
    if (mode != kModeLP) return;
    int32_t s = wave_sum(ncand);
    if (lane_id() == 0 && s) atomicAdd(&A.ctl->ncand, s);
}

// Orders this wave's LDS accesses before and after the point (the wave runs
// in lockstep, so completing its own LDS operations is enough).
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Rows of <= 8 entries: 8 lanes, one entry each; a lane's connectivity to its
// entry's part is summed over the group with shuffles (no LDS).
__global__ void __launch_bounds__(kRowBS) k_jet_select_s(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A)
{
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    int64_t gi = gindex<8>();
    int lane = glane<8>();
    int32_t ncand = 0;
    if (gi < cnt) {
        int32_t v = ids[gi];
        int32_t p = A.where[v];
        bool eligible = sel_eligible(A, mode, maxwgt, opt, v, p);
        eid_t j = A.xadj[v] + lane;
        int32_t q = -1, w = 0;
        if (eligible && j < A.xadj[v + 1]) { q = A.where[A.adjncy[j]]; w = A.adjwgt[j]; }
        int32_t conn = w;
#pragma unroll
        for (int o = 1; o < 8; o++) {
            int32_t qq = __shfl_xor(q, o, 8), ww = __shfl_xor(w, o, 8);
            if (qq == q) conn += ww;
        }
        int32_t pc = group_sum<8>(q == p ? w : 0);
        PartKey best{0, -1};
        if (q >= 0 && q != p && sel_target(A, mode, maxwgt, q)) best = PartKey{conn, q};
        best = group_best<8>(best, BetterPart());
        if (lane == 0) ncand = sel_decide(A, C, mode, v, p, eligible, pc, best);
    }
    sel_count(A, mode, ncand);
}

// Rows of 9..512 entries: one wave per row, entries cached in registers (8 per
// lane), a k-int LDS accumulator per wave.  Waves loop over rows on their own.
__global__ void __launch_bounds__(kRowBS) k_jet_select_w(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A)
{
    extern __shared__ int32_t acc[];
    constexpr int S = (int)(kBinMid / kWave);
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k, lane = lane_id();
    const int wv = threadIdx.x / kWave, nw = kRowBS / kWave;
    int32_t *my = acc + (size_t)wv * k;
    for (int i = lane; i < k; i += kWave) my[i] = 0;
    wave_lds_sync();
    int32_t ncand = 0;
    for (int64_t gi = (int64_t)blockIdx.x * nw + wv; gi < cnt; gi += (int64_t)gridDim.x * nw) {
        int32_t v = ids[gi];
        int32_t p = A.where[v];
        bool eligible = sel_eligible(A, mode, maxwgt, opt, v, p);
        eid_t b = A.xadj[v], e = A.xadj[v + 1];
        int32_t q[S], w[S];
#pragma unroll
        for (int t = 0; t < S; t++) {
            eid_t j = b + lane + (eid_t)t * kWave;
            q[t] = -1;
            w[t] = 0;
            if (eligible && j < e) {
                q[t] = A.where[A.adjncy[j]];
                w[t] = A.adjwgt[j];
                atomicAdd(&my[q[t]], w[t]);
            }
        }
        wave_lds_sync();
        int32_t pc = eligible ? my[p] : 0;
        PartKey best{0, -1};
#pragma unroll
        for (int t = 0; t < S; t++) {
            if (q[t] >= 0 && q[t] != p && sel_target(A, mode, maxwgt, q[t])) {
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
        if (lane == 0) ncand += sel_decide(A, C, mode, v, p, eligible, pc, best);
    }
    sel_count(A, mode, ncand);
}

// Rows of > 512 entries: one block per row and a k-int LDS accumulator.  The
// arg-max and the clearing walk the k parts when the row is longer than k,
// else the row again.
__global__ void __launch_bounds__(kRowBS) k_jet_select_b(const int32_t *__restrict__ ids, int64_t cnt, SelArgs A)
{
    extern __shared__ int32_t acc[];
    const JetCtl &C = *A.ctl;
    if (C.done) return;
    const int32_t mode = C.mode;
    const int64_t maxwgt = C.maxwgt, opt = C.opt;
    const int k = A.k;
    for (int i = threadIdx.x; i < k; i += kRowBS) acc[i] = 0;
    __syncthreads();
    int32_t ncand = 0;
    for (int64_t gi = blockIdx.x; gi < cnt; gi += gridDim.x) {
        int32_t v = ids[gi];
        int32_t p = A.where[v];
        bool eligible = sel_eligible(A, mode, maxwgt, opt, v, p);
        eid_t b = A.xadj[v], e = A.xadj[v + 1];
        bool walk_parts = e - b > (eid_t)k;
        if (eligible)
            for (eid_t j = b + threadIdx.x; j < e; j += kRowBS) atomicAdd(&acc[A.where[A.adjncy[j]]], A.adjwgt[j]);
        __syncthreads();
        int32_t pc = eligible ? acc[p] : 0;
        PartKey best{0, -1};
        if (eligible) {
            if (walk_parts) {
                for (int q = threadIdx.x; q < k; q += kRowBS) {
                    if (q == p || !sel_target(A, mode, maxwgt, q)) continue;
                    PartKey cand{acc[q], q};
                    if (BetterPart()(cand, best)) best = cand;
                }
            } else {
                for (eid_t j = b + threadIdx.x; j < e; j += kRowBS) {
                    int32_t q = A.where[A.adjncy[j]];
                    if (q == p || !sel_target(A, mode, maxwgt, q)) continue;
                    PartKey cand{acc[q], q};
                    if (BetterPart()(cand, best)) best = cand;
                }
            }
        }
        best = group_best<kRowBS>(best, BetterPart());
        if (eligible) {
            if (walk_parts) {
                for (int q = threadIdx.x; q < k; q += kRowBS) acc[q] = 0;
            } else {
                for (eid_t j = b + threadIdx.x; j < e; j += kRowBS) acc[A.where[A.adjncy[j]]] = 0;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) ncand += sel_decide(A, C, mode, v, p, eligible, pc, best);
    }
    sel_count(A, mode, ncand);
}

// Afterburner (afterburner_heuristic, :2137-2178): recompute a candidate's
// gain assuming every higher-priority neighbour (greater gain, ties to the
// smaller id) has already moved.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_afterburner(const int32_t *__restrict__ ids, int64_t cnt,
                                                        const eid_t *__restrict__ xadj,
                                                        const vid_t *__restrict__ adjncy,
                                                        const wgt_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ where,
                                                        const int32_t *__restrict__ dest,
                                                        const int32_t *__restrict__ pregain, const JetCtl *ctl,
                                                        uint8_t *__restrict__ sel)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t v = ids[gi];
    int32_t ig = pregain[v];
    if (ig == kNoGain) {
        if (lane == 0) sel[v] = 0;
        return;
    }
    int32_t p = where[v], best = dest[v];
    int64_t change = 0;
    for (eid_t j = xadj[v] + lane; j < xadj[v + 1]; j += G) {
        int32_t u = adjncy[j];
        int32_t ng = pregain[u];
        if (ng > ig || (ng == ig && u < v)) {
            int32_t w = adjwgt[j];
            int32_t nd = dest[u];
            if (nd == p) change -= w; else if (nd == best) change += w;
            int32_t nw = where[u];
            if (nw == p) change += w; else if (nw == best) change -= w;
        }
    }
    change = group_sum<G>(change);
    if (lane == 0) sel[v] = ((int64_t)ig + change >= 0) ? 1 : 0;
}

// filter_beneficial_moves + set_lock (:2180-2221): with candidates present,
// locks become exactly the selected set; without any, nothing changes.
__global__ void k_lp_finish(int32_t n, const JetCtl *ctl, const uint8_t *__restrict__ sel,
                            const int32_t *__restrict__ where, int32_t *__restrict__ dest,
                            uint8_t *__restrict__ lock)
{
    if (ctl->done || ctl->mode != kModeLP) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    if (ctl->ncand == 0) {
        dest[v] = where[v];
        return;
    }
    uint8_t s = sel[v];
    lock[v] = s;
    if (!s) dest[v] = where[v];
}

// ---------------------------------------------------------------- rebalancers

__global__ void k_iota_u32(int32_t n, uint32_t *__restrict__ val)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) val[v] = (uint32_t)v;
}

// After the sort by (key, id): exclusive prefix weight of each candidate
// inside its part, via an inclusive scan of weights and the part's first
// position.
__global__ void k_rb_weights(int32_t n, int32_t k, const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                             const wgt_t *__restrict__ vwgt, int64_t *__restrict__ w)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    w[i] = key[i] < (uint32_t)k * kMaxBuckets ? vwgt[val[i]] : 0;
}

__global__ void k_rb_pstart(int32_t n, int32_t k, const uint32_t *__restrict__ key, int32_t *__restrict__ pstart)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    uint32_t kk = key[i];
    if (kk >= (uint32_t)k * kMaxBuckets) return;
    uint32_t p = kk / kMaxBuckets;
    if (i == 0 || key[i - 1] / kMaxBuckets != p) pstart[p] = i;
}

// Evict a candidate iff the candidates ahead of it in its part weigh less
// than the part's excess (filter_scores_below_cutoff :3344-3366 for RW,
// find_score_cutoffs/filter_below_cutoffs :2641-2691 for RS).  RW keeps the
// destination chosen in selection; RS only flags the evicted ones.
__global__ void k_rb_evict(int32_t n, int32_t k, const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                           const int64_t *__restrict__ wsc, const wgt_t *__restrict__ vwgt,
                           const int32_t *__restrict__ pstart, const int64_t *__restrict__ pwgts, const JetCtl *ctl,
                           const int32_t *__restrict__ where, int32_t *__restrict__ dest, uint8_t *__restrict__ ev)
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
    if (ctl->mode == kModeRW) {
        if (!evict) dest[v] = where[v];
    } else {
        dest[v] = where[v];
    }
}

// Cookie cutter (:2710-2745): slice the evicted list (in sorted order) into
// consecutive runs, part by part, each filling maxwgt - pwgts.
__global__ void k_rs_cookie(int32_t k, const int64_t *__restrict__ bs, const int64_t *__restrict__ ne_ptr,
                            const int64_t *__restrict__ pwgts, const JetCtl *ctl, int32_t *__restrict__ es,
                            int32_t *__restrict__ ee)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t ne = (int32_t)*ne_ptr;
    int64_t maxwgt = ctl->maxwgt;
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
        int64_t dh = bs[hi] - find, dl = bs[lo] - find;
        dh = dh < 0 ? -dh : dh;
        dl = dl < 0 ? -dl : dl;
        int32_t end = dh < dl ? hi : lo;
        ee[p] = end;
        start = end;
    }
}

// select_dest_parts_rs (:2747-2766) over the compacted evicted list.
__global__ void k_rs_assign(int32_t k, const int64_t *__restrict__ ne_ptr, const int32_t *__restrict__ elist,
                            const uint32_t *__restrict__ val, const int32_t *__restrict__ es,
                            const int32_t *__restrict__ ee, const int32_t *__restrict__ where,
                            int32_t *__restrict__ dest)
{
    int32_t i = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= *ne_ptr) return;
    int32_t lo = 0, hi = k - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (es[mid] <= i) lo = mid; else hi = mid - 1;
    }
    int32_t v = (int32_t)val[elist[i]];
    if (i < ee[lo] && lo != where[v]) dest[v] = lo;
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

// Exact cut change of applying dest[] (edges between two moved vertices are
// seen from both ends, so each side counts once and one-sided edges twice),
// plus part-weight updates (perform_moves, :4034-4155).
// Persistent blocks over the row list; the cut change and the part-weight
// changes are summed in LDS (a k-entry histogram) and flushed with one global
// atomic per block and touched part.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_move_delta(const int32_t *__restrict__ ids, int64_t cnt, int32_t k,
                                                       const eid_t *__restrict__ xadj,
                                                       const vid_t *__restrict__ adjncy,
                                                       const wgt_t *__restrict__ adjwgt,
                                                       const wgt_t *__restrict__ vwgt,
                                                       const int32_t *__restrict__ where,
                                                       const int32_t *__restrict__ dest, JetCtl *ctl,
                                                       int64_t *__restrict__ pwgts)
{
    extern __shared__ unsigned long long pdelta[];
    __shared__ unsigned long long bdelta;
    if (ctl->done) return;
    for (int i = threadIdx.x; i < k; i += kRowBS) pdelta[i] = 0;
    if (threadIdx.x == 0) bdelta = 0;
    __syncthreads();
    constexpr int NG = kRowBS / G;
    const int g = threadIdx.x / G, lane = glane<G>();
    for (int64_t base = (int64_t)blockIdx.x * NG; base < cnt; base += (int64_t)gridDim.x * NG) {
        int64_t gi = base + g;
        int32_t v = gi < cnt ? ids[gi] : 0;
        int32_t pv = gi < cnt ? where[v] : 0, dv = gi < cnt ? dest[v] : 0;
        int64_t d2 = 0;
        if (pv != dv) {
            for (eid_t j = xadj[v] + lane; j < xadj[v + 1]; j += G) {
                int32_t u = adjncy[j];
                int32_t pu = where[u], du = dest[u];
                bool um = pu != du;
                int32_t before = pv != pu, after = dv != (um ? du : pu);
                d2 += (int64_t)(after - before) * adjwgt[j] * (um ? 1 : 2);
            }
        }
        d2 = group_sum<G>(d2);
        if (lane == 0 && pv != dv) {
            if (d2) atomicAdd(&bdelta, (unsigned long long)d2);
            atomicAdd(&pdelta[pv], (unsigned long long)(-(int64_t)vwgt[v]));
            atomicAdd(&pdelta[dv], (unsigned long long)(int64_t)vwgt[v]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += kRowBS)
        if (pdelta[i]) atomicAdd((unsigned long long *)&pwgts[i], pdelta[i]);
    if (threadIdx.x == 0 && bdelta) atomicAdd(&ctl->delta2, bdelta);
}

__global__ void k_move_commit(int32_t n, const JetCtl *ctl, const int32_t *__restrict__ dest, int32_t *__restrict__ where)
{
    if (ctl->done) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) where[v] = dest[v];
}

// Best-partition bookkeeping (:4805-4839) on the heaviest part weight.
__global__ void __launch_bounds__(kBS) k_jet_end(JetCtl *ctl, int32_t k, const int64_t *__restrict__ pwgts)
{
    __shared__ int64_t smax[kBS];
    JetCtl &C = *ctl;
    if (C.done) return;
    int64_t mx = INT64_MIN;
    for (int q = threadIdx.x; q < k; q += kBS) mx = pwgts[q] > mx ? pwgts[q] : mx;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && smax[threadIdx.x + s] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    int64_t curmax = smax[0];
    C.cut += (int64_t)C.delta2 / 2;
    C.iters++;
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

__global__ void k_copy_best(int32_t n, const JetCtl *ctl, const int32_t *__restrict__ where, int32_t *__restrict__ best)
{
    if (ctl->done || !ctl->copy_best) return;
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

template <int MaxG = 256, class F>
void for_bins(const RowBins &b, F f)
{
    if (b.count(0)) f(std::integral_constant<int, 8>{}, b.list(0), b.count(0));
    if (b.count(1)) f(std::integral_constant<int, 64>{}, b.list(1), b.count(1));
    if (b.count(2)) f(std::integral_constant<int, MaxG < 256 ? MaxG : 256>{}, b.list(2), b.count(2));
}

__global__ void k_half(int64_t *d) { *d /= 2; }

__global__ void k_pwgts(int32_t n, const wgt_t *__restrict__ vwgt, const int32_t *__restrict__ where,
                        unsigned long long *pw)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) atomicAdd(&pw[where[v]], (unsigned long long)(int64_t)vwgt[v]);
}

__global__ void k_project(int32_t n, const int32_t *__restrict__ cmap, const int32_t *__restrict__ wc,
                          int32_t *__restrict__ wf)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) wf[v] = wc[cmap[v]];
}

void launch_select(Ctx &c, const DevGraph &g, int32_t k, const int32_t *where, const int64_t *pwgts,
                   const uint8_t *lock, const int32_t *under, JetCtl *ctl, int32_t *dest, int32_t *pregain,
                   uint32_t *key)
{
    const RowBins &b = g.bins;
    SelArgs A{k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, pwgts, lock, under, ctl, dest, pregain, key};
    if (b.count(0)) {
        k_jet_select_s<<<grid_groups(b.count(0), 8), kRowBS, 0, c.stream>>>(b.list(0), b.count(0), A);
        HGP_LAUNCH_CHECK();
    }
    if (b.count(1)) {
        int64_t blocks = std::min<int64_t>((b.count(1) + 3) / 4, 4096);
        k_jet_select_w<<<(unsigned)blocks, kRowBS, sizeof(int32_t) * 4 * (size_t)k, c.stream>>>(b.list(1),
                                                                                               b.count(1), A);
        HGP_LAUNCH_CHECK();
    }
    if (b.count(2)) {
        int64_t blocks = std::min<int64_t>(b.count(2), 4096);
        k_jet_select_b<<<(unsigned)blocks, kRowBS, sizeof(int32_t) * (size_t)k, c.stream>>>(b.list(2), b.count(2),
                                                                                           A);
        HGP_LAUNCH_CHECK();
    }
}

}  // namespace

void edge_cut(Ctx &c, const DevGraph &g, const int32_t *where, int64_t *d_cut)
{
    fill<int64_t>(c, d_cut, 1, 0);
    if (g.n <= 0) return;
    for_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
        constexpr int G = decltype(gc)::value;
        k_cut<G><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt, where,
                                                             (unsigned long long *)d_cut);
        HGP_LAUNCH_CHECK();
    });
    k_half<<<1, 1, 0, c.stream>>>(d_cut);
    HGP_LAUNCH_CHECK();
}

void part_weights(Ctx &c, int32_t n, const wgt_t *vwgt, const int32_t *where, int32_t k, int64_t *pwgts)
{
    fill<int64_t>(c, pwgts, k, 0);
    if (n <= 0) return;
    k_pwgts<<<grid_for(n, kBS), kBS, 0, c.stream>>>(n, vwgt, where, (unsigned long long *)pwgts);
    HGP_LAUNCH_CHECK();
}

void project(Ctx &c, int32_t n, const int32_t *cmap, const int32_t *where_c, int32_t *where_f)
{
    if (n <= 0) return;
    k_project<<<grid_for(n, kBS), kBS, 0, c.stream>>>(n, cmap, where_c, where_f);
    HGP_LAUNCH_CHECK();
}

// hunyuangraph_malloc_krefine (hunyuangraph_GPU_uncoarsen.h:428-508) + k_refine.
RefineStats refine_level(Ctx &c, const DevGraph &g, int32_t k, int64_t tvwgt, bool finest, const Params &p,
                         int32_t *where, int64_t *d_cut)
{
    RefineStats rs;
    int32_t n = g.n;
    if (n <= 0 || k <= 1) return rs;
    if (k > 4096) throw Error{-4, "refine: k > 4096 is not supported"};
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
    int64_t *d_ne = c.arena.ralloc<int64_t>(1);
    fill<uint8_t>(c, lock, n, 0);
    HGP_CHECK_HIP(hipMemcpyAsync(best, where, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    part_weights(c, n, g.vwgt, where, k, pwgts);
    k_jet_init<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts, d_cut, maxwgt, opt, p.patience, finest ? 1 : 3);
    HGP_LAUNCH_CHECK();
    unsigned gv = grid_for(n, kBS);
    for (;;) {
        k_jet_begin<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts, under);
        HGP_LAUNCH_CHECK();
        read_scalars(c, ctl, sizeof(JetCtl));
        JetCtl h = *(JetCtl *)c.pinned;
        if (h.done) break;
        {
            // LP: every row (adjncy, adjwgt, where gather) + xadj, where, lock, dest, pregain per vertex.
            // RW/RS: the same per-vertex traffic; rows only of overloaded parts (not counted).
            bool lp = h.mode == kModeLP;
            KScope ks(c, lp ? "jet_select_lp" : "jet_select_rb", (lp ? 12.0 * (double)g.m : 0.0) + 21.0 * n);
            launch_select(c, g, k, where, pwgts, lock, under, ctl, dest, pregain, key);
        }
        if (h.mode == kModeLP) {
            {
                KScope ks(c, "jet_afterburner", 5.0 * n);
                for_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
                    constexpr int G = decltype(gc)::value;
                    k_afterburner<G><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy,
                                                                                 g.adjwgt, where, dest, pregain,
                                                                                 ctl, sel);
                    HGP_LAUNCH_CHECK();
                });
            }
            k_lp_finish<<<gv, kBS, 0, c.stream>>>(n, ctl, sel, where, dest, lock);
            HGP_LAUNCH_CHECK();
        } else {
            k_iota_u32<<<gv, kBS, 0, c.stream>>>(n, val);
            HGP_LAUNCH_CHECK();
            radix_sort(c, key, val, n, 0, bits_for((uint64_t)k * kMaxBuckets));
            k_rb_weights<<<gv, kBS, 0, c.stream>>>(n, k, key, val, g.vwgt, wsc);
            HGP_LAUNCH_CHECK();
            scan_inclusive(c, wsc, wsc, n);
            k_rb_pstart<<<gv, kBS, 0, c.stream>>>(n, k, key, pstart);
            HGP_LAUNCH_CHECK();
            k_rb_evict<<<gv, kBS, 0, c.stream>>>(n, k, key, val, wsc, g.vwgt, pstart, pwgts, ctl, where, dest, sel);
            HGP_LAUNCH_CHECK();
            if (h.mode == kModeRS) {
                compact_indices(c, sel, n, elist, d_ne);
                k_gather_w<<<gv, kBS, 0, c.stream>>>(d_ne, elist, val, g.vwgt, bs);
                HGP_LAUNCH_CHECK();
                scan_inclusive(c, bs + 1, bs + 1, n);
                k_rs_cookie<<<1, 64, 0, c.stream>>>(k, bs, d_ne, pwgts, ctl, es, ee);
                HGP_LAUNCH_CHECK();
                k_rs_assign<<<gv, kBS, 0, c.stream>>>(k, d_ne, elist, val, es, ee, where, dest);
                HGP_LAUNCH_CHECK();
            }
        }
        {
            KScope ks(c, "jet_move_delta", 8.0 * n);
            for_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
                constexpr int G = decltype(gc)::value;
                int64_t blocks = std::min<int64_t>((nb + kRowBS / G - 1) / (kRowBS / G), 2048);
                k_move_delta<G><<<(unsigned)blocks, kRowBS, sizeof(unsigned long long) * (size_t)k, c.stream>>>(
                    ids, nb, k, g.xadj, g.adjncy, g.adjwgt, g.vwgt, where, dest, ctl, pwgts);
                HGP_LAUNCH_CHECK();
            });
        }
        k_move_commit<<<gv, kBS, 0, c.stream>>>(n, ctl, dest, where);
        HGP_LAUNCH_CHECK();
        k_jet_end<<<1, kBS, 0, c.stream>>>(ctl, k, pwgts);
        HGP_LAUNCH_CHECK();
        k_copy_best<<<gv, kBS, 0, c.stream>>>(n, ctl, where, best);
        HGP_LAUNCH_CHECK();
    }
    rs.iters = ((JetCtl *)c.pinned)->iters;
    HGP_CHECK_HIP(hipMemcpyAsync(where, best, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    k_jet_fini<<<1, 1, 0, c.stream>>>(ctl, d_cut);
    HGP_LAUNCH_CHECK();
    c.arena.right_release(mark);
    return rs;
}

}  // namespace hgp
