// hgp_lp.hip -- the older label-propagation refiners the reference keeps
// beside Jet: the k-way rounds of hunyuangraph_k_refinement_me
// (hunyuangraph_GPU_LP.h:1245-1374) and the 2-way pass of jetLP (:673-781).
//
// k-way.  A round moves, into every part, the unmoved vertex of highest
// positive gain whose best part it is (at most k moves a round); then
// rebalancing rounds move, into every underweight part, the vertex of an
// overweight part with the most edge weight to it.  The reference recomputes
// every vertex's best part in every round, O(m) work for at most k moves.
// Here the candidates (best part, gain) of all vertices stay in HBM and a
// round recomputes only the vertices it moved and their neighbours (listed,
// without duplicates, by the move kernel); a rebalancing round after which a
// part crossed a bound recomputes every vertex.  Rows longer than a threshold
// keep their edge weight per part in an HBM row that the moves update with
// atomics, so a hub's candidate is an O(k) scan instead of a walk of its row.
// A rebalancing round without a move leaves nothing changed, so every later
// round would repeat it: the loop then ends and counts the rounds left.
// Per destination part the winner is the largest (gain, -id) key of one
// 64-bit atomicMax, so the result does not depend on scheduling.
//
// 2-way.  A synchronous flip pass, then, when a side is outside its bounds,
// a transfer of the best-gain prefix of the heavier side (stable radix sort
// by gain, inclusive scan of the weights).
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"

#include <algorithm>

namespace hgp {

namespace {

constexpr int kLpBS = 256;
constexpr int kLpWaves = kLpBS / kWave;
constexpr int kLpTrackBins = 16;          // tracking thresholds 64 << i entries
constexpr int64_t kLpTrackMin = 64;
constexpr size_t kLpTrackCap = (size_t)1 << 30;   // bytes of tracked rows at most

struct LpCtl {
    int32_t done, iters1, done2, iters2;
    int32_t moves[2];                     // moves of the rounds by parity
    int64_t list_n;                       // vertices listed for recomputation
    int32_t full;                         // rebalancing: the bounds changed, recompute every vertex
    int32_t pad;
};

struct LpArgs {
    int32_t n, k;
    const eid_t *xadj;
    const vid_t *adjncy;
    const wgt_t *adjwgt;
    const wgt_t *vwgt;
    int32_t *where;
    int64_t *pw;
    const int32_t *track;                 // row of T, -1 if the vertex walks its row
    int32_t *T;                           // tracked rows: k ints of edge weight per part
    uint32_t *flag;                       // bit 0 moved, bit 1 listed
    int32_t *cgain, *cdest;               // candidate: gain (0: none) and destination
    uint8_t *over, *thin;                 // rebalancing: part above / below its bound
    unsigned long long *best;             // per destination: (gain << 32) | ~id of the winner
    int32_t *list;
    LpCtl *ctl;
};

__device__ inline unsigned long long lp_key(int32_t val, int32_t id)
{
    return ((unsigned long long)(uint32_t)val << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)id);
}

__device__ inline int32_t lp_key_id(unsigned long long key) { return (int32_t)(0xFFFFFFFFu - (uint32_t)key); }

// Counts the rows longer than 64 << i entries, i < kLpTrackBins.
__global__ void __launch_bounds__(kLpBS) k_lp_degcount(int32_t n, const eid_t *__restrict__ xadj,
                                                       unsigned long long *__restrict__ cnt)
{
    __shared__ unsigned int sh[kLpTrackBins];
    if (threadIdx.x < kLpTrackBins) sh[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = xadj[v + 1] - xadj[v];
        for (int i = 0; i < kLpTrackBins && d > (kLpTrackMin << i); i++) atomicAdd(&sh[i], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kLpTrackBins && sh[threadIdx.x])
        atomicAdd(&cnt[threadIdx.x], (unsigned long long)sh[threadIdx.x]);
}

__global__ void k_lp_longrows(int32_t n, const eid_t *__restrict__ xadj, int64_t lim, uint8_t *__restrict__ mark)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) mark[v] = xadj[v + 1] - xadj[v] > lim;
}

__global__ void k_lp_slots(int64_t nt, const int32_t *__restrict__ idx, int32_t *__restrict__ track)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nt) track[idx[s]] = (int32_t)s;
}

// Edge weight per part of the tracked rows (a block per row, k ints of LDS).
__global__ void __launch_bounds__(kLpBS) k_lp_tbuild(const int32_t *__restrict__ idx, int64_t nt, LpArgs A)
{
    extern __shared__ int32_t acc[];
    const int k = A.k;
    for (int64_t s = blockIdx.x; s < nt; s += gridDim.x) {
        for (int p = threadIdx.x; p < k; p += kLpBS) acc[p] = 0;
        __syncthreads();
        const int32_t v = idx[s];
        for (eid_t j = A.xadj[v] + threadIdx.x; j < A.xadj[v + 1]; j += kLpBS)
            atomicAdd(&acc[A.where[A.adjncy[j]]], A.adjwgt[j]);
        __syncthreads();
        int32_t *row = A.T + s * (int64_t)k;
        for (int p = threadIdx.x; p < k; p += kLpBS) row[p] = acc[p];
        __syncthreads();
    }
}

// Candidates of the listed vertices (LIST: the count is the control block's
// list_n, and each vertex leaves the list; else every vertex, in PH 2 only
// after a round that changed the bounds).  PH 1: an unmoved vertex
// whose best other part (largest edge weight, lowest id among equal ones)
// beats its own part by gain > 0.  PH 2: a vertex of an overweight part with
// edge weight to an underweight part other than its own; the best such part,
// gain = that weight.  A wave takes 64 vertices and works through the ones
// with work together: a tracked row scans its k-int row, another row adds its
// entries into the wave's k-int LDS row.
template <int PH, bool LIST>
__global__ void __launch_bounds__(kLpBS) k_lp_cand(const int32_t *__restrict__ ids, int64_t cnt, LpArgs A)
{
    extern __shared__ int32_t acc[];
    const LpCtl &C = *A.ctl;
    if (PH == 1 ? C.done : C.done2) return;
    if (PH == 2 && !LIST && !C.full) return;
    if (LIST) cnt = C.list_n;
    const int k = A.k, lane = lane_id(), wv = threadIdx.x / kWave;
    int32_t *my = acc + (size_t)wv * k;
    for (int i = lane; i < k; i += kWave) my[i] = 0;
    wave_lds_sync();
    for (int64_t base = ((int64_t)blockIdx.x * kLpWaves + wv) * kWave; base < cnt;
         base += (int64_t)gridDim.x * kLpWaves * kWave) {
        const int64_t i = base + lane;
        const int32_t v = i < cnt ? (LIST ? ids[i] : (int32_t)i) : -1;
        bool work = false;
        if (v >= 0) {
            uint32_t f = 0;
            if (LIST || PH == 1) f = A.flag[v];
            if (LIST) A.flag[v] = f & ~2u;
            if (PH == 1) {
                work = !(f & 1u);
            } else {
                work = A.over[A.where[v]] != 0;
                if (!work) A.cgain[v] = 0;
            }
        }
        uint64_t todo = __ballot(work);
        while (todo) {
            const int l = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const int32_t u = __shfl(v, l, kWave);
            const int32_t me = A.where[u];
            const int32_t slot = A.track[u];
            unsigned long long bkey = 0;
            int32_t cme = 0;
            if (slot >= 0) {
                const int32_t *row = A.T + (int64_t)slot * k;
                for (int p = lane; p < k; p += kWave) {
                    const int32_t val = row[p];
                    const bool el = p != me && (PH == 1 || A.thin[p]);
                    if (el && val > 0) {
                        const unsigned long long key = lp_key(val, p);
                        bkey = key > bkey ? key : bkey;
                    }
                }
                if (PH == 1) cme = row[me];
            } else if (A.xadj[u + 1] - A.xadj[u] <= kWave) {
                // one entry per lane, kept in registers over the three passes
                const eid_t j = A.xadj[u] + lane;
                const bool live = j < A.xadj[u + 1];
                int32_t q = -1, w = 0;
                if (live) {
                    q = A.where[A.adjncy[j]];
                    w = A.adjwgt[j];
                }
                if (live && (PH == 1 || (q != me && A.thin[q]))) atomicAdd(&my[q], w);
                wave_lds_sync();
                if (live && q != me && (PH == 1 || A.thin[q])) {
                    const int32_t val = my[q];
                    if (val > 0) bkey = lp_key(val, q);
                }
                if (PH == 1) cme = my[me];
                wave_lds_sync();
                if (live) my[q] = 0;
                wave_lds_sync();
            } else {
                const eid_t b = A.xadj[u], e = A.xadj[u + 1];
                for (eid_t j = b + lane; j < e; j += kWave) {
                    const int32_t q = A.where[A.adjncy[j]];
                    if (PH == 1 || (q != me && A.thin[q])) atomicAdd(&my[q], A.adjwgt[j]);
                }
                wave_lds_sync();
                for (eid_t j = b + lane; j < e; j += kWave) {
                    const int32_t q = A.where[A.adjncy[j]];
                    if (q != me && (PH == 1 || A.thin[q])) {
                        const int32_t val = my[q];
                        if (val > 0) {
                            const unsigned long long key = lp_key(val, q);
                            bkey = key > bkey ? key : bkey;
                        }
                    }
                }
                if (PH == 1) cme = my[me];
                wave_lds_sync();
                for (eid_t j = b + lane; j < e; j += kWave) my[A.where[A.adjncy[j]]] = 0;
                wave_lds_sync();
            }
            bkey = wave_max(bkey);
            if (lane == 0) {
                const int32_t bval = (int32_t)(bkey >> 32);
                int32_t gain = 0;
                if (bkey) gain = PH == 1 ? (bval > cme ? bval - cme : 0) : bval;
                A.cgain[u] = gain;
                if (gain > 0) A.cdest[u] = lp_key_id(bkey);
            }
        }
    }
}

// Per destination part, the candidate of largest (gain, -id): a block folds
// its vertices into k LDS keys, then one atomicMax per part and block.  In
// PH 1 it also ends the loop after a round without moves and resets the
// round's counters.
template <int PH>
__global__ void __launch_bounds__(kLpBS) k_lp_pick(int32_t it, LpArgs A)
{
    extern __shared__ unsigned long long sb[];
    LpCtl &C = *A.ctl;
    if (PH == 1) {
        if (C.done) return;
        if (it > 0 && C.moves[(it + 1) & 1] == 0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) C.done = 1;
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            C.moves[it & 1] = 0;
            C.list_n = 0;
        }
    } else {
        if (C.done2) return;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            C.moves[it & 1] = 0;
            C.list_n = 0;
        }
    }
    const int k = A.k;
    for (int p = threadIdx.x; p < k; p += kLpBS) sb[p] = 0;
    __syncthreads();
    for (int64_t v = (int64_t)blockIdx.x * kLpBS + threadIdx.x; v < A.n; v += (int64_t)gridDim.x * kLpBS) {
        const int32_t g = A.cgain[v];
        if (g > 0) atomicMax(&sb[A.cdest[v]], lp_key(g, (int32_t)v));
    }
    __syncthreads();
    for (int p = threadIdx.x; p < k; p += kLpBS)
        if (sb[p]) atomicMax(&A.best[p], sb[p]);
}

// calculate_move (hunyuangraph_GPU_LP.h:990-1006): block p moves its winner
// into part p, then walks the winner's row: tracked neighbours update their
// part rows, and every neighbour not yet listed is listed once (in PH 2 the
// winner too: its candidate changes with its part).
template <int PH>
__global__ void __launch_bounds__(kLpBS) k_lp_apply(int32_t it, LpArgs A)
{
    LpCtl &C = *A.ctl;
    if (PH == 1 ? C.done : C.done2) return;
    const int32_t p = blockIdx.x;
    const unsigned long long key = A.best[p];
    if (key == 0) return;
    const int32_t v = lp_key_id(key);
    const int32_t a = A.where[v];
    __syncthreads();   // every thread has read best[p] and where[v]
    if (threadIdx.x == 0) {
        const int32_t vw = A.vwgt[v];
        A.best[p] = 0;
        A.where[v] = p;
        A.cgain[v] = 0;
        atomicAdd((unsigned long long *)&A.pw[a], (unsigned long long)(-(int64_t)vw));
        atomicAdd((unsigned long long *)&A.pw[p], (unsigned long long)(int64_t)vw);
        if (PH == 1) {
            atomicOr(&A.flag[v], 1u);
            if (atomicAdd(&C.moves[it & 1], 1) == 0) atomicAdd(&C.iters1, 1);
        } else {
            atomicAdd(&C.moves[it & 1], 1);
            if (!(atomicOr(&A.flag[v], 2u) & 2u)) {
                const unsigned long long pos = atomicAdd((unsigned long long *)&C.list_n, 1ull);
                A.list[pos] = v;
            }
        }
    }
    const int k = A.k, lane = lane_id();
    for (eid_t j = A.xadj[v] + threadIdx.x; j < A.xadj[v + 1]; j += kLpBS) {
        const int32_t u = A.adjncy[j], w = A.adjwgt[j];
        const int32_t s = A.track[u];
        if (s >= 0) {
            int32_t *row = A.T + (int64_t)s * k;
            atomicAdd(&row[a], -w);
            atomicAdd(&row[p], w);
        }
        {
            const bool add = !(atomicOr(&A.flag[u], 2u) & 2u);
            const uint64_t m = __ballot(add);
            if (m) {
                const int lead = __ffsll((long long)m) - 1;
                unsigned long long base = 0;
                if (lane == lead) base = atomicAdd((unsigned long long *)&C.list_n, (unsigned long long)__popcll(m));
                base = __shfl(base, lead, kWave);
                if (add) A.list[base + __popcll(m & lanemask_lt())] = u;
            }
        }
    }
}

// calculate_overweight (hunyuangraph_GPU_LP.h:1008-1026) and the loop test
// of the rebalancing rounds: part weights compared as floats with the float
// bounds; the loop ends when no part is over or none is under.  After a
// round without moves nothing has changed, so the rounds up to the
// threshold would all repeat it: they are counted and the loop ends.  full
// is set when a part's flags changed (or in the first round).
__global__ void __launch_bounds__(kLpBS) k_lp_flags(int32_t it, int32_t thr, LpArgs A, float fmin, float fmax)
{
    LpCtl &C = *A.ctl;
    if (C.done2) return;
    if (it > 0 && C.moves[(it + 1) & 1] == 0) {
        if (threadIdx.x == 0) {
            C.iters2 = thr;
            C.done2 = 1;
        }
        return;
    }
    __shared__ int nover, nthin, changed;
    if (threadIdx.x == 0) nover = nthin = changed = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < A.k; p += kLpBS) {
        const float val = (float)(int32_t)A.pw[p];
        const uint8_t o = val > fmax, t = val < fmin;
        if (o != A.over[p] || t != A.thin[p]) changed = 1;
        A.over[p] = o;
        A.thin[p] = t;
        if (o) atomicAdd(&nover, 1);
        if (t) atomicAdd(&nthin, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        C.full = it == 0 || changed;
        if (nover > 0 && nthin > 0) C.iters2++;
        else C.done2 = 1;
    }
}

// ---------------------------------------------------------------- 2-way

// Edge weight of each vertex's row into its own side (same) and the other
// side; a wave takes 64 vertices, a row of > 64 entries is summed by the
// whole wave, the others by their own lane.  F: flip pass (nw = the side
// after the pass), else the transfer keys (key = other - same on side
// `from`, -m elsewhere, stored as INT32_MAX - key so that an ascending sort
// puts the largest key first).
template <bool F>
__global__ void __launch_bounds__(kLpBS) k_lp2_rows(int32_t n, const eid_t *__restrict__ xadj,
                                                    const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ where, int32_t from,
                                                    int32_t *__restrict__ nw, uint32_t *__restrict__ key,
                                                    uint32_t *__restrict__ val)
{
    const int lane = lane_id();
    const int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) - lane;
    const int64_t i = base + lane;
    const int32_t v = i < n ? (int32_t)i : -1;
    int64_t same = 0, other = 0;
    eid_t b = 0, e = 0;
    int32_t me = 0;
    if (v >= 0) {
        b = xadj[v];
        e = xadj[v + 1];
        me = where[v];
    }
    const bool lng = v >= 0 && e - b > kWave;
    if (v >= 0 && !lng) {
        for (eid_t j = b; j < e; j++) {
            if (where[adjncy[j]] == me) same += adjwgt[j];
            else other += adjwgt[j];
        }
    }
    uint64_t todo = __ballot(lng);
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const eid_t rb = __shfl(b, l, kWave), re = __shfl(e, l, kWave);
        const int32_t rp = __shfl(me, l, kWave);
        int64_t s = 0, o = 0;
        for (eid_t j = rb + lane; j < re; j += kWave) {
            if (where[adjncy[j]] == rp) s += adjwgt[j];
            else o += adjwgt[j];
        }
        s = wave_sum(s);
        o = wave_sum(o);
        if (lane == l) {
            same = s;
            other = o;
        }
    }
    if (v < 0) return;
    if (F) {
        nw[v] = same <= other ? 1 - me : me;
    } else {
        const int64_t k64 = me == from ? other - same : -(int64_t)xadj[n];
        key[v] = (uint32_t)((int64_t)INT32_MAX - k64);
        val[v] = (uint32_t)v;
    }
}

__global__ void k_lp2_weights(int32_t n, const uint32_t *__restrict__ order, const wgt_t *__restrict__ vwgt,
                              int64_t *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = vwgt[order[i]];
}

__global__ void k_lp2_move(int32_t n, const uint32_t *__restrict__ order, const int64_t *__restrict__ prefix,
                           int64_t num, int32_t to, int32_t *__restrict__ where)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && prefix[i] <= num) where[order[i]] = to;
}

// hunyuangraph_compute_log2 (hunyuangraph_common.h:26-31).
int32_t ref_log2(int32_t a)
{
    int32_t i = 1;
    for (; a > 1; i++) a >>= 1;
    return i - 1;
}

}  // namespace

int32_t lp_threshold(int32_t n, int32_t k)
{
    // hunyuangraph_GPU_LP.h:1251 (which divides by zero for one part: no rounds then)
    if (k < 2) return 0;
    return std::max(n / (6000 * ref_log2(k)), 8 * k);
}

void lp_part_bounds(int32_t k, int64_t tvwgt, float *fmin, float *fmax)
{
    // calculate_overweight (hunyuangraph_GPU_LP.h:1008-1026) with
    // tpwgts[0] = (float)(1.0 / nparts) (hunyuangraph_admin.h:21-23)
    const float temp = (float)(1.0 / k);
    const int32_t tv = (int32_t)tvwgt;
    *fmin = (float)(temp / 1.03 * tv);
    *fmax = (float)(temp * 1.03 * tv);
}

RefineStats lp_refine_level(Ctx &c, const DevGraph &g, int32_t k, int64_t tvwgt, int32_t *where, int64_t *d_cut)
{
    RefineStats rs;
    const int32_t n = g.n;
    if (k > 4096) throw Error{-4, "lp refine: k > 4096 is not supported"};
    if (n == 0 || k < 2) {
        edge_cut(c, g, where, d_cut);
        return rs;
    }
    const size_t mark = c.arena.right_mark();
    LpArgs A{};
    A.n = n;
    A.k = k;
    A.xadj = g.xadj;
    A.adjncy = g.adjncy;
    A.adjwgt = g.adjwgt;
    A.vwgt = g.vwgt;
    A.where = where;
    A.pw = c.arena.ralloc<int64_t>((size_t)k);
    A.best = c.arena.ralloc<unsigned long long>((size_t)k);
    A.over = c.arena.ralloc<uint8_t>((size_t)k);
    A.thin = c.arena.ralloc<uint8_t>((size_t)k);
    A.flag = c.arena.ralloc<uint32_t>((size_t)n);
    A.cgain = c.arena.ralloc<int32_t>((size_t)n);
    A.cdest = c.arena.ralloc<int32_t>((size_t)n);
    A.list = c.arena.ralloc<int32_t>((size_t)n);
    int32_t *track = c.arena.ralloc<int32_t>((size_t)n);
    A.track = track;
    A.ctl = c.arena.ralloc<LpCtl>(1);
    unsigned long long *cnt = c.arena.ralloc<unsigned long long>(kLpTrackBins);
    HGP_CHECK_HIP(hipMemsetAsync(A.best, 0, sizeof(unsigned long long) * (size_t)k, c.stream));
    HGP_CHECK_HIP(hipMemsetAsync(A.flag, 0, sizeof(uint32_t) * (size_t)n, c.stream));
    HGP_CHECK_HIP(hipMemsetAsync(A.cgain, 0, sizeof(int32_t) * (size_t)n, c.stream));
    HGP_CHECK_HIP(hipMemsetAsync(A.ctl, 0, sizeof(LpCtl), c.stream));
    HGP_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * kLpTrackBins, c.stream));
    fill<int32_t>(c, track, n, -1);
    part_weights(c, n, g.vwgt, where, k, A.pw);

    // Rows to track: the longest ones whose k-int rows fit the budget.
    k_lp_degcount<<<std::min<unsigned>(grid_for(n, kLpBS), 1024), kLpBS, 0, c.stream>>>(n, g.xadj, cnt);
    HGP_LAUNCH_CHECK();
    read_scalars(c, cnt, sizeof(unsigned long long) * kLpTrackBins);
    std::vector<int64_t> cnts(c.pinned, c.pinned + kLpTrackBins);
    const size_t room = c.arena.right_mark() - c.arena.left_mark();
    const size_t budget = std::min(kLpTrackCap, room / 2);
    int64_t lim = -1, nt = 0;
    for (int i = 0; i < kLpTrackBins; i++) {
        if ((size_t)cnts[i] * (size_t)k * sizeof(int32_t) + (size_t)cnts[i] * 8 <= budget) {
            lim = kLpTrackMin << i;
            nt = cnts[i];
            break;
        }
    }
    A.T = nullptr;
    if (nt > 0) {
        uint8_t *lm = c.arena.ralloc<uint8_t>((size_t)n);
        int32_t *idx = c.arena.ralloc<int32_t>((size_t)nt);
        int64_t *dcount = c.arena.ralloc<int64_t>(1);
        A.T = c.arena.ralloc<int32_t>((size_t)nt * (size_t)k);
        k_lp_longrows<<<grid_for(n, kLpBS), kLpBS, 0, c.stream>>>(n, g.xadj, lim, lm);
        HGP_LAUNCH_CHECK();
        compact_indices(c, lm, n, idx, dcount);
        k_lp_slots<<<grid_for(nt, kLpBS), kLpBS, 0, c.stream>>>(nt, idx, track);
        HGP_LAUNCH_CHECK();
        k_lp_tbuild<<<(unsigned)std::min<int64_t>(nt, 4096), kLpBS, sizeof(int32_t) * (size_t)k, c.stream>>>(idx, nt,
                                                                                                             A);
        HGP_LAUNCH_CHECK();
    }

    const int32_t thr = lp_threshold(n, k);
    const size_t lds_cand = sizeof(int32_t) * (size_t)k * kLpWaves;
    const size_t lds_pick = sizeof(unsigned long long) * (size_t)k;
    const unsigned g_all = std::min<unsigned>(grid_for(n, kLpBS), 2048);
    const unsigned g_list = std::min<unsigned>(grid_for(n, kLpBS), 2048);
    constexpr int kChunk = 16;

    // first loop (hunyuangraph_GPU_LP.h:1283-1308)
    k_lp_cand<1, false><<<g_all, kLpBS, lds_cand, c.stream>>>(nullptr, n, A);
    HGP_LAUNCH_CHECK();
    for (int32_t it = 0; it < thr; it++) {
        k_lp_pick<1><<<g_all, kLpBS, lds_pick, c.stream>>>(it, A);
        k_lp_apply<1><<<(unsigned)k, kLpBS, 0, c.stream>>>(it, A);
        k_lp_cand<1, true><<<g_list, kLpBS, lds_cand, c.stream>>>(A.list, 0, A);
        HGP_LAUNCH_CHECK();
        if ((it + 1) % kChunk == 0 && it + 1 < thr) {
            read_scalars(c, A.ctl, sizeof(LpCtl));
            if (((const LpCtl *)c.pinned)->done) break;
        }
    }

    // rebalancing loop (:1313-1364)
    float fmin, fmax;
    lp_part_bounds(k, tvwgt, &fmin, &fmax);
    for (int32_t it = 0; it < thr; it++) {
        k_lp_flags<<<1, kLpBS, 0, c.stream>>>(it, thr, A, fmin, fmax);
        k_lp_cand<2, true><<<g_list, kLpBS, lds_cand, c.stream>>>(A.list, 0, A);
        k_lp_cand<2, false><<<g_all, kLpBS, lds_cand, c.stream>>>(nullptr, n, A);
        k_lp_pick<2><<<g_all, kLpBS, lds_pick, c.stream>>>(it, A);
        k_lp_apply<2><<<(unsigned)k, kLpBS, 0, c.stream>>>(it, A);
        HGP_LAUNCH_CHECK();
        if ((it + 1) % kChunk == 0 || it + 1 == thr) {
            read_scalars(c, A.ctl, sizeof(LpCtl));
            if (((const LpCtl *)c.pinned)->done2) break;
        }
    }
    read_scalars(c, A.ctl, sizeof(LpCtl));
    const LpCtl hc = *(const LpCtl *)c.pinned;
    rs.iters = hc.iters1 + hc.iters2;
    c.arena.right_release(mark);
    edge_cut(c, g, where, d_cut);
    return rs;
}

bool lp2_refine(Ctx &c, const DevGraph &g, int32_t *where)
{
    const int32_t n = g.n;
    if (n == 0) return false;
    const size_t mark = c.arena.right_mark();
    int32_t *nw = c.arena.ralloc<int32_t>((size_t)n);
    int64_t *pw = c.arena.ralloc<int64_t>(2);
    k_lp2_rows<true><<<grid_for(n, kLpBS), kLpBS, 0, c.stream>>>(n, g.xadj, g.adjncy, g.adjwgt, where, 0, nw,
                                                                 nullptr, nullptr);
    HGP_LAUNCH_CHECK();
    HGP_CHECK_HIP(hipMemcpyAsync(where, nw, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    part_weights(c, n, g.vwgt, where, 2, pw);
    read_scalars(c, pw, sizeof(int64_t) * 2);
    const int64_t p0 = c.pinned[0], p1 = c.pinned[1];
    // jetLP (hunyuangraph_GPU_LP.h:724-728): both sides within
    // [tvwgt * 0.5 / 1.03, tvwgt * 0.5 * 1.03], as doubles
    const int32_t tv = (int32_t)(p0 + p1);
    const double lo = tv * 0.5 / 1.03, hi = tv * 0.5 * 1.03;
    bool ran = false;
    if (!((double)p0 >= lo && (double)p0 <= hi && (double)p1 >= lo && (double)p1 <= hi)) {
        const int32_t from = p0 > p1 ? 0 : 1, to = 1 - from;
        const int64_t num = ((from == 0 ? p0 : p1) - (from == 0 ? p1 : p0)) / 2;
        uint32_t *key = c.arena.ralloc<uint32_t>((size_t)n);
        uint32_t *val = c.arena.ralloc<uint32_t>((size_t)n);
        int64_t *prefix = c.arena.ralloc<int64_t>((size_t)n);
        k_lp2_rows<false><<<grid_for(n, kLpBS), kLpBS, 0, c.stream>>>(n, g.xadj, g.adjncy, g.adjwgt, where, from,
                                                                      nullptr, key, val);
        HGP_LAUNCH_CHECK();
        radix_sort(c, key, val, n, 0, 32);
        k_lp2_weights<<<grid_for(n, kLpBS), kLpBS, 0, c.stream>>>(n, val, g.vwgt, prefix);
        HGP_LAUNCH_CHECK();
        scan_inclusive(c, prefix, prefix, n);
        k_lp2_move<<<grid_for(n, kLpBS), kLpBS, 0, c.stream>>>(n, val, prefix, num, to, where);
        HGP_LAUNCH_CHECK();
        ran = true;
    }
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    c.arena.right_release(mark);
    return ran;
}

}  // namespace hgp
