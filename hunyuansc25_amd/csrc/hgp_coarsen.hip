// hgp_coarsen.hip -- one coarsening level: deterministic handshake heavy-edge
// matching with the reference's fallback passes, cmap, and contraction into
// the canonical coarse CSR (rows sorted by coarse id, no self-loops,
// duplicates merged).
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"

#include <algorithm>
#include <type_traits>

namespace hgp {

void attach_bins(Ctx &c, DevGraph &g)
{
    g.bins = row_bins(c, g.n, g.xadj, nullptr, true);
    g.segs = row_segs(c, g.xadj, g.bins);
}

int32_t coarsen_threshold(int32_t k)
{
    int32_t t = k * 8;
    if (t > 1024) {
        t = k * 2;
        if (t < 1024) t = 1024;
    }
    return t;
}

int32_t coarsen_maxvwgt(int64_t tvwgt, int32_t thr)
{
    int64_t v = (3 * tvwgt) / (2 * (int64_t)thr);
    return v > INT32_MAX ? INT32_MAX : (int32_t)v;
}

namespace {

constexpr int kBS = 256;

// Edge order of the handshake: weight desc, hash desc, lo asc, hi asc.
__device__ inline bool better(int32_t w, uint32_t h, int32_t lo, int32_t hi, int32_t bw, uint32_t bh, int32_t blo,
                              int32_t bhi)
{
    if (w != bw) return w > bw;
    if (h != bh) return h > bh;
    if (lo != blo) return lo < blo;
    return hi < bhi;
}

struct EdgeKey {
    int32_t w, lo, hi;
    uint32_t h;
};

struct BetterEdge {
    __device__ bool operator()(const EdgeKey &a, const EdgeKey &b) const
    {
        return better(a.w, a.h, a.lo, a.hi, b.w, b.h, b.lo, b.hi);
    }
};

// Candidates of a free vertex only drop out over the rounds of a level (a
// neighbour that is matched, dead or too heavy for v stays so while v is
// free), so if last round's choice of v is still free it is still v's best:
// v proposes to it again without reading its row.
__device__ inline bool keeps_proposal(int32_t first, int32_t v, const int32_t *__restrict__ avail,
                                      const int32_t *__restrict__ prop)
{
    if (first) return false;
    int32_t u = prop[v];
    return u >= 0 && avail[u] >= 0;
}

// Handshake proposal (replaces random_match_conflict, hunyuangraph_GPU_match.h:454-510,
// and set_receive_send_topk_one, :800-861): every free vertex proposes to its
// best free neighbour u with vwgt[u]+vwgt[v] < maxvwgt.  G lanes per row.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_propose(const int32_t *__restrict__ ids, int64_t cnt,
                                                    const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ avail, int32_t maxvwgt,
                                                    uint64_t s1, int32_t *__restrict__ prop,
                                                    const int32_t *prev_progress, int32_t first,
                                                    unsigned long long *kb)
{
    if (prev_progress && *prev_progress == 0) return;
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    uint64_t kbytes = 0;
    if (gi < cnt) {   // whole groups take the same branches
        int32_t v = ids[gi];
        int32_t wv = avail[v];
        if (wv < 0) {
            if (lane == 0) prop[v] = -1;
        } else if (keeps_proposal(first, v, avail, prop)) {
            kbytes = 0;
        } else {
            EdgeKey best{-1, 0, 0, 0u};
            eid_t beg = xadj[v], end = xadj[v + 1];
            const uint64_t tv = mix64(s1 ^ (uint64_t)(uint32_t)v);
            // entries per lane and phase: a whole class-w row for a wave, a whole
            // class-m row for a 16-lane group, one entry otherwise
            constexpr int T = G == kWave ? (int)(kBinMid / kWave) : (G == 16 ? (int)(kBinHalf / 16) : 1);
            for (eid_t c0 = beg; c0 < end; c0 += (eid_t)T * G) {
                int32_t u[T], w[T], wu[T];
                load_entries<T, G>(adjncy, adjwgt, c0, end, lane, u, w);
                gather_slots<T>(avail, u, wu, -1);
#pragma unroll
                for (int t = 0; t < T; t++) {
                    if (u[t] < 0 || u[t] == v || wu[t] < 0) continue;
                    if ((int64_t)wu[t] + wv >= maxvwgt) continue;
                    EdgeKey e{w[t], u[t] < v ? u[t] : v, u[t] < v ? v : u[t], 0u};
                    e.h = edge_hash_row(s1, tv, v, u[t]);
                    if (BetterEdge()(e, best)) best = e;
                }
            }
            best = group_best<G>(best, BetterEdge());
            if (lane == 0) {
                prop[v] = best.w < 0 ? -1 : (best.lo == v ? best.hi : best.lo);
                kbytes = 16 + 12 * (uint64_t)(end - beg);   // xadj; adjncy, adjwgt, avail[u]
            }
        }
    }
    kb_flush(kb, kbytes);
}

// Class-2 rows by segments (a hub row is shared by several blocks): the
// segment of a single-segment row writes the proposal; the segments of a
// multi row leave their best edge in sbest for k_propose_fin.
__global__ void __launch_bounds__(kSegBlk) k_propose_seg(RowSegs S, const eid_t *__restrict__ xadj,
                                                        const vid_t *__restrict__ adjncy,
                                                        const wgt_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ avail, int32_t maxvwgt,
                                                        uint64_t s1, int32_t *__restrict__ prop,
                                                        EdgeKey *__restrict__ sbest, const int32_t *prev_progress,
                                                        int32_t first, unsigned long long *kb)
{
    if (prev_progress && *prev_progress == 0) return;
    uint64_t kbytes = 0;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t mi = S.mi[s];
        int32_t wv = avail[v];
        if (wv < 0) {
            if (threadIdx.x == 0 && mi < 0) prop[v] = -1;
            continue;
        }
        if (keeps_proposal(first, v, avail, prop)) continue;   // (k_propose_fin makes the same test)
        if (mi >= 0 && !first) {
            // The same argument per segment: a segment whose best edge still
            // leads to a free vertex (or that had none) keeps it in sbest.
            const EdgeKey sb = sbest[s];
            if (sb.w < 0 || avail[sb.lo == v ? sb.hi : sb.lo] >= 0) continue;
        }
        EdgeKey best{-1, 0, 0, 0u};
        eid_t j0 = S.j0[s], j1 = min(j0 + (eid_t)kSegLen, xadj[v + 1]);
        const uint64_t tv = mix64(s1 ^ (uint64_t)(uint32_t)v);
        constexpr int T = 8;   // entries per thread and phase
        for (eid_t c0 = j0; c0 < j1; c0 += (eid_t)T * kSegBlk) {
            int32_t u[T], w[T], wu[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, j1, threadIdx.x, u, w);
            gather_slots<T>(avail, u, wu, -1);
#pragma unroll
            for (int t = 0; t < T; t++) {
                if (u[t] < 0 || u[t] == v || wu[t] < 0) continue;
                if ((int64_t)wu[t] + wv >= maxvwgt) continue;
                EdgeKey e{w[t], u[t] < v ? u[t] : v, u[t] < v ? v : u[t], 0u};
                e.h = edge_hash_row(s1, tv, v, u[t]);
                if (BetterEdge()(e, best)) best = e;
            }
        }
        best = group_best<kSegBlk>(best, BetterEdge());
        if (threadIdx.x == 0) {
            if (mi < 0) prop[v] = best.w < 0 ? -1 : (best.lo == v ? best.hi : best.lo);
            else sbest[s] = best;
            kbytes += 8 + 12 * (uint64_t)(j1 - j0);   // xadj[v+1]; adjncy, adjwgt, avail[u]
        }
    }
    kb_flush(kb, kbytes);
}

__global__ void k_propose_fin(RowSegs S, const int32_t *__restrict__ avail, const EdgeKey *__restrict__ sbest,
                              int32_t *__restrict__ prop, const int32_t *prev_progress, int32_t first)
{
    if (prev_progress && *prev_progress == 0) return;
    int32_t t = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= S.nmulti) return;
    int32_t v = S.mrow[t];
    if (avail[v] < 0) {   // the segments skipped this row: no fresh best edges
        prop[v] = -1;
        return;
    }
    if (keeps_proposal(first, v, avail, prop)) return;   // skipped by the segments too
    EdgeKey best{-1, 0, 0, 0u};
    for (int64_t s = S.mseg[2 * t]; s < S.mseg[2 * t + 1]; s++)
        if (BetterEdge()(sbest[s], best)) best = sbest[s];
    prop[v] = best.w < 0 ? -1 : (best.lo == v ? best.hi : best.lo);
}

// Mutual proposals match.  avail[v] is vwgt of a vertex that may still be
// matched in a later round, -1 otherwise: once matched, or once it had no
// candidate (a free vertex without a free neighbour that fits has none in
// later rounds either, and no neighbour can pick it), so later rounds skip
// its row and the proposals read one array per neighbour.
__global__ void k_accept(int32_t n, const int32_t *__restrict__ prop, int32_t *__restrict__ match,
                         int32_t *__restrict__ avail, int32_t *progress, const int32_t *prev_progress)
{
    if (prev_progress && *prev_progress == 0) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    bool got = false;
    if (v < n && match[v] == -1) {
        int32_t u = prop[v];
        if (u < 0) {
            avail[v] = -1;
        } else if (prop[u] == v) {
            match[v] = u;
            avail[v] = -1;
            got = true;
        }
    }
    if (__syncthreads_or(got) && threadIdx.x == 0) *progress = 1;   // only "any progress" is read
}

__global__ void __launch_bounds__(kBS) k_count_matched(int32_t n, const int32_t *__restrict__ match,
                                                       unsigned long long *cnt)
{
    __shared__ uint32_t part[kBS / kWave];
    uint32_t c = 0;
    for (int64_t v = (int64_t)blockIdx.x * kBS + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBS)
        c += match[v] >= 0;
    c = wave_sum(c);
    if (lane_id() == 0) part[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kBS / kWave; i++) t += part[i];
        if (t) atomicAdd(cnt, (unsigned long long)t);
    }
}

// ---- leaf (hunyuangraph_GPU_match.h:1374-1423) and relative (:1597-1680)
// matches: vertices that share a hub are paired consecutively in increasing
// id order, (0,1), (2,3), ...; a pair whose weight is not below maxvwgt stays
// unmatched.  hub[v] = -1 for vertices that take no part.
__global__ void k_leaf_hub(int32_t n, const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                           const int32_t *__restrict__ match, int32_t *__restrict__ hub, uint8_t *__restrict__ f)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    bool leaf = match[v] == -1 && xadj[v + 1] - xadj[v] == 1;
    hub[v] = leaf ? adjncy[xadj[v]] : -1;
    f[v] = leaf;
}

__global__ void k_hub_flag(int32_t n, const int32_t *__restrict__ hub, uint8_t *__restrict__ f)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) f[v] = hub[v] >= 0;
}

__global__ void k_hub_keys(const int64_t *__restrict__ cnt, const int32_t *__restrict__ list,
                           const int32_t *__restrict__ hub, uint32_t *__restrict__ key, uint32_t *__restrict__ val)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *cnt) return;
    int32_t v = list[i];
    key[i] = (uint32_t)hub[v];
    val[i] = (uint32_t)v;
}

__global__ void k_hub_heads(const int64_t *__restrict__ cnt, const uint32_t *__restrict__ key,
                            int32_t *__restrict__ h)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *cnt) return;
    h[i] = (i == 0 || key[i - 1] != key[i]) ? 1 : 0;
}

__global__ void k_hub_start(const int64_t *__restrict__ cnt, const int32_t *__restrict__ gid,
                            int32_t *__restrict__ start)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *cnt) return;
    if (i == 0 || gid[i - 1] != gid[i]) start[gid[i] - 1] = (int32_t)i;
}

__global__ void k_hub_pair(const int64_t *__restrict__ cnt, const uint32_t *__restrict__ key,
                           const uint32_t *__restrict__ val, const int32_t *__restrict__ gid,
                           const int32_t *__restrict__ start, const wgt_t *__restrict__ vwgt, int32_t maxvwgt,
                           int32_t *__restrict__ match)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nl = *cnt;
    if (i + 1 >= nl) return;
    if (((i - start[gid[i] - 1]) & 1) != 0 || key[i + 1] != key[i]) return;
    int32_t u = (int32_t)val[i], v = (int32_t)val[i + 1];
    if ((int64_t)vwgt[u] + vwgt[v] < maxvwgt) {
        match[u] = v;
        match[v] = u;
    }
}

// ---- isolate / forced matches: consecutive pairs of an id-ordered list
// (hunyuangraph_GPU_match.h:1577-1594, :1682-1716)
__global__ void k_flag_deg0(int32_t n, const eid_t *__restrict__ xadj, uint8_t *__restrict__ f)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) f[v] = xadj[v + 1] == xadj[v];
}

__global__ void k_flag_unmatched(int32_t n, const int32_t *__restrict__ match, uint8_t *__restrict__ f)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) f[v] = match[v] == -1;
}

// isolate: odd list positions pair with the previous entry if the odd one is
// still unmatched; forced: same but both are unmatched by construction.
__global__ void k_pair_list(int64_t cap, const int64_t *__restrict__ count, const int32_t *__restrict__ list,
                            const wgt_t *__restrict__ vwgt, int32_t maxvwgt, int32_t *match)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *count || (i & 1) == 0) return;
    int32_t x = list[i], y = list[i - 1];
    if (match[x] != -1) return;
    if ((int64_t)vwgt[x] + vwgt[y] < maxvwgt) { match[x] = y; match[y] = x; }
}

// ---- twin matches (hunyuangraph_GPU_match.h:1519-1574), restated as exact
// grouping by (degree, row sum) with consecutive pairing in id order.
// Row sums of adjncy for the unmatched rows (the twin key), by degree class:
// G lanes per row for rows of <= 512 entries, segments for longer rows
// (adding into rsum, which starts at zero).
template <int G>
__global__ void __launch_bounds__(kRowBS) k_rsum_g(const int32_t *__restrict__ ids, int64_t cnt,
                                                   const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                   const wgt_t *__restrict__ adjwgt, const int32_t *__restrict__ match,
                                                   int64_t *__restrict__ rsum)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;   // whole groups
    int32_t v = ids[gi];
    if (match[v] != -1) return;
    constexpr int T = G == kWave ? (int)(kBinMid / kWave) : 1;   // entries per lane and phase
    int64_t sum = 0;
    eid_t b = xadj[v], e = xadj[v + 1];
    for (eid_t c0 = b; c0 < e; c0 += (eid_t)T * G) {
        int32_t u[T], w[T];
        load_entries<T, G>(adjncy, adjwgt, c0, e, lane, u, w);
#pragma unroll
        for (int t = 0; t < T; t++) sum += u[t] >= 0 ? u[t] : 0;
    }
    sum = group_sum<G>(sum);
    if (lane == 0) rsum[v] = sum;
}

__global__ void __launch_bounds__(kSegBlk) k_rsum_seg(RowSegs S, const eid_t *__restrict__ xadj,
                                                     const vid_t *__restrict__ adjncy,
                                                     const wgt_t *__restrict__ adjwgt,
                                                     const int32_t *__restrict__ match,
                                                     unsigned long long *__restrict__ rsum)
{
    constexpr int T = 8;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        if (match[v] != -1) continue;   // block-uniform
        eid_t j0 = S.j0[s], j1 = min(j0 + (eid_t)kSegLen, xadj[v + 1]);
        int64_t sum = 0;
        for (eid_t c0 = j0; c0 < j1; c0 += (eid_t)T * kSegBlk) {
            int32_t u[T], w[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, j1, threadIdx.x, u, w);
#pragma unroll
            for (int t = 0; t < T; t++) sum += u[t] >= 0 ? u[t] : 0;
        }
        sum = group_sum<kSegBlk>(sum);
        if (threadIdx.x == 0 && sum) atomicAdd(&rsum[v], (unsigned long long)sum);
    }
}

// Twin sort keys of the unmatched list: the row sum, with the vertex id.
__global__ void k_twin_keys(int64_t cap, const int64_t *__restrict__ count, const int32_t *__restrict__ list,
                            const int64_t *__restrict__ rsum, uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *count) return;
    int32_t v = list[i];
    key[i] = (uint64_t)rsum[v];
    val[i] = (uint32_t)v;
}

__global__ void k_degkey(int64_t cap, const int64_t *__restrict__ count, const uint32_t *__restrict__ val,
                         const eid_t *__restrict__ xadj, uint32_t *__restrict__ key)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *count) return;
    int32_t v = (int32_t)val[i];
    key[i] = (uint32_t)(xadj[v + 1] - xadj[v]);
}

__device__ inline bool twin_same(int32_t a, int32_t b, const eid_t *xadj, const int64_t *rsum)
{
    return xadj[a + 1] - xadj[a] == xadj[b + 1] - xadj[b] && rsum[a] == rsum[b];
}

__global__ void k_twin_bound(int64_t cap, const int64_t *__restrict__ count, const uint32_t *__restrict__ val,
                             const eid_t *__restrict__ xadj, const int64_t *__restrict__ rsum, int32_t *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    if (i >= *count) { b[i] = 0; return; }
    b[i] = (i == 0 || !twin_same((int32_t)val[i], (int32_t)val[i - 1], xadj, rsum)) ? 1 : 0;
}

__global__ void k_twin_start(int64_t cap, const int64_t *__restrict__ count, const int32_t *__restrict__ gid,
                             const uint32_t *__restrict__ val, const eid_t *__restrict__ xadj,
                             const int64_t *__restrict__ rsum, int32_t *__restrict__ start)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *count) return;
    if (i == 0 || !twin_same((int32_t)val[i], (int32_t)val[i - 1], xadj, rsum)) start[gid[i] - 1] = (int32_t)i;
}

__global__ void k_twin_pair(int64_t cap, const int64_t *__restrict__ count, const int32_t *__restrict__ gid,
                            const int32_t *__restrict__ start, const uint32_t *__restrict__ val,
                            const eid_t *__restrict__ xadj, const int64_t *__restrict__ rsum,
                            const wgt_t *__restrict__ vwgt, int32_t maxvwgt, int32_t *match)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c = *count;
    if (i >= cap || i + 1 >= c) return;
    int64_t rank = i - start[gid[i] - 1];
    if (rank & 1) return;
    int32_t x = (int32_t)val[i], y = (int32_t)val[i + 1];
    if (!twin_same(x, y, xadj, rsum)) return;
    if ((int64_t)vwgt[x] + vwgt[y] < maxvwgt) { match[x] = y; match[y] = x; }
}

// ---- relative matches (hunyuangraph_GPU_match.h:1597-1680): dist = the
// neighbour of largest degree, then larger edge weight, then first in row.
struct DistKey {
    int32_t l, w, r, u;
};

struct BetterDist {
    __device__ bool operator()(const DistKey &a, const DistKey &b) const
    {
        if (a.l != b.l) return a.l > b.l;
        if (a.w != b.w) return a.w > b.w;
        return a.r < b.r;
    }
};

template <int G>
__global__ void __launch_bounds__(kRowBS) k_rel_dist(const int32_t *__restrict__ ids, int64_t cnt,
                                                     const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                     const wgt_t *__restrict__ adjwgt,
                                                     const int32_t *__restrict__ match, int32_t *__restrict__ dist)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t v = ids[gi];
    if (match[v] != -1) {
        if (lane == 0) dist[v] = -1;
        return;
    }
    constexpr int T = G == kWave ? (int)(kBinMid / kWave) : 1;   // entries per lane and phase
    DistKey best{-1, 0, 0, -1};
    eid_t b = xadj[v], e = xadj[v + 1];
    for (eid_t c0 = b; c0 < e; c0 += (eid_t)T * G) {
        int32_t u[T], w[T];
        int64_t d0[T], d1[T];
        load_entries<T, G>(adjncy, adjwgt, c0, e, lane, u, w);
#pragma unroll
        for (int t = 0; t < T; t++) {
            d0[t] = u[t] >= 0 ? xadj[u[t]] : 0;
            d1[t] = u[t] >= 0 ? xadj[u[t] + 1] : 0;
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (u[t] < 0) continue;
            DistKey k{(int32_t)(d1[t] - d0[t]), w[t], (int32_t)(c0 + lane + (eid_t)t * G - b), u[t]};
            if (BetterDist()(k, best)) best = k;
        }
    }
    best = group_best<G>(best, BetterDist());
    if (lane == 0) dist[v] = best.u;
}

// Long rows by segments: a single-segment row writes dist, a multi row leaves
// the best of each segment in dbest for k_rel_dist_fin.  (Degrees are
// differences of xadj entries, exact in 32 bits for rows of < 2^31 entries.)
__global__ void __launch_bounds__(kSegBlk) k_rel_dist_seg(RowSegs S, const eid_t *__restrict__ xadj,
                                                         const vid_t *__restrict__ adjncy,
                                                         const wgt_t *__restrict__ adjwgt,
                                                         const int32_t *__restrict__ match,
                                                         int32_t *__restrict__ dist, DistKey *__restrict__ dbest)
{
    constexpr int T = 8;
    for (int64_t s = blockIdx.x; s < S.nseg; s += gridDim.x) {
        int32_t v = S.row[s];
        int32_t mi = S.mi[s];
        if (match[v] != -1) {   // block-uniform
            if (threadIdx.x == 0 && mi < 0) dist[v] = -1;
            continue;
        }
        const eid_t b = xadj[v];
        eid_t j0 = S.j0[s], j1 = min(j0 + (eid_t)kSegLen, xadj[v + 1]);
        DistKey best{-1, 0, 0, -1};
        for (eid_t c0 = j0; c0 < j1; c0 += (eid_t)T * kSegBlk) {
            int32_t u[T], w[T];
            int64_t d0[T], d1[T];
            load_entries<T, kSegBlk>(adjncy, adjwgt, c0, j1, threadIdx.x, u, w);
#pragma unroll
            for (int t = 0; t < T; t++) {
                d0[t] = u[t] >= 0 ? xadj[u[t]] : 0;
                d1[t] = u[t] >= 0 ? xadj[u[t] + 1] : 0;
            }
#pragma unroll
            for (int t = 0; t < T; t++) {
                if (u[t] < 0) continue;
                DistKey k{(int32_t)(d1[t] - d0[t]), w[t], (int32_t)(c0 + threadIdx.x + (eid_t)t * kSegBlk - b), u[t]};
                if (BetterDist()(k, best)) best = k;
            }
        }
        best = group_best<kSegBlk>(best, BetterDist());
        if (threadIdx.x == 0) {
            if (mi < 0) dist[v] = best.u;
            else dbest[s] = best;
        }
    }
}

__global__ void k_rel_dist_fin(RowSegs S, const int32_t *__restrict__ match, const DistKey *__restrict__ dbest,
                               int32_t *__restrict__ dist)
{
    int32_t t = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= S.nmulti) return;
    int32_t v = S.mrow[t];
    if (match[v] != -1) {
        dist[v] = -1;
        return;
    }
    DistKey best{-1, 0, 0, -1};
    for (int64_t s = S.mseg[2 * t]; s < S.mseg[2 * t + 1]; s++)
        if (BetterDist()(dbest[s], best)) best = dbest[s];
    dist[v] = best.u;
}

// ---- cmap (resolve_conflict_12 + scan + resolve_conflict_4's copy)
__global__ void k_norm_flag(int32_t n, int32_t *__restrict__ match, int32_t *__restrict__ f)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int32_t u = match[v];
    if (u < 0) { u = v; match[v] = v; }
    f[v] = (v > 0 && v <= u) ? 1 : 0;
}

__global__ void k_cmap_fix(int32_t n, const int32_t *__restrict__ match, int32_t *cmap)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int32_t u = match[v];
    if (v > u) cmap[v] = cmap[u];
}

// ---- contraction
__global__ void k_rep(int32_t n, const eid_t *__restrict__ xadj, const wgt_t *__restrict__ vwgt,
                      const int32_t *__restrict__ match, const int32_t *__restrict__ cmap, int64_t *__restrict__ len,
                      wgt_t *__restrict__ cvwgt, int32_t *__restrict__ rep)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int32_t u = match[v];
    if (v > u) return;
    int32_t c = cmap[v];
    int64_t l = xadj[v + 1] - xadj[v];
    wgt_t w = vwgt[v];
    if (u != v) { l += xadj[u + 1] - xadj[u]; w += vwgt[u]; }
    len[c] = l;
    cvwgt[c] = w;
    rep[c] = v;
}

// ---- segmented sort-and-merge of the coarse rows
//
// Coarse row c is the concatenation of the rows of its fine vertices, renamed
// by cmap.  Each row is sorted by coarse id, entries equal to c are dropped
// and equal ids are merged by summing weights; the merged row is written at
// toff[c] (the concatenation's offset) of tadj/twgt and its length to ccnt[c].
// Rows of <= 64 entries sort in registers (bitonic over xor-shuffles), rows of
// <= kSegBig entries in LDS (radix, one block per row), longer rows through
// the device radix sort.

constexpr uint32_t kSent = 0xFFFFFFFFu;   // dropped entry; sorts last
constexpr int kSegWaveMax = 512;   // rows of 65..512 entries: one wave, up to 8 entries per lane
constexpr int kClsSegLong = 5;      // contraction class of the rows longer than kSegBig
constexpr int kSegLds = 4096;       // rows of <= kSegLds entries: a 512-thread block, <= 8 entries per thread
constexpr int kSegBS = 512;
constexpr int kSegE = kSegLds / kSegBS;
constexpr int kSegBig = 16384;      // rows of <= kSegBig entries: a 1024-thread block, <= 16 per thread (144 KB LDS)
constexpr int kSegBigBS = 1024;

// Entry i of coarse row c's concatenation (fine rows of v, then of u).
struct CatRow {
    eid_t bv, dv, bu;
    __device__ CatRow(const eid_t *xadj, int32_t v, int32_t u)
    {
        bv = xadj[v];
        dv = xadj[v + 1] - bv;
        bu = u != v ? xadj[u] : 0;
    }
    __device__ eid_t at(eid_t i) const { return i < dv ? bv + i : bu + (i - dv); }
};

template <int G>
__device__ inline void group_bitonic(uint32_t &key, int32_t &val)
{
    const int lane = glane<G>();
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint32_t ok = __shfl_xor(key, j, G);
            int32_t ov = __shfl_xor(val, j, G);
            bool asc = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool take = (lower == asc) ? ok < key : ok > key;
            if (take) {
                key = ok;
                val = ov;
            }
        }
    }
}

template <int G>
__device__ inline int32_t group_inclusive_scan(int32_t v)
{
    const int lane = glane<G>();
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        int32_t t = __shfl_up(v, o, G);
        if (lane >= o) v += t;
    }
    return v;
}

template <int G>
__device__ inline void seg_reg_row(int32_t c, int lane, const int32_t *__restrict__ rep,
                                   const int32_t *__restrict__ match, const eid_t *__restrict__ xadj,
                                   const vid_t *__restrict__ adjncy, const wgt_t *__restrict__ adjwgt,
                                   const int32_t *__restrict__ cmap, const int64_t *__restrict__ toff,
                                   int32_t *__restrict__ tadj, int32_t *__restrict__ twgt, int32_t *__restrict__ ccnt,
                                   uint64_t &kbytes)
{
    int32_t v = rep[c], u = match[v];
    CatRow row(xadj, v, u);
    eid_t len = row.dv + (u != v ? xadj[u + 1] - row.bu : 0);
    uint32_t key = kSent;
    int32_t w = 0;
    if (lane < len) {
        eid_t j = row.at(lane);
        int32_t cc = cmap[adjncy[j]];
        if (cc != c) {
            key = (uint32_t)cc;
            w = adjwgt[j];
        }
    }
    group_bitonic<G>(key, w);
    uint32_t nxt = __shfl_down(key, 1, G);
    bool tail = key != kSent && (lane == G - 1 || nxt != key);
    int32_t incl = group_inclusive_scan<G>(w);
    uint64_t tm = group_ballot<G>(tail);
    uint64_t below = tm & group_lanemask_lt<G>();
    int32_t prev = __shfl(incl, below ? 63 - __clzll(below) : 0, G);
    if (tail) {
        int64_t o = toff[c] + __popcll(below);
        tadj[o] = (int32_t)key;
        twgt[o] = incl - (below ? prev : 0);
    }
    if (lane == 0) {
        ccnt[c] = __popcll(tm);
        kbytes += 8 * (uint64_t)__popcll(tm);
    }
}

// Rows of <= G entries, one entry per lane.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_seg_reg(const int32_t *__restrict__ ids, int64_t cnt,
                                                    const int32_t *__restrict__ rep, const int32_t *__restrict__ match,
                                                    const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt, const int32_t *__restrict__ cmap,
                                                    const int64_t *__restrict__ toff, int32_t *__restrict__ tadj,
                                                    int32_t *__restrict__ twgt, int32_t *__restrict__ ccnt,
                                                    unsigned long long *kb)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    uint64_t kbytes = 0;
    if (gi < cnt) {   // whole groups take the same branches
        seg_reg_row<G>(ids[gi], lane, rep, match, xadj, adjncy, adjwgt, cmap, toff, tadj, twgt, ccnt, kbytes);
    }
    kb_flush(kb, kbytes);
}

// Bitonic network over R*64 entries held R per lane (entry e = r*64 + lane):
// partners closer than 64 are exchanged with xor-shuffles, farther ones are
// register swaps inside the lane.
template <int R, int KR>
__device__ inline void wave_bitonic(uint32_t (&key)[KR], int32_t (&val)[KR])
{
    const int lane = lane_id();
    constexpr int N = R * kWave;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int r2 = r ^ (j / kWave);
                    if (r2 > r) {
                        bool asc = ((r * kWave) & k) == 0;   // k >= 128: the lane bits do not matter
                        bool sw = asc ? key[r] > key[r2] : key[r] < key[r2];
                        if (sw) {
                            uint32_t tk = key[r];
                            key[r] = key[r2];
                            key[r2] = tk;
                            int32_t tv = val[r];
                            val[r] = val[r2];
                            val[r2] = tv;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    uint32_t ok = __shfl_xor(key[r], j, kWave);
                    int32_t ov = __shfl_xor(val[r], j, kWave);
                    bool asc = ((r * kWave + lane) & k) == 0;
                    bool lower = (lane & j) == 0;
                    bool take = (lower == asc) ? ok < key[r] : ok > key[r];
                    if (take) {
                        key[r] = ok;
                        val[r] = ov;
                    }
                }
            }
        }
    }
}

// Stable LSD radix sort of a block's P = (BS/64)*R*64 entries (registers
// 0..R-1 of RMAX in use) by the low 8*passes bits of the key (8-bit digits),
// entry w*C + r*64 + lane (C = R*64) held in register r of lane `lane` of
// wave w.  Each pass after the first ranks a wave's entries by digit
// (entries of equal digit keep their order: register-major, then lane) from
// 8 ballots per register and a per-wave count of each digit (wc, 256 ints
// per wave); the first pass, whose input order is arbitrary anyway, takes
// its ranks from LDS atomics on the same counts.  The counts become scatter
// bases by a scan over (digit, wave), and the entries scatter into sk/sw; the entries are read back into
// registers for the next pass.  The last pass leaves the sorted entries in
// sk/sw, P of them.  Keys beyond the low bits (the sentinel) must sort last
// on those bits alone.
template <int RMAX, int BS>
__device__ inline void block_radix(uint32_t (&key)[RMAX], int32_t (&val)[RMAX], int R, uint32_t *sk, int32_t *sw,
                                   int32_t *wc, int32_t *dsum, int passes)
{
    constexpr int NW = BS / kWave;
    const int C = R * kWave;
    const int w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t lt = lanemask_lt();
    int32_t *my = wc + w * 256;
    for (int pass = 0; pass < passes; pass++) {
        const int shift = 8 * pass;
        for (int i = lane; i < 256; i += kWave) my[i] = 0;
        wave_lds_sync();
        int32_t rk[RMAX];
        if (pass == 0) {
            // the first pass may order equal digits any way (the input order
            // is arbitrary anyway): ranks from LDS atomics
#pragma unroll
            for (int r = 0; r < RMAX; r++) {
                if (r >= R) break;
                rk[r] = atomicAdd(&my[(key[r] >> shift) & 255u], 1);
            }
            wave_lds_sync();
        } else {
#pragma unroll
            for (int r = 0; r < RMAX; r++) {
                if (r >= R) break;   // R is the same for the whole block
                const uint32_t d = (key[r] >> shift) & 255u;
                uint64_t peers = ~0ull;   // lanes whose register r holds the same digit
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const bool bit = (d >> b) & 1u;
                    const uint64_t bb = __ballot(bit);
                    peers &= bit ? bb : ~bb;
                }
                const int32_t base = my[d];
                rk[r] = base + __popcll(peers & lt);
                wave_lds_sync();   // every lane has its base
                if ((peers & lt) == 0) my[d] = base + __popcll(peers);
                wave_lds_sync();
            }
        }
        __syncthreads();
        // base of (digit d, wave x): entries of smaller digits, then of digit d in waves < x
        const int d = threadIdx.x & 255;
        int32_t run = 0;
        if (threadIdx.x < 256) {
#pragma unroll
            for (int x = 0; x < NW; x++) {
                const int32_t c = wc[x * 256 + d];
                wc[x * 256 + d] = run;
                run += c;
            }
        }
        const int32_t incl = wave_inclusive_scan(run);
        if (threadIdx.x < 256 && lane == kWave - 1) dsum[w] = incl;
        __syncthreads();
        if (threadIdx.x < 256) {
            int32_t pre = incl - run;
            for (int x = 0; x < w; x++) pre += dsum[x];
#pragma unroll
            for (int x = 0; x < NW; x++) wc[x * 256 + d] += pre;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RMAX; r++) {
            if (r >= R) break;
            const int32_t at = wc[w * 256 + ((key[r] >> shift) & 255u)] + rk[r];
            sk[at] = key[r];
            sw[at] = val[r];
        }
        __syncthreads();
        if (pass + 1 < passes) {
#pragma unroll
            for (int r = 0; r < RMAX; r++) {
                if (r >= R) break;
                key[r] = sk[w * C + r * kWave + lane];
                val[r] = sw[w * C + r * kWave + lane];
            }
        }
    }
}

// Rows of 65..kSegWaveMax entries: one wave per row, sorted in registers
// (padded to 128, 256 or 512 entries); tails ranked by ballots, weights by
// wave scans carried across the registers.
__global__ void __launch_bounds__(kRowBS) k_seg_wave(const int32_t *__restrict__ ids, int64_t cnt,
                                                     const int32_t *__restrict__ rep,
                                                     const int32_t *__restrict__ match,
                                                     const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                     const wgt_t *__restrict__ adjwgt,
                                                     const int32_t *__restrict__ cmap,
                                                     const int64_t *__restrict__ toff, int32_t *__restrict__ tadj,
                                                     int32_t *__restrict__ twgt, int32_t *__restrict__ ccnt,
                                                     unsigned long long *kb)
{
    constexpr int KR = kSegWaveMax / kWave;
    const int lane = lane_id();
    const int64_t gi = ((int64_t)blockIdx.x * kRowBS + threadIdx.x) / kWave;
    uint64_t kbytes = 0;
    if (gi < cnt) {   // whole waves
        int32_t c = ids[gi];
        int32_t v = rep[c], u = match[v];
        CatRow row(xadj, v, u);
        int len = (int)(row.dv + (u != v ? xadj[u + 1] - row.bu : 0));
        uint32_t key[KR];
        int32_t val[KR];
        int32_t nb[KR];
#pragma unroll
        for (int r = 0; r < KR; r++) {   // every index and weight load first, then the cmap gathers
            int i = r * kWave + lane;
            nb[r] = -1;
            val[r] = 0;
            if (i < len) {
                eid_t j = row.at(i);
                nb[r] = adjncy[j];
                val[r] = adjwgt[j];
            }
        }
#pragma unroll
        for (int r = 0; r < KR; r++) {
            int32_t cc = nb[r] >= 0 ? cmap[nb[r]] : c;
            key[r] = cc != c ? (uint32_t)cc : kSent;
            val[r] = cc != c ? val[r] : 0;
        }
        int R = len <= 128 ? 2 : (len <= 256 ? 4 : 8);
        if (R == 2) wave_bitonic<2>(key, val);
        else if (R == 4) wave_bitonic<4>(key, val);
        else wave_bitonic<8>(key, val);
        const int64_t o = toff[c];
        int32_t carry = 0, lastcum = 0, rank = 0;
#pragma unroll
        for (int r = 0; r < KR; r++) {
            if (r < R) {
                uint32_t nxt = __shfl_down(key[r], 1, kWave);
                uint32_t head_next = __shfl(key[r + 1 < KR ? r + 1 : r], 0, kWave);
                if (lane == kWave - 1) nxt = r + 1 < R ? head_next : kSent;
                bool tail = key[r] != kSent && nxt != key[r];
                int32_t incl = wave_inclusive_scan(val[r]) + carry;
                uint64_t tm = __ballot(tail);
                uint64_t below = tm & lanemask_lt();
                int32_t pv = __shfl(incl, below ? 63 - __clzll(below) : 0, kWave);
                int32_t last = __shfl(incl, tm ? 63 - __clzll(tm) : 0, kWave);
                if (tail) {
                    int64_t at = o + rank + __popcll(below);
                    tadj[at] = (int32_t)key[r];
                    twgt[at] = incl - (below ? pv : lastcum);
                }
                rank += __popcll(tm);
                if (tm) lastcum = last;
                carry = __shfl(incl, kWave - 1, kWave);
            }
        }
        if (lane == 0) {
            ccnt[c] = rank;
            kbytes = 8 * (uint64_t)rank;
        }
    }
    kb_flush(kb, kbytes);
}

// Rows of up to BS*RMAX entries: one block of BS threads per row, radix sort
// through LDS (padded to BS*RMAX, half or a quarter of it; `passes` 8-bit
// digits cover every coarse id), then tails (last entry of each run of equal ids) are ranked by
// a block scan and their inclusive weight prefixes differenced.
template <int BS, int RMAX>
__global__ void __launch_bounds__(BS) k_seg_lds(const int32_t *__restrict__ ids, int64_t cnt,
                                                    const int32_t *__restrict__ rep,
                                                    const int32_t *__restrict__ match,
                                                    const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ cmap, const int64_t *__restrict__ toff,
                                                    int32_t *__restrict__ tadj, int32_t *__restrict__ twgt,
                                                    int32_t *__restrict__ ccnt, int passes, unsigned long long *kb)
{
    constexpr int PMAX = BS * RMAX;
    __shared__ uint32_t sk[PMAX];
    uint64_t kbytes = 0;
    __shared__ int32_t sw[PMAX];
    __shared__ int32_t wc[BS / kWave * 256];
    __shared__ int32_t dsum[4];
    __shared__ int64_t sh[BS / kWave + 1];
    const int tid = threadIdx.x;
    for (int64_t gi = blockIdx.x; gi < cnt; gi += gridDim.x) {
        int32_t c = ids[gi];
        int32_t v = rep[c], u = match[v];
        CatRow row(xadj, v, u);
        int len = (int)(row.dv + (u != v ? xadj[u + 1] - row.bu : 0));
        const int P = len <= PMAX / 4 ? PMAX / 4 : (len <= PMAX / 2 ? PMAX / 2 : PMAX);
        {
            // entries straight into registers (entry w*C + r*64 + lane), sorted by the block
            const int R = P / BS, w = tid / kWave, lane = lane_id();
            uint32_t key[RMAX];
            int32_t val[RMAX];
            int32_t nb[RMAX];
#pragma unroll
            for (int r = 0; r < RMAX; r++) {   // every index and weight load first, then the cmap gathers
                int i = w * R * kWave + r * kWave + lane;
                nb[r] = -1;
                val[r] = 0;
                if (r < R && i < len) {
                    eid_t j = row.at(i);
                    nb[r] = adjncy[j];
                    val[r] = adjwgt[j];
                }
            }
#pragma unroll
            for (int r = 0; r < RMAX; r++) {
                int32_t cc = nb[r] >= 0 ? cmap[nb[r]] : c;
                key[r] = cc != c ? (uint32_t)cc : kSent;
                val[r] = cc != c ? val[r] : 0;
            }
            block_radix<RMAX, BS>(key, val, R, sk, sw, wc, dsum, passes);
        }
        // Thread tid owns entries [tid*E, tid*E + E), E = P / BS (one each when P < BS).
        const int E = P >= BS ? P / BS : 1;
        const int first = tid * E;
        uint32_t kk[RMAX];
        int32_t ww[RMAX];
        uint32_t after = first + E < P ? sk[first + E] : kSent;
#pragma unroll
        for (int e = 0; e < RMAX; e++) {
            int i = first + e;
            bool own = e < E && i < P;
            kk[e] = own ? sk[i] : kSent;
            ww[e] = own ? sw[i] : 0;
        }
        int32_t nt = 0, wsum = 0;
#pragma unroll
        for (int e = 0; e < RMAX; e++) {
            uint32_t nx = e + 1 < E ? kk[e + 1 < RMAX ? e + 1 : e] : after;
            if (e < E && kk[e] != kSent && nx != kk[e]) nt++;
            wsum += ww[e];
        }
        int64_t tot;
        int64_t ex = block_exclusive_scan(((int64_t)nt << 32) + wsum, sh, &tot);   // weight sums < 2^31
        int32_t rank = (int32_t)(ex >> 32), run = (int32_t)(ex & 0xFFFFFFFFll);
        const int32_t ntot = (int32_t)(tot >> 32);
        __syncthreads();   // every entry has been read: sk/sw now hold the merged row
#pragma unroll
        for (int e = 0; e < RMAX; e++) {
            uint32_t nx = e + 1 < E ? kk[e + 1 < RMAX ? e + 1 : e] : after;
            run += ww[e];
            if (e < E && kk[e] != kSent && nx != kk[e]) {
                sk[rank] = kk[e];
                sw[rank] = run;
                rank++;
            }
        }
        __syncthreads();
        const int64_t o = toff[c];
        for (int r = tid; r < ntot; r += BS) {
            tadj[o + r] = (int32_t)sk[r];
            twgt[o + r] = sw[r] - (r > 0 ? sw[r - 1] : 0);
        }
        if (tid == 0) {
            ccnt[c] = ntot;
            kbytes += 8 * (uint64_t)ntot;
        }
        __syncthreads();
    }
    kb_flush(kb, kbytes);
}

// Rows longer than kSegBig, by coarse-id ranges: the entries of each long row
// are counted and scattered into bins (row, id >> kBktBits); each bin is then
// merged in a dense LDS array of kBkt weights with a presence bitmap and
// emitted in id order.  Bins of one row are consecutive, so a row's merged
// entries are its bins' outputs in bin order.  Rows are canonical whatever
// the order of the atomics.
constexpr int kBktBits = 13;
constexpr int kBkt = 1 << kBktBits;      // coarse ids per bin
constexpr int64_t kMaxBins = 16ll << 20; // beyond this, the device radix sort path

// Long rows are cut into chunks of kLbChunk entries (one block each), so a
// hub row is spread over many blocks.  Bin counts and bin ranges are taken
// with one global atomic per chunk and bin, after an LDS histogram.
constexpr int kLbChunk = 4096;
constexpr int kLbMaxBk = 1024;   // bins per row (cn <= kLbMaxBk * kBkt) for the chunk kernels' LDS

__global__ void k_lb_nchunks(int64_t nl, const int32_t *__restrict__ lids, const int64_t *__restrict__ len,
                             int64_t *__restrict__ nch)
{
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li < nl) nch[li] = (len[lids[li]] + kLbChunk - 1) / kLbChunk;
}

__global__ void k_lb_chunks(int64_t nl, const int64_t *__restrict__ choff, int32_t *__restrict__ ch_li,
                            int32_t *__restrict__ ch_i0)
{
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    for (int64_t ch = choff[li]; ch < choff[li + 1]; ch++) {
        ch_li[ch] = (int32_t)li;
        ch_i0[ch] = (int32_t)((ch - choff[li]) * kLbChunk);
    }
}

__global__ void __launch_bounds__(kRowBS) k_lb_count(int64_t nch, const int32_t *__restrict__ ch_li,
                                                     const int32_t *__restrict__ ch_i0,
                                                     const int32_t *__restrict__ lids, int32_t nbk,
                                                     const int32_t *__restrict__ rep,
                                                     const int32_t *__restrict__ match,
                                                     const eid_t *__restrict__ xadj,
                                                     const vid_t *__restrict__ adjncy,
                                                     const int32_t *__restrict__ cmap, uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t hist[kLbMaxBk];
    const int tid = threadIdx.x;
    for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        for (int b = tid; b < nbk; b += kRowBS) hist[b] = 0;
        __syncthreads();
        int64_t li = ch_li[ch];
        int32_t c = lids[li];
        int32_t v = rep[c], u = match[v];
        CatRow row(xadj, v, u);
        eid_t len = row.dv + (u != v ? xadj[u + 1] - row.bu : 0);
        eid_t i0 = ch_i0[ch], i1 = min(i0 + (eid_t)kLbChunk, len);
        for (eid_t i = i0 + tid; i < i1; i += kRowBS) {
            int32_t cc = cmap[adjncy[row.at(i)]];
            if (cc != c) atomicAdd(&hist[cc >> kBktBits], 1u);
        }
        __syncthreads();
        uint32_t *bins = cnt + li * nbk;
        for (int b = tid; b < nbk; b += kRowBS)
            if (hist[b]) atomicAdd(&bins[b], hist[b]);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kRowBS) k_lb_scatter(int64_t nch, const int32_t *__restrict__ ch_li,
                                                       const int32_t *__restrict__ ch_i0,
                                                       const int32_t *__restrict__ lids, int32_t nbk,
                                                       const int32_t *__restrict__ rep,
                                                       const int32_t *__restrict__ match,
                                                       const eid_t *__restrict__ xadj,
                                                       const vid_t *__restrict__ adjncy,
                                                       const wgt_t *__restrict__ adjwgt,
                                                       const int32_t *__restrict__ cmap,
                                                       const int64_t *__restrict__ bstart, uint32_t *__restrict__ cur,
                                                       uint16_t *__restrict__ lo, int32_t *__restrict__ w)
{
    __shared__ uint32_t hist[kLbMaxBk];
    __shared__ int64_t gbase[kLbMaxBk];
    __shared__ int32_t ecc[kLbChunk];
    __shared__ int32_t ew[kLbChunk];
    const int tid = threadIdx.x;
    for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        for (int b = tid; b < nbk; b += kRowBS) hist[b] = 0;
        __syncthreads();
        int64_t li = ch_li[ch];
        int32_t c = lids[li];
        int32_t v = rep[c], u = match[v];
        CatRow row(xadj, v, u);
        eid_t len = row.dv + (u != v ? xadj[u + 1] - row.bu : 0);
        int32_t i0 = ch_i0[ch];
        int32_t ne = (int32_t)min((eid_t)kLbChunk, len - i0);
        for (int k = tid; k < ne; k += kRowBS) {
            eid_t j = row.at(i0 + k);
            int32_t cc = cmap[adjncy[j]];
            ecc[k] = cc;
            ew[k] = adjwgt[j];
            if (cc != c) atomicAdd(&hist[cc >> kBktBits], 1u);
        }
        __syncthreads();
        const int64_t b0 = li * nbk;
        for (int b = tid; b < nbk; b += kRowBS) {
            uint32_t h = hist[b];
            gbase[b] = h ? bstart[b0 + b] + (int64_t)atomicAdd(&cur[b0 + b], h) : 0;
            hist[b] = 0;
        }
        __syncthreads();
        for (int k = tid; k < ne; k += kRowBS) {
            int32_t cc = ecc[k];
            if (cc == c) continue;
            int b = cc >> kBktBits;
            int64_t pos = gbase[b] + atomicAdd(&hist[b], 1u);
            lo[pos] = (uint16_t)(cc & (kBkt - 1));
            w[pos] = ew[k];
        }
        __syncthreads();
    }
}

// One block per non-empty bin (persistent over bins); 256 threads own the 256
// bitmap words of a bin.  The merged entries are written at the bin's own
// start (a bin never grows) and their number to ucnt.
__global__ void __launch_bounds__(kRowBS) k_lb_merge(int64_t nbins, int32_t nbk, const int64_t *__restrict__ bstart,
                                                     const uint16_t *__restrict__ lo, const int32_t *__restrict__ w,
                                                     uint32_t *__restrict__ ucnt, int32_t *__restrict__ oadj,
                                                     int32_t *__restrict__ ow, unsigned long long *kb)
{
    static_assert(kBkt / 32 == kRowBS, "one bitmap word per thread");
    __shared__ int32_t dw[kBkt];
    __shared__ uint32_t bm[kBkt / 32];
    __shared__ int64_t sh[kRowBS / kWave + 1];
    const int tid = threadIdx.x;
    for (int i = tid; i < kBkt; i += kRowBS) dw[i] = 0;
    bm[tid] = 0;
    __syncthreads();
    uint64_t kbytes = 0;
    for (int64_t b = blockIdx.x; b < nbins; b += gridDim.x) {
        int64_t s = bstart[b], e = bstart[b + 1];
        if (s == e) {
            if (tid == 0) ucnt[b] = 0;
            continue;
        }
        for (int64_t i = s + tid; i < e; i += kRowBS) {
            uint32_t x = lo[i];
            atomicAdd(&dw[x], w[i]);
            atomicOr(&bm[x >> 5], 1u << (x & 31));
        }
        __syncthreads();
        uint32_t word = bm[tid];
        int64_t tot;
        int64_t ex = block_exclusive_scan((int64_t)__popc(word), sh, &tot);
        const int32_t base = (int32_t)((b % nbk) << kBktBits);
        int64_t at = s + ex;
        while (word) {
            int bit = __ffs(word) - 1;
            word &= word - 1;
            int idx = tid * 32 + bit;
            oadj[at] = base + idx;
            ow[at] = dw[idx];
            dw[idx] = 0;
            at++;
        }
        bm[tid] = 0;
        if (tid == 0) {
            ucnt[b] = (uint32_t)tot;
            kbytes += 16 + 6 * (uint64_t)(e - s) + 8 * (uint64_t)tot;
        }
        __syncthreads();
    }
    kb_flush(kb, kbytes);
}

// Bin outputs to their row's place at toff; one 64-lane group per bin.
__global__ void __launch_bounds__(kRowBS) k_lb_copy(int64_t nbins, int32_t nbk, const int32_t *__restrict__ lids,
                                                    const int64_t *__restrict__ bstart,
                                                    const uint32_t *__restrict__ ucnt,
                                                    const int64_t *__restrict__ uoff,
                                                    const int32_t *__restrict__ oadj, const int32_t *__restrict__ ow,
                                                    const int64_t *__restrict__ toff, int32_t *__restrict__ tadj,
                                                    int32_t *__restrict__ twgt)
{
    const int64_t b = gindex<kWave>();
    if (b >= nbins) return;
    uint32_t nu = ucnt[b];
    if (nu == 0) return;
    int64_t li = b / nbk;
    int64_t dst = toff[lids[li]] + (uoff[b] - uoff[li * nbk]);
    int64_t src = bstart[b];
    for (uint32_t i = lane_id(); i < nu; i += kWave) {
        tadj[dst + i] = oadj[src + i];
        twgt[dst + i] = ow[src + i];
    }
}

__global__ void k_lb_rowcnt(int64_t nl, int32_t nbk, const int32_t *__restrict__ lids,
                            const int64_t *__restrict__ uoff, int32_t *__restrict__ ccnt)
{
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li < nl) ccnt[lids[li]] = (int32_t)(uoff[(li + 1) * nbk] - uoff[li * nbk]);
}

// Rows longer than kSegBig: keys (list index << 32 | coarse id) of all such
// rows, laid out at loff, go through the device radix sort.
__global__ void __launch_bounds__(kRowBS) k_long_gather(const int32_t *__restrict__ ids, int64_t cnt, int32_t cn,
                                                        const int32_t *__restrict__ rep,
                                                        const int32_t *__restrict__ match,
                                                        const eid_t *__restrict__ xadj,
                                                        const vid_t *__restrict__ adjncy,
                                                        const wgt_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ cmap,
                                                        const int64_t *__restrict__ loff, uint64_t *__restrict__ key,
                                                        uint32_t *__restrict__ val)
{
    for (int64_t li = blockIdx.x; li < cnt; li += gridDim.x) {
        int32_t c = ids[li];
        int32_t v = rep[c], u = match[v];
        CatRow row(xadj, v, u);
        eid_t len = row.dv + (u != v ? xadj[u + 1] - row.bu : 0);
        uint64_t hi = (uint64_t)li << 32;
        int64_t o = loff[li];
        for (eid_t i = threadIdx.x; i < len; i += blockDim.x) {
            eid_t j = row.at(i);
            int32_t cc = cmap[adjncy[j]];
            key[o + i] = hi | (uint32_t)(cc == c ? cn : cc);
            val[o + i] = (uint32_t)adjwgt[j];
        }
    }
}

__global__ void k_long_len(int64_t cnt, const int32_t *__restrict__ ids, const int64_t *__restrict__ len,
                           int64_t *__restrict__ llen)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) llen[i] = len[ids[i]];
}

__global__ void k_heads(int64_t m, int32_t cn, const uint64_t *__restrict__ key, int32_t *__restrict__ h)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint64_t k = key[i];
    bool valid = (uint32_t)k != (uint32_t)cn;
    h[i] = (valid && (i == 0 || key[i - 1] != k)) ? 1 : 0;
}

__global__ void k_run_start(int64_t m, const int32_t *__restrict__ hsc, int32_t *__restrict__ start)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int32_t prev = i > 0 ? hsc[i - 1] : 0;
    if (hsc[i] != prev) start[hsc[i] - 1] = (int32_t)i;
}

// Last entry of each run writes the merged entry at its row's toff.
__global__ void k_long_write(int64_t tl, int32_t cn, const int32_t *__restrict__ ids, const uint64_t *__restrict__ key,
                             const int32_t *__restrict__ hsc, const int32_t *__restrict__ wsc,
                             const int32_t *__restrict__ start, const int64_t *__restrict__ loff,
                             const int64_t *__restrict__ toff, int32_t *__restrict__ tadj,
                             int32_t *__restrict__ twgt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tl) return;
    uint64_t k = key[i];
    uint32_t cc = (uint32_t)k;
    if (cc == (uint32_t)cn) return;
    if (i + 1 < tl && key[i + 1] == k) return;
    int64_t li = (int64_t)(k >> 32);
    int64_t b = loff[li];
    int32_t before = b > 0 ? hsc[b - 1] : 0;
    int32_t r = hsc[i] - 1;
    int32_t s = start[r];
    int64_t o = toff[ids[li]] + (r - before);
    tadj[o] = (int32_t)cc;
    twgt[o] = wsc[i] - (s > 0 ? wsc[s - 1] : 0);
}

__global__ void k_long_count(int64_t cnt, const int32_t *__restrict__ ids, const int64_t *__restrict__ loff,
                             const int32_t *__restrict__ hsc, int32_t *__restrict__ ccnt)
{
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= cnt) return;
    int64_t b = loff[li], e = loff[li + 1];
    int32_t before = b > 0 ? hsc[b - 1] : 0;
    ccnt[ids[li]] = (e > b ? hsc[e - 1] : before) - before;
}

// Moves each merged row from toff[c] to cxadj[c], four loads per lane in
// flight before the stores.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_seg_compact(const int32_t *__restrict__ ids, int64_t cnt,
                                                        const int64_t *__restrict__ toff,
                                                        const eid_t *__restrict__ cxadj,
                                                        const int32_t *__restrict__ tadj,
                                                        const int32_t *__restrict__ twgt, vid_t *__restrict__ cadjncy,
                                                        wgt_t *__restrict__ cadjwgt, unsigned long long *kb)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    uint64_t kbytes = 0;
    if (gi < cnt) {
        int32_t c = ids[gi];
        int64_t src = toff[c];
        eid_t dst = cxadj[c], l = cxadj[c + 1] - dst;
        eid_t j = lane;
        for (; j + 3 * G < l; j += 4 * G) {
            int32_t a[4], w[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                a[t] = tadj[src + j + t * G];
                w[t] = twgt[src + j + t * G];
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                cadjncy[dst + j + t * G] = a[t];
                cadjwgt[dst + j + t * G] = w[t];
            }
        }
        for (; j < l; j += G) {
            cadjncy[dst + j] = tadj[src + j];
            cadjwgt[dst + j] = twgt[src + j];
        }
        if (lane == 0) kbytes = 16 * (uint64_t)l;
    }
    kb_flush(kb, kbytes);
}

int64_t count_matched(Ctx &c, int32_t n, const int32_t *match, int64_t *d_cnt)
{
    fill<int64_t>(c, d_cnt, 1, 0);
    unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)n + kBS - 1) / kBS, 1024);
    k_count_matched<<<blocks, kBS, 0, c.stream>>>(n, match, (unsigned long long *)d_cnt);
    HGP_LAUNCH_CHECK();
    read_scalars(c, d_cnt, sizeof(int64_t));
    return c.pinned[0];
}

// Pairs the vertices with hub[v] >= 0 that share a hub, consecutively in
// increasing id order: compaction keeps id order and the radix sort by hub is
// stable.  f (n bytes) is scratch.
void pair_by_hub(Ctx &c, int32_t n, const int32_t *hub, uint8_t *f, const wgt_t *vwgt, int32_t maxvwgt,
                 int32_t *match)
{
    size_t mark = c.arena.right_mark();
    int32_t *list = c.arena.ralloc<int32_t>((size_t)n);
    int64_t *d_nl = c.arena.ralloc<int64_t>(1);
    compact_indices(c, f, n, list, d_nl);
    read_scalars(c, d_nl, sizeof(int64_t));
    int64_t nl = c.pinned[0];
    if (nl >= 2) {
        uint32_t *key = c.arena.ralloc<uint32_t>((size_t)nl);
        uint32_t *val = c.arena.ralloc<uint32_t>((size_t)nl);
        int32_t *gid = c.arena.ralloc<int32_t>((size_t)nl);
        int32_t *start = c.arena.ralloc<int32_t>((size_t)nl);
        unsigned g = grid_for(nl, kBS);
        k_hub_keys<<<g, kBS, 0, c.stream>>>(d_nl, list, hub, key, val);
        HGP_LAUNCH_CHECK();
        radix_sort(c, key, val, nl, 0, bits_for((uint64_t)(n > 0 ? n - 1 : 0)));
        k_hub_heads<<<g, kBS, 0, c.stream>>>(d_nl, key, gid);
        HGP_LAUNCH_CHECK();
        scan_inclusive(c, gid, gid, nl);
        k_hub_start<<<g, kBS, 0, c.stream>>>(d_nl, gid, start);
        HGP_LAUNCH_CHECK();
        k_hub_pair<<<g, kBS, 0, c.stream>>>(d_nl, key, val, gid, start, vwgt, maxvwgt, match);
        HGP_LAUNCH_CHECK();
    }
    c.arena.right_release(mark);
}

inline bool need_more(int64_t matched, int32_t n) { return 4 * matched < 3 * (int64_t)n; }

}  // namespace

// hunyuangraph_gpu_match (hunyuangraph_GPU_match.h:1846-3214) restated with
// deterministic rules: handshake rounds, then the fallback passes while fewer
// than 75% of vertices are matched (host-gated on an exact device count).
int64_t match_level(Ctx &c, const DevGraph &g, int32_t maxvwgt, uint64_t seed, int32_t level, int32_t rounds,
                    int32_t *match)
{
    int32_t n = g.n;
    if (n <= 0) return 0;
    size_t mark = c.arena.right_mark();
    int32_t *prop = c.arena.ralloc<int32_t>((size_t)n);
    int32_t *avail = c.arena.ralloc<int32_t>((size_t)n);
    EdgeKey *sbest = c.arena.ralloc<EdgeKey>((size_t)g.segs.nseg);
    int32_t *progress = c.arena.ralloc<int32_t>((size_t)(rounds > 0 ? rounds : 1));
    int64_t *d_cnt = c.arena.ralloc<int64_t>(2);
    fill<int32_t>(c, match, n, -1);
    HGP_CHECK_HIP(hipMemcpyAsync(avail, g.vwgt, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    fill<int32_t>(c, progress, rounds > 0 ? rounds : 1, 0);
    unsigned gv = grid_for(n, kBS);
    const uint64_t s1 = edge_key(seed, level);
    for (int r = 0; r < rounds; r++) {
        const int32_t *prev = r > 0 ? progress + r - 1 : nullptr;
        {
            // fixed bytes: ids, avail[v], prop per row (24 per segment); each free row adds its own in the kernel
            static const char *names[3] = {"k_propose_s", "k_propose_m", "k_propose_w"};
            if (g.bins.few() && g.bins.short_rows() > 0) {
                int64_t cnt = g.bins.short_rows();
                KScope ks(c, "k_propose_few", 12.0 * (double)cnt, true);
                k_propose<64><<<grid_groups(cnt, 64), kRowBS, 0, c.stream>>>(g.bins.list(0), cnt, g.xadj, g.adjncy,
                                                                             g.adjwgt, avail, maxvwgt, s1, prop, prev,
                                                                             r == 0, ks.dev());
                HGP_LAUNCH_CHECK();
            }
            for (int cl = 0; cl < 3 && !g.bins.few(); cl++) {
                int64_t cnt = g.bins.count(cl);
                if (!cnt) continue;
                KScope ks(c, names[cl], 12.0 * (double)cnt, true);
                const int32_t *ids = g.bins.list(cl);
                if (cl == 0)
                    k_propose<8><<<grid_groups(cnt, 8), kRowBS, 0, c.stream>>>(ids, cnt, g.xadj, g.adjncy, g.adjwgt,
                                                                               avail, maxvwgt, s1, prop, prev,
                                                                               r == 0, ks.dev());
                else if (cl == 1)
                    k_propose<16><<<grid_groups(cnt, 16), kRowBS, 0, c.stream>>>(ids, cnt, g.xadj, g.adjncy, g.adjwgt,
                                                                                 avail, maxvwgt, s1, prop, prev,
                                                                                 r == 0, ks.dev());
                else
                    k_propose<64><<<grid_groups(cnt, 64), kRowBS, 0, c.stream>>>(ids, cnt, g.xadj, g.adjncy, g.adjwgt,
                                                                                 avail, maxvwgt, s1, prop, prev,
                                                                                 r == 0, ks.dev());
                HGP_LAUNCH_CHECK();
            }
            if (g.segs.nseg) {
                {
                    KScope ks(c, "k_propose_seg", 24.0 * (double)g.segs.nseg, true);
                    k_propose_seg<<<(unsigned)std::min<int64_t>(g.segs.nseg, 32768), kSegBlk, 0, c.stream>>>(
                        g.segs, g.xadj, g.adjncy, g.adjwgt, avail, maxvwgt, s1, prop, sbest, prev, r == 0, ks.dev());
                    HGP_LAUNCH_CHECK();
                }
                if (g.segs.nmulti) {
                    k_propose_fin<<<grid_for(g.segs.nmulti, kBS), kBS, 0, c.stream>>>(g.segs, avail, sbest, prop, prev,
                                                                                     r == 0);
                    HGP_LAUNCH_CHECK();
                }
            }
        }
        {
            KScope ks(c, "k_accept", 16.0 * n);   // match, prop, prop[u]; match and avail writes
            k_accept<<<gv, kBS, 0, c.stream>>>(n, prop, match, avail, progress + r, prev);
            HGP_LAUNCH_CHECK();
        }
    }
    int64_t matched = count_matched(c, n, match, d_cnt);
    if (need_more(matched, n)) {
        size_t m2 = c.arena.right_mark();
        int32_t *hub = c.arena.ralloc<int32_t>((size_t)n);
        uint8_t *f = c.arena.ralloc<uint8_t>((size_t)n);
        k_leaf_hub<<<gv, kBS, 0, c.stream>>>(n, g.xadj, g.adjncy, match, hub, f);
        HGP_LAUNCH_CHECK();
        pair_by_hub(c, n, hub, f, g.vwgt, maxvwgt, match);
        c.arena.right_release(m2);
        matched = count_matched(c, n, match, d_cnt);
    }
    if (need_more(matched, n)) {
        size_t m2 = c.arena.right_mark();
        uint8_t *f = c.arena.ralloc<uint8_t>((size_t)n);
        int32_t *list = c.arena.ralloc<int32_t>((size_t)n);
        k_flag_deg0<<<gv, kBS, 0, c.stream>>>(n, g.xadj, f);
        HGP_LAUNCH_CHECK();
        compact_indices(c, f, n, list, d_cnt + 1);
        k_pair_list<<<gv, kBS, 0, c.stream>>>(n, d_cnt + 1, list, g.vwgt, maxvwgt, match);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(m2);
        matched = count_matched(c, n, match, d_cnt);
    }
    if (need_more(matched, n)) {
        size_t m2 = c.arena.right_mark();
        uint8_t *f = c.arena.ralloc<uint8_t>((size_t)n);
        int32_t *list = c.arena.ralloc<int32_t>((size_t)n);
        uint64_t *key = c.arena.ralloc<uint64_t>((size_t)n);
        uint32_t *val = c.arena.ralloc<uint32_t>((size_t)n);
        uint32_t *key32 = c.arena.ralloc<uint32_t>((size_t)n);
        int64_t *rsum = c.arena.ralloc<int64_t>((size_t)n);
        int32_t *gid = c.arena.ralloc<int32_t>((size_t)n);
        int32_t *start = c.arena.ralloc<int32_t>((size_t)n);
        k_flag_unmatched<<<gv, kBS, 0, c.stream>>>(n, match, f);
        HGP_LAUNCH_CHECK();
        compact_indices(c, f, n, list, d_cnt + 1);
        read_scalars(c, d_cnt + 1, sizeof(int64_t));
        int64_t nc = c.pinned[0];
        if (nc > 1) {
            unsigned gc = grid_for(nc, kBS);
            fill<int64_t>(c, rsum, n, 0);
            for (int cl = 0; cl < 3; cl++) {
                int64_t nb = g.bins.count(cl);
                if (!nb) continue;
                const int32_t *ids = g.bins.list(cl);
                if (cl == 0)
                    k_rsum_g<8><<<grid_groups(nb, 8), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt,
                                                                            match, rsum);
                else if (cl == 1)
                    k_rsum_g<32><<<grid_groups(nb, 32), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt,
                                                                              match, rsum);
                else
                    k_rsum_g<64><<<grid_groups(nb, 64), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt,
                                                                              match, rsum);
                HGP_LAUNCH_CHECK();
            }
            if (g.segs.nseg) {
                k_rsum_seg<<<(unsigned)std::min<int64_t>(g.segs.nseg, 16384), kSegBlk, 0, c.stream>>>(
                    g.segs, g.xadj, g.adjncy, g.adjwgt, match, (unsigned long long *)rsum);
                HGP_LAUNCH_CHECK();
            }
            k_twin_keys<<<gc, kBS, 0, c.stream>>>(nc, d_cnt + 1, list, rsum, key, val);
            HGP_LAUNCH_CHECK();
            radix_sort(c, key, val, nc, 0, bits_for((uint64_t)(n > 0 ? n - 1 : 0) * (uint64_t)(g.m > 0 ? g.m : 1)));
            k_degkey<<<gc, kBS, 0, c.stream>>>(nc, d_cnt + 1, val, g.xadj, key32);
            HGP_LAUNCH_CHECK();
            radix_sort(c, key32, val, nc, 0, bits_for((uint64_t)g.m));
            k_twin_bound<<<gc, kBS, 0, c.stream>>>(nc, d_cnt + 1, val, g.xadj, rsum, gid);
            HGP_LAUNCH_CHECK();
            scan_inclusive(c, gid, gid, nc);
            k_twin_start<<<gc, kBS, 0, c.stream>>>(nc, d_cnt + 1, gid, val, g.xadj, rsum, start);
            HGP_LAUNCH_CHECK();
            k_twin_pair<<<gc, kBS, 0, c.stream>>>(nc, d_cnt + 1, gid, start, val, g.xadj, rsum, g.vwgt, maxvwgt,
                                                  match);
            HGP_LAUNCH_CHECK();
        }
        c.arena.right_release(m2);
        matched = count_matched(c, n, match, d_cnt);
    }
    if (need_more(matched, n)) {
        size_t m2 = c.arena.right_mark();
        int32_t *dist = c.arena.ralloc<int32_t>((size_t)n);
        uint8_t *f = c.arena.ralloc<uint8_t>((size_t)n);
        for (int cl = 0; cl < 3; cl++) {
            int64_t nb = g.bins.count(cl);
            if (!nb) continue;
            const int32_t *ids = g.bins.list(cl);
            if (cl == 0)
                k_rel_dist<8><<<grid_groups(nb, 8), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt, match,
                                                                          dist);
            else if (cl == 1)
                k_rel_dist<32><<<grid_groups(nb, 32), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt,
                                                                            match, dist);
            else
                k_rel_dist<64><<<grid_groups(nb, 64), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt,
                                                                            match, dist);
            HGP_LAUNCH_CHECK();
        }
        if (g.segs.nseg) {
            DistKey *dbest = c.arena.ralloc<DistKey>((size_t)g.segs.nseg);
            k_rel_dist_seg<<<(unsigned)std::min<int64_t>(g.segs.nseg, 16384), kSegBlk, 0, c.stream>>>(
                g.segs, g.xadj, g.adjncy, g.adjwgt, match, dist, dbest);
            HGP_LAUNCH_CHECK();
            if (g.segs.nmulti) {
                k_rel_dist_fin<<<grid_for(g.segs.nmulti, kBS), kBS, 0, c.stream>>>(g.segs, match, dbest, dist);
                HGP_LAUNCH_CHECK();
            }
        }
        k_hub_flag<<<gv, kBS, 0, c.stream>>>(n, dist, f);
        HGP_LAUNCH_CHECK();
        pair_by_hub(c, n, dist, f, g.vwgt, maxvwgt, match);
        c.arena.right_release(m2);
        matched = count_matched(c, n, match, d_cnt);
    }
    if (need_more(matched, n)) {
        size_t m2 = c.arena.right_mark();
        uint8_t *f = c.arena.ralloc<uint8_t>((size_t)n);
        int32_t *list = c.arena.ralloc<int32_t>((size_t)n);
        k_flag_unmatched<<<gv, kBS, 0, c.stream>>>(n, match, f);
        HGP_LAUNCH_CHECK();
        compact_indices(c, f, n, list, d_cnt + 1);
        k_pair_list<<<gv, kBS, 0, c.stream>>>(n, d_cnt + 1, list, g.vwgt, maxvwgt, match);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(m2);
        matched = count_matched(c, n, match, d_cnt);
    }
    c.arena.right_release(mark);
    return matched;
}

int32_t build_cmap(Ctx &c, int32_t n, int32_t *match, int32_t *cmap)
{
    if (n <= 0) return 0;
    unsigned gv = grid_for(n, kBS);
    k_norm_flag<<<gv, kBS, 0, c.stream>>>(n, match, cmap);
    HGP_LAUNCH_CHECK();
    scan_inclusive(c, cmap, cmap, n);
    // cnvtxs = number of representatives = the scan's last value, read before
    // partners copy their representative's id (vertex n-1 may be a partner).
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned, cmap + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    k_cmap_fix<<<gv, kBS, 0, c.stream>>>(n, match, cmap);
    HGP_LAUNCH_CHECK();
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    return ((int32_t *)c.pinned)[0] + 1;
}

// hunyuangraph_gpu_create_cgraph (hunyuangraph_GPU_contraction.h:715-1247):
// gather the concatenated rows of each coarse vertex, stable-radix-sort the
// (row, coarse neighbour) keys, merge equal runs by summing weights (scan
// differences) and drop internal entries.
DevGraph contract(Ctx &c, const DevGraph &g, const int32_t *match, const int32_t *cmap, int32_t cn)
{
    DevGraph cg;
    cg.n = cn;
    cg.xadj = c.arena.lalloc<eid_t>((size_t)cn + 1);
    cg.vwgt = c.arena.lalloc<wgt_t>((size_t)cn);
    int64_t m = g.m;
    size_t mark = c.arena.right_mark();
    int64_t *len = c.arena.ralloc<int64_t>((size_t)cn);
    int32_t *rep = c.arena.ralloc<int32_t>((size_t)cn);
    int64_t *toff = c.arena.ralloc<int64_t>((size_t)cn + 1);
    unsigned gv = grid_for(g.n, kBS);
    k_rep<<<gv, kBS, 0, c.stream>>>(g.n, g.xadj, g.vwgt, match, cmap, len, cg.vwgt, rep);
    HGP_LAUNCH_CHECK();
    if (m <= 0) {
        cg.adjncy = c.arena.lalloc<vid_t>(1);
        cg.adjwgt = c.arena.lalloc<wgt_t>(1);
        fill<int64_t>(c, cg.xadj, (int64_t)cn + 1, 0);
        cg.m = 0;
        c.arena.right_release(mark);
        return cg;
    }
    offsets_from_counts(c, len, toff, cn);   // the concatenations fill [0, m) exactly
    int32_t *tadj = c.arena.ralloc<int32_t>((size_t)m);
    int32_t *twgt = c.arena.ralloc<int32_t>((size_t)m);
    int32_t *ccnt = c.arena.ralloc<int32_t>((size_t)cn);
    RowBins cb = row_bins(c, cn, nullptr, len, false, 8, 64, kSegWaveMax, kSegLds, kSegBig);
    {
        // per coarse row: ids, rep, match, 2 xadj pairs, toff, count (56 B); per fine entry: adjncy, adjwgt,
        // cmap gather (12 B); the kernels count the merged entries they write (8 B each)
        auto bytes = [&](int cl) { return 56.0 * (double)cb.count(cl) + 12.0 * (double)cb.ent[cl]; };
        if (cb.count(0)) {
            KScope ks(c, "k_seg_reg8", bytes(0), true);
            k_seg_reg<8><<<grid_groups(cb.count(0), 8), kRowBS, 0, c.stream>>>(
                cb.list(0), cb.count(0), rep, match, g.xadj, g.adjncy, g.adjwgt, cmap, toff, tadj, twgt, ccnt, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (cb.count(1)) {
            KScope ks(c, "k_seg_reg64", bytes(1), true);
            k_seg_reg<64><<<grid_groups(cb.count(1), 64), kRowBS, 0, c.stream>>>(
                cb.list(1), cb.count(1), rep, match, g.xadj, g.adjncy, g.adjwgt, cmap, toff, tadj, twgt, ccnt, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (cb.count(2)) {
            KScope ks(c, "k_seg_wave", bytes(2), true);
            k_seg_wave<<<grid_groups(cb.count(2), kWave), kRowBS, 0, c.stream>>>(
                cb.list(2), cb.count(2), rep, match, g.xadj, g.adjncy, g.adjwgt, cmap, toff, tadj, twgt, ccnt,
                ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (cb.count(3)) {
            KScope ks(c, "k_seg_lds", bytes(3), true);
            unsigned blocks = (unsigned)std::min<int64_t>(cb.count(3), 8192);
            const int passes = std::max(1, (bits_for((uint64_t)cn) + 7) / 8);   // coarse ids < 2^(8*passes) - 1
            k_seg_lds<kSegBS, kSegE><<<blocks, kSegBS, 0, c.stream>>>(cb.list(3), cb.count(3), rep, match, g.xadj,
                                                                      g.adjncy, g.adjwgt, cmap, toff, tadj, twgt,
                                                                      ccnt, passes, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        if (cb.count(4)) {
            KScope ks(c, "k_seg_lds_big", bytes(4), true);
            unsigned blocks = (unsigned)std::min<int64_t>(cb.count(4), 4096);
            const int passes = std::max(1, (bits_for((uint64_t)cn) + 7) / 8);
            k_seg_lds<kSegBigBS, kSegBig / kSegBigBS><<<blocks, kSegBigBS, 0, c.stream>>>(
                cb.list(4), cb.count(4), rep, match, g.xadj, g.adjncy, g.adjwgt, cmap, toff, tadj, twgt, ccnt, passes,
                ks.dev());
            HGP_LAUNCH_CHECK();
        }
    }
    const int32_t nbk = (int32_t)(((int64_t)cn + kBkt - 1) >> kBktBits);
    if (int64_t nl = cb.count(kClsSegLong); nl > 0 && nl * (int64_t)nbk <= kMaxBins && nbk <= kLbMaxBk) {
        // long rows by coarse-id bins
        size_t m2 = c.arena.right_mark();
        const int32_t *lids = cb.list(kClsSegLong);
        const int64_t nbins = nl * (int64_t)nbk, tl = cb.ent[kClsSegLong];
        int64_t *nchr = c.arena.ralloc<int64_t>((size_t)nl);
        int64_t *choff = c.arena.ralloc<int64_t>((size_t)nl + 1);
        k_lb_nchunks<<<grid_for(nl, kBS), kBS, 0, c.stream>>>(nl, lids, len, nchr);
        HGP_LAUNCH_CHECK();
        offsets_from_counts(c, nchr, choff, nl);
        read_scalars(c, choff + nl, sizeof(int64_t));
        const int64_t nch = c.pinned[0];
        int32_t *ch_li = c.arena.ralloc<int32_t>((size_t)nch);
        int32_t *ch_i0 = c.arena.ralloc<int32_t>((size_t)nch);
        k_lb_chunks<<<grid_for(nl, kBS), kBS, 0, c.stream>>>(nl, choff, ch_li, ch_i0);
        HGP_LAUNCH_CHECK();
        uint32_t *cnt = c.arena.ralloc<uint32_t>((size_t)nbins);
        int64_t *bstart = c.arena.ralloc<int64_t>((size_t)nbins + 1);
        int64_t *uoff = c.arena.ralloc<int64_t>((size_t)nbins + 1);
        uint16_t *lo = c.arena.ralloc<uint16_t>((size_t)tl);
        int32_t *w = c.arena.ralloc<int32_t>((size_t)tl);
        int32_t *oadj = c.arena.ralloc<int32_t>((size_t)tl);
        int32_t *ow = c.arena.ralloc<int32_t>((size_t)tl);
        unsigned ch_grid = (unsigned)std::min<int64_t>(nch, 16384);
        fill<uint32_t>(c, cnt, nbins, 0u);
        {
            // per entry: adjncy, cmap gather; per chunk and bin: one count atomic
            KScope ks(c, "k_lb_count", 8.0 * (double)tl + 56.0 * nch + 4.0 * (double)nbins);
            k_lb_count<<<ch_grid, kRowBS, 0, c.stream>>>(nch, ch_li, ch_i0, lids, nbk, rep, match, g.xadj, g.adjncy,
                                                        cmap, cnt);
            HGP_LAUNCH_CHECK();
        }
        offsets_from_counts(c, (const int32_t *)cnt, bstart, nbins);
        fill<uint32_t>(c, cnt, nbins, 0u);
        {
            // per entry: adjncy, adjwgt, cmap gather, 6 B out; per chunk and bin: bstart and a range atomic
            KScope ks(c, "k_lb_scatter", 18.0 * (double)tl + 56.0 * nch + 12.0 * (double)nbins);
            k_lb_scatter<<<ch_grid, kRowBS, 0, c.stream>>>(nch, ch_li, ch_i0, lids, nbk, rep, match, g.xadj, g.adjncy,
                                                          g.adjwgt, cmap, bstart, cnt, lo, w);
            HGP_LAUNCH_CHECK();
        }
        {
            KScope ks(c, "k_lb_merge", 0.0, true);
            k_lb_merge<<<(unsigned)std::min<int64_t>(nbins, 8192), kRowBS, 0, c.stream>>>(
                nbins, nbk, bstart, lo, w, cnt, oadj, ow, ks.dev());
            HGP_LAUNCH_CHECK();
        }
        offsets_from_counts(c, (const int32_t *)cnt, uoff, nbins);
        k_lb_copy<<<grid_groups(nbins, kWave), kRowBS, 0, c.stream>>>(nbins, nbk, lids, bstart, cnt, uoff, oadj, ow,
                                                                      toff, tadj, twgt);
        HGP_LAUNCH_CHECK();
        k_lb_rowcnt<<<grid_for(nl, kBS), kBS, 0, c.stream>>>(nl, nbk, lids, uoff, ccnt);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(m2);
    } else if (nl > 0) {
        // more bins than kMaxBins: the device radix sort over (row, coarse id) keys
        size_t m2 = c.arena.right_mark();
        const int32_t *lids = cb.list(kClsSegLong);
        int64_t *llen = c.arena.ralloc<int64_t>((size_t)nl);
        int64_t *loff = c.arena.ralloc<int64_t>((size_t)nl + 1);
        k_long_len<<<grid_for(nl, kBS), kBS, 0, c.stream>>>(nl, lids, len, llen);
        HGP_LAUNCH_CHECK();
        offsets_from_counts(c, llen, loff, nl);
        read_scalars(c, loff + nl, sizeof(int64_t));
        int64_t tl = c.pinned[0];
        uint64_t *key = c.arena.ralloc<uint64_t>((size_t)tl);
        uint32_t *val = c.arena.ralloc<uint32_t>((size_t)tl);
        int32_t *hsc = c.arena.ralloc<int32_t>((size_t)tl);
        int32_t *wsc = c.arena.ralloc<int32_t>((size_t)tl);
        int32_t *start = c.arena.ralloc<int32_t>((size_t)tl);
        {
            KScope ks(c, "k_long_gather", 24.0 * (double)tl + 44.0 * nl);   // + 12 B in per entry, key and val out
            k_long_gather<<<(unsigned)std::min<int64_t>(nl, 8192), kRowBS, 0, c.stream>>>(
                lids, nl, cn, rep, match, g.xadj, g.adjncy, g.adjwgt, cmap, loff, key, val);
            HGP_LAUNCH_CHECK();
        }
        radix_sort(c, key, val, tl, 0, bits_for((uint64_t)cn));
        if (nl > 1) radix_sort(c, key, val, tl, 32, 32 + bits_for((uint64_t)(nl - 1)));
        unsigned gt = grid_for(tl, kBS);
        k_heads<<<gt, kBS, 0, c.stream>>>(tl, cn, key, hsc);
        HGP_LAUNCH_CHECK();
        scan_inclusive(c, hsc, hsc, tl);
        scan_inclusive(c, (const int32_t *)val, wsc, tl);
        k_run_start<<<gt, kBS, 0, c.stream>>>(tl, hsc, start);
        HGP_LAUNCH_CHECK();
        k_long_write<<<gt, kBS, 0, c.stream>>>(tl, cn, lids, key, hsc, wsc, start, loff, toff, tadj, twgt);
        HGP_LAUNCH_CHECK();
        k_long_count<<<grid_for(nl, kBS), kBS, 0, c.stream>>>(nl, lids, loff, hsc, ccnt);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(m2);
    }
    offsets_from_counts(c, ccnt, cg.xadj, cn);
    read_scalars(c, cg.xadj + cn, sizeof(int64_t));
    int64_t cm = c.pinned[0];
    cg.adjncy = c.arena.lalloc<vid_t>((size_t)(cm > 0 ? cm : 1));
    cg.adjwgt = c.arena.lalloc<wgt_t>((size_t)(cm > 0 ? cm : 1));
    {
        // classes 2..5 are consecutive in the id list: one launch of 256-lane groups
        // rows of <= 8 / <= 512 / longer pre-merge entries: 8 lanes, a wave, a block
        const int64_t cnt[3] = {cb.count(0), cb.count(1) + cb.count(2), cb.off[kClsSegLong + 1] - cb.off[3]};
        const int32_t *lst[3] = {cb.list(0), cb.list(1), cb.list(3)};
        static const char *names[3] = {"k_seg_compact8", "k_seg_compact64", "k_seg_compact256"};
        for (int cl = 0; cl < 3; cl++) {
            if (!cnt[cl]) continue;
            // per row: ids, toff, cxadj pair (28 B); the kernel counts the merged row read and written (16 B per entry)
            KScope ks(c, names[cl], 28.0 * (double)cnt[cl], true);
            if (cl == 0)
                k_seg_compact<8><<<grid_groups(cnt[cl], 8), kRowBS, 0, c.stream>>>(
                    lst[cl], cnt[cl], toff, cg.xadj, tadj, twgt, cg.adjncy, cg.adjwgt, ks.dev());
            else if (cl == 1)
                k_seg_compact<64><<<grid_groups(cnt[cl], 64), kRowBS, 0, c.stream>>>(
                    lst[cl], cnt[cl], toff, cg.xadj, tadj, twgt, cg.adjncy, cg.adjwgt, ks.dev());
            else
                k_seg_compact<256><<<grid_groups(cnt[cl], 256), kRowBS, 0, c.stream>>>(
                    lst[cl], cnt[cl], toff, cg.xadj, tadj, twgt, cg.adjncy, cg.adjwgt, ks.dev());
            HGP_LAUNCH_CHECK();
        }
    }
    cg.m = cm;
    c.arena.right_release(mark);
    return cg;
}

}  // namespace hgp
