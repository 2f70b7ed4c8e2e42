// hgp_coarsen.hip -- one coarsening level: deterministic handshake heavy-edge
// matching with the reference's fallback passes, cmap, and contraction into
// the canonical coarse CSR (rows sorted by coarse id, no self-loops,
// duplicates merged).
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"

#include <type_traits>

namespace hgp {

void attach_bins(Ctx &c, DevGraph &g) { g.bins = row_bins(c, g.n, g.xadj, nullptr, true); }

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
constexpr int kG = 16;          // lanes per vertex in row-scanning kernels

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

// Handshake proposal (replaces random_match_conflict, hunyuangraph_GPU_match.h:454-510,
// and set_receive_send_topk_one, :800-861): every free vertex proposes to its
// best free neighbour u with vwgt[u]+vwgt[v] < maxvwgt.  G lanes per row.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_propose(const int32_t *__restrict__ ids, int64_t cnt,
                                                    const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                    const wgt_t *__restrict__ adjwgt, const wgt_t *__restrict__ vwgt,
                                                    const int32_t *__restrict__ match, int32_t maxvwgt,
                                                    uint64_t seed, int32_t level, int32_t *__restrict__ prop,
                                                    const int32_t *prev_progress)
{
    if (prev_progress && *prev_progress == 0) return;
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t v = ids[gi];
    if (match[v] != -1) {
        if (lane == 0) prop[v] = -1;
        return;
    }
    int32_t wv = vwgt[v];
    EdgeKey best{-1, 0, 0, 0u};
    eid_t end = xadj[v + 1];
    for (eid_t j = xadj[v] + lane; j < end; j += G) {
        int32_t u = adjncy[j];
        if (u == v || match[u] != -1) continue;
        if ((int64_t)vwgt[u] + wv >= maxvwgt) continue;
        EdgeKey e{adjwgt[j], u < v ? u : v, u < v ? v : u, 0u};
        e.h = edge_hash(seed, level, e.lo, e.hi);
        if (BetterEdge()(e, best)) best = e;
    }
    best = group_best<G>(best, BetterEdge());
    if (lane == 0) prop[v] = best.w < 0 ? -1 : (best.lo == v ? best.hi : best.lo);
}

__global__ void k_accept(int32_t n, const int32_t *__restrict__ prop, int32_t *__restrict__ match, int32_t *progress,
                         const int32_t *prev_progress)
{
    if (prev_progress && *prev_progress == 0) return;
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n || match[v] != -1) return;
    int32_t u = prop[v];
    if (u >= 0 && prop[u] == v) {
        match[v] = u;
        atomicAdd(progress, 1);
    }
}

__global__ void k_count_matched(int32_t n, const int32_t *__restrict__ match, unsigned long long *cnt)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    uint64_t b = __ballot(v < n && match[v] >= 0);
    if (lane_id() == 0 && b) atomicAdd(cnt, (unsigned long long)__popcll(b));
}

// ---- leaf matches (hunyuangraph_GPU_match.h:1374-1423)
__global__ void k_leaf_count(int32_t n, const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                             const int32_t *__restrict__ match, int32_t *__restrict__ cnt)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n || match[v] != -1) return;
    if (xadj[v + 1] - xadj[v] == 1) atomicAdd(&cnt[adjncy[xadj[v]]], 1);
}

// Consecutive pairing in row order: the eligible entries of hub h's row form
// a sequence that is paired (0,1), (2,3), ...; a pair whose weight is not
// below maxvwgt is skipped.  A G-lane group walks the row in chunks; lane
// ranks come from ballots, the unpaired last entry of a chunk is carried.
// Leaf (hunyuangraph_GPU_match.h:1374-1423): eligible = unmatched degree-1
// neighbour.  Relative (:1597-1680): eligible = unmatched neighbour whose
// dist is h.  Each eligible vertex belongs to exactly one hub.
template <int G, bool kLeaf>
__global__ void __launch_bounds__(kRowBS) k_pair_row(const int32_t *__restrict__ ids, int64_t cnt,
                                                     const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                     const wgt_t *__restrict__ vwgt, int32_t maxvwgt,
                                                     const int32_t *__restrict__ dist,
                                                     const int32_t *__restrict__ hubcnt, int32_t *match)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t h = ids[gi];
    if (hubcnt[h] <= 1) return;
    int32_t pending = -1;
    eid_t e = xadj[h + 1];
    for (eid_t s = xadj[h]; s < e; s += G) {
        eid_t j = s + lane;
        int32_t u = j < e ? adjncy[j] : -1;
        bool elig = false;
        if (u >= 0 && match[u] == -1) {
            if (kLeaf) elig = xadj[u + 1] - xadj[u] == 1;
            else elig = dist[u] == h;
        }
        uint64_t mask = group_ballot<G>(elig);
        int npend = pending >= 0 ? 1 : 0;
        // shuffles run on every lane of the group (a shuffle from an inactive
        // lane would not deliver its value): the previous eligible lane's
        // vertex, and the last eligible vertex of the chunk
        uint64_t below = mask & group_lanemask_lt<G>();
        int32_t prev_u = __shfl(u, below ? 63 - __clzll((long long)below) : 0, G);
        int32_t last_u = __shfl(u, mask ? 63 - __clzll((long long)mask) : 0, G);
        if (elig) {
            int t = __popcll(below) + npend;   // index in this chunk's sequence
            if (t & 1) {
                int32_t partner = (t == 1 && npend) ? pending : prev_u;
                if ((int64_t)vwgt[partner] + vwgt[u] < maxvwgt) { match[partner] = u; match[u] = partner; }
            }
        }
        // uniform update of the carried entry
        int total = __popcll(mask) + npend;
        if (total & 1) pending = mask ? last_u : pending;
        else pending = -1;
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
__global__ void __launch_bounds__(kBS) k_rowsum(int64_t cap, const int64_t *__restrict__ count,
                                                const int32_t *__restrict__ list, const eid_t *__restrict__ xadj,
                                                const vid_t *__restrict__ adjncy, uint64_t *__restrict__ key,
                                                uint32_t *__restrict__ val, int64_t *__restrict__ rsum)
{
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gt / kG;
    int lane = (int)(threadIdx.x % kG);
    if (i >= cap || i >= *count) return;
    int32_t v = list[i];
    int64_t s = 0;
    for (eid_t j = xadj[v] + lane; j < xadj[v + 1]; j += kG) s += adjncy[j];
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kG);
    if (lane == 0) {
        key[i] = (uint64_t)s;
        val[i] = (uint32_t)v;
        rsum[v] = s;
    }
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
                                                     const int32_t *__restrict__ match, int32_t *__restrict__ dist,
                                                     int32_t *__restrict__ hubcnt)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t v = ids[gi];
    if (match[v] != -1) {
        if (lane == 0) dist[v] = -1;
        return;
    }
    DistKey best{-1, 0, 0, -1};
    eid_t b = xadj[v], e = xadj[v + 1];
    for (eid_t j = b + lane; j < e; j += G) {
        int32_t u = adjncy[j];
        DistKey k{(int32_t)(xadj[u + 1] - xadj[u]), adjwgt[j], (int32_t)(j - b), u};
        if (BetterDist()(k, best)) best = k;
    }
    best = group_best<G>(best, BetterDist());
    if (lane == 0) {
        dist[v] = best.u;
        if (best.u >= 0) atomicAdd(&hubcnt[best.u], 1);
    }
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

// Concatenated rows of both endpoints renamed by cmap; key = row << 32 | cc,
// with the row's own id replaced by cn so internal entries sort last.
template <int G>
__global__ void __launch_bounds__(kRowBS) k_gather(const int32_t *__restrict__ ids, int64_t cnt, int32_t cn,
                                                   const int32_t *__restrict__ rep, const int32_t *__restrict__ match,
                                                   const eid_t *__restrict__ xadj, const vid_t *__restrict__ adjncy,
                                                   const wgt_t *__restrict__ adjwgt, const int32_t *__restrict__ cmap,
                                                   const int64_t *__restrict__ toff, uint64_t *__restrict__ key,
                                                   uint32_t *__restrict__ val)
{
    int64_t gi = gindex<G>();
    int lane = glane<G>();
    if (gi >= cnt) return;
    int32_t c = ids[gi];
    int32_t v = rep[c], u = match[v];
    int64_t o = toff[c];
    uint64_t hi = (uint64_t)(uint32_t)c << 32;
    eid_t b = xadj[v], e = xadj[v + 1];
    for (eid_t j = b + lane; j < e; j += G) {
        int32_t cc = cmap[adjncy[j]];
        key[o + (j - b)] = hi | (uint32_t)(cc == c ? cn : cc);
        val[o + (j - b)] = (uint32_t)adjwgt[j];
    }
    if (u != v) {
        o += e - b;
        b = xadj[u];
        e = xadj[u + 1];
        for (eid_t j = b + lane; j < e; j += G) {
            int32_t cc = cmap[adjncy[j]];
            key[o + (j - b)] = hi | (uint32_t)(cc == c ? cn : cc);
            val[o + (j - b)] = (uint32_t)adjwgt[j];
        }
    }
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

__global__ void k_run_write(int64_t m, int32_t cn, const uint64_t *__restrict__ key, const int32_t *__restrict__ hsc,
                            const int32_t *__restrict__ wsc, const int32_t *__restrict__ start,
                            vid_t *__restrict__ cadjncy, wgt_t *__restrict__ cadjwgt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint64_t k = key[i];
    uint32_t cc = (uint32_t)k;
    if (cc == (uint32_t)cn) return;
    if (i + 1 < m && key[i + 1] == k) return;          // not the last of its run
    int32_t r = hsc[i] - 1;
    int32_t s = start[r];
    cadjncy[r] = (vid_t)cc;
    cadjwgt[r] = wsc[i] - (s > 0 ? wsc[s - 1] : 0);
}

__global__ void k_cxadj(int32_t cn, int64_t cm, const int64_t *__restrict__ toff, const int32_t *__restrict__ hsc,
                        eid_t *__restrict__ cxadj)
{
    int32_t c = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (c > cn) return;
    if (c == cn) { cxadj[c] = cm; return; }
    int64_t t = toff[c];
    cxadj[c] = t > 0 ? hsc[t - 1] : 0;
}

int64_t count_matched(Ctx &c, int32_t n, const int32_t *match, int64_t *d_cnt)
{
    fill<int64_t>(c, d_cnt, 1, 0);
    k_count_matched<<<grid_for(n, kBS), kBS, 0, c.stream>>>(n, match, (unsigned long long *)d_cnt);
    HGP_LAUNCH_CHECK();
    read_scalars(c, d_cnt, sizeof(int64_t));
    return c.pinned[0];
}

inline bool need_more(int64_t matched, int32_t n) { return 4 * matched < 3 * (int64_t)n; }

// Calls f(integral_constant<G>, ids, count) for every non-empty degree class;
// MaxG caps the group size (ballot-based kernels need G <= 64).
template <int MaxG = 256, class F>
void for_bins(const RowBins &b, F f)
{
    if (b.count(0)) f(std::integral_constant<int, 8>{}, b.list(0), b.count(0));
    if (b.count(1)) f(std::integral_constant<int, 64>{}, b.list(1), b.count(1));
    if (b.count(2)) f(std::integral_constant<int, MaxG < 256 ? MaxG : 256>{}, b.list(2), b.count(2));
}

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
    int32_t *progress = c.arena.ralloc<int32_t>((size_t)(rounds > 0 ? rounds : 1));
    int64_t *d_cnt = c.arena.ralloc<int64_t>(2);
    fill<int32_t>(c, match, n, -1);
    fill<int32_t>(c, progress, rounds > 0 ? rounds : 1, 0);
    unsigned gv = grid_for(n, kBS);
    for (int r = 0; r < rounds; r++) {
        const int32_t *prev = r > 0 ? progress + r - 1 : nullptr;
        {
            // per level: rows (adjncy, adjwgt) + gathers of match/vwgt per entry; xadj, match, vwgt, prop per vertex
            KScope ks(c, "match_propose", 16.0 * (double)g.m + 28.0 * n);
            for_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t cnt) {
                constexpr int G = decltype(gc)::value;
                k_propose<G><<<grid_groups(cnt, G), kRowBS, 0, c.stream>>>(ids, cnt, g.xadj, g.adjncy, g.adjwgt,
                                                                           g.vwgt, match, maxvwgt, seed, level,
                                                                           prop, prev);
                HGP_LAUNCH_CHECK();
            });
        }
        k_accept<<<gv, kBS, 0, c.stream>>>(n, prop, match, progress + r, prev);
        HGP_LAUNCH_CHECK();
    }
    int64_t matched = count_matched(c, n, match, d_cnt);
    if (need_more(matched, n)) {
        int32_t *cnt = prop;   // reuse
        fill<int32_t>(c, cnt, n, 0);
        k_leaf_count<<<gv, kBS, 0, c.stream>>>(n, g.xadj, g.adjncy, match, cnt);
        HGP_LAUNCH_CHECK();
        for_bins<64>(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
            constexpr int G = decltype(gc)::value;
            k_pair_row<G, true><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.vwgt,
                                                                            maxvwgt, nullptr, cnt, match);
            HGP_LAUNCH_CHECK();
        });
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
            k_rowsum<<<grid_for(nc * kG, kBS), kBS, 0, c.stream>>>(nc, d_cnt + 1, list, g.xadj, g.adjncy, key, val,
                                                                   rsum);
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
        int32_t *cnt = prop;
        fill<int32_t>(c, cnt, n, 0);
        for_bins(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
            constexpr int G = decltype(gc)::value;
            k_rel_dist<G><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.adjwgt, match,
                                                                       dist, cnt);
            HGP_LAUNCH_CHECK();
        });
        for_bins<64>(g.bins, [&](auto gc, const int32_t *ids, int64_t nb) {
            constexpr int G = decltype(gc)::value;
            k_pair_row<G, false><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, g.xadj, g.adjncy, g.vwgt,
                                                                             maxvwgt, dist, cnt, match);
            HGP_LAUNCH_CHECK();
        });
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
    offsets_from_counts(c, len, toff, cn);
    int64_t cm = 0;
    if (m > 0) {
        uint64_t *key = c.arena.ralloc<uint64_t>((size_t)m);
        uint32_t *val = c.arena.ralloc<uint32_t>((size_t)m);
        {
            // per fine entry: adjncy, adjwgt, cmap gather, key+val write; per coarse vertex: rep, match, toff, 2 xadj
            RowBins cb = row_bins(c, cn, nullptr, len, false);
            KScope ks(c, "contract_gather", 24.0 * (double)m + 40.0 * cn);
            for_bins(cb, [&](auto gc, const int32_t *ids, int64_t nb) {
                constexpr int G = decltype(gc)::value;
                k_gather<G><<<grid_groups(nb, G), kRowBS, 0, c.stream>>>(ids, nb, cn, rep, match, g.xadj, g.adjncy,
                                                                         g.adjwgt, cmap, toff, key, val);
                HGP_LAUNCH_CHECK();
            });
        }
        int bc = bits_for((uint64_t)cn);
        int br = bits_for((uint64_t)(cn > 0 ? cn - 1 : 0));
        radix_sort(c, key, val, m, 0, bc);
        if (br > 0) radix_sort(c, key, val, m, 32, 32 + br);
        int32_t *hsc = c.arena.ralloc<int32_t>((size_t)m);
        int32_t *wsc = c.arena.ralloc<int32_t>((size_t)m);
        int32_t *start = c.arena.ralloc<int32_t>((size_t)m);
        unsigned gm = grid_for(m, kBS);
        k_heads<<<gm, kBS, 0, c.stream>>>(m, cn, key, hsc);
        HGP_LAUNCH_CHECK();
        scan_inclusive(c, hsc, hsc, m);
        scan_inclusive(c, (const int32_t *)val, wsc, m);
        k_run_start<<<gm, kBS, 0, c.stream>>>(m, hsc, start);
        HGP_LAUNCH_CHECK();
        read_scalars(c, hsc + m - 1, sizeof(int32_t));
        cm = ((int32_t *)c.pinned)[0];
        cg.adjncy = c.arena.lalloc<vid_t>((size_t)cm);
        cg.adjwgt = c.arena.lalloc<wgt_t>((size_t)cm);
        k_run_write<<<gm, kBS, 0, c.stream>>>(m, cn, key, hsc, wsc, start, cg.adjncy, cg.adjwgt);
        HGP_LAUNCH_CHECK();
        k_cxadj<<<grid_for((int64_t)cn + 1, kBS), kBS, 0, c.stream>>>(cn, cm, toff, hsc, cg.xadj);
        HGP_LAUNCH_CHECK();
    } else {
        cg.adjncy = c.arena.lalloc<vid_t>(1);
        cg.adjwgt = c.arena.lalloc<wgt_t>(1);
        fill<int64_t>(c, cg.xadj, (int64_t)cn + 1, 0);
    }
    cg.m = cm;
    c.arena.right_release(mark);
    return cg;
}

}  // namespace hgp
