// hgp_initpart.hip -- recursive-bisection initial partition of the coarsest
// graph (hunyuangraph_gpu_initialpartition / _RecursiveBisection,
// hunyuangraph_GPU_initialpartition.h:3913-4496).
//
// The recursion is run depth by depth.  All bisection nodes of one depth live
// in one "forest": every vertex of the coarsest graph sits in the contiguous
// range of the tree node that owns it, rows hold forest positions, and edges
// never cross nodes.  One launch runs every trial of every node of the depth
// (one wave per GGGP + FM trial), so k-1 bisections take ceil(log2 k) launches
// and no host round-trips.  Tree node ids are heap ids (root 1, children 2i
// and 2i+1), which also key the trial random streams.
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"

#include <cmath>

namespace hgp {

namespace {

constexpr int kTrialBytesPerVertex = 14;   // where, moved (u8) + ed, id, queue (i32)
// heap FM: + position index and two heaps of (key, vertex), n + 1 slots each (i32)
constexpr int kHeapBytesPerVertex = 20;
constexpr int kLdsBudget = 60 * 1024;
// fm_kind -1 keeps the scan: on the dense coarsest graphs this pipeline
// produces (R-MAT 2^22: 995 vertices, average degree ~300) the queues took
// 188 ms against 7 ms, a move's row being ~300 serial queue updates
constexpr int32_t kAutoHeapMinVertices = INT32_MAX;

__host__ __device__ inline int64_t trial_bytes(int64_t n, bool heap)
{
    int64_t b = ((2 * n + 3) & ~3ll) + 12 * n;
    if (heap) b += 4 * n + 16 * (n + 1);
    return b;
}

struct Node {
    int32_t start, end, k, fpart;
    int64_t tv, onemin, onemax, ideal0;
    int32_t T, trial_off, best, cnt0;
    int64_t woff;       // offset of the node's trial-where block (bytes)
    int64_t soff;       // offset of the node's global scratch block (bytes)
};

__global__ void k_ip_init(int32_t n, const wgt_t *__restrict__ vwgt, int32_t *__restrict__ node_of,
                          int32_t *__restrict__ label, wgt_t *__restrict__ fvw, Node *nodes)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) {
        node_of[v] = 1;
        label[v] = v;
        fvw[v] = vwgt[v];
    }
    if (v == 0) {
        nodes[1].start = 0;
        nodes[1].end = n;
    }
}

// Per active node: total weight and the bisection bounds
// (:3986-3987 with tp1 = kr/k, ub = 1.03^(1/ln k) computed on the host).
__global__ void __launch_bounds__(256) k_ip_setup(const int32_t *__restrict__ act, Node *nodes,
                                                  const wgt_t *__restrict__ fvw, double ub, int32_t trials)
{
    __shared__ int64_t sh[4];
    Node &nd = nodes[act[blockIdx.x]];
    int64_t s = 0;
    for (int32_t v = nd.start + threadIdx.x; v < nd.end; v += blockDim.x) s += fvw[v];
    s = wave_sum(s);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t tv = sh[0] + sh[1] + sh[2] + sh[3];
        int32_t k = nd.k, kl = k >> 1, kr = k - kl;
        double t1 = ((double)tv * (double)kr) / (double)k;
        nd.tv = tv;
        nd.onemax = (int64_t)floor(ub * t1);
        nd.onemin = (int64_t)floor(t1 / ub);
        nd.ideal0 = (tv * kl) / k;
        int32_t nn = nd.end - nd.start;
        nd.T = trials < nn ? trials : nn;
        nd.best = -1;
    }
}

// Exclusive offsets of trials and scratch over the active nodes (one block).
__global__ void k_ip_offsets(const int32_t *__restrict__ act, int32_t nact, Node *nodes, int32_t lds_cap,
                             bool heap, int32_t *total)
{
    if (threadIdx.x != 0) return;
    int32_t t = 0;
    int64_t w = 0, s = 0;
    for (int32_t a = 0; a < nact; a++) {
        Node &nd = nodes[act[a]];
        int64_t nn = nd.end - nd.start;
        nd.trial_off = t;
        nd.woff = w;
        nd.soff = s;
        t += nd.T;
        w += (int64_t)nd.T * nn;
        if (nn > lds_cap) s += (int64_t)nd.T * ((trial_bytes(nn, heap) + 255) & ~255ll);
    }
    *total = t;
}

struct TrialMem {
    uint8_t *where, *moved;
    int32_t *ed, *id, *queue;
    int32_t *loc, *hkey[2], *hval[2];   // heap FM only
};

// ---- device priority queue (hunyuangraph_GPU_priorityqueue.h:27-225), run by
// one lane: a max-heap of vertices with a position index.  An element moves
// above another when its key is strictly greater, or, with `ties`, equal with
// a smaller vertex id (the top is then the unique (gain desc, id asc)
// maximum); a removal refills the hole from the last slot.
struct Heap {
    int32_t *key, *val, *loc;
    int32_t cnt;
    bool ties;
    __device__ bool above(int32_t ka, int32_t va, int32_t kb, int32_t vb) const
    {
        return ka > kb || (ties && ka == kb && va < vb);
    }
    __device__ void place(int32_t i, int32_t k, int32_t v)
    {
        key[i] = k;
        val[i] = v;
        loc[v] = i;
    }
    __device__ void up(int32_t i, int32_t k, int32_t v)
    {
        while (i > 0) {
            int32_t j = (i - 1) >> 1;
            int32_t kj = key[j], vj = val[j];
            if (!above(k, v, kj, vj)) break;
            place(i, kj, vj);
            i = j;
        }
        place(i, k, v);
    }
    __device__ void down(int32_t i, int32_t k, int32_t v)
    {
        int32_t j;
        while ((j = 2 * i + 1) < cnt) {
            int32_t kj = key[j], vj = val[j];
            if (j + 1 < cnt) {
                int32_t k1 = key[j + 1], v1 = val[j + 1];
                if (above(kj, vj, k, v)) {
                    if (above(k1, v1, kj, vj)) { j++; kj = k1; vj = v1; }
                } else if (above(k1, v1, k, v)) {
                    j++; kj = k1; vj = v1;
                } else {
                    break;
                }
            } else if (!above(kj, vj, k, v)) {
                break;
            }
            place(i, kj, vj);
            i = j;
        }
        place(i, k, v);
    }
    __device__ void insert(int32_t v, int32_t k) { up(cnt++, k, v); }
    __device__ int32_t top()
    {
        if (cnt == 0) return -1;
        int32_t t = val[0];
        loc[t] = -1;
        if (--cnt > 0) down(0, key[cnt], val[cnt]);
        return t;
    }
    __device__ void remove(int32_t v)
    {
        int32_t i = loc[v];
        loc[v] = -1;
        if (--cnt > 0 && val[cnt] != v) {
            int32_t nk = key[cnt], nv = val[cnt];
            if (above(nk, nv, key[i], v)) up(i, nk, nv);
            else down(i, nk, nv);
        }
    }
    __device__ void update(int32_t v, int32_t k)
    {
        int32_t i = loc[v];
        if (k > key[i]) up(i, k, v);
        else down(i, k, v);
    }
};

__device__ inline void fence_wave()
{
    __threadfence_block();
    __syncthreads();
}

// One GGGP + FM trial per wave (hunyuangraph_gpu_BFS_warp_2wayrefine_noedit[_sampling],
// :2426-3659).  Lane 0 runs the serial BFS; vertex scans, the FM arg-max
// (gain desc, id asc) and neighbour updates are spread over the 64 lanes.
// The boundary flag is derived: bnd(v) = deg(v) == 0 || ed[v] > 0, which is
// the invariant the reference's incremental updates maintain.
template <bool HEAP>
__global__ void __launch_bounds__(kWave) k_ip_trials(const int32_t *__restrict__ act, int32_t nact,
                                                     const Node *__restrict__ nodes, const int32_t *total,
                                                     const eid_t *__restrict__ fx, const vid_t *__restrict__ fadj,
                                                     const wgt_t *__restrict__ fw, const wgt_t *__restrict__ fvw,
                                                     uint64_t seed, int32_t fm_passes, int32_t fm_kind,
                                                     int32_t lds_cap,
                                                     uint8_t *__restrict__ trial_where, char *__restrict__ gscratch,
                                                     int64_t *__restrict__ tcut)
{
    extern __shared__ char lds[];
    int32_t gw = blockIdx.x;
    if (gw >= *total) return;
    int lane = threadIdx.x;
    int32_t lo = 0, hi = nact - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (nodes[act[mid]].trial_off <= gw) lo = mid; else hi = mid - 1;
    }
    const int32_t nid = act[lo];
    const Node nd = nodes[nid];
    const int32_t t = gw - nd.trial_off;
    const int32_t n = nd.end - nd.start, s0 = nd.start;
    char *base;
    if (n <= lds_cap) base = lds;
    else base = gscratch + nd.soff + (int64_t)t * ((trial_bytes(n, HEAP) + 255) & ~255ll);
    TrialMem M;
    M.where = (uint8_t *)base;
    M.moved = M.where + n;
    M.ed = (int32_t *)(base + (((int64_t)2 * n + 3) & ~3ll));
    M.id = M.ed + n;
    M.queue = M.id + n;
    if (HEAP) {
        M.loc = M.queue + n;
        M.hkey[0] = M.loc + n;
        M.hval[0] = M.hkey[0] + (n + 1);
        M.hkey[1] = M.hval[0] + (n + 1);
        M.hval[1] = M.hkey[1] + (n + 1);
    }
    const int64_t tv = nd.tv, onemin = nd.onemin, onemax = nd.onemax, ideal0 = nd.ideal0;

    for (int32_t i = lane; i < n; i += kWave) { M.where[i] = 1; M.moved[i] = 0; }
    fence_wave();

    // GGGP: BFS from the start vertex, moving vertices to side 0 in BFS order.
    // Every lane runs the same control flow on the same values; the rows of
    // dequeued vertices are scanned 64 entries at a time and the unvisited
    // neighbours appended in row order (ballot ranks), as a serial scan would.
    int64_t pw0 = 0, pw1 = tv;
    {
        int32_t draw = 0, start;
        if (nd.T >= n) {
            start = t;
        } else {
            int32_t len = n / nd.T, rem = n % nd.T, left;
            if (t < rem) { len++; left = t * len; }
            else left = rem * (len + 1) + (t - rem) * len;
            start = left + (int32_t)(trial_rng(seed, (uint32_t)nid, t, draw++) % (uint64_t)len);
        }
        int32_t first = 0, last = 1, nleft = n - 1, drain = 0;
        if (lane == 0) {
            M.queue[0] = start;
            M.moved[start] = 1;
        }
        fence_wave();
        for (;;) {
            if (first == last) {
                if (nleft == 0 || drain) break;
                // restart from the r-th unvisited vertex in id order
                int32_t r = (int32_t)(trial_rng(seed, (uint32_t)nid, t, draw++) % (uint64_t)nleft);
                int32_t v = -1;
                for (int32_t b0 = 0; b0 < n; b0 += kWave) {
                    int32_t u = b0 + lane;
                    bool un = u < n && M.moved[u] == 0;
                    uint64_t bal = __ballot(un);
                    int32_t cnt = __popcll(bal);
                    if (r < cnt) {
                        uint64_t pick = __ballot(un && __popcll(bal & lanemask_lt()) == r);
                        v = b0 + (__ffsll((long long)pick) - 1);
                        break;
                    }
                    r -= cnt;
                }
                if (lane == 0) {
                    M.queue[0] = v;
                    M.moved[v] = 1;
                }
                fence_wave();
                first = 0;
                last = 1;
                nleft--;
            }
            int32_t v = M.queue[first++];
            int32_t wv = fvw[s0 + v];
            if (pw0 > 0 && pw1 - wv < onemin) { drain = 1; continue; }
            if (lane == 0) M.where[v] = 0;
            pw0 += wv;
            pw1 -= wv;
            if (pw1 <= onemax) break;
            drain = 0;
            // once every vertex is queued a row has nothing fresh to append
            eid_t rb = fx[s0 + v], re = nleft > 0 ? fx[s0 + v + 1] : rb;
            for (eid_t j0 = rb; j0 < re; j0 += kWave) {
                eid_t j = j0 + lane;
                int32_t x = j < re ? fadj[j] - s0 : -1;
                bool fresh = x >= 0 && M.moved[x] == 0;
                uint64_t bal = __ballot(fresh);
                if (fresh) {
                    M.queue[last + __popcll(bal & lanemask_lt())] = x;
                    M.moved[x] = 1;
                }
                last += __popcll(bal);
                nleft -= __popcll(bal);
            }
            fence_wave();
        }
    }
    fence_wave();

    // ed / id and the cut (compute_v_ed_id_bnd, :23-62).  From here on
    // moved[v] holds bit 0 = moved in this pass, bit 1 = v has no edges, so
    // the FM arg-max scan reads only this trial's arrays.
    int64_t cutacc = 0;
    for (int32_t v = lane; v < n; v += kWave) {
        int32_t e = 0, in = 0;
        uint8_t me = M.where[v];
        eid_t rb = fx[s0 + v], re = fx[s0 + v + 1];
        eid_t j = rb;
        for (; j + 8 <= re; j += 8) {   // eight entries' loads in flight
            int32_t x[8], w[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                x[t] = fadj[j + t];
                w[t] = fw[j + t];
            }
#pragma unroll
            for (int t = 0; t < 8; t++) {
                if (M.where[x[t] - s0] != me) e += w[t]; else in += w[t];
            }
        }
        for (; j < re; j++) {
            if (M.where[fadj[j] - s0] != me) e += fw[j]; else in += fw[j];
        }
        M.ed[v] = e;
        M.id[v] = in;
        M.moved[v] = rb == re ? 2 : 0;
        cutacc += e;
    }
    int64_t cut = wave_sum(cutacc) / 2;
    fence_wave();

    int64_t pw[2] = {pw0, pw1};
    const int64_t ideal[2] = {ideal0, tv - ideal0};
    int64_t limit = n / 100;
    if (limit < 15) limit = 15;
    if (limit > 100) limit = 100;
    int32_t *swaps = M.queue;

    if (HEAP) {
        // FM driven by two priority queues (FM_2way_cut_refinement,
        // hunyuangraph_GPU_initialpartition.h:160-300).  Lane 0 runs every
        // queue operation in the serial order; the lanes share the gain
        // updates of a move's row.  fm_kind 1: (gain desc, id asc) keys and
        // the part bounds of the scan FM below, so the moves are the same;
        // fm_kind 2: the reference's queue order and balance rule
        // (|ideal0 - pw0| <= origdiff + avgvwgt, weights applied before the
        // test as hunyuangraph_CPU_2wayrefine.h:133-145 does).
        Heap q[2];
        for (int s = 0; s < 2; s++) {
            q[s].key = M.hkey[s];
            q[s].val = M.hval[s];
            q[s].loc = M.loc;
            q[s].cnt = 0;
            q[s].ties = fm_kind == 1;
        }
        const bool live_rule = fm_kind == 1;
        const int64_t a20 = (pw[0] + pw[1]) / 20, a2n = 2 * (pw[0] + pw[1]) / n;
        const int64_t avgvwgt = a20 < a2n ? a20 : a2n;
        const int64_t origdiff = ideal[0] - pw[0] < 0 ? pw[0] - ideal[0] : ideal[0] - pw[0];
        for (int32_t pass = 0; pass < fm_passes; pass++) {
            int64_t mincutorder = -1;
            int64_t newcut = cut, mincut = cut, initcut = cut;
            int64_t d0 = ideal[0] - pw[0];
            int64_t mindiff = d0 < 0 ? -d0 : d0;
            if (lane == 0) {
                q[0].cnt = q[1].cnt = 0;
                for (int32_t v = 0; v < n; v++) {
                    int32_t e = M.ed[v];
                    if (e > 0 || (M.moved[v] & 2)) q[M.where[v]].insert(v, e - M.id[v]);
                }
            }
            int32_t nswaps;
            for (nswaps = 0; nswaps < n; nswaps++) {
                int from = (ideal[0] - pw[0] < ideal[1] - pw[1]) ? 0 : 1;
                int to = from ^ 1;
                int32_t bv = -1, bg = 0;
                if (lane == 0) {
                    bv = q[from].top();
                    if (bv >= 0) bg = M.ed[bv] - M.id[bv];
                }
                bv = __shfl(bv, 0, kWave);
                bg = __shfl(bg, 0, kWave);
                if (bv == -1) break;
                int32_t wv = fvw[s0 + bv];
                newcut -= bg;
                bool accept;
                if (live_rule) {
                    int64_t dd = ideal[0] - pw[0];
                    dd = dd < 0 ? -dd : dd;
                    accept = (newcut < mincut && pw[to] + wv <= onemax && pw[from] - wv <= onemax) ||
                             (newcut == mincut && dd < mindiff);
                    if (accept) mindiff = dd;
                    pw[to] += wv;
                    pw[from] -= wv;
                } else {
                    pw[to] += wv;
                    pw[from] -= wv;
                    int64_t dd = ideal[0] - pw[0];
                    dd = dd < 0 ? -dd : dd;
                    accept = (newcut < mincut && dd <= origdiff + avgvwgt) || (newcut == mincut && dd < mindiff);
                    if (accept) mindiff = dd;
                }
                if (accept) {
                    mincut = newcut;
                    mincutorder = nswaps;
                } else if (nswaps - mincutorder > limit) {
                    newcut += bg;
                    pw[from] += wv;
                    pw[to] -= wv;
                    break;
                }
                if (lane == 0) {
                    M.where[bv] = (uint8_t)to;
                    M.moved[bv] |= 1;
                    swaps[nswaps] = bv;
                    int32_t tmp = M.ed[bv];
                    M.ed[bv] = M.id[bv];
                    M.id[bv] = tmp;
                }
                fence_wave();
                const eid_t rb = fx[s0 + bv], re = fx[s0 + bv + 1];
                for (eid_t j = rb + lane; j < re; j += kWave) {
                    int32_t x = fadj[j] - s0;
                    int32_t kw = (M.where[x] == from) ? -fw[j] : fw[j];
                    M.id[x] += kw;     // rows hold no duplicates: one lane per neighbour
                    M.ed[x] -= kw;
                }
                fence_wave();
                if (lane == 0) {
                    // queue upkeep in row order: a neighbour that left the
                    // boundary is deleted, one that stays is re-keyed, one
                    // that joined is inserted (moved vertices stay out)
                    for (eid_t j = rb; j < re; j++) {
                        int32_t x = fadj[j] - s0;
                        uint8_t st = M.moved[x];
                        if (st & 1) continue;
                        int32_t e = M.ed[x];
                        int32_t kw = (M.where[x] == from) ? -fw[j] : fw[j];
                        bool was = e + kw > 0;
                        int sx = M.where[x];
                        if (was) {
                            if (e == 0) q[sx].remove(x);
                            else q[sx].update(x, e - M.id[x]);
                        } else if (e > 0) {
                            q[sx].insert(x, e - M.id[x]);
                        }
                    }
                }
                fence_wave();
            }
            for (int32_t i = lane; i < nswaps; i += kWave) M.moved[swaps[i]] &= 2;
            fence_wave();
            for (nswaps--; nswaps > mincutorder; nswaps--) {
                int32_t v = swaps[nswaps];
                int from = M.where[v];
                int to = from ^ 1;
                int32_t wv = fvw[s0 + v];
                fence_wave();
                if (lane == 0) {
                    M.where[v] = (uint8_t)to;
                    int32_t tmp = M.ed[v];
                    M.ed[v] = M.id[v];
                    M.id[v] = tmp;
                }
                pw[to] += wv;
                pw[from] -= wv;
                fence_wave();
                for (eid_t j = fx[s0 + v] + lane; j < fx[s0 + v + 1]; j += kWave) {
                    int32_t x = fadj[j] - s0;
                    int32_t kw = (M.where[x] == from) ? -fw[j] : fw[j];
                    M.id[x] += kw;
                    M.ed[x] -= kw;
                }
                fence_wave();
            }
            cut = mincut;
            if (mincutorder <= 0 || mincut == initcut) break;
        }
    }

    for (int32_t pass = 0; pass < (HEAP ? 0 : fm_passes); pass++) {
        int64_t mincutorder = -1;
        int64_t newcut = cut, mincut = cut, initcut = cut;
        int64_t d0 = ideal[0] - pw[0];
        int64_t mindiff = d0 < 0 ? -d0 : d0;
        int32_t nswaps;
        for (nswaps = 0; nswaps < n; nswaps++) {
            int from = (ideal[0] - pw[0] < ideal[1] - pw[1]) ? 0 : 1;
            int to = from ^ 1;
            int32_t bv = -1, bg = 0;
            for (int32_t u = lane; u < n; u += kWave) {
                uint8_t st = M.moved[u];
                if (M.where[u] != from || (st & 1)) continue;
                int32_t e = M.ed[u];
                bool bnd = e > 0 || (st & 2);
                if (!bnd) continue;
                int32_t g = e - M.id[u];
                if (bv == -1 || g > bg) { bv = u; bg = g; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                int32_t ov = __shfl_xor(bv, o, kWave), og = __shfl_xor(bg, o, kWave);
                if (ov != -1 && (bv == -1 || og > bg || (og == bg && ov < bv))) { bv = ov; bg = og; }
            }
            if (bv == -1) break;
            int32_t wv = fvw[s0 + bv];
            newcut -= bg;
            int64_t dd = ideal[0] - pw[0];
            dd = dd < 0 ? -dd : dd;
            if ((newcut < mincut && pw[to] + wv <= onemax && pw[from] - wv <= onemax) ||
                (newcut == mincut && dd < mindiff)) {
                mincut = newcut;
                mindiff = dd;
                mincutorder = nswaps;
            } else if (nswaps - mincutorder > limit) {
                newcut += bg;
                break;
            }
            pw[to] += wv;
            pw[from] -= wv;
            if (lane == 0) {
                M.where[bv] = (uint8_t)to;
                M.moved[bv] |= 1;
                swaps[nswaps] = bv;
                int32_t tmp = M.ed[bv];
                M.ed[bv] = M.id[bv];
                M.id[bv] = tmp;
            }
            fence_wave();
            for (eid_t j = fx[s0 + bv] + lane; j < fx[s0 + bv + 1]; j += kWave) {
                int32_t x = fadj[j] - s0;
                int32_t kw = (M.where[x] == from) ? -fw[j] : fw[j];
                atomicAdd(&M.id[x], kw);
                atomicAdd(&M.ed[x], -kw);
            }
            fence_wave();
        }
        for (int32_t i = lane; i < nswaps; i += kWave) M.moved[swaps[i]] &= 2;
        fence_wave();
        for (nswaps--; nswaps > mincutorder; nswaps--) {
            int32_t v = swaps[nswaps];
            int from = M.where[v];
            int to = from ^ 1;
            int32_t wv = fvw[s0 + v];
            fence_wave();
            if (lane == 0) {
                M.where[v] = (uint8_t)to;
                int32_t tmp = M.ed[v];
                M.ed[v] = M.id[v];
                M.id[v] = tmp;
            }
            pw[to] += wv;
            pw[from] -= wv;
            fence_wave();
            for (eid_t j = fx[s0 + v] + lane; j < fx[s0 + v + 1]; j += kWave) {
                int32_t x = fadj[j] - s0;
                int32_t kw = (M.where[x] == from) ? -fw[j] : fw[j];
                atomicAdd(&M.id[x], kw);
                atomicAdd(&M.ed[x], -kw);
            }
            fence_wave();
        }
        cut = mincut;
        if (mincutorder <= 0 || mincut == initcut) break;
    }

    uint8_t *out = trial_where + nd.woff + (int64_t)t * n;
    for (int32_t i = lane; i < n; i += kWave) out[i] = M.where[i];
    if (lane == 0) tcut[gw] = cut;
}

// Arg-min cut over a node's trials, first minimum wins (:3713-3812).
__global__ void __launch_bounds__(256) k_ip_select(const int32_t *__restrict__ act, Node *nodes,
                                                   const int64_t *__restrict__ tcut)
{
    __shared__ int64_t sc[256];
    __shared__ int32_t si[256];
    Node &nd = nodes[act[blockIdx.x]];
    int64_t bc = INT64_MAX;
    int32_t bi = -1;
    for (int32_t t = threadIdx.x; t < nd.T; t += blockDim.x) {
        int64_t c = tcut[nd.trial_off + t];
        if (c < bc) { bc = c; bi = t; }
    }
    sc[threadIdx.x] = bc;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            int64_t oc = sc[threadIdx.x + s];
            int32_t oi = si[threadIdx.x + s];
            if (oi != -1 && (si[threadIdx.x] == -1 || oc < sc[threadIdx.x] ||
                             (oc == sc[threadIdx.x] && oi < si[threadIdx.x]))) {
                sc[threadIdx.x] = oc;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) nd.best = si[0];
}

// Write the winning bisection through the labels (:3896-3902) and record
// each vertex's side for the split (-1 for vertices of finished nodes).
__global__ void k_ip_side(int32_t n, const int32_t *__restrict__ node_of, const Node *__restrict__ nodes,
                          const uint8_t *__restrict__ trial_where, const int32_t *__restrict__ label,
                          int32_t *__restrict__ part, int8_t *__restrict__ side)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    const Node &nd = nodes[node_of[v]];
    if (nd.k < 2 || nd.best < 0) { side[v] = -1; return; }
    int32_t nn = nd.end - nd.start;
    uint8_t w = trial_where[nd.woff + (int64_t)nd.best * nn + (v - nd.start)];
    part[label[v]] = nd.fpart + w;
    side[v] = (int8_t)w;
}

__global__ void __launch_bounds__(256) k_ip_count0(const int32_t *__restrict__ act, Node *nodes, int32_t ntab,
                                                   const int8_t *__restrict__ side)
{
    __shared__ int32_t sh[4];
    int32_t id = act[blockIdx.x];
    Node &nd = nodes[id];
    int32_t s = 0;
    for (int32_t v = nd.start + threadIdx.x; v < nd.end; v += blockDim.x) s += side[v] == 0;
    s = wave_sum(s);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t c0 = sh[0] + sh[1] + sh[2] + sh[3];
        nd.cnt0 = c0;
        if (2 * id + 1 < ntab) {
            nodes[2 * id].start = nd.start;
            nodes[2 * id].end = nd.start + c0;
            nodes[2 * id + 1].start = nd.start + c0;
            nodes[2 * id + 1].end = nd.end;
        }
    }
}

__global__ void k_ip_f0(int32_t n, const int8_t *__restrict__ side, int32_t *__restrict__ f0)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v < n) f0[v] = side[v] == 0;
}

// Order-preserving split (hunyuangraph_GPU_splitgraph.h:802-944): side-0
// vertices first, then side-1, each in their old order, inside the node's
// range.  Rows keep only same-side entries; finished nodes drop their rows.
__global__ void k_ip_pos(int32_t n, const int32_t *__restrict__ node_of, const Node *__restrict__ nodes,
                         const int8_t *__restrict__ side, const int32_t *__restrict__ s0, const eid_t *__restrict__ fx,
                         const vid_t *__restrict__ fadj, int32_t *__restrict__ pos, int32_t *__restrict__ nnode_of,
                         int32_t *__restrict__ nlabel, const int32_t *__restrict__ label, wgt_t *__restrict__ nfvw,
                         const wgt_t *__restrict__ fvw, int64_t *__restrict__ ndeg)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int32_t id = node_of[v];
    int8_t sd = side[v];
    int32_t p;
    int64_t d = 0;
    if (sd < 0) {
        p = v;
        nnode_of[p] = id;
    } else {
        const Node &nd = nodes[id];
        int32_t before = nd.start > 0 ? s0[nd.start - 1] : 0;
        int32_t rank0 = (s0[v] - (sd == 0 ? 1 : 0)) - before;
        int32_t rank1 = (v - nd.start) - rank0;
        p = sd == 0 ? nd.start + rank0 : nd.start + nd.cnt0 + rank1;
        nnode_of[p] = 2 * id + sd;
        eid_t j = fx[v];
        const eid_t je = fx[v + 1];
        for (; j + 8 <= je; j += 8) {   // eight entries' loads in flight
            int32_t x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) x[t] = fadj[j + t];
#pragma unroll
            for (int t = 0; t < 8; t++) d += side[x[t]] == sd;
        }
        for (; j < je; j++) d += side[fadj[j]] == sd;
    }
    pos[v] = p;
    nlabel[p] = label[v];
    nfvw[p] = fvw[v];
    ndeg[p] = d;
}

__global__ void k_ip_edges(int32_t n, const int8_t *__restrict__ side, const int32_t *__restrict__ pos,
                           const eid_t *__restrict__ fx, const vid_t *__restrict__ fadj, const wgt_t *__restrict__ fw,
                           const eid_t *__restrict__ nfx, vid_t *__restrict__ nfadj, wgt_t *__restrict__ nfw)
{
    int32_t v = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= n) return;
    int8_t sd = side[v];
    if (sd < 0) return;
    eid_t o = nfx[pos[v]];
    eid_t j = fx[v];
    const eid_t je = fx[v + 1];
    for (; j + 8 <= je; j += 8) {   // eight entries' loads in flight, written in row order
        int32_t x[8], w[8], sx[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            x[t] = fadj[j + t];
            w[t] = fw[j + t];
        }
#pragma unroll
        for (int t = 0; t < 8; t++) sx[t] = side[x[t]] == sd ? pos[x[t]] : -1;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (sx[t] < 0) continue;
            nfadj[o] = sx[t];
            nfw[o] = w[t];
            o++;
        }
    }
    for (; j < je; j++) {
        int32_t x = fadj[j];
        if (side[x] != sd) continue;
        nfadj[o] = pos[x];
        nfw[o] = fw[j];
        o++;
    }
}

struct TreeNode { int32_t id, k, fpart; };

// fm_kind -1 picks the queue-driven FM (the same moves as the scan) once the
// coarsest graph is large enough for O(log n) queue steps to beat an O(n/64)
// scan per move.
bool fm_uses_heap(int32_t fm_kind, int32_t n)
{
    if (fm_kind >= 0) return fm_kind != 0;
    return n >= kAutoHeapMinVertices;
}

}  // namespace

void initial_partition(Ctx &c, const DevGraph &g, int32_t k, const Params &p, int32_t *where)
{
    int32_t n = g.n;
    if (n <= 0) return;
    fill<int32_t>(c, where, n, 0);
    if (k <= 1) return;

    // tree shape depends on k only
    std::vector<std::vector<TreeNode>> depth_nodes;
    std::vector<TreeNode> cur{{1, k, 0}};
    int32_t maxid = 1;
    while (!cur.empty()) {
        depth_nodes.push_back(cur);
        std::vector<TreeNode> nxt;
        for (auto &t : cur) {
            int32_t kl = t.k >> 1, kr = t.k - kl;
            maxid = std::max(maxid, 2 * t.id + 1);
            if (t.k > 2) {
                if (kl >= 2) nxt.push_back({2 * t.id, kl, t.fpart});
                if (kr >= 2) nxt.push_back({2 * t.id + 1, kr, t.fpart + kl});
            }
        }
        cur.swap(nxt);
    }
    int32_t ntab = maxid + 1;
    std::vector<Node> htab(ntab);
    for (auto &t : htab) { t.start = t.end = 0; t.k = 1; t.fpart = 0; t.best = -1; t.T = 0; }
    {
        // k and fpart of every tree node (leaves included)
        std::vector<TreeNode> stack{{1, k, 0}};
        while (!stack.empty()) {
            TreeNode t = stack.back();
            stack.pop_back();
            if (t.id >= ntab) continue;
            htab[t.id].k = t.k;
            htab[t.id].fpart = t.fpart;
            if (t.k >= 2) {
                stack.push_back({2 * t.id, t.k >> 1, t.fpart});
                stack.push_back({2 * t.id + 1, t.k - (t.k >> 1), t.fpart + (t.k >> 1)});
            }
        }
    }

    double ub = std::pow((double)p.ub_num / (double)p.ub_den, 1.0 / std::log((double)k));
    const bool heap = fm_uses_heap(p.fm_kind, n);
    const int32_t fm_kind = p.fm_kind < 0 ? (heap ? 1 : 0) : p.fm_kind;
    int32_t lds_cap = (kLdsBudget - 64) / (kTrialBytesPerVertex + (heap ? kHeapBytesPerVertex : 0));
    int32_t lds_n = std::min(n, lds_cap);
    size_t lds_bytes = (size_t)trial_bytes(lds_n, heap) + 16;

    size_t mark = c.arena.right_mark();
    Node *nodes = c.arena.ralloc<Node>((size_t)ntab);
    // active node ids of every depth, back to back
    std::vector<int32_t> hact;
    std::vector<size_t> act_off;
    for (auto &dn : depth_nodes) {
        act_off.push_back(hact.size());
        for (auto &t : dn) hact.push_back(t.id);
    }
    int32_t *act_all = c.arena.ralloc<int32_t>(hact.size());
    upload_staged(c, nodes, htab.data(), sizeof(Node) * (size_t)ntab);
    upload_staged(c, act_all, hact.data(), sizeof(int32_t) * hact.size());
    int32_t *total = c.arena.ralloc<int32_t>(1);
    int64_t maxtrials = (int64_t)std::min(p.trials, n);
    uint8_t *trial_where = c.arena.ralloc<uint8_t>((size_t)(maxtrials * n));
    char *gscratch = nullptr;
    if (n > lds_cap) {
        // sum over nodes of T * round256(trial_bytes(nn)): trial_bytes is affine
        // in nn with a constant below 20, and at most n / lds_cap + 1 nodes
        // use global scratch
        size_t gs = (size_t)maxtrials * ((size_t)trial_bytes(n, heap) + 512 * ((size_t)(n / lds_cap) + 1));
        gscratch = c.arena.ralloc<char>(gs);
    }
    // trials of one depth: sum of min(trials, n_node) <= n
    int64_t *tcut = c.arena.ralloc<int64_t>((size_t)n + 1);
    int32_t *node_of[2], *label[2];
    wgt_t *fvw[2], *fw[2];
    eid_t *fx[2];
    vid_t *fadj[2];
    int64_t mm = g.m > 0 ? g.m : 1;
    for (int b = 0; b < 2; b++) {
        node_of[b] = c.arena.ralloc<int32_t>((size_t)n);
        label[b] = c.arena.ralloc<int32_t>((size_t)n);
        fvw[b] = c.arena.ralloc<wgt_t>((size_t)n);
        fx[b] = c.arena.ralloc<eid_t>((size_t)n + 1);
        fadj[b] = c.arena.ralloc<vid_t>((size_t)mm);
        fw[b] = c.arena.ralloc<wgt_t>((size_t)mm);
    }
    int8_t *side = c.arena.ralloc<int8_t>((size_t)n);
    int32_t *f0 = c.arena.ralloc<int32_t>((size_t)n);
    int32_t *pos = c.arena.ralloc<int32_t>((size_t)n);
    int64_t *ndeg = c.arena.ralloc<int64_t>((size_t)n);

    unsigned gv = grid_for(n, 256);
    k_ip_init<<<gv, 256, 0, c.stream>>>(n, g.vwgt, node_of[0], label[0], fvw[0], nodes);
    HGP_LAUNCH_CHECK();
    HGP_CHECK_HIP(hipMemcpyAsync(fx[0], g.xadj, sizeof(eid_t) * ((size_t)n + 1), hipMemcpyDeviceToDevice, c.stream));
    if (g.m > 0) {
        HGP_CHECK_HIP(hipMemcpyAsync(fadj[0], g.adjncy, sizeof(vid_t) * (size_t)g.m, hipMemcpyDeviceToDevice, c.stream));
        HGP_CHECK_HIP(hipMemcpyAsync(fw[0], g.adjwgt, sizeof(wgt_t) * (size_t)g.m, hipMemcpyDeviceToDevice, c.stream));
    }

    int cb = 0;
    for (size_t d = 0; d < depth_nodes.size(); d++) {
        const auto &dn = depth_nodes[d];
        int32_t nact = (int32_t)dn.size();
        const int32_t *act = act_all + act_off[d];
        k_ip_setup<<<nact, 256, 0, c.stream>>>(act, nodes, fvw[cb], ub, p.trials);
        HGP_LAUNCH_CHECK();
        k_ip_offsets<<<1, 64, 0, c.stream>>>(act, nact, nodes, lds_cap, heap, total);
        HGP_LAUNCH_CHECK();
        int64_t waves = std::min<int64_t>((int64_t)nact * p.trials, n);
        KScope ks(c, "k_ip_trials", 0.0);
        auto trials_kernel = heap ? k_ip_trials<true> : k_ip_trials<false>;
        trials_kernel<<<(unsigned)std::max<int64_t>(waves, 1), kWave, lds_bytes, c.stream>>>(
            act, nact, nodes, total, fx[cb], fadj[cb], fw[cb], fvw[cb], p.seed, p.fm_passes, fm_kind, lds_cap,
            trial_where, gscratch, tcut);
        HGP_LAUNCH_CHECK();
        k_ip_select<<<nact, 256, 0, c.stream>>>(act, nodes, tcut);
        HGP_LAUNCH_CHECK();
        k_ip_side<<<gv, 256, 0, c.stream>>>(n, node_of[cb], nodes, trial_where, label[cb], where, side);
        HGP_LAUNCH_CHECK();
        if (d + 1 == depth_nodes.size()) break;
        k_ip_count0<<<nact, 256, 0, c.stream>>>(act, nodes, ntab, side);
        HGP_LAUNCH_CHECK();
        k_ip_f0<<<gv, 256, 0, c.stream>>>(n, side, f0);
        HGP_LAUNCH_CHECK();
        scan_inclusive(c, f0, f0, n);
        int nb = cb ^ 1;
        k_ip_pos<<<gv, 256, 0, c.stream>>>(n, node_of[cb], nodes, side, f0, fx[cb], fadj[cb], pos, node_of[nb],
                                           label[nb], label[cb], fvw[nb], fvw[cb], ndeg);
        HGP_LAUNCH_CHECK();
        offsets_from_counts(c, ndeg, fx[nb], n);
        k_ip_edges<<<gv, 256, 0, c.stream>>>(n, side, pos, fx[cb], fadj[cb], fw[cb], fx[nb], fadj[nb], fw[nb]);
        HGP_LAUNCH_CHECK();
        cb = nb;
    }
    c.arena.right_release(mark);
}

}  // namespace hgp
