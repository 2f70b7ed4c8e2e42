// hgp_check.hip -- device-side checks of a CSR graph and the evaluation of a
// device-resident partition (edge cut, part weights).
//
// The reference checks a device graph with check_graph
// (hunyuangraph_GPU_match.h:1718-1800: vertex weights >= 1, offsets inside
// [0, m], neighbour ids in range, every edge present in both rows with equal
// weight), printing one line per fault.  Here the check returns the first
// fault (smallest vertex) of each kind to the host, in two launches: the row
// pass proves the offsets monotone before the entry pass reads any row, so a
// malformed xadj never sends a kernel outside adjncy.
//
// Entries are walked in tiles of kEPT x 64 consecutive entries, one tile per
// wave: lane l takes entries l, l + 64, ... (each load instruction reads 64
// consecutive entries).  Two lanes find the rows of the tile's ends by a
// binary search over xadj; every lane then finds the row of its first entry
// by a search between those two rows (a few cached lines) and follows the
// rows forward, so hub rows cost no more than short ones and xadj is read
// about once.
#include "hgp_check.hpp"
#include "hgp_host.hpp"

#include <algorithm>
#include <climits>
#include <mutex>

namespace hgp {

namespace {

constexpr int kChkBS = 256;
constexpr int kEPT = 16;                        // entries per lane and tile
constexpr int64_t kWaveTile = (int64_t)kWave * kEPT;
constexpr int64_t kTile = (int64_t)kChkBS * kEPT;
constexpr unsigned kEntryGrid = 2048;           // grid of the entry walkers (grid-stride)

// Row v with xadj[v] <= e < xadj[v + 1] (xadj monotone, xadj[0] = 0, e < xadj[n]).
__device__ inline int32_t row_of(const eid_t *__restrict__ xadj, int32_t n, int64_t e)
{
    int32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int32_t mid = lo + (hi - lo) / 2;
        if (xadj[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// Calls f(v, j) for this lane's entries j of the wave's tile [t0, t1)
// (t0 + lane, t0 + lane + 64, ...; row v).  Every lane of the wave must call
// it.  The row index stays in [0, n - 1], whatever xadj holds.
template <class F>
__device__ inline void walk_tile(const eid_t *__restrict__ xadj, int32_t n, int64_t t0, int64_t t1, F f)
{
    const int lane = lane_id();
    int32_t r = 0;
    if (lane == 0) r = row_of(xadj, n, t0);
    if (lane == 1) r = row_of(xadj, n, t1 - 1);
    const int32_t r0 = __shfl(r, 0, kWave), r1 = __shfl(r, 1, kWave);
    int64_t j = t0 + lane;
    if (j >= t1) return;
    int32_t lo = r0, hi = r1 + 1;   // the row of j lies in [r0, r1] when xadj is monotone
    while (hi - lo > 1) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (xadj[mid] <= j) lo = mid; else hi = mid;
    }
    int32_t v = lo;
    eid_t nx = xadj[v + 1];
    for (; j < t1; j += kWave) {
        while (nx <= j && v < n - 1) nx = xadj[++v + 1];
        f(v, j);
    }
}

template <class T>
__device__ inline T block_sum(T v)
{
    __shared__ T sh[kChkBS / kWave];
    v = wave_sum(v);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
    __syncthreads();
    T r = 0;
#pragma unroll
    for (int i = 0; i < kChkBS / kWave; i++) r += sh[i];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ graph check

__global__ void k_check_clear(CheckState *s)
{
    int i = threadIdx.x;
    if (i < kChkReasons) s->first[i] = ULLONG_MAX;
    if (i == 0) { s->wsum = 0; s->h[0] = 0; s->h[1] = 0; }
}

__global__ void __launch_bounds__(kChkBS) k_check_rows(int32_t n, const eid_t *__restrict__ xadj,
                                                      const wgt_t *__restrict__ vwgt, CheckState *s)
{
    int64_t v = (int64_t)blockIdx.x * kChkBS + threadIdx.x;
    if (v >= n) return;
    eid_t a = xadj[v], b = xadj[v + 1];
    if (v == 0 && a != 0) atomicMin(&s->first[kChkXadj0], 0ull);
    if (b < a) atomicMin(&s->first[kChkMonotone], (unsigned long long)v);
    if (vwgt && vwgt[v] < 1) atomicMin(&s->first[kChkVwgt], (unsigned long long)v);
}

// Edge (lo, hi, w) hashed twice (two SplitMix64 streams); an entry of row
// v < u adds the hash, of row v > u subtracts it, so a symmetric graph sums
// to zero in both.
__device__ inline void edge_hashes(int32_t lo, int32_t hi, int32_t w, uint64_t &h0, uint64_t &h1)
{
    uint64_t key = ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
    uint64_t a = mix64(key ^ 0x5bd1e9955bd1e995ull);
    h0 = mix64(a ^ (uint64_t)(uint32_t)w);
    h1 = mix64(mix64(key ^ 0xc2b2ae3d27d4eb4full) + (uint64_t)(uint32_t)w);
}

__global__ void __launch_bounds__(kChkBS) k_check_entries(int32_t n, int64_t m, const eid_t *__restrict__ xadj,
                                                         const vid_t *__restrict__ adjncy,
                                                         const wgt_t *__restrict__ adjwgt, int symmetric,
                                                         CheckState *s)
{
    uint64_t wsum = 0, h0 = 0, h1 = 0;
    const int64_t wave = (int64_t)blockIdx.x * (kChkBS / kWave) + threadIdx.x / kWave;
    for (int64_t t0 = wave * kWaveTile; t0 < m; t0 += (int64_t)gridDim.x * kTile) {
        const int64_t t1 = min(t0 + kWaveTile, m);
        walk_tile(xadj, n, t0, t1, [&](int32_t v, int64_t j) {
            int32_t u = adjncy[j];
            int32_t w = adjwgt ? adjwgt[j] : 1;
            if (u < 0 || u >= n) atomicMin(&s->first[kChkNeighbour], (unsigned long long)v);
            else if (u == v) atomicMin(&s->first[kChkSelfLoop], (unsigned long long)v);
            if (w < 1) atomicMin(&s->first[kChkAdjwgt], (unsigned long long)v);
            wsum += (uint64_t)(int64_t)w;
            if (symmetric && u >= 0 && u < n && u != v) {
                uint64_t a, b;
                edge_hashes(min(u, v), max(u, v), w, a, b);
                if (v < u) { h0 += a; h1 += b; } else { h0 -= a; h1 -= b; }
            }
        });
    }
    wsum = block_sum(wsum);
    h0 = block_sum(h0);
    h1 = block_sum(h1);
    if (threadIdx.x == 0) {
        if (wsum) atomicAdd(&s->wsum, (unsigned long long)wsum);
        if (h0) atomicAdd(&s->h[0], (unsigned long long)h0);
        if (h1) atomicAdd(&s->h[1], (unsigned long long)h1);
    }
}

// ------------------------------------------------------------------ evaluation

__global__ void k_eval_clear(int64_t *scr, int32_t k)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < kEvalHead + (int64_t)k;
         i += (int64_t)gridDim.x * blockDim.x)
        scr[i] = 0;
}

// Part weights: LDS bins per block, one atomic per block and touched part;
// vertices whose part id is outside [0, k) are counted instead.
__global__ void __launch_bounds__(kChkBS) k_eval_vertices(int32_t n, int32_t k, const wgt_t *__restrict__ vwgt,
                                                         const int32_t *__restrict__ part, int64_t *scr)
{
    extern __shared__ unsigned long long bins[];
    for (int i = threadIdx.x; i < k; i += kChkBS) bins[i] = 0;
    __syncthreads();
    uint64_t bad = 0;
    for (int64_t v = (int64_t)blockIdx.x * kChkBS + threadIdx.x; v < n; v += (int64_t)gridDim.x * kChkBS) {
        int32_t p = part[v];
        if (p < 0 || p >= k) bad++;
        else atomicAdd(&bins[p], (unsigned long long)(int64_t)(vwgt ? vwgt[v] : 1));
    }
    __syncthreads();
    unsigned long long *pw = (unsigned long long *)(scr + kEvalHead);
    for (int i = threadIdx.x; i < k; i += kChkBS)
        if (bins[i]) atomicAdd(&pw[i], bins[i]);
    bad = block_sum(bad);
    if (threadIdx.x == 0 && bad) atomicAdd((unsigned long long *)&scr[1], (unsigned long long)bad);
}

// Cut over the directed entries (each undirected edge twice); entries whose
// neighbour id is outside [0, n) are counted instead of gathered.
__global__ void __launch_bounds__(kChkBS) k_eval_entries(int32_t n, const eid_t *__restrict__ xadj,
                                                        const vid_t *__restrict__ adjncy,
                                                        const wgt_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ part, int64_t *scr)
{
    const int64_t m = xadj[n];
    uint64_t cut = 0, bad = 0;
    const int64_t wave = (int64_t)blockIdx.x * (kChkBS / kWave) + threadIdx.x / kWave;
    for (int64_t t0 = wave * kWaveTile; t0 < m; t0 += (int64_t)gridDim.x * kTile) {
        const int64_t t1 = min(t0 + kWaveTile, m);
        int32_t pv = -1, cv = -1;
        walk_tile(xadj, n, t0, t1, [&](int32_t v, int64_t j) {
            if (v != cv) { cv = v; pv = part[v]; }
            int32_t u = adjncy[j];
            if (u < 0 || u >= n) { bad++; return; }
            if (part[u] != pv) cut += (uint64_t)(int64_t)(adjwgt ? adjwgt[j] : 1);
        });
    }
    cut = block_sum(cut);
    bad = block_sum(bad);
    if (threadIdx.x == 0) {
        if (cut) atomicAdd((unsigned long long *)&scr[0], (unsigned long long)cut);
        if (bad) atomicAdd((unsigned long long *)&scr[2], (unsigned long long)bad);
    }
}

// out = {cut, heaviest part, bad part ids, bad neighbour ids}.
__global__ void __launch_bounds__(kChkBS) k_eval_final(int32_t k, const int64_t *scr, int64_t *out)
{
    int64_t mx = 0;
    for (int i = threadIdx.x; i < k; i += kChkBS) mx = max(mx, scr[kEvalHead + i]);
    mx = wave_max(mx);
    __shared__ int64_t sh[kChkBS / kWave];
    if (lane_id() == 0) sh[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kChkBS / kWave; i++) mx = max(mx, sh[i]);
        out[0] = scr[0] / 2;
        out[1] = mx;
        out[2] = scr[1];
        out[3] = scr[2];
    }
}

}  // namespace

void check_graph_launch(hipStream_t st, int32_t n, int64_t m, const eid_t *xadj, const vid_t *adjncy,
                        const wgt_t *vwgt, const wgt_t *adjwgt, bool symmetric, CheckState *s, bool rows)
{
    if (rows) {
        k_check_clear<<<1, 64, 0, st>>>(s);
        HGP_LAUNCH_CHECK();
        if (n > 0) {
            k_check_rows<<<grid_for(n, kChkBS), kChkBS, 0, st>>>(n, xadj, vwgt, s);
            HGP_LAUNCH_CHECK();
        }
        return;
    }
    if (m > 0) {
        unsigned g = (unsigned)std::min<int64_t>((m + kTile - 1) / kTile, kEntryGrid);
        k_check_entries<<<g, kChkBS, 0, st>>>(n, m, xadj, adjncy, adjwgt, symmetric ? 1 : 0, s);
        HGP_LAUNCH_CHECK();
    }
}

void evaluate_launch(hipStream_t st, int32_t n, const eid_t *xadj, const vid_t *adjncy, const wgt_t *adjwgt,
                     const wgt_t *vwgt, const int32_t *part, int32_t k, int64_t *scr, int64_t *out)
{
    k_eval_clear<<<grid_for(kEvalHead + (int64_t)k, kChkBS), kChkBS, 0, st>>>(scr, k);
    HGP_LAUNCH_CHECK();
    if (n > 0) {
        unsigned gv = (unsigned)std::min<int64_t>(((int64_t)n + kChkBS - 1) / kChkBS, 1024);
        k_eval_vertices<<<gv, kChkBS, sizeof(unsigned long long) * (size_t)k, st>>>(n, k, vwgt, part, scr);
        HGP_LAUNCH_CHECK();
        k_eval_entries<<<kEntryGrid, kChkBS, 0, st>>>(n, xadj, adjncy, adjwgt, part, scr);
        HGP_LAUNCH_CHECK();
    }
    k_eval_final<<<1, kChkBS, 0, st>>>(k, scr, out);
    HGP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ fixed scratch

// Evaluation scratch of the entry point that allocates nothing: kEvalHead
// accumulators and up to kEvalMaxParts part weights (32.8 KB per device,
// held until the process ends).  Host code never passes `scratch` itself to
// a kernel: its name on the host is not a device address.  The address on
// each device comes from hipGetSymbolAddress, taken once per device with
// that device current.
__device__ __align__(256) int64_t scratch[kEvalHead + kEvalMaxParts];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu[kScratchDevices];
int64_t *g_scratch_ptr[kScratchDevices] = {};
}  // namespace

void evaluate_fixed_scratch(int32_t device, int32_t n, const eid_t *xadj, const vid_t *adjncy,
                            const wgt_t *adjwgt, const wgt_t *vwgt, const int32_t *part, int32_t k,
                            int64_t *out)
{
    if (device < 0 || device >= kScratchDevices) throw Error{-1, "device ordinal out of range"};
    int count = 0;
    HGP_CHECK_HIP(hipGetDeviceCount(&count));
    if (device >= count) throw Error{-1, "device ordinal out of range"};
    // one call at a time per device: the scratch is shared by every call on it
    std::lock_guard<std::mutex> lk(g_scratch_mu[device]);
    HGP_CHECK_HIP(hipSetDevice(device));
    if (!g_scratch_ptr[device]) {
        int64_t *p = nullptr;
        HGP_CHECK_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error{-2, "no device address for the evaluation scratch"};
        g_scratch_ptr[device] = p;
    }
    // the null stream of `device`; the clear kernel zeroes the scratch, and
    // the call waits, so the next call finds no work of this one in flight
    evaluate_launch(nullptr, n, xadj, adjncy, adjwgt, vwgt, part, k, g_scratch_ptr[device], out);
    HGP_CHECK_HIP(hipStreamSynchronize(nullptr));
}

namespace {
const char *const kReasonText[kChkReasons] = {
    "xadj[0] != 0",
    "xadj is not monotone at vertex ",
    "vertex weight < 1 at vertex ",
    "neighbour id out of range in row ",
    "self-loop at vertex ",
    "edge weight < 1 in row ",
};
}  // namespace

void check_graph_device(Ctx &c, int32_t n, const eid_t *xadj, const vid_t *adjncy, const wgt_t *vwgt,
                        const wgt_t *adjwgt, bool symmetric)
{
    if (n < 0) throw Error{-1, "nvtxs < 0"};
    if (n == 0) return;
    if (!xadj) throw Error{-1, "xadj is NULL"};
    int64_t m = 0;
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned, xadj + n, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    m = c.pinned[0];
    if (m < 0) throw Error{-1, "xadj[n] < 0"};
    if (m > INT32_MAX) throw Error{-4, "more than 2^31-1 directed entries"};
    if (m > 0 && !adjncy) throw Error{-1, "adjncy is NULL"};
    size_t mark = c.arena.right_mark();
    CheckState *s = c.arena.ralloc<CheckState>(1);
    auto report = [&](int lo, int hi) {
        read_scalars(c, s, sizeof(CheckState));
        const CheckState *h = (const CheckState *)c.pinned;
        for (int r = lo; r < hi; r++)
            if (h->first[r] != ULLONG_MAX)
                throw Error{-1, r == kChkXadj0 ? std::string(kReasonText[r])
                                               : kReasonText[r] + std::to_string(h->first[r])};
        return *h;
    };
    check_graph_launch(c.stream, n, m, xadj, adjncy, vwgt, adjwgt, symmetric, s, true);
    report(0, kChkNeighbour);
    check_graph_launch(c.stream, n, m, xadj, adjncy, vwgt, adjwgt, symmetric, s, false);
    CheckState h = report(kChkNeighbour, kChkReasons);
    c.arena.right_release(mark);
    if (h.wsum > (unsigned long long)INT32_MAX) throw Error{-4, "total edge weight exceeds 2^31-1"};
    if (symmetric && (h.h[0] != 0 || h.h[1] != 0)) throw Error{-1, "graph is not symmetric"};
}

}  // namespace hgp
