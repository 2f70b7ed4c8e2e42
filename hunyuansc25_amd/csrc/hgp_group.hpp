// hgp_group.hpp -- lane groups of G threads working on one CSR row.
//
// Row-parallel kernels take their vertices from degree classes (RowBins):
// rows of <= 8 entries get 8-lane groups, rows of <= 512 entries one wave,
// longer rows one block per segment.  Group reductions are xor-shuffles
// inside a wave for G <= 64 and go through LDS across the block's waves for
// a whole block (G > 64).  Every reduction here uses a strict total order (or an integer
// sum), so results do not depend on which lane saw which entry.
#pragma once

#include "hgp_common.hpp"

namespace hgp {

constexpr int kRowBS = 256;             // block size of row-parallel kernels
// Degree classes of the rows of a level: 0 = <= 8 entries (8 lanes per row),
// 1 = <= 32 (32 lanes), 2 = <= 512 (one wave), 3 = longer (segments of
// kSegLen entries, a block of kSegBlk threads each).
constexpr int64_t kBinSmall = 8;
constexpr int64_t kBinHalf = 32;
constexpr int64_t kBinMid = 512;
constexpr int kClsLong = 3;

// A level with at most Tuning::few_rows rows in classes 0-2 runs their row
// kernels as one wave-per-row launch over the three classes (consecutive in
// the id list): on such levels a row gets a wave of its own anyway, and one
// launch instead of three saves launch latency, which dominates there.

struct RowBins {
    int32_t *ids = nullptr;   // row ids grouped by class (order inside a class is arbitrary)
    int64_t off[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t ent[6] = {0, 0, 0, 0, 0, 0};   // sum of the row lengths of each class
    int64_t count(int c) const { return off[c + 1] - off[c]; }
    const int32_t *list(int c) const { return ids + off[c]; }
    int64_t short_rows() const { return off[3] - off[0]; }   // classes 0-2
    bool few() const { return short_rows() <= tuning().few_rows; }
};

// Classes of the rows of xadj (or of the lengths len, if xadj is null): class
// 0 holds lengths <= t0, 1 <= t1, 2 <= t2, 3 <= t3, 4 <= t4, 5 the rest.  ids
// is allocated on the left end of the arena when `left`, else on the right.
RowBins row_bins(Ctx &c, int32_t n, const eid_t *xadj, const int64_t *len, bool left, int64_t t0 = kBinSmall,
                 int64_t t1 = kBinHalf, int64_t t2 = kBinMid, int64_t t3 = INT64_MAX, int64_t t4 = INT64_MAX);

// Rows of class kClsLong cut into segments of at most kSegLen entries, so that a hub
// row is shared by several blocks.  A segment of a single-segment row is
// finished by the segment kernel itself; the segments of a "multi" row leave
// partial results that a second kernel combines per row.  A segment kernel's
// block has kSegBlk threads (8 entries per thread and load phase): small
// blocks keep many segments in flight per CU, which is what bounds these
// kernels (each segment is a short chain of dependent loads).
constexpr int64_t kSegLen = 2048;
constexpr int kSegBlk = 128;

struct RowSegs {
    int64_t nseg = 0;
    int32_t *row = nullptr;    // vertex of each segment
    int64_t *j0 = nullptr;     // first entry of each segment (CSR offset)
    int32_t *mi = nullptr;     // multi-row index of each segment, -1 for a single-segment row
    int32_t nmulti = 0;
    int32_t *mrow = nullptr;   // vertex of each multi row
    int64_t *mseg = nullptr;   // segments of multi row t: [mseg[2t], mseg[2t+1])
};

// Segments of the class-kClsLong rows of b (arrays on the left end of the arena).
RowSegs row_segs(Ctx &c, const eid_t *xadj, const RowBins &b);

template <int G>
__device__ inline int glane() { return (int)(threadIdx.x % G); }

template <int G>
__device__ inline int64_t gindex() { return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; }

template <class T>
__device__ inline T shfl_xor_pod(T v, int o, int width)
{
    static_assert(sizeof(T) % 4 == 0, "pod shuffle needs 4-byte multiples");
    T r;
    const int32_t *p = reinterpret_cast<const int32_t *>(&v);
    int32_t *q = reinterpret_cast<int32_t *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) q[i] = __shfl_xor(p[i], o, width);
    return r;
}

// All lanes of the group receive the best value under `better` (a strict
// total order on distinct values).
template <int G, class T, class Better>
__device__ inline T group_best(T v, Better better)
{
    if constexpr (G <= kWave) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            T u = shfl_xor_pod(v, o, G);
            if (better(u, v)) v = u;
        }
        return v;
    } else {
        __shared__ T sh[G / kWave];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            T u = shfl_xor_pod(v, o, kWave);
            if (better(u, v)) v = u;
        }
        if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
        __syncthreads();
        T r = sh[0];
#pragma unroll
        for (int i = 1; i < G / kWave; i++)
            if (better(sh[i], r)) r = sh[i];
        __syncthreads();
        return r;
    }
}

template <int G, class T>
__device__ inline T group_sum(T v)
{
    if constexpr (G <= kWave) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
        return v;
    } else {
        __shared__ T sh[G / kWave];
        v = wave_sum(v);
        if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
        __syncthreads();
        T r = 0;
#pragma unroll
        for (int i = 0; i < G / kWave; i++) r += sh[i];
        __syncthreads();
        return r;
    }
}

// Ballot restricted to the G lanes of the caller's group (G <= 64), bit i =
// group lane i.
template <int G>
__device__ inline uint64_t group_ballot(bool p)
{
    uint64_t b = __ballot(p);
    if constexpr (G == kWave) return b;
    int base = lane_id() & ~(G - 1);
    return (b >> base) & ((1ull << G) - 1);
}

template <int G>
__device__ inline uint64_t group_lanemask_lt()
{
    int l = glane<G>();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Phased row walks.  A group of L lanes takes T entries per lane of the row
// range [b, e): j = b + lane + t*L.  Every index and weight load is issued
// before the first one is used, and the gathers through the indices follow
// back to back, so a lane keeps T loads in flight where an entry-by-entry
// loop waits out two memory round trips per entry.  Slots past e hold
// neighbour -1 and weight 0.
template <int T, int L>
__device__ inline void load_entries(const vid_t *__restrict__ adjncy, const wgt_t *__restrict__ adjwgt, eid_t b,
                                    eid_t e, int lane, int32_t (&u)[T], int32_t (&w)[T])
{
#pragma unroll
    for (int t = 0; t < T; t++) {
        eid_t j = b + lane + (eid_t)t * L;
        u[t] = -1;
        w[t] = 0;
        if (j < e) {   // streamed once: kept out of the way of the gathered arrays in L2
            u[t] = __builtin_nontemporal_load(adjncy + j);
            w[t] = __builtin_nontemporal_load(adjwgt + j);
        }
    }
}

// x[t] = a[u[t]] for the live slots (u[t] >= 0), dflt for the others.
template <int T, class A>
__device__ inline void gather_slots(const A *__restrict__ a, const int32_t (&u)[T], int32_t (&x)[T], int32_t dflt)
{
#pragma unroll
    for (int t = 0; t < T; t++) x[t] = u[t] >= 0 ? (int32_t)a[u[t]] : dflt;
}

inline unsigned grid_groups(int64_t groups, int G)
{
    return grid_for(groups * G, kRowBS);
}

}  // namespace hgp
