// hgp_common.hpp -- shared types, error handling, device arena and wave64
// helpers of the MI355X partitioner.
#pragma once

#include <hip/hip_runtime.h>

#include "hgp_hostcheck.hpp"

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace hgp {

using eid_t = int64_t;   // edge offsets (CSR xadj), 64-bit so m may exceed 2^31
using vid_t = int32_t;   // vertex ids
using wgt_t = int32_t;   // vertex and edge weights (int32, as in the reference)

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxBuckets = 50;    // Jet gain buckets (hunyuangraph_GPU_krefine.h:2473)
constexpr int kMidBucket = 25;

// ------------------------------------------------------------------ errors

void clear_error();

#define HGP_CHECK_HIP(expr)                                                        \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) throw ::hgp::Error{-2, std::string(#expr) + ": " +   \
                                                     hipGetErrorString(e_)};       \
    } while (0)

#define HGP_LAUNCH_CHECK() HGP_CHECK_HIP(hipGetLastError())

// ------------------------------------------------------------------ arena

// Two-ended stack over one device allocation (after hunyuangraph_GPU_memory.h:34-376):
// the left end holds data that lives across pipeline phases (the level
// hierarchy), the right end holds per-phase temporaries.  Frees are LIFO per
// end; a mark/release pair frees everything above a mark at once.  A request
// that does not fit throws Error{-3}.
class Arena {
  public:
    Arena() = default;
    ~Arena();
    void reserve(size_t bytes);          // (re)allocate the block if smaller
    void *left(size_t bytes);
    void *right(size_t bytes);
    size_t left_mark() const { return lo_; }
    size_t right_mark() const { return hi_; }
    void left_release(size_t mark);
    void right_release(size_t mark);
    void reset();                        // free everything (keeps the block)
    void free_block();                   // return the block to HIP (nothing may be live)
    size_t capacity() const { return cap_; }
    size_t high_water() const { return peak_; }
    char *base() const { return base_; }  // left allocations start here (left_mark() bytes in use)

    template <class T> T *lalloc(size_t n) { return (T *)left(sizeof(T) * (n ? n : 1)); }
    template <class T> T *ralloc(size_t n) { return (T *)right(sizeof(T) * (n ? n : 1)); }

  private:
    char *base_ = nullptr;
    size_t cap_ = 0, lo_ = 0, hi_ = 0, peak_ = 0;
};

// ------------------------------------------------------------------ context

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    int64_t *pinned = nullptr;     // pinned host buffer: scalar read-backs + upload staging
    hipEvent_t ev[8] = {};         // 0-3: phase times of a partition; 7: read_scalars_start
};

constexpr size_t kPinnedBytes = 1 << 20;
constexpr size_t kStageOffset = 64 << 10;   // staging area starts here

Ctx &context(int device);                        // per-device default context (lazily created)
std::unique_ptr<Ctx> create_context(int device);   // a private context (own stream, arena, staging)
void destroy_context(std::unique_ptr<Ctx> c);      // releases a private context's resources

// ------------------------------------------------------------------ tuning
// Scheduling knobs: launch shapes and path thresholds that change how fast a
// partition is computed, never what it is (every path computes the same
// exact result).  Defaults are the values measured best by hand; the
// autotuner (hunyuansc25_amd/autotune.py) measures them on the device in use
// and sets them through hgp_set_tuning.
struct Tuning {
    int32_t tiny_max = 512;      // levels of <= this many vertices (k <= 256) run the Jet loop in one block
    int64_t few_rows = 8192;     // levels with <= this many rows of <= 512 entries run them in one launch
    int64_t fill_waves = 8192;   // waves the wave-per-row kernels spread a level's rows over
    int32_t seg_grid = 32768;    // grid cap of the segment kernels
    int32_t batch_lanes = 8;     // lanes of hgp_partition_batch when the caller passes <= 0
    int32_t hub_track = 0;       // 1: long rows keep their connectivity rows current through a level's
                                 // iterations (k <= 512) instead of rebuilding them in each selection
    int32_t graph_max_vertices = 0;   // levels of <= this many vertices replay their Jet launch
                                      // sequences as HIP graphs (0: never)
    int32_t batch_fuse_max = 8192;    // hgp_partition_batch refines the levels of <= this many vertices
                                      // of all graphs in one launch (0: per graph; at most 16384)
};
Tuning &tuning();

// ------------------------------------------------------------------ kernel timer
// Optional per-kernel timing with HIP events on the context's stream: a
// KScope records an event pair around the launch in its scope and the
// algorithmic HBM bytes the launch must move: a fixed part known on the host
// plus, for kernels whose work depends on the data (rows of free, eligible or
// moved vertices), a device counter the kernel adds its row bytes to.  Off
// unless enabled.
struct KRec {
    const char *name;
    hipEvent_t a, b;
    double bytes;
    int64_t slot;                // device counter slot, -1 if none
};

struct KTimer {
    bool on = false;
    std::mutex mu;               // scopes may close on several host threads
    std::vector<KRec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    unsigned long long *dcnt = nullptr;   // device counters (bytes)
    int64_t ncnt = 0, cnt_cap = 0;
    hipEvent_t get();
    int64_t counter_slot();      // -1 when the counters are exhausted
};

KTimer &ktimer();

// Optional trace of the Jet refinement loops: one record per read of the
// control block (before each iteration, and the final read), holding the
// level's vertex count, the iterations done, the mode of the coming
// iteration, the cut so far and whether the loop has ended.  Off unless
// enabled.
struct JetTrace {
    bool on = false;
    std::mutex mu;
    std::vector<int64_t> recs;   // kJetTraceFields per record
    // one iteration's state: where (before its moves) and dest (the moves)
    // of the level with dump_n vertices at iteration dump_iter
    int64_t dump_n = -1;
    int32_t dump_iter = -1;
    std::vector<int32_t> where, dest;
};
constexpr int kJetTraceFields = 5;
JetTrace &jet_trace();

class KScope {
  public:
    KScope(Ctx &c, const char *name, double bytes, bool counted = false);
    ~KScope();
    // Device counter for the launch's data-dependent bytes (nullptr when off).
    unsigned long long *dev() const { return dev_; }

  private:
    Ctx &c_;
    bool on_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
    const char *name_;
    double bytes_;
    int64_t slot_ = -1;
    unsigned long long *dev_ = nullptr;
};

// Copies `bytes` from device to the pinned buffer and waits.
void read_scalars(Ctx &c, const void *dsrc, size_t bytes);

// The same copy in two halves: start queues it (and an event behind it),
// wait blocks until it has landed; work queued in between keeps the device
// busy meanwhile.  The pinned buffer must not be reused in between.
void read_scalars_start(Ctx &c, const void *dsrc, size_t bytes);
void read_scalars_wait(Ctx &c);

// Host-to-device copy through the pinned staging area; waits for the copy, so
// the host source may die right after the call.
void upload_staged(Ctx &c, void *ddst, const void *hsrc, size_t bytes);

// ------------------------------------------------------------------ hashing
// Identical to the oracle's generators (oracle_dga.c): SplitMix64 finaliser.

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t rng3(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return mix64(mix64(mix64(seed ^ a) ^ b) ^ c);
}

__host__ __device__ inline uint32_t edge_hash(uint64_t seed, int32_t level, int32_t lo, int32_t hi)
{
    return (uint32_t)(rng3(seed, 0x100000000ull + (uint64_t)(uint32_t)level, (uint64_t)(uint32_t)lo,
                           (uint64_t)(uint32_t)hi) >> 32);
}

// edge_hash split for the proposal loops: the level's key s1 is computed once
// per launch and mix64(s1 ^ v) once per row, so an edge (v, u) costs one mix64
// when u > v and two when u < v (the same value as edge_hash).
__host__ __device__ inline uint64_t edge_key(uint64_t seed, int32_t level)
{
    return mix64(seed ^ (0x100000000ull + (uint64_t)(uint32_t)level));
}

__device__ inline uint32_t edge_hash_row(uint64_t s1, uint64_t tv, int32_t v, int32_t u)
{
    uint64_t z = u > v ? mix64(tv ^ (uint64_t)(uint32_t)u) : mix64(mix64(s1 ^ (uint64_t)(uint32_t)u) ^ (uint64_t)(uint32_t)v);
    return (uint32_t)(z >> 32);
}

__host__ __device__ inline uint64_t trial_rng(uint64_t seed, uint32_t node, int32_t trial, int32_t draw)
{
    return rng3(seed, 0x200000000ull + node, (uint64_t)(uint32_t)trial, (uint64_t)(uint32_t)draw);
}

// ------------------------------------------------------------------ launch

inline unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 0x7fffffff) g = 0x7fffffff;
    return (unsigned)g;
}

// ------------------------------------------------------------------ device helpers

__device__ inline int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

__device__ inline uint64_t lanemask_lt()
{
    int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Orders a wave's LDS accesses around this point (the wave's own lanes only).
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <class T>
__device__ inline T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <class T>
__device__ inline T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, kWave);
        v = u > v ? u : v;
    }
    return v;
}

// Inclusive wave scan (Hillis-Steele over 64 lanes).
template <class T>
__device__ inline T wave_inclusive_scan(T v)
{
    int l = lane_id();
    for (int o = 1; o < kWave; o <<= 1) {
        T u = __shfl_up(v, o, kWave);
        if (l >= o) v += u;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (blockDim.x <= 1024,
// multiple of 64).  `sh` needs blockDim.x/64 + 1 elements.  Returns the
// exclusive prefix; *total receives the block total.
template <class T>
__device__ inline T block_exclusive_scan(T v, T *sh, T *total)
{
    int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    T inc = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1) sh[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int i = 0; i < nw; i++) { T t = sh[i]; sh[i] = acc; acc += t; }
        sh[nw] = acc;
    }
    __syncthreads();
    T r = sh[w] + inc - v;
    if (total) *total = sh[nw];
    __syncthreads();
    return r;
}

// Kernel-timer byte counters come in groups of kKbLanes, spread over by block
// index, so that the timed kernels do not contend on one address.
constexpr int kKbLanes = 64;

// Adds each thread's byte count to the launch's kernel-timer counters: the
// block sums in LDS, then one atomic per block.  Every thread of the block
// must call it (kb may be null: timer off, then it returns at once).
__device__ inline void kb_flush(unsigned long long *kb, uint64_t v)
{
    if (!kb) return;
    __shared__ unsigned long long kb_acc;
    if (threadIdx.x == 0) kb_acc = 0;
    __syncthreads();
    v = wave_sum(v);
    if (lane_id() == 0 && v) atomicAdd(&kb_acc, (unsigned long long)v);
    __syncthreads();
    if (threadIdx.x == 0 && kb_acc) atomicAdd(&kb[blockIdx.x % kKbLanes], kb_acc);
}

}  // namespace hgp
