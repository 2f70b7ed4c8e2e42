// hgp_prims.hip -- scans, stable radix sort, fills, reductions (wave64).
#include "hgp_prims.hpp"

#include <algorithm>
#include "hgp_group.hpp"

namespace hgp {

namespace {

constexpr int kScanBS = 256;    // threads per block
constexpr int kScanR = 8;       // rounds per tile: tile = 2048 items
constexpr int64_t kRawTiles = 2048;   // multi-tile scans of up to this many tiles take two launches

// Per-tile sums of a striped tile (item r*256+t of the tile in round r).
template <class T>
__global__ void __launch_bounds__(kScanBS) k_tile_reduce(const T *__restrict__ in, int64_t n, T *__restrict__ sums)
{
    __shared__ T sh[kScanBS / kWave];
    int64_t base = (int64_t)blockIdx.x * (kScanBS * kScanR);
    T acc = 0;
#pragma unroll
    for (int r = 0; r < kScanR; r++) {
        int64_t i = base + r * kScanBS + threadIdx.x;
        if (i < n) acc += in[i];
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < kScanBS / kWave; w++) t += sh[w];
        sums[blockIdx.x] = t;
    }
}

// Inclusive scan of each tile plus the inclusive prefix of the previous
// tiles (incl_sums[tile-1]).
template <class T>
__global__ void __launch_bounds__(kScanBS) k_tile_scan(const T *in, T *out, int64_t n, const T *__restrict__ sums,
                                                      int raw)
{
    __shared__ T sh[kScanBS / kWave + 1];
    int64_t base = (int64_t)blockIdx.x * (kScanBS * kScanR);
    T carry = 0;
    if (sums && blockIdx.x > 0) {
        if (raw) {
            // sums[t] = total of tile t (not scanned): the block adds those of
            // the tiles before its own (few tiles, so this costs less than a launch)
            T part = 0;
            for (int i = threadIdx.x; i < (int)blockIdx.x; i += kScanBS) part += sums[i];
            block_exclusive_scan(part, sh, &carry);
        } else {
            carry = sums[blockIdx.x - 1];   // inclusive scan of the tile totals
        }
    }
    for (int r = 0; r < kScanR; r++) {
        int64_t i = base + r * kScanBS + threadIdx.x;
        T v = i < n ? in[i] : (T)0;
        T tot;
        T ex = block_exclusive_scan(v, sh, &tot);
        if (i < n) out[i] = carry + ex + v;
        carry += tot;
        if (base + (r + 1) * kScanBS >= n) break;   // uniform across the block
    }
}

template <class T>
void scan_impl(Ctx &c, const T *in, T *out, int64_t n)
{
    if (n <= 0) return;
    int64_t tile = kScanBS * kScanR;
    int64_t tiles = (n + tile - 1) / tile;
    if (tiles == 1) {
        k_tile_scan<T><<<1, kScanBS, 0, c.stream>>>(in, out, n, nullptr, 0);
        HGP_LAUNCH_CHECK();
        return;
    }
    size_t mark = c.arena.right_mark();
    T *sums = c.arena.ralloc<T>((size_t)tiles);
    k_tile_reduce<T><<<(unsigned)tiles, kScanBS, 0, c.stream>>>(in, n, sums);
    HGP_LAUNCH_CHECK();
    // up to kRawTiles tiles each block sums the tile totals before it itself;
    // beyond, the totals are scanned first (the per-block sums would grow
    // with the square of the tile count)
    const bool raw = tiles <= kRawTiles;
    if (!raw) scan_impl<T>(c, sums, sums, tiles);
    {
        KScope ks(c, sizeof(T) == 8 ? "k_tile_scan<i64>" : "k_tile_scan<i32>", 2.0 * sizeof(T) * (double)n);
        k_tile_scan<T><<<(unsigned)tiles, kScanBS, 0, c.stream>>>(in, out, n, sums, raw ? 1 : 0);
        HGP_LAUNCH_CHECK();
    }
    c.arena.right_release(mark);
}

template <class TI, class TO>
__global__ void k_copy_shift(const TI *__restrict__ in, TO *__restrict__ out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i + 1] = (TO)in[i];
    if (i == 0) out[0] = 0;
}

template <class T>
__global__ void k_fill(T *p, int64_t n, T v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_iota(int32_t *p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__global__ void k_reduce_i32(const int32_t *__restrict__ in, int64_t n, unsigned long long *out)
{
    __shared__ int64_t sh[4];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t acc = 0;
    for (; i < n; i += stride) acc += in[i];
    acc = wave_sum(acc);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) t += sh[w];
        atomicAdd(out, (unsigned long long)t);
    }
}

// ------------------------------------------------------------ radix sort

constexpr int kRsBS = 256;
constexpr int kRsIPT = 16;                 // rounds per tile
constexpr int kRsTile = kRsBS * kRsIPT;    // 4096 items
constexpr int kRsWaves = kRsBS / kWave;
constexpr int kRsFuseTiles = 64;           // up to this many tiles the scatter scans the histogram itself

template <class K>
__global__ void __launch_bounds__(kRsBS) k_rs_hist(const K *__restrict__ keys, int64_t n, int shift, int ntiles,
                                                   uint32_t *__restrict__ hist)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsIPT; r++) {
        int64_t i = base + r * kRsBS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// One round (kRsBS items from `first`) of the stable scatter: items are
// ranked in index order; ranks inside a wave come from an 8-ballot digit
// match, ranks across waves from the LDS counters wcnt.  base_d[d] holds the
// next output slot of digit d and is advanced past the round's items.
// Without vin an item's value is its index (the first pass of a sort that
// returns the permutation).
template <class K>
__device__ inline void scatter_round(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                     K *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n, int shift,
                                     int64_t first, uint32_t *base_d, uint32_t (*wcnt)[256])
{
    int t = threadIdx.x, w = t / kWave;
    int64_t i = first + t;
    bool valid = i < n;
    K key = valid ? kin[i] : (K)0;
    uint32_t val = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++) wcnt[ww][t] = 0;
    __syncthreads();
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    uint64_t lower = m & lanemask_lt();
    uint32_t rank = (uint32_t)__popcll(lower);
    if (valid && lower == 0) wcnt[w][d] = (uint32_t)__popcll(m);
    __syncthreads();
    {
        uint32_t acc = base_d[t];
#pragma unroll
        for (int ww = 0; ww < kRsWaves; ww++) {
            uint32_t c = wcnt[ww][t];
            wcnt[ww][t] = acc;
            acc += c;
        }
        base_d[t] = acc;
    }
    __syncthreads();
    if (valid) {
        uint32_t pos = wcnt[w][d] + rank;
        kout[pos] = key;
        vout[pos] = val;
    }
    __syncthreads();
}

// Stable scatter of one tile.  incl = inclusive scan of hist (digit-major),
// so the tile's first slot for digit d is incl[d*ntiles+tile] -
// hist[d*ntiles+tile].  Without incl (at most kRsFuseTiles tiles) the block
// computes the same from hist: thread d sums digit d over all tiles and over
// the tiles before its own, and a block scan over the digits adds the rest.
template <class K>
__global__ void __launch_bounds__(kRsBS) k_rs_scatter(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                      K *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n,
                                                      int shift, int ntiles, const uint32_t *__restrict__ hist,
                                                      const uint32_t *__restrict__ incl)
{
    __shared__ uint32_t base_d[256];
    __shared__ uint32_t wcnt[kRsWaves][256];
    int t = threadIdx.x;
    if (incl) {
        int64_t h = (int64_t)t * ntiles + blockIdx.x;
        base_d[t] = incl[h] - hist[h];
    } else {
        __shared__ uint32_t wsh[kRsWaves + 1];
        uint32_t tot = 0, before = 0;
        const uint32_t *row = hist + (int64_t)t * ntiles;
        for (int tt = 0; tt < ntiles; tt++) {
            const uint32_t x = row[tt];
            tot += x;
            before += tt < (int)blockIdx.x ? x : 0u;
        }
        base_d[t] = block_exclusive_scan<uint32_t>(tot, wsh, nullptr) + before;
    }
    int64_t tbase = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsIPT; r++) {
        if (tbase + (int64_t)r * kRsBS >= n) break;      // uniform across the block
        scatter_round<K>(kin, vin, kout, vout, n, shift, tbase + (int64_t)r * kRsBS, base_d, wcnt);
    }
}

// All passes of a sort of at most one tile in a single block: per pass a
// digit histogram in LDS, its exclusive scan and the stable scatter into the
// other buffer; the sorted items end in keys/vals.  The passes read what the
// block wrote in the pass before (ordered by the barriers at block scope).
template <class K>
__global__ void __launch_bounds__(kRsBS) k_rs_small(K *keys, uint32_t *vals, K *k2, uint32_t *v2, int32_t n,
                                                    int lo, int hi, int index_vals)
{
    __shared__ uint32_t cnt[256];
    __shared__ uint32_t base_d[256];
    __shared__ uint32_t wcnt[kRsWaves][256];
    __shared__ uint32_t wsh[kRsWaves + 1];
    int t = threadIdx.x;
    K *src_k = keys, *dst_k = k2;
    uint32_t *src_v = vals, *dst_v = v2;
    for (int shift = lo; shift < hi; shift += 8) {
        cnt[t] = 0;
        __syncthreads();
        for (int32_t i = t; i < n; i += kRsBS) atomicAdd(&cnt[(uint32_t)(src_k[i] >> shift) & 255u], 1u);
        __syncthreads();
        base_d[t] = block_exclusive_scan<uint32_t>(cnt[t], wsh, nullptr);
        __syncthreads();
        const uint32_t *vin = shift == lo && index_vals ? nullptr : src_v;
        for (int32_t first = 0; first < n; first += kRsBS)
            scatter_round<K>(src_k, vin, dst_k, dst_v, n, shift, first, base_d, wcnt);
        K *tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t *tv = src_v; src_v = dst_v; dst_v = tv;
    }
    if (src_k != keys) {
        for (int32_t i = t; i < n; i += kRsBS) {
            keys[i] = src_k[i];
            vals[i] = src_v[i];
        }
    }
}

template <class K>
void radix_impl(Ctx &c, K *keys, uint32_t *vals, int64_t n, int lo, int hi, bool index_vals = false)
{
    if (index_vals && (n == 1 || (n > 1 && hi <= lo))) iota(c, (int32_t *)vals, n);
    if (n <= 1 || hi <= lo) return;
    if (n > 0x7fffffffll) throw Error{-4, "radix_sort: more than 2^31 items"};
    size_t mark = c.arena.right_mark();
    int ntiles = (int)((n + kRsTile - 1) / kRsTile);
    K *k2 = c.arena.ralloc<K>((size_t)n);
    uint32_t *v2 = c.arena.ralloc<uint32_t>((size_t)n);
    if (ntiles == 1) {
        KScope ks(c, sizeof(K) == 8 ? "k_rs_small<u64>" : "k_rs_small<u32>",
                  (double)((hi - lo + 7) / 8) * 3.0 * (double)(sizeof(K) + 4) * (double)n);
        k_rs_small<K><<<1, kRsBS, 0, c.stream>>>(keys, vals, k2, v2, (int32_t)n, lo, hi, index_vals ? 1 : 0);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(mark);
        return;
    }
    uint32_t *hist = c.arena.ralloc<uint32_t>((size_t)ntiles * 256);
    uint32_t *incl = c.arena.ralloc<uint32_t>((size_t)ntiles * 256);
    K *src_k = keys, *dst_k = k2;
    uint32_t *src_v = vals, *dst_v = v2;
    int passes = 0;
    for (int shift = lo; shift < hi; shift += 8) {
        {
            KScope ks(c, sizeof(K) == 8 ? "k_rs_hist<u64>" : "k_rs_hist<u32>", (double)sizeof(K) * (double)n);
            k_rs_hist<K><<<ntiles, kRsBS, 0, c.stream>>>(src_k, n, shift, ntiles, hist);
            HGP_LAUNCH_CHECK();
        }
        const bool fuse = ntiles <= kRsFuseTiles;
        if (!fuse) scan_impl<int32_t>(c, (const int32_t *)hist, (int32_t *)incl, (int64_t)ntiles * 256);
        {
            KScope ks(c, sizeof(K) == 8 ? "k_rs_scatter<u64>" : "k_rs_scatter<u32>",
                      2.0 * (double)(sizeof(K) + 4) * (double)n);
            const uint32_t *vin = shift == lo && index_vals ? nullptr : src_v;
            k_rs_scatter<K><<<ntiles, kRsBS, 0, c.stream>>>(src_k, vin, dst_k, dst_v, n, shift, ntiles, hist,
                                                            fuse ? nullptr : incl);
            HGP_LAUNCH_CHECK();
        }
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
        passes++;
    }
    if (passes & 1) {
        HGP_CHECK_HIP(hipMemcpyAsync(keys, src_k, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
        HGP_CHECK_HIP(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    }
    c.arena.right_release(mark);
}

__global__ void k_flag_to_i32(const uint8_t *__restrict__ f, int64_t n, int32_t *__restrict__ o)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = f[i] != 0;
}

__global__ void k_compact(const uint8_t *__restrict__ f, const int32_t *__restrict__ incl, int64_t n,
                          int32_t *__restrict__ idx, int64_t *count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && f[i]) idx[incl[i] - 1] = (int32_t)i;
    if (i == n - 1) *count = incl[i];
}

// Flag compaction in two launches (up to kRawTiles tiles): per-tile flag
// counts, then each block adds the counts of the tiles before its own, ranks
// its tile's flags with block scans and writes their indices in order.
__global__ void __launch_bounds__(kScanBS) k_flag_tile_count(const uint8_t *__restrict__ f, int64_t n,
                                                             int32_t *__restrict__ tsum)
{
    __shared__ int32_t sh[kScanBS / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * (kScanBS * kScanR);
    int32_t cnt = 0;
#pragma unroll
    for (int r = 0; r < kScanR; r++) {
        const int64_t i = base + r * kScanBS + threadIdx.x;
        cnt += (i < n && f[i]) ? 1 : 0;
    }
    int32_t tot;
    block_exclusive_scan(cnt, sh, &tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBS) k_flag_compact(const uint8_t *__restrict__ f, int64_t n,
                                                          const int32_t *__restrict__ tsum,
                                                          int32_t *__restrict__ idx, int64_t *count)
{
    __shared__ int32_t sh[kScanBS / kWave + 1];
    const int64_t base = (int64_t)blockIdx.x * (kScanBS * kScanR);
    int32_t part = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += kScanBS) part += tsum[i];
    int32_t carry;
    block_exclusive_scan(part, sh, &carry);
    for (int r = 0; r < kScanR; r++) {
        const int64_t i = base + r * kScanBS + threadIdx.x;
        const int32_t v = (i < n && f[i]) ? 1 : 0;
        int32_t tot;
        const int32_t ex = block_exclusive_scan(v, sh, &tot);
        if (v) idx[carry + ex] = (int32_t)i;
        carry += tot;
        if (base + (r + 1) * kScanBS >= n) break;   // uniform across the block
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count = carry;
}

// ------------------------------------------------------------ degree classes

constexpr int kNumCls = 6;

struct BinThr {
    int64_t t0, t1, t2, t3, t4;
    __device__ int cls(int64_t d) const
    {
        return d <= t0 ? 0 : (d <= t1 ? 1 : (d <= t2 ? 2 : (d <= t3 ? 3 : (d <= t4 ? 4 : 5))));
    }
};

// Class counts: per-block LDS counters, one atomic per block and class.
__global__ void __launch_bounds__(256) k_bin_count(int32_t n, const eid_t *__restrict__ xadj,
                                                   const int64_t *__restrict__ len, BinThr t, unsigned long long *cnt)
{
    __shared__ uint32_t bc[kNumCls];
    __shared__ unsigned long long be[kNumCls];
    if (threadIdx.x < kNumCls) {
        bc[threadIdx.x] = 0;
        be[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int64_t v0 = (int64_t)blockIdx.x * 256; v0 < n; v0 += (int64_t)gridDim.x * 256) {
        int64_t v = v0 + threadIdx.x;
        int64_t d = v < n ? (xadj ? xadj[v + 1] - xadj[v] : len[v]) : 0;
        int cls = v < n ? t.cls(d) : -1;
#pragma unroll
        for (int c = 0; c < kNumCls; c++) {
            uint64_t b = __ballot(cls == c);
            uint64_t e = wave_sum((uint64_t)(cls == c ? d : 0));
            if (lane_id() == 0 && b) {
                atomicAdd(&bc[c], (uint32_t)__popcll(b));
                atomicAdd(&be[c], (unsigned long long)e);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kNumCls && bc[threadIdx.x]) {
        atomicAdd(&cnt[threadIdx.x], (unsigned long long)bc[threadIdx.x]);
        atomicAdd(&cnt[2 * kNumCls + threadIdx.x], be[threadIdx.x]);
    }
}

struct BinBase {
    int64_t o[kNumCls];
};

// Each block takes its slots of every class with one atomic per class, then
// hands them out to its waves through LDS.
__global__ void __launch_bounds__(256) k_bin_fill(int32_t n, const eid_t *__restrict__ xadj,
                                                  const int64_t *__restrict__ len, BinThr t, BinBase base,
                                                  unsigned long long *cursor, int32_t *__restrict__ ids)
{
    __shared__ uint32_t wc[kNumCls][256 / kWave];
    __shared__ unsigned long long bb[kNumCls];
    const int w = threadIdx.x / kWave;
    for (int64_t v0 = (int64_t)blockIdx.x * 256; v0 < n; v0 += (int64_t)gridDim.x * 256) {
        int64_t v = v0 + threadIdx.x;
        int cls = v < n ? t.cls(xadj ? xadj[v + 1] - xadj[v] : len[v]) : -1;
        uint64_t bal[kNumCls];
#pragma unroll
        for (int c = 0; c < kNumCls; c++) {
            bal[c] = __ballot(cls == c);
            if (lane_id() == 0) wc[c][w] = (uint32_t)__popcll(bal[c]);
        }
        __syncthreads();
        if (threadIdx.x < kNumCls) {
            int c = threadIdx.x;
            uint32_t tot = 0;
            for (int i = 0; i < 256 / kWave; i++) {
                uint32_t x = wc[c][i];
                wc[c][i] = tot;
                tot += x;
            }
            bb[c] = tot ? atomicAdd(&cursor[c], (unsigned long long)tot) : 0;
        }
        __syncthreads();
        if (cls >= 0)
            ids[base.o[cls] + (int64_t)bb[cls] + wc[cls][w] + __popcll(bal[cls] & lanemask_lt())] = (int32_t)v;
        __syncthreads();
    }
}

__global__ void k_seg_count(int64_t nl, const int32_t *__restrict__ list, const eid_t *__restrict__ xadj,
                            int64_t *__restrict__ cnt, uint8_t *__restrict__ multi, int32_t *__restrict__ mimap)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    int32_t v = list[i];
    int64_t c = (xadj[v + 1] - xadj[v] + kSegLen - 1) / kSegLen;
    cnt[i] = c;
    multi[i] = c > 1;
    mimap[i] = -1;
}

__global__ void k_seg_multi(int64_t nm, const int64_t *__restrict__ mlist_n, const int32_t *__restrict__ mlist,
                            const int32_t *__restrict__ list, const int64_t *__restrict__ segoff,
                            int32_t *__restrict__ mimap, int32_t *__restrict__ mrow, int64_t *__restrict__ mseg)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm || t >= *mlist_n) return;
    int32_t li = mlist[t];
    mimap[li] = (int32_t)t;
    mrow[t] = list[li];
    mseg[2 * t] = segoff[li];
    mseg[2 * t + 1] = segoff[li + 1];
}

__global__ void k_seg_fill(int64_t nl, const int32_t *__restrict__ list, const eid_t *__restrict__ xadj,
                           const int64_t *__restrict__ segoff, const int32_t *__restrict__ mimap,
                           int32_t *__restrict__ row, int64_t *__restrict__ j0, int32_t *__restrict__ mi)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    int32_t v = list[i];
    eid_t b = xadj[v];
    for (int64_t s = segoff[i]; s < segoff[i + 1]; s++) {
        row[s] = v;
        j0[s] = b + (s - segoff[i]) * kSegLen;
        mi[s] = mimap[i];
    }
}

}  // namespace

RowBins row_bins(Ctx &c, int32_t n, const eid_t *xadj, const int64_t *len, bool left, int64_t t0, int64_t t1,
                 int64_t t2, int64_t t3, int64_t t4)
{
    RowBins b;
    b.ids = left ? c.arena.lalloc<int32_t>((size_t)n) : c.arena.ralloc<int32_t>((size_t)n);
    if (n <= 0) return b;
    size_t mark = c.arena.right_mark();
    unsigned long long *cnt = c.arena.ralloc<unsigned long long>(3 * kNumCls);   // counts, cursors, entries
    fill<uint32_t>(c, (uint32_t *)cnt, 6 * kNumCls, 0u);
    unsigned g = (unsigned)std::min<int64_t>(((int64_t)n + 255) / 256, 1024);
    BinThr thr{t0, t1, t2, t3, t4};
    k_bin_count<<<g, 256, 0, c.stream>>>(n, xadj, len, thr, cnt);
    HGP_LAUNCH_CHECK();
    read_scalars(c, cnt, 3 * kNumCls * sizeof(int64_t));
    BinBase base;
    b.off[0] = 0;
    for (int i = 0; i < kNumCls; i++) {
        base.o[i] = b.off[i];
        b.off[i + 1] = b.off[i] + c.pinned[i];
        b.ent[i] = c.pinned[2 * kNumCls + i];
    }
    k_bin_fill<<<g, 256, 0, c.stream>>>(n, xadj, len, thr, base, cnt + kNumCls, b.ids);
    HGP_LAUNCH_CHECK();
    c.arena.right_release(mark);
    return b;
}

RowSegs row_segs(Ctx &c, const eid_t *xadj, const RowBins &b)
{
    RowSegs r;
    int64_t nl = b.count(kClsLong);
    if (nl <= 0) return r;
    const int32_t *list = b.list(kClsLong);
    int32_t *mimap = c.arena.lalloc<int32_t>((size_t)nl);
    size_t mark = c.arena.right_mark();
    int64_t *cnt = c.arena.ralloc<int64_t>((size_t)nl);
    int64_t *segoff = c.arena.ralloc<int64_t>((size_t)nl + 1);
    uint8_t *multi = c.arena.ralloc<uint8_t>((size_t)nl);
    int32_t *mlist = c.arena.ralloc<int32_t>((size_t)nl);
    int64_t *d_nm = c.arena.ralloc<int64_t>(1);
    unsigned g = grid_for(nl, 256);
    k_seg_count<<<g, 256, 0, c.stream>>>(nl, list, xadj, cnt, multi, mimap);
    HGP_LAUNCH_CHECK();
    offsets_from_counts(c, cnt, segoff, nl);
    compact_indices(c, multi, nl, mlist, d_nm);
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned, segoff + nl, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    HGP_CHECK_HIP(hipMemcpyAsync(c.pinned + 1, d_nm, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    r.nseg = c.pinned[0];
    r.nmulti = (int32_t)c.pinned[1];
    r.row = c.arena.lalloc<int32_t>((size_t)r.nseg);
    r.j0 = c.arena.lalloc<int64_t>((size_t)r.nseg);
    r.mi = c.arena.lalloc<int32_t>((size_t)r.nseg);
    r.mrow = c.arena.lalloc<int32_t>((size_t)r.nmulti);
    r.mseg = c.arena.lalloc<int64_t>(2 * (size_t)r.nmulti);
    if (r.nmulti > 0) {
        k_seg_multi<<<grid_for(r.nmulti, 256), 256, 0, c.stream>>>(r.nmulti, d_nm, mlist, list, segoff, mimap, r.mrow,
                                                                   r.mseg);
        HGP_LAUNCH_CHECK();
    }
    k_seg_fill<<<g, 256, 0, c.stream>>>(nl, list, xadj, segoff, mimap, r.row, r.j0, r.mi);
    HGP_LAUNCH_CHECK();
    c.arena.right_release(mark);
    return r;
}

void scan_inclusive(Ctx &c, const int32_t *in, int32_t *out, int64_t n) { scan_impl<int32_t>(c, in, out, n); }
void scan_inclusive(Ctx &c, const int64_t *in, int64_t *out, int64_t n) { scan_impl<int64_t>(c, in, out, n); }

void offsets_from_counts(Ctx &c, const int64_t *cnt, int64_t *off, int64_t n)
{
    k_copy_shift<int64_t, int64_t><<<grid_for(n > 0 ? n : 1, 256), 256, 0, c.stream>>>(cnt, off, n);
    HGP_LAUNCH_CHECK();
    scan_impl<int64_t>(c, off + 1, off + 1, n);
}

void offsets_from_counts(Ctx &c, const int32_t *cnt, int64_t *off, int64_t n)
{
    k_copy_shift<int32_t, int64_t><<<grid_for(n > 0 ? n : 1, 256), 256, 0, c.stream>>>(cnt, off, n);
    HGP_LAUNCH_CHECK();
    scan_impl<int64_t>(c, off + 1, off + 1, n);
}

void offsets_from_counts(Ctx &c, const int32_t *cnt, int32_t *off, int64_t n)
{
    k_copy_shift<int32_t, int32_t><<<grid_for(n > 0 ? n : 1, 256), 256, 0, c.stream>>>(cnt, off, n);
    HGP_LAUNCH_CHECK();
    scan_impl<int32_t>(c, off + 1, off + 1, n);
}

void radix_sort(Ctx &c, uint32_t *keys, uint32_t *vals, int64_t n, int lo, int hi)
{
    radix_impl<uint32_t>(c, keys, vals, n, lo, hi);
}

void radix_sort(Ctx &c, uint64_t *keys, uint32_t *vals, int64_t n, int lo, int hi)
{
    radix_impl<uint64_t>(c, keys, vals, n, lo, hi);
}

void radix_sort_index(Ctx &c, uint32_t *keys, uint32_t *idx, int64_t n, int lo, int hi)
{
    radix_impl<uint32_t>(c, keys, idx, n, lo, hi, true);
}

template <class T>
void fill(Ctx &c, T *p, int64_t n, T v)
{
    if (n <= 0) return;
    unsigned g = grid_for(n, 256);
    if (g > 65536) g = 65536;
    k_fill<T><<<g, 256, 0, c.stream>>>(p, n, v);
    HGP_LAUNCH_CHECK();
}

template void fill<int32_t>(Ctx &, int32_t *, int64_t, int32_t);
template void fill<int64_t>(Ctx &, int64_t *, int64_t, int64_t);
template void fill<uint8_t>(Ctx &, uint8_t *, int64_t, uint8_t);
template void fill<uint32_t>(Ctx &, uint32_t *, int64_t, uint32_t);

void iota(Ctx &c, int32_t *p, int64_t n)
{
    if (n <= 0) return;
    k_iota<<<grid_for(n, 256), 256, 0, c.stream>>>(p, n);
    HGP_LAUNCH_CHECK();
}

void reduce_sum(Ctx &c, const int32_t *in, int64_t n, int64_t *out)
{
    fill<int64_t>(c, out, 1, 0);
    if (n <= 0) return;
    unsigned g = grid_for(n, 256);
    if (g > 4096) g = 4096;
    k_reduce_i32<<<g, 256, 0, c.stream>>>(in, n, (unsigned long long *)out);
    HGP_LAUNCH_CHECK();
}

void compact_indices(Ctx &c, const uint8_t *flag, int64_t n, int32_t *idx, int64_t *count)
{
    if (n <= 0) {
        fill<int64_t>(c, count, 1, 0);
        return;
    }
    size_t mark = c.arena.right_mark();
    const int64_t tiles = (n + kScanBS * kScanR - 1) / (kScanBS * kScanR);
    if (tiles <= kRawTiles) {
        int32_t *tsum = c.arena.ralloc<int32_t>((size_t)tiles);
        k_flag_tile_count<<<(unsigned)tiles, kScanBS, 0, c.stream>>>(flag, n, tsum);
        HGP_LAUNCH_CHECK();
        k_flag_compact<<<(unsigned)tiles, kScanBS, 0, c.stream>>>(flag, n, tsum, idx, count);
        HGP_LAUNCH_CHECK();
        c.arena.right_release(mark);
        return;
    }
    int32_t *incl = c.arena.ralloc<int32_t>((size_t)n);
    k_flag_to_i32<<<grid_for(n, 256), 256, 0, c.stream>>>(flag, n, incl);
    HGP_LAUNCH_CHECK();
    scan_impl<int32_t>(c, incl, incl, n);
    k_compact<<<grid_for(n, 256), 256, 0, c.stream>>>(flag, incl, n, idx, count);
    HGP_LAUNCH_CHECK();
    c.arena.right_release(mark);
}

int bits_for(uint64_t maxval)
{
    int b = 0;
    while (b < 64 && (maxval >> b) != 0) b++;
    return b;
}

}  // namespace hgp
