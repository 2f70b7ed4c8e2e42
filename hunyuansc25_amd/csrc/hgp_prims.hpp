// hgp_prims.hpp -- device-wide primitives: scans, stable radix sort, fills,
// reductions.  All run on the context's stream; temporaries come from the
// right end of the arena and are released before return (stream order makes
// the reuse safe).
#pragma once

#include "hgp_common.hpp"

namespace hgp {

// Inclusive scan, in == out allowed (hunyuangraph_GPU_prefixsum.h:141-188).
void scan_inclusive(Ctx &c, const int32_t *in, int32_t *out, int64_t n);
void scan_inclusive(Ctx &c, const int64_t *in, int64_t *out, int64_t n);

// off[0] = 0, off[i+1] = sum(cnt[0..i]); off has n+1 entries.
void offsets_from_counts(Ctx &c, const int64_t *cnt, int64_t *off, int64_t n);
void offsets_from_counts(Ctx &c, const int32_t *cnt, int64_t *off, int64_t n);
void offsets_from_counts(Ctx &c, const int32_t *cnt, int32_t *off, int64_t n);

// Stable LSD radix sort of (key, val) pairs by key bits [lo, hi) of each
// listed range, in place.  Ranges are processed least significant first.
void radix_sort(Ctx &c, uint32_t *keys, uint32_t *vals, int64_t n, int lo, int hi);
void radix_sort(Ctx &c, uint64_t *keys, uint32_t *vals, int64_t n, int lo, int hi);
// The same sort of keys alone; idx receives the permutation (the original
// index of each sorted key), as a sort of (key, index) pairs would.
void radix_sort_index(Ctx &c, uint32_t *keys, uint32_t *idx, int64_t n, int lo, int hi);

template <class T> void fill(Ctx &c, T *p, int64_t n, T v);
void iota(Ctx &c, int32_t *p, int64_t n);

// Sum of an int32 array into an int64 device scalar (overwrites *out).
void reduce_sum(Ctx &c, const int32_t *in, int64_t n, int64_t *out);

// Stable compaction of the indices i < n with flag[i] != 0 into idx (in
// increasing order); *count (device int64) receives the number kept.
void compact_indices(Ctx &c, const uint8_t *flag, int64_t n, int32_t *idx, int64_t *count);

int bits_for(uint64_t maxval);   // number of bits needed to hold maxval

}  // namespace hgp
