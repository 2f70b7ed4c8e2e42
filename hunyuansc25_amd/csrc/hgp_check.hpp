// hgp_check.hpp -- device-side CSR checks and partition evaluation.
#pragma once

#include "hgp_common.hpp"

namespace hgp {

// Kinds of CSR fault, in the order they are reported.  The first three are
// found by the row pass (offsets and vertex weights), the others by the
// entry pass, which runs only on monotone offsets.
enum : int {
    kChkXadj0 = 0,
    kChkMonotone,
    kChkVwgt,
    kChkNeighbour,
    kChkSelfLoop,
    kChkAdjwgt,
    kChkReasons
};

// Device record of one check: the smallest vertex showing each fault
// (ULLONG_MAX: none), the total edge weight over the directed entries and
// two hash sums that are zero for a symmetric graph.
struct CheckState {
    unsigned long long first[8];
    unsigned long long wsum;
    unsigned long long h[2];
};

// Launches the row pass (rows = true; clears s first) or the entry pass on
// stream st.  m = xadj[n].
void check_graph_launch(hipStream_t st, int32_t n, int64_t m, const eid_t *xadj, const vid_t *adjncy,
                        const wgt_t *vwgt, const wgt_t *adjwgt, bool symmetric, CheckState *s, bool rows);

// Checks a device CSR (xadj[0] = 0, monotone offsets, vertex weights >= 1,
// neighbour ids in [0, n), no self-loops, edge weights >= 1, total edge
// weight < 2^31, and with `symmetric` every edge in both rows with the same
// weight); throws Error{-1} naming the first fault, Error{-4} past the
// supported sizes.  Waits for the device (two read-backs).
void check_graph_device(Ctx &c, int32_t n, const eid_t *xadj, const vid_t *adjncy, const wgt_t *vwgt,
                        const wgt_t *adjwgt, bool symmetric);

// Evaluation scratch: kEvalHead int64 (cut accumulator, bad part ids, bad
// neighbour ids, pad) followed by k part weights.
constexpr int64_t kEvalHead = 4;
constexpr int32_t kEvalMaxParts = 4096;

// Queues the evaluation of part[] on stream st: out (device, 4 int64) =
// {edge cut, heaviest part weight, vertices with a part id outside [0, k),
// entries with a neighbour id outside [0, n)}.  scr holds kEvalHead + k
// int64 and is overwritten.  vwgt/adjwgt may be null (unit weights).
void evaluate_launch(hipStream_t st, int32_t n, const eid_t *xadj, const vid_t *adjncy, const wgt_t *adjwgt,
                     const wgt_t *vwgt, const int32_t *part, int32_t k, int64_t *scr, int64_t *out);

// The same evaluation with the fixed per-device scratch instead of an arena
// (k <= kEvalMaxParts, checked by the caller): no device memory is
// allocated.  Runs on the null stream of `device` and waits for it; calls on
// one device are serialised by a lock.
void evaluate_fixed_scratch(int32_t device, int32_t n, const eid_t *xadj, const vid_t *adjncy,
                            const wgt_t *adjwgt, const wgt_t *vwgt, const int32_t *part, int32_t k,
                            int64_t *out);

}  // namespace hgp
