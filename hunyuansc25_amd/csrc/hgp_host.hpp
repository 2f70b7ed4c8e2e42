// hgp_host.hpp -- host-side job plumbing shared by the C entry points: error
// capture, arena sizing and the validated host-array job.
#pragma once

#include "hunyuangraph.h"

#include "hgp_common.hpp"
#include "hgp_graph.hpp"

#include <algorithm>
#include <cstdint>
#include <new>
#include <stdexcept>

namespace hgp {

const char *last_error_text();

// One graph given as host arrays (the arguments of hgp_partition).
struct HostJob {
    int32_t n;
    const int64_t *xadj;
    const int32_t *adjncy, *vwgt, *adjwgt;
    int32_t k;
    int32_t *part;
    hgp_stats *stats;
};

// A coarsening hierarchy held outside any arena, in one device allocation:
// levels 1..L, every level's degree classes and the fine-to-coarse maps.
// lv[0]'s CSR arrays are the resident input graph's (not part of buf).
// Nothing reads it but the uncoarsening, which only reads, so one
// hierarchy serves concurrent partitions of its graph.
struct CachedHierarchy {
    void *buf = nullptr;
    size_t bytes = 0;
    std::vector<DevGraph> lv;
    std::vector<int32_t *> cmaps;
    int64_t tvwgt = 0;
};

// What selects a hierarchy: the graph, the matching's seed and rounds, and
// the coarsening threshold (a function of k: every k in 128..512 coarsens to
// 1024 vertices, hunyuangraph_partitiongraph.h:298-302).
struct HierKey {
    int64_t graph;
    uint64_t seed;
    int32_t threshold, rounds;
    bool operator<(const HierKey &o) const
    {
        if (graph != o.graph) return graph < o.graph;
        if (seed != o.seed) return seed < o.seed;
        if (threshold != o.threshold) return threshold < o.threshold;
        return rounds < o.rounds;
    }
};

// Where the pipeline finds and keeps hierarchies (the serving loop's cache).
class HierarchyStore {
  public:
    virtual ~HierarchyStore() = default;
    // The kept hierarchy of key, held until release; nullptr if none.
    virtual const CachedHierarchy *acquire(const HierKey &key) = 0;
    virtual void release(const CachedHierarchy *h) = 0;
    // Whether a hierarchy of `bytes` may be kept; reserves the bytes if so.
    virtual bool reserve(size_t bytes) = 0;
    virtual void unreserve(size_t bytes) = 0;
    // Takes a new hierarchy (its bytes reserved); held by the caller until release.
    virtual const CachedHierarchy *insert(const HierKey &key, CachedHierarchy &&h) = 0;
};

// Frees a hierarchy's device allocation.
void free_hierarchy(CachedHierarchy &h);

// Partitions a device-resident graph g0 (arrays outside the arena) into k
// parts on context c and copies part[] to the host; the coarsening comes
// from `store` when it holds the hierarchy of `key`, else it is built and
// offered to the store.  Same result as exec_job on the same graph.
void exec_resident(Ctx &c, const hgp_options &o, const DevGraph &g0, int32_t k, int32_t *part, hgp_stats *stats,
                   HierarchyStore *store, int64_t graph, size_t bytes, size_t max_bytes = SIZE_MAX);

// Throws Error{-1/-4} for invalid options or k.
void check_options(const hgp_options &o, int32_t k);

// Arena bytes a partition of a resident graph of n vertices and m entries
// starts with.
size_t resident_job_bytes(int32_t n, int64_t m, int32_t k);

// Throws Error{-1/-4} for invalid options or CSR; touches no device.
void check_job(const hgp_options &o, const HostJob &j);
// Device arena bytes the job is run with first.
size_t job_bytes(const HostJob &j);
// Upload, partition, download on context c (after check_job), with an arena
// of `bytes` that may grow up to max_bytes when the run needs more.
void exec_job(Ctx &c, const hgp_options &o, const HostJob &j, size_t bytes, size_t max_bytes = SIZE_MAX);

// Runs f, turning exceptions into a status code and the thread's error text.
template <class F>
int guarded(F f)
{
    clear_error();
    try {
        f();
        return HGP_OK;
    } catch (const Error &e) {
        set_error(e.code, e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error(-3, "host out of memory");
        return HGP_ERR_OOM;
    } catch (const std::exception &e) {
        set_error(-2, e.what());
        return HGP_ERR_DEVICE;
    }
}

// Runs body with an arena of at least `bytes`; on arena overflow the arena
// grows to twice its size or to 5/4 of what the failed request needed,
// whichever is larger (up to max_bytes, three times at most), and the body
// is re-run (the body must be restartable).  Growing by the need lets a run
// whose coarsest graph stays large (its bisection trials' scratch grows with
// the square of that size) succeed on the first retry.
template <class F>
void with_arena(Ctx &c, size_t bytes, F body, size_t max_bytes = SIZE_MAX)
{
    for (int attempt = 0;; attempt++) {
        c.arena.reset();
        c.arena.reserve(bytes);
        try {
            body();
            HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
            c.arena.reset();
            return;
        } catch (const Error &e) {
            (void)hipStreamSynchronize(c.stream);
            c.arena.reset();
            if (!e.overflow || attempt >= 3 || bytes >= max_bytes) throw;
            size_t grow = std::max(bytes <= SIZE_MAX / 2 ? 2 * bytes : SIZE_MAX, e.need / 4 * 5 + (64 << 20));
            if (grow > max_bytes) {
                if (e.need > max_bytes) throw;
                grow = max_bytes;
            }
            bytes = grow;
        }
    }
}

}  // namespace hgp
