// hgp_api.cpp -- pipeline orchestration (coarsen -> initial partition ->
// uncoarsen/refine) and the C ABI.
#include "hunyuangraph.h"

#include "hgp_check.hpp"
#include "hgp_graph.hpp"
#include "hgp_prims.hpp"
#include "hgp_host.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

namespace hgp {

bool alloc_host_graph(hgp_graph *g, int32_t n, int64_t m);   // hgp_gen.cpp

namespace {

std::mutex g_api_mu;   // one call at a time per process on the default contexts

struct RunStats {
    hgp_stats *st;
};

// Algorithmic HBM bytes of a run: per level l with n_l vertices and m_l
// directed entries (x = 8, int64 offsets):
//   CSR_l = 8(n_l+1) + 8 m_l + 4 n_l
//   coarsening (l < L): 2 CSR_l + 8 m_l + 8 n_l + CSR_{l+1}
//   refinement: (1 + I_l)(CSR_l + 4 m_l + 8 n_l)
//   projection (l < L): 12 n_l
double alg_bytes(const std::vector<int64_t> &ns, const std::vector<int64_t> &ms, const std::vector<int32_t> &iters)
{
    auto csr = [&](size_t l) { return 8.0 * (ns[l] + 1) + 8.0 * ms[l] + 4.0 * ns[l]; };
    double b = 0;
    size_t L = ns.size() - 1;
    for (size_t l = 0; l <= L; l++) {
        if (l < L) b += 2 * csr(l) + 8.0 * ms[l] + 8.0 * ns[l] + csr(l + 1) + 12.0 * ns[l];
        b += (1.0 + iters[l]) * (csr(l) + 4.0 * ms[l] + 8.0 * ns[l]);
    }
    return b;
}

size_t estimate_arena(int32_t n, int64_t m, int32_t k, bool host_copy)
{
    // The level stack dominates: power-law graphs keep most of their entries
    // for several levels (the sum over levels reaches ~10 m), plus the
    // contraction temporaries of the largest level and the k-int rows of hub
    // rows cut into segments (at most m/4096 rows).
    size_t b = (size_t)112 * (size_t)m + (size_t)256 * (size_t)n + ((size_t)64 << 20) + (size_t)k * 4096 +
               ((size_t)m / 1024 + 1) * (size_t)k;
    if (host_copy) b += (size_t)8 * (size_t)(n + 1) + (size_t)8 * (size_t)m + (size_t)8 * (size_t)n;
    return b;
}

// Per-level probes of the coarsening loop (Params::level_stats): an event
// after each level is built, and the total edge weight of every level, taken
// after the loop so that the reductions stay out of the level times.
struct LevelProbe {
    bool on = false;
    std::vector<hipEvent_t> ev;
    int64_t *wsum = nullptr;   // device, one per level
    explicit LevelProbe(bool enabled) : on(enabled) {}
    ~LevelProbe()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void mark(Ctx &c)
    {
        if (!on) return;
        hipEvent_t e;
        HGP_CHECK_HIP(hipEventCreate(&e));
        ev.push_back(e);
        HGP_CHECK_HIP(hipEventRecord(e, c.stream));
    }
    void weigh(Ctx &c, const std::vector<DevGraph> &lv)
    {
        if (!on) return;
        wsum = c.arena.lalloc<int64_t>(lv.size());
        for (size_t l = 0; l < lv.size(); l++) reduce_sum(c, lv[l].adjwgt, lv[l].m, wsum + l);
    }
    // after the stream has passed the last mark
    void fill(Ctx &c, hgp_stats *st, size_t levels)
    {
        if (!on || !st) return;
        std::vector<int64_t> w(levels);
        HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
        HGP_CHECK_HIP(hipMemcpy(w.data(), wsum, sizeof(int64_t) * levels, hipMemcpyDeviceToHost));
        for (size_t l = 0; l < levels && l < HGP_MAX_LEVELS; l++) {
            st->level_adjwgt[l] = w[l];
            float ms = 0;
            if (l + 1 < ev.size()) HGP_CHECK_HIP(hipEventElapsedTime(&ms, ev[l], ev[l + 1]));
            st->level_ms[l] = ms;
        }
    }
};

// hunyuangarph_coarsen (hunyuangraph_GPU_coarsen.h:163-253): do-while with
// the stop rule n > thr && n < 0.85 n_finer && m > n/2 (exact integers).  lv
// holds level 0 on entry and every level on return; cmaps[l] maps level l to
// level l+1.
void coarsen_levels(Ctx &c, int32_t k, const Params &p, int64_t tvwgt, std::vector<DevGraph> &lv,
                    std::vector<int32_t *> &cmaps, LevelProbe &probe)
{
    int32_t thr = coarsen_threshold(k);
    int32_t maxvwgt = coarsen_maxvwgt(tvwgt, thr);
    probe.mark(c);
    attach_bins(c, lv[0]);
    for (;;) {
        const DevGraph f = lv.back();   // a copy: lv.push_back below may reallocate
        int32_t *cmap = c.arena.lalloc<int32_t>((size_t)f.n);
        size_t rmark = c.arena.right_mark();
        int32_t *match = c.arena.ralloc<int32_t>((size_t)f.n);
        int32_t level = (int32_t)lv.size() - 1;
        match_level(c, f, maxvwgt, p.seed, level, p.match_rounds, match);
        int32_t cn = build_cmap(c, f.n, match, cmap);
        DevGraph cg = contract(c, f, match, cmap, cn);
        c.arena.right_release(rmark);
        attach_bins(c, cg);
        probe.mark(c);
        cmaps.push_back(cmap);
        lv.push_back(cg);
        int32_t nf = f.n;
        bool go = cg.n > thr && 100 * (int64_t)cg.n < 85 * (int64_t)nf && cg.m > cg.n / 2;
        if (!go || (int)lv.size() >= HGP_MAX_LEVELS - 1) break;
    }
}

int64_t total_vwgt(Ctx &c, const DevGraph &g, int64_t *dsc)
{
    reduce_sum(c, g.vwgt, g.n, dsc);
    read_scalars(c, dsc, sizeof(int64_t));
    return c.pinned[0];
}

// Copies the coarsest level to the host graph h (arrays allocated by the
// caller) and fills the coarsening statistics.
void copy_coarsest(Ctx &c, const DevGraph &cg, hgp_graph &h, const std::vector<DevGraph> &lv, int64_t tvwgt,
                   LevelProbe &probe, hgp_stats *stats)
{
    probe.weigh(c, lv);
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    HGP_CHECK_HIP(hipMemcpy(h.xadj, cg.xadj, sizeof(int64_t) * ((size_t)cg.n + 1), hipMemcpyDeviceToHost));
    HGP_CHECK_HIP(hipMemcpy(h.adjncy, cg.adjncy, sizeof(int32_t) * (size_t)cg.m, hipMemcpyDeviceToHost));
    HGP_CHECK_HIP(hipMemcpy(h.adjwgt, cg.adjwgt, sizeof(int32_t) * (size_t)cg.m, hipMemcpyDeviceToHost));
    HGP_CHECK_HIP(hipMemcpy(h.vwgt, cg.vwgt, sizeof(int32_t) * (size_t)cg.n, hipMemcpyDeviceToHost));
    if (!stats) return;
    float t01 = 0;
    HGP_CHECK_HIP(hipEventElapsedTime(&t01, c.ev[0], c.ev[1]));
    stats->tvwgt = tvwgt;
    stats->levels = (int32_t)lv.size() - 1;
    stats->coarsest_n = cg.n;
    stats->coarsest_m = cg.m;
    stats->ms_coarsen = t01;
    stats->ms_total = t01;
    for (size_t l = 0; l < lv.size() && l < HGP_MAX_LEVELS; l++) {
        stats->level_n[l] = lv[l].n;
        stats->level_m[l] = lv[l].m;
    }
    stats->arena_bytes = (int64_t)c.arena.capacity();
    stats->arena_peak = (int64_t)c.arena.high_water();
    probe.fill(c, stats, lv.size());
}

// Copies the left-end region [m0, m1) of c's arena, which holds the levels
// built by coarsen_levels, into a device allocation of its own and points a
// copy of the level records at it.  Returns false (nothing kept) if the
// allocation fails.
bool snapshot_hierarchy(Ctx &c, const std::vector<DevGraph> &lv, const std::vector<int32_t *> &cmaps, size_t m0,
                        size_t m1, int64_t tvwgt, CachedHierarchy &h)
{
    const size_t bytes = m1 - m0;
    void *buf = nullptr;
    if (hipMalloc(&buf, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    char *lo = c.arena.base() + m0, *hi = c.arena.base() + m1;
    HGP_CHECK_HIP(hipMemcpyAsync(buf, lo, bytes, hipMemcpyDeviceToDevice, c.stream));
    // the copy completes before another stream can read the hierarchy
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    auto rb = [&](auto *ptr) {
        char *q = (char *)ptr;
        return (q >= lo && q < hi) ? (decltype(ptr))((char *)buf + (q - lo)) : ptr;
    };
    h.buf = buf;
    h.bytes = bytes;
    h.tvwgt = tvwgt;
    h.lv = lv;
    for (DevGraph &g : h.lv) {
        g.xadj = rb(g.xadj);
        g.adjncy = rb(g.adjncy);
        g.adjwgt = rb(g.adjwgt);
        g.vwgt = rb(g.vwgt);
        g.bins.ids = rb(g.bins.ids);
        g.segs.row = rb(g.segs.row);
        g.segs.j0 = rb(g.segs.j0);
        g.segs.mi = rb(g.segs.mi);
        g.segs.mrow = rb(g.segs.mrow);
        g.segs.mseg = rb(g.segs.mseg);
    }
    h.cmaps.clear();
    for (int32_t *cm : cmaps) h.cmaps.push_back(rb(cm));
    return true;
}

// The multilevel pipeline on a device-resident level-0 graph
// (hunyuangraph_kway_partition, hunyuangraph_partitiongraph.h:13-149).  With
// a store, the coarsening of `key` is taken from it when kept there, else
// built and offered to it.
void run_pipeline(Ctx &c, const DevGraph &g0, int32_t k, const Params &p, int32_t *d_part, hgp_stats *st,
                  HierarchyStore *store = nullptr, const HierKey *key = nullptr)
{
    int32_t n = g0.n;
    int64_t *dsc = c.arena.lalloc<int64_t>(8);
    HGP_CHECK_HIP(hipEventRecord(c.ev[0], c.stream));
    const CachedHierarchy *held = nullptr;
    if (store && k > 1 && !p.level_stats) held = store->acquire(*key);
    struct Release {
        HierarchyStore *s;
        const CachedHierarchy *&h;
        ~Release() { if (h) s->release(h); }
    } release_guard{store, held};
    int64_t tvwgt = held ? held->tvwgt : total_vwgt(c, g0, dsc);
    if (k == 1) {
        fill<int32_t>(c, d_part, n, 0);
        HGP_CHECK_HIP(hipEventRecord(c.ev[1], c.stream));
        HGP_CHECK_HIP(hipEventRecord(c.ev[2], c.stream));
        HGP_CHECK_HIP(hipEventRecord(c.ev[3], c.stream));
        HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
        if (st) { st->tvwgt = tvwgt; st->max_pwgt = tvwgt; }
        return;
    }
    std::vector<DevGraph> lv{g0};
    std::vector<int32_t *> cmaps;
    LevelProbe probe(p.level_stats && st);
    if (held) {
        lv = held->lv;
        cmaps = held->cmaps;
    } else {
        const size_t m0 = c.arena.left_mark();
        coarsen_levels(c, k, p, tvwgt, lv, cmaps, probe);
        const size_t m1 = c.arena.left_mark();
        if (store && !p.level_stats && store->reserve(m1 - m0)) {
            CachedHierarchy h;
            if (snapshot_hierarchy(c, lv, cmaps, m0, m1, tvwgt, h)) held = store->insert(*key, std::move(h));
            else store->unreserve(m1 - m0);
        }
    }
    HGP_CHECK_HIP(hipEventRecord(c.ev[1], c.stream));
    int32_t L = (int32_t)lv.size() - 1;
    const DevGraph &cg = lv[L];
    int32_t *where = c.arena.lalloc<int32_t>((size_t)cg.n);
    initial_partition(c, cg, k, p, where);
    int64_t *d_cut = dsc + 1;
    edge_cut(c, cg, where, d_cut);
    HGP_CHECK_HIP(hipEventRecord(c.ev[2], c.stream));
    read_scalars(c, d_cut, sizeof(int64_t));
    int64_t init_cut = c.pinned[0];
    std::vector<int32_t> iters(lv.size(), 0);
    for (int32_t l = L; l >= 0; l--) {
        RefineStats rs = refine_level(c, lv[l], k, tvwgt, l == 0, p, where, d_cut);
        iters[l] = rs.iters;
        if (l == 0) break;
        int32_t *wf = (l - 1 == 0) ? d_part : c.arena.lalloc<int32_t>((size_t)lv[l - 1].n);
        project(c, lv[l - 1].n, cmaps[l - 1], where, wf);
        where = wf;
    }
    if (L == 0) HGP_CHECK_HIP(hipMemcpyAsync(d_part, where, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, c.stream));
    HGP_CHECK_HIP(hipEventRecord(c.ev[3], c.stream));
    // final statistics
    int64_t *pw = c.arena.lalloc<int64_t>((size_t)k);
    part_weights(c, n, g0.vwgt, d_part, k, pw);
    std::vector<int64_t> hpw(k);
    HGP_CHECK_HIP(hipMemcpyAsync(hpw.data(), pw, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToHost, c.stream));
    read_scalars(c, d_cut, sizeof(int64_t));
    if (st) {
        float t01 = 0, t12 = 0, t23 = 0;
        HGP_CHECK_HIP(hipEventElapsedTime(&t01, c.ev[0], c.ev[1]));
        HGP_CHECK_HIP(hipEventElapsedTime(&t12, c.ev[1], c.ev[2]));
        HGP_CHECK_HIP(hipEventElapsedTime(&t23, c.ev[2], c.ev[3]));
        st->cut = c.pinned[0];
        int64_t mx = 0;
        for (int64_t x : hpw) mx = x > mx ? x : mx;
        st->max_pwgt = mx;
        st->tvwgt = tvwgt;
        st->levels = L;
        st->coarsest_n = cg.n;
        st->coarsest_m = cg.m;
        st->init_cut = init_cut;
        st->ms_coarsen = t01;
        st->ms_init = t12;
        st->ms_uncoarsen = t23;
        st->ms_total = (double)t01 + t12 + t23;
        std::vector<int64_t> ns, ms;
        int64_t tot = 0;
        for (size_t l = 0; l < lv.size(); l++) {
            ns.push_back(lv[l].n);
            ms.push_back(lv[l].m);
            if (l < HGP_MAX_LEVELS) {
                st->level_n[l] = lv[l].n;
                st->level_m[l] = lv[l].m;
                st->level_iters[l] = iters[l];
            }
            tot += iters[l];
        }
        st->refine_iters = tot;
        st->alg_bytes = alg_bytes(ns, ms, iters);
        st->arena_bytes = (int64_t)c.arena.capacity();
        st->arena_peak = (int64_t)c.arena.high_water();
    }
    if (probe.on) {
        probe.weigh(c, lv);
        probe.fill(c, st, lv.size());
    }
}

Params params_from(const hgp_options *o)
{
    Params p;
    if (!o) return p;
    p.seed = o->seed;
    p.ub_num = o->ub_num;
    p.ub_den = o->ub_den;
    p.match_rounds = o->match_rounds;
    p.trials = o->trials;
    p.fm_passes = o->fm_passes;
    p.patience = o->patience;
    p.level_stats = o->level_stats != 0;
    p.fm_kind = o->fm_kind;
    p.refiner = o->refiner;
    return p;
}

void check_params(const Params &p, int32_t k)
{
    if (k < 1) throw Error{-1, "nparts must be >= 1"};
    if (k > 4096) throw Error{-4, "nparts > 4096 is not supported"};
    if (p.ub_num <= 0 || p.ub_den <= 0 || p.ub_num < p.ub_den) throw Error{-1, "balance bound must be >= 1"};
    if (p.match_rounds < 0 || p.trials < 1 || p.fm_passes < 0 || p.patience < 1 || p.fm_kind < -1 ||
        p.fm_kind > 2 || p.refiner < 0 || p.refiner > 1)
        throw Error{-1, "invalid option values"};
}

template <class T>
T *to_dev(Ctx &c, const T *h, size_t n)
{
    T *d = c.arena.lalloc<T>(n);
    if (n) HGP_CHECK_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, c.stream));
    return d;
}

template <class T>
void to_host(Ctx &c, T *h, const T *d, size_t n)
{
    if (n) HGP_CHECK_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, c.stream));
}

DevGraph upload(Ctx &c, int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                const int32_t *adjwgt)
{
    DevGraph g;
    g.n = n;
    g.m = xadj[n];
    g.xadj = to_dev(c, xadj, (size_t)n + 1);
    g.adjncy = to_dev(c, adjncy, (size_t)g.m);
    if (adjwgt) g.adjwgt = to_dev(c, adjwgt, (size_t)g.m);
    else { g.adjwgt = c.arena.lalloc<wgt_t>((size_t)g.m); fill<int32_t>(c, g.adjwgt, g.m, 1); }
    if (vwgt) g.vwgt = to_dev(c, vwgt, (size_t)n);
    else { g.vwgt = c.arena.lalloc<wgt_t>((size_t)n); fill<int32_t>(c, g.vwgt, n, 1); }
    return g;
}

// Private contexts of the batch lanes, per device, kept for reuse (their
// arenas keep their size between calls).
std::vector<Ctx *> lane_contexts(int device, int count)
{
    static std::map<int, std::vector<std::unique_ptr<Ctx>>> pool;
    auto &v = pool[device];
    while ((int)v.size() < count) v.push_back(create_context(device));
    std::vector<Ctx *> out;
    for (int i = 0; i < count; i++) out.push_back(v[i].get());
    return out;
}

// ---------------------------------------------------------------- batched mode, small levels fused

// One graph of a fused batch group: its hierarchy, the partition buffers of
// the levels the fused launch writes, and the one-block loop's scratch, all
// in one device buffer (buf) that outlives the lane's arena between phases.
struct FusedGraph {
    int32_t i = -1;
    bool fused = false;          // false: finished in phase A (or failed)
    void *buf = nullptr;
    std::vector<DevGraph> lv;
    std::vector<int32_t *> cmaps;
    std::vector<int32_t *> where;   // by level: lo - 1 .. hi
    TinyScratch s;
    int64_t tvwgt = 0, init_cut = 0;
    int32_t lo = 0, hi = 0;
    double ms_a = 0;
};

// Copies the arena's left region [m0, m1) into its own device buffer and
// points every array of f into the copy.
void detach_fused(Ctx &c, size_t m0, size_t m1, FusedGraph &f)
{
    const size_t bytes = m1 - m0;
    HGP_CHECK_HIP(hipMalloc(&f.buf, bytes ? bytes : 1));
    char *lo = c.arena.base() + m0, *hi = c.arena.base() + m1;
    HGP_CHECK_HIP(hipMemcpyAsync(f.buf, lo, bytes, hipMemcpyDeviceToDevice, c.stream));
    HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
    auto rb = [&](auto *&ptr) {
        char *q = (char *)ptr;
        if (q >= lo && q < hi) ptr = (std::remove_reference_t<decltype(ptr)>)((char *)f.buf + (q - lo));
    };
    for (DevGraph &g : f.lv) {
        rb(g.xadj);
        rb(g.adjncy);
        rb(g.adjwgt);
        rb(g.vwgt);
        rb(g.bins.ids);
        rb(g.segs.row);
        rb(g.segs.j0);
        rb(g.segs.mi);
        rb(g.segs.mrow);
        rb(g.segs.mseg);
    }
    for (int32_t *&p : f.cmaps) rb(p);
    for (int32_t *&p : f.where) rb(p);
    TinyScratch &t = f.s;
    rb(t.dest), rb(t.pregain), rb(t.best), rb(t.elist), rb(t.under), rb(t.es), rb(t.ee), rb(t.pstart);
    rb(t.iters), rb(t.lock), rb(t.sel), rb(t.w8), rb(t.w16), rb(t.pd16), rb(t.pd32), rb(t.key), rb(t.val);
    rb(t.k2), rb(t.v2), rb(t.ab), rb(t.wsc), rb(t.bs), rb(t.pwgts), rb(t.cut), rb(t.ctl);
}

// Statistics of a fused batch graph from its level-0 partition part0.
void finish_fused_stats(Ctx &c, const HostJob &j, const FusedGraph &f, const int32_t *part0, const int64_t *d_cut,
                        const std::vector<int32_t> &iters)
{
    if (!j.stats) return;
    hgp_stats *st = j.stats;
    int64_t *pw = c.arena.lalloc<int64_t>((size_t)j.k);
    part_weights(c, j.n, f.lv[0].vwgt, part0, j.k, pw);
    std::vector<int64_t> hpw((size_t)j.k);
    HGP_CHECK_HIP(hipMemcpyAsync(hpw.data(), pw, sizeof(int64_t) * (size_t)j.k, hipMemcpyDeviceToHost, c.stream));
    read_scalars(c, d_cut, sizeof(int64_t));
    st->cut = c.pinned[0];
    int64_t mx = 0;
    for (int64_t x : hpw) mx = x > mx ? x : mx;
    st->max_pwgt = mx;
    st->tvwgt = f.tvwgt;
    const int32_t L = (int32_t)f.lv.size() - 1;
    st->levels = L;
    st->coarsest_n = f.lv[(size_t)L].n;
    st->coarsest_m = f.lv[(size_t)L].m;
    st->init_cut = f.init_cut;
    std::vector<int64_t> ns, ms;
    int64_t tot = 0;
    for (size_t l = 0; l < f.lv.size(); l++) {
        ns.push_back(f.lv[l].n);
        ms.push_back(f.lv[l].m);
        if (l < HGP_MAX_LEVELS) {
            st->level_n[l] = f.lv[l].n;
            st->level_m[l] = f.lv[l].m;
            st->level_iters[l] = iters[l];
        }
        tot += iters[l];
    }
    st->refine_iters = tot;
    st->alg_bytes = alg_bytes(ns, ms, iters);
    st->arena_bytes = (int64_t)c.arena.capacity();
    st->arena_peak = (int64_t)c.arena.high_water();
}

// Phase A of a fused batch graph: upload, coarsening, initial partition; the
// levels of at most fuse_n vertices at the coarse end are left to the fused
// launch.  A graph without such a level (or with k = 1, or k above the
// one-block loop's range) is finished here.
void fused_phase_a(Ctx &c, const hgp_options &o, const HostJob &j, int32_t fuse_n, FusedGraph &f)
{
    Params p = params_from(&o);
    HGP_CHECK_HIP(hipSetDevice(c.device));
    if (j.n == 0) return;
    if (j.k == 1 || j.k > kTinyMaxK) {
        exec_job(c, o, j, job_bytes(j));
        return;
    }
    with_arena(c, job_bytes(j), [&] {
        auto t0 = std::chrono::steady_clock::now();
        const size_t m0 = c.arena.left_mark();
        f.lv.assign(1, upload(c, j.n, j.xadj, j.adjncy, j.vwgt, j.adjwgt));
        f.cmaps.clear();
        int64_t *dsc = c.arena.lalloc<int64_t>(8);
        f.tvwgt = total_vwgt(c, f.lv[0], dsc);
        LevelProbe probe(false);
        coarsen_levels(c, j.k, p, f.tvwgt, f.lv, f.cmaps, probe);
        const int32_t L = (int32_t)f.lv.size() - 1;
        int32_t lo = L + 1;
        while (lo > 0 && f.lv[(size_t)lo - 1].n <= fuse_n) lo--;
        f.where.assign((size_t)L + 1, nullptr);
        f.where[(size_t)L] = c.arena.lalloc<int32_t>((size_t)f.lv[(size_t)L].n);
        initial_partition(c, f.lv[(size_t)L], j.k, p, f.where[(size_t)L]);
        if (lo > L) {
            // no small level: the rest of the pipeline here
            int64_t *d_cut = dsc + 1;
            edge_cut(c, f.lv[(size_t)L], f.where[(size_t)L], d_cut);
            read_scalars(c, d_cut, sizeof(int64_t));
            f.init_cut = c.pinned[0];
            int32_t *where = f.where[(size_t)L];
            std::vector<int32_t> iters(f.lv.size(), 0);
            for (int32_t l = L; l >= 0; l--) {
                iters[(size_t)l] = refine_level(c, f.lv[(size_t)l], j.k, f.tvwgt, l == 0, p, where, d_cut).iters;
                if (l == 0) break;
                int32_t *wf = c.arena.lalloc<int32_t>((size_t)f.lv[(size_t)l - 1].n);
                project(c, f.lv[(size_t)l - 1].n, f.cmaps[(size_t)l - 1], where, wf);
                where = wf;
            }
            finish_fused_stats(c, j, f, where, d_cut, iters);
            to_host(c, j.part, where, (size_t)j.n);
            return;
        }
        f.s = alloc_tiny_scratch(c, f.lv[(size_t)lo].n, j.k);
        edge_cut(c, f.lv[(size_t)L], f.where[(size_t)L], f.s.cut);
        read_scalars(c, f.s.cut, sizeof(int64_t));
        f.init_cut = c.pinned[0];
        for (int32_t l = std::max(lo - 1, 0); l < L; l++)
            f.where[(size_t)l] = c.arena.lalloc<int32_t>((size_t)f.lv[(size_t)l].n);
        f.lo = lo;
        f.hi = L;
        detach_fused(c, m0, c.arena.left_mark(), f);
        f.fused = true;
        f.ms_a = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

// Phase C: the levels finer than the fused ones, final statistics, the
// partition to the host; frees the graph's buffer.
void fused_phase_c(Ctx &c, const hgp_options &o, const HostJob &j, FusedGraph &f)
{
    Params p = params_from(&o);
    HGP_CHECK_HIP(hipSetDevice(c.device));
    with_arena(c, resident_job_bytes(j.n, j.xadj[j.n], j.k), [&] {
        int64_t *d_cut = c.arena.lalloc<int64_t>(1);
        HGP_CHECK_HIP(hipMemcpyAsync(d_cut, f.s.cut, sizeof(int64_t), hipMemcpyDeviceToDevice, c.stream));
        std::vector<int32_t> iters(f.lv.size(), 0);
        HGP_CHECK_HIP(hipMemcpyAsync(iters.data(), f.s.iters, sizeof(int32_t) * iters.size(), hipMemcpyDeviceToHost,
                                     c.stream));
        HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
        int32_t *where = f.lo > 0 ? f.where[(size_t)f.lo - 1] : f.where[0];
        for (int32_t l = f.lo - 1; l >= 0; l--) {
            iters[(size_t)l] = refine_level(c, f.lv[(size_t)l], j.k, f.tvwgt, l == 0, p, where, d_cut).iters;
            if (l == 0) break;
            int32_t *wf = c.arena.lalloc<int32_t>((size_t)f.lv[(size_t)l - 1].n);
            project(c, f.lv[(size_t)l - 1].n, f.cmaps[(size_t)l - 1], where, wf);
            where = wf;
        }
        finish_fused_stats(c, j, f, where, d_cut, iters);
        to_host(c, j.part, where, (size_t)j.n);
    });
}

}  // namespace

size_t job_bytes(const HostJob &j) { return estimate_arena(j.n, j.n > 0 ? j.xadj[j.n] : 0, j.k, true); }

void check_job(const hgp_options &o, const HostJob &j)
{
    Params p = params_from(&o);
    check_params(p, j.k);
    if (j.stats) memset(j.stats, 0, sizeof(*j.stats));
    validate_host(j.n, j.xadj, j.adjncy, j.vwgt, j.adjwgt, o.validate != 0);
    if (j.n > 0 && !j.part) throw Error{-1, "part is NULL"};
}

void free_hierarchy(CachedHierarchy &h)
{
    if (h.buf) (void)hipFree(h.buf);
    h.buf = nullptr;
    h.bytes = 0;
    h.lv.clear();
    h.cmaps.clear();
}

size_t resident_job_bytes(int32_t n, int64_t m, int32_t k) { return estimate_arena(n, m, k, false); }

void check_options(const hgp_options &o, int32_t k) { check_params(params_from(&o), k); }

void exec_resident(Ctx &c, const hgp_options &o, const DevGraph &g0, int32_t k, int32_t *part, hgp_stats *stats,
                   HierarchyStore *store, int64_t graph, size_t bytes, size_t max_bytes)
{
    if (g0.n == 0) return;
    Params p = params_from(&o);
    check_params(p, k);
    if (stats) memset(stats, 0, sizeof(*stats));
    HierKey key{graph, p.seed, coarsen_threshold(k), p.match_rounds};
    HGP_CHECK_HIP(hipSetDevice(c.device));
    with_arena(c, bytes, [&] {
        int32_t *d_part = c.arena.lalloc<int32_t>((size_t)g0.n);
        run_pipeline(c, g0, k, p, d_part, stats, store, &key);
        to_host(c, part, d_part, (size_t)g0.n);
    }, max_bytes);
}

void exec_job(Ctx &c, const hgp_options &o, const HostJob &j, size_t bytes, size_t max_bytes)
{
    if (j.n == 0) return;
    Params p = params_from(&o);
    HGP_CHECK_HIP(hipSetDevice(c.device));
    with_arena(c, bytes, [&] {
        DevGraph g = upload(c, j.n, j.xadj, j.adjncy, j.vwgt, j.adjwgt);
        int32_t *d_part = c.arena.lalloc<int32_t>((size_t)j.n);
        run_pipeline(c, g, j.k, p, d_part, j.stats);
        to_host(c, j.part, d_part, (size_t)j.n);
    }, max_bytes);
}

}  // namespace hgp

using namespace hgp;

extern "C" {

const char *hgp_last_error(void) { return hgp::last_error_text(); }

const char *hgp_version(void) { return "hgp 0.2.0 (gfx950)"; }

int hgp_device_count(void)
{
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cnt;
}

void hgp_default_options(hgp_options *o)
{
    if (!o) return;
    o->seed = 1;
    o->ub_num = 103;
    o->ub_den = 100;
    o->match_rounds = 8;
    o->trials = 1024;
    o->fm_passes = 10;
    o->patience = 6;
    o->fm_kind = -1;
    o->validate = 0;
    o->device = 0;
    o->level_stats = 0;
    o->refiner = 0;
}

int hgp_partition(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                  const int32_t *adjwgt, int32_t k, const hgp_options *opt, int32_t *part, hgp_stats *stats)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        std::lock_guard<std::mutex> lk(g_api_mu);
        HostJob j{n, xadj, adjncy, vwgt, adjwgt, k, part, stats};
        check_job(o, j);
        if (n > 0) exec_job(context(o.device), o, j, job_bytes(j));
    });
}

int hgp_partition_batch(int32_t count, const int32_t *n, const int64_t *const *xadj, const int32_t *const *adjncy,
                        const int32_t *const *vwgt, const int32_t *const *adjwgt, const int32_t *k,
                        const hgp_options *opt, int32_t lanes, int32_t *const *part, int32_t *status,
                        hgp_stats *stats)
{
    return guarded([&] {
        if (count < 0) throw Error{-1, "count < 0"};
        if (count == 0) return;
        if (!n || !xadj || !adjncy || !k || !part) throw Error{-1, "NULL array argument"};
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        int nl = lanes > 0 ? lanes : tuning().batch_lanes;
        nl = std::min(std::min(nl, 64), count);
        std::lock_guard<std::mutex> lk(g_api_mu);
        std::vector<Ctx *> ctx = lane_contexts(o.device, nl);
        // largest graphs first, so the lanes finish together
        std::vector<int32_t> order((size_t)count);
        std::iota(order.begin(), order.end(), 0);
        auto size_of = [&](int32_t i) { return xadj[i] && n[i] >= 0 ? xadj[i][n[i]] + n[i] : 0; };
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return size_of(a) > size_of(b); });
        std::vector<int32_t> st((size_t)count, 0);
        std::vector<std::string> msg((size_t)count);
        auto job_of = [&](int32_t i) {
            if (!xadj[i]) throw Error{-1, "xadj is NULL"};
            return HostJob{n[i], xadj[i], adjncy[i], vwgt ? vwgt[i] : nullptr, adjwgt ? adjwgt[i] : nullptr, k[i],
                           part[i], stats ? stats + i : nullptr};
        };
        // runs body(c, t) for t in [0, cnt) over the lanes
        auto over_lanes = [&](int32_t cnt, auto body) {
            std::atomic<int32_t> next{0};
            auto work = [&](Ctx *c) {
                for (int32_t t; (t = next.fetch_add(1)) < cnt;) body(c, t);
            };
            std::vector<std::thread> th;
            for (int l = 1; l < std::min(nl, (int)cnt); l++) th.emplace_back(work, ctx[l]);
            work(ctx[0]);
            for (auto &t : th) t.join();
        };
        const int32_t fuse_n = std::min(tuning().batch_fuse_max, kTinyMaxN);
        if (fuse_n > 0 && !o.level_stats && o.refiner == 0) {
            // Groups of graphs (largest first, up to ~8 GB of hierarchies):
            // phase A coarsens and initially partitions each graph on the
            // lanes and keeps its hierarchy; one launch refines the small
            // levels of every graph of the group (a block per graph);
            // phase C refines the rest on the lanes.
            const size_t budget = (size_t)8 << 30;
            for (int32_t g0 = 0; g0 < count;) {
                int32_t g1 = g0;
                size_t gb = 0;
                while (g1 < count && (g1 == g0 || gb < budget)) {
                    const int32_t i = order[(size_t)g1];
                    gb += xadj[i] && n[i] >= 0 ? (size_t)(24 * (xadj[i][n[i]] + 2 * (int64_t)n[i])) + (1 << 20) : 0;
                    g1++;
                }
                std::vector<FusedGraph> fg((size_t)(g1 - g0));
                over_lanes(g1 - g0, [&](Ctx *c, int32_t t) {
                    const int32_t i = order[(size_t)(g0 + t)];
                    fg[(size_t)t].i = i;
                    int rc = guarded([&] {
                        HostJob j = job_of(i);
                        check_job(o, j);
                        fused_phase_a(*c, o, j, fuse_n, fg[(size_t)t]);
                    });
                    st[i] = rc;
                    if (rc != HGP_OK) msg[i] = last_error_text();
                });
                std::vector<TinyRun> runs;
                std::vector<size_t> idx;
                for (size_t t = 0; t < fg.size(); t++) {
                    FusedGraph &f = fg[t];
                    if (!f.fused || st[f.i] != HGP_OK) continue;
                    TinyRun r;
                    r.k = k[f.i];
                    r.hi = f.hi;
                    r.lo = f.lo;
                    r.tvwgt = f.tvwgt;
                    r.p = params_from(&o);
                    r.lv = f.lv;
                    r.where = f.where;
                    r.cmap.assign(f.cmaps.begin(), f.cmaps.end());
                    r.s = f.s;
                    runs.push_back(std::move(r));
                    idx.push_back(t);
                }
                int brc = guarded([&] {
                    HGP_CHECK_HIP(hipSetDevice(ctx[0]->device));
                    with_arena(*ctx[0], (sizeof(TinyRun) + 2048) * (runs.size() + 1) * 4 + (4 << 20),
                               [&] { tiny_batch(*ctx[0], runs); });
                });
                if (brc != HGP_OK)
                    for (size_t t : idx) {
                        st[fg[t].i] = brc;
                        msg[fg[t].i] = last_error_text();
                    }
                over_lanes((int32_t)idx.size(), [&](Ctx *c, int32_t t) {
                    FusedGraph &f = fg[idx[(size_t)t]];
                    if (st[f.i] != HGP_OK) return;
                    int rc = guarded([&] { fused_phase_c(*c, o, job_of(f.i), f); });
                    st[f.i] = rc;
                    if (rc != HGP_OK) msg[f.i] = last_error_text();
                });
                for (FusedGraph &f : fg)
                    if (f.buf) (void)hipFree(f.buf);
                g0 = g1;
            }
            if (status) std::copy(st.begin(), st.end(), status);
            for (int32_t i = 0; i < count; i++)
                if (st[i] != HGP_OK) throw Error{st[i], "graph " + std::to_string(i) + ": " + msg[i]};
            return;
        }
        std::atomic<int32_t> next{0};
        auto work = [&](Ctx *c) {
            for (;;) {
                int32_t t = next.fetch_add(1);
                if (t >= count) break;
                int32_t i = order[t];
                int rc = guarded([&] {
                    if (!xadj[i]) throw Error{-1, "xadj is NULL"};
                    HostJob j{n[i], xadj[i], adjncy[i], vwgt ? vwgt[i] : nullptr, adjwgt ? adjwgt[i] : nullptr, k[i],
                              part[i], stats ? stats + i : nullptr};
                    check_job(o, j);
                    exec_job(*c, o, j, job_bytes(j));
                });
                st[i] = rc;
                if (rc != HGP_OK) msg[i] = last_error_text();
            }
        };
        std::vector<std::thread> th;
        for (int l = 1; l < nl; l++) th.emplace_back(work, ctx[l]);
        work(ctx[0]);
        for (auto &t : th) t.join();
        if (status) std::copy(st.begin(), st.end(), status);
        for (int32_t i = 0; i < count; i++)
            if (st[i] != HGP_OK) throw Error{st[i], "graph " + std::to_string(i) + ": " + msg[i]};
    });
}

int hgp_partition_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_vwgt,
                         const int32_t *d_adjwgt, int32_t k, const hgp_options *opt, int32_t *d_part,
                         hgp_stats *stats)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        Params p = params_from(&o);
        check_params(p, k);
        if (stats) memset(stats, 0, sizeof(*stats));
        if (n < 0) throw Error{-1, "nvtxs < 0"};
        if (n == 0) return;
        if (!d_xadj || !d_adjncy || !d_vwgt || !d_adjwgt || !d_part) throw Error{-1, "NULL device pointer"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(o.device);
        HGP_CHECK_HIP(hipSetDevice(c.device));
        int64_t m = 0;
        HGP_CHECK_HIP(hipMemcpy(&m, d_xadj + n, sizeof(int64_t), hipMemcpyDeviceToHost));
        if (m < 0) throw Error{-1, "xadj[n] < 0"};
        if (m > INT32_MAX) throw Error{-4, "more than 2^31-1 directed entries"};
        with_arena(c, estimate_arena(n, m, k, false), [&] {
            // the CSR invariants of check_graph (hunyuangraph_GPU_match.h:1718-1800),
            // checked on the device before any kernel gathers through it
            check_graph_device(c, n, d_xadj, d_adjncy, d_vwgt, d_adjwgt, o.validate != 0);
            DevGraph g;
            g.n = n;
            g.m = m;
            g.xadj = const_cast<eid_t *>(d_xadj);
            g.adjncy = const_cast<vid_t *>(d_adjncy);
            g.adjwgt = const_cast<wgt_t *>(d_adjwgt);
            g.vwgt = const_cast<wgt_t *>(d_vwgt);
            run_pipeline(c, g, k, p, d_part, stats);
        });
    });
}

int hgp_check_graph_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_vwgt,
                           const int32_t *d_adjwgt, int32_t symmetric, int32_t device)
{
    return guarded([&] {
        if (n < 0) throw Error{-1, "nvtxs < 0"};
        if (n == 0) return;
        if (!d_xadj) throw Error{-1, "xadj is NULL"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(device);
        HGP_CHECK_HIP(hipSetDevice(c.device));
        with_arena(c, 1 << 20, [&] { check_graph_device(c, n, d_xadj, d_adjncy, d_vwgt, d_adjwgt, symmetric != 0); });
    });
}

int hgp_evaluate_device(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_adjwgt,
                        const int32_t *d_vwgt, const int32_t *d_part, int32_t k, int32_t device, int64_t *d_out)
{
    return guarded([&] {
        if (n < 0) throw Error{-1, "nvtxs < 0"};
        if (k < 1) throw Error{-1, "nparts must be >= 1"};
        if (k > kEvalMaxParts) throw Error{-4, "nparts > 4096 is not supported"};
        if (!d_out) throw Error{-1, "out is NULL"};
        if (n > 0 && (!d_xadj || !d_adjncy || !d_part)) throw Error{-1, "NULL device pointer"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(device);
        HGP_CHECK_HIP(hipSetDevice(c.device));
        with_arena(c, sizeof(int64_t) * ((size_t)kEvalHead + (size_t)k) + 4096, [&] {
            int64_t *scr = c.arena.ralloc<int64_t>((size_t)kEvalHead + (size_t)k);
            evaluate_launch(c.stream, n, d_xadj, d_adjncy, d_adjwgt, d_vwgt, d_part, k, scr, d_out);
        });
    });
}

int hgp_evaluate_device_noalloc(int32_t n, const int64_t *d_xadj, const int32_t *d_adjncy, const int32_t *d_adjwgt,
                                const int32_t *d_vwgt, const int32_t *d_part, int32_t k, int32_t device,
                                int64_t *d_out)
{
    return guarded([&] {
        if (n < 0) throw Error{-1, "nvtxs < 0"};
        if (k < 1) throw Error{-1, "nparts must be >= 1"};
        if (k > kEvalMaxParts) throw Error{-4, "nparts > 4096 is not supported"};
        if (!d_out) throw Error{-1, "out is NULL"};
        if (n > 0 && (!d_xadj || !d_adjncy || !d_part)) throw Error{-1, "NULL device pointer"};
        evaluate_fixed_scratch(device, n, d_xadj, d_adjncy, d_adjwgt, d_vwgt, d_part, k, d_out);
    });
}

int hgp_coarsen(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt, const int32_t *adjwgt,
                int32_t k, const hgp_options *opt, hgp_graph *coarsest, hgp_stats *stats)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        Params p = params_from(&o);
        check_params(p, k);
        if (!coarsest) throw Error{-1, "coarsest is NULL"};
        memset(coarsest, 0, sizeof(*coarsest));
        if (stats) memset(stats, 0, sizeof(*stats));
        validate_host(n, xadj, adjncy, vwgt, adjwgt, o.validate != 0);
        if (n == 0) throw Error{-1, "the graph has no vertices"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(o.device);
        HGP_CHECK_HIP(hipSetDevice(c.device));
        with_arena(c, estimate_arena(n, xadj[n], k, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            int64_t *dsc = c.arena.lalloc<int64_t>(8);
            HGP_CHECK_HIP(hipEventRecord(c.ev[0], c.stream));
            int64_t tvwgt = total_vwgt(c, g, dsc);
            std::vector<DevGraph> lv{g};
            std::vector<int32_t *> cmaps;
            LevelProbe probe(p.level_stats && stats);
            coarsen_levels(c, k, p, tvwgt, lv, cmaps, probe);
            HGP_CHECK_HIP(hipEventRecord(c.ev[1], c.stream));
            const DevGraph &cg = lv.back();
            hgp_graph h{};
            const bool got = alloc_host_graph(&h, cg.n, cg.m);
            try {
                if (!got) throw Error{-3, "host out of memory"};
                copy_coarsest(c, cg, h, lv, tvwgt, probe, stats);
            } catch (...) {
                hgp_graph_free(&h);
                throw;
            }
            *coarsest = h;
        });
    });
}

void hunyuangraph_PartitionGraph(int *nvtxs, int *xadj, int *adjncy, int *vwgt, int *adjwgt, int *nparts,
                                 float *tpwgts, float *ubvec, int *part)
{
    (void)tpwgts;   // uniform target weights, as in the reference (hunyuangraph_admin.h:19-22)
    guarded([&] {
        if (!nvtxs || !xadj || !nparts) throw Error{-1, "NULL argument"};
        int32_t n = *nvtxs;
        std::vector<int64_t> x64((size_t)n + 1);
        for (int32_t i = 0; i <= n; i++) x64[i] = xadj[i];
        hgp_options o;
        hgp_default_options(&o);
        if (ubvec && *ubvec > 1.0f) {
            o.ub_num = (int32_t)std::lround((double)*ubvec * 1000.0);
            o.ub_den = 1000;
        }
        int rc = hgp_partition(n, x64.data(), adjncy, vwgt, adjwgt, *nparts, &o, part, nullptr);
        if (rc != HGP_OK) throw Error{rc, hgp_last_error()};
    });
}

// ---------------------------------------------------------------- tuning

int hgp_get_tuning(hgp_tuning *t)
{
    return guarded([&] {
        if (!t) throw Error{-1, "tuning is NULL"};
        const Tuning &x = tuning();
        t->tiny_max_vertices = x.tiny_max;
        t->few_rows = (int32_t)x.few_rows;
        t->fill_waves = (int32_t)x.fill_waves;
        t->seg_grid = x.seg_grid;
        t->batch_lanes = x.batch_lanes;
        t->hub_track = x.hub_track;
        t->graph_max_vertices = x.graph_max_vertices;
        t->batch_fuse_max_vertices = x.batch_fuse_max;
    });
}

int hgp_set_tuning(const hgp_tuning *t)
{
    return guarded([&] {
        if (!t) throw Error{-1, "tuning is NULL"};
        if (t->tiny_max_vertices < 0 || t->tiny_max_vertices > 4096 || t->few_rows < 0 || t->fill_waves < 1 ||
            t->seg_grid < 1 || t->batch_lanes < 1 || t->batch_lanes > 64 || t->hub_track < 0 || t->hub_track > 1 ||
            t->graph_max_vertices < 0 || t->batch_fuse_max_vertices < 0 || t->batch_fuse_max_vertices > kTinyMaxN)
            throw Error{-1, "tuning value out of range"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Tuning &x = tuning();
        x.tiny_max = t->tiny_max_vertices;
        x.few_rows = t->few_rows;
        x.fill_waves = t->fill_waves;
        x.seg_grid = t->seg_grid;
        x.batch_lanes = t->batch_lanes;
        x.hub_track = t->hub_track;
        x.graph_max_vertices = t->graph_max_vertices;
        x.batch_fuse_max = t->batch_fuse_max_vertices;
    });
}

// ---------------------------------------------------------------- Jet trace

int hgp_jet_trace_enable(int on)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(jet_trace().mu);
        jet_trace().on = on != 0;
        jet_trace().recs.clear();
    });
}

int hgp_jet_trace_dump(int64_t n, int32_t iter)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(jet_trace().mu);
        jet_trace().dump_n = n;
        jet_trace().dump_iter = iter;
        jet_trace().where.clear();
        jet_trace().dest.clear();
    });
}

int64_t hgp_jet_trace_dumped(int32_t *where, int32_t *dest, int64_t cap)
{
    std::lock_guard<std::mutex> lk(jet_trace().mu);
    const JetTrace &t = jet_trace();
    int64_t c = std::min<int64_t>(cap, (int64_t)t.where.size());
    if (where && c > 0) std::copy(t.where.begin(), t.where.begin() + c, where);
    if (dest && c > 0) std::copy(t.dest.begin(), t.dest.begin() + c, dest);
    return (int64_t)t.where.size();
}

int64_t hgp_jet_trace(int64_t *buf, int64_t cap)
{
    std::lock_guard<std::mutex> lk(jet_trace().mu);
    const std::vector<int64_t> &r = jet_trace().recs;
    if (buf && cap > 0) std::copy(r.begin(), r.begin() + std::min<int64_t>(cap, (int64_t)r.size()), buf);
    return (int64_t)r.size();
}

// ---------------------------------------------------------------- kernel timer

int hgp_ktimer_enable(int on)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_api_mu);
        KTimer &t = ktimer();
        t.on = on != 0;
        t.recs.clear();
        t.used = 0;
        t.ncnt = 0;
    });
}

int64_t hgp_ktimer_report(char *buf, int64_t len)
{
    std::string out = "{";
    try {
        std::lock_guard<std::mutex> lk(g_api_mu);
        KTimer &t = ktimer();
        struct Agg { int64_t n = 0; double ms = 0, bytes = 0; };
        std::vector<std::pair<std::string, Agg>> agg;
        std::vector<unsigned long long> cnt((size_t)t.ncnt);
        if (t.ncnt > 0) {
            HGP_CHECK_HIP(hipDeviceSynchronize());
            HGP_CHECK_HIP(hipMemcpy(cnt.data(), t.dcnt, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost));
        }
        for (auto &r : t.recs) {
            HGP_CHECK_HIP(hipEventSynchronize(r.b));
            float ms = 0;
            HGP_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = std::find_if(agg.begin(), agg.end(), [&](auto &x) { return x.first == r.name; });
            if (it == agg.end()) { agg.push_back({r.name, Agg{}}); it = agg.end() - 1; }
            it->second.n++;
            it->second.ms += ms;
            double dyn = 0;
            if (r.slot >= 0)
                for (int i = 0; i < kKbLanes; i++) dyn += (double)cnt[(size_t)r.slot + i];
            it->second.bytes += r.bytes + dyn;
        }
        char tmp[256];
        for (size_t i = 0; i < agg.size(); i++) {
            snprintf(tmp, sizeof(tmp), "%s\"%s\": [%lld, %.6f, %.1f]", i ? ", " : "", agg[i].first.c_str(),
                     (long long)agg[i].second.n, agg[i].second.ms, agg[i].second.bytes);
            out += tmp;
        }
    } catch (const Error &e) {
        set_error(e.code, e.msg);
        return e.code;
    }
    out += "}";
    if (buf && len > 0) {
        size_t nb = std::min((size_t)len - 1, out.size());
        memcpy(buf, out.data(), nb);
        buf[nb] = 0;
    }
    return (int64_t)out.size() + 1;
}

// ---------------------------------------------------------------- stages

int hgp_scan_i32(const int32_t *in, int32_t *out, int64_t n)
{
    return guarded([&] {
        if (n < 0) throw Error{-1, "n < 0"};
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, (size_t)n * 12 + (16 << 20), [&] {
            int32_t *d = to_dev(c, in, (size_t)n);
            scan_inclusive(c, d, d, n);
            to_host(c, out, d, (size_t)n);
        });
    });
}

int hgp_scan_i64(const int64_t *in, int64_t *out, int64_t n)
{
    return guarded([&] {
        if (n < 0) throw Error{-1, "n < 0"};
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, (size_t)n * 24 + (16 << 20), [&] {
            int64_t *d = to_dev(c, in, (size_t)n);
            scan_inclusive(c, d, d, n);
            to_host(c, out, d, (size_t)n);
        });
    });
}

int hgp_part_weights(int32_t n, const int32_t *vwgt, const int32_t *where, int32_t k, int64_t *pwgts)
{
    return guarded([&] {
        if (n < 0 || k < 1) throw Error{-1, "bad sizes"};
        for (int32_t v = 0; v < n; v++)
            if (where[v] < 0 || where[v] >= k) throw Error{-1, "part id out of range"};
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, (size_t)n * 16 + (size_t)k * 16 + (16 << 20), [&] {
            int32_t *dv = to_dev(c, vwgt, (size_t)n);
            int32_t *dw = to_dev(c, where, (size_t)n);
            int64_t *dp = c.arena.lalloc<int64_t>((size_t)k);
            part_weights(c, n, dv, dw, k, dp);
            to_host(c, pwgts, dp, (size_t)k);
        });
    });
}

int hgp_edge_cut(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt, const int32_t *where,
                 int64_t *cut)
{
    return guarded([&] {
        validate_host(n, xadj, adjncy, nullptr, adjwgt, false);
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, estimate_arena(n, xadj[n], 1, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, nullptr, adjwgt);
            attach_bins(c, g);
            int32_t *dw = to_dev(c, where, (size_t)n);
            int64_t *dc = c.arena.lalloc<int64_t>(1);
            edge_cut(c, g, dw, dc);
            to_host(c, cut, dc, 1);
        });
    });
}

int hgp_match(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt, const int32_t *adjwgt,
              int32_t maxvwgt, uint64_t seed, int32_t level, int32_t rounds, int32_t *match, int64_t *matched)
{
    return guarded([&] {
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        if (n == 0) { if (matched) *matched = 0; return; }
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, estimate_arena(n, xadj[n], 1, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            attach_bins(c, g);
            int32_t *dm = c.arena.lalloc<int32_t>((size_t)n);
            int64_t mc = match_level(c, g, maxvwgt, seed, level, rounds, dm);
            to_host(c, match, dm, (size_t)n);
            if (matched) *matched = mc;
        });
    });
}

int hgp_contract(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt, const int32_t *vwgt,
                 int32_t *match, int32_t *cmap, int32_t *cn, int64_t *cxadj, int32_t *cadjncy, int32_t *cadjwgt,
                 int32_t *cvwgt, int64_t *cm)
{
    return guarded([&] {
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        for (int32_t v = 0; v < n; v++) {
            int32_t u = match[v];
            if (u < -1 || u >= n) throw Error{-1, "match entry out of range"};
            if (u >= 0 && u != v && match[u] != v) throw Error{-1, "match is not symmetric"};
        }
        if (n == 0) { *cn = 0; *cm = 0; cxadj[0] = 0; return; }
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(0);
        with_arena(c, estimate_arena(n, xadj[n], 1, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            int32_t *dm = to_dev(c, match, (size_t)n);
            int32_t *dc = c.arena.lalloc<int32_t>((size_t)n);
            int32_t ncn = build_cmap(c, n, dm, dc);
            DevGraph cg = contract(c, g, dm, dc, ncn);
            HGP_CHECK_HIP(hipStreamSynchronize(c.stream));
            to_host(c, match, dm, (size_t)n);
            to_host(c, cmap, dc, (size_t)n);
            to_host(c, cxadj, cg.xadj, (size_t)ncn + 1);
            to_host(c, cadjncy, cg.adjncy, (size_t)cg.m);
            to_host(c, cadjwgt, cg.adjwgt, (size_t)cg.m);
            to_host(c, cvwgt, cg.vwgt, (size_t)ncn);
            *cn = ncn;
            *cm = cg.m;
        });
    });
}

int hgp_initpart(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt, const int32_t *adjwgt,
                 int32_t k, const hgp_options *opt, int32_t *where)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        Params p = params_from(&o);
        check_params(p, k);
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(o.device);
        size_t extra = (size_t)std::min(p.trials, n) * (size_t)n * 16;
        with_arena(c, estimate_arena(n, xadj[n], k, true) + extra, [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            int32_t *dw = c.arena.lalloc<int32_t>((size_t)n);
            initial_partition(c, g, k, p, dw);
            to_host(c, where, dw, (size_t)n);
        });
    });
}

int hgp_refine(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt, const int32_t *adjwgt,
               int32_t k, int64_t tvwgt, int32_t is_finest, const hgp_options *opt, int32_t *where, int64_t *cut,
               int32_t *iters)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        Params p = params_from(&o);
        check_params(p, k);
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        for (int32_t v = 0; v < n; v++)
            if (where[v] < 0 || where[v] >= k) throw Error{-1, "part id out of range"};
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(o.device);
        with_arena(c, estimate_arena(n, xadj[n], k, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            attach_bins(c, g);
            int32_t *dw = to_dev(c, where, (size_t)n);
            int64_t *dc = to_dev(c, cut, 1);
            RefineStats rs = refine_level(c, g, k, tvwgt, is_finest != 0, p, dw, dc);
            to_host(c, where, dw, (size_t)n);
            to_host(c, cut, dc, 1);
            if (iters) *iters = rs.iters;
        });
    });
}

int hgp_refine_2way_lp(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                       const int32_t *adjwgt, const hgp_options *opt, int32_t *where, int32_t *ran)
{
    return guarded([&] {
        hgp_options o;
        if (opt) o = *opt; else hgp_default_options(&o);
        validate_host(n, xadj, adjncy, vwgt, adjwgt, false);
        for (int32_t v = 0; v < n; v++)
            if (where[v] < 0 || where[v] > 1) throw Error{-1, "side out of range (0 or 1)"};
        if (ran) *ran = 0;
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(g_api_mu);
        Ctx &c = context(o.device);
        with_arena(c, estimate_arena(n, xadj[n], 2, true), [&] {
            DevGraph g = upload(c, n, xadj, adjncy, vwgt, adjwgt);
            int32_t *dw = to_dev(c, where, (size_t)n);
            const bool r = lp2_refine(c, g, dw);
            to_host(c, where, dw, (size_t)n);
            if (ran) *ran = r ? 1 : 0;
        });
    });
}

}  // extern "C"
