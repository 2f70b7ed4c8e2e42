// hgp_gen.cpp -- seeded synthetic graph generators (host code).
//
// Every generator is deterministic in its arguments: randomness comes from a
// counter-based SplitMix64 stream, so the same seed gives the same CSR on any
// machine and with any thread count.  Rows are sorted by neighbour id; there
// are no self-loops and no duplicate entries.
#include "hunyuangraph.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include <chrono>
#include <cstdio>

namespace hgp {

// The arrays of a library-allocated host graph (n vertices, m directed
// entries); false if any allocation failed (hgp_graph_free releases the rest).
bool alloc_host_graph(hgp_graph *g, int32_t n, int64_t m)
{
    g->n = n;
    g->m = m;
    g->xadj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    g->adjncy = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    g->adjwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    g->vwgt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    return g->xadj && g->adjncy && g->adjwgt && g->vwgt;
}

}  // namespace hgp

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed ^ mix64(stream)) ^ i);
}

inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool gen_timing() { return getenv("HGP_GEN_TIMING") != nullptr; }

#define GEN_T(label) do { if (gen_timing()) fprintf(stderr, "gen %-12s %.3f s\n", label, now_s() - t0); } while (0)

int nthreads()
{
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 4;
    if (t > 32) t = 32;
    return (int)t;
}

template <class F>
void parallel_for(int64_t n, F f)
{
    int nt = nthreads();
    if (n < 1 << 16) nt = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=, &f] { for (int64_t i = lo; i < hi; i++) f(i); });
    }
    for (auto &x : th) x.join();
}

// LSD radix sort of 64-bit keys (8 passes of 8 bits, skipping constant digits).
void radix_sort_u64(std::vector<uint64_t> &a)
{
    std::vector<uint64_t> b(a.size());
    for (int shift = 0; shift < 64; shift += 8) {
        size_t cnt[256] = {0};
        for (uint64_t x : a) cnt[(x >> shift) & 255]++;
        bool trivial = false;
        for (int d = 0; d < 256; d++) if (cnt[d] == a.size()) trivial = true;
        if (trivial) continue;
        size_t off = 0;
        for (int d = 0; d < 256; d++) { size_t c = cnt[d]; cnt[d] = off; off += c; }
        for (uint64_t x : a) b[cnt[(x >> shift) & 255]++] = x;
        a.swap(b);
    }
}

// Builds a symmetric CSR from unique undirected pairs (lo < hi) packed as
// lo << 32 | hi; rows come out sorted, so the result does not depend on the
// order of the pairs or on thread scheduling.
bool csr_from_pairs(int32_t n, const std::vector<uint64_t> &pairs, hgp_graph *g)
{
    int64_t P = (int64_t)pairs.size();
    std::vector<std::atomic<int64_t>> deg(n);
    parallel_for(n, [&](int64_t v) { deg[v].store(0, std::memory_order_relaxed); });
    parallel_for(P, [&](int64_t i) {
        deg[pairs[i] >> 32].fetch_add(1, std::memory_order_relaxed);
        deg[pairs[i] & 0xffffffffu].fetch_add(1, std::memory_order_relaxed);
    });
    if (!hgp::alloc_host_graph(g, n, 2 * P)) {
        hgp_graph_free(g);
        return false;
    }
    g->xadj[0] = 0;
    for (int32_t v = 0; v < n; v++) g->xadj[v + 1] = g->xadj[v] + deg[v].load(std::memory_order_relaxed);
    parallel_for(n, [&](int64_t v) { deg[v].store(g->xadj[v], std::memory_order_relaxed); });
    parallel_for(P, [&](int64_t i) {
        int32_t a = (int32_t)(pairs[i] >> 32), b = (int32_t)(pairs[i] & 0xffffffffu);
        g->adjncy[deg[a].fetch_add(1, std::memory_order_relaxed)] = b;
        g->adjncy[deg[b].fetch_add(1, std::memory_order_relaxed)] = a;
    });
    parallel_for(n, [&](int64_t v) {
        std::sort(g->adjncy + g->xadj[v], g->adjncy + g->xadj[v + 1]);
        g->vwgt[v] = 1;
    });
    parallel_for(2 * P, [&](int64_t j) { g->adjwgt[j] = 1; });
    return true;
}

}  // namespace

extern "C" {

void hgp_graph_free(hgp_graph *g)
{
    if (!g) return;
    free(g->xadj);
    free(g->adjncy);
    free(g->adjwgt);
    free(g->vwgt);
    memset(g, 0, sizeof(*g));
}

int hgp_gen_grid2d(int32_t rows, int32_t cols, hgp_graph *g)
{
    if (!g || rows < 1 || cols < 1 || (int64_t)rows * cols > INT32_MAX) return HGP_ERR_INVALID;
    int32_t n = rows * cols;
    int64_t m = 2 * ((int64_t)(rows - 1) * cols + (int64_t)rows * (cols - 1));
    if (!hgp::alloc_host_graph(g, n, m)) return HGP_ERR_OOM;
    int64_t e = 0;
    g->xadj[0] = 0;
    for (int32_t r = 0; r < rows; r++) {
        for (int32_t c = 0; c < cols; c++) {
            int32_t v = r * cols + c;
            if (r > 0) g->adjncy[e++] = v - cols;
            if (c > 0) g->adjncy[e++] = v - 1;
            if (c + 1 < cols) g->adjncy[e++] = v + 1;
            if (r + 1 < rows) g->adjncy[e++] = v + cols;
            g->xadj[v + 1] = e;
            g->vwgt[v] = 1;
        }
    }
    for (int64_t j = 0; j < m; j++) g->adjwgt[j] = 1;
    return HGP_OK;
}

int hgp_gen_rgg3d(int32_t n, double radius, uint64_t seed, int32_t vw_lo, int32_t vw_hi,
                  int32_t ew_lo, int32_t ew_hi, hgp_graph *g)
{
    if (!g || n < 1 || !(radius > 0) || radius >= 1) return HGP_ERR_INVALID;
    std::vector<double> P((size_t)n * 3);
    parallel_for(n, [&](int64_t i) {
        for (int d = 0; d < 3; d++) P[3 * i + d] = unit(rnd(seed, 1 + d, (uint64_t)i));
    });
    int32_t nc = (int32_t)std::floor(1.0 / radius);
    if (nc < 1) nc = 1;
    if (nc > 1024) nc = 1024;
    int64_t ncell = (int64_t)nc * nc * nc;
    auto cell_of = [&](int64_t i) {
        int32_t c[3];
        for (int d = 0; d < 3; d++) {
            int32_t x = (int32_t)(P[3 * i + d] * nc);
            c[d] = x >= nc ? nc - 1 : x;
        }
        return ((int64_t)c[0] * nc + c[1]) * nc + c[2];
    };
    std::vector<int64_t> cstart(ncell + 1, 0);
    std::vector<int64_t> cid(n);
    for (int32_t i = 0; i < n; i++) { cid[i] = cell_of(i); cstart[cid[i] + 1]++; }
    for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
    std::vector<int32_t> cpts(n);
    {
        std::vector<int64_t> pos(cstart.begin(), cstart.end() - 1);
        for (int32_t i = 0; i < n; i++) cpts[pos[cid[i]]++] = i;
    }
    double r2 = radius * radius;
    std::vector<int64_t> deg(n + 1, 0);
    auto scan_nbrs = [&](int32_t i, auto emit) {
        int64_t c = cid[i];
        int32_t cx = (int32_t)(c / ((int64_t)nc * nc)), cy = (int32_t)((c / nc) % nc), cz = (int32_t)(c % nc);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int32_t x = cx + dx, y = cy + dy, z = cz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) continue;
                    int64_t cc = ((int64_t)x * nc + y) * nc + z;
                    for (int64_t t = cstart[cc]; t < cstart[cc + 1]; t++) {
                        int32_t j = cpts[t];
                        if (j == i) continue;
                        double d0 = P[3 * i] - P[3 * j], d1 = P[3 * i + 1] - P[3 * j + 1],
                               d2 = P[3 * i + 2] - P[3 * j + 2];
                        if (d0 * d0 + d1 * d1 + d2 * d2 <= r2) emit(j);
                    }
                }
    };
    parallel_for(n, [&](int64_t i) {
        int64_t d = 0;
        scan_nbrs((int32_t)i, [&](int32_t) { d++; });
        deg[i] = d;
    });
    int64_t m = 0;
    for (int32_t i = 0; i < n; i++) m += deg[i];
    if (!hgp::alloc_host_graph(g, n, m)) return HGP_ERR_OOM;
    g->xadj[0] = 0;
    for (int32_t i = 0; i < n; i++) g->xadj[i + 1] = g->xadj[i] + deg[i];
    bool wv = vw_hi > 0 && vw_hi >= vw_lo && vw_lo > 0;
    bool we = ew_hi > 0 && ew_hi >= ew_lo && ew_lo > 0;
    parallel_for(n, [&](int64_t i) {
        int64_t e = g->xadj[i];
        scan_nbrs((int32_t)i, [&](int32_t j) { g->adjncy[e++] = j; });
        std::sort(g->adjncy + g->xadj[i], g->adjncy + g->xadj[i + 1]);
        for (int64_t t = g->xadj[i]; t < g->xadj[i + 1]; t++) {
            int32_t j = g->adjncy[t];
            uint64_t lo = (uint64_t)std::min<int64_t>(i, j), hi = (uint64_t)std::max<int64_t>(i, j);
            g->adjwgt[t] = we ? ew_lo + (int32_t)(rnd(seed, 7, (lo << 32) | hi) % (uint64_t)(ew_hi - ew_lo + 1)) : 1;
        }
        g->vwgt[i] = wv ? vw_lo + (int32_t)(rnd(seed, 5, (uint64_t)i) % (uint64_t)(vw_hi - vw_lo + 1)) : 1;
    });
    return HGP_OK;
}

int hgp_gen_rmat(int32_t scale, int64_t nedges, double a, double b, double c, uint64_t seed,
                 hgp_graph *g)
{
    if (!g || scale < 1 || scale > 30 || nedges < 1) return HGP_ERR_INVALID;
    int64_t n = (int64_t)1 << scale;
    if ((double)nedges > 0.5 * (double)n * (double)(n - 1)) return HGP_ERR_INVALID;
    // quadrant thresholds as 16-bit fractions
    uint32_t ta = (uint32_t)std::lround(a * 65536.0), tb = (uint32_t)std::lround((a + b) * 65536.0),
             tc = (uint32_t)std::lround((a + b + c) * 65536.0);
    auto draw = [&](uint64_t i) -> uint64_t {
        uint64_t u = 0, v = 0, r = 0;
        for (int lvl = 0; lvl < scale; lvl++) {
            if ((lvl & 3) == 0) r = rnd(seed, 11, i * 8 + (uint64_t)(lvl >> 2));
            uint32_t x = (uint32_t)(r & 0xffff);
            r >>= 16;
            uint64_t bu = 0, bv = 0;
            if (x < ta) {} else if (x < tb) bv = 1; else if (x < tc) bu = 1; else { bu = 1; bv = 1; }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u == v) return ~0ull;
        uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
        return (lo << 32) | hi;
    };
    double t0 = now_s();
    // The edge set is the first nedges distinct pairs in draw order.  Draws are
    // made in batches; every distinct pair keeps the index of its first draw.
    std::vector<uint64_t> keys;      // distinct pairs
    std::vector<uint64_t> firsts;    // first draw index, parallel to keys
    uint64_t drawn = 0;
    const int nt = nthreads();
    while ((int64_t)keys.size() < nedges) {
        uint64_t want = (uint64_t)((double)(nedges - (int64_t)keys.size()) * 1.3) + 1024;
        std::vector<uint64_t> batch(want);
        parallel_for((int64_t)want, [&](int64_t i) { batch[i] = draw(drawn + (uint64_t)i); });
        GEN_T("draw");
        // MSD buckets by the top key bits over (kept pairs, new draws), in parallel
        const int bb = 10, shift = 32 + scale - bb;   // keys are lo << 32 | hi, lo < 2^scale
        const size_t nb = (size_t)1 << bb;
        const size_t K = keys.size(), N = K + want;
        auto key_at = [&](size_t i) { return i < K ? keys[i] : batch[i - K]; };
        auto idx_at = [&](size_t i) { return i < K ? firsts[i] : drawn + (i - K); };
        std::vector<std::vector<size_t>> cnt(nt, std::vector<size_t>(nb + 1, 0));
        auto chunk = [&](int t, size_t &lo, size_t &hi) { lo = N * t / nt; hi = N * (t + 1) / nt; };
        {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++)
                th.emplace_back([&, t] {
                    size_t lo, hi;
                    chunk(t, lo, hi);
                    for (size_t i = lo; i < hi; i++) {
                        uint64_t kk = key_at(i);
                        if (kk != ~0ull) cnt[t][kk >> shift]++;
                    }
                });
            for (auto &x : th) x.join();
        }
        std::vector<size_t> bstart(nb + 1, 0);
        size_t off = 0;
        for (size_t b = 0; b < nb; b++) {
            bstart[b] = off;
            for (int t = 0; t < nt; t++) { size_t c = cnt[t][b]; cnt[t][b] = off; off += c; }
        }
        bstart[nb] = off;
        std::vector<std::pair<uint64_t, uint64_t>> P(off);
        {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++)
                th.emplace_back([&, t] {
                    size_t lo, hi;
                    chunk(t, lo, hi);
                    for (size_t i = lo; i < hi; i++) {
                        uint64_t kk = key_at(i);
                        if (kk != ~0ull) P[cnt[t][kk >> shift]++] = {kk, idx_at(i)};
                    }
                });
            for (auto &x : th) x.join();
        }
        batch.clear();
        batch.shrink_to_fit();
        GEN_T("pack");
        // sort each bucket by (pair, index) and keep the first entry of each pair
        std::vector<size_t> ucnt(nb, 0);
        {
            std::atomic<size_t> next(0);
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++)
                th.emplace_back([&] {
                    for (size_t b; (b = next.fetch_add(1)) < nb;) {
                        auto lo = P.begin() + bstart[b], hi = P.begin() + bstart[b + 1];
                        std::sort(lo, hi);
                        size_t u = 0;
                        for (auto it = lo; it != hi; ++it)
                            if (it == lo || it->first != (it - 1)->first) u++;
                        ucnt[b] = u;
                    }
                });
            for (auto &x : th) x.join();
        }
        GEN_T("sort");
        std::vector<size_t> uoff(nb + 1, 0);
        for (size_t b = 0; b < nb; b++) uoff[b + 1] = uoff[b] + ucnt[b];
        keys.assign(uoff[nb], 0);
        firsts.assign(uoff[nb], 0);
        parallel_for((int64_t)nb, [&](int64_t b) {
            size_t o = uoff[b];
            auto lo = P.begin() + bstart[b], hi = P.begin() + bstart[b + 1];
            for (auto it = lo; it != hi; ++it)
                if (it == lo || it->first != (it - 1)->first) { keys[o] = it->first; firsts[o] = it->second; o++; }
        });
        drawn += want;
    }
    GEN_T("dedup");
    if ((int64_t)keys.size() > nedges) {
        // keep the nedges pairs with the smallest first-draw index: mark first
        // draws in a bitmap over draw indices and find the cut-off index
        std::vector<uint64_t> bits((drawn + 63) / 64, 0);
        for (uint64_t f : firsts) bits[f >> 6] |= 1ull << (f & 63);
        int64_t need = nedges, acc = 0;
        uint64_t cutoff = drawn;
        for (size_t w = 0; w < bits.size(); w++) {
            int c = __builtin_popcountll(bits[w]);
            if (acc + c >= need) {
                uint64_t x = bits[w];
                for (int b = 0; b < 64; b++) {
                    if (x >> b & 1) {
                        if (++acc == need) { cutoff = w * 64 + b + 1; break; }
                    }
                }
                break;
            }
            acc += c;
        }
        std::vector<uint64_t> kept;
        kept.reserve(nedges);
        for (size_t i = 0; i < keys.size(); i++) if (firsts[i] < cutoff) kept.push_back(keys[i]);
        keys.swap(kept);
    }
    GEN_T("trim");
    // seeded relabelling (Fisher-Yates)
    std::vector<int32_t> perm(n);
    for (int64_t v = 0; v < n; v++) perm[v] = (int32_t)v;
    for (int64_t v = n - 1; v > 0; v--) {
        int64_t j = (int64_t)(rnd(seed, 13, (uint64_t)v) % (uint64_t)(v + 1));
        std::swap(perm[v], perm[j]);
    }
    parallel_for((int64_t)keys.size(), [&](int64_t i) {
        uint64_t p = keys[i];
        uint64_t u = (uint64_t)perm[p >> 32], v = (uint64_t)perm[p & 0xffffffffu];
        keys[i] = u < v ? (u << 32) | v : (v << 32) | u;
    });
    GEN_T("relabel");
    if (!csr_from_pairs((int32_t)n, keys, g)) return HGP_ERR_OOM;
    GEN_T("csr");
    return HGP_OK;
}

int hgp_gen_banded(int32_t n, int32_t bw, double frac, uint64_t seed, hgp_graph *g)
{
    if (!g || n < 2 || bw < 1) return HGP_ERR_INVALID;
    std::vector<uint64_t> pairs;
    for (int32_t i = 0; i < n; i++)
        for (int32_t j = i + 1; j <= i + bw && j < n; j++) pairs.push_back(((uint64_t)i << 32) | (uint64_t)j);
    int64_t band = (int64_t)pairs.size();
    int64_t extra = (int64_t)std::llround(frac * (double)band);
    int64_t avail = (int64_t)n * (n - 1) / 2 - band;
    if (extra > avail) extra = avail;
    std::vector<uint64_t> off;
    uint64_t i = 0;
    while ((int64_t)off.size() < extra) {
        uint64_t r = rnd(seed, 17, i++);
        int32_t x = (int32_t)((r >> 32) % (uint64_t)n), y = (int32_t)((r & 0xffffffffu) % (uint64_t)n);
        if (x == y) continue;
        int32_t lo = std::min(x, y), hi = std::max(x, y);
        if (hi - lo <= bw) continue;
        off.push_back(((uint64_t)lo << 32) | (uint64_t)hi);
        if ((int64_t)off.size() == extra) {
            radix_sort_u64(off);
            off.erase(std::unique(off.begin(), off.end()), off.end());
        }
    }
    pairs.insert(pairs.end(), off.begin(), off.end());
    radix_sort_u64(pairs);
    if (!csr_from_pairs(n, pairs, g)) return HGP_ERR_OOM;
    return HGP_OK;
}

}  // extern "C"
