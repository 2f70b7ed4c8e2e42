// Runs the library's host-only code (graph generators, graph-file reader and
// writers, the host CSR check) and the oracle's C restatements under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/native/Makefile, target `sanitize`).  Arguments: a scratch directory,
// then graph files (.graph / .mtx) to read.  Prints "ok" and exits 0 when
// every check passes; a sanitizer report aborts the run.
#include "hunyuangraph.h"
#include "hgp_hostcheck.hpp"
#include "hgp_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace hgp {
// the library's error slot lives in its device runtime; a plain one here
void set_error(int, const std::string &) {}
}  // namespace hgp

#define CHECK(c)                                                                    \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);    \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

static bool throws(int32_t n, const int64_t *x, const int32_t *a, const int32_t *vw, const int32_t *aw, bool sym)
{
    try {
        hgp::validate_host(n, x, a, vw, aw, sym);
    } catch (const hgp::Error &) {
        return true;
    }
    return false;
}

static void exercise(const hgp_graph &g, const std::string &dir, int k)
{
    CHECK(!throws(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, true));
    // write and read back with both weight kinds
    std::string path = dir + "/roundtrip.graph";
    CHECK(hgp_write_graph(path.c_str(), &g, 11) == HGP_OK);
    hgp_graph h{};
    CHECK(hgp_read_graph(path.c_str(), &h) == HGP_OK);
    CHECK(h.n == g.n && h.m == g.m);
    CHECK(std::memcmp(h.xadj, g.xadj, sizeof(int64_t) * (size_t)(g.n + 1)) == 0);
    CHECK(std::memcmp(h.adjncy, g.adjncy, sizeof(int32_t) * (size_t)g.m) == 0);
    CHECK(std::memcmp(h.adjwgt, g.adjwgt, sizeof(int32_t) * (size_t)g.m) == 0);
    CHECK(std::memcmp(h.vwgt, g.vwgt, sizeof(int32_t) * (size_t)g.n) == 0);
    hgp_graph_free(&h);
    // the oracle's pipelines on it
    std::vector<int32_t> part((size_t)g.n), part2((size_t)g.n);
    oracle_opts o;
    oracle_default_opts(&o);
    o.trials = 64;
    oracle_stats st;
    CHECK(oracle_dga_partition(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, k, &o, part.data(), &st) == 0);
    CHECK(st.cut == oracle_edge_cut(g.n, g.xadj, g.adjncy, g.adjwgt, part.data()));
    int64_t cut = oracle_cpu_rb_partition(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, k, 1, part2.data());
    CHECK(cut == oracle_edge_cut(g.n, g.xadj, g.adjncy, g.adjwgt, part2.data()));
    std::string pf = dir + "/roundtrip.graph";
    CHECK(hgp_write_partition(pf.c_str(), part.data(), g.n, k) == HGP_OK);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    std::string dir = argv[1];
    hgp_graph g{};
    CHECK(hgp_gen_grid2d(20, 30, &g) == HGP_OK);
    exercise(g, dir, 8);
    hgp_graph_free(&g);
    double r = std::cbrt(12.0 * 3.0 / (4.0 * M_PI * 3000.0));
    CHECK(hgp_gen_rgg3d(3000, r, 5, 1, 4, 1, 16, &g) == HGP_OK);
    exercise(g, dir, 16);
    hgp_graph_free(&g);
    CHECK(hgp_gen_rmat(12, 1 << 14, 0.57, 0.19, 0.19, 3, &g) == HGP_OK);
    exercise(g, dir, 32);
    // the host check's faults
    std::vector<int32_t> bad(g.adjncy, g.adjncy + g.m);
    bad[5] = g.n;
    CHECK(throws(g.n, g.xadj, bad.data(), g.vwgt, g.adjwgt, false));
    bad[5] = -1;
    CHECK(throws(g.n, g.xadj, bad.data(), g.vwgt, g.adjwgt, false));
    std::vector<int32_t> w(g.adjwgt, g.adjwgt + g.m);
    w[7] += 1;   // asymmetric weight
    CHECK(!throws(g.n, g.xadj, g.adjncy, g.vwgt, w.data(), false));
    CHECK(throws(g.n, g.xadj, g.adjncy, g.vwgt, w.data(), true));
    std::vector<int64_t> x(g.xadj, g.xadj + g.n + 1);
    x[3] = x[4] + 1;
    CHECK(throws(g.n, x.data(), g.adjncy, g.vwgt, g.adjwgt, false));
    hgp_graph_free(&g);
    CHECK(hgp_gen_banded(2000, 8, 0.1, 4, &g) == HGP_OK);
    exercise(g, dir, 4);
    hgp_graph_free(&g);
    // graph files given on the command line
    for (int i = 2; i < argc; i++) {
        std::string p = argv[i];
        bool mtx = p.size() > 4 && p.compare(p.size() - 4, 4, ".mtx") == 0;
        hgp_graph f{};
        CHECK((mtx ? hgp_read_mtx(p.c_str(), &f) : hgp_read_graph(p.c_str(), &f)) == HGP_OK);
        if (f.n > 1) exercise(f, dir, 2);
        hgp_graph_free(&f);
    }
    // files that must be refused
    std::string junk = dir + "/junk.graph";
    FILE *fp = std::fopen(junk.c_str(), "w");
    std::fputs("3 2\n2 9\n1\n\n", fp);
    std::fclose(fp);
    hgp_graph f{};
    CHECK(hgp_read_graph(junk.c_str(), &f) == HGP_ERR_INVALID);
    CHECK(hgp_read_graph((dir + "/missing.graph").c_str(), &f) != HGP_OK);
    std::printf("ok\n");
    return 0;
}
