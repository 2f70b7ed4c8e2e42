// hgp_cli.cpp -- the `hunyuangraph` command: read a graph file, partition it
// on the GPU, report, optionally write the partition.
//
// Mirrors the reference's program (hunyuangraph.cu:5-124): the same
// positional arguments (`<graph> <nparts> [pool]`), best of two runs with a
// host recount of edge cut and imbalance, the print_time_all block of each
// run (hunyuangraph_timer.h:435-447) and the best_* times.  The reference's
// study builds (-DFIGURE9_SUM, -DFIGURE9_TIME, -DFIGURE10_CGRAPH,
// -DFIGURE10_EXHAUSTIVE / _SAMPLING, -DFIGURE14_EDGECUT) are run-time modes
// here (--study).  Host code over the library's C ABI only.
#include "hunyuangraph.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const char *kUsage =
    "usage: hunyuangraph <graph> <nparts> [pool] [options]\n"
    "  <graph>          METIS .graph file, or a Matrix Market .mtx file (converted to\n"
    "                   the graph of pattern(A + A^T) without the diagonal)\n"
    "  <nparts>         number of parts\n"
    "  [pool]           0|1, accepted for the reference's third argument; the library\n"
    "                   always partitions inside its own device arena\n"
    "options:\n"
    "  --runs R         partitions to run, best cut kept (default 2)\n"
    "  --seed S         seed of the first run; run i uses S+i (default 1)\n"
    "  --ub X           balance bound (default 1.03)\n"
    "  --write          write the best partition to <graph>.part.<nparts>\n"
    "  --study MODE     stop after a phase and print what the reference's study builds print:\n"
    "     coarsen-sum         per level: vertices, entries, entries/vertex, edge-weight sum\n"
    "     coarsen-time        per level: coarsening time\n"
    "     coarsest-graph[=F]  write the coarsest graph as a fmt 011 file F (default graph.txt)\n"
    "     init-exhaustive     initial partition of the input graph, a trial from every vertex\n"
    "     init-sampling       initial partition of the input graph, sampled trials\n"
    "     edgecut             refinement patience 12, and print best_edgecut\n";

struct Args {
    std::string path;
    int32_t k = 0;
    int pool = 1;
    int runs = 2;
    uint64_t seed = 1;
    double ub = 1.03;
    bool write = false;
    std::string study;
    std::string study_file = "graph.txt";
};

[[noreturn]] void usage_error(const std::string &msg)
{
    fprintf(stderr, "hunyuangraph: %s\n%s", msg.c_str(), kUsage);
    exit(2);
}

bool parse_long(const char *s, long long &out)
{
    if (!s || !*s) return false;
    char *end = nullptr;
    out = strtoll(s, &end, 10);
    return end && *end == 0;
}

Args parse_args(int argc, char **argv)
{
    Args a;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        auto value = [&](const char *name) -> const char * {
            if (i + 1 >= argc) usage_error(std::string(name) + " needs a value");
            return argv[++i];
        };
        long long v;
        if (s == "-h" || s == "--help") {
            fputs(kUsage, stdout);
            exit(0);
        } else if (s == "--runs") {
            if (!parse_long(value("--runs"), v) || v < 1 || v > 1000) usage_error("--runs must be 1..1000");
            a.runs = (int)v;
        } else if (s == "--seed") {
            if (!parse_long(value("--seed"), v) || v < 0) usage_error("--seed must be a non-negative integer");
            a.seed = (uint64_t)v;
        } else if (s == "--ub") {
            const char *x = value("--ub");
            char *end = nullptr;
            a.ub = strtod(x, &end);
            if (!end || *end || !(a.ub >= 1.0) || a.ub > 100.0) usage_error("--ub must be a number >= 1");
        } else if (s == "--write") {
            a.write = true;
        } else if (s == "--study") {
            std::string m = value("--study");
            size_t eq = m.find('=');
            if (eq != std::string::npos) {
                a.study_file = m.substr(eq + 1);
                m = m.substr(0, eq);
                if (m != "coarsest-graph" || a.study_file.empty()) usage_error("bad --study value");
            }
            if (m != "coarsen-sum" && m != "coarsen-time" && m != "coarsest-graph" && m != "init-exhaustive" &&
                m != "init-sampling" && m != "edgecut")
                usage_error("unknown --study mode '" + m + "'");
            a.study = m;
        } else if (s.size() > 1 && s[0] == '-' && !(s[1] >= '0' && s[1] <= '9')) {
            usage_error("unknown option " + s);
        } else {
            pos.push_back(s);
        }
    }
    if (pos.size() < 2 || pos.size() > 3) usage_error("expected <graph> <nparts> [pool]");
    a.path = pos[0];
    long long v;
    if (!parse_long(pos[1].c_str(), v) || v < 1 || v > INT32_MAX) usage_error("<nparts> must be a positive integer");
    a.k = (int32_t)v;
    if (pos.size() == 3) {
        if (!parse_long(pos[2].c_str(), v) || (v != 0 && v != 1)) usage_error("[pool] must be 0 or 1");
        a.pool = (int)v;
    }
    return a;
}

bool ends_with(const std::string &s, const char *suf)
{
    size_t n = strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; i++)
        if (tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

int fail(const char *what)
{
    fprintf(stderr, "hunyuangraph: %s: %s\n", what, hgp_last_error());
    return 1;
}

// Host recount (hunyuangraph_computecut_cpu / hunyuangraph_compute_imbalance_cpu,
// hunyuangraph_graph.h:144-171): each cut edge is seen from both ends.
int64_t host_cut(const hgp_graph &g, const std::vector<int32_t> &part)
{
    int64_t cut = 0;
    for (int32_t v = 0; v < g.n; v++)
        for (int64_t j = g.xadj[v]; j < g.xadj[v + 1]; j++)
            if (part[v] != part[g.adjncy[j]]) cut += g.adjwgt[j];
    return cut / 2;
}

// max part weight over the average part weight
double host_imbalance(const hgp_graph &g, const std::vector<int32_t> &part, int32_t k)
{
    std::vector<int64_t> pw((size_t)k, 0);
    int64_t tv = 0;
    for (int32_t v = 0; v < g.n; v++) {
        pw[(size_t)part[v]] += g.vwgt[v];
        tv += g.vwgt[v];
    }
    int64_t mx = *std::max_element(pw.begin(), pw.end());
    return tv > 0 ? (double)mx * k / (double)tv : 0.0;
}

void print_run(double all, const hgp_stats &st, int64_t cut, double imb)
{
    printf("---------------------------------------------------------\n");
    printf("Hunyuangraph-Partition-end\n");
    printf("Hunyuangraph_Partition_time= %10.2lf ms\n", all);
    printf("------Coarsen_time=          %10.2lf ms\n", st.ms_coarsen);
    printf("------Init_time=             %10.2lf ms\n", st.ms_init);
    printf("------Uncoarsen_time=        %10.2lf ms\n", st.ms_uncoarsen);
    printf("------else_time=             %10.2lf ms\n", all - (st.ms_coarsen + st.ms_init + st.ms_uncoarsen));
    printf("edge-cut=                    %10lld\n", (long long)cut);
    printf("imbalance=                   %10.3f\n", imb);
    printf("---------------------------------------------------------\n");
}

int run_coarsen_study(const Args &a, const hgp_graph &g, const hgp_options &o)
{
    hgp_options oo = o;
    oo.level_stats = 1;
    hgp_graph cg;
    hgp_stats st;
    if (hgp_coarsen(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, a.k, &oo, &cg, &st) != HGP_OK) return fail("coarsening");
    int rc = 0;
    if (a.study == "coarsen-sum") {
        // the reference prints each coarse level as it is built
        for (int32_t l = 1; l <= st.levels && l < HGP_MAX_LEVELS; l++)
            printf("level %2d: nvtxs %10lld nedges %10lld nedges/nvtxs=%7.2lf adjwgtsum %12lld\n", l,
                   (long long)st.level_n[l], (long long)st.level_m[l],
                   (double)st.level_m[l] / (double)std::max<int64_t>(st.level_n[l], 1), (long long)st.level_adjwgt[l]);
    } else if (a.study == "coarsen-time") {
        for (int32_t l = 0; l < st.levels && l < HGP_MAX_LEVELS; l++)
            printf("level=%3d time=%10.3lf ms\n", l + 1, st.level_ms[l]);
    } else {
        if (hgp_write_graph(a.study_file.c_str(), &cg, 11) != HGP_OK) rc = fail("writing the coarsest graph");
        else printf("coarsest graph: %d vertices, %lld edges -> %s\n", cg.n, (long long)(cg.m / 2), a.study_file.c_str());
    }
    printf("Coarsen_time=              %10.3lf ms\n", st.ms_coarsen);
    hgp_graph_free(&cg);
    return rc;
}

int run_init_study(const Args &a, const hgp_graph &g, const hgp_options &o)
{
    hgp_options oo = o;
    if (a.study == "init-exhaustive") {
        // a trial from every vertex: n trials of n-vertex state
        if (g.n > 65536) {
            fprintf(stderr, "hunyuangraph: init-exhaustive is meant for coarsest-size graphs (<= 65536 vertices)\n");
            return 1;
        }
        oo.trials = g.n;
    }
    std::vector<int32_t> where((size_t)g.n);
    auto t0 = std::chrono::steady_clock::now();
    if (hgp_initpart(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, a.k, &oo, where.data()) != HGP_OK)
        return fail("initial partition");
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    printf("trials=%d edgecut=%10lld imbalance=%10.3f\n", std::min(oo.trials, g.n), (long long)host_cut(g, where),
           host_imbalance(g, where, a.k));
    printf("Init_time=                 %10.3lf ms\n", ms);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Args a = parse_args(argc, argv);
    hgp_graph g;
    int rc = ends_with(a.path, ".mtx") ? hgp_read_mtx(a.path.c_str(), &g) : hgp_read_graph(a.path.c_str(), &g);
    if (rc != HGP_OK) return fail(a.path.c_str());
    printf("graph:%s %d %lld %d %d\n", a.path.c_str(), g.n, (long long)g.m, a.k, a.pool);
    fflush(stdout);

    hgp_options o;
    hgp_default_options(&o);
    int32_t num = (int32_t)(a.ub * 1000.0 + 0.5), den = 1000;   // the bound as a reduced fraction
    for (int32_t x = num, y = den; y;) {
        int32_t t = x % y;
        x = y;
        y = t;
        if (!y) { num /= x; den /= x; }
    }
    o.ub_num = num;
    o.ub_den = den;
    if (a.study == "edgecut") o.patience = 12;   // the reference's -DFIGURE14_EDGECUT loop bound

    int status = 0;
    if (a.study.rfind("coarse", 0) == 0) {
        status = run_coarsen_study(a, g, o);
    } else if (a.study.rfind("init-", 0) == 0) {
        status = run_init_study(a, g, o);
    } else {
        std::vector<int32_t> part((size_t)g.n), best((size_t)g.n);
        int64_t best_cut = INT64_MAX;
        double best_all = 1e300, best_c = 1e300, best_i = 1e300, best_u = 1e300;
        for (int r = 0; r < a.runs; r++) {
            o.seed = a.seed + (uint64_t)r;
            hgp_stats st;
            auto t0 = std::chrono::steady_clock::now();
            if (hgp_partition(g.n, g.xadj, g.adjncy, g.vwgt, g.adjwgt, a.k, &o, part.data(), &st) != HGP_OK) {
                status = fail("partition");
                break;
            }
            double all = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            int64_t cut = host_cut(g, part);
            double imb = host_imbalance(g, part, a.k);
            print_run(all, st, cut, imb);
            fflush(stdout);
            if (cut != st.cut) {
                fprintf(stderr, "hunyuangraph: host recount %lld differs from the library's cut %lld\n",
                        (long long)cut, (long long)st.cut);
                status = 1;
                break;
            }
            best_all = std::min(best_all, all);
            best_c = std::min(best_c, st.ms_coarsen);
            best_i = std::min(best_i, st.ms_init);
            best_u = std::min(best_u, st.ms_uncoarsen);
            if (cut < best_cut) {
                best_cut = cut;
                best = part;
            }
        }
        if (status == 0) {
            printf("best_alltime=         %10.3lf\n", best_all);
            printf("best_coarsentime=     %10.3lf\n", best_c);
            printf("best_inittime=        %10.3lf\n", best_i);
            printf("best_uncoarsentime=   %10.3lf\n", best_u);
            if (a.study == "edgecut") printf("best_edgecut=         %10lld\n", (long long)best_cut);
            if (a.write) {
                if (hgp_write_partition(a.path.c_str(), best.data(), g.n, a.k) != HGP_OK) status = fail("writing the partition");
                else printf("partition written to %s.part.%d\n", a.path.c_str(), a.k);
            }
        }
    }
    hgp_graph_free(&g);
    return status;
}
