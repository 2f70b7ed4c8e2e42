// hgp_io.cpp -- graph and partition files (host code).
//
// METIS .graph files (hunyuangraph_readgraph, hunyuangraph_io.h:25-186), the
// partition file the reference CLI writes (hunyuangraph_writetofile,
// hunyuangraph_io.h:189-205), and Matrix Market matrices turned into graphs
// the way the reference's converter does (matrix_to_graph/matrix_to_graph.c:
// pattern of A + A^T without the diagonal, neighbours ascending, weights
// dropped).  Errors are returned as status codes with a message in
// hgp_last_error(); nothing exits the process.
#include "hunyuangraph.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace hgp {
void set_error(int code, const std::string &msg);
bool alloc_host_graph(hgp_graph *g, int32_t n, int64_t m);   // hgp_gen.cpp
}

namespace {

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw IoError{code, msg}; }

// A whole text file in memory (graph files are read once, front to back).
std::string slurp(const char *path)
{
    if (!path) fail(HGP_ERR_INVALID, "path is NULL");
    FILE *f = fopen(path, "rb");
    if (!f) fail(HGP_ERR_IO, std::string("cannot open ") + path + ": " + strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, r);
    bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) fail(HGP_ERR_IO, std::string("read error on ") + path);
    return s;
}

// Line cursor over the file text.
struct Lines {
    const std::string &s;
    size_t pos = 0;
    int64_t lineno = 0;
    explicit Lines(const std::string &t) : s(t) {}
    // next line (without the newline); false at the end of the text
    bool next(const char *&b, const char *&e)
    {
        if (pos >= s.size()) return false;
        size_t nl = s.find('\n', pos);
        if (nl == std::string::npos) nl = s.size();
        b = s.data() + pos;
        e = s.data() + nl;
        if (e > b && e[-1] == '\r') e--;
        pos = nl + 1;
        lineno++;
        return true;
    }
    // next line that is not a comment ('%' in the first column)
    bool next_data(const char *&b, const char *&e)
    {
        while (next(b, e))
            if (!(e > b && b[0] == '%')) return true;
        return false;
    }
};

// Reads a decimal integer at p the way strtol does: leading white space
// skipped, optional sign, digits; whatever follows the digits is left for the
// next call.  Returns false (p unchanged) when no number starts there.
bool parse_int(const char *&p, const char *e, int64_t &out)
{
    const char *q = p;
    while (q < e && (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\v' || *q == '\f')) q++;
    bool neg = false;
    if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
    if (q >= e || *q < '0' || *q > '9') return false;
    int64_t v = 0;
    while (q < e && *q >= '0' && *q <= '9') {
        v = v * 10 + (*q - '0');
        if (v > ((int64_t)1 << 40)) fail(HGP_ERR_INVALID, "number out of range");
        q++;
    }
    out = neg ? -v : v;
    p = q;
    return true;
}

std::string at(const Lines &L) { return " (line " + std::to_string(L.lineno) + ")"; }

void read_metis(const char *path, hgp_graph *g)
{
    std::string text = slurp(path);
    Lines L(text);
    const char *b, *e;
    if (!L.next_data(b, e)) fail(HGP_ERR_INVALID, "premature end of file: no header line");
    int64_t nv = 0, ne = 0, fmt = 0;
    const char *p = b;
    if (!parse_int(p, e, nv) || !parse_int(p, e, ne))
        fail(HGP_ERR_INVALID, "the header does not give the numbers of vertices and edges");
    if (!parse_int(p, e, fmt)) fmt = 0;
    int64_t ncon = 1;
    if (parse_int(p, e, ncon) && ncon != 1) fail(HGP_ERR_UNSUPPORTED, "only one vertex weight per vertex is supported");
    if (nv <= 0 || ne <= 0) fail(HGP_ERR_INVALID, "the numbers of vertices and edges must be positive");
    if (nv > INT32_MAX) fail(HGP_ERR_UNSUPPORTED, "more than 2^31-1 vertices");
    if (fmt > 111) fail(HGP_ERR_INVALID, "cannot read this file format [fmt=" + std::to_string(fmt) + "]");
    if (fmt < 0) fail(HGP_ERR_INVALID, "the file format must not be negative [fmt=" + std::to_string(fmt) + "]");
    // the three digits of fmt (zero-padded) flag vertex sizes, vertex weights
    // and edge weights; a digit other than 1 leaves its field out
    const bool has_vsize = fmt / 100 == 1, has_vw = (fmt / 10) % 10 == 1, has_ew = fmt % 10 == 1;
    const int64_t m = 2 * ne;
    if (!hgp::alloc_host_graph(g, (int32_t)nv, m)) fail(HGP_ERR_OOM, "host out of memory");
    int64_t k = 0;
    g->xadj[0] = 0;
    for (int32_t v = 0; v < (int32_t)nv; v++) {
        if (!L.next_data(b, e))
            fail(HGP_ERR_INVALID, "premature end of file while reading vertex " + std::to_string(v + 1));
        p = b;
        int64_t x;
        if (has_vsize && !parse_int(p, e, x))   // vertex size: read and ignored
            fail(HGP_ERR_INVALID, "vertex " + std::to_string(v + 1) + " has no size" + at(L));
        g->vwgt[v] = 1;
        if (has_vw) {
            if (!parse_int(p, e, x))
                fail(HGP_ERR_INVALID, "vertex " + std::to_string(v + 1) + " has no weight" + at(L));
            if (x < 0 || x > INT32_MAX)
                fail(HGP_ERR_INVALID, "the weight of vertex " + std::to_string(v + 1) + " must be >= 0" + at(L));
            g->vwgt[v] = (int32_t)x;
        }
        while (parse_int(p, e, x)) {
            if (x < 1 || x > nv)
                fail(HGP_ERR_INVALID, "edge " + std::to_string(x) + " of vertex " + std::to_string(v + 1) +
                                          " is out of bounds" + at(L));
            int64_t w = 1;
            if (has_ew) {
                if (!parse_int(p, e, w))
                    fail(HGP_ERR_INVALID, "premature end of line for vertex " + std::to_string(v + 1) + at(L));
                if (w <= 0 || w > INT32_MAX)
                    fail(HGP_ERR_INVALID, "the weight of edge (" + std::to_string(v + 1) + ", " + std::to_string(x) +
                                              ") must be positive" + at(L));
            }
            if (k == m)
                fail(HGP_ERR_INVALID, "there are more edges in the file than the " + std::to_string(ne) + " specified");
            g->adjncy[k] = (int32_t)(x - 1);
            g->adjwgt[k] = (int32_t)w;
            k++;
        }
        g->xadj[v + 1] = k;
    }
    if (k != m)
        fail(HGP_ERR_INVALID, "the header specifies " + std::to_string(ne) + " edges but the rows hold " +
                                  std::to_string(k) + " entries (" + std::to_string(k / 2) +
                                  " edges; each edge is counted once in the header)");
}

// Matrix Market coordinate file -> symmetric graph.
void read_mtx(const char *path, hgp_graph *g)
{
    std::string text = slurp(path);
    Lines L(text);
    const char *b, *e;
    if (!L.next(b, e)) fail(HGP_ERR_INVALID, "empty file");
    std::string banner(b, e);
    std::string low = banner;
    for (char &c : low) c = (char)tolower((unsigned char)c);
    if (low.rfind("%%matrixmarket", 0) != 0) fail(HGP_ERR_INVALID, "not a Matrix Market file (no banner)");
    if (low.find("matrix") == std::string::npos || low.find("coordinate") == std::string::npos)
        fail(HGP_ERR_UNSUPPORTED, "only coordinate matrices are supported");
    // values (real, integer, complex or none) are dropped: the graph keeps the
    // pattern only, and a symmetric or hermitian file that lists one triangle
    // gets the other back from A + A^T
    if (low.find("pattern") == std::string::npos && low.find("real") == std::string::npos &&
        low.find("integer") == std::string::npos && low.find("complex") == std::string::npos)
        fail(HGP_ERR_UNSUPPORTED, "unknown field type in the banner");
    if (!L.next_data(b, e)) fail(HGP_ERR_INVALID, "no size line");
    const char *p = b;
    int64_t rows, cols, nnz;
    if (!parse_int(p, e, rows) || !parse_int(p, e, cols) || !parse_int(p, e, nnz))
        fail(HGP_ERR_INVALID, "the size line must give rows, columns and entries");
    if (rows != cols) fail(HGP_ERR_INVALID, "the matrix must be square to define a graph");
    if (rows <= 0 || rows > INT32_MAX || nnz < 0) fail(HGP_ERR_INVALID, "bad matrix size");
    const int32_t n = (int32_t)rows;
    // both orientations of every off-diagonal entry (A + A^T), as (row, col) keys
    std::vector<uint64_t> key;
    key.reserve((size_t)(2 * nnz));
    for (int64_t t = 0; t < nnz; t++) {
        if (!L.next_data(b, e)) fail(HGP_ERR_INVALID, "premature end of file at entry " + std::to_string(t + 1));
        p = b;
        int64_t i, j;
        if (!parse_int(p, e, i) || !parse_int(p, e, j))
            fail(HGP_ERR_INVALID, "entry " + std::to_string(t + 1) + " needs a row and a column" + at(L));
        if (i < 1 || i > n || j < 1 || j > n)
            fail(HGP_ERR_INVALID, "entry " + std::to_string(t + 1) + " is out of bounds" + at(L));
        if (i == j) continue;
        key.push_back((uint64_t)(i - 1) << 32 | (uint64_t)(j - 1));
        key.push_back((uint64_t)(j - 1) << 32 | (uint64_t)(i - 1));
    }
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    const int64_t m = (int64_t)key.size();
    if (!hgp::alloc_host_graph(g, n, m)) fail(HGP_ERR_OOM, "host out of memory");
    std::fill(g->xadj, g->xadj + n + 1, 0);
    for (int64_t t = 0; t < m; t++) {
        g->xadj[(key[t] >> 32) + 1]++;
        g->adjncy[t] = (int32_t)(key[t] & 0xFFFFFFFFu);
        g->adjwgt[t] = 1;
    }
    for (int32_t v = 0; v < n; v++) {
        g->xadj[v + 1] += g->xadj[v];
        g->vwgt[v] = 1;
    }
}

template <class F>
int guarded(hgp_graph *g, F f)
{
    try {
        f();
        hgp::set_error(HGP_OK, "");
        return HGP_OK;
    } catch (const IoError &x) {
        if (g) hgp_graph_free(g);
        hgp::set_error(x.code, x.msg);
        return x.code;
    } catch (const std::bad_alloc &) {
        if (g) hgp_graph_free(g);
        hgp::set_error(HGP_ERR_OOM, "host out of memory");
        return HGP_ERR_OOM;
    }
}

}  // namespace

extern "C" {

int hgp_read_graph(const char *path, hgp_graph *g)
{
    if (!g) {
        hgp::set_error(HGP_ERR_INVALID, "graph is NULL");
        return HGP_ERR_INVALID;
    }
    memset(g, 0, sizeof(*g));
    return guarded(g, [&] { read_metis(path, g); });
}

int hgp_read_mtx(const char *path, hgp_graph *g)
{
    if (!g) {
        hgp::set_error(HGP_ERR_INVALID, "graph is NULL");
        return HGP_ERR_INVALID;
    }
    memset(g, 0, sizeof(*g));
    return guarded(g, [&] { read_mtx(path, g); });
}

int hgp_write_graph(const char *path, const hgp_graph *g, int32_t weights)
{
    return guarded(nullptr, [&] {
        if (!path || !g || g->n < 0 || (g->n > 0 && !g->xadj)) fail(HGP_ERR_INVALID, "bad arguments");
        if (weights < 0 || weights > 11 || weights % 10 > 1) fail(HGP_ERR_INVALID, "weights must be 0, 1, 10 or 11");
        const bool vw = weights / 10 == 1, ew = weights % 10 == 1;
        if ((vw && !g->vwgt) || (ew && g->m > 0 && !g->adjwgt)) fail(HGP_ERR_INVALID, "weights requested but absent");
        FILE *f = fopen(path, "wb");
        if (!f) fail(HGP_ERR_IO, std::string("cannot open ") + path + ": " + strerror(errno));
        std::string line;
        bool ok = true;
        auto put = [&](const std::string &s) { ok = ok && fwrite(s.data(), 1, s.size(), f) == s.size(); };
        line = std::to_string(g->n) + " " + std::to_string(g->m / 2);
        if (weights) line += " " + std::to_string(weights);
        put(line + "\n");
        for (int32_t v = 0; v < g->n; v++) {
            line.clear();
            if (vw) line += std::to_string(g->vwgt[v]);
            for (int64_t j = g->xadj[v]; j < g->xadj[v + 1]; j++) {
                if (!line.empty()) line += ' ';
                line += std::to_string(g->adjncy[j] + 1);
                if (ew) line += ' ' + std::to_string(g->adjwgt[j]);
            }
            put(line + "\n");
        }
        ok = (fclose(f) == 0) && ok;
        if (!ok) fail(HGP_ERR_IO, std::string("write error on ") + path);
    });
}

int hgp_write_partition(const char *fname, const int32_t *part, int32_t n, int32_t nparts)
{
    return guarded(nullptr, [&] {
        if (!fname || n < 0 || (n > 0 && !part)) fail(HGP_ERR_INVALID, "bad arguments");
        const std::string path = std::string(fname) + ".part." + std::to_string(nparts);
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) fail(HGP_ERR_IO, std::string("cannot open ") + path + ": " + strerror(errno));
        std::string buf;
        bool ok = true;
        for (int32_t v = 0; v < n; v++) {
            buf += std::to_string(part[v]);
            buf += '\n';
            if (buf.size() > (1 << 20)) {
                ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
                buf.clear();
            }
        }
        ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
        ok = (fclose(f) == 0) && ok;
        if (!ok) fail(HGP_ERR_IO, "write error on " + path);
    });
}

}  // extern "C"
