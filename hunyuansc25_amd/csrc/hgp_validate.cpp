// hgp_validate.cpp -- host-side CSR check of the host-array entry points
// (plain C++; no device code).
#include "hgp_hostcheck.hpp"

#include <algorithm>
#include <climits>
#include <utility>
#include <vector>

namespace hgp {

void validate_host(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                   const int32_t *adjwgt, bool symmetric)
{
    if (n < 0) throw Error{-1, "nvtxs < 0"};
    if (!xadj) throw Error{-1, "xadj is NULL"};
    if (xadj[0] != 0) throw Error{-1, "xadj[0] != 0"};
    for (int32_t v = 0; v < n; v++)
        if (xadj[v + 1] < xadj[v]) throw Error{-1, "xadj is not monotone at vertex " + std::to_string(v)};
    int64_t m = xadj[n];
    if (m > INT32_MAX) throw Error{-4, "more than 2^31-1 directed entries"};
    if (m > 0 && !adjncy) throw Error{-1, "adjncy is NULL"};
    int64_t wsum = 0;
    for (int32_t v = 0; v < n; v++) {
        if (vwgt && vwgt[v] <= 0) throw Error{-1, "vertex weight <= 0 at vertex " + std::to_string(v)};
        for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
            int32_t u = adjncy[j];
            if (u < 0 || u >= n) throw Error{-1, "neighbour id out of range in row " + std::to_string(v)};
            if (u == v) throw Error{-1, "self-loop at vertex " + std::to_string(v)};
            int32_t w = adjwgt ? adjwgt[j] : 1;
            if (w <= 0) throw Error{-1, "edge weight <= 0 in row " + std::to_string(v)};
            wsum += w;
        }
    }
    if (wsum > INT32_MAX) throw Error{-4, "total edge weight exceeds 2^31-1"};
    if (symmetric) {
        // every (v,u,w) must have a matching (u,v,w): compare sorted triples
        std::vector<std::pair<uint64_t, int32_t>> a, b;
        a.reserve((size_t)m);
        b.reserve((size_t)m);
        for (int32_t v = 0; v < n; v++)
            for (int64_t j = xadj[v]; j < xadj[v + 1]; j++) {
                int32_t u = adjncy[j], w = adjwgt ? adjwgt[j] : 1;
                a.push_back({((uint64_t)v << 32) | (uint32_t)u, w});
                b.push_back({((uint64_t)u << 32) | (uint32_t)v, w});
            }
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        if (a != b) throw Error{-1, "graph is not symmetric"};
    }
}

}  // namespace hgp
