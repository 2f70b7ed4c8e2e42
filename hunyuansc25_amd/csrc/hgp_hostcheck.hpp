// hgp_hostcheck.hpp -- the library's error type and the host-side CSR check.
// Plain C++ (no HIP), shared by the device code's headers and the host-only
// sources, so the host-only parts build without the HIP headers too.
#pragma once

#include <cstdint>
#include <string>

namespace hgp {

struct Error {
    int code;
    std::string msg;
    bool overflow = false;   // an arena request did not fit (a larger arena may succeed)
    size_t need = 0;         // with overflow: arena bytes in use plus the request that failed
};

void set_error(int code, const std::string &msg);

// Host-side CSR checks (the invariants of check_graph,
// hunyuangraph_GPU_match.h:1718-1800): xadj[0] = 0, monotone offsets,
// vertex weights >= 1, neighbour ids in [0, n), no self-loops, edge weights
// >= 1, at most 2^31-1 entries and total edge weight; with `symmetric`, every
// (v, u, w) has its (u, v, w).  Throws Error{-1} naming the first fault,
// Error{-4} past the supported sizes.
void validate_host(int32_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *vwgt,
                   const int32_t *adjwgt, bool symmetric);

}  // namespace hgp
