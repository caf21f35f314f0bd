// Device record types shared by host-side workspace code and kernels.
#pragma once
#include <stdint.h>

namespace cent {

// Source-parallel step: one 16-byte record per vertex and slot, so a
// neighbour's BFS level and path count (later coef) share one cache line.
struct __align__(16) PsRec {
    int dist;        // BFS level, -1 = unreached
    uint32_t npred;  // predecessor count (per-predecessor formula only)
    double sig;      // path count, then coef = (1 + delta) / sigma
};

}  // namespace cent
