// Device record types shared by host-side workspace code and kernels.
#pragma once
#include <stdint.h>

namespace cent {

// Source-parallel step: one record per vertex and slot, so a neighbour's
// BFS level and path count (later coef) share one cache line.  Narrow
// form: 16 bytes, the slot's BC partials in a separate array; wide form: 32
// bytes holding the BC partial too, so the backward's update of a vertex
// (its coef and its BC partial) touches one line.
struct __align__(16) PsRec {
    int dist;        // epoch-tagged BFS level (-1: never reached)
    uint32_t npred;  // predecessor count and successor bits
    double sig;      // path count, then coef = (1 + delta) / sigma
};
struct __align__(32) PsRecW {
    int dist;
    uint32_t npred;
    double sig;
    double bc;       // the slot's BC partial over its sources
    double pad;
};
static_assert(sizeof(PsRec) == 16 && sizeof(PsRecW) == 32, "record sizes");

}  // namespace cent
