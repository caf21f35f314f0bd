// Source-parallel level kernels: one workgroup runs the whole traversal of
// one source at a time, level by level with workgroup barriers, so a call is
// a single launch with no host synchronisation per level.  This is the form
// for high-diameter graphs (grids, road-like meshes), where the 64 lanes of
// a bit-parallel batch sit at different levels and share no work: the dense
// step would rescan every vertex on each of hundreds of levels, while here a
// source touches each reached vertex and edge a constant number of times.
// Each workgroup slot keeps its own per-vertex arrays and takes sources s,
// s + G, s + 2G, ... in turn.
//
// Betweenness (k_ps_bc), per slot:
//   R[v]      record (PsRec, 16 bytes, or PsRecW, 32 bytes), so a
//             neighbour's level and path count share one cache line (and,
//             wide, a vertex's coef and BC partial):
//     .dist   int32   epoch-tagged BFS level: epoch of the slot's current
//                     source << 16 | level; any other value = unreached,
//                     so nothing is restored between sources
//     .npred  uint32  low 28 bits: predecessor count (per-predecessor
//                     formula); high 4 bits: which of a row of at most 4
//                     edges lead to level + 1 (its successors), so the
//                     backward reads only the successors' records
//     .sig    double  path count; the backward overwrites it in place with
//                     coef = (1 + delta) / sigma once v's delta is known
//     .bc     double  wide form: this slot's BC partial over its sources
//   bcp[v]    double  narrow form: the same partials, a separate array
//   Q[]       int32   BFS order; level l is Q[loff[l] .. loff[l+1])
// Forward, level d: each vertex u of the level scans its row once.  A
// neighbour at level d - 1 adds its path count to u's (a pull, in row
// order: sigma[u] = sum of sigma[pred], Seq_Brandes/Brandes.c:286-291), and
// an unreached neighbour is claimed for level d + 1 by compare-and-swap on
// its tag (appended to Q).  u's count is written once, by u's thread, and
// read by level d + 1 after the level's barrier: no FP64 atomics, and the
// record line of u is not read back before it is written.  This is the
// level step of the reference's queue BFS (Brandes.c:274-298) for one
// source.
// Backward, level d from the deepest up: v sums coef[w] over its neighbours
// w at level d + 1 in row order (only the records its successor bits name,
// for a short row), delta = sigma * sum (Brandes.c:317-327 in textbook
// form), then coef[v] replaces sig[v] and its BC partial gets delta.  The
// wide form pays where the slots' arrays are far beyond the Infinity Cache
// (a 2048^2 grid at 256 slots: BC kernel 995 -> 920 ms) and costs where
// they nearly fit (R-MAT 2^14 at 1024 slots: 46 -> 51 ms).  The
// host-side reduce adds the slots in slot order, so BC is deterministic for
// a given G.
//
// Closeness only (k_ps_cc), per slot:
//   V         bitmap, n bits: vertices reached so far.  Claims are atomic ORs
//             on its 32-bit words, and the test "already reached?" on an
//             edge reads one map word: one bit per vertex instead of a
//             16-byte record (2048^2 grid, 1024 sources: 495 -> 304 ms).
//   Q[]       BFS order of the current and the next level.
// Farness is the level of each claimed vertex, summed by its claimer.  The
// same bit maps do not pay for betweenness: a pull of path counts over
// level maps rescans every row once more and scatters over three more maps
// (1338 -> 1547 ms on the same grid), so BC keeps the records above.
//
// Rows longer than kPsBigDeg are queued in LDS and scanned by a whole
// wavefront (lane-strided edges, butterfly sum), so one hub row does not
// serialise a level.
//
// Rows of at most kPsE edges (nearly every vertex of a grid or mesh) are
// read from a 16-byte inline row per vertex, one load that holds the whole
// row, instead of a row-offset line and a column-id line: on the 2048^2
// grid that is about a third of the lines a traversal misses in L2.
//
// Path counts that grow past 2^512 (deep graphs with many parallel
// shortest paths, e.g. chains of diamonds) would leave FP64 a few levels
// later.  A source whose forward writes such a count is undone (its
// distance sums taken back, its records left to the next epoch) and run
// again in scaled form: once a level's counts are written, their largest
// exponent E is known and, when E >= 512, the next level multiplies the
// sums it reads from the level by 2^(256 - E) (exact, a power of two), the
// shift being kept per level in the slot's table; the backward multiplies that
// level's dependencies by the same factor (delta_d = sigma~_d * 2^-k_d *
// sum coef~_{d+1}, with coef~ = (1 + delta) / sigma~ of the stored counts).
// Sources whose counts stay below 2^512 run as before, bit for bit.
#pragma once
#include <type_traits>
#include "records.h"
#include "wave.h"

namespace cent {

#ifndef CENT_PS_BLOCK
#define CENT_PS_BLOCK 512
#endif
constexpr int kPsBlock = CENT_PS_BLOCK;  // 8 waves per workgroup
constexpr int kPsWaves = kPsBlock / kWave;
constexpr int kPsBigDeg = 64;   // longer rows are scanned by a wavefront
constexpr int kPsBig = 1024;    // LDS capacity of the long-row list per level
constexpr int kPsV = 3;         // rows a thread works on at once
constexpr int kPsE = 4;         // edges per row issued together
constexpr int kPsHubE = 8;      // edges per lane issued together on a long row

struct PsArgs {
    int n;
    int nw;                               // closeness: 32-bit words per slot map, a multiple of 4
    const int* __restrict__ rp;
    const int* __restrict__ col;
    // [n] inline rows: the neighbours of a vertex of degree <= kPsE (-1
    // padded), kEllLong in .x for a longer row (read from rp / col); null:
    // every row from rp / col
    const int4* __restrict__ ell;
    const int* __restrict__ src;          // [k] device ids
    int k;
    void* __restrict__ rec;               // BC: [G][n] PsRec or PsRecW
    double* __restrict__ bcp;             // BC, narrow records: [G][n] BC partials
    uint32_t* __restrict__ bits;          // closeness: [G][nw]
    int* __restrict__ Q;                  // [G][n]
    int* __restrict__ loff;               // BC: [G][2 lcap]: level offsets, then per-level count shifts
    int lcap;
    int64_t* __restrict__ far;            // [k]
    int64_t* __restrict__ dist_sum;       // [n] or null
    const int64_t* __restrict__ weight;   // [n] or null
    unsigned long long* __restrict__ cnt; // [2]: reached pairs, DAG edges
    unsigned long long* __restrict__ depth_sum;
    int* __restrict__ overflow;           // path count left FP64, or depth > lcap
};

// Loads of words that other waves change by atomics: agent-scope relaxed
// atomic loads read the L2 copy, never a stale L1 line.
__device__ __forceinline__ int ld_dist(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_sig(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_npred(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct PsLocal {
    int64_t far = 0;
    unsigned long long reached = 0, dag = 0;
};

// ============================================================ betweenness
// Levels carry the slot's source epoch: R[v].dist = ep << kPsLvBits | level
// for a vertex the slot's current source has reached, any other value (an
// older epoch, or -1 from the call's initial fill) for one it has not.  The
// slot bumps ep for every source it takes, so no record is restored between
// sources; before ep would pass kPsEpochMax the slot sets its levels back
// to -1 and starts again at epoch 1.
constexpr int kPsLvBits = 16;               // levels below 2^16 (the level table holds at most 65536)
constexpr int kPsEpochMax = (1 << 15) - 1;  // keeps every tag non-negative
constexpr int kNoEdge = INT32_MIN;          // tag of a staged slot with no edge: never stored
constexpr int kNoLevel = INT32_MIN + 1;     // predecessor tag at level 0: never stored, never kNoEdge

__device__ __forceinline__ int ps_tag(int ep, int l) { return (ep << kPsLvBits) | l; }

// npred word: predecessor count and the successor bits of a short row.
constexpr int kPsSuccShift = 28;
constexpr uint32_t kPsCountMask = (1u << kPsSuccShift) - 1u;
__device__ __forceinline__ bool ps_seen(int t, int ep) { return t >= 0 && (t >> kPsLvBits) == ep; }

// Bookkeeping of a vertex w claimed for level d + 1: appended to the next
// level's part of Q, its level added to the source's farness (and to w's
// distance sum when asked for).
__device__ __forceinline__ void ps_claimed(const PsArgs& a, int* Q, int w, int d, int le, int* s_cnt, PsLocal& L) {
    Q[le + atomicAdd(s_cnt, 1)] = w;
    L.far += a.weight ? (int64_t)(d + 1) * a.weight[w] : (int64_t)(d + 1);
    L.reached++;
    if (a.dist_sum) atomicAdd((unsigned long long*)&a.dist_sum[w], (unsigned long long)(d + 1));
}

// Edge u -> w while u (level d) is expanded: w at level d - 1 (tag tp) adds
// its path count to u's sum; an unreached w is claimed for level d + 1
// (tag tn).
template <class Rec>
__device__ __forceinline__ void ps_visit(const PsArgs& a, Rec* R, int* Q, int w, int ep, int d, int tp, int tn,
                                         int le, int* s_cnt, PsLocal& L, double& acc, uint32_t& np) {
    const int t = ld_dist(&R[w].dist);
    if (t == tp) {
        acc += ld_sig(&R[w].sig);
        ++np;
    } else if (!ps_seen(t, ep) && atomicCAS(&R[w].dist, t, tn) == t) {
        ps_claimed(a, Q, w, d, le, s_cnt, L);
    }
}

// Expansion of up to kPsV rows at once (edges past the first kPsE of a row
// finish in a plain loop): all column loads, then all level loads, then the
// predecessors' path counts and the claims, then the bookkeeping, so one
// thread keeps up to kPsV * kPsE independent chains in flight instead of
// walking its rows edge by edge.  acc[m] / np[m]: row m's sum of its
// predecessors' path counts (in row order) and their number.
constexpr int kEllLong = INT32_MIN;  // inline-row marker: the row is read from rp / col
static_assert(kPsE == 4, "inline rows hold kPsE = 4 neighbours");

// Row of u for the staged expansions: wi = its neighbours from the inline
// row (rb = dg = 0, no tail), or wi.x = kEllLong and rb / dg its CSR row.
// u < 0: an empty row.  A staged row is one int4 register quadruple (an int
// array here left dead stores to scratch memory behind).
__device__ __forceinline__ void ps_row(const PsArgs& a, int u, int4& wi, int& rb, int& dg) {
    rb = 0;
    dg = 0;
    wi = make_int4(-1, -1, -1, -1);
    if (u < 0) return;
    if (a.ell) {
        const int4 e4 = a.ell[u];
        if (e4.x != kEllLong) {
            wi = e4;
            return;
        }
    }
    wi.x = kEllLong;
    rb = a.rp[u];
    dg = a.rp[u + 1] - rb;
}

// Component e (< 4) of v.
__device__ __forceinline__ int ps_lane4(const int4& v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}

// Neighbour e (< kPsE) of a row staged by ps_row, -1 past its end.
__device__ __forceinline__ int ps_nbr(const PsArgs& a, const int4& wi, int rb, int dg, int e) {
    return wi.x != kEllLong ? ps_lane4(wi, e) : (e < dg ? a.col[rb + e] : -1);
}

// True for a row staged by ps_row whose edges all sit in the staged slots
// (an inline row, or a CSR row of at most kPsE edges): its successor bits
// cover the whole row.
__device__ __forceinline__ bool ps_short(const int4& wi, int dg) { return wi.x != kEllLong || dg <= kPsE; }

template <class Rec>
__device__ __forceinline__ void ps_expand_rec(const PsArgs& a, Rec* R, int* Q, const int4* wi,
                                              const int* rb, const int* dg, int ep, int d, int tp, int le,
                                              int* s_cnt, PsLocal& L, double* acc, uint32_t* np, uint32_t* sb) {
    const int tn = ps_tag(ep, d + 1);
    int w[kPsV][kPsE], t[kPsV][kPsE];
    double x[kPsV][kPsE];
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) w[m][e] = ps_nbr(a, wi[m], rb[m], dg[m], e);
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) t[m][e] = w[m][e] >= 0 ? ld_dist(&R[w[m][e]].dist) : kNoEdge;
#pragma unroll
    for (int m = 0; m < kPsV; ++m) {
        sb[m] = 0u;
#pragma unroll
        for (int e = 0; e < kPsE; ++e) {
            const bool pred = w[m][e] >= 0 && t[m][e] == tp;
            x[m][e] = pred ? ld_sig(&R[w[m][e]].sig) : 0.0;
            // An unreached neighbour is claimed; whichever claim wins, it is
            // at level d + 1 once the level is done, a successor, as is one
            // already claimed.  t becomes the claim's old value: stale
            // (not of this epoch) when this claim won, tn when another won.
            const bool tr = w[m][e] >= 0 && !pred && !ps_seen(t[m][e], ep);
            if (tr || t[m][e] == tn) sb[m] |= 1u << e;
            if (tr) t[m][e] = atomicCAS(&R[w[m][e]].dist, t[m][e], tn);
        }
    }
#pragma unroll
    for (int m = 0; m < kPsV; ++m) {
        double s = 0.0;
        uint32_t c = 0;
#pragma unroll
        for (int e = 0; e < kPsE; ++e) {
            if (w[m][e] < 0) continue;
            if (t[m][e] == tp) {
                s += x[m][e];
                ++c;
            } else if (!ps_seen(t[m][e], ep)) {  // this claim won
                ps_claimed(a, Q, w[m][e], d, le, s_cnt, L);
            }
        }
        for (int e = kPsE; e < dg[m]; ++e) ps_visit(a, R, Q, a.col[rb[m] + e], ep, d, tp, tn, le, s_cnt, L, s, c);
        acc[m] = s;
        np[m] = c;
    }
}

// Expansion of a long row [rb, re) for one lane of a wavefront: edges j,
// j + 64, ... taken kPsHubE at a time, staged like ps_expand_rec.  Returns
// the lane's partial sum of predecessor path counts (its edges in order);
// np counts them.
template <class Rec>
__device__ __forceinline__ double ps_expand_long(const PsArgs& a, Rec* R, int* Q, int rb, int re, int lane, int ep,
                                                 int d, int tp, int le, int* s_cnt, PsLocal& L, uint32_t& np) {
    const int tn = ps_tag(ep, d + 1);
    double acc = 0.0;
    for (int j0 = rb + lane; j0 < re; j0 += kWave * kPsHubE) {
        int w[kPsHubE], t[kPsHubE];
        double x[kPsHubE];
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) {
            const int j = j0 + e * kWave;
            w[e] = j < re ? a.col[j] : -1;
        }
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) t[e] = w[e] >= 0 ? ld_dist(&R[w[e]].dist) : kNoEdge;
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) {
            const bool pred = w[e] >= 0 && t[e] == tp;
            x[e] = pred ? ld_sig(&R[w[e]].sig) : 0.0;
            if (w[e] >= 0 && !pred && !ps_seen(t[e], ep)) t[e] = atomicCAS(&R[w[e]].dist, t[e], tn);  // as ps_expand_rec
        }
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) {
            if (w[e] < 0) continue;
            if (t[e] == tp) {
                acc += x[e];
                ++np;
            } else if (!ps_seen(t[e], ep)) {  // this claim won
                ps_claimed(a, Q, w[e], d, le, s_cnt, L);
            }
        }
    }
    return acc;
}

// Backward partial sum of a long row [rb, re) for one lane: coef[w] over the
// lane's edges j, j + 64, ... to neighbours with tag tl, kPsHubE edges'
// loads in flight at a time; fixed order, so the wavefront sum that follows
// is deterministic.
template <class Rec>
__device__ __forceinline__ double ps_pull_long(const PsArgs& a, const Rec* R, int rb, int re, int lane, int tl) {
    double acc = 0.0;
    for (int j0 = rb + lane; j0 < re; j0 += kWave * kPsHubE) {
        int w[kPsHubE], lev[kPsHubE];
        double x[kPsHubE];
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) {
            const int j = j0 + e * kWave;
            w[e] = j < re ? a.col[j] : -1;
        }
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) lev[e] = w[e] >= 0 ? ld_dist(&R[w[e]].dist) : kNoEdge;
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) x[e] = lev[e] == tl ? ld_sig(&R[w[e]].sig) : 0.0;
#pragma unroll
        for (int e = 0; e < kPsHubE; ++e) acc += x[e];
    }
    return acc;
}

// Backward sums of up to kPsV rows at once: acc[m] = sum over row m's
// neighbours w with tag tl of coef[w] (held in sig), in row order.  A short
// row with its successor bits (mk[m] >= 0) reads only its successors'
// counts; any other row (mk[m] < 0) tests every neighbour's level.
template <class Rec>
__device__ __forceinline__ void ps_pull_rec(const PsArgs& a, const Rec* R, const int4* wi, const int* rb,
                                             const int* dg, const int* mk, int tl, double* acc) {
    int w[kPsV][kPsE], lev[kPsV][kPsE];
    double x[kPsV][kPsE];
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) w[m][e] = ps_nbr(a, wi[m], rb[m], dg[m], e);
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) lev[m][e] = mk[m] < 0 && w[m][e] >= 0 ? ld_dist(&R[w[m][e]].dist) : kNoEdge;
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) {
            const bool on = mk[m] >= 0 ? w[m][e] >= 0 && ((mk[m] >> e) & 1) : lev[m][e] == tl;
            x[m][e] = on ? ld_sig(&R[w[m][e]].sig) : 0.0;
        }
#pragma unroll
    for (int m = 0; m < kPsV; ++m) {
        double t = 0.0;
#pragma unroll
        for (int e = 0; e < kPsE; ++e) t += x[m][e];
        for (int e = kPsE; e < dg[m]; ++e) {
            const int y = a.col[rb[m] + e];
            if (ld_dist(&R[y].dist) == tl) t += ld_sig(&R[y].sig);
        }
        acc[m] = t;
    }
}

template <bool COMPAT, bool WIDE>
__global__ __launch_bounds__(kPsBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_ps_bc(PsArgs a) {
    using Rec = typename std::conditional<WIDE, PsRecW, PsRec>::type;
    __shared__ int s_cnt[3];     // appends of level d go to s_cnt[d % 3]
    __shared__ int s_nbig[3];    // long rows of level d: s_big[..s_nbig[d % 3])
    __shared__ int s_emax[3];    // scaled form: largest count exponent written at level d
    __shared__ int s_big[kPsBig];
    __shared__ int64_t s_red64[kPsWaves][3];
    __shared__ int s_large;      // a count of the unscaled form reached 2^512
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid >> 6;
    const size_t so = (size_t)blockIdx.x * (size_t)a.n;
    Rec* R = static_cast<Rec*>(a.rec) + so;
    double* bcp = WIDE ? nullptr : a.bcp + so;
    auto bc_add = [&](int v, double x) {
        if constexpr (WIDE) R[v].bc += x;
        else bcp[v] += x;
    };
    int* Q = a.Q + so;
    int* loff = a.loff + (size_t)blockIdx.x * (size_t)a.lcap * 2;
    int* lsh = loff + a.lcap;    // scaled form: shift k_d of level d's counts
    int scaled_si = -1;          // the source now being redone in scaled form
    int ep = 0;                  // epoch of the slot's current source
    for (int si = blockIdx.x; si < a.k; si += gridDim.x) {
        if (ep == kPsEpochMax) {  // epochs spent: every level back to unreached
            for (int v = tid; v < a.n; v += kPsBlock) R[v].dist = -1;
            ep = 0;
            __syncthreads();
        }
        ++ep;
        const int s = a.src[si];
        const bool scaled = si == scaled_si;
        if (tid == 0) {
            s_large = 0;
            Q[0] = s;
            R[s].dist = ps_tag(ep, 0);
            R[s].sig = 1.0;
            loff[0] = 0;
            loff[1] = 1;
            s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
            s_nbig[0] = s_nbig[1] = s_nbig[2] = 0;
            s_emax[0] = s_emax[1] = s_emax[2] = 0;
        }
        __syncthreads();
        PsLocal L;
        int lb = 0, le = 1, d = 0;
        bool deep = false;
        double scp = 1.0;  // scale of the counts of level d - 1, which level d sums
        // ---------------- forward
        // Level d: every vertex u of the level scans its row once; it sums
        // the path counts of its neighbours at level d - 1 (all final: they
        // were written during level d - 1) into its own count, and claims
        // its unreached neighbours for level d + 1.  No count is added to
        // by atomics, and u's count is written once, by u's thread.
        for (;;) {
            const int c = d % 3;
            if (tid == 0) {  // free the buffers level d + 1 will use (last read two levels ago)
                s_cnt[(d + 1) % 3] = 0;
                s_nbig[(d + 1) % 3] = 0;
                s_emax[(d + 1) % 3] = 0;
            }
            const int tp = d > 0 ? ps_tag(ep, d - 1) : kNoLevel;
            // u's count from its predecessors, with its predecessor count and
            // successor bits
            auto finish = [&](int u, double acc, uint32_t np, uint32_t sb) {
                const double sg = acc * scp;  // exact: scp is a power of two
                R[u].sig = sg;
                R[u].npred = np | (sb << kPsSuccShift);
                if (!scaled) {
                    if (sg >= 0x1p512) s_large = 1;
                } else if (dexp(sg) >= kSigBig) {
                    atomicMax(&s_emax[c], dexp(sg));
                }
            };
            for (int i0 = lb + tid; i0 < le; i0 += kPsBlock * kPsV) {
                int u[kPsV], rb[kPsV], dg[kPsV];
                int4 wi[kPsV];
#pragma unroll
                for (int m = 0; m < kPsV; ++m) {
                    const int i = i0 + m * kPsBlock;
                    u[m] = i < le ? Q[i] : -1;
                }
#pragma unroll
                for (int m = 0; m < kPsV; ++m) ps_row(a, u[m], wi[m], rb[m], dg[m]);
                bool big[kPsV];
#pragma unroll
                for (int m = 0; m < kPsV; ++m) {
                    big[m] = false;
                    if (dg[m] > kPsBigDeg) {
                        const int slot = atomicAdd(&s_nbig[c], 1);
                        if (slot < kPsBig) {
                            s_big[slot] = u[m];
                            dg[m] = 0;
                            big[m] = true;
                        }
                    }
                }
                double acc[kPsV];
                uint32_t np[kPsV], sb[kPsV];
                ps_expand_rec(a, R, Q, wi, rb, dg, ep, d, tp, le, &s_cnt[c], L, acc, np, sb);
#pragma unroll
                for (int m = 0; m < kPsV; ++m) {
                    if (u[m] < 0 || big[m]) continue;
                    L.dag += np[m];
                    if (d > 0) finish(u[m], acc[m], np[m], sb[m]);
                }
            }
            __syncthreads();
            const int nbig = min(s_nbig[c], kPsBig);
            for (int t = wid; t < nbig; t += kPsWaves) {
                const int u = s_big[t];
                uint32_t np = 0;
                const double acc = wave_sum(ps_expand_long(a, R, Q, a.rp[u], a.rp[u + 1], lane, ep, d, tp, le,
                                                           &s_cnt[c], L, np));
                L.dag += np;
                const uint32_t npt = (uint32_t)wave_sum_i64((int64_t)np);
                if (lane == 0 && d > 0) finish(u, acc, npt, 0u);
            }
            __syncthreads();
            const int nnew = s_cnt[c];
            if (s_large) {  // undo this pass of the source; it runs again in scaled form
                le += nnew;
                break;
            }
            if (scaled) {  // scale of level d's counts, for level d + 1's sums and the backward
                const int kd = s_emax[c] >= kSigBig ? s_emax[c] - kSigTo : 0;
                if (tid == 0) lsh[d] = kd;
                scp = __longlong_as_double((long long)(1023 - kd) << 52);
            }
            if (nnew == 0) break;
            if (d + 3 > a.lcap) {  // level offsets would not fit: the host reruns another way
                deep = true;
                le += nnew;
                break;
            }
            lb = le;
            le += nnew;
            ++d;
            if (tid == 0) loff[d + 1] = le;
        }
        const int depth = d;
        const int tail = le;
        if (s_large) {
            // this pass's distance sums taken back (integers, so the redo
            // leaves them exactly as one pass would); its records are left
            // to the next epoch
            if (a.dist_sum)
                for (int i = tid; i < tail; i += kPsBlock) {
                    const int v = Q[i];
                    const int dv = R[v].dist & ((1 << kPsLvBits) - 1);
                    if (dv > 0) atomicAdd((unsigned long long*)&a.dist_sum[v], (unsigned long long)(-(int64_t)dv));
                }
            __syncthreads();
            scaled_si = si;
            si -= gridDim.x;  // the same source again
            continue;
        }
        if (deep && tid == 0) atomicOr(a.overflow, 2);
        // ---------------- backward
        if (!deep) {
            if (tid == 0) s_nbig[0] = s_nbig[1] = s_nbig[2] = 0;  // the forward's lists are spent
            __syncthreads();
            for (int dd = depth; dd >= 1; --dd) {
                const int c = dd % 3;
                if (tid == 0) s_nbig[(dd + 2) % 3] = 0;  // the list level dd - 1 will use
                const int b0 = loff[dd], b1 = loff[dd + 1];
                const int tl = ps_tag(ep, dd + 1);
                const double sc = scaled ? __longlong_as_double((long long)(1023 - lsh[dd]) << 52) : 1.0;
                for (int i0 = b0 + tid; i0 < b1; i0 += kPsBlock * kPsV) {
                    int v[kPsV], rb[kPsV], dg[kPsV];
                    double sv[kPsV];
#pragma unroll
                    for (int m = 0; m < kPsV; ++m) {
                        const int i = i0 + m * kPsBlock;
                        v[m] = i < b1 ? Q[i] : -1;
                    }
                    int4 wi[kPsV];
                    int mk[kPsV];
                    uint32_t pw[kPsV];
#pragma unroll
                    for (int m = 0; m < kPsV; ++m) {
                        ps_row(a, v[m], wi[m], rb[m], dg[m]);
                        sv[m] = v[m] >= 0 ? ld_sig(&R[v[m]].sig) : 1.0;
                        pw[m] = v[m] >= 0 ? ld_npred(&R[v[m]].npred) : 0u;
                        mk[m] = v[m] >= 0 && ps_short(wi[m], dg[m]) ? (int)(pw[m] >> kPsSuccShift) : -1;
                    }
                    // long rows always go to the wavefronts (listed while the
                    // list has room), so each row's sum has one order whoever
                    // takes it and BC has the same bits on every run
                    bool big[kPsV];
#pragma unroll
                    for (int m = 0; m < kPsV; ++m) {
                        big[m] = dg[m] > kPsBigDeg;
                        if (big[m]) {
                            const int slot = atomicAdd(&s_nbig[c], 1);
                            if (slot < kPsBig) s_big[slot] = v[m];
                            dg[m] = 0;
                        }
                    }
                    double acc[kPsV];
                    ps_pull_rec(a, R, wi, rb, dg, mk, tl, acc);
#pragma unroll
                    for (int m = 0; m < kPsV; ++m) {
                        if (v[m] < 0 || big[m]) continue;
                        const double dl = (sv[m] * acc[m]) * sc;
                        if (__builtin_isinf(sv[m])) atomicOr(a.overflow, 1);
                        R[v[m]].sig = (1.0 + dl) / sv[m];
                        bc_add(v[m], COMPAT ? dl * (double)(pw[m] & kPsCountMask) : dl);
                    }
                }
                __syncthreads();
                auto long_row = [&](int v) {
                    const double acc = wave_sum(ps_pull_long(a, R, a.rp[v], a.rp[v + 1], lane, tl));
                    if (lane == 0) {
                        const double sv = ld_sig(&R[v].sig);
                        const double dl = (sv * acc) * sc;
                        if (__builtin_isinf(sv)) atomicOr(a.overflow, 1);
                        R[v].sig = (1.0 + dl) / sv;
                        bc_add(v, COMPAT ? dl * (double)(ld_npred(&R[v].npred) & kPsCountMask) : dl);
                    }
                };
                const int nbig = s_nbig[c];
                if (nbig <= kPsBig) {
                    for (int t = wid; t < nbig; t += kPsWaves) long_row(s_big[t]);
                } else {  // the list overflowed: the wavefronts take the level's long rows from Q
                    for (int i = b0 + wid; i < b1; i += kPsWaves) {
                        const int v = Q[i];
                        if (a.rp[v + 1] - a.rp[v] > kPsBigDeg) long_row(v);
                    }
                }
                __syncthreads();
            }
        }
        // ---------------- per-source totals
        int64_t far = L.far;
        int64_t rch = (int64_t)L.reached, dag = (int64_t)L.dag;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            far += (int64_t)__shfl_xor((long long)far, o);
            rch += (int64_t)__shfl_xor((long long)rch, o);
            dag += (int64_t)__shfl_xor((long long)dag, o);
        }
        if (lane == 0) {
            s_red64[wid][0] = far;
            s_red64[wid][1] = rch;
            s_red64[wid][2] = dag;
        }
        __syncthreads();
        if (tid == 0) {
            int64_t f = 0, r = 0, g = 0;
            for (int w = 0; w < kPsWaves; ++w) {
                f += s_red64[w][0];
                r += s_red64[w][1];
                g += s_red64[w][2];
            }
            a.far[si] = f;
            atomicAdd(&a.cnt[0], (unsigned long long)r);
            atomicAdd(&a.cnt[1], (unsigned long long)g);
            atomicAdd(a.depth_sum, (unsigned long long)depth);
        }
        __syncthreads();  // totals read before the next source writes s_red64
    }
}


// ============================================================ closeness only
__device__ __forceinline__ bool bit_in(const uint32_t* M, int v) {
    return (ld_npred(M + (v >> 5)) >> (v & 31)) & 1u;
}
// True for the one caller that flips v's bit.
__device__ __forceinline__ bool bit_claim(uint32_t* M, int v) {
    const uint32_t m = 1u << (v & 31);
    return (atomicOr(M + (v >> 5), m) & m) == 0u;
}

// Expands up to kPsV rows at once, as ps_expand_rec: all column loads, then
// all map words, then all claims, then the bookkeeping.
__device__ __forceinline__ void ps_expand_bits(const PsArgs& a, uint32_t* V, int* Q, const int4* wi,
                                               const int* rb, const int* dg, int d, int le, int* s_cnt, PsLocal& L) {
    int w[kPsV][kPsE];
    bool seen[kPsV][kPsE], got[kPsV][kPsE];
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) w[m][e] = ps_nbr(a, wi[m], rb[m], dg[m], e);
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) seen[m][e] = w[m][e] < 0 || bit_in(V, w[m][e]);
#pragma unroll
    for (int m = 0; m < kPsV; ++m)
#pragma unroll
        for (int e = 0; e < kPsE; ++e) got[m][e] = !seen[m][e] && bit_claim(V, w[m][e]);
#pragma unroll
    for (int m = 0; m < kPsV; ++m) {
#pragma unroll
        for (int e = 0; e < kPsE; ++e)
            if (got[m][e]) ps_claimed(a, Q, w[m][e], d, le, s_cnt, L);
        for (int e = kPsE; e < dg[m]; ++e) {
            const int x = a.col[rb[m] + e];
            if (!bit_in(V, x) && bit_claim(V, x)) ps_claimed(a, Q, x, d, le, s_cnt, L);
        }
    }
}

__global__ __launch_bounds__(kPsBlock) void k_ps_cc(PsArgs a) {
    __shared__ int s_cnt[3];     // appends of level d go to s_cnt[d % 3]
    __shared__ int s_nbig[3];    // long rows of level d: s_big[..s_nbig[d % 3])
    __shared__ int s_big[kPsBig];
    __shared__ int64_t s_red64[kPsWaves][2];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid >> 6;
    uint32_t* V = a.bits + (size_t)blockIdx.x * (size_t)a.nw;
    int* Q = a.Q + (size_t)blockIdx.x * (size_t)a.n;
    for (int si = blockIdx.x; si < a.k; si += gridDim.x) {
        const int s = a.src[si];
        {   // the slot's map empty (16-byte aligned: nw is a multiple of 4)
            uint4* w4 = (uint4*)V;
            for (int i = tid; i < a.nw / 4; i += kPsBlock) w4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        if (tid == 0) {
            Q[0] = s;
            atomicOr(V + (s >> 5), 1u << (s & 31));
            s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
            s_nbig[0] = s_nbig[1] = s_nbig[2] = 0;
        }
        __syncthreads();
        PsLocal L;
        int lb = 0, le = 1, d = 0;
        for (;;) {
            const int c = d % 3;
            if (tid == 0) {  // free the counters level d + 1 will use (last read two levels ago)
                s_cnt[(d + 1) % 3] = 0;
                s_nbig[(d + 1) % 3] = 0;
            }
            for (int i0 = lb + tid; i0 < le; i0 += kPsBlock * kPsV) {
                int rb[kPsV], dg[kPsV];
                int4 wi[kPsV];
#pragma unroll
                for (int m = 0; m < kPsV; ++m) {
                    const int i = i0 + m * kPsBlock;
                    const int u = i < le ? Q[i] : -1;
                    ps_row(a, u, wi[m], rb[m], dg[m]);
                    if (dg[m] > kPsBigDeg) {
                        const int slot = atomicAdd(&s_nbig[c], 1);
                        if (slot < kPsBig) {
                            s_big[slot] = u;
                            dg[m] = 0;
                        }
                    }
                }
                ps_expand_bits(a, V, Q, wi, rb, dg, d, le, &s_cnt[c], L);
            }
            __syncthreads();
            const int nbig = min(s_nbig[c], kPsBig);
            for (int t = wid; t < nbig; t += kPsWaves) {
                const int u = s_big[t];
                const int re = a.rp[u + 1];
                for (int j0 = a.rp[u] + lane; j0 < re; j0 += kWave * kPsHubE) {  // staged as ps_expand_bits
                    int w[kPsHubE];
                    bool seen[kPsHubE];
#pragma unroll
                    for (int e = 0; e < kPsHubE; ++e) {
                        const int j = j0 + e * kWave;
                        w[e] = j < re ? a.col[j] : -1;
                    }
#pragma unroll
                    for (int e = 0; e < kPsHubE; ++e) seen[e] = w[e] < 0 || bit_in(V, w[e]);
#pragma unroll
                    for (int e = 0; e < kPsHubE; ++e)
                        if (!seen[e] && bit_claim(V, w[e])) ps_claimed(a, Q, w[e], d, le, &s_cnt[c], L);
                }
            }
            __syncthreads();
            const int nnew = s_cnt[c];
            if (nnew == 0) break;
            lb = le;
            le += nnew;
            ++d;
        }
        int64_t far = L.far, rch = (int64_t)L.reached;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            far += (int64_t)__shfl_xor((long long)far, o);
            rch += (int64_t)__shfl_xor((long long)rch, o);
        }
        if (lane == 0) {
            s_red64[wid][0] = far;
            s_red64[wid][1] = rch;
        }
        __syncthreads();
        if (tid == 0) {
            int64_t f = 0, r = 0;
            for (int w = 0; w < kPsWaves; ++w) {
                f += s_red64[w][0];
                r += s_red64[w][1];
            }
            a.far[si] = f;
            atomicAdd(&a.cnt[0], (unsigned long long)r);
            atomicAdd(a.depth_sum, (unsigned long long)d);
        }
        __syncthreads();  // totals read before the next source's clears
    }
}

// bc[v] = sum over slots g = 0 .. G-1 of the slot's BC partial of v, in
// slot order.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_ps_reduce(const void* __restrict__ rec, const double* __restrict__ bcp,
                                                   int G, int n, double* __restrict__ bc) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double t = 0.0;
    for (int g = 0; g < G; ++g) {
        const size_t i = (size_t)g * (size_t)n + v;
        t += WIDE ? static_cast<const PsRecW*>(rec)[i].bc : bcp[i];
    }
    bc[v] += t;
}

// Betweenness slot records to their idle state: never reached, BC partial 0.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_ps_init(void* __restrict__ rec, double* __restrict__ bcp, size_t total) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        if constexpr (WIDE) {
            static_cast<PsRecW*>(rec)[i] = PsRecW{-1, 0u, 0.0, 0.0, 0.0};
        } else {
            static_cast<PsRec*>(rec)[i] = PsRec{-1, 0u, 0.0};
            bcp[i] = 0.0;
        }
    }
}

}  // namespace cent
