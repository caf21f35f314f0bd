// Closeness level step over W side-by-side 64-source batches (W = 1, 2, 4,
// 8, 16, 32 or 64; the batches of a launch group are numbered b = 0 .. nb - 1
// <= W - 1).  The masks are interleaved by vertex:
//   vis[v * W + b]  lanes of batch b that have reached v
//   F[v * W + b]    lanes of batch b at the current distance
// so one scan of v's row serves every batch, and a neighbour's masks for all
// batches are one contiguous W x 8-byte read (one 128-byte cache line at
// W = 16, four at W = 64) instead of W reads from W separate planes.
//
// Lane layout inside a wave: b = lane % W is the batch, e = lane / W the edge
// slot, so a row is read E = 64 / W edges at a time and each lane ORs the
// masks of its batch.  Same pull as the reference's queue BFS level
// (Seq_CC/CC.cpp:81-98), for 64 * W sources at once; only which lanes reach
// a vertex matters, so a row stops once every candidate lane is found.
#pragma once
#include "msbfs_kernels.h"

namespace cent {

constexpr int kCcwUnroll = 4;  // edge groups whose loads are issued together
// The same in the hub-segment kernel, whose long rows gain from more loads
// in flight: all-vertex closeness on a 174k-vertex R-MAT 45.5 -> 44.7 ms
// reduced, 79.3 -> 77.0 ms plain at 8 (12 and 16: the same as 8; 8 in the
// main kernel too: slower).
constexpr int kCcwUnrollHub = 8;

// Workgroup size of the kernels that keep a per-workgroup LDS table of lane
// partials (k_ccw, k_ccw_hubfin): W x 64 counters, 32 KB at W = 64.  Each
// workgroup adds its table into its own slot of part_far at the end of every
// level, and at W = 64 the table, not the registers, limits the waves a CU
// holds (4 a SIMD at 256 threads).  There 512-thread workgroups write half
// the partials and hold 6 waves a SIMD: all-vertex closeness on a 174k-vertex
// R-MAT 48.8 -> 45.5 ms reduced, 83.6 -> 79.0 ms plain (1024 threads: no
// gain).  At W = 32 the two sizes measure the same, and at W = 16 512
// threads lose (4096 sources on R-MAT 2^20: 8.45 -> 8.9 ms): a workgroup
// waits at its flush for its slowest wave.
constexpr int kCcwBigBlock = 512;
constexpr int ccw_block(int W) { return W >= 64 ? kCcwBigBlock : kBlock; }

struct CcwArgs {
    int hub_deg;
    int nchunk;
    const int* __restrict__ chunk_beg;
    const int* __restrict__ rp;
    const int* __restrict__ col;
    uint64_t* __restrict__ vis;            // [N * W]
    const uint64_t* __restrict__ Fd;       // [N * W], distance d
    uint64_t* __restrict__ Fn;             // [N * W], distance d + 1 (zeroed)
    int64_t* __restrict__ dist_sum;        // [n] or null
    const int64_t* __restrict__ weight;    // [n] vertex weights or null (== 1)
    int nb;                                // batches in the group
    uint64_t A_last;                       // active lanes of batch nb - 1
    int dnext;                             // d + 1
    int pslots;                            // partial slots per batch
    int64_t* __restrict__ part_far;        // [nb][pslots][64] lane partials
    unsigned long long* __restrict__ part_cnt;  // [nb][pslots][2]
    int* __restrict__ flag;                // set when anything was reached
    const int* __restrict__ live;          // level d non-empty (null: always)
    // hub rows (longer than hub_deg), cut into edge segments
    int nseg;
    const int* __restrict__ seg_v;
    const int* __restrict__ seg_beg;
    const int* __restrict__ seg_end;
    int seg_base;
    int nhub;
    const int* __restrict__ hub_v;
    const int* __restrict__ hub_seg_off;
    uint64_t* __restrict__ part_or;        // [nseg][W]
};

// Active lanes of batch b (0 past the group's last batch).
__device__ __forceinline__ uint64_t ccw_active(const CcwArgs& a, int b) {
    return b < a.nb - 1 ? ~0ull : (b == a.nb - 1 ? a.A_last : 0ull);
}

// OR over the lanes of one batch (lanes b, b + W, b + 2W, ...).
template <int W>
__device__ __forceinline__ uint64_t ccw_batch_or(uint64_t x) {
#pragma unroll
    for (int o = W; o < kWave; o <<= 1) x |= (uint64_t)__shfl_xor((unsigned long long)x, o);
    return x;
}

// Lanes of batch b that reach some neighbour in edges [beg, end) of Fd,
// restricted to cand; the result is in every lane of batch b.
template <int W, int U = kCcwUnroll>
__device__ __forceinline__ uint64_t ccw_scan(const CcwArgs& a, int beg, int end, uint64_t cand, int lane) {
    constexpr int E = kWave / W;
    const int b = lane & (W - 1), e = lane / W;
    uint64_t orf = 0;
    for (int j0 = beg; j0 < end; j0 += E * U) {
        int u[U];
#pragma unroll
        for (int t = 0; t < U; ++t) {
            const int j = j0 + t * E + e;
            u[t] = j < end ? a.col[j] : -1;
        }
#pragma unroll
        for (int t = 0; t < U; ++t)
            if (u[t] >= 0) orf |= a.Fd[(size_t)u[t] * W + b];
        orf &= cand;
        if (j0 + E * U < end) {
            const uint64_t g = ccw_batch_or<W>(orf);
            if (ballot(g != cand) == 0) return g;  // every candidate lane found
        }
    }
    return ccw_batch_or<W>(orf);
}

// Weights reached per source, kept per lane as bit-sliced counters: plane k
// holds bit k of the counter of each source the lane covers (lane (b, e):
// sources e * W .. e * W + W - 1 of batch b, bit t for source e * W + t).
// Adding weight w for the new lanes nm is a ripple-carry add of nm into the
// planes at each set bit of w, a few register operations where one LDS
// atomic per new (vertex, source) pair cost about a fifth of the closeness
// step.  A carry out of the top plane, and the bits of w above it, go to
// the block's LDS partials directly; the planes are added into them once,
// at the end of the kernel (ccw_count_flush).
constexpr int kCcwPlanes = 8;
// Groups narrower than this keep one LDS atomic per new pair: a lane then
// covers few sources (one at W = 1, the per-item schedule's width) and the
// ripple over the planes costs more than its atomics.
constexpr int kCcwSlicedFrom = 8;

struct CcwCount {
    uint64_t p[kCcwPlanes] = {};
    bool used = false;  // anything counted (else the flush is skipped)
};

template <int W>
__device__ __forceinline__ void ccw_count(CcwCount& c, long long (*s_far)[kWave], uint64_t nm, long long w,
                                          int lane) {
    const int b = lane & (W - 1), e = lane / W;
    const uint64_t bits = W == 64 ? nm : (nm >> (e * W)) & ((1ull << (W & 63)) - 1ull);
    if constexpr (W < kCcwSlicedFrom) {
        for (uint64_t m = bits; m; m &= m - 1)
            atomicAdd((unsigned long long*)&s_far[b][e * W + __builtin_ctzll(m)], (unsigned long long)w);
        return;
    }
    c.used = true;
    uint64_t carry_out = 0;
    if (w == 1) {
        uint64_t carry = bits;
#pragma unroll
        for (int k = 0; k < kCcwPlanes; ++k) {
            const uint64_t t = c.p[k] & carry;
            c.p[k] ^= carry;
            carry = t;
        }
        carry_out = carry;
    } else {
#pragma unroll
        for (int j = 0; j < kCcwPlanes; ++j) {
            uint64_t carry = ((w >> j) & 1) ? bits : 0ull;
#pragma unroll
            for (int k = j; k < kCcwPlanes; ++k) {
                const uint64_t t = c.p[k] & carry;
                c.p[k] ^= carry;
                carry = t;
            }
            // a carry out of plane j's ripple is worth 2^kCcwPlanes like any other
            for (uint64_t m = carry; m; m &= m - 1)
                atomicAdd((unsigned long long*)&s_far[b][e * W + __builtin_ctzll(m)], 1ull << kCcwPlanes);
        }
        const unsigned long long hi = (unsigned long long)(w >> kCcwPlanes) << kCcwPlanes;
        if (hi)
            for (uint64_t m = bits; m; m &= m - 1) atomicAdd((unsigned long long*)&s_far[b][e * W + __builtin_ctzll(m)], hi);
    }
    for (uint64_t m = carry_out; m; m &= m - 1)
        atomicAdd((unsigned long long*)&s_far[b][e * W + __builtin_ctzll(m)], 1ull << kCcwPlanes);
}

// The lane's counters into the block's LDS partials.
template <int W>
__device__ __forceinline__ void ccw_count_flush(const CcwCount& c, long long (*s_far)[kWave], int lane) {
    const int b = lane & (W - 1), e = lane / W;
    if constexpr (W < kCcwSlicedFrom) return;
    if (ballot(c.used) == 0) return;
    for (int t = 0; t < (W == 64 ? 64 : W); ++t) {
        unsigned long long cnt = 0;
#pragma unroll
        for (int k = 0; k < kCcwPlanes; ++k) cnt |= ((c.p[k] >> t) & 1ull) << k;
        if (cnt) atomicAdd((unsigned long long*)&s_far[b][e * W + t], cnt);
    }
}

// Finishes vertex v: writes the new masks, vis, the lane partials and the
// distance sum; returns the number of new (source, v) pairs (lane 0).
template <int W>
__device__ __forceinline__ unsigned long long ccw_done(const CcwArgs& a, int v, uint64_t old, uint64_t nm,
                                                       CcwCount& cnt, long long (*s_far)[kWave], int lane) {
    const int b = lane & (W - 1), e = lane / W;
    if (e == 0) {
        a.Fn[(size_t)v * W + b] = nm;  // every entry of Fn is written each level (no clear)
        if (nm) a.vis[(size_t)v * W + b] = old | nm;
    }
    if (ballot(nm != 0) == 0) return 0;
    ccw_count<W>(cnt, s_far, nm, a.weight ? (long long)a.weight[v] : 1ll, lane);
    const long long pc = (long long)wave_sum_i64(e == 0 ? (int64_t)__popcll(nm) : 0);
    if (lane == 0 && a.dist_sum) a.dist_sum[v] += (int64_t)a.dnext * pc;
    return (unsigned long long)pc;
}

// Block flush: batch b's partials into its slot `slot` (plain adds; the slot
// belongs to this block).
template <int W>
__device__ __forceinline__ void ccw_flush(const CcwArgs& a, long long (*s_far)[kWave], unsigned long long* s_new,
                                          int slot) {
    __syncthreads();
    for (int i = threadIdx.x; i < W * kWave; i += blockDim.x) {
        const int b = i / kWave, s = i % kWave;
        if (b < a.nb && s_far[b][s])
            a.part_far[((size_t)b * a.pslots + slot) * kWave + s] += (int64_t)a.dnext * s_far[b][s];
    }
    if (threadIdx.x == 0 && *s_new) {
        a.part_cnt[(size_t)slot * 2] += *s_new;
        *a.flag = 1;
    }
}

template <int W>
__device__ __forceinline__ void ccw_block_init(long long (*s_far)[kWave], unsigned long long* s_new) {
    for (int i = threadIdx.x; i < W * kWave; i += blockDim.x) s_far[i / kWave][i % kWave] = 0;
    if (threadIdx.x == 0) *s_new = 0;
    __syncthreads();
}

// Held to 64 VGPRs (8 waves a SIMD; a few registers of the count and flush
// spill) against the unconstrained 63-115: all-vertex closeness on a
// 174k-vertex R-MAT 44.8 -> 43.9 ms reduced, 77.4 -> 76.1 ms plain; sampled
// closeness on R-MAT 2^20 from 8192 sources (W = 32) 16.5-17.3 -> 13.9-14.3
// ms, 16384 (W = 64) 26.7-28.5 -> 23.6-23.8 ms, 2048 and 4096 the same.
template <int W>
__global__ __launch_bounds__(ccw_block(W)) __attribute__((amdgpu_waves_per_eu(8))) void k_ccw(CcwArgs a) {
    if (a.live && *a.live == 0) return;  // uniform over the grid
    __shared__ long long s_far[W][kWave];
    __shared__ unsigned long long s_new;
    ccw_block_init<W>(s_far, &s_new);
    constexpr int WPB = ccw_block(W) / kWave;
    const int lane = lane_id();
    const int bl = lane & (W - 1);
    const uint64_t Ab = ccw_active(a, bl);
    const int gw = blockIdx.x * WPB + (threadIdx.x >> 6);
    const int nw = gridDim.x * WPB;
    unsigned long long nn = 0;
    CcwCount cnt;
    for (int c = gw; c < a.nchunk; c += nw) {
        const int base = a.chunk_beg[c];
        const int vl = base + lane;
        bool need = false;
        int rbl = 0, rel = 0;
        if (vl < a.chunk_beg[c + 1]) {
            rbl = a.rp[vl];
            rel = a.rp[vl + 1];
            if (rel - rbl <= a.hub_deg) {  // hub rows: k_ccw_hubseg
                const uint64_t* vv = a.vis + (size_t)vl * W;
#pragma unroll
                for (int b = 0; b < W; ++b) need |= (ccw_active(a, b) & ~vv[b]) != 0;
            }
            // Fn is not cleared between levels: vertices without a scan (and
            // hub rows, which k_ccw_hubfin may overwrite) get zero masks here
            if (!need) {
                uint64_t* fn = a.Fn + (size_t)vl * W;
#pragma unroll
                for (int b = 0; b < W; ++b) fn[b] = 0;
            }
        }
        uint64_t work = ballot(need);
        while (work) {
            const int j = __builtin_ctzll(work);
            work &= work - 1;
            const int v = base + j;
            const uint64_t old = a.vis[(size_t)v * W + bl];
            const uint64_t nm = ccw_scan<W>(a, rl32(rbl, j), rl32(rel, j), Ab & ~old, lane);
            nn += ccw_done<W>(a, v, old, nm, cnt, s_far, lane);
        }
    }
    ccw_count_flush<W>(cnt, s_far, lane);
    if (lane == 0 && nn) atomicAdd(&s_new, nn);
    ccw_flush<W>(a, s_far, &s_new, blockIdx.x);
}

// Hub rows: one wave per edge segment ORs the masks of its edges into
// part_or[segment][b]; k_ccw_hubfin ORs a hub's segments.
template <int W>
__global__ __launch_bounds__(kBlock) void k_ccw_hubseg(CcwArgs a) {
    if (a.live && *a.live == 0) return;
    const int lane = lane_id();
    const int bl = lane & (W - 1);
    const uint64_t Ab = ccw_active(a, bl);
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    for (int i = gw; i < a.nseg; i += nw) {
        const int v = a.seg_v[i];
        const uint64_t cand = Ab & ~a.vis[(size_t)v * W + bl];
        if (ballot(cand != 0) == 0) continue;  // k_ccw_hubfin skips the vertex too
        const uint64_t nm = ccw_scan<W, kCcwUnrollHub>(a, a.seg_beg[i], a.seg_end[i], cand, lane);
        if (lane < W) a.part_or[(size_t)i * W + bl] = nm;
    }
}

template <int W>
__global__ __launch_bounds__(ccw_block(W)) void k_ccw_hubfin(CcwArgs a) {
    if (a.live && *a.live == 0) return;
    __shared__ long long s_far[W][kWave];
    __shared__ unsigned long long s_new;
    ccw_block_init<W>(s_far, &s_new);
    constexpr int E = kWave / W;
    constexpr int WPB = ccw_block(W) / kWave;
    const int lane = lane_id();
    const int bl = lane & (W - 1), el = lane / W;
    const uint64_t Ab = ccw_active(a, bl);
    const int gw = blockIdx.x * WPB + (threadIdx.x >> 6);
    const int nw = gridDim.x * WPB;
    unsigned long long nn = 0;
    CcwCount cnt;
    for (int k = gw; k < a.nhub; k += nw) {
        const int v = a.hub_v[k];
        const uint64_t old = a.vis[(size_t)v * W + bl];
        if (ballot((Ab & ~old) != 0) == 0) continue;
        uint64_t orp = 0;
        for (int i = a.hub_seg_off[k] - a.seg_base + el; i < a.hub_seg_off[k + 1] - a.seg_base; i += E)
            orp |= a.part_or[(size_t)i * W + bl];
        const uint64_t nm = ccw_batch_or<W>(orp);
        nn += ccw_done<W>(a, v, old, nm, cnt, s_far, lane);
    }
    ccw_count_flush<W>(cnt, s_far, lane);
    if (lane == 0 && nn) atomicAdd(&s_new, nn);
    ccw_flush<W>(a, s_far, &s_new, blockIdx.x);
}

// Seeds the group: lane s of batch b = blockIdx.x starts at src[b * 64 + s].
// Only vis is seeded: level 1 is pushed from the source list, so no kernel
// reads the level-0 plane.
template <int W>
__global__ void k_ccw_seed(const int* __restrict__ src, int nb, int k_last, uint64_t* __restrict__ vis) {
    const int b = blockIdx.x, s = threadIdx.x;
    const int kb = b == nb - 1 ? k_last : kWave;
    if (s < kb) atomicOr((unsigned long long*)&vis[(size_t)src[b * kWave + s] * W + b], 1ull << s);
}

// Level 1 by push from the sources' rows (as k_push1_or / k_push1_fin for the
// betweenness step): batch b = blockIdx.y, its source rows flattened over
// blockIdx.x.  The lowest lane of a vertex's final mask finishes it.
template <int W>
__global__ __launch_bounds__(kBlock) void k_ccw_push_or(CcwArgs a, const int* __restrict__ src0) {
    __shared__ int s_pre[kWave + 1];
    __shared__ int s_beg[kWave];
    const int b = blockIdx.y;
    const int* src = src0 + (size_t)b * kWave;
    const int E = push_rows(ccw_active(a, b), a.rp, src, s_pre, s_beg);
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < E; e += gridDim.x * kBlock) {
        const int s = push_lane(s_pre, e);
        const int u = a.col[s_beg[s] + e - s_pre[s]];
        if ((a.vis[(size_t)u * W + b] >> s) & 1) continue;  // a self-loop of lane s's source
        atomicOr((unsigned long long*)&a.Fn[(size_t)u * W + b], 1ull << s);
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_ccw_push_fin(CcwArgs a, const int* __restrict__ src0) {
    __shared__ int s_pre[kWave + 1];
    __shared__ int s_beg[kWave];
    __shared__ unsigned long long s_far[kWave];
    __shared__ unsigned long long s_new;
    const int b = blockIdx.y;
    const int* src = src0 + (size_t)b * kWave;
    if (threadIdx.x < kWave) s_far[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_new = 0;
    const int E = push_rows(ccw_active(a, b), a.rp, src, s_pre, s_beg);  // synchronises the block
    unsigned long long nn = 0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < E; e += gridDim.x * kBlock) {
        const int s = push_lane(s_pre, e);
        const int u = a.col[s_beg[s] + e - s_pre[s]];
        const uint64_t nm = a.Fn[(size_t)u * W + b];
        if (!((nm >> s) & 1) || __builtin_ctzll(nm) != s) continue;  // finished by its lowest lane
        // ... once: a repeated edge gives that lane two threads for u, and the
        // first to set the new lanes in vis finishes it (vis held none of them)
        if (atomicOr((unsigned long long*)&a.vis[(size_t)u * W + b], (unsigned long long)nm) & nm) continue;
        const unsigned long long w = a.weight ? (unsigned long long)a.weight[u] : 1ull;
        for (uint64_t m = nm; m; m &= m - 1) atomicAdd(&s_far[__builtin_ctzll(m)], w);
        const int pc = __popcll(nm);
        if (a.dist_sum) atomicAdd((unsigned long long*)&a.dist_sum[u], (unsigned long long)pc);
        nn += (unsigned long long)pc;
    }
    if (nn) atomicAdd(&s_new, nn);
    __syncthreads();
    const int slot = blockIdx.x;
    if (threadIdx.x < kWave && s_far[threadIdx.x])
        a.part_far[((size_t)b * a.pslots + slot) * kWave + threadIdx.x] += (int64_t)s_far[threadIdx.x];
    if (threadIdx.x == 0 && s_new) {
        a.part_cnt[((size_t)b * a.pslots + slot) * 2] += s_new;
        *a.flag = 1;
    }
}

}  // namespace cent
