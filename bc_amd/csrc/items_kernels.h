// Persistent multi-source closeness over a queue of work items.
//
// An item is (packed id range [lo, hi), sources s0 .. s0+ns-1 inside it); a
// BFS from any of its sources never leaves the range (the range is a union
// of whole biconnected blocks of the packed block graph).  One workgroup
// takes one item at a time from a device counter and runs the whole
// level-synchronous 64-lane bit-parallel BFS for it with its frontier masks
// in LDS, so thousands of small blocks need no host round trip per level —
// the batched form of the per-source, per-level launches of the reference's
// GPU closeness (parallel_CC/para_CC.cu:175-364).
//
// Output: D[s] = sum_y d(s, y) * W(y) for every source s of every item
// (assigned; each packed vertex is the source of exactly one item).
#pragma once
#include "wave.h"

namespace cent {

constexpr int kItemLds = 2048;  // largest item range kept in LDS
constexpr int kItemBlock = 256;

struct ItemArgs {
    const int* __restrict__ rp;
    const int* __restrict__ col;
    const int64_t* __restrict__ weight;  // W per packed vertex
    const int4* __restrict__ items;      // (lo, hi, s0, ns)
    const int* __restrict__ order;       // item visiting order
    int nitems;
    int* __restrict__ head;              // queue counter (zeroed before launch)
    int64_t* __restrict__ D;             // [np]
};

__global__ __launch_bounds__(kItemBlock) void k_cc_items(ItemArgs a) {
    __shared__ uint64_t s_vis[kItemLds];
    __shared__ uint64_t s_F[2][kItemLds];
    __shared__ int64_t s_far[kItemBlock / kWave][kWave];
    __shared__ int s_item;
    __shared__ int s_flag;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wid = tid >> 6;
    constexpr int kW = kItemBlock / kWave;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(a.head, 1);
        __syncthreads();
        const int q = s_item;
        if (q >= a.nitems) break;
        const int4 it = a.items[a.order[q]];
        const int lo = it.x, R = it.y - it.x, s0 = it.z, ns = it.w;
        const uint64_t A = (ns >= 64) ? ~0ull : ((1ull << ns) - 1);
        for (int i = tid; i < R; i += kItemBlock) { s_vis[i] = 0; s_F[0][i] = 0; s_F[1][i] = 0; }
        __syncthreads();
        if (tid < ns) {
            s_F[0][s0 - lo + tid] = 1ull << tid;
            s_vis[s0 - lo + tid] = 1ull << tid;
        }
        int64_t far = 0;
        int cur = 0;
        for (int d = 0;; ++d) {
            if (tid == 0) s_flag = 0;
            __syncthreads();
            const uint64_t* Fc = s_F[cur];
            uint64_t* Fn = s_F[cur ^ 1];
            bool any = false;
            for (int base = wid * kWave; base < R; base += kW * kWave) {
                const int i = base + lane;
                uint64_t visl = A;
                int rb = 0, re = 0;
                if (i < R) {
                    visl = s_vis[i];
                    if (visl != A) { rb = a.rp[lo + i]; re = a.rp[lo + i + 1]; }
                }
                uint64_t work = ballot(visl != A);
                while (work) {
                    const int b = __builtin_ctzll(work);
                    work &= work - 1;
                    const uint64_t old = rl64(visl, b);
                    const uint64_t cand = A & ~old;
                    const int beg = rl32(rb, b), end = rl32(re, b);
                    uint64_t orf = 0;
                    for (int j = beg + lane; j < end; j += kWave) orf |= Fc[a.col[j] - lo] & cand;
                    const uint64_t nm = wave_or(orf);
                    if (nm) {
                        if ((nm >> lane) & 1) far += (int64_t)(d + 1) * a.weight[lo + base + b];
                        if (lane == 0) {
                            Fn[base + b] = nm;
                            s_vis[base + b] = old | nm;
                        }
                        any = true;
                    }
                }
            }
            if (any && lane == 0) s_flag = 1;
            __syncthreads();
            if (!s_flag) break;
            // the buffer read this level is cleared for the level after next
            for (int i = tid; i < R; i += kItemBlock) s_F[cur][i] = 0;
            cur ^= 1;
            __syncthreads();
        }
        s_far[wid][lane] = far;
        __syncthreads();
        if (wid == 0 && lane < ns) {
            int64_t t = 0;
#pragma unroll
            for (int w = 0; w < kW; ++w) t += s_far[w][lane];
            a.D[s0 + lane] = t;
        }
        __syncthreads();
    }
}

}  // namespace cent
