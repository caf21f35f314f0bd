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
// Threads per item workgroup: 16 waves, each walking the item's vertices one
// at a time (measured: 4 waves took 2.2x (closeness) and 2.5x (BC) as long).
constexpr int kBcItemBlock = 1024;
constexpr int kCcItemBlock = 1024;

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

__global__ __launch_bounds__(kCcItemBlock) void k_cc_items(ItemArgs a) {
    __shared__ uint64_t s_vis[kItemLds];
    __shared__ uint64_t s_F[2][kItemLds];
    __shared__ int64_t s_far[kCcItemBlock / kWave][kWave];
    __shared__ int s_item;
    __shared__ int s_flag;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wid = tid >> 6;
    constexpr int kW = kCcItemBlock / kWave;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(a.head, 1);
        __syncthreads();
        const int q = s_item;
        if (q >= a.nitems) break;
        const int4 it = a.items[a.order[q]];
        const int lo = it.x, R = it.y - it.x, s0 = it.z, ns = it.w;
        const uint64_t A = (ns >= 64) ? ~0ull : ((1ull << ns) - 1);
        for (int i = tid; i < R; i += kCcItemBlock) { s_vis[i] = 0; s_F[0][i] = 0; s_F[1][i] = 0; }
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
            for (int i = tid; i < R; i += kCcItemBlock) s_F[cur][i] = 0;
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

namespace cent {

// Persistent weighted Brandes over the same work items (reduced graphs).
// Frontier masks stay in LDS as in k_cc_items; the per-(vertex, lane) path
// counts, coefficients and levels of the item's range live in a per-
// workgroup slice of HBM scratch (64 entries a vertex of the largest item).  Outputs:
//   D[s]                      weighted farness of every source (assigned)
//   bcpart[poff[item] + i]    the item's weighted dependencies at vertex lo+i
//                             (added across the item's levels; the host sums
//                             the items of a block in item order)
// An item whose BFS gets deeper than 254 levels is abandoned and listed in
// deep[] for the host to run on the level-synchronous kernels.
struct BcItemArgs {
    const int* __restrict__ rp;
    const int* __restrict__ col;
    const int64_t* __restrict__ weight;
    const int4* __restrict__ items;
    const int* __restrict__ order;
    const int64_t* __restrict__ poff;     // per item index
    int nitems;
    int* __restrict__ head;
    int64_t* __restrict__ D;
    double* __restrict__ bcpart;
    double* __restrict__ sig;             // [grid][stride]
    double* __restrict__ coef;            // [grid][stride]
    uint8_t* __restrict__ lvl;            // [grid][stride], 255 = unreached
    size_t stride;                        // entries a workgroup: 64 x the largest item range
    int* __restrict__ deep;               // [nitems] item indices
    int* __restrict__ ndeep;
    int* __restrict__ overflow;           // set when a path count is inf
    // sampled sources: item sources are srcs[s0 .. s0 + ns) (packed ids)
    // instead of s0 .. s0 + ns - 1, each with source weight swgt[packed id]
    // (null: weight); D and the partials stay indexed by list position
    const int* __restrict__ srcs;
    const int64_t* __restrict__ swgt;
};

constexpr int kItemMaxDepth = 254;

__global__ __launch_bounds__(kBcItemBlock) void k_bc_items(BcItemArgs a) {
    __shared__ uint64_t s_vis[kItemLds];
    __shared__ uint64_t s_F[2][kItemLds];
    __shared__ int64_t s_far[kBcItemBlock / kWave][kWave];
    __shared__ int s_item;
    __shared__ int s_flag;
    __shared__ int s_deep;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wid = tid >> 6;
    constexpr int kW = kBcItemBlock / kWave;
    double* __restrict__ sig = a.sig + blockIdx.x * a.stride;
    double* __restrict__ coef = a.coef + blockIdx.x * a.stride;
    uint8_t* __restrict__ lvl = a.lvl + blockIdx.x * a.stride;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(a.head, 1);
        __syncthreads();
        const int q = s_item;
        if (q >= a.nitems) break;
        const int itx = a.order[q];
        const int4 it = a.items[itx];
        const int lo = it.x, R = it.y - it.x, s0 = it.z, ns = it.w;
        const uint64_t A = (ns >= 64) ? ~0ull : ((1ull << ns) - 1);
        double* __restrict__ part = a.bcpart + a.poff[itx];
        for (int i = tid; i < R; i += kBcItemBlock) { s_vis[i] = 0; s_F[0][i] = 0; s_F[1][i] = 0; part[i] = 0.0; }
        for (int i = tid; i < R * (kWave / 8); i += kBcItemBlock) reinterpret_cast<uint64_t*>(lvl)[i] = ~0ull;
        if (tid == 0) s_deep = 0;
        __syncthreads();
        // this lane's source (packed id) and its weight as a source
        const int my_src = lane < ns ? (a.srcs ? a.srcs[s0 + lane] : s0 + lane) : lo;
        const double my_sw = lane < ns ? (double)(a.swgt ? a.swgt[my_src] : a.weight[my_src]) : 0.0;
        if (tid < ns) {
            const int i = my_src - lo;
            atomicOr((unsigned long long*)&s_F[0][i], 1ull << tid);  // repeated sources share a vertex
            atomicOr((unsigned long long*)&s_vis[i], 1ull << tid);
            sig[(size_t)i * kWave + tid] = 1.0;
            lvl[(size_t)i * kWave + tid] = 0;
        }
        int64_t far = 0;
        int cur = 0, depth = 0;
        // ---- forward
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
                    double acc = 0.0;
                    uint32_t cnt = 0;
                    for (int j0 = beg; j0 < end; j0 += kWave) {
                        const int j = j0 + lane;
                        int u = 0;
                        uint64_t f = 0;
                        if (j < end) {
                            u = a.col[j] - lo;
                            f = Fc[u] & cand;
                        }
                        uint64_t ball = ballot(f != 0);
                        while (ball) {
                            const int c = __builtin_ctzll(ball);
                            ball &= ball - 1;
                            const int uc = rl32(u, c);
                            const uint64_t fc = rl64(f, c);
                            if ((fc >> lane) & 1) {
                                acc += sig[(size_t)uc * kWave + lane];
                                ++cnt;
                            }
                        }
                    }
                    const uint64_t nm = ballot(cnt != 0);
                    if (nm) {
                        const int vi = base + b;
                        if ((nm >> lane) & 1) {
                            sig[(size_t)vi * kWave + lane] = acc;
                            if (__builtin_isinf(acc)) *a.overflow = 1;
                            lvl[(size_t)vi * kWave + lane] = (uint8_t)(d + 1);
                            far += (int64_t)(d + 1) * a.weight[lo + vi];
                        }
                        if (lane == 0) {
                            Fn[vi] = nm;
                            s_vis[vi] = old | nm;
                        }
                        any = true;
                    }
                }
            }
            if (any && lane == 0) s_flag = 1;
            __syncthreads();
            if (!s_flag) break;
            depth = d + 1;
            if (depth >= kItemMaxDepth) {
                if (tid == 0) s_deep = 1;
                __syncthreads();
                break;
            }
            for (int i = tid; i < R; i += kBcItemBlock) s_F[cur][i] = 0;
            cur ^= 1;
            __syncthreads();
        }
        if (s_deep) {
            if (tid == 0) a.deep[atomicAdd(a.ndeep, 1)] = itx;
            __syncthreads();
            continue;
        }
        // Fc = s_F[cur] holds level `depth`, the other buffer is all zero (level depth + 1)
        // ---- backward, level d uses masks of d (Fd) and d + 1 (Fu)
        for (int d = depth; d >= 1; --d) {
            const uint64_t* Fd = s_F[cur];
            const uint64_t* Fu = s_F[cur ^ 1];
            for (int base = wid * kWave; base < R; base += kW * kWave) {
                const int i = base + lane;
                uint64_t fl = 0;
                int rb = 0, re = 0;
                if (i < R) {
                    fl = Fd[i];
                    if (fl) { rb = a.rp[lo + i]; re = a.rp[lo + i + 1]; }
                }
                uint64_t work = ballot(fl != 0);
                while (work) {
                    const int b = __builtin_ctzll(work);
                    work &= work - 1;
                    const uint64_t m = rl64(fl, b);
                    const int beg = rl32(rb, b), end = rl32(re, b);
                    double acc = 0.0;
                    for (int j0 = beg; j0 < end; j0 += kWave) {
                        const int j = j0 + lane;
                        int w = 0;
                        uint64_t g = 0;
                        if (j < end) {
                            w = a.col[j] - lo;
                            g = Fu[w] & m;
                        }
                        uint64_t ball = ballot(g != 0);
                        while (ball) {
                            const int c = __builtin_ctzll(ball);
                            ball &= ball - 1;
                            const int wc = rl32(w, c);
                            const uint64_t gc = rl64(g, c);
                            if ((gc >> lane) & 1) acc += coef[(size_t)wc * kWave + lane];
                        }
                    }
                    const int vi = base + b;
                    double contrib = 0.0;
                    if ((m >> lane) & 1) {
                        const size_t idx = (size_t)vi * kWave + lane;
                        const double sg = sig[idx];
                        const double dl = sg * acc;
                        coef[idx] = ((double)a.weight[lo + vi] + dl) / sg;
                        contrib = dl * my_sw;
                    }
                    const double tot = wave_sum(contrib);
                    if (lane == 0) part[vi] += tot;
                }
            }
            __syncthreads();
            // the buffer of level d + 1 becomes level d - 1, rebuilt from lvl
            uint64_t* Fnew = s_F[cur ^ 1];
            for (int base = wid * kWave; base < R; base += kW * kWave) {
                for (int b = 0; b < kWave && base + b < R; ++b) {
                    const uint64_t mk = ballot(lane < ns && lvl[(size_t)(base + b) * kWave + lane] == (uint8_t)(d - 1));
                    if (lane == 0) Fnew[base + b] = mk;
                }
            }
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
