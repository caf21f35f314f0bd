// Frontier-queue (sparse) form of the 64-lane level kernels, for graphs whose
// BFS runs hundreds of levels with small frontiers (grids, road-like graphs),
// where a dense pass over all vertices per level would dominate.
//
// Per batch the engine keeps, besides vis / sigma / coef:
//   Q[]  / QF[]   concatenated per-level frontier lists: vertex and its mask
//                 (level l occupies Q[qoff[l] .. qoff[l+1]); each vertex
//                 appears at most once per level, so |Q| <= 64 n)
//   Fd[2][n]      dense masks of two adjacent levels (ping-pong)
//   stamp[n]      last expansion stamp, dedups the candidate list
// Every level runs three kernels — expand (frontier -> candidates), pull
// (candidates -> next frontier) and advance (one thread: level bookkeeping)
// — that read the level index and list sizes from device memory, so a run of
// levels can be captured once in a hipGraph and replayed; the host looks at
// the done flag once per replay.  Appends to the candidate and frontier lists
// are aggregated per wavefront (ballot, one atomicAdd, prefix popcount).
#pragma once
#include "msbfs_kernels.h"

namespace cent {

// ctl[] slots (device int array)
enum {
    kCtlLevel = 0,    // current frontier level d
    kCtlDone = 1,     // forward: no new frontier; backward: reached level 0
    kCtlNCand = 2,    // candidates appended this level
    kCtlNNext = 3,    // next-frontier entries appended this level
    kCtlStamp = 4,    // running expansion stamp (never reset)
    kCtlOverflow = 5, // path count became inf, or a list ran out of room
    kCtlDeepest = 6,  // deepest non-empty level (set when the forward ends)
    kCtlBwdLevel = 7, // backward: level being processed
    kCtlTooDeep = 8,  // the BFS needs more levels than the path-count exponent table
    kCtlSlots = 9
};

struct SpArgs {
    int n;
    const int* __restrict__ rp;
    const int* __restrict__ col;
    uint64_t* __restrict__ vis;
    uint64_t* __restrict__ Fd2;           // [2][n] dense masks, level parity
    double* __restrict__ sigma;
    double* __restrict__ coef;
    uint32_t* __restrict__ npred;
    double* __restrict__ bc;
    int64_t* __restrict__ dist_sum;
    const int64_t* __restrict__ weight;
    int* __restrict__ Q;
    uint64_t* __restrict__ QF;
    int* __restrict__ qoff;               // [levels + 2]
    int qcap;                             // capacity of Q / QF
    int qoff_cap;                         // capacity of qoff
    int* __restrict__ cand;               // [n]
    int* __restrict__ stamp;              // [n]
    int* __restrict__ ctl;                // [kCtlSlots]
    const uint64_t* __restrict__ Aptr;    // active lanes of the batch
    int64_t* __restrict__ part_far;       // [grid][64]
    unsigned long long* __restrict__ part_cnt;  // [grid][2]
    int* __restrict__ lexp;               // [lexp_levels][64] path-count exponents (betweenness)
    int lexp_levels;
};

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

// Wave-aggregated append of the lanes in `mask` to list[base_ctr...].
// Returns this lane's slot (valid only if its bit is set).
__device__ __forceinline__ int wave_append(uint64_t mask, int* ctr, int lane) {
    int base = 0;
    if (lane == 0 && mask) base = atomicAdd(ctr, __popcll(mask));
    base = __shfl(base, 0);
    return base + __popcll(mask & lanemask_lt(lane));
}

// Expand level d: clear the other dense array's entries of level d-1 (they
// become level d+1's masks), then put every not-fully-reached neighbour of
// the frontier on the candidate list once.
__global__ __launch_bounds__(kBlock) void k_sp_expand(SpArgs a) {
    if (a.ctl[kCtlDone]) return;
    const int d = a.ctl[kCtlLevel];
    const int lane = lane_id();
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nthr = gridDim.x * blockDim.x;
    const uint64_t A = *a.Aptr;
    uint64_t* Fn = a.Fd2 + (size_t)a.n * ((d + 1) & 1);
    if (d >= 1) {
        for (int i = a.qoff[d - 1] + gtid; i < a.qoff[d]; i += nthr) Fn[a.Q[i]] = 0;
    }
    const int S = a.ctl[kCtlStamp];
    const int beg = a.qoff[d], end = a.qoff[d + 1];
    // wave-uniform trip count so every lane reaches the ballots
    for (int i0 = beg + (gtid - lane); i0 < end; i0 += nthr) {
        const int i = i0 + lane;
        int rb = 0, re = 0;
        if (i < end) {
            const int u = a.Q[i];
            rb = a.rp[u];
            re = a.rp[u + 1];
        }
        int len = re - rb;
        int maxlen = len;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o));
        for (int t = 0; t < maxlen; ++t) {
            bool want = false;
            int v = 0;
            if (t < len) {
                v = a.col[rb + t];
                if (a.vis[v] != A && a.stamp[v] != S) want = (atomicExch(&a.stamp[v], S) != S);
            }
            const uint64_t m = ballot(want);
            if (!m) continue;
            const int slot = wave_append(m, &a.ctl[kCtlNCand], lane);
            if (want) a.cand[slot] = v;
        }
    }
}

// Pull for the candidates of level d: the dense kernel's per-vertex step over
// an explicit vertex list; new frontier vertices are appended to Q / QF.
template <bool SIGMA, bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_sp_pull(SpArgs a) {
    if (a.ctl[kCtlDone]) return;
    const int d = a.ctl[kCtlLevel];
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const uint64_t A = *a.Aptr;
    const uint64_t* Fd = a.Fd2 + (size_t)a.n * (d & 1);
    uint64_t* Fn = a.Fd2 + (size_t)a.n * ((d + 1) & 1);
    const int nc = a.ctl[kCtlNCand];
    const int qbase = a.qoff[d + 1];
    FwdArgs f{};  // the shared per-vertex epilogue works on FwdArgs
    f.n = a.n; f.rp = a.rp; f.col = a.col; f.vis = a.vis; f.Fd = Fd; f.Fn = Fn; f.sigma = a.sigma;
    f.npred = a.npred; f.dist_sum = a.dist_sum; f.weight = a.weight; f.A = A; f.dnext = d + 1;
    f.lexp = SIGMA ? a.lexp : nullptr;
    const double sc = SIGMA ? level_scale(a.lexp, d, lane) : 1.0;
    FwdAccum r;
    for (int base = gw * kWave; base < nc; base += nw * kWave) {
        const int i = base + lane;
        int v = 0;
        uint64_t visl = A;
        int rb = 0, re = 0;
        if (i < nc) {
            v = a.cand[i];
            visl = a.vis[v];
            if (visl != A) { rb = a.rp[v]; re = a.rp[v + 1]; }
        }
        uint64_t work = ballot(visl != A);
        uint64_t newbits = 0;
        uint64_t my_nm = 0;
        while (work) {
            const int b = __builtin_ctzll(work);
            work &= work - 1;
            const int vb = rl32(v, b);
            const uint64_t old = rl64(visl, b);
            double acc = 0.0;
            uint32_t cnt = 0;
            uint64_t orf = 0;
            gather_fwd<SIGMA, false>(a.col, Fd, a.sigma, rl32(rb, b), rl32(re, b), A & ~old, lane, acc, cnt, orf);
            uint64_t nm;
            if constexpr (SIGMA) nm = ballot(cnt != 0);
            else nm = wave_or(orf);
            fwd_vertex_done<SIGMA, COMPAT, false>(f, vb, old, acc, cnt, nm, lane, r, sc);
            if (nm) {
                newbits |= 1ull << b;
                if (lane == b) my_nm = nm;
            }
        }
        if (newbits) {
            const int slot = wave_append(newbits, &a.ctl[kCtlNNext], lane);
            if ((newbits >> lane) & 1) {
                const int q = qbase + slot;
                if (q < a.qcap) { a.Q[q] = v; a.QF[q] = my_nm; }
                else a.ctl[kCtlOverflow] = 1;
            }
        }
    }
    if (r.ovf) a.ctl[kCtlOverflow] = 1;
    // per-block lane partials (farness, reached pairs, dag edges)
    __shared__ int64_t s_far[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_cnt[kWavesPerBlock][2];
    const int wid = threadIdx.x >> 6;
    s_far[wid][lane] = r.far;
    const unsigned long long dag = (unsigned long long)wave_sum_i64((int64_t)r.dag);
    const unsigned long long nn = (unsigned long long)wave_sum_i64((int64_t)r.nnew);
    if (lane == 0) { s_cnt[wid][0] = nn; s_cnt[wid][1] = dag; }
    __syncthreads();
    if (wid == 0) {
        int64_t t = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) t += s_far[w][lane];
        if (t) a.part_far[(size_t)blockIdx.x * kWave + lane] += t;
        if (lane < 2) {
            unsigned long long c = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) c += s_cnt[w][lane];
            if (c) a.part_cnt[(size_t)blockIdx.x * 2 + lane] += c;
        }
    }
}

// One thread: close level d.
__global__ void k_sp_advance(SpArgs a) {
    if (a.ctl[kCtlDone]) return;
    const int d = a.ctl[kCtlLevel];
    const int cnt = a.ctl[kCtlNNext];
    int nq = a.qoff[d + 1] + cnt;
    if (nq > a.qcap) { a.ctl[kCtlOverflow] = 1; nq = a.qcap; }
    a.ctl[kCtlNCand] = 0;
    a.ctl[kCtlNNext] = 0;
    a.ctl[kCtlStamp] += 1;
    const bool deep = a.lexp && d + 2 >= a.lexp_levels;
    if (cnt == 0 || a.ctl[kCtlOverflow] || d + 3 >= a.qoff_cap || deep) {
        if (cnt != 0) a.ctl[deep ? kCtlTooDeep : kCtlOverflow] = 1;
        a.ctl[kCtlDone] = 1;
        a.ctl[kCtlDeepest] = d;
        a.qoff[d + 2] = a.qoff[d + 1];  // empty level d+1 (qoff_cap >= d + 3 always holds here)
        return;
    }
    a.qoff[d + 2] = nq;
    a.ctl[kCtlLevel] = d + 1;
}

// Backward level d = ctl[kCtlBwdLevel] (>= 1) over Q_d: the dense kernel's
// successor pull, reading level d+1's masks from Fd2[(d+1) & 1].  The same
// launch clears level d+2's entries from Fd2[d & 1].
template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_sp_bwd(SpArgs a) {
    const int d = a.ctl[kCtlBwdLevel];
    if (d < 1) return;
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const uint64_t* Fn = a.Fd2 + (size_t)a.n * ((d + 1) & 1);
    uint64_t* Fo = a.Fd2 + (size_t)a.n * (d & 1);
    BwdArgs b{};
    b.sigma = a.sigma; b.coef = a.coef; b.npred = a.npred; b.bc = a.bc;
    const double sc = level_scale(a.lexp, d, lane);
    const int beg = a.qoff[d], end = a.qoff[d + 1];
    for (int base = beg + gw * kWave; base < end; base += nw * kWave) {
        const int i = base + lane;
        int v = 0, rb = 0, re = 0;
        uint64_t fl = 0;
        if (i < end) {
            v = a.Q[i];
            fl = a.QF[i];
            rb = a.rp[v];
            re = a.rp[v + 1];
        }
        uint64_t work = ballot(fl != 0);
        while (work) {
            const int bb = __builtin_ctzll(work);
            work &= work - 1;
            const uint64_t m = rl64(fl, bb);
            const double acc = gather_bwd<false>(a.col, Fn, a.coef, rl32(rb, bb), rl32(re, bb), m, lane);
            bwd_vertex_done<false>(b, rl32(v, bb), m, 0, acc, lane, COMPAT, sc);
        }
    }
    if (d + 2 <= a.ctl[kCtlDeepest] + 1) {
        const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
        const int nthr = gridDim.x * blockDim.x;
        for (int i = a.qoff[d + 2] + gtid; i < a.qoff[d + 3]; i += nthr) Fo[a.Q[i]] = 0;
    }
}

// Starts the backward at the deepest level.
__global__ void k_sp_bwd_init(SpArgs a) { a.ctl[kCtlBwdLevel] = a.ctl[kCtlDeepest]; }

// Backward bookkeeping: write level d's masks into Fd2[d & 1] (they are
// level d-1's successors), then step to d-1.
__global__ __launch_bounds__(kBlock) void k_sp_bwd_next(SpArgs a) {
    const int d = a.ctl[kCtlBwdLevel];
    if (d < 1) return;
    uint64_t* Fo = a.Fd2 + (size_t)a.n * (d & 1);
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nthr = gridDim.x * blockDim.x;
    for (int i = a.qoff[d] + gtid; i < a.qoff[d + 1]; i += nthr) Fo[a.Q[i]] = a.QF[i];
}

__global__ void k_sp_bwd_step(SpArgs a) {
    const int d = a.ctl[kCtlBwdLevel];
    if (d < 1) return;
    a.ctl[kCtlBwdLevel] = d - 1;
}

// Seeds a sparse batch: Q_0 = the distinct sources with their lane masks.
// One wavefront; lane s holds source s.
__global__ void k_sp_seed(SpArgs a, const int* __restrict__ src, int k, bool with_sigma, uint64_t A) {
    const int lane = threadIdx.x;
    if (lane == 0) *const_cast<uint64_t*>(a.Aptr) = A;
    const int s = (lane < k) ? src[lane] : -1;
    if (lane < k) {
        atomicOr((unsigned long long*)&a.vis[s], 1ull << lane);
        atomicOr((unsigned long long*)&a.Fd2[s], 1ull << lane);
        if (with_sigma) a.sigma[(size_t)s * kWave + lane] = 1.0;
    }
    __syncthreads();
    // the first lane holding each distinct source appends it (full-wave shuffles)
    bool first = lane < k;
    for (int o = 0; o < kWave; ++o) {
        const int so = __shfl(s, o);
        if (o < lane && so == s) first = false;
    }
    const uint64_t m = ballot(first);
    const int slot = __popcll(m & lanemask_lt(lane));
    if (first) {
        a.Q[slot] = s;
        a.QF[slot] = a.Fd2[s];
    }
    if (lane == 0) {
        a.qoff[0] = 0;
        a.qoff[1] = __popcll(m);
        a.ctl[kCtlLevel] = 0;
        a.ctl[kCtlDone] = 0;
        a.ctl[kCtlNCand] = 0;
        a.ctl[kCtlNNext] = 0;
        a.ctl[kCtlOverflow] = 0;
        a.ctl[kCtlDeepest] = 0;
        a.ctl[kCtlBwdLevel] = 0;
        a.ctl[kCtlTooDeep] = 0;
    }
}

// End of batch: clear vis and both dense mask arrays for every vertex the
// batch reached (all of them are on some frontier list), and the path-count
// exponents of its levels.
__global__ __launch_bounds__(kBlock) void k_sp_clear(SpArgs a) {
    const int end = a.qoff[a.ctl[kCtlDeepest] + 1];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nthr = gridDim.x * blockDim.x;
    if (a.lexp) {
        const int ne = min(a.ctl[kCtlDeepest] + 2, a.lexp_levels) * kWave;
        for (int i = gtid; i < ne; i += nthr) a.lexp[i] = 0;
    }
    for (int i = gtid; i < end; i += nthr) {
        const int v = a.Q[i];
        a.vis[v] = 0;
        a.Fd2[v] = 0;
        a.Fd2[(size_t)a.n + v] = 0;
    }
}

}  // namespace cent
