// Multi-source level kernels: one wavefront carries 64 BFS sources, lane s is
// source s of the batch.  Per vertex v the engine keeps
//   vis[v]        uint64  lanes that have reached v
//   F[d][v]       uint64  lanes that reach v at distance d (one mask per level)
//   sigma[v][64]  double  shortest-path counts, lane-contiguous (512 B rows)
//   coef[v][64]   double  (1 + delta) / sigma, written by the backward pass
// so a neighbour's frontier test is one 8-byte load for all 64 sources and a
// path-count gather is one coalesced 512-byte row read.
//
// Forward (pull): every vertex with an unreached lane scans its neighbours u,
// keeps f = F[d][u] & ~vis[v]; lane s sums sigma[u][s] over the neighbours
// whose f has bit s.  Lanes that summed anything are the new frontier.  This
// is the level step of the reference's queue BFS (Seq_Brandes/Brandes.c:274-298,
// Seq_CC/CC.cpp:81-98), run for 64 sources at once.
//
// Backward (pull over successors): a vertex at level d in lane s sums
// coef[w][s] over neighbours w at level d+1 in lane s, then
// delta = sigma * sum (Brandes' dependency, Brandes.c:317-327 in textbook
// form) and coef = (1 + delta) / sigma for the level above.  The BC of v gets
// the lane sum of delta (or of delta * #predecessors in the reference's
// per-predecessor formula), one deterministic butterfly reduction per level.
#pragma once
#include "wave.h"

namespace cent {

constexpr int kBlock = 256;               // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

struct FwdArgs {
    int n;
    const int* __restrict__ rp;
    const int* __restrict__ col;
    uint64_t* __restrict__ vis;
    const uint64_t* __restrict__ Fd;      // frontier at distance d
    uint64_t* __restrict__ Fn;            // frontier at distance d + 1 (zeroed)
    double* __restrict__ sigma;           // [n][64]
    uint32_t* __restrict__ npred;         // [n][64], reference-compat BC only
    int64_t* __restrict__ dist_sum;       // [n] or null
    const int64_t* __restrict__ weight;   // [n] vertex weights or null (== 1)
    uint64_t A;                           // active lanes of the batch
    int dnext;                            // d + 1
    int64_t* __restrict__ part_far;       // [grid][64] per-block lane partials
    unsigned long long* __restrict__ part_cnt;  // [grid][2]: reached pairs, dag edges
    int* __restrict__ flag;               // set to 1 when anything was reached
    int* __restrict__ overflow;           // set to 1 when a path count is inf
};

template <bool SIGMA, bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_fwd(FwdArgs a) {
    __shared__ int64_t s_far[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_cnt[kWavesPerBlock][2];
    const int lane = lane_id();
    const int wid = threadIdx.x >> 6;
    const int gw = blockIdx.x * kWavesPerBlock + wid;
    const int nw = gridDim.x * kWavesPerBlock;
    const uint64_t A = a.A;
    int64_t my_far = 0;
    unsigned long long my_new = 0, my_dag = 0;
    bool ovf = false;

    for (int base = gw * kWave; base < a.n; base += nw * kWave) {
        const int vl = base + lane;
        uint64_t visl = A;
        if (vl < a.n) visl = a.vis[vl];
        uint64_t work = ballot(visl != A);
        if (!work) continue;
        int rbl = 0, rel = 0;
        if (vl < a.n && visl != A) { rbl = a.rp[vl]; rel = a.rp[vl + 1]; }
        while (work) {
            const int b = __builtin_ctzll(work);
            work &= work - 1;
            const int v = base + b;
            const uint64_t old = rl64(visl, b);
            const uint64_t cand = A & ~old;
            const int beg = rl32(rbl, b), end = rl32(rel, b);
            double acc = 0.0;
            uint32_t cnt = 0;
            uint64_t orf = 0;
            for (int j0 = beg; j0 < end; j0 += kWave) {
                const int j = j0 + lane;
                int u = 0;
                uint64_t f = 0;
                if (j < end) {
                    u = a.col[j];
                    f = a.Fd[u] & cand;
                }
                if constexpr (SIGMA) {
                    uint64_t ball = ballot(f != 0);
                    while (ball) {
                        const int c = __builtin_ctzll(ball);
                        ball &= ball - 1;
                        const int uc = rl32(u, c);
                        const uint64_t fc = rl64(f, c);
                        if ((fc >> lane) & 1) {
                            acc += a.sigma[(size_t)uc * kWave + lane];
                            ++cnt;
                        }
                    }
                } else {
                    orf |= f;
                }
            }
            uint64_t nm;
            if constexpr (SIGMA) nm = ballot(cnt != 0);
            else nm = wave_or(orf);
            if (nm) {
                if ((nm >> lane) & 1) {
                    if constexpr (SIGMA) {
                        a.sigma[(size_t)v * kWave + lane] = acc;
                        if (__builtin_isinf(acc)) ovf = true;
                    }
                    if constexpr (COMPAT) a.npred[(size_t)v * kWave + lane] = cnt;
                    my_far += a.weight ? (int64_t)a.dnext * a.weight[v] : (int64_t)a.dnext;
                    my_dag += cnt;
                }
                if (lane == 0) {
                    a.Fn[v] = nm;
                    a.vis[v] = old | nm;
                    const int pc = __popcll(nm);
                    if (a.dist_sum) a.dist_sum[v] += (int64_t)a.dnext * pc;
                    my_new += (unsigned long long)pc;
                }
            }
        }
    }
    // block-level reduction of the lane partials, one plain store per block
    s_far[wid][lane] = my_far;
    const unsigned long long dag = (unsigned long long)wave_sum_i64((int64_t)my_dag);
    if (lane == 0) { s_cnt[wid][0] = my_new; s_cnt[wid][1] = dag; }
    if (ovf) *a.overflow = 1;
    __syncthreads();
    if (wid == 0) {
        int64_t f = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) f += s_far[w][lane];
        if (f) a.part_far[(size_t)blockIdx.x * kWave + lane] += f;
        if (lane < 2) {
            unsigned long long c = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) c += s_cnt[w][lane];
            if (c) {
                a.part_cnt[(size_t)blockIdx.x * 2 + lane] += c;
                if (lane == 0) *a.flag = 1;
            }
        }
    }
}

struct BwdArgs {
    int n;
    const int* __restrict__ rp;
    const int* __restrict__ col;
    const uint64_t* __restrict__ Fd;      // level d masks
    const uint64_t* __restrict__ Fn;      // level d + 1 masks
    const double* __restrict__ sigma;
    double* __restrict__ coef;
    const uint32_t* __restrict__ npred;
    double* __restrict__ bc;
};

template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_bwd(BwdArgs a) {
    const int lane = lane_id();
    const int wid = threadIdx.x >> 6;
    const int gw = blockIdx.x * kWavesPerBlock + wid;
    const int nw = gridDim.x * kWavesPerBlock;
    for (int base = gw * kWave; base < a.n; base += nw * kWave) {
        const int vl = base + lane;
        uint64_t fl = 0;
        if (vl < a.n) fl = a.Fd[vl];
        uint64_t work = ballot(fl != 0);
        if (!work) continue;
        int rbl = 0, rel = 0;
        if (fl) { rbl = a.rp[vl]; rel = a.rp[vl + 1]; }
        while (work) {
            const int b = __builtin_ctzll(work);
            work &= work - 1;
            const int v = base + b;
            const uint64_t m = rl64(fl, b);
            const int beg = rl32(rbl, b), end = rl32(rel, b);
            double acc = 0.0;
            for (int j0 = beg; j0 < end; j0 += kWave) {
                const int j = j0 + lane;
                int w = 0;
                uint64_t g = 0;
                if (j < end) {
                    w = a.col[j];
                    g = a.Fn[w] & m;
                }
                uint64_t ball = ballot(g != 0);
                while (ball) {
                    const int c = __builtin_ctzll(ball);
                    ball &= ball - 1;
                    const int wc = rl32(w, c);
                    const uint64_t gc = rl64(g, c);
                    if ((gc >> lane) & 1) acc += a.coef[(size_t)wc * kWave + lane];
                }
            }
            double contrib = 0.0;
            if ((m >> lane) & 1) {
                const size_t idx = (size_t)v * kWave + lane;
                const double sg = a.sigma[idx];
                const double dl = sg * acc;
                a.coef[idx] = (1.0 + dl) / sg;
                contrib = COMPAT ? dl * (double)a.npred[idx] : dl;
            }
            const double tot = wave_sum(contrib);
            if (lane == 0) a.bc[v] += tot;
        }
    }
}

// Seeds a batch: lane s starts at src[s] (duplicates allowed, each its own lane).
__global__ void k_seed(const int* __restrict__ src, int k, uint64_t* __restrict__ F0,
                       uint64_t* __restrict__ vis, double* __restrict__ sigma, bool with_sigma) {
    const int lane = threadIdx.x;
    if (lane >= k) return;
    const int s = src[lane];
    atomicOr((unsigned long long*)&F0[s], 1ull << lane);
    atomicOr((unsigned long long*)&vis[s], 1ull << lane);
    if (with_sigma) sigma[(size_t)s * kWave + lane] = 1.0;
}

// Sums the per-block lane partials of a batch into far[0..k) and the counters.
__global__ void k_reduce_batch(const int64_t* __restrict__ part_far,
                               const unsigned long long* __restrict__ part_cnt, int nblocks, int k,
                               int64_t* __restrict__ far, unsigned long long* __restrict__ cnt_out) {
    const int lane = threadIdx.x;  // 64 threads
    int64_t f = 0;
    for (int b = 0; b < nblocks; ++b) f += part_far[(size_t)b * kWave + lane];
    if (lane < k) far[lane] = f;
    if (lane < 2) {
        unsigned long long c = 0;
        for (int b = 0; b < nblocks; ++b) c += part_cnt[(size_t)b * 2 + lane];
        cnt_out[lane] += c;
    }
}

}  // namespace cent
