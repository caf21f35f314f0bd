// Multi-source level kernels of the dense betweenness step: one wavefront
// carries 64 BFS sources, lane s is source s of the batch.  Per vertex v the
// engine keeps
//   vis[v]        uint64  lanes that have reached v
//   F[d][v]       {mask, base} record: lanes that reach v at distance d and
//                 the number of lanes that reached it earlier (one plane per
//                 level, kept for the backward pass)
//   sigma[v][64]  double  shortest-path counts, level-sorted (below); the
//                         backward writes coef = (1 + delta) / sigma over
//                         each entry in place
// so a neighbour's frontier test is one 16-byte load for all 64 sources and a
// path-count gather reads one contiguous slice of a 512-byte row.  The
// closeness step (masks only, up to 64 batches interleaved) is in
// ccw_kernels.h; the frontier-queue step reuses the gathers and the vertex
// epilogue below with plain masks and lane-indexed rows.
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
//
// Path-count scaling: counts grow exponentially with depth on some graphs
// (a chain of k diamonds has 2^k shortest paths) and would leave the FP64
// range.  Each level's stored counts carry a per-(level, lane) power-of-two
// scale: when a new count reaches 2^512 its exponent goes into lexp[d][lane]
// (atomic max); the next level multiplies its gathered sums by
// 2^(256 - lexp[d][lane]) and the backward multiplies the level-d
// dependencies by the same factor, because
//   sigma~_{d+1} = sum sigma~_d * 2^-k_d   and   delta_d = sigma~_d * 2^-k_d * sum coef~_{d+1}
// with coef~ = (1 + delta) / sigma~ computed from the stored counts.
// Multiplying by a power of two is exact, and the factor is 1.0 unless a
// level reached 2^512, so results are bit-identical to unscaled runs
// wherever those stay finite.
//
// Load balance: a wavefront takes one vertex chunk at a time — at most 64
// consecutive vertices whose rows hold at most a fixed edge budget — and
// rows longer than the hub threshold are cut into fixed edge
// segments, each scanned by its own wavefront (k_*_hubseg) into a per-segment
// lane partial; k_*_hubfin then adds a hub's partials in segment order, so
// the sums do not depend on scheduling.  The main kernels skip those rows.
#pragma once
#include "wave.h"

namespace cent {

constexpr int kBlock = 256;               // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kUnroll = 8;                // row gathers in flight per wave
constexpr int kUnrollHub = 8;             // the same in the hub-segment kernels

struct HubArgs {
    int nseg;                       // number of hub edge segments
    const int* __restrict__ seg_v;  // [nseg] owning vertex
    const int* __restrict__ seg_beg;
    const int* __restrict__ seg_end;
    int seg_base;                         // hub_seg_off value of segment 0 of this view
    int nhub;
    const int* __restrict__ hub_v;        // [nhub]
    const int* __restrict__ hub_seg_off;  // [nhub + 1]
    double* __restrict__ part_acc;        // [nseg][64]
    uint32_t* __restrict__ part_cnt;      // [nseg][64]
};

struct FwdArgs {
    int n;
    int hub_deg;                          // rows longer than this are hub rows
    int nchunk;                           // vertex chunks: <= 64 vertices, bounded edges
    const int* __restrict__ chunk_beg;    // [nchunk + 1]
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
    // Levels are enqueued ahead of the host's check: a launch whose level
    // is empty (*live == 0, written by the previous level) returns at once.
    const int* __restrict__ live;
    // [levels][64] largest exponent of a stored path count per level and
    // lane, 0 below 2^512 (null: no scaling, closeness steps)
    int* __restrict__ lexp;
    // Level 2 pushed from the level-1 frontier (k_push2_*) or pulled (k_fwd
    // and its hub kernels): push_e counts the directed edges of the level-1
    // vertices' rows (k_push1_fin adds them), and level 2 is pushed while
    // that count is below push_lim.  push_lim 0: pulled.
    unsigned long long* __restrict__ push_e;
    unsigned long long push_lim;
};

// True where the batch's level 2 is pushed (see FwdArgs::push_lim).
__device__ __forceinline__ bool push2_on(const FwdArgs& a) { return a.push_lim && *a.push_e < a.push_lim; }

struct BwdArgs {
    int n;
    int hub_deg;
    int nchunk;
    const int* __restrict__ chunk_beg;
    const int* __restrict__ rp;
    const int* __restrict__ col;
    const uint64_t* __restrict__ Fd;      // level d masks
    const uint64_t* __restrict__ Fn;      // level d + 1 masks
    // coef may be sigma itself (the dense step writes coef over sigma in
    // place: a lane's entry of v is read as sigma and rewritten as coef at
    // v's level for that lane, while the level's gathers read the entries
    // of lanes one level deeper), so neither is restrict.
    const double* sigma;
    double* coef;
    const uint32_t* __restrict__ npred;
    double* __restrict__ bc;
    // weighted dependencies (reduced graphs): target weight wgt[w] replaces the
    // 1 in (1 + delta) / sigma, and lane s's contribution is scaled by its
    // source's weight wgt[bsrc[s]].  Null: unweighted.
    const int64_t* __restrict__ wgt;
    const int* __restrict__ bsrc;
    // source weights when they differ from the target weights (sampled
    // sources through the reduction), indexed like wgt; null: wgt
    const int64_t* __restrict__ swgt;
    // Levels enqueued ahead of the forward's end: a launch for an empty
    // level (*live == 0) returns at once.  Null: always run.
    const int* __restrict__ live;
    // path-count exponents of the forward (see FwdArgs) and the level d of
    // the vertices this launch finishes
    const int* __restrict__ lexp;
    int d;
};

// Factor that brings level d's stored counts of this lane to about 2^kSigTo
// (1.0 when the level stayed below 2^kSigBig or lexp is null).
__device__ __forceinline__ double level_scale(const int* __restrict__ lexp, int d, int lane) {
    if (!lexp) return 1.0;
    const int e = lexp[(size_t)d * kWave + lane];
    return e >= kSigBig ? __longlong_as_double((long long)(1023 + kSigTo - e) << 52) : 1.0;
}

// ---------------------------------------------------------------- gathers
// Level-sorted rows (LS, the dense betweenness step): a vertex's 64 sigma /
// coef entries are kept in order of the level at which each lane reached
// it, so one level's lanes form one contiguous slice of the 512-byte row
// and a gather of that level touches one or two cache lines instead of four.
// Lane s of v at level d sits at row position base_d(v) + rank of s in
// F_d(v), base_d(v) = number of lanes that reached v earlier.  The frontier
// entry of (d, v) is then a 16-byte record {F_d(v), base_d(v)}, read in the
// same load as the mask.  Without LS (frontier-queue step): plain masks and
// lane-indexed rows.
struct FRec {
    uint64_t mask;
    uint64_t base;
};

// Number of set bits of m below the calling lane.
__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <bool LS>
__device__ __forceinline__ void load_front(const uint64_t* __restrict__ F, int v, uint64_t& mask, int& base) {
    if constexpr (LS) {
        const FRec r = reinterpret_cast<const FRec*>(F)[v];
        mask = r.mask;
        base = (int)r.base;
    } else {
        mask = F[v];
        base = 0;
    }
}

// Forward scan of edges [beg, end): lane s sums sigma[u][s] (SIGMA) over the
// neighbours u with bit s in Fd[u] & cand, or ORs the masks (closeness only).
template <bool SIGMA, bool LS, int U = kUnroll>
__device__ __forceinline__ void gather_fwd(const int* __restrict__ col, const uint64_t* __restrict__ Fd,
                                           const double* __restrict__ sigma, int beg, int end,
                                           uint64_t cand, int lane, double& acc, uint32_t& cnt,
                                           uint64_t& orf) {
    for (int j0 = beg; j0 < end; j0 += kWave) {
        const int j = j0 + lane;
        int u = 0, bu = 0;
        uint64_t fm = 0;
        if (j < end) {
            u = col[j];
            load_front<LS>(Fd, u, fm, bu);
        }
        const uint64_t f = fm & cand;
        if constexpr (SIGMA) {
            uint64_t ball = ballot(f != 0);
            while (ball) {
                double x[U];
#pragma unroll
                for (int t = 0; t < U; ++t) {
                    x[t] = 0.0;
                    if (ball) {
                        const int c = __builtin_ctzll(ball);
                        ball &= ball - 1;
                        const int uc = rl32(u, c);
                        const uint64_t fmc = rl64(fm, c);
                        if (((fmc & cand) >> lane) & 1) {
                            const int pos = LS ? rl32(bu, c) + lane_rank(fmc) : lane;
                            x[t] = sigma[(size_t)uc * kWave + pos];
                            ++cnt;
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < U; ++t) acc += x[t];
            }
        } else {
            // closeness needs only which lanes reach v, not how often: stop
            // once every candidate lane is found (in the degree-ordered
            // device layout a row lists its highest-degree neighbours, the
            // earliest reached, first)
            orf |= f;
            if (j0 + kWave < end && (wave_or(orf) & cand) == cand) break;
        }
    }
}

// Backward scan of edges [beg, end): lane s sums coef[w][s] over neighbours w
// with bit s in Fn[w] & m.
template <bool LS, int U = kUnroll>
__device__ __forceinline__ double gather_bwd(const int* __restrict__ col, const uint64_t* __restrict__ Fn,
                                             const double* __restrict__ coef, int beg, int end,
                                             uint64_t m, int lane) {
    double acc = 0.0;
    for (int j0 = beg; j0 < end; j0 += kWave) {
        const int j = j0 + lane;
        int w = 0, bw = 0;
        uint64_t fm = 0;
        if (j < end) {
            w = col[j];
            load_front<LS>(Fn, w, fm, bw);
        }
        uint64_t ball = ballot((fm & m) != 0);
        while (ball) {
            double x[U];
#pragma unroll
            for (int t = 0; t < U; ++t) {
                x[t] = 0.0;
                if (ball) {
                    const int c = __builtin_ctzll(ball);
                    ball &= ball - 1;
                    const int wc = rl32(w, c);
                    const uint64_t fmc = rl64(fm, c);
                    if (((fmc & m) >> lane) & 1) {
                        const int pos = LS ? rl32(bw, c) + lane_rank(fmc) : lane;
                        x[t] = coef[(size_t)wc * kWave + pos];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < U; ++t) acc += x[t];
        }
    }
    return acc;
}

// ---------------------------------------------------------------- epilogues
struct FwdAccum {
    int64_t far = 0;
    unsigned long long nnew = 0, dag = 0;
    bool ovf = false;
};

template <bool SIGMA, bool COMPAT, bool LS>
__device__ __forceinline__ void fwd_vertex_done(const FwdArgs& a, int v, uint64_t old, double acc,
                                                uint32_t cnt, uint64_t orf_all, int lane, FwdAccum& r,
                                                double sc = 1.0) {
    uint64_t nm;
    if constexpr (SIGMA) nm = ballot(cnt != 0);
    else nm = orf_all;
    if (!nm) {
        // level-sorted planes are not cleared between levels: every record
        // of the next plane is written, an empty one here
        if constexpr (LS)
            if (lane == 0) reinterpret_cast<FRec*>(a.Fn)[v] = FRec{0, 0};
        return;
    }
    const int base = LS ? __popcll(old) : 0;  // lanes that reached v earlier
    if ((nm >> lane) & 1) {
        const size_t idx = (size_t)v * kWave + (LS ? base + lane_rank(nm) : lane);
        if constexpr (SIGMA) {
            acc *= sc;  // level d's scale (exact: a power of two)
            a.sigma[idx] = acc;
            if (acc >= 0x1p512 && a.lexp) atomicMax(&a.lexp[(size_t)a.dnext * kWave + lane], dexp(acc));
            if (!(acc <= 0x1.fffffffffffffp1023)) r.ovf = true;  // inf or NaN
        }
        if constexpr (COMPAT) a.npred[idx] = cnt;
        r.far += a.weight ? (int64_t)a.dnext * a.weight[v] : (int64_t)a.dnext;
        r.dag += cnt;
    }
    if (lane == 0) {
        if constexpr (LS) reinterpret_cast<FRec*>(a.Fn)[v] = FRec{nm, (uint64_t)base};
        else a.Fn[v] = nm;
        a.vis[v] = old | nm;
        const int pc = __popcll(nm);
        if (a.dist_sum) a.dist_sum[v] += (int64_t)a.dnext * pc;  // one wave per vertex and level
        r.nnew += (unsigned long long)pc;
    }
}

// Per-block reduction of the forward partials: one plain add per block slot.
__device__ __forceinline__ void fwd_block_flush(const FwdArgs& a, const FwdAccum& r, int slot) {
    __shared__ int64_t s_far[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_cnt[kWavesPerBlock][2];
    const int lane = lane_id();
    const int wid = threadIdx.x >> 6;
    s_far[wid][lane] = r.far;
    const unsigned long long dag = (unsigned long long)wave_sum_i64((int64_t)r.dag);
    const unsigned long long nn = (unsigned long long)wave_sum_i64((int64_t)r.nnew);
    if (lane == 0) { s_cnt[wid][0] = nn; s_cnt[wid][1] = dag; }
    if (r.ovf) *a.overflow = 1;
    __syncthreads();
    if (wid == 0) {
        int64_t f = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) f += s_far[w][lane];
        if (f) a.part_far[(size_t)slot * kWave + lane] += f;
        if (lane < 2) {
            unsigned long long c = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) c += s_cnt[w][lane];
            if (c) {
                a.part_cnt[(size_t)slot * 2 + lane] += c;
                if (lane == 0) *a.flag = 1;
            }
        }
    }
}

template <bool LS>
__device__ __forceinline__ void bwd_vertex_done(const BwdArgs& a, int v, uint64_t m, int base, double acc,
                                                int lane, bool compat, double sc = 1.0) {
    double contrib = 0.0;
    if ((m >> lane) & 1) {
        const size_t idx = (size_t)v * kWave + (LS ? base + lane_rank(m) : lane);
        const double sg = a.sigma[idx];
        const double dl = (sg * acc) * sc;  // level d's scale, as in the forward
        if (a.wgt) {
            a.coef[idx] = ((double)a.wgt[v] + dl) / sg;
            contrib = dl * (double)(a.swgt ? a.swgt : a.wgt)[a.bsrc[lane]];
        } else {
            a.coef[idx] = (1.0 + dl) / sg;
            contrib = compat ? dl * (double)a.npred[idx] : dl;
        }
    }
    const double tot = wave_sum(contrib);
    if (lane == 0) a.bc[v] += tot;
}

// ---------------------------------------------------------------- forward
// Betweenness levels >= 2 (path counts, level-sorted rows; the closeness
// step has its own interleaved kernels).
template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_fwd(FwdArgs a) {
    if ((a.live && *a.live == 0) || push2_on(a)) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const uint64_t A = a.A;
    const double sc = level_scale(a.lexp, a.dnext - 1, lane);
    FwdAccum r;
    for (int c = gw; c < a.nchunk; c += nw) {
        const int base = a.chunk_beg[c];
        const int vl = base + lane;
        uint64_t visl = A;
        int rbl = 0, rel = 0;
        if (vl < a.chunk_beg[c + 1]) {
            visl = a.vis[vl];
            if (visl != A) {
                rbl = a.rp[vl];
                rel = a.rp[vl + 1];
                if (rel - rbl > a.hub_deg) visl = A;  // hub rows: k_fwd_hubseg
            }
            // the next plane is not cleared between levels: vertices without
            // a scan here (and hub rows, which k_fwd_hubfin may overwrite)
            // get an empty record
            if (visl == A) reinterpret_cast<FRec*>(a.Fn)[vl] = FRec{0, 0};
        }
        uint64_t work = ballot(visl != A);
        while (work) {
            const int b = __builtin_ctzll(work);
            work &= work - 1;
            const int v = base + b;
            const uint64_t old = rl64(visl, b);
            double acc = 0.0;
            uint32_t cnt = 0;
            uint64_t orf = 0;
            gather_fwd<true, true>(a.col, a.Fd, a.sigma, rl32(rbl, b), rl32(rel, b), A & ~old, lane, acc, cnt, orf);
            fwd_vertex_done<true, COMPAT, true>(a, v, old, acc, cnt, 0, lane, r, sc);
        }
    }
    fwd_block_flush(a, r, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_fwd_hubseg(FwdArgs a, HubArgs h) {
    if ((a.live && *a.live == 0) || push2_on(a)) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    for (int i = gw; i < h.nseg; i += nw) {
        const int v = h.seg_v[i];
        const uint64_t cand = a.A & ~a.vis[v];
        if (!cand) continue;  // k_fwd_hubfin skips the vertex too
        double acc = 0.0;
        uint32_t cnt = 0;
        uint64_t orf = 0;
        gather_fwd<true, true, kUnrollHub>(a.col, a.Fd, a.sigma, h.seg_beg[i], h.seg_end[i], cand, lane, acc, cnt, orf);
        h.part_acc[(size_t)i * kWave + lane] = acc;
        h.part_cnt[(size_t)i * kWave + lane] = cnt;
    }
}

template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_fwd_hubfin(FwdArgs a, HubArgs h) {
    if ((a.live && *a.live == 0) || push2_on(a)) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const double sc = level_scale(a.lexp, a.dnext - 1, lane);
    FwdAccum r;
    for (int k = gw; k < h.nhub; k += nw) {
        const int v = h.hub_v[k];
        const uint64_t old = a.vis[v];
        if (!(a.A & ~old)) continue;
        double acc = 0.0;
        uint32_t cnt = 0;
        for (int i = h.hub_seg_off[k] - h.seg_base; i < h.hub_seg_off[k + 1] - h.seg_base; ++i) {
            acc += h.part_acc[(size_t)i * kWave + lane];
            cnt += h.part_cnt[(size_t)i * kWave + lane];
        }
        fwd_vertex_done<true, COMPAT, true>(a, v, old, acc, cnt, 0, lane, r, sc);
    }
    fwd_block_flush(a, r, blockIdx.x);
}

// ---------------------------------------------------------------- level 1 push
// Level 1 of a batch is its sources' neighbourhoods.  The pull step would scan
// every row of the graph to find them; the push step scans only the batch's
// source rows, flattened over the grid (edge e of the concatenated rows goes
// to thread e mod grid size), so a hub source is spread over all waves.
//   k_push1_or   ORs lane s into the level-1 mask of every neighbour of s;
//   k_push1_fin  then finishes each level-1 vertex once, in the thread that
//                holds the edge from the lowest lane of its final mask:
//                record base, vis, path count 1 and predecessor count 1 per
//                lane (the one shortest path is the edge), distance sums and
//                the lane partials.
// The results are those of the pull step: at level 1 every reached lane has
// exactly one shortest path and one predecessor.

// Wave 0 writes the prefix sums of the active lanes' source degrees
// (s_pre[0..64]) and the row starts (s_beg); returns the edge total.
__device__ __forceinline__ int push_rows(uint64_t A, const int* __restrict__ rp, const int* __restrict__ src,
                                         int* s_pre, int* s_beg) {
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        int beg = 0, deg = 0;
        if ((A >> lane) & 1) {
            const int sv = src[lane];
            beg = rp[sv];
            deg = rp[sv + 1] - beg;
        }
        int x = deg;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        s_pre[lane + 1] = x;
        if (lane == 0) s_pre[0] = 0;
        s_beg[lane] = beg;
    }
    __syncthreads();
    return s_pre[kWave];
}

// The lane whose row holds flattened edge e (s_pre[lane] <= e < s_pre[lane + 1]).
__device__ __forceinline__ int push_lane(const int* s_pre, int e) {
    int lo = 0, hi = kWave;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int mid = (lo + hi) >> 1;
        if (s_pre[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_push1_or(FwdArgs a, const int* __restrict__ src) {
    __shared__ int s_pre[kWave + 1];
    __shared__ int s_beg[kWave];
    const int E = push_rows(a.A, a.rp, src, s_pre, s_beg);
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < E; e += gridDim.x * kBlock) {
        const int s = push_lane(s_pre, e);
        const int u = a.col[s_beg[s] + e - s_pre[s]];
        if ((a.vis[u] >> s) & 1) continue;  // a self-loop of lane s's source
        atomicOr((unsigned long long*)&reinterpret_cast<FRec*>(a.Fn)[u].mask, 1ull << s);
    }
}

template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_push1_fin(FwdArgs a, const int* __restrict__ src) {
    __shared__ int s_pre[kWave + 1];
    __shared__ int s_beg[kWave];
    __shared__ unsigned long long s_far[kWave];
    __shared__ unsigned long long s_new, s_edges;
    if (threadIdx.x < kWave) s_far[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_new = s_edges = 0;
    const int E = push_rows(a.A, a.rp, src, s_pre, s_beg);  // synchronises the block
    unsigned long long nn = 0, ne = 0;
    FRec* Fn = reinterpret_cast<FRec*>(a.Fn);
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < E; e += gridDim.x * kBlock) {
        const int s = push_lane(s_pre, e);
        const int u = a.col[s_beg[s] + e - s_pre[s]];
        const uint64_t nm = Fn[u].mask;
        if (!((nm >> s) & 1) || __builtin_ctzll(nm) != s) continue;  // finished by its lowest lane
        const uint64_t old = a.vis[u];
        const int base = __popcll(old);
        Fn[u].base = (uint64_t)base;
        a.vis[u] = old | nm;
        const unsigned long long w = a.weight ? (unsigned long long)a.weight[u] : 1ull;
        int r = 0;
        for (uint64_t m = nm; m; m &= m - 1, ++r) {
            const size_t idx = (size_t)u * kWave + base + r;
            a.sigma[idx] = 1.0;
            if constexpr (COMPAT) a.npred[idx] = 1u;
            atomicAdd(&s_far[__builtin_ctzll(m)], w);
        }
        const int pc = __popcll(nm);
        if (a.dist_sum) a.dist_sum[u] += (int64_t)pc;  // u's finishing thread only
        nn += (unsigned long long)pc;
        ne += (unsigned long long)(a.rp[u + 1] - a.rp[u]);  // u's row: level 2's push work
    }
    if (nn) atomicAdd(&s_new, nn);
    if (ne) atomicAdd(&s_edges, ne);
    __syncthreads();
    if (threadIdx.x == 0 && s_edges && a.push_e) atomicAdd(a.push_e, s_edges);
    const int slot = blockIdx.x;
    if (threadIdx.x < kWave && s_far[threadIdx.x])
        a.part_far[(size_t)slot * kWave + threadIdx.x] += (int64_t)s_far[threadIdx.x];
    if (threadIdx.x == 0 && s_new) {
        a.part_cnt[(size_t)slot * 2] += s_new;
        a.part_cnt[(size_t)slot * 2 + 1] += s_new;  // one shortest-path DAG edge per reached pair
        *a.flag = 1;
    }
}

// ---------------------------------------------------------------- level 2 push
// Level 2 of a batch by push from the level-1 frontier, where the level-1
// vertices' rows hold few of the graph's edges (push2_on; on R-MAT 2^20
// about 15%: the pull would scan every row for them):
//   k_push2_or (+ _hub)   ORs into the level-2 mask of every neighbour w of a
//                         level-1 vertex u the lanes of u that have not
//                         reached w (the level-2 plane starts zeroed);
//   k_push2_fin           finishes every level-2 vertex as the pull would:
//                         record base, vis, zeroed path-count slots,
//                         distance sums and lane partials;
//   k_push2_sig (+ _hub)  adds, over the same edges, the path count of lane
//                         s at u into lane s's slot at w, for the lanes of u
//                         that reach w at level 2.
// Level-1 counts are 1 per lane (one path: the edge), so a level-2 count is
// a small integer and its sum exact in any order: the results are the
// pull's, bit for bit.  Multigraphs, whose level 1 is pulled, are pulled
// here too.

// ORs row [beg, end) of level-1 vertex u (lanes m1) into level 2.
__device__ __forceinline__ void push2_or_row(const FwdArgs& a, int beg, int end, uint64_t m1, int lane) {
    FRec* Fn = reinterpret_cast<FRec*>(a.Fn);
    for (int j = beg + lane; j < end; j += kWave) {
        const int w = a.col[j];
        const uint64_t nm = m1 & ~a.vis[w];
        if (nm) atomicOr((unsigned long long*)&Fn[w].mask, (unsigned long long)nm);
    }
}

__global__ __launch_bounds__(kBlock) void k_push2_or(FwdArgs a) {
    if ((a.live && *a.live == 0) || !push2_on(a)) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const FRec* Fd = reinterpret_cast<const FRec*>(a.Fd);
    for (int c = gw; c < a.nchunk; c += nw) {
        const int v = a.chunk_beg[c] + lane;
        uint64_t m1 = 0;
        int rb = 0, re = 0;
        if (v < a.chunk_beg[c + 1]) {
            m1 = Fd[v].mask;
            if (m1) {
                rb = a.rp[v];
                re = a.rp[v + 1];
                if (re - rb > a.hub_deg) m1 = 0;  // hub rows: k_push2_or_hub
            }
        }
        for (uint64_t work = ballot(m1 != 0); work; work &= work - 1) {
            const int b = __builtin_ctzll(work);
            push2_or_row(a, rl32(rb, b), rl32(re, b), rl64(m1, b), lane);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_push2_or_hub(FwdArgs a, HubArgs h) {
    if ((a.live && *a.live == 0) || !push2_on(a)) return;
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const FRec* Fd = reinterpret_cast<const FRec*>(a.Fd);
    for (int i = gw; i < h.nseg; i += nw) {
        const uint64_t m1 = Fd[h.seg_v[i]].mask;  // one address: the same in every lane
        if (m1) push2_or_row(a, h.seg_beg[i], h.seg_end[i], m1, lane);
    }
}

// Launched on at most the forward step's G workgroups (its lane partials go
// to the block slots).
template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_push2_fin(FwdArgs a) {
    if ((a.live && *a.live == 0) || !push2_on(a)) return;
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    FRec* Fn = reinterpret_cast<FRec*>(a.Fn);
    FwdAccum r;
    for (int c = gw; c < a.nchunk; c += nw) {
        const int v0 = a.chunk_beg[c];
        const int vl = v0 + lane;
        uint64_t nm = 0, old = 0;
        if (vl < a.chunk_beg[c + 1]) {
            nm = Fn[vl].mask;
            if (nm) old = a.vis[vl];
        }
        for (uint64_t work = ballot(nm != 0); work; work &= work - 1) {
            const int b = __builtin_ctzll(work);
            const int v = v0 + b;
            const uint64_t m = rl64(nm, b), o = rl64(old, b);
            const int base = __popcll(o);  // lanes that reached v earlier
            if ((m >> lane) & 1) {
                const size_t idx = (size_t)v * kWave + base + lane_rank(m);
                a.sigma[idx] = 0.0;
                if constexpr (COMPAT) a.npred[idx] = 0u;
                r.far += a.weight ? (int64_t)a.dnext * a.weight[v] : (int64_t)a.dnext;
            }
            if (lane == 0) {
                Fn[v].base = (uint64_t)base;
                a.vis[v] = o | m;
                const int pc = __popcll(m);
                if (a.dist_sum) a.dist_sum[v] += (int64_t)a.dnext * pc;
                r.nnew += (unsigned long long)pc;
            }
        }
    }
    fwd_block_flush(a, r, blockIdx.x);
}

// Adds the path counts of level-1 vertex u (lanes m1 at row base b1) over
// row [beg, end) into the level-2 slots; returns the DAG edges it adds (the
// same in every lane).
template <bool COMPAT>
__device__ __forceinline__ unsigned long long push2_sig_row(const FwdArgs& a, int u, int beg, int end,
                                                            uint64_t m1, int b1, int lane) {
    const FRec* Fn = reinterpret_cast<const FRec*>(a.Fn);
    const bool mine = (m1 >> lane) & 1;
    const double su = mine ? a.sigma[(size_t)u * kWave + b1 + lane_rank(m1)] : 0.0;
    unsigned long long dag = 0;
    for (int j0 = beg; j0 < end; j0 += kWave) {
        const int j = j0 + lane;
        int w = 0, bw = 0;
        uint64_t mw = 0;
        if (j < end) {
            w = a.col[j];
            const FRec f = Fn[w];
            mw = f.mask;
            bw = (int)f.base;
        }
        for (uint64_t ball = ballot((mw & m1) != 0); ball; ball &= ball - 1) {
            const int c = __builtin_ctzll(ball);
            const int wc = rl32(w, c);
            const uint64_t mc = rl64(mw, c);
            if (mine && ((mc >> lane) & 1)) {
                const size_t idx = (size_t)wc * kWave + rl32(bw, c) + lane_rank(mc);
                unsafeAtomicAdd(&a.sigma[idx], su);
                if constexpr (COMPAT) atomicAdd(&a.npred[idx], 1u);
            }
            dag += (unsigned long long)__popcll(mc & m1);
        }
    }
    return dag;
}

// Launched on at most the forward step's G workgroups, like k_push2_fin.
template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_push2_sig(FwdArgs a) {
    if ((a.live && *a.live == 0) || !push2_on(a)) return;
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const FRec* Fd = reinterpret_cast<const FRec*>(a.Fd);
    FwdAccum r;
    for (int c = gw; c < a.nchunk; c += nw) {
        const int v = a.chunk_beg[c] + lane;
        uint64_t m1 = 0;
        int rb = 0, re = 0, b1 = 0;
        if (v < a.chunk_beg[c + 1]) {
            const FRec f = Fd[v];
            m1 = f.mask;
            b1 = (int)f.base;
            if (m1) {
                rb = a.rp[v];
                re = a.rp[v + 1];
                if (re - rb > a.hub_deg) m1 = 0;  // hub rows: k_push2_sig_hub
            }
        }
        for (uint64_t work = ballot(m1 != 0); work; work &= work - 1) {
            const int b = __builtin_ctzll(work);
            r.dag += push2_sig_row<COMPAT>(a, a.chunk_beg[c] + b, rl32(rb, b), rl32(re, b), rl64(m1, b),
                                           rl32(b1, b), lane);
        }
    }
    if (lane != 0) r.dag = 0;  // one count per wave
    fwd_block_flush(a, r, blockIdx.x);
}

template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_push2_sig_hub(FwdArgs a, HubArgs h) {
    if ((a.live && *a.live == 0) || !push2_on(a)) return;
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const FRec* Fd = reinterpret_cast<const FRec*>(a.Fd);
    FwdAccum r;
    for (int i = gw; i < h.nseg; i += nw) {
        const int u = h.seg_v[i];
        const FRec f = Fd[u];
        if (f.mask) r.dag += push2_sig_row<COMPAT>(a, u, h.seg_beg[i], h.seg_end[i], f.mask, (int)f.base, lane);
    }
    if (lane != 0) r.dag = 0;
    fwd_block_flush(a, r, blockIdx.x);
}

// ---------------------------------------------------------------- backward
template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_bwd(BwdArgs a) {
    if (a.live && *a.live == 0) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const double sc = level_scale(a.lexp, a.d, lane);
    for (int c = gw; c < a.nchunk; c += nw) {
        const int base = a.chunk_beg[c];
        const int vl = base + lane;
        uint64_t fl = 0;
        int rbl = 0, rel = 0, bl = 0;
        if (vl < a.chunk_beg[c + 1]) {
            load_front<true>(a.Fd, vl, fl, bl);
            if (fl) {
                rbl = a.rp[vl];
                rel = a.rp[vl + 1];
                if (rel - rbl > a.hub_deg) fl = 0;  // hub rows: k_bwd_hubseg
            }
        }
        uint64_t work = ballot(fl != 0);
        while (work) {
            const int b = __builtin_ctzll(work);
            work &= work - 1;
            const uint64_t m = rl64(fl, b);
            const double acc = gather_bwd<true>(a.col, a.Fn, a.coef, rl32(rbl, b), rl32(rel, b), m, lane);
            bwd_vertex_done<true>(a, base + b, m, rl32(bl, b), acc, lane, COMPAT, sc);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_bwd_hubseg(BwdArgs a, HubArgs h) {
    if (a.live && *a.live == 0) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    for (int i = gw; i < h.nseg; i += nw) {
        uint64_t m;
        int bm;
        load_front<true>(a.Fd, h.seg_v[i], m, bm);
        if (!m) continue;
        h.part_acc[(size_t)i * kWave + lane] = gather_bwd<true, kUnrollHub>(a.col, a.Fn, a.coef, h.seg_beg[i], h.seg_end[i], m, lane);
    }
}

template <bool COMPAT>
__global__ __launch_bounds__(kBlock) void k_bwd_hubfin(BwdArgs a, HubArgs h) {
    if (a.live && *a.live == 0) return;  // uniform over the grid
    const int lane = lane_id();
    const int gw = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int nw = gridDim.x * kWavesPerBlock;
    const double sc = level_scale(a.lexp, a.d, lane);
    for (int k = gw; k < h.nhub; k += nw) {
        const int v = h.hub_v[k];
        uint64_t m;
        int bm;
        load_front<true>(a.Fd, v, m, bm);
        if (!m) continue;
        double acc = 0.0;
        for (int i = h.hub_seg_off[k] - h.seg_base; i < h.hub_seg_off[k + 1] - h.seg_base; ++i)
            acc += h.part_acc[(size_t)i * kWave + lane];
        bwd_vertex_done<true>(a, v, m, bm, acc, lane, COMPAT, sc);
    }
}

// ---------------------------------------------------------------- batch setup
// Seeds a batch: lane s starts at src[s] (duplicates allowed, each its own
// lane).  With path counts (betweenness) F0 holds {mask, base 0} records and
// the sources' rows are level-sorted (base 0); the caller zeroes F0.
__global__ void k_seed(const int* __restrict__ src, int k, uint64_t* __restrict__ F0,
                       uint64_t* __restrict__ vis, double* __restrict__ sigma, bool with_sigma) {
    const int lane = threadIdx.x;
    const int s = (lane < k) ? src[lane] : 0;
    uint64_t* m0 = with_sigma ? &reinterpret_cast<FRec*>(F0)[s].mask : &F0[s];
    if (lane < k) {
        atomicOr((unsigned long long*)m0, 1ull << lane);
        atomicOr((unsigned long long*)&vis[s], 1ull << lane);
    }
    __syncthreads();
    if (lane < k && with_sigma) sigma[(size_t)s * kWave + lane_rank(*m0)] = 1.0;
}

// End of a launch group enqueued with a level bound `bound` (levels 0 ..
// bound - 1 run forward): lvf[i] says level i is non-empty.  Adds the depth
// times the group's batch count to out[0], counts a bound that did not hold
// (level `bound` reached: the group is incomplete) in out[1], and the level
// launches that did work in out[2] (forward) and out[3] (backward).
__global__ void k_level_check(const int* __restrict__ lvf, int bound, int nbatch, bool backward,
                              unsigned long long* __restrict__ out) {
    int i = 1;
    while (i <= bound && lvf[i]) ++i;
    const unsigned long long depth = (unsigned long long)(i - 1);
    atomicAdd(&out[0], depth * (unsigned long long)nbatch);
    if (i > bound) atomicAdd(&out[1], 1ull);
    atomicAdd(&out[2], depth + 1);
    if (backward) atomicAdd(&out[3], depth);
}

// Sums the per-block lane partials of the batches of a launch group into
// far[0..k) and the counters: batch b = blockIdx.y has nblocks partial rows
// from row b * nblocks on, 64 sources (k_last for the last batch) and its
// farness at far + 64 b.  Several workgroups share a batch's rows
// (workgroup g takes rows g*16 + r, then every gridDim.x*16-th); each adds
// its lane totals to far / cnt_out with 64-bit atomics.  The partials are
// integers, so the total does not depend on the order.  far must be zeroed
// before the launch.
__global__ __launch_bounds__(1024) void k_reduce_batch(const int64_t* __restrict__ part_far,
                                                       const unsigned long long* __restrict__ part_cnt,
                                                       int nblocks, int k_last, int64_t* __restrict__ far,
                                                       unsigned long long* __restrict__ cnt_out) {
    constexpr int kRows = 1024 / kWave;
    __shared__ int64_t s_far[kRows][kWave];
    __shared__ unsigned long long s_c[kRows][2];
    const int lane = threadIdx.x & 63;
    const int row = threadIdx.x >> 6;
    const int bb = blockIdx.y;
    const int k = bb == (int)gridDim.y - 1 ? k_last : kWave;
    part_far += (size_t)bb * nblocks * kWave;
    part_cnt += (size_t)bb * nblocks * 2;
    far += (size_t)bb * kWave;
    int64_t f = 0;
    unsigned long long c = 0;
    for (int b = blockIdx.x * kRows + row; b < nblocks; b += gridDim.x * kRows) {
        f += part_far[(size_t)b * kWave + lane];
        if (lane < 2) c += part_cnt[(size_t)b * 2 + lane];
    }
    s_far[row][lane] = f;
    if (lane < 2) s_c[row][lane] = c;
    __syncthreads();
    if (row == 0) {
        int64_t t = 0;
        for (int r = 0; r < kRows; ++r) t += s_far[r][lane];
        if (lane < k && t) atomicAdd((unsigned long long*)&far[lane], (unsigned long long)t);
        if (lane < 2) {
            unsigned long long u = 0;
            for (int r = 0; r < kRows; ++r) u += s_c[r][lane];
            if (u) atomicAdd(&cnt_out[lane], u);
        }
    }
}

}  // namespace cent
