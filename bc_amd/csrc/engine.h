// Internal declarations shared by the engine's translation units
// (engine.hip: contexts, graphs, probe and reference-shaped calls;
// levels.hip: the level steps, plain calls and the autotuner; reduced.hip:
// the D1/AP-reduced form).  Not installed; the C-ABI is include/cent.h.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cent.h"
#include "../Lib/reduce.h"
#include "records.h"
#include "wave.h"

using namespace cent;

// Engine functions shared between the translation units of libcent.so; not
// part of the C-ABI, so not exported.
#define CENT_INTERNAL __attribute__((visibility("hidden")))

#define HIPTRY(expr)                                                           \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            set_error(ctx, e_, #expr);                                         \
            return (e_ == hipErrorOutOfMemory) ? CENT_E_OOM : CENT_E_DEVICE;   \
        }                                                                      \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int st_ = (expr);          \
        if (st_) return st_;       \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
};

struct cent_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char last_error[256] = {0};
    cent_stats stats{};
    // workspace
    DevBuf<uint64_t> vis, F, hp_or;
    DevBuf<double> sigma, coef, bc, hp_acc;
    DevBuf<uint32_t> npred, hp_cnt;
    DevBuf<int64_t> dist_sum, far, part_far;
    DevBuf<unsigned long long> part_cnt, cnt;
    DevBuf<int> src, flags, ord;
    // frontier-queue mode
    DevBuf<uint64_t> Fd2, QF, Adev;
    DevBuf<int> Q, qoff, cand, stamp, ctl;
    // reduced closeness / betweenness per-run buffers
    DevBuf<int64_t> red_D;
    DevBuf<int> red_head;
    DevBuf<double> red_part, red_bcp, it_sig, it_coef;
    DevBuf<uint8_t> it_lvl;
    DevBuf<int> red_deep;
    // sampled sources through the reduction: listed copies, their source
    // weights, and the call's items (range, list offset, count)
    DevBuf<int> red_src, red_order;
    DevBuf<int64_t> red_swgt, red_poff;
    DevBuf<int4> red_items;
    // device scratch of the graph reduction (cent_reduce_gpu), grown to the
    // largest request and reused by later calls
    DevBuf<char> red_arena;
    // source-parallel mode: per-slot arrays
    DevBuf<char> ps_rec;     // PsRec or PsRecW records, by byte
    DevBuf<double> ps_bcp;   // narrow records: the slots' BC partials
    DevBuf<uint32_t> ps_bits;
    DevBuf<int> ps_Q, ps_loff;
    DevBuf<unsigned long long> ps_depth;
    size_t F_levels = 0;     // levels the F buffer holds for the current N
    size_t F_n = 0;
    int* h_flags = nullptr;  // pinned, kHostFlags ints
    void* h_stage = nullptr; // pinned staging for a call's inputs and results
    size_t h_stage_bytes = 0;
    DevBuf<int> lexp;        // [levels][64] path-count exponents of a betweenness batch (scaling)
    DevBuf<unsigned long long> push_e;  // edges of a betweenness batch's level-1 rows (level-2 push switch)
    DevBuf<int> lvf;         // per-level "non-empty" flags of an enqueued run of levels
    DevBuf<unsigned long long> lvstat;  // bounded batches: [0] summed depth, [1] bound failures
    int depth_hint = 0;      // depth of the last dense batch (length of the next run)
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // second stream of the dense step: a peer context (own stream and
    // workspace, same GPU), created on first use; ev_join[0] orders the
    // peer after this stream's set-up, ev_join[1] this stream after the peer
    cent_ctx* peer = nullptr;
    hipEvent_t ev_join[2] = {nullptr, nullptr};
};

// Pinned host ints of a context (level flags of a run, status words).
constexpr int kHostFlags = 64;

// Rows longer than kHubDeg are scanned as kHubDeg-edge segments, one
// wavefront each (see msbfs_kernels.h).
constexpr int kHubDeg = 256;
// Edge budget of one wavefront's vertex chunk in the main level kernels.
constexpr int kChunkEdges = 1024;

static inline void set_error(cent_ctx* ctx, hipError_t e, const char* what) {
    if (!ctx) return;
    snprintf(ctx->last_error, sizeof(ctx->last_error), "%s: %s", what, hipGetErrorString(e));
}

// Capacity ensure() gives a buffer asked for n elements (headroom, so a
// slightly larger next call does not re-allocate).
static inline size_t ensure_cap(size_t n) { return n + n / 4 + 64; }

template <class T>
static inline int ensure(cent_ctx* ctx, DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return CENT_OK;
    if (b.p) { HIPTRY(hipFree(b.p)); b.p = nullptr; b.cap = 0; }
    size_t cap = ensure_cap(n);
    HIPTRY(hipMalloc((void**)&b.p, cap * sizeof(T)));
    b.cap = cap;
    return CENT_OK;
}

// Pinned host staging of at least `bytes` bytes (kept and grown per context):
// copies between it and device memory run at full PCIe rate, where pageable
// buffers go through the runtime's small staging chunks.
static inline int ensure_stage(cent_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->h_stage_bytes) return CENT_OK;
    if (ctx->h_stage) HIPTRY(hipHostFree(ctx->h_stage));
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
    const size_t cap = bytes + bytes / 4 + 4096;
    HIPTRY(hipHostMalloc(&ctx->h_stage, cap, hipHostMallocDefault));
    ctx->h_stage_bytes = cap;
    return CENT_OK;
}

// Exact-size variant for the large per-slot arrays (no growth slack).
template <class T>
static inline int ensure_exact(cent_ctx* ctx, DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return CENT_OK;
    if (b.p) { HIPTRY(hipFree(b.p)); b.p = nullptr; b.cap = 0; }
    HIPTRY(hipMalloc((void**)&b.p, (n > 0 ? n : 1) * sizeof(T)));
    b.cap = n > 0 ? n : 1;
    return CENT_OK;
}

template <class T>
static inline void release(DevBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

static inline int check_device(cent_ctx* ctx) {
    HIPTRY(hipSetDevice(ctx->device));
    return CENT_OK;
}

// ------------------------------------------------------------ device CSR
// Device copy of a CSR plus its scheduling metadata: vertex chunks (<= 64
// vertices, <= kChunkEdges non-hub edges; optionally restarted at given
// block boundaries) and hub segments.  Host copies of the chunk and hub
// offsets allow sub-range views.
struct DevCSR {
    int n = 0;
    int64_t nnz = 0;
    int* rp = nullptr;
    int* col = nullptr;
    int hub_deg = kHubDeg;
    int chunk_edges = kChunkEdges;
    int nchunk = 0;
    int* chunk_beg = nullptr;
    int nseg = 0, nhub = 0;
    int* hub_buf = nullptr;
    int *seg_v = nullptr, *seg_beg = nullptr, *seg_end = nullptr, *hub_v = nullptr, *hub_seg_off = nullptr;
    std::vector<int> h_chunk_beg, h_hub_v, h_hub_seg_off;
    std::vector<int64_t> h_chunk_rp;  // row offset at each chunk start: a view's edge count
    bool multi = false;  // rows may repeat a neighbour or list their own vertex
    int4* ell = nullptr; // [n] inline rows of the source-parallel step (see persource_kernels.h), or null
};

// Failure injection for tests of the error paths (cent_debug_fail_upload):
// while positive, counts down at every upload() below, and the upload that
// brings it to zero fails with CENT_E_OOM before allocating anything.
CENT_INTERNAL extern std::atomic<int> g_fail_upload;

template <class T>
static inline int upload(cent_ctx* ctx, T** dst, const T* src, size_t n) {
    if (g_fail_upload.load(std::memory_order_relaxed) > 0 && g_fail_upload.fetch_sub(1) == 1) return CENT_E_OOM;
    HIPTRY(hipMalloc((void**)dst, sizeof(T) * (n > 0 ? n : 1)));
    if (n) HIPTRY(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return CENT_OK;
}

// Chunk and hub tables, device copy (engine.hip).
CENT_INTERNAL void devcsr_free(DevCSR& d);
CENT_INTERNAL int devcsr_schedule(cent_ctx* ctx, const int32_t* rp, const int32_t* bounds, int32_t nbounds,
                                  int hub_deg, int chunk_edges, DevCSR& d);
CENT_INTERNAL int devcsr_build(cent_ctx* ctx, const int32_t* rp, const int32_t* col, int32_t n, int64_t nnz,
                               const int32_t* bounds, int32_t nbounds, DevCSR& d);

// A sub-range [lo, hi) of a DevCSR whose boundaries are chunk boundaries.
struct GView {
    const DevCSR* d;
    int lo, hi;
    int c0, nchunk;       // chunk range
    int s0, nseg;         // hub segment range
    int h0, nhub;         // hub range
};

static inline GView full_view(const DevCSR& d) {
    return GView{&d, 0, d.n, 0, d.nchunk, 0, d.nseg, 0, d.nhub};
}

static inline GView sub_view(const DevCSR& d, int lo, int hi) {
    GView v{&d, lo, hi, 0, 0, 0, 0, 0, 0};
    const auto& cb = d.h_chunk_beg;
    v.c0 = (int)(std::lower_bound(cb.begin(), cb.end(), lo) - cb.begin());
    const int c1 = (int)(std::lower_bound(cb.begin(), cb.end(), hi) - cb.begin());
    v.nchunk = c1 - v.c0;
    const auto& hv = d.h_hub_v;
    v.h0 = (int)(std::lower_bound(hv.begin(), hv.end(), lo) - hv.begin());
    const int h1 = (int)(std::lower_bound(hv.begin(), hv.end(), hi) - hv.begin());
    v.nhub = h1 - v.h0;
    v.s0 = d.h_hub_seg_off[(size_t)v.h0];
    v.nseg = d.h_hub_seg_off[(size_t)h1] - v.s0;
    return v;
}

struct cent_graph {
    int device = 0;                   // GPU holding the device copy (the graph does not keep its context)
    int32_t n = 0;
    int64_t nnz = 0;
    DevCSR d;
    std::vector<int32_t> comp;        // component id per vertex (host, caller ids)
    std::vector<int64_t> comp_nnz;    // directed edges per component
    std::vector<int32_t> new_id;      // device id of each caller id (empty: identity)
    int32_t* d_new_id = nullptr;      // the same table in device memory (null: identity)
    int32_t ecc_probe = 0;            // BFS depth from the highest-degree vertex
    int32_t width_probe = 0;          // widest level of that BFS
    std::vector<int32_t> h_rp;        // device-order row offsets (host), for re-scheduling
    std::vector<int32_t> probe_dist;  // BFS distance from the highest-degree vertex (caller ids, -1 unreached)
    std::vector<int32_t> probe_order; // that BFS's visit order (caller ids; dropped after the relabel)
    // [2 n] batch grouping key per device id: 4-bit BFS distances to the
    // kLandmarks highest-degree vertices (empty: deep graphs, which run the
    // source-parallel step)
    std::vector<uint64_t> lm_key;
    // [n] with lm_key: each device id's highest-degree neighbour (the lowest
    // id among equals; -1 for none), the grouping key's tie-break
    std::vector<int32_t> lm_top;
    bool bfs_order = false;           // device ids are a BFS order (else degree order or the caller's)
    bool multi = false;               // parallel edges or self-loops (CENT_CSR_RAW input)
    cent_tune_sig sig{};              // autotuner signature
    // launch defaults when the call's opts leave them open: [0] betweenness,
    // [1] closeness
    int32_t grid_waves[2] = {0, 0};
    int32_t frontier[2] = {CENT_FRONTIER_AUTO, CENT_FRONTIER_AUTO};
    int32_t push_pct[2] = {0, 0};
};

// Level step for a plain call that neither the call nor the graph's tuned
// settings choose.  Deep graphs (BFS depth above kDeepLevels from the
// highest-degree vertex, probed at upload) go source-parallel: a 64-lane
// batch there spreads its lanes over different levels and shares no work.
// So does betweenness on graphs of at most kSmallBcVertices vertices, where
// the 64-lane step's per-level launches outweigh the traversals (the tuned
// table picks the same for R-MAT up to 2^13 vertices and for block trees up
// to 113k; a 1554-vertex R-MAT with no table entry: 1024-source BC 12.9 ms
// on the 64-lane step).  Otherwise the dense 64-lane pull.  The
// frontier-queue step is chosen only on request (or by the autotuner).
constexpr int kDeepLevels = 48;
constexpr int kSmallBcVertices = 16384;
constexpr int kLandmarks = 32;  // landmarks of the batch grouping key

static inline int level_mode(const cent_graph* g, const cent_opts* o, bool with_bc) {
    const int m = o ? o->frontier : CENT_FRONTIER_AUTO;
    if (m != CENT_FRONTIER_AUTO) return m;
    if (g->ecc_probe > kDeepLevels || (with_bc && g->n <= kSmallBcVertices)) return CENT_FRONTIER_SOURCE;
    return CENT_FRONTIER_DENSE;
}

// ------------------------------------------------------------ timing
static inline hipEvent_t next_event(cent_ctx* ctx) {
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

struct KernelTimer {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fwd, bwd;
};

static inline int timer_begin(cent_ctx* ctx, KernelTimer& tm, hipEvent_t* e0) {
    *e0 = nullptr;
    if (!tm.on) return CENT_OK;
    *e0 = next_event(ctx);
    if (!*e0) return CENT_E_DEVICE;
    HIPTRY(hipEventRecord(*e0, ctx->stream));
    return CENT_OK;
}

static inline int timer_end(cent_ctx* ctx, KernelTimer& tm, hipEvent_t e0, bool fwd) {
    if (!tm.on) return CENT_OK;
    hipEvent_t e1 = next_event(ctx);
    if (!e1) return CENT_E_DEVICE;
    HIPTRY(hipEventRecord(e1, ctx->stream));
    (fwd ? tm.fwd : tm.bwd).push_back({e0, e1});
    return CENT_OK;
}

static inline void timer_collect(cent_ctx* ctx, KernelTimer& tm) {
    float ms;
    for (auto& p : tm.fwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_fwd += ms;
    for (auto& p : tm.bwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_bwd += ms;
    tm.fwd.clear();
    tm.bwd.clear();
    ctx->ev_used = 0;
    if (ctx->peer) ctx->peer->ev_used = 0;
}

// ------------------------------------------------------------ level loops (levels.hip)
struct BatchSpec {
    bool with_bc = false;
    bool compat = false;
    const int64_t* weight = nullptr;  // device, per vertex of the DevCSR, or null
    int64_t* dist_sum = nullptr;      // device [n] or null
    double* bc = nullptr;             // device [n] (with_bc)
    const int64_t* bwgt = nullptr;    // weighted dependencies (reduced graphs), or null
    const int64_t* swgt = nullptr;    // source weights if not bwgt (sampled sources), or null
    const cent_opts* opts = nullptr;
};

// Closeness batches of 64 sources that one dense level launch carries at
// most (32: a vertex's masks for all of them fill two 128-byte cache lines,
// and one row scan serves 2048 sources).
constexpr int kMultiBatch = 64;

// The context's peer (second stream and workspace on the same GPU), created
// on first use, with the two join events.
CENT_INTERNAL int peer_get(cent_ctx* ctx, cent_ctx** out);
// Zeroes the context's sticky path-count overflow flag (flags[1]), which
// every level step of a call sets and the call reads at its end.
CENT_INTERNAL int level_flags_reset(cent_ctx* ctx);
// Batches of the k sources d_src[0..k) (device ids) inside the view: one
// 64-lane batch for BC (k <= 64), up to kMultiBatch for closeness.
// bound > 0: every BFS of the batch is known to end within `bound` levels
// (levels 0 .. bound - 1 carry frontiers), so the whole batch is enqueued
// without a host check; its depth and a bound failure are added to
// ctx->lvstat on the device.  bound == 0: levels in host-checked runs.
CENT_INTERNAL int level_batch(cent_ctx* ctx, const GView& v, const BatchSpec& sp, const int* d_src, int k,
                              int64_t* far_out, KernelTimer& tm, int64_t* levels, int bound = 0);

// ------------------------------------------------------------ workspace estimate
// Device bytes of a context's buffers, added in the order a call allocates
// them (a buffer that grows is re-allocated before the old one is freed):
// `cur` is what the context holds after the call, `peak` the most at once.
struct WsEst {
    size_t cur = 0, peak = 0;
    void add(size_t b) { cur += b; peak = std::max(peak, cur); }
    void sub(size_t b) { cur -= b; }
};
// Dense 64-lane step on a DevCSR of N vertices with nseg hub segments
// (ensure_workspace, then grow_levels until the level store holds `levels`
// levels for betweenness): nb batches a launch group, G workgroups a launch.
CENT_INTERNAL void ws_dense(WsEst& w, size_t N, size_t nseg, bool with_bc, bool compat, int nb, int G,
                            size_t levels);
// Workgroups of a dense level launch over nchunk vertex chunks (grid_waves:
// the call's or the graph's cap, 0 = default).
CENT_INTERNAL int ws_grid(size_t nchunk, int grid_waves);
// Level store a dense betweenness call may need on a graph whose probe BFS
// (from the highest-degree vertex) is `depth` levels deep (-1: unknown, the
// depth up to which the dense step is chosen by default).
CENT_INTERNAL size_t ws_levels(int32_t depth);
