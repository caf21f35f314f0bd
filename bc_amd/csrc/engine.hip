// Host side of the centrality engine: contexts, device graphs, the level
// loop of the 64-source batches, the reduced (D1 + AP) closeness schedule,
// and the C-ABI declared in include/cent.h.
//
// A call processes its sources in batches of 64 (one wavefront lane per
// source).  Per batch: seed, forward levels until a level reaches nothing
// (one flag read per level), then — for betweenness — backward levels from
// the deepest level to 1.  All kernels of a call run in order on the
// context's stream, so the per-vertex BC sums are accumulated in a fixed
// order and the result does not depend on timing.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cent.h"
#include "../Lib/reduce.h"
#include "items_kernels.h"
#include "msbfs_kernels.h"
#include "persource_kernels.h"
#include "sparse_kernels.h"

using namespace cent;

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
    DevBuf<int> src, flags;
    // frontier-queue mode
    DevBuf<uint64_t> Fd2, QF, Adev;
    DevBuf<int> Q, qoff, cand, stamp, ctl;
    // reduced closeness / betweenness per-run buffers
    DevBuf<int64_t> red_D;
    DevBuf<int> red_head;
    DevBuf<double> red_part, red_bcp, it_sig, it_coef;
    DevBuf<uint8_t> it_lvl;
    DevBuf<int> red_deep;
    // source-parallel mode: per-slot arrays
    DevBuf<PsRec> ps_rec;
    DevBuf<int> ps_Q, ps_loff;
    DevBuf<double> ps_bcp;
    DevBuf<unsigned long long> ps_depth;
    size_t F_levels = 0;     // levels the F buffer holds for the current N
    size_t F_n = 0;
    int* h_flags = nullptr;  // pinned
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
};

// Rows longer than kHubDeg are scanned as kHubDeg-edge segments, one
// wavefront each (see msbfs_kernels.h).
constexpr int kHubDeg = 256;
// Edge budget of one wavefront's vertex chunk in the main level kernels.
constexpr int kChunkEdges = 1024;

static void set_error(cent_ctx* ctx, hipError_t e, const char* what) {
    if (!ctx) return;
    snprintf(ctx->last_error, sizeof(ctx->last_error), "%s: %s", what, hipGetErrorString(e));
}

template <class T>
static int ensure(cent_ctx* ctx, DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return CENT_OK;
    if (b.p) { HIPTRY(hipFree(b.p)); b.p = nullptr; b.cap = 0; }
    size_t cap = n + n / 4 + 64;
    HIPTRY(hipMalloc((void**)&b.p, cap * sizeof(T)));
    b.cap = cap;
    return CENT_OK;
}

// Exact-size variant for the large per-slot arrays (no growth slack).
template <class T>
static int ensure_exact(cent_ctx* ctx, DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return CENT_OK;
    if (b.p) { HIPTRY(hipFree(b.p)); b.p = nullptr; b.cap = 0; }
    HIPTRY(hipMalloc((void**)&b.p, (n > 0 ? n : 1) * sizeof(T)));
    b.cap = n > 0 ? n : 1;
    return CENT_OK;
}

template <class T>
static void release(DevBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

static int check_device(cent_ctx* ctx) {
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
};

static void devcsr_free(DevCSR& d) {
    if (d.rp) (void)hipFree(d.rp);
    if (d.col) (void)hipFree(d.col);
    if (d.chunk_beg) (void)hipFree(d.chunk_beg);
    if (d.hub_buf) (void)hipFree(d.hub_buf);
    d = DevCSR();
}

template <class T>
static int upload(cent_ctx* ctx, T** dst, const T* src, size_t n) {
    HIPTRY(hipMalloc((void**)dst, sizeof(T) * (n > 0 ? n : 1)));
    if (n) HIPTRY(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return CENT_OK;
}

// Chunk and hub-segment tables of d for rows rp (host copy of the device
// row offsets), replacing any previous ones.  bounds: optional sorted block
// starts (chunks never straddle a boundary).
static int devcsr_schedule(cent_ctx* ctx, const int32_t* rp, const int32_t* bounds, int32_t nbounds,
                           int hub_deg, int chunk_edges, DevCSR& d) {
    (void)ctx;
    const int32_t n = d.n;
    if (d.chunk_beg) (void)hipFree(d.chunk_beg);
    if (d.hub_buf) (void)hipFree(d.hub_buf);
    d.chunk_beg = nullptr;
    d.hub_buf = nullptr;
    d.seg_v = d.seg_beg = d.seg_end = d.hub_v = d.hub_seg_off = nullptr;
    d.hub_deg = hub_deg;
    d.chunk_edges = chunk_edges;
    // chunks
    std::vector<int>& cb = d.h_chunk_beg;
    cb.clear();
    cb.push_back(0);
    int cnt = 0, bi = 0;
    int64_t edges = 0;
    for (int32_t v = 0; v < n; ++v) {
        bool brk = false;
        while (bounds && bi < nbounds && bounds[bi] <= v) {
            if (bounds[bi] == v && v > 0) brk = true;
            ++bi;
        }
        int dg = rp[v + 1] - rp[v];
        if (dg > d.hub_deg) dg = 0;
        if (cnt > 0 && (brk || cnt == kWave || edges + dg > d.chunk_edges)) {
            cb.push_back(v);
            cnt = 0;
            edges = 0;
        }
        ++cnt;
        edges += dg;
    }
    if (n > 0) cb.push_back(n);
    d.nchunk = (int)cb.size() - 1;
    TRY(upload(ctx, &d.chunk_beg, cb.data(), cb.size()));
    // hub segments
    std::vector<int> sv, sb, se;
    d.h_hub_v.clear();
    d.h_hub_seg_off.assign(1, 0);
    for (int32_t v = 0; v < n; ++v) {
        const int b = rp[v], e = rp[v + 1];
        if (e - b <= d.hub_deg) continue;
        d.h_hub_v.push_back(v);
        for (int x = b; x < e; x += d.hub_deg) {
            sv.push_back(v);
            sb.push_back(x);
            se.push_back(e - x > d.hub_deg ? x + d.hub_deg : e);
        }
        d.h_hub_seg_off.push_back((int)sv.size());
    }
    d.nseg = (int)sv.size();
    d.nhub = (int)d.h_hub_v.size();
    if (d.nhub) {
        std::vector<int> all;
        all.insert(all.end(), sv.begin(), sv.end());
        all.insert(all.end(), sb.begin(), sb.end());
        all.insert(all.end(), se.begin(), se.end());
        all.insert(all.end(), d.h_hub_v.begin(), d.h_hub_v.end());
        all.insert(all.end(), d.h_hub_seg_off.begin(), d.h_hub_seg_off.end());
        TRY(upload(ctx, &d.hub_buf, all.data(), all.size()));
        d.seg_v = d.hub_buf;
        d.seg_beg = d.seg_v + d.nseg;
        d.seg_end = d.seg_beg + d.nseg;
        d.hub_v = d.seg_end + d.nseg;
        d.hub_seg_off = d.hub_v + d.nhub;
    }
    return CENT_OK;
}

static int devcsr_build(cent_ctx* ctx, const int32_t* rp, const int32_t* col, int32_t n, int64_t nnz,
                        const int32_t* bounds, int32_t nbounds, DevCSR& d) {
    d.n = n;
    d.nnz = nnz;
    TRY(upload(ctx, &d.rp, rp, (size_t)n + 1));
    TRY(upload(ctx, &d.col, col, (size_t)nnz));
    return devcsr_schedule(ctx, rp, bounds, nbounds, kHubDeg, kChunkEdges, d);
}

// A sub-range [lo, hi) of a DevCSR whose boundaries are chunk boundaries.
struct GView {
    const DevCSR* d;
    int lo, hi;
    int c0, nchunk;       // chunk range
    int s0, nseg;         // hub segment range
    int h0, nhub;         // hub range
};

static GView full_view(const DevCSR& d) {
    return GView{&d, 0, d.n, 0, d.nchunk, 0, d.nseg, 0, d.nhub};
}

static GView sub_view(const DevCSR& d, int lo, int hi) {
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
    cent_ctx* ctx = nullptr;
    int32_t n = 0;
    int64_t nnz = 0;
    DevCSR d;
    std::vector<int32_t> comp;        // component id per vertex (host, caller ids)
    std::vector<int64_t> comp_nnz;    // directed edges per component
    std::vector<int32_t> new_id;      // device id of each caller id (empty: identity)
    int32_t ecc_probe = 0;            // BFS depth from the highest-degree vertex
    int32_t width_probe = 0;          // widest level of that BFS
    std::vector<int32_t> h_rp;        // device-order row offsets (host), for re-scheduling
    std::vector<int32_t> probe_dist;  // BFS distance from the highest-degree vertex (caller ids, -1 unreached)
    cent_tune_sig sig{};              // autotuner signature
    // launch defaults when the call's opts leave them open: [0] betweenness,
    // [1] closeness
    int32_t grid_waves[2] = {0, 0};
    int32_t frontier[2] = {CENT_FRONTIER_AUTO, CENT_FRONTIER_AUTO};
};

// Level step for a plain call.  Deep graphs (BFS depth above kDeepLevels
// from the highest-degree vertex, probed at upload) go source-parallel: a
// 64-lane batch there spreads its lanes over different levels and shares
// no work.  Otherwise the dense 64-lane pull.  The frontier-queue step is
// chosen only on request (or by the autotuner).
constexpr int kDeepLevels = 48;

static int level_mode(const cent_graph* g, const cent_opts* o) {
    const int m = o ? o->frontier : CENT_FRONTIER_AUTO;
    if (m != CENT_FRONTIER_AUTO) return m;
    return g->ecc_probe > kDeepLevels ? CENT_FRONTIER_SOURCE : CENT_FRONTIER_DENSE;
}

extern "C" int cent_ctx_create(int32_t device, cent_ctx** out) {
    if (!out) return CENT_E_INVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return CENT_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CENT_E_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CENT_E_DEVICE;
    cent_ctx* ctx = new (std::nothrow) cent_ctx();
    if (!ctx) return CENT_E_OOM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CENT_E_DEVICE;
    }
    if (hipHostMalloc((void**)&ctx->h_flags, sizeof(int) * 2, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CENT_E_DEVICE;
    }
    *out = ctx;
    return CENT_OK;
}

extern "C" void cent_ctx_destroy(cent_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx->vis); release(ctx->F); release(ctx->hp_or); release(ctx->sigma); release(ctx->coef);
    release(ctx->bc); release(ctx->hp_acc); release(ctx->npred); release(ctx->hp_cnt);
    release(ctx->dist_sum); release(ctx->far); release(ctx->part_far); release(ctx->part_cnt);
    release(ctx->cnt); release(ctx->src); release(ctx->flags);
    release(ctx->Fd2); release(ctx->QF); release(ctx->Adev); release(ctx->Q); release(ctx->qoff);
    release(ctx->cand); release(ctx->stamp); release(ctx->ctl);
    release(ctx->red_D); release(ctx->red_head);
    release(ctx->red_part); release(ctx->red_bcp); release(ctx->it_sig); release(ctx->it_coef);
    release(ctx->it_lvl); release(ctx->red_deep);
    release(ctx->ps_rec); release(ctx->ps_Q); release(ctx->ps_loff); release(ctx->ps_bcp);
    release(ctx->ps_depth);
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int cent_ctx_stats(const cent_ctx* ctx, cent_stats* out) {
    if (!ctx || !out) return CENT_E_INVAL;
    *out = ctx->stats;
    return CENT_OK;
}

extern "C" int cent_ctx_reset_stats(cent_ctx* ctx) {
    if (!ctx) return CENT_E_INVAL;
    ctx->stats = cent_stats{};
    return CENT_OK;
}

extern "C" int64_t cent_ctx_workspace_bytes(const cent_ctx* ctx) {
    if (!ctx) return 0;
    return (int64_t)(8 * (ctx->vis.cap + ctx->F.cap + ctx->hp_or.cap + ctx->sigma.cap + ctx->coef.cap +
                          ctx->bc.cap + ctx->hp_acc.cap + ctx->dist_sum.cap + ctx->far.cap + ctx->part_far.cap +
                          ctx->part_cnt.cap + ctx->cnt.cap + ctx->Fd2.cap + ctx->QF.cap + ctx->Adev.cap +
                          ctx->red_D.cap + ctx->red_part.cap + ctx->red_bcp.cap + ctx->it_sig.cap +
                          ctx->it_coef.cap + 2 * ctx->ps_rec.cap + ctx->ps_bcp.cap + ctx->ps_depth.cap) +
                     ctx->it_lvl.cap +
                     4 * (ctx->npred.cap + ctx->hp_cnt.cap + ctx->src.cap + ctx->flags.cap + ctx->Q.cap +
                          ctx->qoff.cap + ctx->cand.cap + ctx->stamp.cap + ctx->ctl.cap + ctx->red_head.cap +
                          ctx->red_deep.cap + ctx->ps_Q.cap + ctx->ps_loff.cap));
}

extern "C" int cent_ctx_trim(cent_ctx* ctx) {
    if (!ctx) return CENT_E_INVAL;
    TRY(check_device(ctx));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    release(ctx->vis); release(ctx->F); release(ctx->hp_or); release(ctx->sigma); release(ctx->coef);
    release(ctx->bc); release(ctx->hp_acc); release(ctx->npred); release(ctx->hp_cnt);
    release(ctx->dist_sum); release(ctx->far); release(ctx->part_far); release(ctx->part_cnt);
    release(ctx->cnt); release(ctx->src); release(ctx->flags);
    release(ctx->Fd2); release(ctx->QF); release(ctx->Adev); release(ctx->Q); release(ctx->qoff);
    release(ctx->cand); release(ctx->stamp); release(ctx->ctl);
    release(ctx->red_D); release(ctx->red_head);
    release(ctx->red_part); release(ctx->red_bcp); release(ctx->it_sig); release(ctx->it_coef);
    release(ctx->it_lvl); release(ctx->red_deep);
    release(ctx->ps_rec); release(ctx->ps_Q); release(ctx->ps_loff); release(ctx->ps_bcp);
    release(ctx->ps_depth);
    ctx->F_levels = 0;
    ctx->F_n = 0;
    return CENT_OK;
}

extern "C" const char* cent_ctx_last_error(const cent_ctx* ctx) {
    return ctx ? ctx->last_error : "null context";
}

// ------------------------------------------------------------ graphs
static int validate_csr(const int32_t* row_ptr, const int32_t* col, int32_t n, int64_t nnz) {
    if (n < 0 || nnz < 0 || nnz > INT32_MAX || !row_ptr || (nnz > 0 && !col)) return CENT_E_INVAL;
    if (row_ptr[0] != 0 || row_ptr[n] != nnz) return CENT_E_INVAL;
    for (int32_t v = 0; v < n; ++v)
        if (row_ptr[v + 1] < row_ptr[v]) return CENT_E_INVAL;
    for (int64_t j = 0; j < nnz; ++j)
        if (col[j] < 0 || col[j] >= n) return CENT_E_INVAL;
    return CENT_OK;
}

extern "C" void cent_graph_free(cent_graph* g) {
    if (!g) return;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    devcsr_free(g->d);
    delete g;
}

extern "C" int cent_graph_create(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col, int32_t n,
                                 int64_t nnz, int32_t flags, cent_graph** out) {
    if (!ctx || !out) return CENT_E_INVAL;
    *out = nullptr;
    TRY(validate_csr(row_ptr, col, n, nnz));
    TRY(check_device(ctx));
    cent_graph* g = new (std::nothrow) cent_graph();
    if (!g) return CENT_E_OOM;
    g->ctx = ctx;
    g->n = n;
    g->nnz = nnz;
    // component ids + per-component directed edge counts, for TEPS accounting
    g->comp.assign((size_t)n, -1);
    std::vector<int32_t> q((size_t)(n > 0 ? n : 1));
    int32_t nc = 0;
    for (int32_t s = 0; s < n; ++s) {
        if (g->comp[s] >= 0) continue;
        size_t h = 0, t = 0;
        int64_t e = 0;
        q[t++] = s;
        g->comp[s] = nc;
        while (h < t) {
            int32_t x = q[h++];
            e += row_ptr[x + 1] - row_ptr[x];
            for (int32_t j = row_ptr[x]; j < row_ptr[x + 1]; ++j)
                if (g->comp[col[j]] < 0) { g->comp[col[j]] = nc; q[t++] = col[j]; }
        }
        g->comp_nnz.push_back(e);
        ++nc;
    }
    // Device layout: vertices relabelled by descending degree (ties by id), so
    // the rows of high-degree vertices — the most gathered sigma/coef rows —
    // share a small, cache-resident address range.  Rows are kept sorted by
    // the new ids.  CENT_GRAPH_KEEP_ORDER keeps the caller's ids.
    std::vector<int32_t> drp, dcol;
    const int32_t* urp = row_ptr;
    const int32_t* ucol = col;
    if (!(flags & CENT_GRAPH_KEEP_ORDER) && n > 1) {
        std::vector<int32_t> order((size_t)n);
        for (int32_t v = 0; v < n; ++v) order[(size_t)v] = v;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return (row_ptr[x + 1] - row_ptr[x]) > (row_ptr[y + 1] - row_ptr[y]);
        });
        g->new_id.resize((size_t)n);
        for (int32_t i = 0; i < n; ++i) g->new_id[(size_t)order[(size_t)i]] = i;
        drp.resize((size_t)n + 1);
        dcol.resize((size_t)(nnz > 0 ? nnz : 1));
        drp[0] = 0;
        for (int32_t i = 0; i < n; ++i) {
            const int32_t v = order[(size_t)i];
            int32_t o = drp[(size_t)i];
            for (int32_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) dcol[(size_t)o++] = g->new_id[(size_t)col[j]];
            std::sort(dcol.begin() + drp[(size_t)i], dcol.begin() + o);
            drp[(size_t)i + 1] = o;
        }
        urp = drp.data();
        ucol = dcol.data();
    }
    // eccentricity probe from the highest-degree vertex (host BFS, O(n + m))
    if (n > 0) {
        int32_t s0 = 0;
        for (int32_t v = 1; v < n; ++v)
            if (row_ptr[v + 1] - row_ptr[v] > row_ptr[s0 + 1] - row_ptr[s0]) s0 = v;
        std::vector<int32_t> dist((size_t)n, -1);
        size_t h = 0, t = 0;
        q[t++] = s0;
        dist[(size_t)s0] = 0;
        while (h < t) {
            const int32_t x = q[h++];
            for (int32_t j = row_ptr[x]; j < row_ptr[x + 1]; ++j)
                if (dist[(size_t)col[j]] < 0) { dist[(size_t)col[j]] = dist[(size_t)x] + 1; q[t++] = col[j]; }
        }
        g->ecc_probe = dist[(size_t)q[t - 1]];
        g->probe_dist = dist;  // batch grouping key (caller ids)
        std::vector<int32_t> width((size_t)g->ecc_probe + 1, 0);
        for (size_t i = 0; i < t; ++i) width[(size_t)dist[(size_t)q[i]]]++;
        g->width_probe = *std::max_element(width.begin(), width.end());
    }
    int st = devcsr_build(ctx, urp, ucol, n, nnz, nullptr, 0, g->d);
    if (st) {
        cent_graph_free(g);
        return st;
    }
    g->h_rp.assign(urp, urp + (size_t)n + 1);
    (void)cent_tune_signature(g->h_rp.data(), n, nnz, g->ecc_probe, &g->sig);
    if (!(flags & CENT_GRAPH_NO_TUNE)) {
        // closeness first, so the betweenness entry's chunk/hub layout wins
        for (int32_t what : {CENT_TUNE_CC, CENT_TUNE_BC}) {
            cent_tune_params tp;
            if (cent_tune_table_get(&g->sig, what, &tp) != CENT_OK) continue;
            tp.ms = 0.0;
            st = cent_graph_set_layout(g, what, &tp);
            if (st) {
                cent_graph_free(g);
                return st;
            }
        }
    }
    *out = g;
    return CENT_OK;
}

// ------------------------------------------------------------ tuning hooks
static bool what_ok(int32_t what, bool both_ok) {
    return what == CENT_TUNE_BC || what == CENT_TUNE_CC || (both_ok && what == (CENT_TUNE_BC | CENT_TUNE_CC));
}

extern "C" int cent_graph_get_layout(const cent_graph* g, int32_t what, cent_tune_params* out) {
    if (!g || !out || !what_ok(what, false)) return CENT_E_INVAL;
    const int i = what == CENT_TUNE_BC ? 0 : 1;
    out->hub_deg = g->d.hub_deg;
    out->chunk_edges = g->d.chunk_edges;
    out->grid_waves = g->grid_waves[i];
    out->frontier = g->frontier[i];
    out->ms = 0.0;
    return CENT_OK;
}

extern "C" int cent_graph_set_layout(cent_graph* g, int32_t what, const cent_tune_params* p) {
    if (!g || !what_ok(what, true)) return CENT_E_INVAL;
    TRY(cent_tune_params_check(p));
    if (p->hub_deg != g->d.hub_deg || p->chunk_edges != g->d.chunk_edges) {
        cent_ctx* ctx = g->ctx;
        TRY(check_device(ctx));
        HIPTRY(hipDeviceSynchronize());  // no launch may still read the old tables
        TRY(devcsr_schedule(ctx, g->h_rp.data(), nullptr, 0, p->hub_deg, p->chunk_edges, g->d));
    }
    for (int i = 0; i < 2; ++i)
        if (what & (i == 0 ? CENT_TUNE_BC : CENT_TUNE_CC)) {
            g->grid_waves[i] = p->grid_waves;
            g->frontier[i] = p->frontier;
        }
    return CENT_OK;
}

extern "C" int cent_graph_signature(const cent_graph* g, cent_tune_sig* out) {
    if (!g || !out) return CENT_E_INVAL;
    *out = g->sig;
    return CENT_OK;
}

extern "C" int cent_graph_from_csr(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col,
                                   int32_t n, int64_t nnz, cent_graph** out) {
    return cent_graph_create(ctx, row_ptr, col, n, nnz, 0, out);
}

extern "C" int cent_graph_upload(cent_ctx* ctx, const struct CSR* csr, cent_graph** out) {
    if (!csr) return CENT_E_INVAL;
    return cent_graph_from_csr(ctx, csr->csrV, csr->csrE, csr->csrVSize - 1, csr->csrESize, out);
}

// ------------------------------------------------------------ timing
static hipEvent_t next_event(cent_ctx* ctx) {
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

static int timer_begin(cent_ctx* ctx, KernelTimer& tm, hipEvent_t* e0) {
    *e0 = nullptr;
    if (!tm.on) return CENT_OK;
    *e0 = next_event(ctx);
    if (!*e0) return CENT_E_DEVICE;
    HIPTRY(hipEventRecord(*e0, ctx->stream));
    return CENT_OK;
}

static int timer_end(cent_ctx* ctx, KernelTimer& tm, hipEvent_t e0, bool fwd) {
    if (!tm.on) return CENT_OK;
    hipEvent_t e1 = next_event(ctx);
    if (!e1) return CENT_E_DEVICE;
    HIPTRY(hipEventRecord(e1, ctx->stream));
    (fwd ? tm.fwd : tm.bwd).push_back({e0, e1});
    return CENT_OK;
}

static void timer_collect(cent_ctx* ctx, KernelTimer& tm) {
    float ms;
    for (auto& p : tm.fwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_fwd += ms;
    for (auto& p : tm.bwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_bwd += ms;
    tm.fwd.clear();
    tm.bwd.clear();
    ctx->ev_used = 0;
}

// ------------------------------------------------------------ level loop
static int grid_for(int chunks, const cent_opts* o) {
    int blocks = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    int cap = (o && o->grid_waves > 0) ? (o->grid_waves + kWavesPerBlock - 1) / kWavesPerBlock : 2048;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return blocks;
}

struct BatchSpec {
    bool with_bc = false;
    bool compat = false;
    const int64_t* weight = nullptr;  // device, per vertex of the DevCSR, or null
    int64_t* dist_sum = nullptr;      // device [n] or null
    double* bc = nullptr;             // device [n] (with_bc)
    const int64_t* bwgt = nullptr;    // weighted dependencies (reduced graphs), or null
    const cent_opts* opts = nullptr;
};

// Workspace for a DevCSR with N vertices; the F store grows on demand.
static int ensure_workspace(cent_ctx* ctx, const DevCSR& d, const BatchSpec& sp, int G, int nb) {
    const size_t N = (size_t)(d.n > 0 ? d.n : 1);
    TRY(ensure(ctx, ctx->vis, N * (size_t)nb));
    // BC keeps one mask array per level; closeness a level-parity pair per batch
    const size_t want = sp.with_bc ? N * 17 : N * (2 * (size_t)nb + 1);
    if (ctx->F_n != N || ctx->F.cap < want) {
        TRY(ensure(ctx, ctx->F, want));
        ctx->F_n = N;
    }
    ctx->F_levels = ctx->F.cap / N - 1;
    if (sp.with_bc) {
        TRY(ensure(ctx, ctx->sigma, N * kWave));
        TRY(ensure(ctx, ctx->coef, N * kWave));
        if (sp.compat) TRY(ensure(ctx, ctx->npred, N * kWave));
    }
    TRY(ensure(ctx, ctx->part_far, (size_t)G * kWave * nb));
    TRY(ensure(ctx, ctx->part_cnt, (size_t)G * 2 * nb));
    TRY(ensure(ctx, ctx->cnt, 2));
    TRY(ensure(ctx, ctx->flags, 2));
    if (d.nhub) {
        TRY(ensure(ctx, ctx->hp_acc, (size_t)d.nseg * kWave));
        TRY(ensure(ctx, ctx->hp_cnt, (size_t)d.nseg * kWave));
        TRY(ensure(ctx, ctx->hp_or, (size_t)d.nseg * nb));
    }
    return CENT_OK;
}

static int grow_levels(cent_ctx* ctx, size_t N, int keep_levels) {
    const size_t ncap = ctx->F_levels * 2;
    uint64_t* nf = nullptr;
    HIPTRY(hipMalloc((void**)&nf, sizeof(uint64_t) * N * (ncap + 1)));
    HIPTRY(hipMemcpyAsync(nf, ctx->F.p, sizeof(uint64_t) * N * (size_t)keep_levels, hipMemcpyDeviceToDevice,
                          ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    HIPTRY(hipFree(ctx->F.p));
    ctx->F.p = nf;
    ctx->F.cap = N * (ncap + 1);
    ctx->F_levels = ncap;
    return CENT_OK;
}

// Totals of batch b from its G per-workgroup partials: far_out[0..kb)
// assigned, ctx->cnt accumulated.  One reducing workgroup per 128 partial
// rows, at most 32.
static int reduce_batch(cent_ctx* ctx, int G, int kb, int64_t* far_out, int b = 0) {
    hipStream_t s = ctx->stream;
    HIPTRY(hipMemsetAsync(far_out, 0, sizeof(int64_t) * (size_t)(kb > 0 ? kb : 1), s));
    const int rg = std::max(1, std::min(32, G / 128));
    k_reduce_batch<<<rg, 1024, 0, s>>>(ctx->part_far.p + (size_t)b * G * kWave, ctx->part_cnt.p + (size_t)b * G * 2,
                                       G, kb, far_out, ctx->cnt.p);
    HIPTRY(hipGetLastError());
    return CENT_OK;
}

// Batches of the k sources d_src[0..k) (device ids) whose BFS stays inside
// the view's vertex range: one 64-lane batch for BC (k <= 64), up to
// kMultiBatch batches of 64 consecutive sources side by side for closeness
// (each level is one launch over all of them, so launches and host
// round trips are shared).  far_out (device, k entries) gets the weighted
// distance sums.  Adds the depth to *levels once per batch.
constexpr int kMultiBatch = 8;

static int level_batch(cent_ctx* ctx, const GView& v, const BatchSpec& sp, const int* d_src, int k,
                       int64_t* far_out, KernelTimer& tm, int64_t* levels) {
    const DevCSR& d = *v.d;
    hipStream_t s = ctx->stream;
    const size_t N = (size_t)(d.n > 0 ? d.n : 1);
    const size_t R = (size_t)(v.hi - v.lo);
    const int nb = (k + kWave - 1) / kWave;
    if (nb < 1 || nb > kMultiBatch || (sp.with_bc && nb != 1)) return CENT_E_INVAL;
    const int G = grid_for(v.nchunk, sp.opts);
    TRY(ensure_workspace(ctx, d, sp, G, nb));
    const int kl = k - (nb - 1) * kWave;  // sources of the last batch
    const uint64_t A_last = (kl == 64) ? ~0ull : ((1ull << kl) - 1);
    const uint64_t A = nb == 1 ? A_last : ~0ull;
    const size_t planes = sp.with_bc ? 1 : (size_t)nb;  // closeness: batch b's level-parity plane p at (p*nb + b)*N
    for (int b = 0; b < nb; ++b) {
        HIPTRY(hipMemsetAsync(ctx->vis.p + (size_t)b * N + v.lo, 0, sizeof(uint64_t) * R, s));
        HIPTRY(hipMemsetAsync(ctx->F.p + (size_t)b * N + v.lo, 0, sizeof(uint64_t) * R, s));
        HIPTRY(hipMemsetAsync(ctx->F.p + (planes + b) * N + v.lo, 0, sizeof(uint64_t) * R, s));
    }
    HIPTRY(hipMemsetAsync(ctx->part_far.p, 0, sizeof(int64_t) * (size_t)G * kWave * nb, s));
    HIPTRY(hipMemsetAsync(ctx->part_cnt.p, 0, sizeof(unsigned long long) * (size_t)G * 2 * nb, s));
    HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
    for (int b = 0; b < nb; ++b)
        k_seed<<<1, kWave, 0, s>>>(d_src + b * kWave, b == nb - 1 ? kl : kWave, ctx->F.p + (size_t)b * N,
                                   ctx->vis.p + (size_t)b * N, sp.with_bc ? ctx->sigma.p : nullptr, sp.with_bc);
    HIPTRY(hipGetLastError());

    HubArgs ha{};
    int Gseg = 1, Ghub = 1;
    if (v.nhub) {
        ha.nseg = v.nseg; ha.seg_v = d.seg_v + v.s0; ha.seg_beg = d.seg_beg + v.s0; ha.seg_end = d.seg_end + v.s0;
        ha.seg_base = v.s0;
        ha.nhub = v.nhub; ha.hub_v = d.hub_v + v.h0; ha.hub_seg_off = d.hub_seg_off + v.h0;
        ha.part_acc = ctx->hp_acc.p; ha.part_cnt = ctx->hp_cnt.p; ha.part_or = ctx->hp_or.p;
        Gseg = std::min((v.nseg + kWavesPerBlock - 1) / kWavesPerBlock, 4096);
        Ghub = std::min((v.nhub + kWavesPerBlock - 1) / kWavesPerBlock, G);  // hubfin flushes into G slots
    }
    const bool bc = sp.with_bc, compat = sp.compat;
    int dl = 0;  // current frontier distance
    for (;;) {
        if (bc && (size_t)(dl + 2) > ctx->F_levels) TRY(grow_levels(ctx, N, dl + 1));
        const uint64_t* Fd = ctx->F.p + N * (size_t)(bc ? dl : planes * (dl & 1));
        uint64_t* Fn = ctx->F.p + N * (size_t)(bc ? dl + 1 : planes * ((dl + 1) & 1));
        if (!bc || dl + 1 > 1)
            for (int b = 0; b < nb; ++b)
                HIPTRY(hipMemsetAsync(Fn + (size_t)b * N + v.lo, 0, sizeof(uint64_t) * R, s));
        HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int), s));
        FwdArgs fa{};
        fa.bstride = N; fa.pslots = G; fa.A_last = A_last; fa.dist_atomic = nb > 1;
        fa.n = d.n; fa.hub_deg = d.hub_deg; fa.nchunk = v.nchunk; fa.chunk_beg = d.chunk_beg + v.c0;
        fa.rp = d.rp; fa.col = d.col; fa.vis = ctx->vis.p; fa.Fd = Fd; fa.Fn = Fn;
        fa.sigma = bc ? ctx->sigma.p : nullptr;
        fa.npred = compat ? ctx->npred.p : nullptr;
        fa.dist_sum = sp.dist_sum;
        fa.weight = sp.weight;
        fa.A = A; fa.dnext = dl + 1;
        fa.part_far = ctx->part_far.p; fa.part_cnt = ctx->part_cnt.p;
        fa.flag = ctx->flags.p; fa.overflow = ctx->flags.p + 1;
        hipEvent_t e0;
        TRY(timer_begin(ctx, tm, &e0));
        if (bc) {
            if (compat) k_fwd<true, true><<<G, kBlock, 0, s>>>(fa);
            else k_fwd<true, false><<<G, kBlock, 0, s>>>(fa);
        } else {
            k_fwd<false, false><<<dim3(G, nb), kBlock, 0, s>>>(fa);
        }
        HIPTRY(hipGetLastError());
        if (v.nhub) {
            if (bc) {
                k_fwd_hubseg<true><<<Gseg, kBlock, 0, s>>>(fa, ha);
                if (compat) k_fwd_hubfin<true, true><<<Ghub, kBlock, 0, s>>>(fa, ha);
                else k_fwd_hubfin<true, false><<<Ghub, kBlock, 0, s>>>(fa, ha);
            } else {
                k_fwd_hubseg<false><<<dim3(Gseg, nb), kBlock, 0, s>>>(fa, ha);
                k_fwd_hubfin<false, false><<<dim3(Ghub, nb), kBlock, 0, s>>>(fa, ha);
            }
            HIPTRY(hipGetLastError());
        }
        TRY(timer_end(ctx, tm, e0, true));
        ctx->stats.launches_fwd++;
        HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->flags.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (ctx->h_flags[1]) return CENT_E_OVERFLOW;
        if (!ctx->h_flags[0]) break;
        ++dl;
    }
    *levels += (int64_t)dl * nb;  // deepest non-empty level, per batch
    if (bc) {
        for (int lv = dl; lv >= 1; --lv) {
            BwdArgs ba;
            ba.n = d.n; ba.hub_deg = d.hub_deg; ba.nchunk = v.nchunk; ba.chunk_beg = d.chunk_beg + v.c0;
            ba.rp = d.rp; ba.col = d.col;
            ba.Fd = ctx->F.p + N * (size_t)lv;
            ba.Fn = ctx->F.p + N * (size_t)(lv + 1);
            ba.sigma = ctx->sigma.p; ba.coef = ctx->coef.p;
            ba.npred = compat ? ctx->npred.p : nullptr;
            ba.bc = sp.bc;
            ba.wgt = sp.bwgt;
            ba.bsrc = d_src;
            hipEvent_t e0;
            TRY(timer_begin(ctx, tm, &e0));
            if (compat) k_bwd<true><<<G, kBlock, 0, s>>>(ba);
            else k_bwd<false><<<G, kBlock, 0, s>>>(ba);
            HIPTRY(hipGetLastError());
            if (v.nhub) {
                k_bwd_hubseg<<<Gseg, kBlock, 0, s>>>(ba, ha);
                if (compat) k_bwd_hubfin<true><<<Ghub, kBlock, 0, s>>>(ba, ha);
                else k_bwd_hubfin<false><<<Ghub, kBlock, 0, s>>>(ba, ha);
                HIPTRY(hipGetLastError());
            }
            TRY(timer_end(ctx, tm, e0, false));
            ctx->stats.launches_bwd++;
        }
    }
    for (int b = 0; b < nb; ++b)
        TRY(reduce_batch(ctx, G, b == nb - 1 ? kl : kWave, far_out + b * kWave, b));
    ctx->stats.batches += nb;
    return CENT_OK;
}


// ------------------------------------------------------------ frontier-queue mode
constexpr int kSpReplay = 16;  // levels per hipGraph replay

struct SpPlan {
    SpArgs a{};
    int G = 1;                 // blocks of the list kernels
    hipGraph_t gf = nullptr, gb = nullptr;
    hipGraphExec_t xf = nullptr, xb = nullptr;
    ~SpPlan() {
        if (xf) (void)hipGraphExecDestroy(xf);
        if (xb) (void)hipGraphExecDestroy(xb);
        if (gf) (void)hipGraphDestroy(gf);
        if (gb) (void)hipGraphDestroy(gb);
    }
};

static int sp_prepare(cent_ctx* ctx, const DevCSR& d, const BatchSpec& sp, SpPlan& pl) {
    const size_t N = (size_t)(d.n > 0 ? d.n : 1);
    pl.G = std::max(1, std::min((int)((N + kBlock - 1) / kBlock), 1024));
    TRY(ensure(ctx, ctx->vis, N));
    TRY(ensure(ctx, ctx->Fd2, 2 * N));
    const size_t qcap = std::min<size_t>(N * kWave, (size_t)INT32_MAX);
    TRY(ensure(ctx, ctx->Q, qcap));
    TRY(ensure(ctx, ctx->QF, qcap));
    TRY(ensure(ctx, ctx->qoff, N + 3));
    TRY(ensure(ctx, ctx->cand, N));
    TRY(ensure(ctx, ctx->stamp, N));
    TRY(ensure(ctx, ctx->ctl, kCtlSlots));
    TRY(ensure(ctx, ctx->Adev, 1));
    TRY(ensure(ctx, ctx->part_far, (size_t)pl.G * kWave));
    TRY(ensure(ctx, ctx->part_cnt, (size_t)pl.G * 2));
    TRY(ensure(ctx, ctx->cnt, 2));
    if (sp.with_bc) {
        TRY(ensure(ctx, ctx->sigma, N * kWave));
        TRY(ensure(ctx, ctx->coef, N * kWave));
        if (sp.compat) TRY(ensure(ctx, ctx->npred, N * kWave));
    }
    hipStream_t s = ctx->stream;
    // all-zero vis / masks between batches (batches clear what they touched)
    HIPTRY(hipMemsetAsync(ctx->vis.p, 0, sizeof(uint64_t) * N, s));
    HIPTRY(hipMemsetAsync(ctx->Fd2.p, 0, sizeof(uint64_t) * 2 * N, s));
    // stamps restart with the call's level counter (ctl), so reset them too
    HIPTRY(hipMemsetAsync(ctx->stamp.p, 0xff, sizeof(int) * N, s));
    HIPTRY(hipMemsetAsync(ctx->ctl.p, 0, sizeof(int) * kCtlSlots, s));
    SpArgs& a = pl.a;
    a.n = d.n; a.rp = d.rp; a.col = d.col; a.vis = ctx->vis.p; a.Fd2 = ctx->Fd2.p;
    a.sigma = sp.with_bc ? ctx->sigma.p : nullptr; a.coef = sp.with_bc ? ctx->coef.p : nullptr;
    a.npred = sp.compat ? ctx->npred.p : nullptr; a.bc = sp.bc; a.dist_sum = sp.dist_sum;
    a.weight = sp.weight; a.Q = ctx->Q.p; a.QF = ctx->QF.p; a.qoff = ctx->qoff.p;
    a.qcap = (int)qcap; a.qoff_cap = (int)(N + 3); a.cand = ctx->cand.p; a.stamp = ctx->stamp.p;
    a.ctl = ctx->ctl.p; a.Aptr = ctx->Adev.p; a.part_far = ctx->part_far.p; a.part_cnt = ctx->part_cnt.p;
    // capture kSpReplay levels of the forward and of the backward once per call
    const int G = pl.G;
    HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    for (int l = 0; l < kSpReplay; ++l) {
        k_sp_expand<<<G, kBlock, 0, s>>>(a);
        if (sp.with_bc) {
            if (sp.compat) k_sp_pull<true, true><<<G, kBlock, 0, s>>>(a);
            else k_sp_pull<true, false><<<G, kBlock, 0, s>>>(a);
        } else {
            k_sp_pull<false, false><<<G, kBlock, 0, s>>>(a);
        }
        k_sp_advance<<<1, 1, 0, s>>>(a);
    }
    HIPTRY(hipStreamEndCapture(s, &pl.gf));
    HIPTRY(hipGraphInstantiate(&pl.xf, pl.gf, nullptr, nullptr, 0));
    if (sp.with_bc) {
        HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        for (int l = 0; l < kSpReplay; ++l) {
            if (sp.compat) k_sp_bwd<true><<<G, kBlock, 0, s>>>(a);
            else k_sp_bwd<false><<<G, kBlock, 0, s>>>(a);
            k_sp_bwd_next<<<G, kBlock, 0, s>>>(a);
            k_sp_bwd_step<<<1, 1, 0, s>>>(a);
        }
        HIPTRY(hipStreamEndCapture(s, &pl.gb));
        HIPTRY(hipGraphInstantiate(&pl.xb, pl.gb, nullptr, nullptr, 0));
    }
    return CENT_OK;
}

// One batch of kb <= 64 sources in frontier-queue mode (whole graph).
static int level_batch_sparse(cent_ctx* ctx, SpPlan& pl, const BatchSpec& sp, const int* d_src, int kb,
                              int64_t* far_out, KernelTimer& tm, int64_t* levels) {
    hipStream_t s = ctx->stream;
    const int G = pl.G;
    const uint64_t A = (kb == 64) ? ~0ull : ((1ull << kb) - 1);
    HIPTRY(hipMemsetAsync(ctx->part_far.p, 0, sizeof(int64_t) * (size_t)G * kWave, s));
    HIPTRY(hipMemsetAsync(ctx->part_cnt.p, 0, sizeof(unsigned long long) * (size_t)G * 2, s));
    k_sp_seed<<<1, kWave, 0, s>>>(pl.a, d_src, kb, sp.with_bc, A);
    HIPTRY(hipGetLastError());
    int deepest = 0;
    for (;;) {
        hipEvent_t e0;
        TRY(timer_begin(ctx, tm, &e0));
        HIPTRY(hipGraphLaunch(pl.xf, s));
        TRY(timer_end(ctx, tm, e0, true));
        ctx->stats.launches_fwd += kSpReplay;
        HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->ctl.p + kCtlDone, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (ctx->h_flags[0]) break;
    }
    int ctl_h[kCtlSlots];
    HIPTRY(hipMemcpyAsync(ctl_h, ctx->ctl.p, sizeof(ctl_h), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (ctl_h[kCtlOverflow]) return CENT_E_OVERFLOW;
    deepest = ctl_h[kCtlDeepest];
    *levels += deepest;
    if (sp.with_bc && deepest >= 1) {
        k_sp_bwd_init<<<1, 1, 0, s>>>(pl.a);
        const int replays = (deepest + kSpReplay - 1) / kSpReplay;
        for (int r = 0; r < replays; ++r) {
            hipEvent_t e0;
            TRY(timer_begin(ctx, tm, &e0));
            HIPTRY(hipGraphLaunch(pl.xb, s));
            TRY(timer_end(ctx, tm, e0, false));
            ctx->stats.launches_bwd += kSpReplay;
        }
    }
    k_sp_clear<<<G, kBlock, 0, s>>>(pl.a);
    TRY(reduce_batch(ctx, G, kb, far_out));
    ctx->stats.batches++;
    return CENT_OK;
}

// ------------------------------------------------------------ source-parallel mode
// All k sources of a plain call in one launch of k_ps: G workgroup slots,
// each with its own per-vertex arrays, take sources s, s + G, ... in turn.
// G is bounded by k, by the slot cap (opts->max_slots, else grid_waves / 8,
// else 256: more concurrent traversals than that only contend for the
// memory system) and by a memory budget (40% of free device memory, at most
// 32 GiB).  Returns CENT_E_UNSUPPORTED, with nothing accumulated into the BC
// output, if some source's BFS is deeper than the level table: the caller
// then runs the batch form.
constexpr int kPsDefaultSlots = 256;

static int run_persource(cent_ctx* ctx, const cent_graph* g, const BatchSpec& sp, int32_t k, KernelTimer& tm,
                         int64_t* levels) {
    const DevCSR& d = g->d;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)(d.n > 0 ? d.n : 1);
    const bool bc = sp.with_bc, compat = sp.compat;
    if (sp.bwgt || k < 1) return CENT_E_UNSUPPORTED;
    const int lcap = (int)std::min<size_t>(n + 2, 65536);
    const size_t per_slot = n * (sizeof(PsRec) + 4 + (bc ? 8 : 0)) + (size_t)lcap * 4;
    size_t cap = kPsDefaultSlots;
    if (sp.opts && sp.opts->max_slots > 0) cap = (size_t)sp.opts->max_slots;
    else if (sp.opts && sp.opts->grid_waves > 0) cap = std::max(1, sp.opts->grid_waves / kPsWaves);
    size_t G = std::min((size_t)k, cap);
    size_t freeb = 0, totalb = 0;
    HIPTRY(hipMemGetInfo(&freeb, &totalb));
    const size_t have = sizeof(PsRec) * ctx->ps_rec.cap + 4 * (ctx->ps_Q.cap + ctx->ps_loff.cap) +
                        8 * ctx->ps_bcp.cap;
    const size_t budget = std::min<size_t>((size_t)32 << 30, (freeb + have) / 5 * 2);
    G = std::max<size_t>(1, std::min(G, budget / per_slot));
    TRY(ensure_exact(ctx, ctx->ps_rec, n * G));
    TRY(ensure_exact(ctx, ctx->ps_Q, n * G));
    TRY(ensure_exact(ctx, ctx->ps_loff, (size_t)lcap * G));
    if (bc) TRY(ensure_exact(ctx, ctx->ps_bcp, n * G));
    TRY(ensure(ctx, ctx->ps_depth, 1));
    TRY(ensure(ctx, ctx->flags, 2));
    const size_t total = n * G;
    const int ig = (int)std::min<size_t>((total + 255) / 256, 16384);
    k_ps_init<<<ig, 256, 0, s>>>(ctx->ps_rec.p, bc ? ctx->ps_bcp.p : nullptr, total);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemsetAsync(ctx->ps_depth.p, 0, sizeof(unsigned long long), s));
    HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
    PsArgs a{};
    a.n = d.n; a.rp = d.rp; a.col = d.col; a.src = ctx->src.p; a.k = k;
    a.rec = ctx->ps_rec.p; a.Q = ctx->ps_Q.p;
    a.loff = ctx->ps_loff.p; a.lcap = lcap;
    a.bcp = bc ? ctx->ps_bcp.p : nullptr;
    a.far = ctx->far.p; a.dist_sum = sp.dist_sum; a.weight = sp.weight;
    a.cnt = ctx->cnt.p; a.depth_sum = ctx->ps_depth.p; a.overflow = ctx->flags.p;
    hipEvent_t e0;
    TRY(timer_begin(ctx, tm, &e0));
    if (!bc) k_ps<false, false><<<(int)G, kPsBlock, 0, s>>>(a);
    else if (compat) k_ps<true, true><<<(int)G, kPsBlock, 0, s>>>(a);
    else k_ps<true, false><<<(int)G, kPsBlock, 0, s>>>(a);
    HIPTRY(hipGetLastError());
    TRY(timer_end(ctx, tm, e0, true));
    ctx->stats.launches_fwd++;
    unsigned long long depth = 0;
    HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->flags.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(&depth, ctx->ps_depth.p, sizeof(depth), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (ctx->h_flags[0] & 2) return CENT_E_UNSUPPORTED;
    if (ctx->h_flags[0] & 1) return CENT_E_OVERFLOW;
    if (bc) {
        k_ps_reduce<<<(int)((n + 255) / 256), 256, 0, s>>>(ctx->ps_bcp.p, (int)G, d.n, sp.bc);
        HIPTRY(hipGetLastError());
    }
    *levels += (int64_t)depth;
    ctx->stats.batches += k;
    return CENT_OK;
}

// Runs all batches of one plain call.  with_bc: Brandes (sigma + backward);
// otherwise closeness only (bitmask BFS).
static int run_batches(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, bool with_bc,
                       int bc_mode, double* bc, int64_t* farness, int64_t* dist_sum, const cent_opts* opts) {
    const int n = g->n;
    const size_t N = (size_t)(n > 0 ? n : 1);
    hipStream_t s = ctx->stream;
    // the graph's tuned launch settings fill what the call leaves open
    cent_opts eff{};
    if (opts) eff = *opts;
    if (eff.grid_waves == 0) eff.grid_waves = g->grid_waves[with_bc ? 0 : 1];
    if (eff.frontier == CENT_FRONTIER_AUTO) eff.frontier = g->frontier[with_bc ? 0 : 1];
    opts = &eff;
    BatchSpec sp;
    sp.with_bc = with_bc;
    sp.compat = with_bc && bc_mode == CENT_BC_REF_COMPAT;
    sp.opts = opts;
    if (with_bc) {
        TRY(ensure(ctx, ctx->bc, N));
        sp.bc = ctx->bc.p;
        HIPTRY(hipMemsetAsync(ctx->bc.p, 0, sizeof(double) * N, s));
    }
    if (dist_sum) {
        TRY(ensure(ctx, ctx->dist_sum, N));
        sp.dist_sum = ctx->dist_sum.p;
        HIPTRY(hipMemsetAsync(ctx->dist_sum.p, 0, sizeof(int64_t) * N, s));
    }
    TRY(ensure(ctx, ctx->far, (size_t)(k > 0 ? k : 1)));
    TRY(ensure(ctx, ctx->src, (size_t)(k > 0 ? k : 1)));
    TRY(ensure(ctx, ctx->cnt, 2));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
    // Batch order of the sources: the 64-lane steps group sources by BFS
    // distance to the highest-degree vertex, then by degree, so the lanes
    // of a batch tend to reach a vertex at the same levels (fewer
    // (vertex, level) rows touched).  order[i] = caller index of slot i.
    const int lm = k > 0 ? level_mode(g, opts) : CENT_FRONTIER_DENSE;
    std::vector<int32_t> order;
    if (k > kWave && lm != CENT_FRONTIER_SOURCE && !(opts->flags & CENT_OPT_SOURCE_ORDER) &&
        !g->probe_dist.empty()) {
        order.resize((size_t)k);
        for (int32_t i = 0; i < k; ++i) order[(size_t)i] = i;
        auto deg = [&](int32_t v) {
            const int32_t x = g->new_id.empty() ? v : g->new_id[(size_t)v];
            return g->h_rp[(size_t)x + 1] - g->h_rp[(size_t)x];
        };
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            const int32_t sa = src[a], sb = src[b];
            const int32_t da = g->probe_dist[(size_t)sa], db = g->probe_dist[(size_t)sb];
            if (da != db) return da < db;
            const int32_t ga = deg(sa), gb = deg(sb);
            if (ga != gb) return ga > gb;
            return sa < sb;
        });
    }
    std::vector<int32_t> dsrc;
    const int32_t* up_src = src;
    if ((!g->new_id.empty() || !order.empty()) && k > 0) {
        dsrc.resize((size_t)k);
        for (int32_t i = 0; i < k; ++i) {
            const int32_t s0 = order.empty() ? src[i] : src[order[(size_t)i]];
            dsrc[(size_t)i] = g->new_id.empty() ? s0 : g->new_id[(size_t)s0];
        }
        up_src = dsrc.data();
    }
    if (k > 0) HIPTRY(hipMemcpyAsync(ctx->src.p, up_src, sizeof(int) * (size_t)k, hipMemcpyHostToDevice, s));
    HIPTRY(hipStreamSynchronize(s));  // dsrc is a pageable host buffer

    KernelTimer tm;
    tm.on = opts && (opts->flags & CENT_OPT_TIMING);
    ctx->ev_used = 0;
    const GView view = full_view(g->d);
    int64_t levels = 0;
    const int lmode = lm;
    bool batched = true;
    if (lmode == CENT_FRONTIER_SOURCE) {
        const int st = run_persource(ctx, g, sp, k, tm, &levels);
        if (st == CENT_OK) {
            batched = false;
        } else if (st == CENT_E_UNSUPPORTED) {  // too deep for the level table: redo as batches
            HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
            if (sp.dist_sum) HIPTRY(hipMemsetAsync(sp.dist_sum, 0, sizeof(int64_t) * N, s));
        } else {
            return st;
        }
    }
    const bool sparse = batched && lmode == CENT_FRONTIER_SPARSE;
    SpPlan plan;
    if (sparse) TRY(sp_prepare(ctx, g->d, sp, plan));
    // closeness on the dense step takes up to kMultiBatch batches per launch
    const int32_t step = (!with_bc && !sparse) ? kWave * kMultiBatch : kWave;
    for (int32_t b0 = 0; batched && b0 < k; b0 += step) {
        const int kb = (k - b0) < step ? (k - b0) : step;
        if (sparse) TRY(level_batch_sparse(ctx, plan, sp, ctx->src.p + b0, kb, ctx->far.p + b0, tm, &levels));
        else TRY(level_batch(ctx, view, sp, ctx->src.p + b0, kb, ctx->far.p + b0, tm, &levels));
        if (tm.on && ctx->ev_used > 4096) {  // keep the event pool bounded
            HIPTRY(hipStreamSynchronize(s));
            timer_collect(ctx, tm);
        }
    }
    // results back to the host
    std::vector<double> hb;
    std::vector<int64_t> hd;
    unsigned long long hc[2] = {0, 0};
    if (with_bc && bc) {
        hb.resize(N);
        HIPTRY(hipMemcpyAsync(hb.data(), ctx->bc.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    }
    if (dist_sum) {
        hd.resize(N);
        HIPTRY(hipMemcpyAsync(hd.data(), ctx->dist_sum.p, sizeof(int64_t) * N, hipMemcpyDeviceToHost, s));
    }
    std::vector<int64_t> hf;
    if (farness && k > 0) {
        if (order.empty()) {
            HIPTRY(hipMemcpyAsync(farness, ctx->far.p, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToHost, s));
        } else {
            hf.resize((size_t)k);
            HIPTRY(hipMemcpyAsync(hf.data(), ctx->far.p, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToHost, s));
        }
    }
    HIPTRY(hipMemcpyAsync(hc, ctx->cnt.p, sizeof(hc), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (!hf.empty())
        for (int32_t i = 0; i < k; ++i) farness[order[(size_t)i]] = hf[(size_t)i];
    const bool remap = !g->new_id.empty();
    if (with_bc && bc)
        for (int v = 0; v < n; ++v) bc[v] += hb[remap ? (size_t)g->new_id[(size_t)v] : (size_t)v];
    if (dist_sum)
        for (int v = 0; v < n; ++v) dist_sum[v] += hd[remap ? (size_t)g->new_id[(size_t)v] : (size_t)v];
    timer_collect(ctx, tm);
    ctx->stats.levels += levels;
    ctx->stats.reached += (int64_t)hc[0] + k;
    ctx->stats.dag_edges += (int64_t)hc[1];
    for (int32_t i = 0; i < k; ++i) ctx->stats.edges_reached += g->comp_nnz[(size_t)g->comp[(size_t)src[i]]];
    return CENT_OK;
}

static int prepare_sources(const cent_graph* g, const int32_t* src, int32_t k, std::vector<int32_t>& all,
                           const int32_t** out_src, int32_t* out_k) {
    if (k == -1) {
        all.resize((size_t)g->n);
        for (int32_t v = 0; v < g->n; ++v) all[(size_t)v] = v;
        *out_src = all.data();
        *out_k = g->n;
        return CENT_OK;
    }
    if (k < 0 || (k > 0 && !src)) return CENT_E_INVAL;
    for (int32_t i = 0; i < k; ++i)
        if (src[i] < 0 || src[i] >= g->n) return CENT_E_INVAL;
    *out_src = src;
    *out_k = k;
    return CENT_OK;
}

extern "C" int cent_bc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, double* bc,
                       int64_t* farness, int64_t* dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->ctx->device != ctx->device) return CENT_E_INVAL;
    int mode = opts ? opts->bc_mode : CENT_BC_TEXTBOOK;
    if (mode != CENT_BC_TEXTBOOK && mode != CENT_BC_REF_COMPAT) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    TRY(prepare_sources(g, src, k, all, &s, &kk));
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    int st = run_batches(ctx, g, s, kk, true, mode, bc, farness, dist_sum, opts);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

extern "C" int cent_cc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, int64_t* farness,
                       int64_t* dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->ctx->device != ctx->device) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    TRY(prepare_sources(g, src, k, all, &s, &kk));
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    int st = run_batches(ctx, g, s, kk, false, 0, nullptr, farness, dist_sum, opts);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

// ------------------------------------------------------------ autotuner
// Coordinate descent over the plain form's launch settings: level step
// first; the source-parallel step then only sweeps its slot count
// (grid_waves / 8) and the frontier-queue step nothing; the dense step gets
// up to two passes over hub_deg, chunk_edges and grid_waves, each swept with
// the others held at their best so far.  A trial is the faster of two
// runs of the sample (the first also warms the workspace).
namespace {
struct Tuner {
    cent_ctx* ctx;
    cent_graph* g;
    const int32_t* src;
    int32_t k;
    int32_t what;
    int32_t budget;  // remaining timed runs (< 0: unlimited)
    std::vector<double> bc;
    std::vector<int64_t> far;

    int run_once(double* ms) {
        std::fill(bc.begin(), bc.end(), 0.0);
        const auto t0 = std::chrono::steady_clock::now();
        if (what == CENT_TUNE_BC) TRY(cent_bc(ctx, g, src, k, bc.data(), far.data(), nullptr, nullptr));
        else TRY(cent_cc(ctx, g, src, k, far.data(), nullptr, nullptr));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return CENT_OK;
    }

    // Time setting p; *ms = +inf when the budget is spent.
    int trial(const cent_tune_params& p, double* ms) {
        *ms = 1e300;
        if (budget == 0) return CENT_OK;
        TRY(cent_graph_set_layout(g, what, &p));
        double a = 0.0, b = 0.0;
        TRY(run_once(&a));
        TRY(run_once(&b));
        if (budget > 0) --budget;
        *ms = a < b ? a : b;
        return CENT_OK;
    }
};
}  // namespace

extern "C" int cent_tune(cent_ctx* ctx, cent_graph* g, const int32_t* src, int32_t k, int32_t what,
                         int32_t max_trials, cent_tune_params* best) {
    if (!ctx || !g || !best || g->ctx->device != ctx->device) return CENT_E_INVAL;
    if (what != CENT_TUNE_BC && what != CENT_TUNE_CC) return CENT_E_INVAL;
    if (k < 1 || k > 256 || !src || max_trials < 0) return CENT_E_INVAL;
    for (int32_t i = 0; i < k; ++i)
        if (src[i] < 0 || src[i] >= g->n) return CENT_E_INVAL;
    const cent_stats saved = ctx->stats;
    Tuner t{ctx, g, src, k, what, max_trials > 0 ? max_trials : -1, {}, {}};
    t.bc.assign((size_t)(g->n > 0 ? g->n : 1), 0.0);
    t.far.assign((size_t)k, 0);
    cent_tune_params orig, cur;
    TRY(cent_graph_get_layout(g, what, &orig));
    cur = orig;
    cur.frontier = CENT_FRONTIER_DENSE;
    double best_ms = 1e300, ms = 0.0;
    int st = CENT_OK;
    auto sweep = [&](int32_t cent_tune_params::*field, std::initializer_list<int32_t> vals) {
        cent_tune_params keep = cur;
        for (int32_t v : vals) {
            if (st) return;
            cent_tune_params p = cur;
            p.*field = v;
            if (p.*field == keep.*field && best_ms < 1e300) continue;  // already timed
            st = t.trial(p, &ms);
            if (!st && ms < best_ms) {
                best_ms = ms;
                keep = p;
            }
        }
        cur = keep;
    };
    sweep(&cent_tune_params::frontier, {CENT_FRONTIER_DENSE, CENT_FRONTIER_SPARSE, CENT_FRONTIER_SOURCE});
    if (!st && cur.frontier == CENT_FRONTIER_SOURCE) {
        // grid_waves / 8 = concurrent source slots; caps above the sample size
        // time the same as the sample size, so only smaller ones are tried
        for (int32_t gw : {kPsWaves * 64, kPsWaves * 128})
            if (gw / kPsWaves <= k) sweep(&cent_tune_params::grid_waves, {gw});
    }
    for (int pass = 0; pass < 2 && !st && cur.frontier == CENT_FRONTIER_DENSE; ++pass) {
        const cent_tune_params before = cur;
        sweep(&cent_tune_params::hub_deg, {128, 256, 512, 1024, 2048, 4096});
        sweep(&cent_tune_params::chunk_edges, {256, 512, 1024, 2048, 4096});
        sweep(&cent_tune_params::grid_waves, {4096, 8192, 16384, 32768});
        if (before.hub_deg == cur.hub_deg && before.chunk_edges == cur.chunk_edges &&
            before.grid_waves == cur.grid_waves)
            break;  // a pass that moves nothing is a fixed point
    }
    ctx->stats = saved;
    if (!st && best_ms >= 1e300) st = CENT_E_INVAL;  // budget too small to time anything
    if (st) {
        (void)cent_graph_set_layout(g, what, &orig);  // leave the graph as it came
        return st;
    }
    cur.ms = best_ms;
    TRY(cent_graph_set_layout(g, what, &cur));
    TRY(cent_tune_table_put(&g->sig, what, &cur));
    *best = cur;
    return CENT_OK;
}

// ------------------------------------------------------------ reduced closeness
// Device copy of a packed block graph, cached on the cent_reduced object
// (freed through its dev_free hook).
struct RedDev {
    cent_ctx* ctx = nullptr;
    DevCSR d;
    int64_t* w = nullptr;     // [np]
    int* iota = nullptr;      // [np] packed ids 0..np-1 (batch sources)
    int4* items = nullptr;    // [nitems]
    int* order = nullptr;     // [nsmall] small items, largest range first
    int nsmall = 0;
    std::vector<int> big;     // items handled by the level-synchronous kernels
    int64_t* poff = nullptr;  // [nitems] offsets of the items' BC partials (small items)
    int64_t part_total = 0;
    std::vector<int64_t> h_poff;
};

static void reddev_free(void* p) {
    RedDev* rd = (RedDev*)p;
    if (!rd) return;
    if (rd->ctx) (void)hipSetDevice(rd->ctx->device);
    devcsr_free(rd->d);
    if (rd->w) (void)hipFree(rd->w);
    if (rd->iota) (void)hipFree(rd->iota);
    if (rd->items) (void)hipFree(rd->items);
    if (rd->order) (void)hipFree(rd->order);
    if (rd->poff) (void)hipFree(rd->poff);

    delete rd;
}

static std::mutex g_reddev_mu;  // first-use creation of a reduced graph's device copy

static int reddev_get(cent_ctx* ctx, cent_reduced* r, RedDev** out) {
    std::lock_guard<std::mutex> lk(g_reddev_mu);
    if (r->dev) {
        RedDev* rd = (RedDev*)r->dev;
        if (rd->ctx->device != ctx->device) return CENT_E_INVAL;
        *out = rd;
        return CENT_OK;
    }
    RedDev* rd = new (std::nothrow) RedDev();
    if (!rd) return CENT_E_OOM;
    rd->ctx = ctx;
    r->dev = rd;
    r->dev_free = reddev_free;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    TRY(devcsr_build(ctx, r->prp, r->pcol, r->np, r->nnzp, r->boff, r->nblocks, rd->d));
    TRY(upload(ctx, &rd->w, r->pw, (size_t)r->np));
    std::vector<int> iota(NP);
    for (size_t i = 0; i < NP; ++i) iota[i] = (int)i;
    TRY(upload(ctx, &rd->iota, iota.data(), NP));
    std::vector<int4> items((size_t)(r->nitems > 0 ? r->nitems : 1));
    std::vector<int> small;
    for (int i = 0; i < r->nitems; ++i) {
        const cent_item& it = r->items[i];
        items[(size_t)i] = make_int4(it.lo, it.hi, it.s0, it.ns);
        if (it.hi - it.lo <= kItemLds) small.push_back(i);
        else rd->big.push_back(i);
    }
    std::stable_sort(small.begin(), small.end(), [&](int a, int b) {
        return (r->items[a].hi - r->items[a].lo) > (r->items[b].hi - r->items[b].lo);
    });
    rd->nsmall = (int)small.size();
    TRY(upload(ctx, &rd->items, items.data(), (size_t)r->nitems));
    TRY(upload(ctx, &rd->order, small.data(), small.size()));
    rd->h_poff.assign((size_t)(r->nitems > 0 ? r->nitems : 1), -1);
    for (int i : small) {
        rd->h_poff[(size_t)i] = rd->part_total;
        rd->part_total += r->items[i].hi - r->items[i].lo;
    }
    TRY(upload(ctx, &rd->poff, rd->h_poff.data(), rd->h_poff.size()));
    *out = rd;
    return CENT_OK;
}

extern "C" int cent_cc_reduced(cent_ctx* ctx, const cent_reduced* rc, int64_t* farness, int32_t sched,
                               const cent_opts* opts) {
    if (!ctx || !rc || !farness) return CENT_E_INVAL;
    if (sched != CENT_SCHED_PERSISTENT && sched != CENT_SCHED_PER_ITEM) return CENT_E_INVAL;
    cent_reduced* r = const_cast<cent_reduced*>(rc);  // the device cache lives on the object
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    RedDev* rd;
    TRY(reddev_get(ctx, r, &rd));
    hipStream_t s = ctx->stream;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    // per-run buffers live in the context, so several contexts can share rd
    TRY(ensure(ctx, ctx->red_D, NP));
    TRY(ensure(ctx, ctx->red_head, 1));
    int64_t* Dd = ctx->red_D.p;
    HIPTRY(hipMemsetAsync(Dd, 0, sizeof(int64_t) * NP, s));
    TRY(ensure(ctx, ctx->cnt, 2));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
    KernelTimer tm;
    tm.on = opts && (opts->flags & CENT_OPT_TIMING);
    ctx->ev_used = 0;
    BatchSpec sp;
    sp.weight = rd->w;
    sp.opts = opts;
    int64_t levels = 0;
    auto run_item = [&](int i) -> int {
        const cent_item& it = r->items[i];
        const GView v = sub_view(rd->d, it.lo, it.hi);
        return level_batch(ctx, v, sp, rd->iota + it.s0, it.ns, Dd + it.s0, tm, &levels);
    };
    if (sched == CENT_SCHED_PER_ITEM) {
        for (int i = 0; i < r->nitems; ++i) TRY(run_item(i));
    } else {
        if (rd->nsmall) {
            HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int), s));
            ItemArgs ia;
            ia.rp = rd->d.rp; ia.col = rd->d.col; ia.weight = rd->w; ia.items = rd->items; ia.order = rd->order;
            ia.nitems = rd->nsmall; ia.head = ctx->red_head.p; ia.D = Dd;
            const int grid = std::min(rd->nsmall, 2048);
            hipEvent_t e0;
            TRY(timer_begin(ctx, tm, &e0));
            k_cc_items<<<grid, kItemBlock, 0, s>>>(ia);
            HIPTRY(hipGetLastError());
            TRY(timer_end(ctx, tm, e0, true));
            ctx->stats.launches_fwd++;
            ctx->stats.batches += rd->nsmall;
        }
        // large items: consecutive full items of one block (contiguous
        // sources) go up to kMultiBatch at a time through one level loop
        for (size_t j = 0; j < rd->big.size();) {
            const cent_item& it = r->items[rd->big[j]];
            int ns = it.ns;
            size_t e = j + 1;
            while (e < rd->big.size() && e - j < (size_t)kMultiBatch) {
                const cent_item& pv = r->items[rd->big[e - 1]];
                const cent_item& nx = r->items[rd->big[e]];
                if (pv.ns != kWave || nx.lo != it.lo || nx.hi != it.hi || nx.s0 != pv.s0 + pv.ns) break;
                ns += nx.ns;
                ++e;
            }
            const GView v = sub_view(rd->d, it.lo, it.hi);
            TRY(level_batch(ctx, v, sp, rd->iota + it.s0, ns, Dd + it.s0, tm, &levels));
            j = e;
        }
    }
    std::vector<int64_t> D(NP);
    HIPTRY(hipMemcpyAsync(D.data(), Dd, sizeof(int64_t) * NP, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    timer_collect(ctx, tm);
    TRY(cent_reduced_finish(r, D.data(), farness));
    ctx->stats.levels += levels;
    ctx->stats.reached += (int64_t)r->n;  // every vertex is a source
    ctx->stats.edges_reached += 2 * r->m;
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CENT_OK;
}


// Betweenness (+ closeness) through the reduction: weighted Brandes from every
// packed vertex inside its block, then the host adds the piece-size terms.
constexpr int kBcItemGrid = 512;

extern "C" int cent_bc_reduced(cent_ctx* ctx, const cent_reduced* rc, double* bc, int64_t* farness,
                               int32_t sched, const cent_opts* opts) {
    if (!ctx || !rc || !bc) return CENT_E_INVAL;
    if (sched != CENT_SCHED_PERSISTENT && sched != CENT_SCHED_PER_ITEM) return CENT_E_INVAL;
    cent_reduced* r = const_cast<cent_reduced*>(rc);
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    RedDev* rd;
    TRY(reddev_get(ctx, r, &rd));
    hipStream_t s = ctx->stream;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    TRY(ensure(ctx, ctx->red_D, NP));
    TRY(ensure(ctx, ctx->red_head, 2));
    TRY(ensure(ctx, ctx->red_bcp, NP));
    TRY(ensure(ctx, ctx->red_deep, (size_t)(r->nitems > 0 ? r->nitems : 1)));
    TRY(ensure(ctx, ctx->cnt, 2));
    int64_t* Dd = ctx->red_D.p;
    HIPTRY(hipMemsetAsync(Dd, 0, sizeof(int64_t) * NP, s));
    HIPTRY(hipMemsetAsync(ctx->red_bcp.p, 0, sizeof(double) * NP, s));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
    KernelTimer tm;
    tm.on = opts && (opts->flags & CENT_OPT_TIMING);
    ctx->ev_used = 0;
    BatchSpec sp;
    sp.with_bc = true;
    sp.weight = rd->w;
    sp.bwgt = rd->w;
    sp.bc = ctx->red_bcp.p;
    sp.opts = opts;
    int64_t levels = 0;
    auto run_item = [&](int i) -> int {
        const cent_item& it = r->items[i];
        const GView v = sub_view(rd->d, it.lo, it.hi);
        return level_batch(ctx, v, sp, rd->iota + it.s0, it.ns, Dd + it.s0, tm, &levels);
    };
    std::vector<int> deep;
    const bool persistent = (sched == CENT_SCHED_PERSISTENT) && rd->nsmall > 0;
    if (persistent) {
        const int grid = std::min(rd->nsmall, kBcItemGrid);
        const size_t slot = (size_t)kItemLds * kWave;
        TRY(ensure(ctx, ctx->it_sig, slot * grid));
        TRY(ensure(ctx, ctx->it_coef, slot * grid));
        TRY(ensure(ctx, ctx->it_lvl, slot * grid));
        TRY(ensure(ctx, ctx->red_part, (size_t)(rd->part_total > 0 ? rd->part_total : 1)));
        TRY(ensure(ctx, ctx->flags, 2));
        HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int) * 2, s));
        HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
        BcItemArgs ba;
        ba.rp = rd->d.rp; ba.col = rd->d.col; ba.weight = rd->w; ba.items = rd->items; ba.order = rd->order;
        ba.poff = rd->poff; ba.nitems = rd->nsmall; ba.head = ctx->red_head.p; ba.D = Dd;
        ba.bcpart = ctx->red_part.p; ba.sig = ctx->it_sig.p; ba.coef = ctx->it_coef.p; ba.lvl = ctx->it_lvl.p;
        ba.deep = ctx->red_deep.p; ba.ndeep = ctx->red_head.p + 1; ba.overflow = ctx->flags.p + 1;
        hipEvent_t e0;
        TRY(timer_begin(ctx, tm, &e0));
        k_bc_items<<<grid, kItemBlock, 0, s>>>(ba);
        HIPTRY(hipGetLastError());
        TRY(timer_end(ctx, tm, e0, true));
        ctx->stats.launches_fwd++;
        ctx->stats.batches += rd->nsmall;
        int hh[2];
        HIPTRY(hipMemcpyAsync(hh, ctx->red_head.p, sizeof(hh), hipMemcpyDeviceToHost, s));
        HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->flags.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (ctx->h_flags[1]) return CENT_E_OVERFLOW;
        if (hh[1] > 0) {
            deep.resize((size_t)hh[1]);
            HIPTRY(hipMemcpy(deep.data(), ctx->red_deep.p, sizeof(int) * (size_t)hh[1], hipMemcpyDeviceToHost));
            std::sort(deep.begin(), deep.end());
        }
        for (int i : rd->big) TRY(run_item(i));
        for (int i : deep) TRY(run_item(i));
    } else {
        for (int i = 0; i < r->nitems; ++i) TRY(run_item(i));
    }
    std::vector<int64_t> D(NP);
    std::vector<double> bcp(NP);
    std::vector<double> part;
    HIPTRY(hipMemcpyAsync(D.data(), Dd, sizeof(int64_t) * NP, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(bcp.data(), ctx->red_bcp.p, sizeof(double) * NP, hipMemcpyDeviceToHost, s));
    if (persistent && rd->part_total > 0) {
        part.resize((size_t)rd->part_total);
        HIPTRY(hipMemcpyAsync(part.data(), ctx->red_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    timer_collect(ctx, tm);
    if (persistent) {
        // add the small items' partials in item order (deep items were redone above)
        std::vector<char> is_deep((size_t)(r->nitems > 0 ? r->nitems : 1), 0);
        for (int i : deep) is_deep[(size_t)i] = 1;
        for (int i = 0; i < r->nitems; ++i) {
            const int64_t o = rd->h_poff[(size_t)i];
            if (o < 0 || is_deep[(size_t)i]) continue;
            const cent_item& it = r->items[i];
            for (int x = it.lo; x < it.hi; ++x) bcp[(size_t)x] += part[(size_t)(o + x - it.lo)];
        }
    }
    TRY(cent_reduced_finish_bc(r, bcp.data(), bc));
    if (farness) TRY(cent_reduced_finish(r, D.data(), farness));
    ctx->stats.levels += levels;
    ctx->stats.reached += (int64_t)r->n;
    ctx->stats.edges_reached += 2 * r->m;
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CENT_OK;
}

// ------------------------------------------------------------ probe
// Plain streaming copy, 16 B per lane per step: the calibration kernel for
// HBM counter readings and the measured bandwidth ceiling of the device.
__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

extern "C" int cent_probe_copy_bandwidth(cent_ctx* ctx, int64_t bytes, int32_t iters, double* gbs) {
    if (!ctx || bytes < 1024 || iters < 1 || !gbs) return CENT_E_INVAL;
    TRY(check_device(ctx));
    const size_t n16 = (size_t)bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    HIPTRY(hipMalloc((void**)&a, n16 * 16));
    if (hipMalloc((void**)&b, n16 * 16) != hipSuccess) { (void)hipFree(a); return CENT_E_OOM; }
    ctx->ev_used = 0;
    hipEvent_t e0 = next_event(ctx), e1 = next_event(ctx);
    int rc = CENT_OK;
    if (!e0 || !e1 || hipMemsetAsync(a, 1, n16 * 16, ctx->stream) != hipSuccess) rc = CENT_E_DEVICE;
    const int grid = 256 * 32;
    if (!rc) {
        k_copy_probe<<<grid, 256, 0, ctx->stream>>>(a, b, n16);  // warm
        (void)hipEventRecord(e0, ctx->stream);
        for (int i = 0; i < iters; ++i) k_copy_probe<<<grid, 256, 0, ctx->stream>>>(a, b, n16);
        (void)hipEventRecord(e1, ctx->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = CENT_E_DEVICE;
        else *gbs = 2.0 * (double)(n16 * 16) * iters / (ms * 1e-3) / 1e9;
    }
    ctx->ev_used = 0;
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

// ------------------------------------------------------------ reference shapes
static std::mutex g_default_mu;
static cent_ctx* g_default = nullptr;

static int default_ctx(cent_ctx** out) {
    if (!g_default) {
        int st = cent_ctx_create(0, &g_default);
        if (st) return st;
    }
    *out = g_default;
    return CENT_OK;
}

static int shim_run(const struct CSR* csr, bool with_bc, int mode, double* BCs, int64_t* CCs) {
    if (!csr || (with_bc && !BCs) || (!with_bc && !CCs)) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(g_default_mu);
    cent_ctx* ctx;
    TRY(default_ctx(&ctx));
    cent_graph* g = nullptr;
    TRY(cent_graph_upload(ctx, csr, &g));
    int32_t k = csr->endNodeID - csr->startNodeID + 1;
    std::vector<int32_t> src((size_t)(k > 0 ? k : 0));
    for (int32_t i = 0; i < k; ++i) src[(size_t)i] = csr->startNodeID + i;
    cent_opts o{};
    o.bc_mode = mode;
    int st;
    if (with_bc) {
        st = cent_bc(ctx, g, src.data(), k, BCs, nullptr, nullptr, &o);
    } else {
        std::vector<int64_t> far((size_t)(k > 0 ? k : 1));
        st = cent_cc(ctx, g, src.data(), k, far.data(), nullptr, &o);
        if (!st)
            for (int32_t i = 0; i < k; ++i) CCs[csr->startNodeID + i] = far[(size_t)i];
    }
    cent_graph_free(g);
    return st;
}

extern "C" int computeBC(const struct CSR* csr, double* BCs) { return shim_run(csr, true, CENT_BC_TEXTBOOK, BCs, nullptr); }
extern "C" int computeBC_compat(const struct CSR* csr, double* BCs) { return shim_run(csr, true, CENT_BC_REF_COMPAT, BCs, nullptr); }
extern "C" int computeCC(const struct CSR* csr, int64_t* CCs) { return shim_run(csr, false, 0, nullptr, CCs); }

extern "C" int compute_D1_AP_CC_shareBased_DegreeOrder(const struct CSR* csr, int64_t* CCs) {
    if (!csr || !CCs) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(g_default_mu);
    cent_ctx* ctx;
    TRY(default_ctx(&ctx));
    cent_reduced* r = nullptr;
    TRY(cent_reduce(csr, &r));
    const int32_t n = csr->csrVSize - 1;
    std::vector<int64_t> far((size_t)(n > 0 ? n : 1));
    int st = cent_cc_reduced(ctx, r, far.data(), CENT_SCHED_PERSISTENT, nullptr);
    if (!st)
        for (int32_t v = csr->startNodeID; v <= csr->endNodeID; ++v) CCs[v] = far[(size_t)v];
    cent_reduced_free(r);
    return st;
}
