// Host side of the centrality engine: contexts, device graphs, the level
// loop of the 64-source batches, and the C-ABI declared in include/cent.h.
//
// A call processes its sources in batches of 64 (one wavefront lane per
// source).  Per batch: seed, forward levels until a level reaches nothing
// (one flag read per level), then — for betweenness — backward levels from
// the deepest level to 1.  All kernels of a call run in order on the
// context's stream, so the per-vertex BC sums are accumulated in a fixed
// order and the result does not depend on timing.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cent.h"
#include "msbfs_kernels.h"

using namespace cent;

#define HIPTRY(expr)                                                           \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            set_error(ctx, e_, #expr);                                         \
            return (e_ == hipErrorOutOfMemory) ? CENT_E_OOM : CENT_E_DEVICE;   \
        }                                                                      \
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
    DevBuf<uint64_t> vis, F;
    DevBuf<double> sigma, coef, bc, hp_acc;
    DevBuf<uint32_t> npred, hp_cnt;
    DevBuf<uint64_t> hp_or;
    DevBuf<int64_t> dist_sum, far, part_far;
    DevBuf<unsigned long long> part_cnt, cnt;
    DevBuf<int> src, flags;
    int* h_flags = nullptr;  // pinned
    size_t h_flags_cap = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
};

// Rows longer than kHubDeg are scanned as kHubDeg-edge segments, one
// wavefront each (see msbfs_kernels.h).
constexpr int kHubDeg = 256;

struct cent_graph {
    cent_ctx* ctx = nullptr;
    int32_t n = 0;
    int64_t nnz = 0;
    int* rp = nullptr;
    int* col = nullptr;
    std::vector<int32_t> comp;        // component id per vertex (host)
    std::vector<int64_t> comp_nnz;    // directed edges per component (host)
    // hub segments (device): seg_v/seg_beg/seg_end [nseg], hub_v [nhub], hub_seg_off [nhub+1]
    int hub_deg = kHubDeg;
    int nseg = 0, nhub = 0;
    int* hub_buf = nullptr;           // one allocation holding the five arrays
    int *seg_v = nullptr, *seg_beg = nullptr, *seg_end = nullptr, *hub_v = nullptr, *hub_seg_off = nullptr;
};

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
    *out = ctx;
    return CENT_OK;
}

extern "C" void cent_ctx_destroy(cent_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx->vis); release(ctx->F); release(ctx->sigma); release(ctx->coef); release(ctx->bc);
    release(ctx->npred); release(ctx->dist_sum); release(ctx->far); release(ctx->part_far);
    release(ctx->part_cnt); release(ctx->cnt); release(ctx->src); release(ctx->flags);
    release(ctx->hp_acc); release(ctx->hp_cnt); release(ctx->hp_or);
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

extern "C" const char* cent_ctx_last_error(const cent_ctx* ctx) {
    return ctx ? ctx->last_error : "null context";
}

extern "C" int cent_graph_from_csr(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col,
                                   int32_t n, int64_t nnz, cent_graph** out) {
    if (!ctx || !out || n < 0 || nnz < 0 || nnz > INT32_MAX || !row_ptr || (nnz > 0 && !col))
        return CENT_E_INVAL;
    *out = nullptr;
    if (row_ptr[0] != 0 || row_ptr[n] != nnz) return CENT_E_INVAL;
    for (int32_t v = 0; v < n; ++v)
        if (row_ptr[v + 1] < row_ptr[v]) return CENT_E_INVAL;
    for (int64_t j = 0; j < nnz; ++j)
        if (col[j] < 0 || col[j] >= n) return CENT_E_INVAL;
    int st = check_device(ctx);
    if (st) return st;
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
    hipError_t e1 = hipMalloc((void**)&g->rp, sizeof(int) * ((size_t)n + 1));
    hipError_t e2 = hipMalloc((void**)&g->col, sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    if (e1 != hipSuccess || e2 != hipSuccess) {
        if (g->rp) (void)hipFree(g->rp);
        if (g->col) (void)hipFree(g->col);
        delete g;
        return CENT_E_OOM;
    }
    if (hipMemcpy(g->rp, row_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nnz > 0 && hipMemcpy(g->col, col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess)) {
        cent_graph_free(g);
        return CENT_E_DEVICE;
    }
    // hub segments
    std::vector<int> sv, sb, se, hv, ho;
    ho.push_back(0);
    for (int32_t v = 0; v < n; ++v) {
        const int b = row_ptr[v], e = row_ptr[v + 1];
        if (e - b <= g->hub_deg) continue;
        hv.push_back(v);
        for (int x = b; x < e; x += g->hub_deg) {
            sv.push_back(v);
            sb.push_back(x);
            se.push_back(x + g->hub_deg < e ? x + g->hub_deg : e);
        }
        ho.push_back((int)sv.size());
    }
    g->nseg = (int)sv.size();
    g->nhub = (int)hv.size();
    if (g->nhub) {
        const size_t tot = 3 * sv.size() + hv.size() + ho.size();
        if (hipMalloc((void**)&g->hub_buf, sizeof(int) * tot) != hipSuccess) {
            cent_graph_free(g);
            return CENT_E_OOM;
        }
        g->seg_v = g->hub_buf;
        g->seg_beg = g->seg_v + g->nseg;
        g->seg_end = g->seg_beg + g->nseg;
        g->hub_v = g->seg_end + g->nseg;
        g->hub_seg_off = g->hub_v + g->nhub;
        std::vector<int> all;
        all.reserve(tot);
        all.insert(all.end(), sv.begin(), sv.end());
        all.insert(all.end(), sb.begin(), sb.end());
        all.insert(all.end(), se.begin(), se.end());
        all.insert(all.end(), hv.begin(), hv.end());
        all.insert(all.end(), ho.begin(), ho.end());
        if (hipMemcpy(g->hub_buf, all.data(), sizeof(int) * tot, hipMemcpyHostToDevice) != hipSuccess) {
            cent_graph_free(g);
            return CENT_E_DEVICE;
        }
    }
    *out = g;
    return CENT_OK;
}

extern "C" int cent_graph_upload(cent_ctx* ctx, const struct CSR* csr, cent_graph** out) {
    if (!csr) return CENT_E_INVAL;
    return cent_graph_from_csr(ctx, csr->csrV, csr->csrE, csr->csrVSize - 1, csr->csrESize, out);
}

extern "C" void cent_graph_free(cent_graph* g) {
    if (!g) return;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    if (g->rp) (void)hipFree(g->rp);
    if (g->col) (void)hipFree(g->col);
    if (g->hub_buf) (void)hipFree(g->hub_buf);
    delete g;
}

// ------------------------------------------------------------------ timing
static hipEvent_t next_event(cent_ctx* ctx) {
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fwd, bwd;
};

// ------------------------------------------------------------------ core
static int grid_for(int n, const cent_opts* o) {
    int chunks = (n + kWave - 1) / kWave;
    int blocks = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    int cap = (o && o->grid_waves > 0) ? (o->grid_waves + kWavesPerBlock - 1) / kWavesPerBlock : 2048;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return blocks;
}

// Runs all batches of one call.  with_bc: Brandes (sigma + backward);
// otherwise closeness only (bitmask BFS).
static int run_batches(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                       bool with_bc, int bc_mode, double* bc, int64_t* farness, int64_t* dist_sum,
                       const cent_opts* opts) {
    const int n = g->n;
    const bool compat = with_bc && bc_mode == CENT_BC_REF_COMPAT;
    const bool timing = opts && (opts->flags & CENT_OPT_TIMING);
    const int G = grid_for(n, opts);
    int st;
    const size_t N = (size_t)(n > 0 ? n : 1);
    int Lcap = with_bc ? 16 : 2;
    if ((st = ensure(ctx, ctx->vis, N))) return st;
    if ((st = ensure(ctx, ctx->F, N * (size_t)(Lcap + 1)))) return st;
    Lcap = (int)(ctx->F.cap / N) - 1;
    if (with_bc) {
        if ((st = ensure(ctx, ctx->sigma, N * kWave))) return st;
        if ((st = ensure(ctx, ctx->coef, N * kWave))) return st;
        if ((st = ensure(ctx, ctx->bc, N))) return st;
        if (compat && (st = ensure(ctx, ctx->npred, N * kWave))) return st;
    }
    if (dist_sum && (st = ensure(ctx, ctx->dist_sum, N))) return st;
    if ((st = ensure(ctx, ctx->far, (size_t)(k > 0 ? k : 1)))) return st;
    if ((st = ensure(ctx, ctx->part_far, (size_t)G * kWave))) return st;
    if ((st = ensure(ctx, ctx->part_cnt, (size_t)G * 2))) return st;
    if ((st = ensure(ctx, ctx->cnt, 2))) return st;
    if ((st = ensure(ctx, ctx->src, (size_t)(k > 0 ? k : 1)))) return st;
    const int max_levels = n + 2;
    // hub segment partials
    HubArgs ha{};
    int Gseg = 1, Ghub = 1;
    if (g->nhub) {
        if ((st = ensure(ctx, ctx->hp_acc, (size_t)g->nseg * kWave))) return st;
        if ((st = ensure(ctx, ctx->hp_cnt, (size_t)g->nseg * kWave))) return st;
        if ((st = ensure(ctx, ctx->hp_or, (size_t)g->nseg))) return st;
        ha.nseg = g->nseg; ha.seg_v = g->seg_v; ha.seg_beg = g->seg_beg; ha.seg_end = g->seg_end;
        ha.nhub = g->nhub; ha.hub_v = g->hub_v; ha.hub_seg_off = g->hub_seg_off;
        ha.part_acc = ctx->hp_acc.p; ha.part_cnt = ctx->hp_cnt.p; ha.part_or = ctx->hp_or.p;
        Gseg = (g->nseg + kWavesPerBlock - 1) / kWavesPerBlock;
        if (Gseg > 4096) Gseg = 4096;
        Ghub = (g->nhub + kWavesPerBlock - 1) / kWavesPerBlock;
        if (Ghub > G) Ghub = G;  // k_fwd_hubfin flushes into the G per-block slots
    }
    if ((st = ensure(ctx, ctx->flags, 2))) return st;
    if (!ctx->h_flags) {
        HIPTRY(hipHostMalloc((void**)&ctx->h_flags, sizeof(int) * 2, hipHostMallocDefault));
        ctx->h_flags_cap = 2;
    }
    hipStream_t s = ctx->stream;
    if (k > 0) HIPTRY(hipMemcpyAsync(ctx->src.p, src, sizeof(int) * (size_t)k, hipMemcpyHostToDevice, s));
    if (with_bc) HIPTRY(hipMemsetAsync(ctx->bc.p, 0, sizeof(double) * N, s));
    if (dist_sum) HIPTRY(hipMemsetAsync(ctx->dist_sum.p, 0, sizeof(int64_t) * N, s));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));

    ctx->ev_used = 0;
    KernelTimer tm;
    int64_t levels_total = 0;
    int64_t batches = 0;
    const bool timing_ok = timing;

    for (int32_t b0 = 0; b0 < k; b0 += kWave) {
        const int kb = (k - b0) < kWave ? (k - b0) : kWave;
        const uint64_t A = (kb == 64) ? ~0ull : ((1ull << kb) - 1);
        HIPTRY(hipMemsetAsync(ctx->vis.p, 0, sizeof(uint64_t) * N, s));
        HIPTRY(hipMemsetAsync(ctx->F.p, 0, sizeof(uint64_t) * N * 2, s));
        HIPTRY(hipMemsetAsync(ctx->part_far.p, 0, sizeof(int64_t) * (size_t)G * kWave, s));
        HIPTRY(hipMemsetAsync(ctx->part_cnt.p, 0, sizeof(unsigned long long) * (size_t)G * 2, s));
        HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
        k_seed<<<1, kWave, 0, s>>>(ctx->src.p + b0, kb, ctx->F.p, ctx->vis.p,
                                   with_bc ? ctx->sigma.p : nullptr, with_bc);
        int d = 0;  // current frontier distance
        for (;;) {
            if (d + 1 >= max_levels) break;
            // frontier buffers: with_bc keeps every level, closeness ping-pongs two
            if (with_bc && d + 2 > Lcap) {
                // grow the level store, keeping levels 0..d+1
                int ncap = Lcap * 2;
                uint64_t* nf = nullptr;
                HIPTRY(hipMalloc((void**)&nf, sizeof(uint64_t) * N * (size_t)(ncap + 1)));
                HIPTRY(hipMemcpyAsync(nf, ctx->F.p, sizeof(uint64_t) * N * (size_t)(d + 1),
                                      hipMemcpyDeviceToDevice, s));
                HIPTRY(hipStreamSynchronize(s));
                HIPTRY(hipFree(ctx->F.p));
                ctx->F.p = nf;
                ctx->F.cap = N * (size_t)(ncap + 1);
                Lcap = ncap;
            }
            const uint64_t* Fd = ctx->F.p + N * (size_t)(with_bc ? d : (d & 1));
            uint64_t* Fn = ctx->F.p + N * (size_t)(with_bc ? d + 1 : ((d + 1) & 1));
            if (with_bc) {
                if (d + 1 > 1) HIPTRY(hipMemsetAsync(Fn, 0, sizeof(uint64_t) * N, s));
            } else {
                HIPTRY(hipMemsetAsync(Fn, 0, sizeof(uint64_t) * N, s));
            }
            HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int), s));
            FwdArgs fa;
            fa.n = n; fa.hub_deg = g->hub_deg; fa.rp = g->rp; fa.col = g->col; fa.vis = ctx->vis.p;
            fa.Fd = Fd; fa.Fn = Fn;
            fa.sigma = with_bc ? ctx->sigma.p : nullptr;
            fa.npred = compat ? ctx->npred.p : nullptr;
            fa.dist_sum = dist_sum ? ctx->dist_sum.p : nullptr;
            fa.weight = nullptr;
            fa.A = A; fa.dnext = d + 1;
            fa.part_far = ctx->part_far.p; fa.part_cnt = ctx->part_cnt.p;
            fa.flag = ctx->flags.p; fa.overflow = ctx->flags.p + 1;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timing_ok) { e0 = next_event(ctx); e1 = next_event(ctx); HIPTRY(hipEventRecord(e0, s)); }
            if (with_bc) {
                if (compat) k_fwd<true, true><<<G, kBlock, 0, s>>>(fa);
                else k_fwd<true, false><<<G, kBlock, 0, s>>>(fa);
            } else {
                k_fwd<false, false><<<G, kBlock, 0, s>>>(fa);
            }
            HIPTRY(hipGetLastError());
            if (g->nhub) {
                if (with_bc) {
                    k_fwd_hubseg<true><<<Gseg, kBlock, 0, s>>>(fa, ha);
                    if (compat) k_fwd_hubfin<true, true><<<Ghub, kBlock, 0, s>>>(fa, ha);
                    else k_fwd_hubfin<true, false><<<Ghub, kBlock, 0, s>>>(fa, ha);
                } else {
                    k_fwd_hubseg<false><<<Gseg, kBlock, 0, s>>>(fa, ha);
                    k_fwd_hubfin<false, false><<<Ghub, kBlock, 0, s>>>(fa, ha);
                }
                HIPTRY(hipGetLastError());
            }
            if (timing_ok) { HIPTRY(hipEventRecord(e1, s)); tm.fwd.push_back({e0, e1}); }
            ctx->stats.launches_fwd++;
            HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->flags.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
            HIPTRY(hipStreamSynchronize(s));
            if (ctx->h_flags[1]) return CENT_E_OVERFLOW;
            if (!ctx->h_flags[0]) break;
            ++d;
        }
        // d = deepest non-empty level
        levels_total += d;
        if (with_bc) {
            for (int lv = d; lv >= 1; --lv) {
                BwdArgs ba;
                ba.n = n; ba.hub_deg = g->hub_deg; ba.rp = g->rp; ba.col = g->col;
                ba.Fd = ctx->F.p + N * (size_t)lv;
                ba.Fn = ctx->F.p + N * (size_t)(lv + 1);
                ba.sigma = ctx->sigma.p; ba.coef = ctx->coef.p;
                ba.npred = compat ? ctx->npred.p : nullptr;
                ba.bc = ctx->bc.p;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (timing_ok) { e0 = next_event(ctx); e1 = next_event(ctx); HIPTRY(hipEventRecord(e0, s)); }
                if (compat) k_bwd<true><<<G, kBlock, 0, s>>>(ba);
                else k_bwd<false><<<G, kBlock, 0, s>>>(ba);
                HIPTRY(hipGetLastError());
                if (g->nhub) {
                    k_bwd_hubseg<<<Gseg, kBlock, 0, s>>>(ba, ha);
                    if (compat) k_bwd_hubfin<true><<<Ghub, kBlock, 0, s>>>(ba, ha);
                    else k_bwd_hubfin<false><<<Ghub, kBlock, 0, s>>>(ba, ha);
                    HIPTRY(hipGetLastError());
                }
                if (timing_ok) { HIPTRY(hipEventRecord(e1, s)); tm.bwd.push_back({e0, e1}); }
                ctx->stats.launches_bwd++;
            }
        }
        k_reduce_batch<<<1, 1024, 0, s>>>(ctx->part_far.p, ctx->part_cnt.p, G, kb, ctx->far.p + b0, ctx->cnt.p);
        HIPTRY(hipGetLastError());
        ++batches;
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
    if (farness && k > 0)
        HIPTRY(hipMemcpyAsync(farness, ctx->far.p, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(hc, ctx->cnt.p, sizeof(hc), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (with_bc && bc)
        for (int v = 0; v < n; ++v) bc[v] += hb[v];
    if (dist_sum)
        for (int v = 0; v < n; ++v) dist_sum[v] += hd[v];
    if (timing_ok) {
        float ms;
        for (auto& p : tm.fwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_fwd += ms;
        for (auto& p : tm.bwd) if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) ctx->stats.ms_bwd += ms;
    }
    ctx->stats.batches += batches;
    ctx->stats.levels += levels_total;
    ctx->stats.reached += (int64_t)hc[0] + k;
    ctx->stats.dag_edges += (int64_t)hc[1];
    for (int32_t i = 0; i < k; ++i) ctx->stats.edges_reached += g->comp_nnz[(size_t)g->comp[src[i]]];
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
    if (!ctx || !g || g->ctx != ctx) return CENT_E_INVAL;
    int mode = opts ? opts->bc_mode : CENT_BC_TEXTBOOK;
    if (mode != CENT_BC_TEXTBOOK && mode != CENT_BC_REF_COMPAT) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    int st = prepare_sources(g, src, k, all, &s, &kk);
    if (st) return st;
    if ((st = check_device(ctx))) return st;
    auto t0 = std::chrono::steady_clock::now();
    st = run_batches(ctx, g, s, kk, true, mode, bc, farness, dist_sum, opts);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

extern "C" int cent_cc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, int64_t* farness,
                       int64_t* dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->ctx != ctx) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    int st = prepare_sources(g, src, k, all, &s, &kk);
    if (st) return st;
    if ((st = check_device(ctx))) return st;
    auto t0 = std::chrono::steady_clock::now();
    st = run_batches(ctx, g, s, kk, false, 0, nullptr, farness, dist_sum, opts);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

extern "C" int cent_ctx_reset_stats(cent_ctx* ctx) {
    if (!ctx) return CENT_E_INVAL;
    ctx->stats = cent_stats{};
    return CENT_OK;
}

// ------------------------------------------- reference-shaped entry points
static std::mutex g_default_mu;
static cent_ctx* g_default = nullptr;

static int default_ctx(cent_ctx** out) {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) {
        int st = cent_ctx_create(0, &g_default);
        if (st) return st;
    }
    *out = g_default;
    return CENT_OK;
}

static int shim_run(const struct CSR* csr, bool with_bc, int mode, double* BCs, int64_t* CCs) {
    if (!csr || (with_bc && !BCs) || (!with_bc && !CCs)) return CENT_E_INVAL;
    cent_ctx* ctx;
    int st = default_ctx(&ctx);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_default_mu);
    cent_graph* g = nullptr;
    if ((st = cent_graph_upload(ctx, csr, &g))) return st;
    int32_t k = csr->endNodeID - csr->startNodeID + 1;
    std::vector<int32_t> src((size_t)(k > 0 ? k : 0));
    for (int32_t i = 0; i < k; ++i) src[(size_t)i] = csr->startNodeID + i;
    cent_opts o{};
    o.bc_mode = mode;
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
