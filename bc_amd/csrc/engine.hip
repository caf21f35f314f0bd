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
#include "engine.h"

std::atomic<int> g_fail_upload{0};

extern "C" int cent_debug_fail_upload(int32_t nth) {
    if (nth < 0) return CENT_E_INVAL;
    g_fail_upload.store(nth);
    return CENT_OK;
}

void devcsr_free(DevCSR& d) {
    if (d.rp) (void)hipFree(d.rp);
    if (d.col) (void)hipFree(d.col);
    if (d.chunk_beg) (void)hipFree(d.chunk_beg);
    if (d.hub_buf) (void)hipFree(d.hub_buf);
    if (d.ell) (void)hipFree(d.ell);
    d = DevCSR();
}


// Chunk and hub-segment tables of d for rows rp (host copy of the device
// row offsets), replacing any previous ones.  bounds: optional sorted block
// starts (chunks never straddle a boundary).
int devcsr_schedule(cent_ctx* ctx, const int32_t* rp, const int32_t* bounds, int32_t nbounds,
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
    d.h_chunk_rp.resize(cb.size());
    for (size_t i = 0; i < cb.size(); ++i) d.h_chunk_rp[i] = rp[cb[i]];
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

int devcsr_build(cent_ctx* ctx, const int32_t* rp, const int32_t* col, int32_t n, int64_t nnz,
                        const int32_t* bounds, int32_t nbounds, DevCSR& d) {
    d.n = n;
    d.nnz = nnz;
    TRY(upload(ctx, &d.rp, rp, (size_t)n + 1));
    TRY(upload(ctx, &d.col, col, (size_t)nnz));
    return devcsr_schedule(ctx, rp, bounds, nbounds, kHubDeg, kChunkEdges, d);
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
    if (hipHostMalloc((void**)&ctx->h_flags, sizeof(int) * kHostFlags, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CENT_E_DEVICE;
    }
    *out = ctx;
    return CENT_OK;
}

extern "C" void cent_ctx_destroy(cent_ctx* ctx) {
    if (!ctx) return;
    if (ctx->peer) cent_ctx_destroy(ctx->peer);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto e : ctx->ev_join) if (e) (void)hipEventDestroy(e);
    release(ctx->vis); release(ctx->F); release(ctx->hp_or); release(ctx->sigma); release(ctx->coef);
    release(ctx->bc); release(ctx->hp_acc); release(ctx->npred); release(ctx->hp_cnt);
    release(ctx->dist_sum); release(ctx->far); release(ctx->part_far); release(ctx->part_cnt);
    release(ctx->cnt); release(ctx->src); release(ctx->flags); release(ctx->ord);
    release(ctx->Fd2); release(ctx->QF); release(ctx->Adev); release(ctx->Q); release(ctx->qoff);
    release(ctx->cand); release(ctx->stamp); release(ctx->ctl);
    release(ctx->red_D); release(ctx->red_head);
    release(ctx->red_part); release(ctx->red_bcp); release(ctx->it_sig); release(ctx->it_coef);
    release(ctx->it_lvl); release(ctx->red_deep);
    release(ctx->ps_rec); release(ctx->ps_bits); release(ctx->ps_Q); release(ctx->ps_loff); release(ctx->ps_bcp);
    release(ctx->ps_depth); release(ctx->lvf); release(ctx->lvstat); release(ctx->red_arena);
    release(ctx->lexp); release(ctx->push_e); release(ctx->red_src); release(ctx->red_order); release(ctx->red_swgt);
    release(ctx->red_poff); release(ctx->red_items);
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
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
    return cent_ctx_workspace_bytes(ctx->peer) + (int64_t)(8 * (ctx->vis.cap + ctx->F.cap + ctx->hp_or.cap + ctx->sigma.cap + ctx->coef.cap +
                          ctx->bc.cap + ctx->hp_acc.cap + ctx->dist_sum.cap + ctx->far.cap + ctx->part_far.cap +
                          ctx->part_cnt.cap + ctx->cnt.cap + ctx->Fd2.cap + ctx->QF.cap + ctx->Adev.cap +
                          ctx->red_D.cap + ctx->red_part.cap + ctx->red_bcp.cap + ctx->it_sig.cap +
                          ctx->it_coef.cap + ctx->ps_rec.cap / 8 + ctx->ps_bcp.cap + ctx->ps_depth.cap +
                          ctx->red_swgt.cap + ctx->red_poff.cap + 2 * ctx->red_items.cap) +
                     ctx->it_lvl.cap + ctx->red_arena.cap +
                     4 * (ctx->npred.cap + ctx->hp_cnt.cap + ctx->src.cap + ctx->flags.cap + ctx->ord.cap + ctx->Q.cap +
                          ctx->qoff.cap + ctx->cand.cap + ctx->stamp.cap + ctx->ctl.cap + ctx->red_head.cap +
                          ctx->red_deep.cap + ctx->ps_Q.cap + ctx->ps_loff.cap + ctx->ps_bits.cap +
                          ctx->lexp.cap + 2 * ctx->push_e.cap + ctx->red_src.cap + ctx->red_order.cap));
}

extern "C" int cent_ctx_trim(cent_ctx* ctx) {
    if (!ctx) return CENT_E_INVAL;
    if (ctx->peer) TRY(cent_ctx_trim(ctx->peer));
    TRY(check_device(ctx));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    release(ctx->vis); release(ctx->F); release(ctx->hp_or); release(ctx->sigma); release(ctx->coef);
    release(ctx->bc); release(ctx->hp_acc); release(ctx->npred); release(ctx->hp_cnt);
    release(ctx->dist_sum); release(ctx->far); release(ctx->part_far); release(ctx->part_cnt);
    release(ctx->cnt); release(ctx->src); release(ctx->flags); release(ctx->ord);
    release(ctx->Fd2); release(ctx->QF); release(ctx->Adev); release(ctx->Q); release(ctx->qoff);
    release(ctx->cand); release(ctx->stamp); release(ctx->ctl);
    release(ctx->red_D); release(ctx->red_head);
    release(ctx->red_part); release(ctx->red_bcp); release(ctx->it_sig); release(ctx->it_coef);
    release(ctx->it_lvl); release(ctx->red_deep);
    release(ctx->ps_rec); release(ctx->ps_bits); release(ctx->ps_Q); release(ctx->ps_loff); release(ctx->ps_bcp);
    release(ctx->ps_depth); release(ctx->lvf); release(ctx->lvstat); release(ctx->red_arena);
    release(ctx->lexp); release(ctx->push_e); release(ctx->red_src); release(ctx->red_order); release(ctx->red_swgt);
    release(ctx->red_poff); release(ctx->red_items);
    if (ctx->h_stage) HIPTRY(hipHostFree(ctx->h_stage));
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
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
    (void)hipSetDevice(g->device);
    devcsr_free(g->d);
    if (g->d_new_id) (void)hipFree(g->d_new_id);
    delete g;
}

// ------------------------------------------------------------ grouping key
// Batch grouping key of every vertex: its BFS distances to the kLandmarks
// highest-degree vertices, 4 bits each (15: 15 or more, or unreached),
// landmark 0 in the top bits of word 0 and landmark 16 in the top bits of
// word 1, so comparing (word 0, word 1) compares the distance vectors in
// landmark order.  Sources with equal or close keys tend to reach every
// vertex at the same levels, so a 64-lane batch of them gathers fewer and
// fuller slices of the level-sorted sigma / coef rows.  The keys come from
// one 32-lane bit-parallel BFS over the device graph at upload.
__global__ void k_lm_seed(const int32_t* __restrict__ lm, int L, uint32_t* __restrict__ vis,
                          uint32_t* __restrict__ F, uint64_t* __restrict__ key) {
    const int j = threadIdx.x;
    if (j < L) {  // the landmarks are distinct vertices: one writer each
        const int v = lm[j];
        vis[v] = 1u << j;
        F[v] = 1u << j;
        key[2 * (size_t)v + (j >> 4)] &= ~(0xFull << (60 - 4 * (j & 15)));
    }
}

// Level d: a wavefront per vertex ORs its neighbours' level-(d-1) masks,
// lane-strided over the row, until every landmark it still lacks is found.
__global__ __launch_bounds__(256) void k_lm_level(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                  int n, uint32_t full, int d, uint32_t* __restrict__ vis,
                                                  const uint32_t* __restrict__ Fd, uint32_t* __restrict__ Fn,
                                                  uint64_t* __restrict__ key, int* __restrict__ flag) {
    const int lane = lane_id();
    const int wpb = blockDim.x / kWave;
    const int nw = gridDim.x * wpb;
    const uint64_t dd = (uint64_t)(d < 15 ? d : 15);
    for (int v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += nw) {
        const uint32_t cand = full & ~vis[v];
        uint32_t orf = 0;
        if (cand) {
            const int b = rp[v], e = rp[v + 1];
            for (int j0 = b; j0 < e; j0 += kWave) {
                const int j = j0 + lane;
                orf |= (uint32_t)wave_or(j < e ? Fd[col[j]] : 0u);
                if ((orf & cand) == cand) break;  // orf is wave-uniform
            }
        }
        const uint32_t nm = orf & cand;
        if (lane == 0) {
            Fn[v] = nm;
            if (nm) {
                vis[v] |= nm;
                uint64_t k[2] = {key[2 * (size_t)v], key[2 * (size_t)v + 1]};
                for (uint32_t m = nm; m; m &= m - 1) {
                    const int j = __builtin_ctz(m);
                    const int sh = 60 - 4 * (j & 15);
                    k[j >> 4] = (k[j >> 4] & ~(0xFull << sh)) | (dd << sh);
                }
                key[2 * (size_t)v] = k[0];
                key[2 * (size_t)v + 1] = k[1];
                *flag = 1;
            }
        }
    }
}

// Fills g->lm_key from the device CSR g->d (device ids; rp / col: its host
// copy), and g->lm_top.  Only shallow graphs get keys: deep ones run the
// source-parallel step, which takes sources in caller order.
static int lm_keys(cent_ctx* ctx, cent_graph* g, const int32_t* rp, const int32_t* col) {
    const int32_t n = g->n;
    if (n < 2 || g->ecc_probe > kDeepLevels) return CENT_OK;
    const int L = n < kLandmarks ? (int)n : kLandmarks;
    std::vector<int32_t> ids((size_t)n);
    for (int32_t v = 0; v < n; ++v) ids[(size_t)v] = v;
    std::partial_sort(ids.begin(), ids.begin() + L, ids.end(), [&](int32_t x, int32_t y) {
        const int32_t dx = rp[x + 1] - rp[x], dy = rp[y + 1] - rp[y];
        return dx != dy ? dx > dy : x < y;
    });
    // one scratch block: vis, two level planes (u32 per vertex), the keys
    // (two u64 per vertex), the landmark list and the level flag
    const size_t nn = (size_t)n;
    const size_t off_key = (12 * nn + 15) / 16 * 16, off_lm = off_key + 16 * nn, off_flag = off_lm + 4 * kLandmarks;
    char* buf = nullptr;
    HIPTRY(hipMalloc((void**)&buf, off_flag + 4));
    uint32_t* vis = reinterpret_cast<uint32_t*>(buf);
    uint32_t* F[2] = {vis + nn, vis + 2 * nn};
    uint64_t* key = reinterpret_cast<uint64_t*>(buf + off_key);
    int32_t* lm = reinterpret_cast<int32_t*>(buf + off_lm);
    int* flag = reinterpret_cast<int*>(buf + off_flag);
    hipStream_t s = ctx->stream;
    int rc = CENT_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == CENT_OK) {
            set_error(ctx, e, what);
            rc = e == hipErrorOutOfMemory ? CENT_E_OOM : CENT_E_DEVICE;
        }
        return rc == CENT_OK;
    };
    const uint32_t full = L == 32 ? 0xffffffffu : ((1u << L) - 1u);
    if (chk(hipMemsetAsync(vis, 0, 8 * nn, s), "hipMemsetAsync(lm vis)") &&
        chk(hipMemsetAsync(key, 0xff, 16 * nn, s), "hipMemsetAsync(lm key)") &&
        chk(hipMemcpyAsync(lm, ids.data(), 4 * (size_t)L, hipMemcpyHostToDevice, s), "hipMemcpyAsync(lm)")) {
        k_lm_seed<<<1, kLandmarks, 0, s>>>(lm, L, vis, F[0], key);
        const int grid = (int)std::min<int64_t>(((int64_t)n + 3) / 4, 16384);
        for (int d = 1; d <= n && rc == CENT_OK; ++d) {
            int h = 0;
            if (!chk(hipMemsetAsync(flag, 0, 4, s), "hipMemsetAsync(lm flag)")) break;
            k_lm_level<<<grid, 256, 0, s>>>(g->d.rp, g->d.col, n, full, d, vis, F[(d - 1) & 1], F[d & 1], key, flag);
            if (!chk(hipGetLastError(), "k_lm_level") ||
                !chk(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(lm flag)") ||
                !chk(hipStreamSynchronize(s), "hipStreamSynchronize(lm)"))
                break;
            if (!h) break;
        }
        if (rc == CENT_OK) {
            g->lm_key.resize(2 * nn);
            if (chk(hipMemcpyAsync(g->lm_key.data(), key, 16 * nn, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(lm key)"))
                chk(hipStreamSynchronize(s), "hipStreamSynchronize(lm)");
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (rc) {
        g->lm_key.clear();
        return rc;
    }
    // Sources hanging off the same high-degree vertex see the graph at
    // nearly the same levels (one level further out than it): batched
    // together they gather fuller slices.
    g->lm_top.assign(nn, -1);
    for (int32_t v = 0; v < n; ++v) {
        int32_t best = -1, bd = -1;
        for (int32_t j = rp[v]; j < rp[v + 1]; ++j) {
            const int32_t w = col[j], dw = rp[w + 1] - rp[w];
            if (dw > bd || (dw == bd && w < best)) {
                bd = dw;
                best = w;
            }
        }
        g->lm_top[(size_t)v] = best;
    }
    return rc;
}

extern "C" int cent_graph_create(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col, int32_t n,
                                 int64_t nnz, int32_t flags, cent_graph** out) {
    if (!ctx || !out) return CENT_E_INVAL;
    *out = nullptr;
    if ((flags & CENT_GRAPH_DEGREE_ORDER) && (flags & CENT_GRAPH_BFS_ORDER)) return CENT_E_INVAL;
    TRY(validate_csr(row_ptr, col, n, nnz));
    TRY(check_device(ctx));
    cent_graph* g = new (std::nothrow) cent_graph();
    if (!g) return CENT_E_OOM;
    g->device = ctx->device;
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
        if (!(flags & CENT_GRAPH_KEEP_ORDER)) g->probe_order.assign(q.begin(), q.begin() + (ptrdiff_t)t);
    }
    // Device layout.  Shallow graphs (probe depth up to kDeepLevels, where the
    // 64-lane steps run) are relabelled by descending degree, ties by id, so
    // the rows of high-degree vertices — the most gathered sigma/coef rows —
    // share a small, cache-resident address range.  Deep graphs (grids,
    // meshes, where one workgroup walks one source level by level) are
    // relabelled in BFS order from the probe's hub, each further component
    // after it in BFS order from its smallest id: a level's frontier and its
    // neighbours then share cache lines (2048^2 grid, 1024 sources: BC
    // 1587 -> 1327 ms, closeness 632 -> 495 ms).  CENT_GRAPH_DEGREE_ORDER /
    // CENT_GRAPH_BFS_ORDER force either; CENT_GRAPH_KEEP_ORDER keeps the
    // caller's ids.  Rows are kept sorted by the new ids.
    std::vector<int32_t> drp, dcol;
    const int32_t* urp = row_ptr;
    const int32_t* ucol = col;
    if (!(flags & CENT_GRAPH_KEEP_ORDER) && n > 1) {
        const bool bfs = (flags & CENT_GRAPH_BFS_ORDER) ||
                         (!(flags & CENT_GRAPH_DEGREE_ORDER) && g->ecc_probe > kDeepLevels);
        std::vector<int32_t> order;
        if (bfs) {
            order.reserve((size_t)n);
            order.assign(g->probe_order.begin(), g->probe_order.end());
            std::vector<char> seen((size_t)n, 0);
            for (int32_t v : order) seen[(size_t)v] = 1;
            for (int32_t s = 0; s < n; ++s) {
                if (seen[(size_t)s]) continue;
                size_t h = order.size();
                order.push_back(s);
                seen[(size_t)s] = 1;
                while (h < order.size()) {
                    const int32_t x = order[h++];
                    for (int32_t j = row_ptr[x]; j < row_ptr[x + 1]; ++j)
                        if (!seen[(size_t)col[j]]) { seen[(size_t)col[j]] = 1; order.push_back(col[j]); }
                }
            }
            g->bfs_order = true;
        } else {
            order.resize((size_t)n);
            for (int32_t v = 0; v < n; ++v) order[(size_t)v] = v;
            std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
                return (row_ptr[x + 1] - row_ptr[x]) > (row_ptr[y + 1] - row_ptr[y]);
            });
        }
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
            // sorted: a repeated neighbour is adjacent
            for (int32_t j = drp[(size_t)i]; j < o; ++j)
                if (dcol[(size_t)j] == i || (j > drp[(size_t)i] && dcol[(size_t)j] == dcol[(size_t)j - 1])) g->multi = true;
            drp[(size_t)i + 1] = o;
        }
        urp = drp.data();
        ucol = dcol.data();
    } else {
        // a multigraph (a CENT_CSR_RAW input: repeated neighbours or
        // self-loops) gets its level-1 path counts from the pull, which
        // counts every parallel edge, instead of the push, which takes one
        // path per neighbour
        std::vector<int32_t> stamp((size_t)(n > 0 ? n : 1), -1);
        for (int32_t v = 0; v < n && !g->multi; ++v)
            for (int32_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j) {
                const int32_t y = col[j];
                if (y == v || stamp[(size_t)y] == v) {
                    g->multi = true;
                    break;
                }
                stamp[(size_t)y] = v;
            }
    }
    g->probe_order.clear();
    g->probe_order.shrink_to_fit();
    int st = devcsr_build(ctx, urp, ucol, n, nnz, nullptr, 0, g->d);
    if (st) {
        cent_graph_free(g);
        return st;
    }
    g->d.multi = g->multi;
    st = lm_keys(ctx, g, urp, ucol);
    if (st) {
        cent_graph_free(g);
        return st;
    }
    {   // inline rows for the source-parallel step: a vertex of degree <= 4 holds its neighbours in one
        // 16-byte row (-1 padded); a longer row is marked and read from the CSR
        std::vector<int4> ell((size_t)(n > 0 ? n : 1));
        for (int32_t v = 0; v < n; ++v) {
            const int32_t b = urp[v], dgv = urp[v + 1] - b;
            int e[4] = {-1, -1, -1, -1};
            if (dgv > 4) e[0] = INT32_MIN;
            else
                for (int32_t j = 0; j < dgv; ++j) e[j] = ucol[b + j];
            ell[(size_t)v] = make_int4(e[0], e[1], e[2], e[3]);
        }
        st = upload(ctx, &g->d.ell, ell.data(), (size_t)n);
        if (st) {
            cent_graph_free(g);
            return st;
        }
    }
    g->h_rp.assign(urp, urp + (size_t)n + 1);
    if (!g->new_id.empty()) {  // for results written to device memory in caller order
        hipError_t e = hipMalloc((void**)&g->d_new_id, sizeof(int32_t) * (size_t)n);
        if (e == hipSuccess)
            e = hipMemcpy(g->d_new_id, g->new_id.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error(ctx, e, "hipMalloc/hipMemcpy(d_new_id)");
            cent_graph_free(g);
            return e == hipErrorOutOfMemory ? CENT_E_OOM : CENT_E_DEVICE;
        }
    }
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
    out->push_pct = g->push_pct[i];
    out->ms = 0.0;
    return CENT_OK;
}

extern "C" int cent_graph_set_layout(cent_graph* g, int32_t what, const cent_tune_params* p) {
    if (!g || !what_ok(what, true)) return CENT_E_INVAL;
    TRY(cent_tune_params_check(p));
    if (p->hub_deg != g->d.hub_deg || p->chunk_edges != g->d.chunk_edges) {
        cent_ctx* ctx = nullptr;  // errors have no context to record them in
        HIPTRY(hipSetDevice(g->device));
        HIPTRY(hipDeviceSynchronize());  // no launch may still read the old tables
        TRY(devcsr_schedule(ctx, g->h_rp.data(), nullptr, 0, p->hub_deg, p->chunk_edges, g->d));
    }
    for (int i = 0; i < 2; ++i)
        if (what & (i == 0 ? CENT_TUNE_BC : CENT_TUNE_CC)) {
            g->grid_waves[i] = p->grid_waves;
            g->frontier[i] = p->frontier;
            g->push_pct[i] = p->push_pct;
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

// Random-gather probe: groups of L lanes read L consecutive doubles at a
// random L*8-byte aligned position (a hash of the group and step), four
// reads in flight per lane, so the rate is that of independent gathers of
// L*8-byte pieces, like the dense step's slices of sigma / coef rows.
__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <int L>
__global__ __launch_bounds__(256) void k_gather_probe(const double* __restrict__ a, uint32_t npieces, int steps,
                                                      double* __restrict__ sink) {
    constexpr int U = 4;
    const uint32_t grp = (blockIdx.x * blockDim.x + threadIdx.x) / L;
    const int sub = (int)(threadIdx.x % L);
    double acc = 0.0;
    for (int it = 0; it < steps; it += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t h = probe_hash(grp * 0x9E3779B9u + (uint32_t)(it + u) * 0x85EBCA6Bu);
            v[u] = a[(size_t)(h % npieces) * L + sub];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    if (acc != 0.0) sink[0] = acc;  // the buffer is zero: never taken, keeps the loads
}

extern "C" int cent_probe_gather_bandwidth(cent_ctx* ctx, int64_t bytes, int32_t piece, int32_t iters, double* gbs) {
    if (!ctx || !gbs || iters < 1 || (piece != 32 && piece != 64 && piece != 128 && piece != 256 && piece != 512) ||
        bytes < 4 * piece || bytes / piece > (int64_t)UINT32_MAX)
        return CENT_E_INVAL;
    TRY(check_device(ctx));
    double* a = nullptr;
    HIPTRY(hipMalloc((void**)&a, (size_t)bytes + 64));
    double* sink = a + bytes / 8;
    ctx->ev_used = 0;
    hipEvent_t e0 = next_event(ctx), e1 = next_event(ctx);
    int rc = CENT_OK;
    if (!e0 || !e1 || hipMemsetAsync(a, 0, (size_t)bytes + 64, ctx->stream) != hipSuccess) rc = CENT_E_DEVICE;
    const int grid = 256 * 32, steps = 64;
    const uint32_t np = (uint32_t)(bytes / piece);
    auto launch = [&] {
        switch (piece) {
            case 32: k_gather_probe<4><<<grid, 256, 0, ctx->stream>>>(a, np, steps, sink); break;
            case 64: k_gather_probe<8><<<grid, 256, 0, ctx->stream>>>(a, np, steps, sink); break;
            case 128: k_gather_probe<16><<<grid, 256, 0, ctx->stream>>>(a, np, steps, sink); break;
            case 256: k_gather_probe<32><<<grid, 256, 0, ctx->stream>>>(a, np, steps, sink); break;
            default: k_gather_probe<64><<<grid, 256, 0, ctx->stream>>>(a, np, steps, sink); break;
        }
    };
    if (!rc) {
        launch();  // warm
        (void)hipEventRecord(e0, ctx->stream);
        for (int i = 0; i < iters; ++i) launch();
        (void)hipEventRecord(e1, ctx->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = CENT_E_DEVICE;
        else *gbs = (double)grid * 256 * steps * 8 * iters / (ms * 1e-3) / 1e9;
    }
    ctx->ev_used = 0;
    (void)hipFree(a);
    return rc;
}

// Line-gather probe: each wave load instruction reads LC random whole
// 128-byte lines (16 lanes a line, 8 bytes a lane; the wave's other lanes
// issue nothing), eight loads in flight per lane.  The memory pipeline's
// rate per load instruction, not per byte, limits loads that touch one or
// two lines, which is what the dense step's slice gathers are.
template <int LC>
__global__ __launch_bounds__(256) void k_gather_lines_probe(const double* __restrict__ a, uint32_t nlines, int steps,
                                                            double* __restrict__ sink) {
    constexpr int U = 8;
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    double acc = 0.0;
    if (lane < 16 * LC) {
        const uint32_t grp = (uint32_t)(lane >> 4);
        const int sub = lane & 15;
        for (int it = 0; it < steps; it += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t h = probe_hash(wave * 0x9E3779B9u + (uint32_t)(it + u) * 0x85EBCA6Bu + grp * 0x27D4EB2Fu);
                v[u] = a[(size_t)(h % nlines) * 16 + sub];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) acc += v[u];
        }
    }
    if (acc != 0.0) sink[0] = acc;  // the buffer is zero: never taken, keeps the loads
}

extern "C" int cent_probe_gather_lines(cent_ctx* ctx, int64_t bytes, int32_t lines, int32_t iters, double* gbs) {
    if (!ctx || !gbs || iters < 1 || (lines != 1 && lines != 2 && lines != 4) || bytes < 128 * 4 ||
        bytes / 128 > (int64_t)UINT32_MAX)
        return CENT_E_INVAL;
    TRY(check_device(ctx));
    double* a = nullptr;
    HIPTRY(hipMalloc((void**)&a, (size_t)bytes + 64));
    double* sink = a + bytes / 8;
    ctx->ev_used = 0;
    hipEvent_t e0 = next_event(ctx), e1 = next_event(ctx);
    int rc = CENT_OK;
    if (!e0 || !e1 || hipMemsetAsync(a, 0, (size_t)bytes + 64, ctx->stream) != hipSuccess) rc = CENT_E_DEVICE;
    const int grid = 256 * 32, steps = 64;
    const uint32_t nl = (uint32_t)(bytes / 128);
    auto launch = [&] {
        switch (lines) {
            case 1: k_gather_lines_probe<1><<<grid, 256, 0, ctx->stream>>>(a, nl, steps, sink); break;
            case 2: k_gather_lines_probe<2><<<grid, 256, 0, ctx->stream>>>(a, nl, steps, sink); break;
            default: k_gather_lines_probe<4><<<grid, 256, 0, ctx->stream>>>(a, nl, steps, sink); break;
        }
    };
    if (!rc) {
        launch();  // warm
        (void)hipEventRecord(e0, ctx->stream);
        for (int i = 0; i < iters; ++i) launch();
        (void)hipEventRecord(e1, ctx->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = CENT_E_DEVICE;
        else *gbs = (double)grid * 4 * steps * lines * 128.0 * iters / (ms * 1e-3) / 1e9;
    }
    ctx->ev_used = 0;
    (void)hipFree(a);
    return rc;
}

// ------------------------------------------------------------ device arrays
extern "C" int cent_device_alloc(cent_ctx* ctx, int64_t bytes, void** out) {
    if (!ctx || !out || bytes < 0) return CENT_E_INVAL;
    *out = nullptr;
    TRY(check_device(ctx));
    HIPTRY(hipMalloc(out, (size_t)(bytes > 0 ? bytes : 1)));
    return CENT_OK;
}

extern "C" int cent_device_free(cent_ctx* ctx, void* p) {
    if (!ctx) return CENT_E_INVAL;
    if (!p) return CENT_OK;
    TRY(check_device(ctx));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    HIPTRY(hipFree(p));
    return CENT_OK;
}

extern "C" int cent_device_copy(cent_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t kind) {
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return CENT_E_INVAL;
    hipMemcpyKind k;
    switch (kind) {
        case CENT_COPY_TO_DEVICE: k = hipMemcpyHostToDevice; break;
        case CENT_COPY_TO_HOST: k = hipMemcpyDeviceToHost; break;
        case CENT_COPY_ON_DEVICE: k = hipMemcpyDeviceToDevice; break;
        default: return CENT_E_INVAL;
    }
    if (bytes == 0) return CENT_OK;
    TRY(check_device(ctx));
    HIPTRY(hipMemcpyAsync(dst, src, (size_t)bytes, k, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    return CENT_OK;
}

extern "C" int cent_device_fill(cent_ctx* ctx, void* p, int32_t byte, int64_t bytes) {
    if (!ctx || bytes < 0 || (bytes > 0 && !p)) return CENT_E_INVAL;
    if (bytes == 0) return CENT_OK;
    TRY(check_device(ctx));
    HIPTRY(hipMemsetAsync(p, byte & 0xff, (size_t)bytes, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    return CENT_OK;
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
