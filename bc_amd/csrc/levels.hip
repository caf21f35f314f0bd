// Plain calls: the three level steps (dense 64-lane pull with hub
// segments, frontier queues replayed from hipGraphs, source-parallel
// workgroups), the batching of a call's sources, cent_bc / cent_cc and the
// autotuner.  Shared types are in engine.h.
#include "engine.h"
#include "msbfs_kernels.h"
#include "ccw_kernels.h"
#include "persource_kernels.h"
#include "sparse_kernels.h"

// ------------------------------------------------------------ level loop
static int grid_for(int chunks, const cent_opts* o) {
    int blocks = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    int cap = (o && o->grid_waves > 0) ? (o->grid_waves + kWavesPerBlock - 1) / kWavesPerBlock : 2048;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return blocks;
}


// Path-count exponent rows (64 ints a level) for levels 0 .. levels + 1,
// keeping the first keep_rows rows when the table grows.
static int ensure_lexp(cent_ctx* ctx, size_t levels, size_t keep_rows) {
    const size_t want = (levels + 2) * kWave;
    if (want <= ctx->lexp.cap) return CENT_OK;
    int* nl = nullptr;
    HIPTRY(hipMalloc((void**)&nl, sizeof(int) * want));
    HIPTRY(hipMemsetAsync(nl, 0, sizeof(int) * want, ctx->stream));
    if (keep_rows && ctx->lexp.p)
        HIPTRY(hipMemcpyAsync(nl, ctx->lexp.p, sizeof(int) * kWave * std::min(keep_rows, ctx->lexp.cap / kWave),
                              hipMemcpyDeviceToDevice, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    if (ctx->lexp.p) HIPTRY(hipFree(ctx->lexp.p));
    ctx->lexp.p = nl;
    ctx->lexp.cap = want;
    return CENT_OK;
}

// Workspace for a DevCSR with N vertices; the F store grows on demand.
// Batches side by side in the closeness step's interleaved masks: nb
// rounded up to a power of two.
static inline int ccw_width(int nb) { return nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : nb <= 8 ? 8 : nb <= 16 ? 16 : nb <= 32 ? 32 : 64; }

static int ensure_workspace(cent_ctx* ctx, const DevCSR& d, const BatchSpec& sp, int G, int nb) {
    const size_t N = (size_t)(d.n > 0 ? d.n : 1);
    const size_t W = sp.with_bc ? 1 : (size_t)ccw_width(nb);
    TRY(ensure(ctx, ctx->vis, N * W));
    // BC keeps one {mask, base} record plane per level (2 words a vertex);
    // closeness a level-parity pair of planes of W interleaved masks
    const size_t fw = sp.with_bc ? 2 : 1;
    const size_t want = sp.with_bc ? 2 * N * 17 : N * (2 * W + 1);
    if (ctx->F_n != N || ctx->F.cap < want) {
        TRY(ensure(ctx, ctx->F, want));
        ctx->F_n = N;
    }
    ctx->F_levels = ctx->F.cap / (N * fw) - 1;
    if (sp.with_bc) {
        TRY(ensure_lexp(ctx, ctx->F_levels, 0));
        TRY(ensure(ctx, ctx->sigma, N * kWave));  // coef is written over it in the backward
        if (sp.compat) TRY(ensure(ctx, ctx->npred, N * kWave));
        TRY(ensure(ctx, ctx->push_e, 1));
    }
    TRY(ensure(ctx, ctx->part_far, (size_t)G * kWave * nb));
    TRY(ensure(ctx, ctx->part_cnt, (size_t)G * 2 * nb));
    TRY(ensure(ctx, ctx->cnt, 2));
    TRY(ensure(ctx, ctx->flags, 2));
    if (d.nhub) {
        TRY(ensure(ctx, ctx->hp_acc, (size_t)d.nseg * kWave));
        TRY(ensure(ctx, ctx->hp_cnt, (size_t)d.nseg * kWave));
        TRY(ensure(ctx, ctx->hp_or, (size_t)d.nseg * W));
    }
    return CENT_OK;
}

// Doubles the level planes of F (PL words each), keeping the first keep_levels.
static int grow_levels(cent_ctx* ctx, size_t PL, int keep_levels) {
    const size_t ncap = ctx->F_levels * 2;
    uint64_t* nf = nullptr;
    HIPTRY(hipMalloc((void**)&nf, sizeof(uint64_t) * PL * (ncap + 1)));
    HIPTRY(hipMemcpyAsync(nf, ctx->F.p, sizeof(uint64_t) * PL * (size_t)keep_levels, hipMemcpyDeviceToDevice,
                          ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    HIPTRY(hipFree(ctx->F.p));
    ctx->F.p = nf;
    ctx->F.cap = PL * (ncap + 1);
    ctx->F_levels = ncap;
    if (ctx->lexp.p) TRY(ensure_lexp(ctx, ncap, (size_t)keep_levels + 1));
    return CENT_OK;
}

// Totals of the nb batches of a launch group from their G per-workgroup
// partials each: far_out[0..k) assigned (k = 64 (nb - 1) + k_last),
// ctx->cnt accumulated.  One launch, grid y = batch; one reducing workgroup
// per 128 partial rows of a batch, at most 32.
static int reduce_batch(cent_ctx* ctx, int G, int nb, int k_last, int64_t* far_out) {
    hipStream_t s = ctx->stream;
    const size_t k = (size_t)(nb - 1) * kWave + (size_t)k_last;
    HIPTRY(hipMemsetAsync(far_out, 0, sizeof(int64_t) * (k > 0 ? k : 1), s));
    const int rg = std::max(1, std::min(32, G / 128));
    k_reduce_batch<<<dim3(rg, nb), 1024, 0, s>>>(ctx->part_far.p, ctx->part_cnt.p, G, k_last, far_out, ctx->cnt.p);
    HIPTRY(hipGetLastError());
    return CENT_OK;
}


int level_flags_reset(cent_ctx* ctx) {
    TRY(ensure(ctx, ctx->flags, 2));
    HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, ctx->stream));
    return CENT_OK;
}

// Zeroes up to eight ranges of 64-bit words in one launch (the clears that
// start a launch group; one launch instead of a fill per buffer).
struct ClearRanges {
    uint64_t* p[8];
    size_t n[8];
    int count;
};

// Level 2 of a betweenness batch is pushed from its level-1 frontier when
// that frontier's rows hold under this share (percent) of the view's
// edges, unless the call or the graph's tuned settings say otherwise.
// Measured one-stream on 2^20-vertex graphs (push always / pulled): an
// Erdos-Renyi graph of average degree 32 152 / 160 ms (its level-1 rows
// hold about 0.1% of the edges), R-MAT with a = 0.45 102 / 105 ms, the
// skewed R-MAT of a = 0.57 249 / 235 ms (level-1 rows about 15%: there
// the push's per-edge atomics cost more than the pull's scan).
constexpr int kPushPctDefault = 2;

__global__ void k_clear(ClearRanges c) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int r = 0; r < c.count; ++r) {
        uint64_t* p = c.p[r];
        const size_t n = c.n[r];
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
    }
}

// Workgroups of the level-1 push launches (their lane partials go to the
// first kPushBlocks of the forward step's block slots).
constexpr int kPushBlocks = 256;

// One closeness level of a launch group on the interleaved step: level 1 by
// push (grid y = batch), later levels by the pull over non-hub rows, then
// the hub segments and their finish.
template <int W>
static void ccw_level(const CcwArgs& ca, int lv, int G, int Gp, int Gseg, int Ghub, bool hubs, int nb,
                      const int* d_src, hipStream_t s) {
    if (lv == 0) {
        k_ccw_push_or<W><<<dim3(Gp, nb), kBlock, 0, s>>>(ca, d_src);
        k_ccw_push_fin<W><<<dim3(Gp, nb), kBlock, 0, s>>>(ca, d_src);
        return;
    }
    k_ccw<W><<<G, ccw_block(W), 0, s>>>(ca);
    if (hubs) {
        k_ccw_hubseg<W><<<Gseg, kBlock, 0, s>>>(ca);
        k_ccw_hubfin<W><<<Ghub, ccw_block(W), 0, s>>>(ca);
    }
}

// Batches of the k sources d_src[0..k) (device ids) whose BFS stays inside
// the view's vertex range: one 64-lane batch for BC (k <= 64), up to
// kMultiBatch batches of 64 consecutive sources for closeness, their masks
// interleaved by vertex (one row scan per level serves all of them).
// far_out (device, k entries) gets the weighted distance sums.  bound > 0:
// the BFS of every source ends within `bound` levels, so all levels are
// enqueued at once with no host check (0: host-checked runs).  Adds the
// depth to *levels once per batch (host-checked runs only; bounded ones
// report it through the context's level statistics).
int level_batch(cent_ctx* ctx, const GView& v, const BatchSpec& sp, const int* d_src, int k,
                int64_t* far_out, KernelTimer& tm, int64_t* levels, int bound) {
    const DevCSR& d = *v.d;
    hipStream_t s = ctx->stream;
    const size_t N = (size_t)(d.n > 0 ? d.n : 1);
    const size_t R = (size_t)(v.hi - v.lo);
    const int nb = (k + kWave - 1) / kWave;
    if (nb < 1 || nb > kMultiBatch || (sp.with_bc && nb != 1)) return CENT_E_INVAL;
    const int W = sp.with_bc ? 1 : ccw_width(nb);
    // workgroups of the level kernels (they hold the partial slots): the
    // closeness kernels run the same waves in ccw_block(W)-thread workgroups
    const int G = sp.with_bc ? grid_for(v.nchunk, sp.opts)
                             : std::max(1, grid_for(v.nchunk, sp.opts) * kBlock / ccw_block(W));
    TRY(ensure_workspace(ctx, d, sp, G, nb));
    if (bound > kHostFlags - 2) bound = 0;  // the level flags of one run must fit
    if (bound > 0 && sp.with_bc)
        while ((size_t)(bound + 1) > ctx->F_levels) TRY(grow_levels(ctx, N * 2, 0));
    const int kl = k - (nb - 1) * kWave;  // sources of the last batch
    const uint64_t A_last = (kl == 64) ? ~0ull : ((1ull << kl) - 1);
    const uint64_t A = nb == 1 ? A_last : ~0ull;
    // F planes: BC level lv at F + lv * PL ({mask, base} records, PL = 2N
    // words); closeness level parity p at F + p * PL, W interleaved masks a
    // vertex (PL = W N words)
    const size_t fw = sp.with_bc ? 2 : (size_t)W;
    const size_t PL = N * fw;
    auto plane = [&](size_t p) { return ctx->F.p + p * PL; };
    // Level 2 of a betweenness batch on a simple graph: pushed from the
    // level-1 frontier while its rows hold fewer than push_lim edges (0:
    // pulled); the batch's kernels decide on the device (push2_on).
    unsigned long long push_lim = 0;
    if (sp.with_bc && !d.multi && v.nchunk > 0) {
        const int pct = sp.opts && sp.opts->push_pct ? sp.opts->push_pct : kPushPctDefault;
        const int64_t ev = d.h_chunk_rp[(size_t)(v.c0 + v.nchunk)] - d.h_chunk_rp[(size_t)v.c0];
        if (pct > 0 && ev > 0) push_lim = (unsigned long long)std::max<int64_t>(1, ev / 100 * pct + ev % 100 * pct / 100);
    }
    {  // vis, the level-1 plane (and the level-0 one for betweenness), the partials
        ClearRanges cr{};
        auto add = [&](void* p, size_t n) { cr.p[cr.count] = static_cast<uint64_t*>(p); cr.n[cr.count++] = n; };
        add(ctx->vis.p + (size_t)W * v.lo, (size_t)W * R);
        if (sp.with_bc) add(plane(0) + fw * v.lo, fw * R);
        add(plane(1) + fw * v.lo, fw * R);
        add(ctx->part_far.p, (size_t)G * kWave * nb);
        add(ctx->part_cnt.p, (size_t)G * 2 * nb);
        if (sp.with_bc)  // the exponent rows of every level this batch can reach: 64 ints, 32 words each
            add(ctx->lexp.p, std::min(ctx->lexp.cap, (ctx->F_levels + 2) * (size_t)kWave) / 2);
        if (push_lim) {  // the level-1 rows' edge count, and the level-2 plane the push ORs into
            add(ctx->push_e.p, 1);
            if (ctx->F_levels >= 2) add(plane(2) + fw * v.lo, fw * R);
        }
        size_t tot = 0;
        for (int i = 0; i < cr.count; ++i) tot = std::max(tot, cr.n[i]);
        const int cg = (int)std::min<size_t>((tot + 255) / 256, 4096);
        k_clear<<<std::max(cg, 1), 256, 0, s>>>(cr);
        HIPTRY(hipGetLastError());
    }
    if (sp.with_bc) {
        k_seed<<<1, kWave, 0, s>>>(d_src, kl, plane(0), ctx->vis.p, ctx->sigma.p, true);
    } else {
        switch (W) {
            case 1: k_ccw_seed<1><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            case 2: k_ccw_seed<2><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            case 4: k_ccw_seed<4><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            case 8: k_ccw_seed<8><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            case 16: k_ccw_seed<16><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            case 32: k_ccw_seed<32><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
            default: k_ccw_seed<64><<<nb, kWave, 0, s>>>(d_src, nb, kl, ctx->vis.p); break;
        }
    }
    HIPTRY(hipGetLastError());

    HubArgs ha{};
    int Gseg = 1, Ghub = 1;
    if (v.nhub) {
        ha.nseg = v.nseg; ha.seg_v = d.seg_v + v.s0; ha.seg_beg = d.seg_beg + v.s0; ha.seg_end = d.seg_end + v.s0;
        ha.seg_base = v.s0;
        ha.nhub = v.nhub; ha.hub_v = d.hub_v + v.h0; ha.hub_seg_off = d.hub_seg_off + v.h0;
        ha.part_acc = ctx->hp_acc.p; ha.part_cnt = ctx->hp_cnt.p;
        Gseg = std::min((v.nseg + kWavesPerBlock - 1) / kWavesPerBlock, 4096);
        const int hw = sp.with_bc ? kWavesPerBlock : ccw_block(W) / kWave;
        Ghub = std::min((v.nhub + hw - 1) / hw, G);  // hubfin flushes into G slots
    }
    const bool bc = sp.with_bc, compat = sp.compat;
    // Forward levels go in runs enqueued without waiting: level d0 + i of a
    // run reads lvf[i] ("level d0 + i is non-empty", lvf[0] = 1) and sets
    // lvf[i + 1]; once the BFS has ended the remaining launches return at
    // once.  The host checks the flags once per run.  A run is as long as
    // the last batch's depth suggests.
    TRY(ensure(ctx, ctx->lvf, (size_t)kHostFlags));
    int dl = 0;  // first level of the current run; on exit, the deepest non-empty level
    for (;;) {
        int run = bound > 0 ? bound : std::max(2, std::min(kHostFlags - 2, ctx->depth_hint + 2 - dl));
        // a grown store keeps the planes up to level dl + 2: levels 1 and 2
        // are cleared when the batch starts (their pushes OR into them)
        if (bc && (size_t)(dl + run + 1) > ctx->F_levels)
            TRY(grow_levels(ctx, PL, (int)std::min<size_t>((size_t)dl + 3, ctx->F_levels + 1)));
        if (bc) run = std::min(run, (int)ctx->F_levels - 1 - dl);
        HIPTRY(hipMemsetAsync(ctx->lvf.p, 0, sizeof(int) * (size_t)(run + 1), s));
        HIPTRY(hipMemsetD32Async((hipDeviceptr_t)ctx->lvf.p, 1, 1, s));
        for (int i = 0; i < run; ++i) {
            const int lv = dl + i;
            const uint64_t* Fd = bc ? ctx->F.p + PL * (size_t)lv : plane(lv & 1);
            uint64_t* Fn = bc ? ctx->F.p + PL * (size_t)(lv + 1) : plane((lv + 1) & 1);
            // no clear of the next plane: the level kernels (betweenness and
            // closeness) write every entry of a level >= 2 plane, and level 1's
            // plane is cleared when the batch starts
            if (!bc) {  // closeness: the interleaved step over the group's batches
                CcwArgs ca{};
                ca.hub_deg = d.hub_deg; ca.nchunk = v.nchunk; ca.chunk_beg = d.chunk_beg + v.c0;
                ca.rp = d.rp; ca.col = d.col; ca.vis = ctx->vis.p; ca.Fd = Fd; ca.Fn = Fn;
                ca.dist_sum = sp.dist_sum; ca.weight = sp.weight;
                ca.nb = nb; ca.A_last = A_last; ca.dnext = lv + 1; ca.pslots = G;
                ca.part_far = ctx->part_far.p; ca.part_cnt = ctx->part_cnt.p;
                ca.flag = ctx->lvf.p + i + 1; ca.live = ctx->lvf.p + i;
                ca.nseg = ha.nseg; ca.seg_v = ha.seg_v; ca.seg_beg = ha.seg_beg; ca.seg_end = ha.seg_end;
                ca.seg_base = ha.seg_base; ca.nhub = ha.nhub; ca.hub_v = ha.hub_v; ca.hub_seg_off = ha.hub_seg_off;
                ca.part_or = ctx->hp_or.p;
                hipEvent_t e0;
                TRY(timer_begin(ctx, tm, &e0));
                const int Gp = std::min(G, kPushBlocks);
                switch (W) {
                    case 1: ccw_level<1>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    case 2: ccw_level<2>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    case 4: ccw_level<4>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    case 8: ccw_level<8>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    case 16: ccw_level<16>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    case 32: ccw_level<32>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                    default: ccw_level<64>(ca, lv, G, Gp, Gseg, Ghub, v.nhub > 0, nb, d_src, s); break;
                }
                HIPTRY(hipGetLastError());
                TRY(timer_end(ctx, tm, e0, true));
                continue;
            }
            FwdArgs fa{};
            fa.n = d.n; fa.hub_deg = d.hub_deg; fa.nchunk = v.nchunk; fa.chunk_beg = d.chunk_beg + v.c0;
            fa.rp = d.rp; fa.col = d.col; fa.vis = ctx->vis.p; fa.Fd = Fd; fa.Fn = Fn;
            fa.sigma = bc ? ctx->sigma.p : nullptr;
            fa.npred = compat ? ctx->npred.p : nullptr;
            fa.dist_sum = sp.dist_sum;
            fa.weight = sp.weight;
            fa.A = A; fa.dnext = lv + 1;
            fa.part_far = ctx->part_far.p; fa.part_cnt = ctx->part_cnt.p;
            fa.flag = ctx->lvf.p + i + 1; fa.overflow = ctx->flags.p + 1;
            fa.live = ctx->lvf.p + i;
            fa.lexp = ctx->lexp.p;
            if (push_lim && lv <= 1) {  // level 1 counts its rows' edges, level 2 reads the count
                fa.push_e = ctx->push_e.p;
                fa.push_lim = lv == 1 ? push_lim : 0;
            }
            hipEvent_t e0;
            TRY(timer_begin(ctx, tm, &e0));
            if (lv == 1 && push_lim) {  // level 2 by push where the batch's level-1 rows are few (push2_on)
                const int Gsh = std::min(Gseg, G);  // k_push2_sig_hub flushes into the G block slots
                k_push2_or<<<G, kBlock, 0, s>>>(fa);
                if (v.nhub) k_push2_or_hub<<<Gseg, kBlock, 0, s>>>(fa, ha);
                if (compat) {
                    k_push2_fin<true><<<G, kBlock, 0, s>>>(fa);
                    k_push2_sig<true><<<G, kBlock, 0, s>>>(fa);
                    if (v.nhub) k_push2_sig_hub<true><<<Gsh, kBlock, 0, s>>>(fa, ha);
                } else {
                    k_push2_fin<false><<<G, kBlock, 0, s>>>(fa);
                    k_push2_sig<false><<<G, kBlock, 0, s>>>(fa);
                    if (v.nhub) k_push2_sig_hub<false><<<Gsh, kBlock, 0, s>>>(fa, ha);
                }
                HIPTRY(hipGetLastError());
            }
            if (lv == 0 && !d.multi) {  // level 1: push from the sources' rows (one path per neighbour)
                const int Gp = std::min(G, kPushBlocks);
                k_push1_or<<<Gp, kBlock, 0, s>>>(fa, d_src);
                if (compat) k_push1_fin<true><<<Gp, kBlock, 0, s>>>(fa, d_src);
                else k_push1_fin<false><<<Gp, kBlock, 0, s>>>(fa, d_src);
                HIPTRY(hipGetLastError());
                TRY(timer_end(ctx, tm, e0, true));
                continue;
            }
            if (compat) k_fwd<true><<<G, kBlock, 0, s>>>(fa);
            else k_fwd<false><<<G, kBlock, 0, s>>>(fa);
            HIPTRY(hipGetLastError());
            if (v.nhub) {
                k_fwd_hubseg<<<Gseg, kBlock, 0, s>>>(fa, ha);
                if (compat) k_fwd_hubfin<true><<<Ghub, kBlock, 0, s>>>(fa, ha);
                else k_fwd_hubfin<false><<<Ghub, kBlock, 0, s>>>(fa, ha);
                HIPTRY(hipGetLastError());
            }
            TRY(timer_end(ctx, tm, e0, true));
        }
        if (bound > 0) {  // no host check: the backward follows at once, level by level gated
            dl = run - 1;
            break;
        }
        HIPTRY(hipMemcpyAsync(ctx->h_flags, ctx->lvf.p, sizeof(int) * (size_t)(run + 1), hipMemcpyDeviceToHost, s));
        HIPTRY(hipMemcpyAsync(ctx->h_flags + kHostFlags - 1, ctx->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (ctx->h_flags[kHostFlags - 1]) return CENT_E_OVERFLOW;
        int i = 1;
        while (i <= run && ctx->h_flags[i]) ++i;  // levels dl .. dl + i - 1 are non-empty
        ctx->stats.launches_fwd += std::min(i, run);
        if (i <= run) {
            dl += i - 1;
            break;
        }
        dl += run;
    }
    if (bound == 0) {
        ctx->depth_hint = dl;
        *levels += (int64_t)dl * nb;  // deepest non-empty level, per batch
    }
    if (bc) {
        for (int lv = dl; lv >= 1; --lv) {
            BwdArgs ba{};
            ba.live = bound > 0 ? ctx->lvf.p + lv : nullptr;
            ba.n = d.n; ba.hub_deg = d.hub_deg; ba.nchunk = v.nchunk; ba.chunk_beg = d.chunk_beg + v.c0;
            ba.rp = d.rp; ba.col = d.col;
            ba.Fd = ctx->F.p + PL * (size_t)lv;
            ba.Fn = ctx->F.p + PL * (size_t)(lv + 1);
            ba.sigma = ctx->sigma.p; ba.coef = ctx->sigma.p;  // coef over sigma, in place
            ba.npred = compat ? ctx->npred.p : nullptr;
            ba.bc = sp.bc;
            ba.wgt = sp.bwgt;
            ba.swgt = sp.swgt;
            ba.bsrc = d_src;
            ba.lexp = ctx->lexp.p;
            ba.d = lv;
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
            if (bound == 0) ctx->stats.launches_bwd++;
        }
    }
    if (bound > 0) {
        k_level_check<<<1, 1, 0, s>>>(ctx->lvf.p, bound, nb, bc, ctx->lvstat.p);
        HIPTRY(hipGetLastError());
    }
    TRY(reduce_batch(ctx, G, nb, kl, far_out));
    ctx->stats.batches += nb;
    return CENT_OK;
}


// ------------------------------------------------------------ frontier-queue mode
constexpr int kSpReplay = 16;  // levels per hipGraph replay
constexpr size_t kSpLevels = 65536;  // deepest BFS of a betweenness batch on this step

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
    // path-count exponents of up to kSpLevels levels (betweenness); a
    // deeper BFS makes the call fall back to the dense step
    const int lexp_levels = (int)std::min<size_t>(N + 3, kSpLevels);
    if (sp.with_bc) {
        TRY(ensure(ctx, ctx->sigma, N * kWave));
        TRY(ensure(ctx, ctx->coef, N * kWave));
        if (sp.compat) TRY(ensure(ctx, ctx->npred, N * kWave));
        TRY(ensure_lexp(ctx, (size_t)lexp_levels, 0));
    }
    hipStream_t s = ctx->stream;
    if (sp.with_bc) HIPTRY(hipMemsetAsync(ctx->lexp.p, 0, sizeof(int) * ctx->lexp.cap, s));
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
    a.lexp = sp.with_bc ? ctx->lexp.p : nullptr;
    a.lexp_levels = lexp_levels;
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
    if (ctl_h[kCtlTooDeep]) return CENT_E_UNSUPPORTED;  // the caller reruns on the dense step
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
    TRY(reduce_batch(ctx, G, 1, kb, far_out));
    ctx->stats.batches++;
    return CENT_OK;
}

// ------------------------------------------------------------ source-parallel mode
// All k sources of a plain call in one launch of k_ps_bc (betweenness) or
// k_ps_cc (closeness only): G workgroup slots, each with its own per-vertex
// arrays, take sources s, s + G, ... in turn.
// G is bounded by k, by the slot cap (opts->max_slots, else grid_waves / 8,
// else 256 — more concurrent traversals than that only contend for the
// memory system — doubled up to 1024 while all slots' arrays stay within
// kPsSlotSetBytes) and by a memory budget (40% of free device memory, at
// most 64 GiB).  Betweenness returns CENT_E_UNSUPPORTED, with nothing
// accumulated into the BC output, if some source's BFS is deeper than the
// level table: the caller then runs the batch form.  Closeness keeps no
// level table and has no depth limit.
constexpr int kPsDefaultSlots = 256;
constexpr int kPsMaxCachedSlots = 1024;
// Measured on MI355X: doubling the slots from 256 pays (by 8-30%) while the
// slots' arrays stay within about 2 GB — R-MAT 2^10-2^14 and block trees of
// 11k-113k vertices, all sources — and costs beyond (the 4.2M-vertex grid,
// 117 MB per slot for BC, is fastest at 256).
constexpr size_t kPsSlotSetBytes = (size_t)2 << 30;
// Betweenness records go wide (PsRecW, the BC partial in the record) when
// the slots' narrow records and partials together pass this size, far
// beyond the 256 MB Infinity Cache; below it they nearly fit there and the
// narrow form's smaller footprint wins.
constexpr size_t kPsWideBytes = (size_t)512 << 20;

// Slot plan of the source-parallel step: slot count, record form and the
// bytes a slot holds.  budget: the most the slots may hold together
// (SIZE_MAX: no limit).
struct PsPlan {
    size_t G;
    bool wide;
    size_t per_slot;
};

static PsPlan ps_plan(size_t n, bool bc, size_t K, int max_slots, int grid_waves, size_t budget) {
    const size_t lcap = std::min<size_t>(n + 2, 65536);
    const size_t nw = ((n + 31) / 32 + 3) / 4 * 4;
    auto slot_bytes = [&](bool wide) {
        return bc ? n * ((wide ? sizeof(PsRecW) : sizeof(PsRec) + 8) + 4) + lcap * 8 : n * 4 + nw * 4;
    };
    size_t per = slot_bytes(false);
    size_t cap = kPsDefaultSlots;
    if (max_slots > 0) cap = (size_t)max_slots;
    else if (grid_waves > 0) cap = (size_t)std::max(1, grid_waves / kPsWaves);
    else
        while (cap < (size_t)kPsMaxCachedSlots && cap * 2 * per <= kPsSlotSetBytes) cap *= 2;
    size_t G = std::max<size_t>(1, std::min({K, cap, budget / per}));
    const bool wide = bc && G * n * (sizeof(PsRec) + 8) > kPsWideBytes;
    if (wide) {
        per = slot_bytes(true);
        G = std::max<size_t>(1, std::min(G, budget / per));
    }
    return PsPlan{G, wide, per};
}

static int run_persource(cent_ctx* ctx, const cent_graph* g, const BatchSpec& sp, int32_t k, KernelTimer& tm,
                         int64_t* levels) {
    const DevCSR& d = g->d;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)(d.n > 0 ? d.n : 1);
    const bool bc = sp.with_bc, compat = sp.compat;
    if (sp.bwgt || k < 1) return CENT_E_UNSUPPORTED;
    const int lcap = (int)std::min<size_t>(n + 2, 65536);
    const size_t nw = ((n + 31) / 32 + 3) / 4 * 4;  // closeness map words, padded for 16-byte clears
    size_t freeb = 0, totalb = 0;
    HIPTRY(hipMemGetInfo(&freeb, &totalb));
    const size_t have = ctx->ps_rec.cap + 8 * ctx->ps_bcp.cap + 4 * (ctx->ps_Q.cap + ctx->ps_loff.cap + ctx->ps_bits.cap);
    const size_t budget = std::min<size_t>((size_t)64 << 30, (freeb + have) / 5 * 2);
    const PsPlan plan = ps_plan(n, bc, (size_t)k, sp.opts ? sp.opts->max_slots : 0,
                                sp.opts ? sp.opts->grid_waves : 0, budget);
    const size_t G = plan.G;
    const bool wide = plan.wide;
    TRY(ensure_exact(ctx, ctx->ps_Q, n * G));
    if (bc) {
        TRY(ensure_exact(ctx, ctx->ps_rec, n * G * (wide ? sizeof(PsRecW) : sizeof(PsRec))));
        if (!wide) TRY(ensure_exact(ctx, ctx->ps_bcp, n * G));
        TRY(ensure_exact(ctx, ctx->ps_loff, (size_t)lcap * 2 * G));  // level offsets + count shifts
    } else {
        TRY(ensure_exact(ctx, ctx->ps_bits, nw * G));
    }
    TRY(ensure(ctx, ctx->ps_depth, 1));
    TRY(ensure(ctx, ctx->flags, 2));
    if (bc) {
        const size_t total = n * G;
        const int ig = (int)std::min<size_t>((total + 255) / 256, 16384);
        if (wide) k_ps_init<true><<<ig, 256, 0, s>>>(ctx->ps_rec.p, nullptr, total);
        else k_ps_init<false><<<ig, 256, 0, s>>>(ctx->ps_rec.p, ctx->ps_bcp.p, total);
        HIPTRY(hipGetLastError());
    }
    HIPTRY(hipMemsetAsync(ctx->ps_depth.p, 0, sizeof(unsigned long long), s));
    HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
    PsArgs a{};
    a.n = d.n; a.nw = (int)nw; a.rp = d.rp; a.col = d.col; a.ell = d.ell; a.src = ctx->src.p; a.k = k;
    a.Q = ctx->ps_Q.p;
    if (bc) {
        a.rec = ctx->ps_rec.p;
        a.bcp = wide ? nullptr : ctx->ps_bcp.p;
        a.loff = ctx->ps_loff.p;
        a.lcap = lcap;
    } else {
        a.bits = ctx->ps_bits.p;
    }
    a.far = ctx->far.p; a.dist_sum = sp.dist_sum; a.weight = sp.weight;
    a.cnt = ctx->cnt.p; a.depth_sum = ctx->ps_depth.p; a.overflow = ctx->flags.p;
    hipEvent_t e0;
    TRY(timer_begin(ctx, tm, &e0));
    if (!bc) k_ps_cc<<<(int)G, kPsBlock, 0, s>>>(a);
    else if (compat && wide) k_ps_bc<true, true><<<(int)G, kPsBlock, 0, s>>>(a);
    else if (compat) k_ps_bc<true, false><<<(int)G, kPsBlock, 0, s>>>(a);
    else if (wide) k_ps_bc<false, true><<<(int)G, kPsBlock, 0, s>>>(a);
    else k_ps_bc<false, false><<<(int)G, kPsBlock, 0, s>>>(a);
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
        if (wide) k_ps_reduce<true><<<(int)((n + 255) / 256), 256, 0, s>>>(ctx->ps_rec.p, nullptr, (int)G, d.n, sp.bc);
        else k_ps_reduce<false><<<(int)((n + 255) / 256), 256, 0, s>>>(ctx->ps_rec.p, ctx->ps_bcp.p, (int)G, d.n, sp.bc);
        HIPTRY(hipGetLastError());
    }
    *levels += (int64_t)depth;
    ctx->stats.batches += k;
    return CENT_OK;
}

// Adds the peer stream's partial results into this stream's: BC and
// distance sums per vertex (null: not computed), the two counters, the level
// statistics and the overflow flag.  One fixed order per vertex (this
// stream's sum, then the peer's), so a call's BC does not depend on timing.
__global__ void k_join(double* __restrict__ bc, const double* __restrict__ pbc, int64_t* __restrict__ ds,
                       const int64_t* __restrict__ pds, int n, unsigned long long* __restrict__ cnt,
                       const unsigned long long* __restrict__ pcnt, unsigned long long* __restrict__ lvst,
                       const unsigned long long* __restrict__ plvst, int* __restrict__ flags,
                       const int* __restrict__ pflags) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        if (bc) bc[v] += pbc[v];
        if (ds) ds[v] += pds[v];
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x < 2) cnt[threadIdx.x] += pcnt[threadIdx.x];
        if (threadIdx.x < 4) lvst[threadIdx.x] += plvst[threadIdx.x];
        if (threadIdx.x == 0) flags[1] |= pflags[1];
    }
}

int peer_get(cent_ctx* ctx, cent_ctx** out) {
    if (!ctx->peer) {
        cent_ctx* p = nullptr;
        TRY(cent_ctx_create(ctx->device, &p));
        ctx->peer = p;
    }
    for (auto& e : ctx->ev_join)
        if (!e) HIPTRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *out = ctx->peer;
    return CENT_OK;
}

// Results into caller-order device arrays: out[v] += in[new_id[v]]
// (identity when new_id is null).
template <typename T>
__global__ void k_out_accum(T* __restrict__ out, const T* __restrict__ in, const int32_t* __restrict__ new_id,
                            int n) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        out[v] += in[new_id ? new_id[v] : v];
}

// out[order[i]] = far[i] (order null: identity).
__global__ void k_out_far(int64_t* __restrict__ out, const int64_t* __restrict__ far,
                          const int32_t* __restrict__ order, int k) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
        out[order ? order[i] : i] = far[i];
}

// Checks that [p, p + bytes) lies inside one device allocation of the
// context's GPU.
static int check_dev_range(cent_ctx* ctx, const void* p, size_t bytes) {
    if (!p) return CENT_OK;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unregistered host pointer: not an error of the stream
        return CENT_E_INVAL;
    }
    if ((a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != ctx->device)
        return CENT_E_INVAL;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return CENT_E_INVAL;
    }
    const char* b = static_cast<const char*>(base);
    const char* q = static_cast<const char*>(p);
    if (q < b || (size_t)(q - b) + bytes > size) return CENT_E_INVAL;
    return CENT_OK;
}

// Runs all batches of one plain call.  with_bc: Brandes (sigma + backward);
// otherwise closeness only (bitmask BFS).  dev_out: bc, farness and
// dist_sum are device arrays of the context's GPU (same meaning and order
// as the host ones), written by the device with no copy to the host.
static int run_batches(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, bool with_bc,
                       int bc_mode, double* bc, int64_t* farness, int64_t* dist_sum, const cent_opts* opts,
                       bool dev_out = false) {
    const int n = g->n;
    const size_t N = (size_t)(n > 0 ? n : 1);
    hipStream_t s = ctx->stream;
    // the graph's tuned launch settings fill what the call leaves open
    cent_opts eff{};
    if (opts) eff = *opts;
    if (eff.grid_waves == 0) eff.grid_waves = g->grid_waves[with_bc ? 0 : 1];
    if (eff.frontier == CENT_FRONTIER_AUTO) eff.frontier = g->frontier[with_bc ? 0 : 1];
    if (eff.push_pct == 0) eff.push_pct = g->push_pct[with_bc ? 0 : 1];
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
    // Batch order of the sources: the 64-lane steps group sources whose
    // lanes tend to reach each vertex at the same levels, so a level's
    // gathers read fewer and fuller slices of the level-sorted rows.  Key:
    // the sum of the source's distances to the graph's landmarks (the
    // highest-degree vertices), then its highest-degree neighbour (that
    // neighbour's degree, then its id: sources hanging off one vertex see
    // the graph at nearly the same levels), then the distances in landmark
    // order, then degree (R-MAT 2^20, 4096 sources: about 12% fewer
    // 128-byte row lines per step than grouping by the distance to the top
    // vertex alone; sources in caller order: 30% more).  Without landmark
    // keys: the BFS distance to the highest-degree vertex, then degree.
    // order[i] = caller index of slot i.
    const int lm = k > 0 ? level_mode(g, opts, with_bc) : CENT_FRONTIER_DENSE;
    std::vector<int32_t> order;
    if (k > kWave && lm != CENT_FRONTIER_SOURCE && !(opts->flags & CENT_OPT_SOURCE_ORDER) &&
        !g->probe_dist.empty()) {
        order.resize((size_t)k);
        for (int32_t i = 0; i < k; ++i) order[(size_t)i] = i;
        struct Key {
            uint64_t k0, k1;
            int32_t sum, deg, tdeg, top;
        };
        std::vector<Key> key((size_t)k);
        for (int32_t i = 0; i < k; ++i) {
            const int32_t s0 = src[i];
            const int32_t x = g->new_id.empty() ? s0 : g->new_id[(size_t)s0];
            Key& q = key[(size_t)i];
            q.deg = g->h_rp[(size_t)x + 1] - g->h_rp[(size_t)x];
            q.tdeg = q.top = 0;
            if (!g->lm_key.empty()) {
                q.k0 = g->lm_key[2 * (size_t)x];
                q.k1 = g->lm_key[2 * (size_t)x + 1];
                q.sum = 0;
                for (int b = 0; b < 64; b += 4) q.sum += (int32_t)((q.k0 >> b) & 15) + (int32_t)((q.k1 >> b) & 15);
                if (!g->lm_top.empty() && g->lm_top[(size_t)x] >= 0) {
                    q.top = g->lm_top[(size_t)x];
                    q.tdeg = g->h_rp[(size_t)q.top + 1] - g->h_rp[(size_t)q.top];
                }
            } else {
                q.k0 = q.k1 = 0;
                q.sum = g->probe_dist[(size_t)s0];
            }
        }
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            const Key &ka = key[(size_t)a], &kb = key[(size_t)b];
            if (ka.sum != kb.sum) return ka.sum < kb.sum;
            if (ka.tdeg != kb.tdeg) return ka.tdeg > kb.tdeg;
            if (ka.top != kb.top) return ka.top < kb.top;
            if (ka.k0 != kb.k0) return ka.k0 < kb.k0;
            if (ka.k1 != kb.k1) return ka.k1 < kb.k1;
            if (ka.deg != kb.deg) return ka.deg > kb.deg;
            return src[a] < src[b];
        });
    }
    // pinned staging: device-order sources in, then BC, distance sums and
    // farness out.  A call that ended early on an error may have left copies
    // of it in flight, so the stream is drained before the host writes it.
    const size_t stage_words = (size_t)(k > 0 ? k : 1) + (with_bc ? N : 0) + (dist_sum ? N : 0);
    HIPTRY(hipStreamSynchronize(s));
    TRY(ensure_stage(ctx, sizeof(int64_t) * stage_words));
    if (k > 0) {
        int32_t* dsrc = static_cast<int32_t*>(ctx->h_stage);
        for (int32_t i = 0; i < k; ++i) {
            const int32_t s0 = order.empty() ? src[i] : src[order[(size_t)i]];
            dsrc[(size_t)i] = g->new_id.empty() ? s0 : g->new_id[(size_t)s0];
        }
        HIPTRY(hipMemcpyAsync(ctx->src.p, dsrc, sizeof(int) * (size_t)k, hipMemcpyHostToDevice, s));
        if (dev_out && farness && !order.empty()) {  // the batch order, for the farness scatter
            TRY(ensure(ctx, ctx->ord, (size_t)k));
            int32_t* dord = dsrc + k;
            std::copy(order.begin(), order.end(), dord);
            HIPTRY(hipMemcpyAsync(ctx->ord.p, dord, sizeof(int) * (size_t)k, hipMemcpyHostToDevice, s));
        }
    }

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
    bool sparse = batched && lmode == CENT_FRONTIER_SPARSE;
    SpPlan plan;
    if (sparse) TRY(sp_prepare(ctx, g->d, sp, plan));
    // closeness on the dense step takes up to kMultiBatch batches per launch
    // group, halved while that leaves fewer than two groups (one per stream)
    // and a group keeps eight or more.  R-MAT 2^20 from 2048 / 4096 / 8192
    // sources: 4.5 / 7.9 / 12.9-13.5 ms, against 5.0 / 8.15 / 13.9 ms when
    // halving down to four groups; all-vertex on a 148k-vertex R-MAT: 32 a
    // group 68.7 ms, 16 85.3 ms, 8 108 ms.
    int gb = kMultiBatch;
    {
        const int nbt = (int)((k + kWave - 1) / kWave);
        while (gb > 8 && (nbt + gb - 1) / gb < 2) gb /= 2;
    }
    const int32_t step = (!with_bc && !sparse) ? kWave * gb : kWave;
    // Level bound of a launch group on the dense step: a BFS from s ends
    // within d(s, hub) + ecc(hub) levels (hub = the highest-degree vertex of
    // the upload probe, ecc its BFS depth), so a group whose sources all lie
    // in the hub's component is enqueued whole, with no host check.  0: no
    // bound known (host-checked runs).
    auto group_bound = [&](int32_t b0, int kb) -> int {
        if (sparse || g->probe_dist.empty()) return 0;
        int mx = 0;
        for (int i = 0; i < kb; ++i) {
            const int32_t s0 = order.empty() ? src[b0 + i] : src[order[(size_t)(b0 + i)]];
            const int32_t d0 = g->probe_dist[(size_t)s0];
            if (d0 < 0) return 0;
            mx = std::max(mx, d0);
        }
        return mx + g->ecc_probe + 1;
    };
    TRY(ensure(ctx, ctx->lvstat, 4));
    // Dense step with two or more launch groups: the groups alternate between
    // this context's stream and its peer's, so the small levels and the
    // tails of one group's launches overlap the other's; the peer's BC,
    // distance sums and counters are added at the end (k_join).
    const bool dual = batched && !sparse && !(opts->flags & CENT_OPT_ONE_STREAM) && k > step;
    cent_ctx* peer = nullptr;
    BatchSpec spp = sp;
    if (dual) {
        TRY(peer_get(ctx, &peer));
        TRY(check_device(peer));
        if (with_bc) {
            TRY(ensure(peer, peer->bc, N));
            spp.bc = peer->bc.p;
        }
        if (sp.dist_sum) {
            TRY(ensure(peer, peer->dist_sum, N));
            spp.dist_sum = peer->dist_sum.p;
        }
        TRY(ensure(peer, peer->cnt, 2));
        TRY(ensure(peer, peer->lvstat, 4));
        peer->ev_used = 0;
    }
    const cent_stats pst0 = peer ? peer->stats : cent_stats{};
    unsigned long long* lvst = reinterpret_cast<unsigned long long*>(ctx->h_flags);  // [4], then the overflow flag
    for (int i = 0; i < 4; ++i) lvst[i] = 0;
    for (int attempt = 0; batched && attempt < 2; ++attempt) {
        // attempt 1 only runs if a bound failed, which a correct probe rules
        // out, or if a frontier-queue batch was deeper than its exponent
        // table: everything is redone with host-checked dense levels on one
        // stream
        const bool two = dual && attempt == 0;
        if (attempt == 1) {
            if (with_bc) HIPTRY(hipMemsetAsync(ctx->bc.p, 0, sizeof(double) * N, s));
            if (sp.dist_sum) HIPTRY(hipMemsetAsync(sp.dist_sum, 0, sizeof(int64_t) * N, s));
            HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
        }
        HIPTRY(hipMemsetAsync(ctx->lvstat.p, 0, sizeof(unsigned long long) * 4, s));
        TRY(level_flags_reset(ctx));
        if (two) {
            hipStream_t ps = peer->stream;
            if (with_bc) HIPTRY(hipMemsetAsync(peer->bc.p, 0, sizeof(double) * N, ps));
            if (sp.dist_sum) HIPTRY(hipMemsetAsync(peer->dist_sum.p, 0, sizeof(int64_t) * N, ps));
            HIPTRY(hipMemsetAsync(peer->cnt.p, 0, sizeof(unsigned long long) * 2, ps));
            HIPTRY(hipMemsetAsync(peer->lvstat.p, 0, sizeof(unsigned long long) * 4, ps));
            TRY(level_flags_reset(peer));
            // the peer's groups read the sources uploaded on this stream
            HIPTRY(hipEventRecord(ctx->ev_join[0], s));
            HIPTRY(hipStreamWaitEvent(ps, ctx->ev_join[0], 0));
        }
        int grp = 0;
        for (int32_t b0 = 0; b0 < k; b0 += step, ++grp) {
            const int kb = (k - b0) < step ? (k - b0) : step;
            if (sparse) {
                const int st = level_batch_sparse(ctx, plan, sp, ctx->src.p + b0, kb, ctx->far.p + b0, tm, &levels);
                if (st == CENT_E_UNSUPPORTED) {  // too deep for the exponent table
                    sparse = false;
                    lvst[1] = 1;  // forces attempt 1 below
                    break;
                }
                TRY(st);
            } else {
                const int bound = attempt == 0 ? group_bound(b0, kb) : 0;
                cent_ctx* x = (two && (grp & 1)) ? peer : ctx;
                TRY(level_batch(x, view, x == ctx ? sp : spp, ctx->src.p + b0, kb, ctx->far.p + b0, tm, &levels,
                                bound));
            }
            if (tm.on && (ctx->ev_used > 4096 || (peer && peer->ev_used > 4096))) {  // keep the event pools bounded
                HIPTRY(hipStreamSynchronize(s));
                if (peer) HIPTRY(hipStreamSynchronize(peer->stream));
                timer_collect(ctx, tm);
            }
        }
        if (two) {
            HIPTRY(hipEventRecord(ctx->ev_join[1], peer->stream));
            HIPTRY(hipStreamWaitEvent(s, ctx->ev_join[1], 0));
            const int jg = (int)std::min<size_t>((N + 255) / 256, 4096);
            k_join<<<jg, 256, 0, s>>>(with_bc ? ctx->bc.p : nullptr, peer->bc.p, sp.dist_sum,
                                      sp.dist_sum ? peer->dist_sum.p : nullptr, n, ctx->cnt.p, peer->cnt.p,
                                      ctx->lvstat.p, peer->lvstat.p, ctx->flags.p, peer->flags.p);
            HIPTRY(hipGetLastError());
        }
        if (!sparse && lmode == CENT_FRONTIER_SPARSE && attempt == 0) {
            levels = 0;
            continue;  // the sparse batches gave up: attempt 1 runs them dense
        }
        HIPTRY(hipMemcpyAsync(lvst, ctx->lvstat.p, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, s));
        HIPTRY(hipMemcpyAsync(ctx->h_flags + 8, ctx->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (ctx->h_flags[8]) return CENT_E_OVERFLOW;
        if (lvst[1] == 0) break;
    }
    if (peer) {  // the peer's launch counts belong to this call
        ctx->stats.batches += peer->stats.batches - pst0.batches;
        ctx->stats.launches_fwd += peer->stats.launches_fwd - pst0.launches_fwd;
        ctx->stats.launches_bwd += peer->stats.launches_bwd - pst0.launches_bwd;
    }
    levels += (int64_t)lvst[0];
    ctx->stats.launches_fwd += (int64_t)lvst[2];
    ctx->stats.launches_bwd += (int64_t)lvst[3];
    unsigned long long* hc = reinterpret_cast<unsigned long long*>(ctx->h_flags);
    if (dev_out) {  // results stay in device memory, in caller order
        const int32_t* nid = g->d_new_id;
        const int rb = 256;
        const int rgn = (int)std::min<size_t>((N + rb - 1) / rb, 8192);
        if (with_bc && bc) k_out_accum<double><<<rgn, rb, 0, s>>>(bc, ctx->bc.p, nid, n);
        if (dist_sum) k_out_accum<int64_t><<<rgn, rb, 0, s>>>(dist_sum, ctx->dist_sum.p, nid, n);
        if (farness && k > 0) {
            const int rgk = (int)std::min<int64_t>(((int64_t)k + rb - 1) / rb, 8192);
            k_out_far<<<rgk, rb, 0, s>>>(farness, ctx->far.p, order.empty() ? nullptr : ctx->ord.p, k);
        }
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(hc, ctx->cnt.p, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
    } else {  // results back to the host, through the pinned staging
        int64_t* hf = static_cast<int64_t*>(ctx->h_stage);  // [k]
        double* hb = reinterpret_cast<double*>(hf + (k > 0 ? k : 1));  // [N] if BC
        int64_t* hd = reinterpret_cast<int64_t*>(hb + (with_bc ? N : 0));  // [N] if distance sums
        if (with_bc && bc) HIPTRY(hipMemcpyAsync(hb, ctx->bc.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
        if (dist_sum) HIPTRY(hipMemcpyAsync(hd, ctx->dist_sum.p, sizeof(int64_t) * N, hipMemcpyDeviceToHost, s));
        if (farness && k > 0) HIPTRY(hipMemcpyAsync(hf, ctx->far.p, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToHost, s));
        HIPTRY(hipMemcpyAsync(hc, ctx->cnt.p, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (farness && k > 0)
            for (int32_t i = 0; i < k; ++i) farness[order.empty() ? i : order[(size_t)i]] = hf[(size_t)i];
        const bool remap = !g->new_id.empty();
        if (with_bc && bc)
            for (int v = 0; v < n; ++v) bc[v] += hb[remap ? (size_t)g->new_id[(size_t)v] : (size_t)v];
        if (dist_sum)
            for (int v = 0; v < n; ++v) dist_sum[v] += hd[remap ? (size_t)g->new_id[(size_t)v] : (size_t)v];
    }
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
    if (!ctx || !g || g->device != ctx->device) return CENT_E_INVAL;
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

extern "C" int cent_bc_device(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, double* d_bc,
                              int64_t* d_farness, int64_t* d_dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->device != ctx->device) return CENT_E_INVAL;
    int mode = opts ? opts->bc_mode : CENT_BC_TEXTBOOK;
    if (mode != CENT_BC_TEXTBOOK && mode != CENT_BC_REF_COMPAT) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    TRY(prepare_sources(g, src, k, all, &s, &kk));
    TRY(check_device(ctx));
    TRY(check_dev_range(ctx, d_bc, sizeof(double) * (size_t)g->n));
    TRY(check_dev_range(ctx, d_farness, sizeof(int64_t) * (size_t)kk));
    TRY(check_dev_range(ctx, d_dist_sum, sizeof(int64_t) * (size_t)g->n));
    auto t0 = std::chrono::steady_clock::now();
    int st = run_batches(ctx, g, s, kk, true, mode, d_bc, d_farness, d_dist_sum, opts, true);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

extern "C" int cent_cc_device(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                              int64_t* d_farness, int64_t* d_dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->device != ctx->device) return CENT_E_INVAL;
    std::vector<int32_t> all;
    const int32_t* s;
    int32_t kk;
    TRY(prepare_sources(g, src, k, all, &s, &kk));
    TRY(check_device(ctx));
    TRY(check_dev_range(ctx, d_farness, sizeof(int64_t) * (size_t)kk));
    TRY(check_dev_range(ctx, d_dist_sum, sizeof(int64_t) * (size_t)g->n));
    auto t0 = std::chrono::steady_clock::now();
    int st = run_batches(ctx, g, s, kk, false, 0, nullptr, d_farness, d_dist_sum, opts, true);
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

extern "C" int cent_cc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k, int64_t* farness,
                       int64_t* dist_sum, const cent_opts* opts) {
    if (!ctx || !g || g->device != ctx->device) return CENT_E_INVAL;
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

// ------------------------------------------------------------ allocation-free closeness
// cent_cc_device_noalloc: cent_cc_device on the source-parallel closeness
// kernel (k_ps_cc) with every device buffer of the call — the source list,
// the counters, the per-slot visited maps and BFS queues, the device-order
// distance sums — carved out of one file-scope device array instead of the
// context's workspace.  Nothing is allocated per call, on the device or the
// host.  The array is per GPU (each device that runs the call holds its
// own copy until the process ends), and every call on a GPU uses that GPU's
// copy: calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocMaxSlots = 256;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of this GPU's scratch (the host-side name is only a
// stand-in): looked up once per device, with that device current.
static int noalloc_scratch(cent_ctx* ctx, unsigned char** out) {
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        HIPTRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return CENT_E_DEVICE;
        g_scratch_ptr[ctx->device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[ctx->device];
    return CENT_OK;
}

// Zeroes n 64-bit words (a kernel rather than a memset node, so the call
// stays free of runtime-internal work).
__global__ void k_zero_words(unsigned long long* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Caller ids to device ids in place (new_id: caller id -> device id).
__global__ void k_remap_ids(int32_t* __restrict__ ids, int k, const int32_t* __restrict__ new_id) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) ids[i] = new_id[ids[i]];
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ids[i] = i (all-vertex calls: every vertex is a source, in caller order).
__global__ void k_iota(int32_t* __restrict__ ids, int k) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) ids[i] = i;
}

extern "C" int cent_cc_device_noalloc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                                      int64_t* d_farness, int64_t* d_dist_sum) {
    if (!ctx || !g || g->device != ctx->device || k < -1 || (k > 0 && !src)) return CENT_E_INVAL;
    const int32_t n = g->n;
    const bool all = k == -1;  // every vertex, farness indexed by vertex
    if (all) k = n;
    for (int32_t i = 0; !all && i < k; ++i)
        if (src[i] < 0 || src[i] >= n) return CENT_E_INVAL;
    TRY(check_device(ctx));
    TRY(check_dev_range(ctx, d_farness, sizeof(int64_t) * (size_t)k));
    TRY(check_dev_range(ctx, d_dist_sum, sizeof(int64_t) * (size_t)n));
    if (k == 0) return CENT_OK;
    // layout: counters | sources | farness (when the caller passes none) |
    // device-order distance sums (relabelled graphs) | per-slot queue + map
    const size_t N = (size_t)(n > 0 ? n : 1);
    const size_t nw = ((N + 31) / 32 + 3) / 4 * 4;
    const bool ds_remap = d_dist_sum && g->d_new_id;
    const size_t o_src = 256;
    const size_t o_far = o_src + align256(sizeof(int32_t) * (size_t)k);
    const size_t o_ds = o_far + (d_farness ? 0 : align256(sizeof(int64_t) * (size_t)k));
    const size_t o_slots = o_ds + (ds_remap ? align256(sizeof(int64_t) * N) : 0);
    const size_t per_slot = align256(sizeof(int32_t) * N) + align256(sizeof(uint32_t) * nw);
    if (o_slots + per_slot > kNoAllocScratchBytes) {
        snprintf(ctx->last_error, sizeof(ctx->last_error),
                 "cent_cc_device_noalloc: %d vertices and %d sources need %zu bytes of scratch, the maximum is %zu",
                 n, k, o_slots + per_slot, kNoAllocScratchBytes);
        return CENT_E_INVAL;
    }
    const size_t G = std::min<size_t>({(size_t)k, (size_t)kNoAllocMaxSlots, (kNoAllocScratchBytes - o_slots) / per_slot});
    unsigned char* base = nullptr;
    TRY(noalloc_scratch(ctx, &base));
    hipStream_t s = ctx->stream;
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base);  // [0..1] counts, [2] depth sum
    int32_t* dsrc = reinterpret_cast<int32_t*>(base + o_src);
    int64_t* far = d_farness ? d_farness : reinterpret_cast<int64_t*>(base + o_far);
    int64_t* ds = ds_remap ? reinterpret_cast<int64_t*>(base + o_ds) : d_dist_sum;
    int* Q = reinterpret_cast<int*>(base + o_slots);
    uint32_t* bits = reinterpret_cast<uint32_t*>(base + o_slots + G * align256(sizeof(int32_t) * N));
    k_zero_words<<<1, 64, 0, s>>>(ctr, 4);
    if (ds_remap) k_zero_words<<<(int)std::min<size_t>((N + 255) / 256, 4096), 256, 0, s>>>(
        reinterpret_cast<unsigned long long*>(ds), N);
    if (all) k_iota<<<(int)std::min<int64_t>(((int64_t)k + 255) / 256, 4096), 256, 0, s>>>(dsrc, k);
    else HIPTRY(hipMemcpyAsync(dsrc, src, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, s));
    if (g->d_new_id) k_remap_ids<<<(int)std::min<int64_t>(((int64_t)k + 255) / 256, 4096), 256, 0, s>>>(dsrc, k, g->d_new_id);
    HIPTRY(hipGetLastError());
    PsArgs a{};
    a.n = n; a.nw = (int)nw; a.rp = g->d.rp; a.col = g->d.col; a.ell = g->d.ell; a.src = dsrc; a.k = k;
    a.Q = Q; a.bits = bits; a.far = far; a.dist_sum = ds; a.weight = nullptr;
    a.cnt = ctr; a.depth_sum = ctr + 2; a.overflow = nullptr;  // closeness sets no flag
    // slot g's queue is Q + g n (the kernel's stride) and its map bits + g nw
    k_ps_cc<<<(int)G, kPsBlock, 0, s>>>(a);
    HIPTRY(hipGetLastError());
    if (ds_remap) {
        const int rgn = (int)std::min<size_t>((N + 255) / 256, 8192);
        k_out_accum<int64_t><<<rgn, 256, 0, s>>>(d_dist_sum, ds, g->d_new_id, n);
        HIPTRY(hipGetLastError());
    }
    HIPTRY(hipStreamSynchronize(s));
    return CENT_OK;
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
    if (!ctx || !g || !best || g->device != ctx->device) return CENT_E_INVAL;
    if (what != CENT_TUNE_BC && what != CENT_TUNE_CC) return CENT_E_INVAL;
    if (k < 1 || k > (what == CENT_TUNE_CC ? 1024 : 256) || !src || max_trials < 0) return CENT_E_INVAL;
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
    // the dense step's push/pull switch at level 2: always pulled, or pushed
    // while the level-1 rows hold under 2 / 10 / 30 % of the edges
    if (!st && what == CENT_TUNE_BC && cur.frontier == CENT_FRONTIER_DENSE)
        sweep(&cent_tune_params::push_pct, {-1, 2, 10, 30});
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

// ------------------------------------------------------------ workspace estimate
// The allocations of run_batches and the level steps, replayed on a WsEst.
void ws_dense(WsEst& w, size_t N, size_t nseg, bool bc, bool compat, int nb, int G, size_t levels) {
    const size_t W = bc ? 1 : (size_t)ccw_width(nb);
    w.add(8 * ensure_cap(N * W));                                   // vis
    size_t fcap = ensure_cap(bc ? 2 * N * 17 : N * (2 * W + 1));     // F
    w.add(8 * fcap);
    size_t fl = fcap / (N * (bc ? 2 : 1)) - 1;
    if (bc) {
        w.add(4 * (fl + 2) * kWave);                                 // lexp
        w.add(8 * ensure_cap(N * kWave));                            // sigma (coef written over it)
        if (compat) w.add(4 * ensure_cap(N * kWave));                // npred
        w.add(8 * ensure_cap(1));                                    // push_e
    }
    w.add(8 * ensure_cap((size_t)G * kWave * nb) + 8 * ensure_cap((size_t)G * 2 * nb));  // part_far, part_cnt
    w.add(8 * ensure_cap(2) + 4 * ensure_cap(2));                    // cnt, flags
    if (nseg) w.add(8 * ensure_cap(nseg * kWave) + 4 * ensure_cap(nseg * kWave) + 8 * ensure_cap(nseg * W));
    while (bc && fl < levels) {                                      // grow_levels
        const size_t ncap = fl * 2, nf = 2 * N * (ncap + 1);
        w.add(8 * nf);
        w.sub(8 * fcap);
        w.add(4 * (ncap + 2) * kWave);
        w.sub(4 * (fl + 2) * kWave);
        fcap = nf;
        fl = ncap;
    }
}

int ws_grid(size_t nchunk, int grid_waves) {
    cent_opts o{};
    o.grid_waves = grid_waves;
    return grid_for((int)std::min<size_t>(nchunk, INT32_MAX), &o);
}

size_t ws_levels(int32_t depth) {
    // a bounded launch group needs its bound + 1 levels (bound <= 2 depth + 1
    // in the probe's component); host-checked runs overshoot the deepest
    // level by up to a run of kHostFlags - 2 levels
    const size_t d = depth >= 0 ? (size_t)depth : (size_t)kDeepLevels;
    return std::max(2 * d + 2, d + kHostFlags);
}

// Peak workspace of one plain call (cent_bc / cent_cc and the _device forms)
// on a graph of n vertices with nseg hub segments and nchunk vertex chunks,
// probe depth `depth` (-1: unknown), tuned grid_waves / frontier for the
// call type (0 / AUTO: none).
static size_t ws_plain(int32_t n, size_t nseg, size_t nchunk, int32_t depth, int32_t gw_tuned,
                       int32_t fr_tuned, bool bc, int32_t k, const cent_opts* opts) {
    const size_t N = (size_t)(n > 0 ? n : 1);
    if (k < 0) k = n;
    const size_t K = (size_t)(k > 0 ? k : 1);
    cent_opts eff{};
    if (opts) eff = *opts;
    if (eff.grid_waves == 0) eff.grid_waves = gw_tuned;
    if (eff.frontier == CENT_FRONTIER_AUTO) eff.frontier = fr_tuned;
    const bool compat = bc && eff.bc_mode == CENT_BC_REF_COMPAT;
    WsEst base;
    if (bc) base.add(8 * ensure_cap(N));                             // bc
    base.add(8 * ensure_cap(N));                                     // dist_sum
    base.add(8 * ensure_cap(K) + 4 * ensure_cap(K) + 4 * ensure_cap(K));  // far, src, ord
    base.add(8 * ensure_cap(2) + 8 * ensure_cap(4) + 4 * ensure_cap(kHostFlags));  // cnt, lvstat, lvf
    std::vector<int> modes;
    if (eff.frontier != CENT_FRONTIER_AUTO) modes.push_back(eff.frontier);
    else if (depth >= 0)
        modes.push_back(depth > kDeepLevels || (bc && n <= kSmallBcVertices) ? CENT_FRONTIER_SOURCE
                                                                           : CENT_FRONTIER_DENSE);
    else
        modes = {CENT_FRONTIER_SOURCE, CENT_FRONTIER_DENSE};
    const size_t levels = ws_levels(depth);
    const int G = ws_grid(nchunk, eff.grid_waves);
    // dense step of the call (with the peer context's copy on two streams)
    auto dense = [&](WsEst& w, bool may_dual) {
        int gb = kMultiBatch;
        const int nbt = (int)((K + kWave - 1) / kWave);
        while (gb > 8 && (nbt + gb - 1) / gb < 2) gb /= 2;  // as run_batches
        const int nb = bc ? 1 : std::min(gb, nbt);
        ws_dense(w, N, nseg, bc, compat, nb, G, levels);
        const size_t step = bc ? (size_t)kWave : (size_t)kWave * gb;
        if (may_dual && !(eff.flags & CENT_OPT_ONE_STREAM) && K > step) {
            WsEst p;
            if (bc) p.add(8 * ensure_cap(N));
            p.add(8 * ensure_cap(N) + 8 * ensure_cap(2) + 8 * ensure_cap(4));
            ws_dense(p, N, nseg, bc, compat, nb, G, levels);
            w.add(p.peak);
        }
    };
    size_t best = 0;
    for (int m : modes) {
        WsEst w = base;
        if (m == CENT_FRONTIER_SOURCE) {
            const size_t lcap = std::min<size_t>(N + 2, 65536);
            // the free-memory budget only lowers the slot count
            const PsPlan plan = ps_plan(N, bc, K, eff.max_slots, eff.grid_waves, SIZE_MAX);
            w.add(plan.per_slot * plan.G + 8 * ensure_cap(1) + 4 * ensure_cap(2));
            // a betweenness BFS deeper than the level table reruns as batches
            if (bc && (depth < 0 ? false : 2 * (size_t)depth + 2 >= lcap)) dense(w, true);
        } else if (m == CENT_FRONTIER_SPARSE) {
            const size_t G2 = (size_t)std::max(1, std::min((int)((N + kBlock - 1) / kBlock), 1024));
            const size_t qcap = std::min<size_t>(N * kWave, (size_t)INT32_MAX);
            w.add(8 * ensure_cap(N) + 8 * ensure_cap(2 * N) + 4 * ensure_cap(qcap) + 8 * ensure_cap(qcap));
            w.add(4 * ensure_cap(N + 3) + 8 * ensure_cap(N) + 4 * ensure_cap(kCtlSlots) + 8 * ensure_cap(1));
            w.add(8 * ensure_cap(G2 * kWave) + 8 * ensure_cap(G2 * 2));
            if (bc) {
                w.add(16 * ensure_cap(N * kWave) + (compat ? 4 * ensure_cap(N * kWave) : 0));
                w.add(4 * (std::min<size_t>(N + 3, kSpLevels) + 2) * kWave);
            }
            // a batch deeper than the exponent table reruns on the dense step
            if (bc && (depth < 0 || 2 * (size_t)depth + 2 >= kSpLevels)) dense(w, false);
        } else {
            dense(w, true);
        }
        best = std::max(best, w.peak);
    }
    return best;
}

static size_t ws_nseg(const int32_t* rp, int32_t n, int hub_deg, size_t* nchunk) {
    size_t nseg = 0;
    for (int32_t v = 0; v < n; ++v) {
        const size_t d = (size_t)(rp[v + 1] - rp[v]);
        if (d > (size_t)hub_deg) nseg += (d + hub_deg - 1) / hub_deg;
    }
    *nchunk = (size_t)(n > 0 ? n : 1);  // at least one vertex a chunk
    return nseg;
}

extern "C" int cent_workspace_estimate(const int32_t* row_ptr, int32_t n, int64_t nnz, int32_t depth,
                                       int32_t gflags, int32_t what, int32_t k, const cent_opts* opts,
                                       int64_t* bytes) {
    if (!bytes || n < 0 || nnz < 0 || (n > 0 && !row_ptr) || (what != CENT_TUNE_BC && what != CENT_TUNE_CC) ||
        k < -1)
        return CENT_E_INVAL;
    // the layout cent_graph_create gives the graph: table entries for the
    // signature (the betweenness entry's hub threshold wins), else defaults
    int hub_deg = kHubDeg;
    cent_tune_params mine{};
    mine.frontier = CENT_FRONTIER_AUTO;
    if (!(gflags & CENT_GRAPH_NO_TUNE) && depth >= 0) {
        cent_tune_sig sig;
        TRY(cent_tune_signature(row_ptr, n, nnz, depth, &sig));
        for (int32_t w : {CENT_TUNE_CC, CENT_TUNE_BC}) {
            cent_tune_params tp;
            if (cent_tune_table_get(&sig, w, &tp) != CENT_OK) continue;
            hub_deg = tp.hub_deg;
            if (w == what) mine = tp;
        }
    }
    size_t nchunk = 0;
    const size_t nseg = ws_nseg(row_ptr, n, hub_deg, &nchunk);
    *bytes = (int64_t)ws_plain(n, nseg, nchunk, depth, mine.grid_waves, mine.frontier, what == CENT_TUNE_BC, k, opts);
    return CENT_OK;
}

extern "C" int cent_graph_workspace_estimate(const cent_graph* g, int32_t what, int32_t k, const cent_opts* opts,
                                             int64_t* bytes) {
    if (!g || !bytes || (what != CENT_TUNE_BC && what != CENT_TUNE_CC) || k < -1) return CENT_E_INVAL;
    const int i = what == CENT_TUNE_BC ? 0 : 1;
    *bytes = (int64_t)ws_plain(g->n, g->d.nhub ? (size_t)g->d.nseg : 0, (size_t)std::max(1, g->d.nchunk),
                               g->ecc_probe, g->grid_waves[i], g->frontier[i], what == CENT_TUNE_BC, k, opts);
    return CENT_OK;
}
