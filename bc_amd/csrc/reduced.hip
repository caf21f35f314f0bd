// The D1/AP-reduced form: device copies of packed block graphs, the
// persistent work-queue kernels over (block, source block) items and the
// level-synchronous path for large items.  Shared types are in engine.h.
#include <map>

#include "engine.h"
#include "items_kernels.h"
#include "share_kernels.h"

// ------------------------------------------------------------ reduced closeness
// Device copy of a packed block graph, cached on the cent_reduced object
// (freed through its dev_free hook).
struct RedDev {
    int device = 0;           // GPU holding it (no context is kept: it may be gone by the next call)
    DevCSR d;
    int64_t* w = nullptr;     // [np]
    int* iota = nullptr;      // [np] packed ids 0..np-1 (batch sources)
    int4* items = nullptr;    // [nitems]
    int* order = nullptr;     // [nsmall] small items, largest range first
    int nsmall = 0;
    int max_small = 1;        // largest range of a small item
    std::vector<int> big;     // items handled by the level-synchronous kernels
    int64_t* poff = nullptr;  // [nitems] offsets of the items' BC partials (small items)
    int64_t part_total = 0;
    std::vector<int64_t> h_poff;
    // level bounds of the big items' blocks: BFS distance of every packed
    // vertex of a big block from the block's highest-degree vertex (-1
    // elsewhere), and that BFS's depth per block (keyed by the block's lo)
    std::vector<int32_t> probe_d;
    std::map<int, int> probe_ecc;
    // neighbour-sharing units over the blocks of at most kShareLds vertices
    // (built on the first CENT_SCHED_SHARED call)
    bool share_built = false;
    int nunits = 0;
    int4* units = nullptr;  // (center, lo, hi, first neighbour)
    int* ucnt = nullptr;
    int* unbr = nullptr;
};

static void reddev_free(void* p) {
    RedDev* rd = (RedDev*)p;
    if (!rd) return;
    (void)hipSetDevice(rd->device);
    devcsr_free(rd->d);
    if (rd->w) (void)hipFree(rd->w);
    if (rd->iota) (void)hipFree(rd->iota);
    if (rd->items) (void)hipFree(rd->items);
    if (rd->order) (void)hipFree(rd->order);
    if (rd->poff) (void)hipFree(rd->poff);
    if (rd->units) (void)hipFree(rd->units);
    if (rd->ucnt) (void)hipFree(rd->ucnt);
    if (rd->unbr) (void)hipFree(rd->unbr);

    delete rd;
}

static std::mutex g_reddev_mu;  // first-use creation of a reduced graph's device copy

static int reddev_build(cent_ctx* ctx, const cent_reduced* r, RedDev* rd);

static int reddev_get(cent_ctx* ctx, cent_reduced* r, RedDev** out) {
    std::lock_guard<std::mutex> lk(g_reddev_mu);
    if (r->dev) {
        RedDev* rd = (RedDev*)r->dev;
        if (rd->device != ctx->device) return CENT_E_INVAL;
        *out = rd;
        return CENT_OK;
    }
    RedDev* rd = new (std::nothrow) RedDev();
    if (!rd) return CENT_E_OOM;
    rd->device = ctx->device;
    // built privately and published on r only once every upload succeeded:
    // a half-built copy (no items uploaded) must never be found by a later call
    const int st = reddev_build(ctx, r, rd);
    if (st != CENT_OK) {
        reddev_free(rd);
        return st;
    }
    r->dev = rd;
    r->dev_free = reddev_free;
    *out = rd;
    return CENT_OK;
}

static int reddev_build(cent_ctx* ctx, const cent_reduced* r, RedDev* rd) {
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
    if (!small.empty()) rd->max_small = std::max(1, r->items[small[0]].hi - r->items[small[0]].lo);
    TRY(upload(ctx, &rd->items, items.data(), (size_t)r->nitems));
    TRY(upload(ctx, &rd->order, small.data(), small.size()));
    rd->h_poff.assign((size_t)(r->nitems > 0 ? r->nitems : 1), -1);
    for (int i : small) {
        rd->h_poff[(size_t)i] = rd->part_total;
        rd->part_total += r->items[i].hi - r->items[i].lo;
    }
    TRY(upload(ctx, &rd->poff, rd->h_poff.data(), rd->h_poff.size()));
    // probe BFS per big block (host, linear in the block)
    rd->probe_d.assign(NP, -1);
    std::vector<int32_t> q;
    for (int i : rd->big) {
        const int lo = r->items[i].lo, hi = r->items[i].hi;
        if (rd->probe_ecc.count(lo)) continue;
        int h = lo;
        for (int x = lo; x < hi; ++x)
            if (r->prp[x + 1] - r->prp[x] > r->prp[h + 1] - r->prp[h]) h = x;
        q.assign(1, h);
        rd->probe_d[(size_t)h] = 0;
        int depth = 0;
        for (size_t qh = 0; qh < q.size(); ++qh) {
            const int x = q[qh];
            depth = rd->probe_d[(size_t)x];
            for (int64_t j = r->prp[x]; j < r->prp[x + 1]; ++j) {
                const int y = r->pcol[j];
                if (rd->probe_d[(size_t)y] < 0) {
                    rd->probe_d[(size_t)y] = depth + 1;
                    q.push_back(y);
                }
            }
        }
        rd->probe_ecc[lo] = depth;
    }
    return CENT_OK;
}

// Units of the neighbour-sharing schedule, in the reference's order
// (CC.cpp:1779-1866): inside each block, sources by decreasing degree (the
// packed order); a source not yet answered with at least 3 unanswered
// neighbours answers up to 32 of them too (first in row order), otherwise
// it is a unit alone.  Blocks larger than kShareLds stay with the 64-lane
// big-item path.  Largest blocks first, for the work queue.
static int reddev_share(cent_ctx* ctx, const cent_reduced* r, RedDev* rd) {
    std::lock_guard<std::mutex> lk(g_reddev_mu);
    if (rd->share_built) return CENT_OK;
    std::vector<int> blocks;
    for (int b = 0; b < r->nblocks; ++b)
        if (r->boff[b + 1] - r->boff[b] <= kShareLds) blocks.push_back(b);
    std::stable_sort(blocks.begin(), blocks.end(), [&](int x, int y) {
        return r->boff[x + 1] - r->boff[x] > r->boff[y + 1] - r->boff[y];
    });
    std::vector<int4> units;
    std::vector<int> cnt, nbr;
    std::vector<uint8_t> done((size_t)(r->np > 0 ? r->np : 1), 0);
    for (int b : blocks) {
        const int lo = r->boff[b], hi = r->boff[b + 1];
        for (int v = lo; v < hi; ++v) {
            if (done[(size_t)v]) continue;
            done[(size_t)v] = 1;
            int undone = 0;
            for (int j = r->prp[v]; j < r->prp[v + 1]; ++j) undone += !done[(size_t)r->pcol[j]];
            const int off = (int)nbr.size();
            int k = 0;
            if (undone >= 3) {
                for (int j = r->prp[v]; j < r->prp[v + 1] && k < kShareMax; ++j) {
                    const int y = r->pcol[j];
                    if (done[(size_t)y]) continue;
                    done[(size_t)y] = 1;
                    nbr.push_back(y);
                    ++k;
                }
            }
            units.push_back(make_int4(v, lo, hi, off));
            cnt.push_back(k);
        }
    }
    int st = upload(ctx, &rd->units, units.data(), units.size());
    if (!st) st = upload(ctx, &rd->ucnt, cnt.data(), cnt.size());
    if (!st) st = upload(ctx, &rd->unbr, nbr.data(), nbr.size());
    if (st) {  // leave no half-built plan behind: the next call rebuilds it
        for (int** p : {(int**)&rd->units, &rd->ucnt, &rd->unbr}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        return st;
    }
    rd->nunits = (int)units.size();
    rd->share_built = true;
    return CENT_OK;
}

// Level bound of a group of big-item sources [s0, s0 + ns) in the block
// starting at lo: d(s, probe) + ecc(probe) + 1 over the group (0: none known).
static int red_bound(const RedDev* rd, int lo, int s0, int ns) {
    auto it = rd->probe_ecc.find(lo);
    if (it == rd->probe_ecc.end()) return 0;
    int mx = 0;
    for (int i = s0; i < s0 + ns; ++i) {
        const int d0 = rd->probe_d[(size_t)i];
        if (d0 < 0) return 0;
        mx = std::max(mx, d0);
    }
    return mx + it->second + 1;
}

// The big items' launch groups (lo, hi, s0, ns): each enqueued whole under
// its block's level bound, the groups alternating between this context's
// stream and its peer's (one stream with CENT_OPT_ONE_STREAM).  Every group
// writes its own slice D[s0 .. s0 + ns), so the streams share no output.
// If a bound ever failed (a correct probe rules it out), all groups are
// redone with host-checked levels on one stream.
static int run_big_groups(cent_ctx* ctx, const RedDev* rd, const BatchSpec& sp, const std::vector<int4>& groups,
                          int64_t* D, const cent_opts* opts, KernelTimer& tm, int64_t* levels) {
    if (groups.empty()) return CENT_OK;
    hipStream_t s = ctx->stream;
    const bool dual = groups.size() >= 2 && !(opts && (opts->flags & CENT_OPT_ONE_STREAM));
    cent_ctx* peer = nullptr;
    if (dual) {
        TRY(peer_get(ctx, &peer));
        TRY(check_device(peer));
        TRY(ensure(peer, peer->lvstat, 4));
        TRY(ensure(peer, peer->cnt, 2));
        peer->ev_used = 0;
    }
    const cent_stats pst0 = peer ? peer->stats : cent_stats{};
    TRY(ensure(ctx, ctx->lvstat, 4));
    unsigned long long* lv = reinterpret_cast<unsigned long long*>(ctx->h_flags);  // [0..3] ctx, [4..7] peer
    for (int attempt = 0; attempt < 2; ++attempt) {
        const bool two = dual && attempt == 0;
        HIPTRY(hipMemsetAsync(ctx->lvstat.p, 0, sizeof(unsigned long long) * 4, s));
        if (two) {
            HIPTRY(hipMemsetAsync(peer->lvstat.p, 0, sizeof(unsigned long long) * 4, peer->stream));
            HIPTRY(hipMemsetAsync(peer->cnt.p, 0, sizeof(unsigned long long) * 2, peer->stream));
            TRY(level_flags_reset(peer));
            // D was cleared on this stream
            HIPTRY(hipEventRecord(ctx->ev_join[0], s));
            HIPTRY(hipStreamWaitEvent(peer->stream, ctx->ev_join[0], 0));
        }
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const int4 g = groups[gi];
            cent_ctx* x = (two && (gi & 1)) ? peer : ctx;
            const int bound = attempt == 0 ? red_bound(rd, g.x, g.z, g.w) : 0;
            const GView v = sub_view(rd->d, g.x, g.y);
            TRY(level_batch(x, v, sp, rd->iota + g.z, g.w, D + g.z, tm, levels, bound));
            if (tm.on && (ctx->ev_used > 4096 || (peer && peer->ev_used > 4096))) {
                HIPTRY(hipStreamSynchronize(s));
                if (peer) HIPTRY(hipStreamSynchronize(peer->stream));
                timer_collect(ctx, tm);
            }
        }
        for (int i = 0; i < 8; ++i) lv[i] = 0;
        if (two) {
            HIPTRY(hipEventRecord(ctx->ev_join[1], peer->stream));
            HIPTRY(hipStreamWaitEvent(s, ctx->ev_join[1], 0));
            HIPTRY(hipMemcpyAsync(lv + 4, peer->lvstat.p, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, s));
        }
        HIPTRY(hipMemcpyAsync(lv, ctx->lvstat.p, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        *levels += (int64_t)(lv[0] + lv[4]);
        ctx->stats.launches_fwd += (int64_t)(lv[2] + lv[6]);
        ctx->stats.launches_bwd += (int64_t)(lv[3] + lv[7]);
        if (lv[1] + lv[5] == 0) break;
    }
    if (peer) ctx->stats.batches += peer->stats.batches - pst0.batches;
    return CENT_OK;
}

extern "C" int cent_cc_reduced(cent_ctx* ctx, const cent_reduced* rc, int64_t* farness, int32_t sched,
                               const cent_opts* opts) {
    if (!ctx || !rc || !farness) return CENT_E_INVAL;
    if (sched != CENT_SCHED_PERSISTENT && sched != CENT_SCHED_PER_ITEM && sched != CENT_SCHED_SHARED)
        return CENT_E_INVAL;
    cent_reduced* r = const_cast<cent_reduced*>(rc);  // the device cache lives on the object
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    RedDev* rd;
    TRY(reddev_get(ctx, r, &rd));
    if (sched == CENT_SCHED_SHARED) TRY(reddev_share(ctx, r, rd));
    hipStream_t s = ctx->stream;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    // per-run buffers live in the context, so several contexts can share rd
    TRY(ensure(ctx, ctx->red_D, NP));
    TRY(ensure(ctx, ctx->red_head, 1));
    int64_t* Dd = ctx->red_D.p;
    HIPTRY(hipMemsetAsync(Dd, 0, sizeof(int64_t) * NP, s));
    TRY(ensure(ctx, ctx->cnt, 2));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
    TRY(level_flags_reset(ctx));
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
        if (sched == CENT_SCHED_SHARED) {
            if (rd->nunits) {
                HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int), s));
                ShareArgs sa;
                sa.rp = rd->d.rp; sa.col = rd->d.col; sa.weight = rd->w; sa.units = rd->units; sa.ucnt = rd->ucnt;
                sa.nbr = rd->unbr; sa.nunits = rd->nunits; sa.head = ctx->red_head.p; sa.D = Dd;
                const int grid = std::min(rd->nunits, 1024);
                hipEvent_t e0;
                TRY(timer_begin(ctx, tm, &e0));
                k_cc_share<<<grid, kShareBlock, 0, s>>>(sa);
                HIPTRY(hipGetLastError());
                TRY(timer_end(ctx, tm, e0, true));
                ctx->stats.launches_fwd++;
                ctx->stats.batches += rd->nunits;
            }
        } else if (rd->nsmall) {
            HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int), s));
            ItemArgs ia;
            ia.rp = rd->d.rp; ia.col = rd->d.col; ia.weight = rd->w; ia.items = rd->items; ia.order = rd->order;
            ia.nitems = rd->nsmall; ia.head = ctx->red_head.p; ia.D = Dd;
            const int grid = std::min(rd->nsmall, 2048);
            hipEvent_t e0;
            TRY(timer_begin(ctx, tm, &e0));
            k_cc_items<<<grid, kCcItemBlock, 0, s>>>(ia);
            HIPTRY(hipGetLastError());
            TRY(timer_end(ctx, tm, e0, true));
            ctx->stats.launches_fwd++;
            ctx->stats.batches += rd->nsmall;
        }
        // large items: consecutive full items of one block (contiguous
        // sources) go up to kMultiBatch at a time through one level loop
        std::vector<int4> groups;  // (lo, hi, s0, ns)
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
            groups.push_back(make_int4(it.lo, it.hi, it.s0, ns));
            j = e;
        }
        TRY(run_big_groups(ctx, rd, sp, groups, Dd, opts, tm, &levels));
    }
    // the distance sums come back through the context's pinned staging
    TRY(ensure_stage(ctx, sizeof(int64_t) * NP));
    int64_t* D = static_cast<int64_t*>(ctx->h_stage);
    HIPTRY(hipMemcpyAsync(D, Dd, sizeof(int64_t) * NP, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    timer_collect(ctx, tm);
    TRY(cent_reduced_finish(r, D, farness));
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
    if (rc->multi) return CENT_E_UNSUPPORTED;  // parallel edges multiply path counts: use cent_bc
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
    TRY(level_flags_reset(ctx));
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
        const size_t slot = (size_t)rd->max_small * kWave;
        TRY(ensure(ctx, ctx->it_sig, slot * grid));
        TRY(ensure(ctx, ctx->it_coef, slot * grid));
        TRY(ensure(ctx, ctx->it_lvl, slot * grid));
        TRY(ensure(ctx, ctx->red_part, (size_t)(rd->part_total > 0 ? rd->part_total : 1)));
        TRY(ensure(ctx, ctx->flags, 2));
        HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int) * 2, s));
        HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
        BcItemArgs ba{};
        ba.rp = rd->d.rp; ba.col = rd->d.col; ba.weight = rd->w; ba.items = rd->items; ba.order = rd->order;
        ba.poff = rd->poff; ba.nitems = rd->nsmall; ba.head = ctx->red_head.p; ba.D = Dd; ba.stride = slot;
        ba.bcpart = ctx->red_part.p; ba.sig = ctx->it_sig.p; ba.coef = ctx->it_coef.p; ba.lvl = ctx->it_lvl.p;
        ba.deep = ctx->red_deep.p; ba.ndeep = ctx->red_head.p + 1; ba.overflow = ctx->flags.p + 1;
        hipEvent_t e0;
        TRY(timer_begin(ctx, tm, &e0));
        k_bc_items<<<grid, kBcItemBlock, 0, s>>>(ba);
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


// Betweenness from a sampled source multiset through the reduction: weighted
// Brandes inside each block from the copies with sources behind them only
// (source weight WS, target weight W), then the host adds the piece terms
// (cent_reduced_bc_sources).  On a graph of many blocks most blocks need one
// or two traversals instead of one per source.
extern "C" int cent_bc_reduced_sources(cent_ctx* ctx, const cent_reduced* rc, const int32_t* src, int32_t k,
                                       double* bc, int64_t* farness, int32_t sched, const cent_opts* opts) {
    if (!ctx || !rc || !bc || k < 0 || (k > 0 && !src)) return CENT_E_INVAL;
    if (sched != CENT_SCHED_PERSISTENT && sched != CENT_SCHED_PER_ITEM) return CENT_E_INVAL;
    for (int32_t i = 0; i < k; ++i)
        if (src[i] < 0 || src[i] >= rc->n) return CENT_E_INVAL;
    if (rc->multi) return CENT_E_UNSUPPORTED;  // parallel edges multiply path counts: use cent_bc
    cent_reduced* r = const_cast<cent_reduced*>(rc);
    TRY(check_device(ctx));
    auto t0 = std::chrono::steady_clock::now();
    RedDev* rd;
    TRY(reddev_get(ctx, r, &rd));
    hipStream_t s = ctx->stream;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    // traversals: copies with sources behind them, plus (for farness) each
    // non-root block's head copy and the head vertex's copy in the parent
    std::vector<int64_t> WS(NP, 0), wsf((size_t)(r->n > 0 ? r->n : 1)), Sc((size_t)(r->ncomp > 0 ? r->ncomp : 1));
    TRY(cent_reduced_source_weights_all(r, src, k, WS.data(), wsf.data(), Sc.data()));
    std::vector<char> need(NP, 0);
    for (int32_t x = 0; x < r->np; ++x) need[(size_t)x] = WS[(size_t)x] > 0;
    if (farness)
        for (int32_t b = 0; b < r->nblocks; ++b)
            if (r->bparent[b] >= 0) {
                need[(size_t)r->bhead[b]] = 1;
                need[(size_t)r->rep[r->porig[r->bhead[b]]]] = 1;
            }
    // items over contiguous block ranges, at most 64 listed copies each;
    // blocks beyond the LDS window go through the level kernels
    std::vector<int> list;
    std::vector<int4> items;       // (lo, hi, offset into list, count)
    std::vector<int4> big;
    for (int32_t b = 0; b < r->nblocks; ++b) {
        const int lo = r->boff[b], hi = r->boff[b + 1];
        const int first = (int)list.size();
        for (int x = lo; x < hi; ++x)
            if (need[(size_t)x]) list.push_back(x);
        const int cnt = (int)list.size() - first;
        if (!cnt) continue;
        for (int o = 0; o < cnt; o += kWave) {
            const int ns = std::min(kWave, cnt - o);
            if (hi - lo > kItemLds) {
                big.push_back(make_int4(lo, hi, first + o, ns));
                continue;
            }
            if (o == 0 && !items.empty()) {  // join the previous item when adjacent and small enough
                int4& p = items.back();
                if (p.y == lo && p.z + p.w == first && p.w + ns <= kWave && hi - p.x <= kItemLds) {
                    p.y = hi;
                    p.w += ns;
                    continue;
                }
            }
            items.push_back(make_int4(lo, hi, first + o, ns));
        }
    }
    const size_t L = list.size() > 0 ? list.size() : 1;
    std::vector<int64_t> poff(items.size() + 1, 0);
    for (size_t i = 0; i < items.size(); ++i) poff[i + 1] = poff[i] + (items[i].y - items[i].x);
    std::vector<int> order(items.size());
    for (size_t i = 0; i < items.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (items[(size_t)a].y - items[(size_t)a].x) > (items[(size_t)b].y - items[(size_t)b].x);
    });
    TRY(ensure(ctx, ctx->red_D, L));
    TRY(ensure(ctx, ctx->red_bcp, NP));
    TRY(ensure(ctx, ctx->red_head, 2));
    TRY(ensure(ctx, ctx->red_src, L));
    TRY(ensure(ctx, ctx->red_swgt, NP));
    TRY(ensure(ctx, ctx->red_items, items.size() + 1));
    TRY(ensure(ctx, ctx->red_order, items.size() + 1));
    TRY(ensure(ctx, ctx->red_poff, items.size() + 1));
    TRY(ensure(ctx, ctx->red_deep, items.size() + 1));
    TRY(ensure(ctx, ctx->cnt, 2));
    TRY(ensure(ctx, ctx->flags, 2));
    int64_t* Dd = ctx->red_D.p;
    HIPTRY(hipMemsetAsync(Dd, 0, sizeof(int64_t) * L, s));
    HIPTRY(hipMemsetAsync(ctx->red_bcp.p, 0, sizeof(double) * NP, s));
    HIPTRY(hipMemsetAsync(ctx->cnt.p, 0, sizeof(unsigned long long) * 2, s));
    HIPTRY(hipMemsetAsync(ctx->red_head.p, 0, sizeof(int) * 2, s));
    HIPTRY(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 2, s));
    if (!list.empty()) HIPTRY(hipMemcpyAsync(ctx->red_src.p, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->red_swgt.p, WS.data(), sizeof(int64_t) * NP, hipMemcpyHostToDevice, s));
    if (!items.empty()) {
        HIPTRY(hipMemcpyAsync(ctx->red_items.p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, s));
        HIPTRY(hipMemcpyAsync(ctx->red_order.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, s));
        HIPTRY(hipMemcpyAsync(ctx->red_poff.p, poff.data(), sizeof(int64_t) * items.size(), hipMemcpyHostToDevice, s));
    }
    TRY(level_flags_reset(ctx));
    KernelTimer tm;
    tm.on = opts && (opts->flags & CENT_OPT_TIMING);
    ctx->ev_used = 0;
    BatchSpec sp;
    sp.with_bc = true;
    sp.weight = rd->w;
    sp.bwgt = rd->w;
    sp.swgt = ctx->red_swgt.p;
    sp.bc = ctx->red_bcp.p;
    sp.opts = opts;
    int64_t levels = 0;
    auto run_range = [&](const int4& it) -> int {
        const GView v = sub_view(rd->d, it.x, it.y);
        return level_batch(ctx, v, sp, ctx->red_src.p + it.z, it.w, Dd + it.z, tm, &levels);
    };
    std::vector<int> deep;
    const bool persistent = sched == CENT_SCHED_PERSISTENT && !items.empty();
    if (persistent) {
        const int grid = std::min((int)items.size(), kBcItemGrid);
        int max_range = 1;
        for (const int4& it : items) max_range = std::max(max_range, it.y - it.x);
        const size_t slot = (size_t)max_range * kWave;
        TRY(ensure(ctx, ctx->it_sig, slot * grid));
        TRY(ensure(ctx, ctx->it_coef, slot * grid));
        TRY(ensure(ctx, ctx->it_lvl, slot * grid));
        TRY(ensure(ctx, ctx->red_part, (size_t)(poff.back() > 0 ? poff.back() : 1)));
        BcItemArgs ba{};
        ba.rp = rd->d.rp; ba.col = rd->d.col; ba.weight = rd->w; ba.items = ctx->red_items.p;
        ba.order = ctx->red_order.p; ba.poff = ctx->red_poff.p; ba.nitems = (int)items.size();
        ba.head = ctx->red_head.p; ba.D = Dd; ba.bcpart = ctx->red_part.p; ba.sig = ctx->it_sig.p; ba.stride = slot;
        ba.coef = ctx->it_coef.p; ba.lvl = ctx->it_lvl.p; ba.deep = ctx->red_deep.p; ba.ndeep = ctx->red_head.p + 1;
        ba.overflow = ctx->flags.p + 1; ba.srcs = ctx->red_src.p; ba.swgt = ctx->red_swgt.p;
        hipEvent_t e0;
        TRY(timer_begin(ctx, tm, &e0));
        k_bc_items<<<grid, kBcItemBlock, 0, s>>>(ba);
        HIPTRY(hipGetLastError());
        TRY(timer_end(ctx, tm, e0, true));
        ctx->stats.launches_fwd++;
        ctx->stats.batches += (int64_t)items.size();
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
        for (int i : deep) TRY(run_range(items[(size_t)i]));
    } else {
        for (const int4& it : items) TRY(run_range(it));
    }
    for (const int4& it : big) TRY(run_range(it));
    std::vector<int64_t> Dl(L);
    std::vector<double> bcp(NP), part;
    HIPTRY(hipMemcpyAsync(Dl.data(), Dd, sizeof(int64_t) * L, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(bcp.data(), ctx->red_bcp.p, sizeof(double) * NP, hipMemcpyDeviceToHost, s));
    if (persistent && poff.back() > 0) {
        part.resize((size_t)poff.back());
        HIPTRY(hipMemcpyAsync(part.data(), ctx->red_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    timer_collect(ctx, tm);
    if (persistent) {  // the items' partials in item order (deep items were redone above)
        std::vector<char> is_deep(items.size() + 1, 0);
        for (int i : deep) is_deep[(size_t)i] = 1;
        for (size_t i = 0; i < items.size(); ++i) {
            if (is_deep[i]) continue;
            const int4& it = items[i];
            for (int x = it.x; x < it.y; ++x) bcp[(size_t)x] += part[(size_t)(poff[i] + x - it.x)];
        }
    }
    TRY(cent_reduced_bc_sources_from(r, WS.data(), wsf.data(), Sc.data(), bcp.data(), bc));
    if (farness && k > 0) {
        std::vector<int64_t> D(NP, 0);
        for (size_t i = 0; i < list.size(); ++i) D[(size_t)list[i]] = Dl[i];
        TRY(cent_reduced_farness_sources(r, D.data(), src, k, farness));
    }
    ctx->stats.levels += levels;
    ctx->stats.reached += k;
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CENT_OK;
}

// ------------------------------------------------------------ workspace estimate
// The allocations of cent_cc_reduced / cent_bc_reduced /
// cent_bc_reduced_sources replayed on a WsEst (levels.hip); the reduced
// graph's own device copy is not included.
extern "C" int cent_reduced_workspace_estimate(const cent_reduced* r, int32_t what, int32_t k, int32_t sched,
                                               const cent_opts* opts, int64_t* bytes) {
    if (!r || !bytes || (what != CENT_TUNE_BC && what != CENT_TUNE_CC) || k < -1 ||
        (sched != CENT_SCHED_PERSISTENT && sched != CENT_SCHED_PER_ITEM && sched != CENT_SCHED_SHARED))
        return CENT_E_INVAL;
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    size_t nseg = 0;
    for (int32_t x = 0; x < r->np; ++x) {
        const size_t d = (size_t)(r->prp[x + 1] - r->prp[x]);
        if (d > (size_t)kHubDeg) nseg += (d + kHubDeg - 1) / kHubDeg;
    }
    size_t nsmall = 0, nbig = 0, max_small = 1;
    for (int32_t i = 0; i < r->nitems; ++i) {
        const size_t R = (size_t)(r->items[i].hi - r->items[i].lo);
        if (R <= (size_t)kItemLds) {
            ++nsmall;
            max_small = std::max(max_small, R);
        } else {
            ++nbig;
        }
    }
    const int G = ws_grid(NP, opts ? opts->grid_waves : 0);
    const size_t levels = ws_levels(-1);
    const bool one = opts && (opts->flags & CENT_OPT_ONE_STREAM);
    WsEst w;
    w.add(8 * ensure_cap(2) + 4 * ensure_cap(2) + 8 * ensure_cap(4) + 4 * ensure_cap(kHostFlags));  // cnt, flags, lvstat, lvf
    if (what == CENT_TUNE_CC) {
        w.add(8 * ensure_cap(NP) + 4 * ensure_cap(1));  // red_D, red_head
        if (sched == CENT_SCHED_PER_ITEM) {
            ws_dense(w, NP, nseg, false, false, 1, G, levels);
        } else if (nbig) {
            // groups of up to kMultiBatch big items, alternating over two streams
            const int nb = (int)std::min<size_t>(nbig, kMultiBatch);
            ws_dense(w, NP, nseg, false, false, nb, G, levels);
            if (nbig >= 2 && !one) {
                WsEst p;
                p.add(8 * ensure_cap(2) + 8 * ensure_cap(4));
                ws_dense(p, NP, nseg, false, false, nb, G, levels);
                w.add(p.peak);
            }
        }
    } else {
        // all vertices (cent_bc_reduced) or k sampled sources
        // (cent_bc_reduced_sources: at most one listed copy a source, plus
        // two a block for the farness terms; items of at most 64 copies)
        const size_t L = k < 0 ? NP : std::min(NP, (size_t)k + 2 * (size_t)std::max(1, r->nblocks));
        const size_t I = k < 0 ? (size_t)std::max(1, r->nitems) : L;
        w.add(8 * ensure_cap(L) + 8 * ensure_cap(NP) + 4 * ensure_cap(2) + 4 * ensure_cap(I));  // red_D, red_bcp, red_head, red_deep
        if (k >= 0) w.add(4 * ensure_cap(L) + 8 * ensure_cap(NP) + 16 * ensure_cap(I + 1) + 4 * ensure_cap(I + 1) +
                          8 * ensure_cap(I + 1));  // red_src, red_swgt, red_items, red_order, red_poff
        const size_t small = k < 0 ? nsmall : I;
        if (sched == CENT_SCHED_PERSISTENT && small) {
            const size_t grid = std::min<size_t>(small, kBcItemGrid);
            // sampled items may join adjacent small blocks up to the LDS window
            const size_t slot = (k < 0 ? max_small : (size_t)kItemLds) * kWave;
            // per-workgroup path counts, coefficients and levels; partial sums
            // over the items' ranges (a block's items each span the block)
            const size_t part = k < 0 ? NP : NP + (size_t)kItemLds / kWave * L;
            w.add(8 * ensure_cap(slot * grid) + 8 * ensure_cap(slot * grid) + ensure_cap(slot * grid) +
                  8 * ensure_cap(part));
        }
        ws_dense(w, NP, nseg, true, false, 1, G, levels);  // big items and items too deep for the kernel
    }
    *bytes = (int64_t)w.peak;
    return CENT_OK;
}
