// Centrality query server: continuous batching of BC / closeness requests on
// one GPU.  (The reference has no serving path; this is the request loop
// around the engine's C-ABI.)
//
// - Requests (graph, kind, source list) queue FIFO up to max_queue.
// - `slots` worker threads, each with its own engine context and HIP stream,
//   take the head request as soon as a slot is free AND its device footprint
//   (the engine's own workspace estimate for the call, and the graph's device
//   copy if not cached) fits in the HBM budget next to everything else held.
//   Requests therefore run concurrently on the device, and a new one is
//   admitted the moment a slot and memory free up, whatever the others are
//   doing.
// - The budget counts what each slot's context holds: while a request runs,
//   the larger of its estimate and the context's cached buffers; between
//   requests, the cached buffers (cent_ctx_workspace_bytes).  A context
//   above its share (budget / slots / 2) is trimmed after its request, and
//   an admission that waits on idle slots' caches has them trimmed.
// - Graphs are registered once (host copy) and uploaded on first use; device
//   copies stay cached and are evicted least-recently-used, when no running
//   request holds them, if the budget needs the room.
// - A request whose footprint alone exceeds the budget fails with CENT_E_OOM
//   instead of blocking the queue.
// - Results are per request and do not depend on what else runs: each
//   request's kernels run in order on its slot's stream, with its own
//   workspace, so its sums are accumulated in the same order alone or mixed.
//
// The DRYRUN flag replaces the engine calls by a sleep of the modelled
// duration (no device is touched, no results are produced); it exists to
// test admission, ordering and accounting on machines without a GPU.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cent.h"

namespace {

using clk = std::chrono::steady_clock;

struct GraphEntry {
    struct CSR* host = nullptr;    // owned copy
    int64_t n = 0, nnz = 0;
    int32_t depth = -1;            // BFS depth from the highest-degree vertex (what the upload probes)
    cent_graph* dev = nullptr;     // cached device CSR
    cent_reduced* red = nullptr;   // host reduction (closeness queries)
    int64_t dev_bytes = 0;         // estimated device bytes of dev + red
    int refs = 0;                  // running requests using the device copy
    uint64_t last_use = 0;
    std::mutex build_mu;           // serialises upload / reduction
};

struct Request {
    int64_t id = 0;
    int32_t gid = 0, kind = 0;
    std::vector<int32_t> src;
    bool all = false;
    int32_t state = CENT_REQ_QUEUED;
    int32_t status = CENT_OK;
    bool cancel = false;
    double t_submit = 0, t_start = 0, t_end = 0;
    int64_t footprint = 0;     // workspace bytes reserved for the request's slot
    int64_t graph_bytes = 0;   // graph bytes added by this request's upload (0 if cached)
    int32_t slot = -1;
    std::vector<double> bc;
    std::vector<int64_t> far;
};

}  // namespace

struct cent_server {
    cent_server_opts o{};
    std::vector<cent_ctx*> ctx;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> reqs;
    std::vector<GraphEntry*> graphs;
    int64_t used_ws = 0, used_graph = 0;  // used_ws = sum of held
    std::vector<int64_t> held;            // per slot: bytes its context may hold (see above)
    std::vector<char> busy;               // per slot: running a request
    bool trim_idle = false;               // an admission waits on idle slots' cached buffers
    int running = 0;
    bool stop = false;
    int64_t next_id = 1;
    uint64_t tick = 0;
    clk::time_point t0 = clk::now();
    cent_server_stats st{};
};

static double now_ms(const cent_server* s) {
    return std::chrono::duration<double, std::milli>(clk::now() - s->t0).count();
}

// Source-parallel requests (deep graphs) run at most this many per-source
// slots, so their workspace stays inside the estimate below (the engine's
// own default: fewer slots leave the GPU idle, e.g. 64 slots take 3x as long
// on a 232k-vertex block tree).
constexpr int32_t kSourceSlots = 256;

// Options every request runs with: at most kSourceSlots per-source slots,
// and one stream (the slots already run side by side; a second stream per
// slot would double its workspace).
static cent_opts request_opts() {
    cent_opts o{};
    o.max_slots = kSourceSlots;
    o.flags = CENT_OPT_ONE_STREAM;
    return o;
}

// Device bytes the request's call may hold in its context besides the graph
// (the engine's estimates, which replay its allocation rules): the plain call
// on the device graph (or, before the upload, on the host rows with the
// graph's probe depth), and for requests that go through the reduction its
// reduced call, or, before the reduction exists, the GPU reduction's scratch
// (run_request raises the reservation once the reduced form is built).
static int64_t workspace_estimate(const GraphEntry* g, const Request* r) {
    const cent_opts o = request_opts();
    const int32_t k = r->all ? -1 : (int32_t)r->src.size();
    const int32_t what = r->kind == CENT_Q_BC ? CENT_TUNE_BC : CENT_TUNE_CC;
    int64_t plain = 0, red = 0;
    if (r->kind == CENT_Q_BC || !r->all) {
        if (g->dev) (void)cent_graph_workspace_estimate(g->dev, what, k, &o, &plain);
        else (void)cent_workspace_estimate(g->host->csrV, (int32_t)g->n, g->nnz, g->depth, 0, what, k, &o, &plain);
    }
    if (r->kind == CENT_Q_BC || r->all) {
        if (g->red) (void)cent_reduced_workspace_estimate(g->red, what, k, CENT_SCHED_PERSISTENT, &o, &red);
        else (void)cent_reduce_gpu_workspace_estimate((int32_t)g->n, g->nnz, &red);
    }
    return std::max(plain, red);
}

static int64_t graph_estimate(const GraphEntry* g) {
    // CSR, then the larger of the upload's grouping-key scratch (28 bytes a
    // vertex, released before the rest is built) and the 16-byte inline
    // rows, the id map and the chunk / hub tables
    return 4 * (g->n + 1) + 4 * g->nnz + std::max<int64_t>(28 * g->n, 20 * g->n + g->n / 4 + g->nnz / 16) +
           (1 << 20);
}

static int64_t reduced_estimate(const cent_reduced* r) {
    cent_reduced_view v;
    cent_reduced_get_view(r, &v);
    return 4 * ((int64_t)v.np + 1) + 4 * v.nnzp + 12 * (int64_t)v.np + 16 * (int64_t)v.nitems + (1 << 20);
}

// Drops least-recently-used unreferenced device copies until `need` more
// bytes fit.  Called with s->mu held.  Returns true if they fit.
static bool make_room(cent_server* s, int64_t need, const GraphEntry* keep) {
    while (s->used_ws + s->used_graph + need > s->o.hbm_budget) {
        GraphEntry* victim = nullptr;
        for (GraphEntry* g : s->graphs)
            if (g != keep && g->dev_bytes > 0 && g->refs == 0 && (!victim || g->last_use < victim->last_use))
                victim = g;
        if (!victim) return false;
        if (victim->dev) cent_graph_free(victim->dev);
        if (victim->red) cent_reduced_drop_device(victim->red);
        victim->dev = nullptr;
        s->used_graph -= victim->dev_bytes;
        victim->dev_bytes = 0;
        s->st.graph_evictions++;
    }
    return true;
}

// Betweenness through the reduction traverses, inside each block, only the
// copies with sources behind them; the plain call traverses the whole graph
// once per source.  The reduced form is taken when its edge count is under
// half the plain one (block trees: a few percent; a graph with one giant
// block: about the same, so plain).
static bool reduced_bc_pays(const GraphEntry* g, const Request* r) {
    cent_reduced_view v;
    if (cent_reduced_get_view(g->red, &v) != CENT_OK || v.np <= 0) return false;
    const int64_t k = r->all ? g->n : (int64_t)r->src.size();
    const double plain = (double)k * (double)std::max<int64_t>(g->nnz, 1);
    double red = 0.0;
    if (r->all) {
        for (int32_t b = 0; b < v.nblocks; ++b)
            red += (double)(v.boff[b + 1] - v.boff[b]) * (double)(v.prp[v.boff[b + 1]] - v.prp[v.boff[b]]);
    } else {
        std::vector<int64_t> ws((size_t)v.np);
        if (cent_reduced_source_weights(g->red, r->src.data(), (int32_t)k, ws.data()) != CENT_OK) return false;
        for (int32_t b = 0; b < v.nblocks; ++b) {
            const double eb = (double)(v.prp[v.boff[b + 1]] - v.prp[v.boff[b]]);
            for (int32_t x = v.boff[b]; x < v.boff[b + 1]; ++x)
                if (ws[(size_t)x] > 0) red += eb;
        }
    }
    return red < 0.5 * plain;
}

// An admission on `slot` waits for memory: idle slots holding cached
// buffers are asked to release them (the worker loop).  Called with s->mu held.
static void ask_idle_trim(cent_server* s, int slot) {
    for (size_t i = 0; i < s->held.size(); ++i)
        if (i != (size_t)slot && !s->busy[i] && s->held[i] > 0) s->trim_idle = true;
    if (s->trim_idle) s->cv_work.notify_all();
}

static void run_request(cent_server* s, int slot, Request* r, GraphEntry* g) {
    int st = CENT_OK;
    const bool dry = (s->o.flags & CENT_SERVER_DRYRUN) != 0;
    cent_ctx* ctx = dry ? nullptr : s->ctx[(size_t)slot];
    bool reduced_now = false;
    {
        std::lock_guard<std::mutex> lk(g->build_mu);
        if (!dry && !g->dev) st = cent_graph_upload(ctx, g->host, &g->dev);
        // the reduction runs on this slot's GPU; the dry run reduces on the
        // host.  All-vertex closeness always goes through it, betweenness
        // when it is cheaper there (below)
        if (!st && ((r->kind == CENT_Q_CC && r->all) || r->kind == CENT_Q_BC) && !g->red) {
            st = dry ? cent_reduce(g->host, &g->red) : cent_reduce_gpu(ctx, g->host, &g->red);
            reduced_now = !st;
        }
    }
    if (reduced_now && !dry) {
        // the reduction's scratch is released; the reservation becomes the
        // reduced call's own estimate, waiting (holding nothing) if it grew
        // beyond what is free
        (void)cent_ctx_trim(ctx);
        const int64_t want = workspace_estimate(g, r);
        std::unique_lock<std::mutex> lk(s->mu);
        if (want > s->held[(size_t)slot]) {
            s->used_ws -= s->held[(size_t)slot];
            s->held[(size_t)slot] = 0;
            if (want + graph_estimate(g) + reduced_estimate(g->red) > s->o.hbm_budget) {
                st = CENT_E_OOM;
            } else {
                while (!make_room(s, want, g)) {
                    ask_idle_trim(s, slot);
                    s->cv_done.wait(lk);
                }
                s->trim_idle = false;
                s->used_ws += want;
                s->held[(size_t)slot] = want;
                r->footprint = want;
            }
        }
    }
    const int64_t k = r->all ? g->n : (int64_t)r->src.size();
    const cent_opts opts = request_opts();
    if (!st) {
        if (dry) {
            // modelled duration: sources x directed edges at 1e9 edge visits/s
            const double us = (double)k * (double)std::max<int64_t>(g->nnz, 1) / 1e3;
            std::this_thread::sleep_for(std::chrono::microseconds((int64_t)std::min(us, 2e6)));
        } else if (r->kind == CENT_Q_BC) {
            r->bc.assign((size_t)g->n, 0.0);
            r->far.assign((size_t)std::max<int64_t>(k, 1), 0);
            st = CENT_E_UNSUPPORTED;
            if (g->red && reduced_bc_pays(g, r))  // (unsupported for a multigraph's reduction)
                st = r->all ? cent_bc_reduced(ctx, g->red, r->bc.data(), r->far.data(), CENT_SCHED_PERSISTENT, &opts)
                            : cent_bc_reduced_sources(ctx, g->red, r->src.data(), (int32_t)k, r->bc.data(),
                                                      r->far.data(), CENT_SCHED_PERSISTENT, &opts);
            if (st == CENT_E_UNSUPPORTED)
                st = cent_bc(ctx, g->dev, r->all ? nullptr : r->src.data(), r->all ? -1 : (int32_t)k, r->bc.data(),
                             r->far.data(), nullptr, &opts);
        } else if (r->all) {
            r->far.assign((size_t)std::max<int64_t>(g->n, 1), 0);
            st = cent_cc_reduced(ctx, g->red, r->far.data(), CENT_SCHED_PERSISTENT, &opts);
        } else {
            r->far.assign((size_t)std::max<int64_t>(k, 1), 0);
            st = cent_cc(ctx, g->dev, r->src.data(), (int32_t)k, r->far.data(), nullptr, &opts);
        }
    }
    // keep a slot's cached workspace within its share of the budget
    if (!dry && cent_ctx_workspace_bytes(ctx) > s->o.hbm_budget / std::max(1, s->o.slots) / 2) cent_ctx_trim(ctx);
    const int64_t keep = dry ? 0 : cent_ctx_workspace_bytes(ctx);
    std::lock_guard<std::mutex> lk(s->mu);
    s->used_ws += keep - s->held[(size_t)slot];  // the slot now holds its cached buffers
    s->held[(size_t)slot] = keep;
    s->busy[(size_t)slot] = 0;
    if (!dry && g->red && (r->kind == CENT_Q_BC || r->all)) {
        // the reduced device copy now exists; account for it once
        const int64_t rb = reduced_estimate(g->red);
        if (g->dev_bytes < graph_estimate(g) + rb) {
            s->used_graph += graph_estimate(g) + rb - g->dev_bytes;
            g->dev_bytes = graph_estimate(g) + rb;
        }
    }
    r->status = st;
    r->t_end = now_ms(s);
    r->state = r->cancel ? CENT_REQ_CANCELLED : (st ? CENT_REQ_FAILED : CENT_REQ_DONE);
    if (r->state == CENT_REQ_DONE) s->st.completed++;
    else if (r->state == CENT_REQ_FAILED) s->st.failed++;
    else s->st.cancelled++;
    g->refs--;
    g->last_use = ++s->tick;
    s->running--;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

static void worker(cent_server* s, int slot) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        Request* r = nullptr;
        GraphEntry* g = nullptr;
        while (!s->stop) {
            if (s->trim_idle && s->held[(size_t)slot] > 0 && s->ctx.size() > (size_t)slot) {
                // another slot's admission waits on cached buffers: release ours
                lk.unlock();
                (void)cent_ctx_trim(s->ctx[(size_t)slot]);
                lk.lock();
                s->used_ws -= s->held[(size_t)slot];
                s->held[(size_t)slot] = 0;
                s->cv_work.notify_all();
                s->cv_done.notify_all();
                continue;
            }
            if (!s->queue.empty()) {
                Request* h = s->queue.front();
                GraphEntry* hg = s->graphs[(size_t)h->gid];
                // this slot's cached buffers stay, so they count toward the reservation
                const int64_t ws = std::max(workspace_estimate(hg, h), s->held[(size_t)slot]);
                const int64_t gb = hg->dev_bytes > 0 ? 0 : graph_estimate(hg);
                if (workspace_estimate(hg, h) + graph_estimate(hg) > s->o.hbm_budget) {
                    // can never fit: fail it instead of blocking the queue
                    s->queue.pop_front();
                    h->state = CENT_REQ_FAILED;
                    h->status = CENT_E_OOM;
                    h->t_start = h->t_end = now_ms(s);
                    s->st.failed++;
                    s->cv_done.notify_all();
                    continue;
                }
                if (make_room(s, ws - s->held[(size_t)slot] + gb, hg)) {
                    s->queue.pop_front();
                    s->trim_idle = false;
                    r = h;
                    g = hg;
                    r->footprint = ws;
                    r->graph_bytes = gb;
                    s->used_ws += ws - s->held[(size_t)slot];
                    s->held[(size_t)slot] = ws;
                    s->busy[(size_t)slot] = 1;
                    if (gb) {
                        s->used_graph += gb;
                        g->dev_bytes = gb;
                        s->st.graph_uploads++;
                    } else {
                        s->st.graph_hits++;
                    }
                    g->refs++;
                    g->last_use = ++s->tick;
                    r->state = CENT_REQ_RUNNING;
                    r->slot = slot;
                    r->t_start = now_ms(s);
                    s->running++;
                    if (s->running > s->st.max_running) s->st.max_running = s->running;
                    break;
                }
                ask_idle_trim(s, slot);  // it waits
            }
            s->cv_work.wait(lk);
        }
        if (!r) return;  // stopping
        lk.unlock();
        run_request(s, slot, r, g);
        lk.lock();
    }
}

extern "C" int cent_server_create(const cent_server_opts* opts, cent_server** out) {
    if (!opts || !out || opts->slots < 1 || opts->slots > 64 || opts->hbm_budget <= 0 || opts->max_queue < 1)
        return CENT_E_INVAL;
    *out = nullptr;
    cent_server* s = new (std::nothrow) cent_server();
    if (!s) return CENT_E_OOM;
    s->o = *opts;
    s->held.assign((size_t)opts->slots, 0);
    s->busy.assign((size_t)opts->slots, 0);
    if (!(opts->flags & CENT_SERVER_DRYRUN)) {
        for (int i = 0; i < opts->slots; ++i) {
            cent_ctx* c = nullptr;
            int st = cent_ctx_create(opts->device, &c);
            if (st) {
                for (cent_ctx* x : s->ctx) cent_ctx_destroy(x);
                delete s;
                return st;
            }
            s->ctx.push_back(c);
        }
    }
    for (int i = 0; i < opts->slots; ++i) s->workers.emplace_back(worker, s, i);
    *out = s;
    return CENT_OK;
}

extern "C" void cent_server_destroy(cent_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->reqs) delete kv.second;
    for (GraphEntry* g : s->graphs) {
        if (g->dev) cent_graph_free(g->dev);
        if (g->red) cent_reduced_free(g->red);
        cent_csr_free(g->host);
        delete g;
    }
    for (cent_ctx* c : s->ctx) cent_ctx_destroy(c);
    delete s;
}

// BFS depth from the highest-degree vertex: the probe cent_graph_create runs
// at upload, which the workspace estimate takes before the graph is there.
static int32_t probe_depth(const struct CSR* g) {
    const int32_t n = g->csrVSize - 1;
    if (n <= 0) return 0;
    int32_t s0 = 0;
    for (int32_t v = 1; v < n; ++v)
        if (g->csrV[v + 1] - g->csrV[v] > g->csrV[s0 + 1] - g->csrV[s0]) s0 = v;
    std::vector<int32_t> dist((size_t)n, -1), q((size_t)n);
    size_t h = 0, t = 0;
    q[t++] = s0;
    dist[(size_t)s0] = 0;
    while (h < t) {
        const int32_t x = q[h++];
        for (int32_t j = g->csrV[x]; j < g->csrV[x + 1]; ++j)
            if (dist[(size_t)g->csrE[j]] < 0) {
                dist[(size_t)g->csrE[j]] = dist[(size_t)x] + 1;
                q[t++] = g->csrE[j];
            }
    }
    return dist[(size_t)q[t - 1]];
}

extern "C" int cent_server_add_graph(cent_server* s, const struct CSR* csr, int32_t* gid) {
    if (!s || !csr || !gid) return CENT_E_INVAL;
    struct CSR* copy = nullptr;
    int st = cent_csr_from_arrays(csr->csrVSize - 1, csr->csrV, csr->csrE, &copy);
    if (st) return st;
    copy->startNodeID = csr->startNodeID;
    copy->endNodeID = csr->endNodeID;
    copy->totalNodeNumber = csr->totalNodeNumber;
    copy->startAtZero = csr->startAtZero;
    GraphEntry* g = new (std::nothrow) GraphEntry();
    if (!g) { cent_csr_free(copy); return CENT_E_OOM; }
    g->host = copy;
    g->n = copy->csrVSize - 1;
    g->nnz = copy->csrESize;
    g->depth = probe_depth(copy);
    std::lock_guard<std::mutex> lk(s->mu);
    *gid = (int32_t)s->graphs.size();
    s->graphs.push_back(g);
    return CENT_OK;
}

extern "C" int cent_server_submit(cent_server* s, int32_t gid, int32_t kind, const int32_t* src, int32_t k,
                                  int64_t* rid) {
    if (!s || !rid || (kind != CENT_Q_BC && kind != CENT_Q_CC) || k < -1 || (k > 0 && !src)) return CENT_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (gid < 0 || (size_t)gid >= s->graphs.size()) return CENT_E_NOTFOUND;
    GraphEntry* g = s->graphs[(size_t)gid];
    for (int32_t i = 0; i < k; ++i)
        if (src[i] < 0 || src[i] >= g->n) return CENT_E_INVAL;
    if ((int)s->queue.size() >= s->o.max_queue) {
        s->st.rejected++;
        return CENT_E_BUSY;
    }
    Request* r = new (std::nothrow) Request();
    if (!r) return CENT_E_OOM;
    r->id = s->next_id++;
    r->gid = gid;
    r->kind = kind;
    r->all = (k == -1);
    if (k > 0) r->src.assign(src, src + k);
    r->t_submit = now_ms(s);
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->st.submitted++;
    if ((int64_t)s->queue.size() > s->st.max_queued) s->st.max_queued = (int64_t)s->queue.size();
    *rid = r->id;
    lk.unlock();
    s->cv_work.notify_one();
    return CENT_OK;
}

static void fill_info(const Request* r, cent_request_info* info) {
    info->state = r->state;
    info->status = r->status;
    info->slot = r->slot;
    info->t_submit_ms = r->t_submit;
    info->t_start_ms = r->t_start;
    info->t_end_ms = r->t_end;
    info->footprint = r->footprint;
}

extern "C" int cent_server_wait(cent_server* s, int64_t rid, int32_t timeout_ms, cent_request_info* info) {
    if (!s) return CENT_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(rid);
    if (it == s->reqs.end()) return CENT_E_NOTFOUND;
    Request* r = it->second;
    auto done = [&] { return r->state >= CENT_REQ_DONE; };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
    if (info) fill_info(r, info);
    return CENT_OK;
}

extern "C" int cent_server_result(cent_server* s, int64_t rid, double* bc, int64_t* farness,
                                  cent_request_info* info) {
    if (!s) return CENT_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(rid);
    if (it == s->reqs.end()) return CENT_E_NOTFOUND;
    Request* r = it->second;
    if (r->state < CENT_REQ_DONE) return CENT_E_BUSY;
    if (info) fill_info(r, info);
    const int st = (r->state == CENT_REQ_CANCELLED) ? CENT_E_CANCELLED : r->status;
    if (r->state == CENT_REQ_DONE) {
        if (bc && !r->bc.empty()) std::memcpy(bc, r->bc.data(), sizeof(double) * r->bc.size());
        const GraphEntry* g = s->graphs[(size_t)r->gid];
        const size_t nf = r->all ? (size_t)g->n : r->src.size();
        if (farness && !r->far.empty()) std::memcpy(farness, r->far.data(), sizeof(int64_t) * nf);
    }
    s->reqs.erase(it);
    delete r;
    return st;
}

extern "C" int cent_server_cancel(cent_server* s, int64_t rid) {
    if (!s) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(rid);
    if (it == s->reqs.end()) return CENT_E_NOTFOUND;
    Request* r = it->second;
    if (r->state == CENT_REQ_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
        r->state = CENT_REQ_CANCELLED;
        r->t_start = r->t_end = now_ms(s);
        s->st.cancelled++;
        s->cv_done.notify_all();
    } else if (r->state == CENT_REQ_RUNNING) {
        r->cancel = true;  // its kernels finish; the result is discarded
    }
    return CENT_OK;
}

extern "C" int cent_server_get_stats(cent_server* s, cent_server_stats* out) {
    if (!s || !out) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->queued = (int64_t)s->queue.size();
    out->running = s->running;
    out->used_workspace = s->used_ws;
    out->used_graph = s->used_graph;
    return CENT_OK;
}
