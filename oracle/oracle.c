/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference's algorithms for the hot path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (bc_amd/) links or calls this file.
 *
 *  oracle_bfs_farness   computeCC, Seq_CC/CC.cpp:37-117 (queue BFS per source,
 *                       farness accumulated in the source's slot, :94), here
 *                       in int64 and with unreachable vertices contributing 0.
 *  oracle_brandes       computeBC, Seq_Brandes/Brandes.c:224-372: queue BFS with
 *                       path counts (:274-298), then reverse-BFS-order
 *                       dependency accumulation (:313-327).  FP64.  Mode 0 is
 *                       textbook Brandes (BC[w] += delta[w] once per source);
 *                       mode 1 is the reference's own formula, which adds
 *                       delta[w] once per predecessor of w (:319-326).
 *  oracle_brandes_f32   mode-1 formula in FP32 exactly as the reference types
 *                       it (Brandes.c:226-227), to check the karate score file.
 *  oracle_weighted_brandes / oracle_weighted_brandes_src
 *                       the same traversal with vertex target weights and
 *                       source weights (per vertex, or per listed source),
 *                       the checker of the reduced (D1/AP) betweenness; with
 *                       unit weights they are oracle_brandes mode 0, which
 *                       the tests check.
 *
 * Pinning (tests/test_oracle.py): closed forms on paths, stars, cycles,
 * complete graphs and full grids; brute-force shortest-path enumeration on
 * tiny random graphs; networkx golden vectors (tests/golden/); the reference's
 * karate_BC_Score.txt (mode 1, FP32) and its closeness known answers
 * (karate CCs[7] = 86, musae_git CCs[7] = 117726); the identity
 * sum_v BC(v) = sum_s farness(s) - (pairs at distance >= 1) on connected graphs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t* dist;
    int32_t* queue;
    double* sigma;
    double* delta;
    int32_t* npred;
} ws_t;

static int ws_alloc(ws_t* w, int32_t n) {
    size_t N = (size_t)(n > 0 ? n : 1);
    w->dist = (int32_t*)malloc(N * sizeof(int32_t));
    w->queue = (int32_t*)malloc(N * sizeof(int32_t));
    w->sigma = (double*)malloc(N * sizeof(double));
    w->delta = (double*)malloc(N * sizeof(double));
    w->npred = (int32_t*)malloc(N * sizeof(int32_t));
    return w->dist && w->queue && w->sigma && w->delta && w->npred;
}
static void ws_free(ws_t* w) {
    free(w->dist); free(w->queue); free(w->sigma); free(w->delta); free(w->npred);
}

/* BFS farness: farness[i] = sum of d(src[i], v) over reached v (assigned);
 * dist_sum[v] += d(src[i], v) for reached v (may be NULL). */
int oracle_bfs_farness(int32_t n, const int32_t* rp, const int32_t* col,
                       const int32_t* src, int32_t k, int64_t* farness, int64_t* dist_sum) {
    ws_t w;
    if (!ws_alloc(&w, n)) { ws_free(&w); return 2; }
    for (int32_t v = 0; v < n; ++v) w.dist[v] = -1;
    for (int32_t i = 0; i < k; ++i) {
        int32_t s = src[i];
        int64_t h = 0, t = 0, far = 0;
        w.queue[t++] = s; w.dist[s] = 0;
        while (h < t) {
            int32_t x = w.queue[h++];
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (w.dist[y] < 0) { w.dist[y] = w.dist[x] + 1; w.queue[t++] = y; far += w.dist[y]; }
            }
        }
        if (farness) farness[i] = far;
        for (int64_t q = 0; q < t; ++q) {
            int32_t y = w.queue[q];
            if (dist_sum) dist_sum[y] += w.dist[y];
            w.dist[y] = -1;
        }
    }
    ws_free(&w);
    return 0;
}

/* One Brandes source into bc (accumulated).  Returns DAG edges. */
static int64_t brandes_one(int32_t n, const int32_t* rp, const int32_t* col, int32_t s,
                           int mode, ws_t* w, double* bc, int64_t* far_out, int64_t* dist_sum) {
    int64_t h = 0, t = 0, dag = 0, far = 0;
    w->queue[t++] = s; w->dist[s] = 0; w->sigma[s] = 1.0; w->npred[s] = 0;
    while (h < t) {
        int32_t x = w->queue[h++];
        int32_t dx = w->dist[x];
        for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
            int32_t y = col[j];
            if (w->dist[y] < 0) {
                w->dist[y] = dx + 1; w->sigma[y] = 0.0; w->npred[y] = 0;
                w->queue[t++] = y; far += dx + 1;
            }
            if (w->dist[y] == dx + 1) { w->sigma[y] += w->sigma[x]; w->npred[y]++; dag++; }
        }
    }
    for (int64_t q = 0; q < t; ++q) w->delta[w->queue[q]] = 0.0;
    /* reverse BFS order: every successor of x is finished before x */
    for (int64_t q = t - 1; q >= 0; --q) {
        int32_t x = w->queue[q];
        int32_t dx = w->dist[x];
        double acc = 0.0;
        for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
            int32_t y = col[j];
            if (w->dist[y] == dx + 1) acc += (1.0 + w->delta[y]) / w->sigma[y];
        }
        w->delta[x] = w->sigma[x] * acc;
        if (x != s) bc[x] += (mode == 1) ? w->delta[x] * (double)w->npred[x] : w->delta[x];
    }
    if (far_out) *far_out = far;
    for (int64_t q = 0; q < t; ++q) {
        int32_t y = w->queue[q];
        if (dist_sum) dist_sum[y] += w->dist[y];
        w->dist[y] = -1;
    }
    return dag;
}

/* Brandes BC from k sources, bc accumulated (+=).  threads <= 1: serial in
 * source order; threads > 1: OpenMP over sources with per-thread partial BC
 * vectors summed in thread order.  Returns total DAG edges via dag_out. */
int oracle_brandes(int32_t n, const int32_t* rp, const int32_t* col, const int32_t* src,
                   int32_t k, int mode, int threads, double* bc, int64_t* farness,
                   int64_t* dist_sum, int64_t* dag_out) {
    int64_t dag = 0;
    if (threads <= 1) {
        ws_t w;
        if (!ws_alloc(&w, n)) { ws_free(&w); return 2; }
        for (int32_t v = 0; v < n; ++v) w.dist[v] = -1;
        for (int32_t i = 0; i < k; ++i)
            dag += brandes_one(n, rp, col, src[i], mode, &w, bc, farness ? &farness[i] : NULL, dist_sum);
        ws_free(&w);
    } else {
#ifdef _OPENMP
        int nt = threads;
        double* part = (double*)calloc((size_t)nt * (size_t)(n > 0 ? n : 1), sizeof(double));
        int64_t* dpart = dist_sum ? (int64_t*)calloc((size_t)nt * (size_t)(n > 0 ? n : 1), sizeof(int64_t)) : NULL;
        if (!part || (dist_sum && !dpart)) { free(part); free(dpart); return 2; }
        int err = 0;
#pragma omp parallel num_threads(nt) reduction(+ : dag)
        {
            int tid = omp_get_thread_num();
            ws_t w;
            if (!ws_alloc(&w, n)) {
#pragma omp atomic write
                err = 1;
            } else {
                for (int32_t v = 0; v < n; ++v) w.dist[v] = -1;
#pragma omp for schedule(dynamic, 1)
                for (int32_t i = 0; i < k; ++i)
                    dag += brandes_one(n, rp, col, src[i], mode, &w, part + (size_t)tid * n,
                                       farness ? &farness[i] : NULL,
                                       dpart ? dpart + (size_t)tid * n : NULL);
            }
            ws_free(&w);
        }
        if (!err) {
            for (int t = 0; t < nt; ++t)
                for (int32_t v = 0; v < n; ++v) {
                    bc[v] += part[(size_t)t * n + v];
                    if (dist_sum) dist_sum[v] += dpart[(size_t)t * n + v];
                }
        }
        free(part); free(dpart);
        if (err) return 2;
#else
        return oracle_brandes(n, rp, col, src, k, mode, 1, bc, farness, dist_sum, dag_out);
#endif
    }
    if (dag_out) *dag_out = dag;
    return 0;
}

/* FP32 restatement of the reference's computeBC arithmetic (per-predecessor
 * accumulation, float sigma/delta/BC; Brandes.c:226-227, 291-295, 319-326).
 * The predecessor visiting order of the reference (pred vectors in discovery
 * order, stack pop order) is followed: predecessors of w are the earlier-
 * dequeued neighbours u with dist[u] + 1 == dist[w], in dequeue order. */
int oracle_brandes_f32(int32_t n, const int32_t* rp, const int32_t* col, const int32_t* src,
                       int32_t k, float* bc) {
    int32_t* dist = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    float* sigma = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float* delta = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int64_t nnz = rp[n];
    int32_t* pred = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    int32_t* npred = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!dist || !queue || !sigma || !delta || !pred || !npred) {
        free(dist); free(queue); free(sigma); free(delta); free(pred); free(npred);
        return 2;
    }
    for (int32_t v = 0; v < n; ++v) dist[v] = -1;
    for (int32_t i = 0; i < k; ++i) {
        int32_t s = src[i];
        int64_t h = 0, t = 0;
        queue[t++] = s; dist[s] = 0; sigma[s] = 1.0f; npred[s] = 0; delta[s] = 0.0f;
        while (h < t) {
            int32_t x = queue[h++];
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (dist[y] < 0) {
                    dist[y] = dist[x] + 1; sigma[y] = 0.0f; npred[y] = 0; delta[y] = 0.0f;
                    queue[t++] = y;
                }
                if (dist[y] == dist[x] + 1) {
                    sigma[y] += sigma[x];
                    pred[rp[y] + npred[y]++] = x;
                }
            }
        }
        for (int64_t q = t - 1; q >= 0; --q) {
            int32_t w = queue[q];
            for (int32_t p = 0; p < npred[w]; ++p) {
                int32_t v = pred[rp[w] + p];
                delta[v] += (sigma[v] / sigma[w]) * (1.0f + delta[w]);
                if (w != s) bc[w] += delta[w];
            }
        }
        for (int64_t q = 0; q < t; ++q) dist[queue[q]] = -1;
    }
    free(dist); free(queue); free(sigma); free(delta); free(pred); free(npred);
    return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Weighted BFS distance sums: far[i] = sum over reached v of d(src[i], v) * w[v].
 * Checker for the packed block graph of the reduced closeness path. */
int oracle_weighted_farness(int32_t n, const int32_t* rp, const int32_t* col, const int64_t* w,
                            const int32_t* src, int32_t k, int64_t* far) {
    int err = 0;
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        int32_t* dist = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        if (!dist || !queue) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = 1;
        } else {
            for (int32_t v = 0; v < n; ++v) dist[v] = -1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
            for (int32_t i = 0; i < k; ++i) {
                int64_t h = 0, t = 0, f = 0;
                queue[t++] = src[i]; dist[src[i]] = 0;
                while (h < t) {
                    int32_t x = queue[h++];
                    for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                        int32_t y = col[j];
                        if (dist[y] < 0) { dist[y] = dist[x] + 1; queue[t++] = y; f += (int64_t)dist[y] * w[y]; }
                    }
                }
                far[i] = f;
                for (int64_t q = 0; q < t; ++q) dist[queue[q]] = -1;
            }
        }
        free(dist); free(queue);
    }
    return err ? 2 : 0;
}

/* Weighted Brandes from listed sources src[i] with their own source weights
 * sw[i] and target weights w[v] (checker for sampled betweenness through the
 * reduction): bc[v] += sw[i] * delta_s(v), delta as below; far[i] (may be
 * NULL) = sum over reached v of d(s, v) w[v].  Serial, bc accumulated. */
int oracle_weighted_brandes_src(int32_t n, const int32_t* rp, const int32_t* col, const int64_t* w,
                                const int32_t* src, const int64_t* sw, int32_t k, double* bc, int64_t* far) {
    ws_t ws;
    if (!ws_alloc(&ws, n)) { ws_free(&ws); return 2; }
    for (int32_t v = 0; v < n; ++v) ws.dist[v] = -1;
    for (int32_t i = 0; i < k; ++i) {
        const int32_t s = src[i];
        int64_t h = 0, t = 0, f = 0;
        ws.queue[t++] = s; ws.dist[s] = 0; ws.sigma[s] = 1.0;
        while (h < t) {
            int32_t x = ws.queue[h++];
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (ws.dist[y] < 0) {
                    ws.dist[y] = ws.dist[x] + 1; ws.sigma[y] = 0.0; ws.queue[t++] = y;
                    f += (int64_t)ws.dist[y] * w[y];
                }
                if (ws.dist[y] == ws.dist[x] + 1) ws.sigma[y] += ws.sigma[x];
            }
        }
        if (far) far[i] = f;
        for (int64_t q = t - 1; q >= 0; --q) {
            int32_t x = ws.queue[q];
            double acc = 0.0;
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (ws.dist[y] == ws.dist[x] + 1) acc += ((double)w[y] + ws.delta[y]) / ws.sigma[y];
            }
            ws.delta[x] = ws.sigma[x] * acc;
            if (x != s) bc[x] += (double)sw[i] * ws.delta[x];
        }
        for (int64_t q = 0; q < t; ++q) ws.dist[ws.queue[q]] = -1;
    }
    ws_free(&ws);
    return 0;
}

/* Weighted Brandes (checker for the reduced BC path): vertex weights w act as
 * multiplicities of sources and targets:
 *   bc[v] += w[s] * delta_s(v),  delta_s(v) = sum_{u: v in pred_s(u)} sigma_v/sigma_u (w[u] + delta_s(u)).
 * With w == 1 this is oracle_brandes mode 0.  Serial, bc accumulated. */
int oracle_weighted_brandes(int32_t n, const int32_t* rp, const int32_t* col, const int64_t* w,
                            const int32_t* src, int32_t k, double* bc, int64_t* far) {
    ws_t ws;
    if (!ws_alloc(&ws, n)) { ws_free(&ws); return 2; }
    for (int32_t v = 0; v < n; ++v) ws.dist[v] = -1;
    for (int32_t i = 0; i < k; ++i) {
        const int32_t s = src[i];
        int64_t h = 0, t = 0, f = 0;
        ws.queue[t++] = s; ws.dist[s] = 0; ws.sigma[s] = 1.0;
        while (h < t) {
            int32_t x = ws.queue[h++];
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (ws.dist[y] < 0) {
                    ws.dist[y] = ws.dist[x] + 1; ws.sigma[y] = 0.0; ws.queue[t++] = y;
                    f += (int64_t)ws.dist[y] * w[y];
                }
                if (ws.dist[y] == ws.dist[x] + 1) ws.sigma[y] += ws.sigma[x];
            }
        }
        if (far) far[i] = f;
        for (int64_t q = t - 1; q >= 0; --q) {
            int32_t x = ws.queue[q];
            double acc = 0.0;
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (ws.dist[y] == ws.dist[x] + 1) acc += ((double)w[y] + ws.delta[y]) / ws.sigma[y];
            }
            ws.delta[x] = ws.sigma[x] * acc;
            if (x != s) bc[x] += (double)w[s] * ws.delta[x];
        }
        for (int64_t q = 0; q < t; ++q) ws.dist[ws.queue[q]] = -1;
    }
    ws_free(&ws);
    return 0;
}
