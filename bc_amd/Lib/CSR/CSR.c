/*
 * Own CSR construction (replaces createCSR, Lib/CSR/CSR.c:9-108, and the
 * adjacency build of buildGraph, Lib/AdjList/AdjList.c:76-92).
 *
 * Differences from the reference, on purpose:
 *  - self-loops and duplicate edges are removed (the reference keeps them),
 *    unless CENT_CSR_RAW asks for the reference's multigraph as it is;
 *  - rows are sorted by neighbour id, not by neighbour degree (AdjList.c:105-131);
 *    this changes traversal order only, never a BFS distance, path count or sum;
 *  - no spare capacity for articulation-point clones: the reduction code
 *    builds separate block CSRs instead of mutating this one.
 */
#include <stdio.h>
#include "../common.h"
#include "../reduce.h"

const char* cent_strerror(int s) {
    switch (s) {
    case CENT_OK: return "ok";
    case CENT_E_INVAL: return "invalid argument";
    case CENT_E_OOM: return "out of memory";
    case CENT_E_DEVICE: return "HIP device error";
    case CENT_E_OVERFLOW: return "shortest-path count overflowed FP64";
    case CENT_E_IO: return "I/O error";
    case CENT_E_BUSY: return "queue full";
    case CENT_E_CANCELLED: return "cancelled";
    case CENT_E_NOTFOUND: return "not found";
    case CENT_E_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

void cent_free(void* p) { free(p); }

int cent_ivec_push(cent_ivec* v, int32_t x) {
    if (v->n == v->cap) {
        int64_t nc = v->cap ? v->cap * 2 : 64;
        int32_t* na = (int32_t*)realloc(v->a, (size_t)nc * sizeof(int32_t));
        if (!na) return CENT_E_OOM;
        v->a = na; v->cap = nc;
    }
    v->a[v->n++] = x;
    return CENT_OK;
}
void cent_ivec_free(cent_ivec* v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }

/* Radix sort of keys (ascending); returns n, or -1 when out of memory. */
int64_t cent_sort_u64(uint64_t* keys, int64_t n) {
    if (n <= 1) return n;
    uint64_t* tmp = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    if (!tmp) return -1;
    uint64_t* src = keys; uint64_t* dst = tmp;
    int64_t cnt[256];
    for (int pass = 0; pass < 8; ++pass) {
        int sh = pass * 8;
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[(src[i] >> sh) & 255]++;
        int constant = 0;
        for (int b = 0; b < 256; ++b) if (cnt[b] == n) { constant = 1; break; }
        if (constant) continue;
        int64_t off = 0;
        for (int b = 0; b < 256; ++b) { int64_t c = cnt[b]; cnt[b] = off; off += c; }
        for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> sh) & 255]++] = src[i];
        uint64_t* t = src; src = dst; dst = t;
    }
    if (src != keys) memcpy(keys, src, (size_t)n * sizeof(uint64_t));
    free(tmp);
    return n;
}

/* Sorts keys and drops repeats; returns the number kept, or -1 (OOM). */
int64_t cent_sort_unique_u64(uint64_t* keys, int64_t n) {
    if (cent_sort_u64(keys, n) < 0) return -1;
    if (n <= 1) return n;
    int64_t w = 1;
    for (int64_t i = 1; i < n; ++i) if (keys[i] != keys[w - 1]) keys[w++] = keys[i];
    return w;
}

struct CSR* cent_csr_alloc(int32_t n, int64_t nnz) {
    struct CSR* g = (struct CSR*)calloc(1, sizeof(struct CSR));
    if (!g) return NULL;
    g->csrV = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    g->csrE = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    g->csrNodesDegree = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!g->csrV || !g->csrE || !g->csrNodesDegree) { cent_csr_free(g); return NULL; }
    g->csrVSize = n + 1;
    g->csrESize = nnz;
    g->startNodeID = 0;
    g->endNodeID = n - 1;
    g->totalNodeNumber = n;
    g->startAtZero = 1;
    return g;
}

void cent_csr_free(struct CSR* g) {
    if (!g) return;
    free(g->csrV); free(g->csrE); free(g->csrNodesDegree); free(g);
}

static void csr_finish(struct CSR* g) {
    int32_t n = g->csrVSize - 1, mx = 0;
    for (int32_t v = 0; v < n; ++v) {
        int32_t d = g->csrV[v + 1] - g->csrV[v];
        g->csrNodesDegree[v] = d;
        if (d > mx) mx = d;
    }
    g->maxDegree = mx;
}

int cent_csr_from_sorted_keys(int32_t n, const uint64_t* keys, int64_t nk, struct CSR** out) {
    if (nk > INT32_MAX) return CENT_E_INVAL;
    struct CSR* g = cent_csr_alloc(n, nk);
    if (!g) return CENT_E_OOM;
    memset(g->csrV, 0, sizeof(int32_t) * ((size_t)n + 1));
    int multi = 0;  /* sorted keys: a repeat is adjacent, a self-loop has equal halves */
    for (int64_t i = 0; i < nk; ++i) {
        g->csrV[(keys[i] >> 32) + 1]++;
        if ((i > 0 && keys[i] == keys[i - 1]) || (keys[i] >> 32) == (keys[i] & 0xffffffffu)) multi = 1;
    }
    for (int32_t v = 0; v < n; ++v) g->csrV[v + 1] += g->csrV[v];
    for (int64_t i = 0; i < nk; ++i) g->csrE[i] = (int32_t)(keys[i] & 0xffffffffu);
    csr_finish(g);
    g->kind = multi ? CENT_CSR_KIND_MULTI : CENT_CSR_KIND_SIMPLE;
    *out = g;
    return CENT_OK;
}

int32_t cent_components(const struct CSR* g, int32_t* comp, int64_t* comp_size) {
    int32_t n = g->csrVSize - 1;
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!q) return -1;
    for (int32_t v = 0; v < n; ++v) comp[v] = -1;
    int32_t nc = 0;
    for (int32_t s = 0; s < n; ++s) {
        if (comp[s] >= 0) continue;
        int64_t h = 0, t = 0;
        q[t++] = s; comp[s] = nc;
        while (h < t) {
            int32_t x = q[h++];
            for (int32_t j = g->csrV[x]; j < g->csrV[x + 1]; ++j) {
                int32_t y = g->csrE[j];
                if (comp[y] < 0) { comp[y] = nc; q[t++] = y; }
            }
        }
        if (comp_size) comp_size[nc] = t;
        nc++;
    }
    free(q);
    return nc;
}

int cent_csr_induced(const struct CSR* g, const uint8_t* keep, struct CSR** out, int32_t** old_id) {
    int32_t n = g->csrVSize - 1;
    int32_t* nid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!nid) return CENT_E_OOM;
    int32_t n2 = 0;
    int64_t nnz2 = 0;
    for (int32_t v = 0; v < n; ++v) {
        if (keep[v]) {
            nid[v] = n2++;
            for (int32_t j = g->csrV[v]; j < g->csrV[v + 1]; ++j) nnz2 += keep[g->csrE[j]] ? 1 : 0;
        } else nid[v] = -1;
    }
    struct CSR* h = cent_csr_alloc(n2, nnz2);
    if (!h) { free(nid); return CENT_E_OOM; }
    int64_t e = 0;
    h->csrV[0] = 0;
    for (int32_t v = 0; v < n; ++v) {
        if (!keep[v]) continue;
        for (int32_t j = g->csrV[v]; j < g->csrV[v + 1]; ++j) {
            int32_t y = g->csrE[j];
            if (keep[y]) h->csrE[e++] = nid[y];
        }
        h->csrV[nid[v] + 1] = (int32_t)e;
    }
    csr_finish(h);
    h->kind = g->kind == CENT_CSR_KIND_SIMPLE ? CENT_CSR_KIND_SIMPLE : CENT_CSR_KIND_UNKNOWN;
    if (old_id) {
        int32_t* o = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n2 > 0 ? n2 : 1));
        if (!o) { free(nid); cent_csr_free(h); return CENT_E_OOM; }
        for (int32_t v = 0; v < n; ++v) if (nid[v] >= 0) o[nid[v]] = v;
        *old_id = o;
    }
    free(nid);
    *out = h;
    return CENT_OK;
}

static int keep_lcc(struct CSR* g, struct CSR** out, int32_t** old_id) {
    int32_t n = g->csrVSize - 1;
    int32_t* comp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t* csz = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    uint8_t* keep = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));
    if (!comp || !csz || !keep) { free(comp); free(csz); free(keep); return CENT_E_OOM; }
    int32_t nc = cent_components(g, comp, csz);
    int32_t best = 0;
    for (int32_t c = 1; c < nc; ++c) if (csz[c] > csz[best]) best = c;
    for (int32_t v = 0; v < n; ++v) keep[v] = (uint8_t)(comp[v] == best);
    int st = cent_csr_induced(g, keep, out, old_id);
    free(comp); free(csz); free(keep);
    return st;
}

int cent_csr_from_edges(int32_t n, const int32_t* u, const int32_t* v, int64_t m,
                        int32_t flags, struct CSR** out, int32_t** old_id) {
    if (!out || n < 0 || m < 0 || (m > 0 && (!u || !v))) return CENT_E_INVAL;
    if ((flags & CENT_CSR_RAW) && (flags & CENT_CSR_LCC)) return CENT_E_INVAL;
    if (old_id) *old_id = NULL;
    const int raw = (flags & CENT_CSR_RAW) != 0;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(2 * m + 1));
    if (!keys) return CENT_E_OOM;
    int64_t nk = 0;
    for (int64_t i = 0; i < m; ++i) {
        int32_t a = u[i], b = v[i];
        if (a < 0 || b < 0 || a >= n || b >= n) { free(keys); return CENT_E_INVAL; }
        if (a == b && !raw) continue;
        /* raw: every line gives both directions, so a self-loop lists its
         * vertex twice in its own row, as the reference's adjacency build
         * (Lib/AdjList/AdjList.c:76-92) does */
        keys[nk++] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
        keys[nk++] = ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
    }
    nk = raw ? cent_sort_u64(keys, nk) : cent_sort_unique_u64(keys, nk);
    if (nk < 0) { free(keys); return CENT_E_OOM; }
    struct CSR* g = NULL;
    int st = cent_csr_from_sorted_keys(n, keys, nk, &g);
    free(keys);
    if (st) return st;
    if (flags & CENT_CSR_LCC) {
        struct CSR* h = NULL;
        st = keep_lcc(g, &h, old_id);
        cent_csr_free(g);
        if (st) return st;
        g = h;
    }
    *out = g;
    return CENT_OK;
}

int cent_csr_from_arrays(int32_t n, const int32_t* row_ptr, const int32_t* col, struct CSR** out) {
    if (!out || n < 0 || !row_ptr || row_ptr[0] != 0) return CENT_E_INVAL;
    for (int32_t i = 0; i < n; ++i) if (row_ptr[i + 1] < row_ptr[i]) return CENT_E_INVAL;
    int64_t nnz = row_ptr[n];
    if (nnz > 0 && !col) return CENT_E_INVAL;
    for (int64_t j = 0; j < nnz; ++j) if (col[j] < 0 || col[j] >= n) return CENT_E_INVAL;
    struct CSR* g = cent_csr_alloc(n, nnz);
    if (!g) return CENT_E_OOM;
    memcpy(g->csrV, row_ptr, sizeof(int32_t) * ((size_t)n + 1));
    if (nnz) memcpy(g->csrE, col, sizeof(int32_t) * (size_t)nnz);
    csr_finish(g);
    g->kind = CENT_CSR_KIND_UNKNOWN;
    const int multi = cent_csr_is_multigraph(g);
    if (multi >= 0) g->kind = multi ? CENT_CSR_KIND_MULTI : CENT_CSR_KIND_SIMPLE;
    *out = g;
    return CENT_OK;
}

/* 1 if some row lists a vertex twice or lists its own vertex (the
 * constructors record it in g->kind; a caller-built CSR is scanned). */
int cent_csr_is_multigraph(const struct CSR* g) {
    if (g->kind == CENT_CSR_KIND_SIMPLE) return 0;
    if (g->kind == CENT_CSR_KIND_MULTI) return 1;
    const int32_t n = g->csrVSize - 1;
    int32_t* stamp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!stamp) return -1;
    for (int32_t v = 0; v < n; ++v) stamp[v] = -1;
    int multi = 0;
    for (int32_t v = 0; v < n && !multi; ++v)
        for (int32_t j = g->csrV[v]; j < g->csrV[v + 1]; ++j) {
            const int32_t y = g->csrE[j];
            if (y == v || stamp[y] == v) { multi = 1; break; }
            stamp[y] = v;
        }
    free(stamp);
    return multi;
}

/* The simple graph under g: self-loops dropped, each neighbour once per row
 * (rows sorted by id). */
int cent_csr_simple(const struct CSR* g, struct CSR** out) {
    const int32_t n = g->csrVSize - 1;
    const int64_t nnz = g->csrV[n];
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nnz > 0 ? nnz : 1));
    if (!keys) return CENT_E_OOM;
    int64_t nk = 0;
    for (int32_t v = 0; v < n; ++v)
        for (int32_t j = g->csrV[v]; j < g->csrV[v + 1]; ++j)
            if (g->csrE[j] != v) keys[nk++] = ((uint64_t)(uint32_t)v << 32) | (uint32_t)g->csrE[j];
    nk = cent_sort_unique_u64(keys, nk);
    if (nk < 0) { free(keys); return CENT_E_OOM; }
    struct CSR* h = NULL;
    int st = cent_csr_from_sorted_keys(n, keys, nk, &h);
    free(keys);
    if (st) return st;
    h->startNodeID = g->startNodeID;
    h->endNodeID = g->endNodeID;
    h->totalNodeNumber = g->totalNodeNumber;
    h->startAtZero = g->startAtZero;
    *out = h;
    return CENT_OK;
}
