/*
 * Seeded synthetic graph generators for the benchmark workloads.  The
 * reference has none (it only loads files, Lib/AdjList/AdjList.c:13); these
 * produce the R-MAT, grid and block-tree graphs the measurements are quoted
 * on, always cleaned (no self-loops, no duplicates) and symmetrised.
 */
#include <math.h>
#include <stdio.h>
#include "../common.h"

static int shuffle_perm(int32_t* p, int32_t n, cent_rng* r) {
    for (int32_t i = 0; i < n; ++i) p[i] = i;
    for (int32_t i = n - 1; i > 0; --i) {
        int32_t j = (int32_t)cent_rng_below(r, (uint64_t)i + 1);
        int32_t t = p[i]; p[i] = p[j]; p[j] = t;
    }
    return CENT_OK;
}

int cent_gen_rmat(int32_t scale, int64_t n_samples, double a, double b, double c,
                  uint64_t seed, int32_t flags, struct CSR** out) {
    if (!out || scale < 1 || scale > 30 || n_samples < 0 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0)
        return CENT_E_INVAL;
    int32_t n = (int32_t)1 << scale;
    cent_rng r;
    cent_rng_seed(&r, seed);
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* u = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_samples > 0 ? n_samples : 1));
    int32_t* v = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_samples > 0 ? n_samples : 1));
    if (!perm || !u || !v) { free(perm); free(u); free(v); return CENT_E_OOM; }
    const double ab = a + b, abc = a + b + c;
    for (int64_t i = 0; i < n_samples; ++i) {
        uint32_t x = 0, y = 0;
        for (int32_t l = 0; l < scale; ++l) {
            double q = cent_rng_uniform(&r);
            uint32_t bx = 0, by = 0;
            if (q < a) { }
            else if (q < ab) by = 1;
            else if (q < abc) bx = 1;
            else { bx = 1; by = 1; }
            x = (x << 1) | bx;
            y = (y << 1) | by;
        }
        u[i] = (int32_t)x; v[i] = (int32_t)y;
    }
    shuffle_perm(perm, n, &r);
    for (int64_t i = 0; i < n_samples; ++i) { u[i] = perm[u[i]]; v[i] = perm[v[i]]; }
    free(perm);
    int st = cent_csr_from_edges(n, u, v, n_samples, flags, out, NULL);
    free(u); free(v);
    return st;
}

int cent_gen_grid(int32_t rows, int32_t cols, double del_frac, int32_t n_shortcuts,
                  uint64_t seed, struct CSR** out) {
    if (!out || rows < 1 || cols < 1 || (int64_t)rows * cols > INT32_MAX / 2 ||
        del_frac < 0 || del_frac > 1 || n_shortcuts < 0)
        return CENT_E_INVAL;
    int32_t n = rows * cols;
    int64_t E = (int64_t)rows * (cols - 1) + (int64_t)(rows - 1) * cols;
    cent_rng r;
    cent_rng_seed(&r, seed);
    int32_t* eu = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    int32_t* ev = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    if (!eu || !ev) { free(eu); free(ev); return CENT_E_OOM; }
    int64_t e = 0;
    for (int32_t i = 0; i < rows; ++i)
        for (int32_t j = 0; j < cols; ++j) {
            int32_t id = i * cols + j;
            if (j + 1 < cols) { eu[e] = id; ev[e] = id + 1; ++e; }
            if (i + 1 < rows) { eu[e] = id; ev[e] = id + cols; ++e; }
        }
    /* remove ndel edges chosen uniformly: partial Fisher-Yates moves them to the tail */
    int64_t ndel = (int64_t)floor(del_frac * (double)E + 0.5);
    for (int64_t t = 0; t < ndel; ++t) {
        int64_t last = E - 1 - t;
        int64_t j = (int64_t)cent_rng_below(&r, (uint64_t)last + 1);
        int32_t tu = eu[j], tv = ev[j];
        eu[j] = eu[last]; ev[j] = ev[last];
        eu[last] = tu; ev[last] = tv;
    }
    int64_t keepE = E - ndel;
    struct CSR* g = NULL;
    int32_t* old = NULL;
    int st = cent_csr_from_edges(n, eu, ev, keepE, CENT_CSR_LCC, &g, &old);
    free(eu); free(ev);
    if (st) return st;
    free(old);
    if (n_shortcuts == 0) { *out = g; return CENT_OK; }
    /* add shortcuts between uniformly drawn distinct vertices of the component */
    int32_t n2 = g->csrVSize - 1;
    int64_t m2 = g->csrESize / 2 + n_shortcuts;
    int32_t* su = (int32_t*)malloc(sizeof(int32_t) * (size_t)m2);
    int32_t* sv = (int32_t*)malloc(sizeof(int32_t) * (size_t)m2);
    if (!su || !sv) { free(su); free(sv); cent_csr_free(g); return CENT_E_OOM; }
    int64_t k = 0;
    for (int32_t x = 0; x < n2; ++x)
        for (int32_t j = g->csrV[x]; j < g->csrV[x + 1]; ++j)
            if (g->csrE[j] > x) { su[k] = x; sv[k] = g->csrE[j]; ++k; }
    for (int32_t t = 0; t < n_shortcuts; ++t) {
        int32_t x = (int32_t)cent_rng_below(&r, (uint64_t)n2), y;
        do { y = (int32_t)cent_rng_below(&r, (uint64_t)n2); } while (y == x && n2 > 1);
        su[k] = x; sv[k] = y; ++k;
    }
    cent_csr_free(g);
    st = cent_csr_from_edges(n2, su, sv, k, 0, out, NULL);
    free(su); free(sv);
    return st;
}

int cent_gen_blocktree(int32_t n_blocks, double alpha, int32_t smin, int32_t smax,
                       double avg_deg, double pendant_frac, double p_geo,
                       uint64_t seed, struct CSR** out) {
    if (!out || n_blocks < 1 || alpha <= 0 || smin < 2 || smax < smin || avg_deg < 0 ||
        pendant_frac < 0 || p_geo <= 0 || p_geo > 1)
        return CENT_E_INVAL;
    cent_rng r;
    cent_rng_seed(&r, seed);
    int32_t* size = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_blocks);
    int64_t* first = (int64_t*)malloc(sizeof(int64_t) * (size_t)n_blocks);
    int32_t* attach = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_blocks);
    if (!size || !first || !attach) { free(size); free(first); free(attach); return CENT_E_OOM; }
    int64_t nv = 0, ne = 0;
    for (int32_t i = 0; i < n_blocks; ++i) {
        double U = 1.0 - cent_rng_uniform(&r); /* (0, 1] */
        double s = floor((double)smin * pow(U, -1.0 / alpha));
        if (s > smax) s = smax;
        if (s < smin) s = smin;
        size[i] = (int32_t)s;
        first[i] = nv;
        nv += (i == 0) ? size[i] : size[i] - 1;
        ne += size[i] + (int64_t)floor(avg_deg * size[i] / 2.0 + 0.5);
    }
    int64_t block_vertices = nv;
    int64_t extra = (int64_t)floor(pendant_frac * (double)block_vertices + 0.5);
    if (nv + extra > INT32_MAX) { free(size); free(first); free(attach); return CENT_E_INVAL; }
    ne += extra;
    int32_t* eu = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
    int32_t* ev = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
    int32_t* loc = NULL;
    int32_t maxs = 0;
    for (int32_t i = 0; i < n_blocks; ++i) if (size[i] > maxs) maxs = size[i];
    loc = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxs);
    if (!eu || !ev || !loc) {
        free(size); free(first); free(attach); free(eu); free(ev); free(loc);
        return CENT_E_OOM;
    }
    int64_t k = 0;
    for (int32_t i = 0; i < n_blocks; ++i) {
        int32_t s = size[i];
        /* global id of local vertex l */
        if (i > 0) {
            int32_t j = (int32_t)cent_rng_below(&r, (uint64_t)i);
            int32_t l = (int32_t)cent_rng_below(&r, (uint64_t)size[j]);
            attach[i] = (j > 0 && l == 0) ? attach[j]
                        : (int32_t)(first[j] + (j > 0 ? l - 1 : l));
        } else attach[i] = -1;
#define GID(l) ((i > 0 && (l) == 0) ? attach[i] : (int32_t)(first[i] + ((i > 0) ? (l) - 1 : (l))))
        /* Hamiltonian cycle over a random order of the block's vertices */
        shuffle_perm(loc, s, &r);
        if (s >= 3)
            for (int32_t t = 0; t < s; ++t) { eu[k] = GID(loc[t]); ev[k] = GID(loc[(t + 1) % s]); ++k; }
        else if (s == 2) { eu[k] = GID(0); ev[k] = GID(1); ++k; }
        /* ER part: round(avg_deg * s / 2) uniform pairs */
        int64_t M = (int64_t)floor(avg_deg * s / 2.0 + 0.5);
        for (int64_t t = 0; t < M && s >= 2; ++t) {
            int32_t x = (int32_t)cent_rng_below(&r, (uint64_t)s), y;
            do { y = (int32_t)cent_rng_below(&r, (uint64_t)s); } while (y == x);
            eu[k] = GID(x); ev[k] = GID(y); ++k;
        }
#undef GID
    }
    /* pendant chains attached to uniformly drawn block vertices */
    int64_t left = extra;
    const double lq = log(1.0 - p_geo);
    while (left > 0) {
        int64_t L = 1;
        if (p_geo < 1.0) {
            double U = 1.0 - cent_rng_uniform(&r);
            L = 1 + (int64_t)floor(log(U) / lq);
        }
        if (L > left) L = left;
        int32_t prev = (int32_t)cent_rng_below(&r, (uint64_t)block_vertices);
        for (int64_t t = 0; t < L; ++t) {
            int32_t x = (int32_t)nv++;
            eu[k] = prev; ev[k] = x; ++k;
            prev = x;
        }
        left -= L;
    }
    int st = cent_csr_from_edges((int32_t)nv, eu, ev, k, 0, out, NULL);
    free(size); free(first); free(attach); free(eu); free(ev); free(loc);
    return st;
}

int cent_sample_sources(int32_t n, int32_t k, uint64_t seed, int32_t* out) {
    if (n < 0 || k < 0 || k > n || (k > 0 && !out)) return CENT_E_INVAL;
    int32_t* p = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!p) return CENT_E_OOM;
    cent_rng r;
    cent_rng_seed(&r, seed);
    for (int32_t i = 0; i < n; ++i) p[i] = i;
    for (int32_t i = 0; i < k; ++i) {
        int32_t j = i + (int32_t)cent_rng_below(&r, (uint64_t)(n - i));
        int32_t t = p[i]; p[i] = p[j]; p[j] = t;
        out[i] = p[i];
    }
    free(p);
    return CENT_OK;
}
