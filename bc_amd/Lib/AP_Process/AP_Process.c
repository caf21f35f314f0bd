/*
 * Articulation-point split of the folded graph — the role of AP_detection,
 * AP_Copy_And_Split and rebuildGraph (Lib/AP_Process/AP_Process.c:6-225,
 * 468-1065, 1141-1321), by a different, linear-time route:
 *
 *  1. Hopcroft–Tarjan biconnected components (iterative DFS with an edge
 *     stack) over the unfolded vertices.  Blocks come out in DFS post-order,
 *     each headed by the vertex it hangs from (its parent articulation
 *     vertex; the head of a component's last block is the DFS root).
 *  2. Block-cut tree weights.  With the tree rooted at each component's last
 *     block, a copy x of vertex v in block B gets
 *         W(x) = w[v] + sum of sub_w(C) over child blocks C hanging at v
 *     and the copy of B's parent articulation vertex gets the rest of the
 *     component, Wcomp - sub_w(B).  The reference's per-part "represent"
 *     values (getPartsInfo, AP_Process.c:383-462) play this role.
 *  3. Packing: each block gets a private, contiguous id range with its own
 *     copy of every articulation vertex, so the packed graph is a disjoint
 *     union of blocks.
 *  4. Work items: consecutive small blocks are grouped into items of at most
 *     64 sources; a larger block gives ceil(size / 64) items over its range.
 *
 * cent_reduced_finish then turns the GPU's weighted distance sums into
 * farness with the block constants K_B (see reduce.h), computed bottom-up and
 * top-down over the block-cut tree — the quantity the reference obtains per
 * articulation point as CC[AP] = total_ff + ff[AP] (AP_Process.c:741) and
 * through the clones' ff values.
 */
#include <stdio.h>
#include "../common.h"
#include "../reduce.h"

void cent_reduced_free(struct cent_reduced* r) {
    if (!r) return;
    if (r->dev && r->dev_free) r->dev_free(r->dev);
    cent_d1_free(&r->d1);
    free(r->prp); free(r->pcol); free(r->pw); free(r->porig); free(r->pblock);
    free(r->boff); free(r->bhead); free(r->bparent); free(r->rep); free(r->vcomp);
    free(r->comp_w); free(r->items);
    free(r->fin_x); free(r->fin_b); free(r->fin_base); free(r->fin_ffsum);
    free(r);
}

int cent_reduced_tables(struct cent_reduced* r) {
    const int32_t n = r->n;
    const size_t N = (size_t)(n > 0 ? n : 1);
    free(r->fin_x); free(r->fin_b); free(r->fin_base); free(r->fin_ffsum);
    r->fin_x = (int32_t*)malloc(N * sizeof(int32_t));
    r->fin_b = (int32_t*)malloc(N * sizeof(int32_t));
    r->fin_base = (int64_t*)malloc(N * sizeof(int64_t));
    r->fin_ffsum = (int64_t*)calloc((size_t)r->nblocks + 1, sizeof(int64_t));
    if (!r->fin_x || !r->fin_b || !r->fin_base || !r->fin_ffsum) return CENT_E_OOM;
    for (int32_t b = 0; b < r->nblocks; ++b)
        for (int32_t i = r->boff[b]; i < r->boff[b + 1]; ++i)
            if (i != r->bhead[b]) r->fin_ffsum[b] += r->d1.ff[r->porig[i]];
    for (int32_t v = 0; v < n; ++v) {
        if (r->d1.parent[v] >= 0) continue;
        const int32_t x = r->rep[v];
        r->fin_x[v] = x;
        r->fin_b[v] = x >= 0 ? r->pblock[x] : -1;
        r->fin_base[v] = x >= 0 ? 0 : r->d1.ff[v];  /* lone: fully folded component */
    }
    /* folded vertices, reverse fold order (a parent before its children):
     * farness(d) = farness(p) + N - 2 w[d] (Seq_CC/CC.cpp:2037-2049) */
    for (int32_t i = r->d1.nfold - 1; i >= 0; --i) {
        const int32_t d = r->d1.order[i];
        const int32_t p = r->d1.parent[d];
        r->fin_x[d] = r->fin_x[p];
        r->fin_b[d] = r->fin_b[p];
        r->fin_base[d] = r->fin_base[p] + r->comp_w[r->vcomp[d]] - 2 * r->d1.w[d];
    }
    return CENT_OK;
}

typedef struct { int32_t u, v; } edge_t;


/* Work items over the packed block graph: consecutive small blocks are
 * grouped up to 64 sources; a block of more than 64 vertices gives one item
 * per 64 consecutive sources over its whole range. */
int cent_reduced_items(struct cent_reduced* r) {
    const int32_t nblk = r->nblocks;
    int64_t cap = r->np / 8 + nblk + 16;
    free(r->items);
    r->items = (struct cent_item*)malloc((size_t)cap * sizeof(struct cent_item));
    if (!r->items) return CENT_E_OOM;
    int32_t ni = 0;
    int32_t b = 0;
    while (b < nblk) {
        const int32_t sz = r->boff[b + 1] - r->boff[b];
        if (sz > 64) {
            for (int32_t s0 = r->boff[b]; s0 < r->boff[b + 1]; s0 += 64) {
                if (ni == cap) break;
                const int32_t ns = (r->boff[b + 1] - s0) < 64 ? (r->boff[b + 1] - s0) : 64;
                r->items[ni].lo = r->boff[b]; r->items[ni].hi = r->boff[b + 1];
                r->items[ni].s0 = s0; r->items[ni].ns = ns;
                ++ni;
            }
            ++b;
        } else {
            int32_t e = b, tot = 0;
            while (e < nblk && tot + (r->boff[e + 1] - r->boff[e]) <= 64) { tot += r->boff[e + 1] - r->boff[e]; ++e; }
            if (ni == cap) break;
            r->items[ni].lo = r->boff[b]; r->items[ni].hi = r->boff[e];
            r->items[ni].s0 = r->boff[b]; r->items[ni].ns = tot;
            ++ni;
            b = e;
        }
    }
    r->nitems = ni;
    return CENT_OK;
}

int cent_reduced_build(const struct CSR* g, struct cent_reduced** out) {
    if (!g || !out) return CENT_E_INVAL;
    *out = NULL;
    const int32_t n = g->csrVSize - 1;
    const size_t N = (size_t)(n > 0 ? n : 1);
    const int32_t* rp = g->csrV;
    const int32_t* col = g->csrE;
    struct cent_reduced* r = (struct cent_reduced*)calloc(1, sizeof(*r));
    if (!r) return CENT_E_OOM;
    r->n = n;
    r->m = g->csrESize / 2;
    int st = cent_d1_fold(g, &r->d1);
    if (st) { free(r); return st; }
    const int32_t* parent1 = r->d1.parent;
    const int64_t* w = r->d1.w;
#define ALIVE(x) (parent1[(x)] < 0)

    /* ---- connected components (of the input graph) and their sizes */
    r->vcomp = (int32_t*)malloc(N * sizeof(int32_t));
    r->comp_w = (int64_t*)malloc(N * sizeof(int64_t));
    r->rep = (int32_t*)malloc(N * sizeof(int32_t));
    if (!r->vcomp || !r->comp_w || !r->rep) { cent_reduced_free(r); return CENT_E_OOM; }
    r->ncomp = cent_components(g, r->vcomp, r->comp_w);
    if (r->ncomp < 0) { cent_reduced_free(r); return CENT_E_OOM; }
    for (int32_t v = 0; v < n; ++v) r->rep[v] = -1;

    /* ---- Hopcroft–Tarjan over unfolded vertices with rdeg > 0 */
    int32_t* disc = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* low = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* dpar = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* it = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* fstack = (int32_t*)malloc(N * sizeof(int32_t));
    edge_t* estack = (edge_t*)malloc((size_t)(g->csrESize / 2 + 1) * sizeof(edge_t));
    int32_t* stamp = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* pblock_of = (int32_t*)malloc(N * sizeof(int32_t)); /* block holding v's tree edge from its DFS parent */
    cent_ivec bverts = {0}, bstart = {0}, bheadv = {0}, bedges_u = {0}, bedges_v = {0}, bestart = {0};
    if (!disc || !low || !dpar || !it || !fstack || !estack || !stamp || !pblock_of) {
        st = CENT_E_OOM;
        goto fail;
    }
    /* disc: -2 folded (skipped), -1 not yet reached, else the DFS time */
    for (int32_t v = 0; v < n; ++v) { disc[v] = ALIVE(v) ? -1 : -2; stamp[v] = -1; pblock_of[v] = -1; }
    int32_t timer = 0, nblk = 0;
    int64_t esp = 0;
    for (int32_t root = 0; root < n; ++root) {
        if (disc[root] != -1 || r->d1.rdeg[root] == 0) continue;
        int32_t fsp = 0;
        fstack[fsp++] = root;
        disc[root] = low[root] = timer++;
        dpar[root] = -1;
        it[root] = rp[root];
        while (fsp > 0) {
            const int32_t v = fstack[fsp - 1];
            /* scan v's remaining edges up to the next tree edge, with v's
             * DFS state in registers */
            const int32_t dv = disc[v], pv = dpar[v], je = rp[v + 1];
            int32_t lv = low[v], j = it[v], x = -1;
            for (; j < je; ++j) {
                const int32_t y = col[j];
                const int32_t dy = disc[y];
                if (dy == -2) continue;
                if (dy == -1) { x = y; ++j; break; }
                if (y != pv && dy < dv) {  /* back edge to an ancestor */
                    estack[esp].u = v; estack[esp].v = y; ++esp;
                    if (dy < lv) lv = dy;
                }
            }
            it[v] = j;
            low[v] = lv;
            if (x >= 0) {
                estack[esp].u = v; estack[esp].v = x; ++esp;
                dpar[x] = v;
                disc[x] = low[x] = timer++;
                it[x] = rp[x];
                fstack[fsp++] = x;
            } else {
                --fsp;
                const int32_t p = dpar[v];
                if (p < 0) continue;
                if (low[v] < low[p]) low[p] = low[v];
                if (low[v] >= disc[p]) {
                    /* pop the block ending at tree edge (p, v); head = p */
                    if ((st = cent_ivec_put(&bstart, (int32_t)bverts.n)) ||
                        (st = cent_ivec_put(&bestart, (int32_t)bedges_u.n)) ||
                        (st = cent_ivec_put(&bheadv, p)))
                        goto fail;
                    for (;;) {
                        --esp;
                        const int32_t a = estack[esp].u, b = estack[esp].v;
                        if ((st = cent_ivec_put(&bedges_u, a)) || (st = cent_ivec_put(&bedges_v, b))) goto fail;
                        if (stamp[a] != nblk) { stamp[a] = nblk; if ((st = cent_ivec_put(&bverts, a))) goto fail; }
                        if (stamp[b] != nblk) { stamp[b] = nblk; if ((st = cent_ivec_put(&bverts, b))) goto fail; }
                        if (dpar[b] == a) pblock_of[b] = nblk;  /* tree edge into b */
                        if (a == p && b == v) break;
                    }
                    ++nblk;
                }
            }
        }
        /* the root's own parent block: the last block it heads */
        pblock_of[root] = nblk - 1;
    }
    if ((st = cent_ivec_push(&bstart, (int32_t)bverts.n)) || (st = cent_ivec_push(&bestart, (int32_t)bedges_u.n)))
        goto fail;

    /* ---- block-cut tree: parent block of each block; children are earlier */
    r->nblocks = nblk;
    r->boff = (int32_t*)malloc(((size_t)nblk + 1) * sizeof(int32_t));
    r->bhead = (int32_t*)malloc(((size_t)nblk + 1) * sizeof(int32_t));
    r->bparent = (int32_t*)malloc(((size_t)nblk + 1) * sizeof(int32_t));
    int64_t* sub_w = (int64_t*)calloc((size_t)nblk + 1, sizeof(int64_t));
    int64_t* hang_w = (int64_t*)malloc(N * sizeof(int64_t)); /* w[v] + sum sub_w of child blocks at v */
    if (!r->boff || !r->bhead || !r->bparent || !sub_w || !hang_w) { free(sub_w); free(hang_w); st = CENT_E_OOM; goto fail; }
    for (int32_t v = 0; v < n; ++v) hang_w[v] = w[v];
    for (int32_t b = 0; b < nblk; ++b) {
        const int32_t h = bheadv.a[b];
        r->bparent[b] = (pblock_of[h] == b) ? -1 : pblock_of[h];
    }
    /* bottom-up (emission order = children first): sub_w(B) = sum of hang_w over B minus its head */
    for (int32_t b = 0; b < nblk; ++b) {
        const int32_t h = bheadv.a[b];
        int64_t s = 0;
        for (int32_t i = bstart.a[b]; i < bstart.a[b + 1]; ++i) {
            const int32_t x = bverts.a[i];
            if (x != h || r->bparent[b] < 0) s += hang_w[x];
        }
        if (r->bparent[b] >= 0) { sub_w[b] = s; hang_w[h] += s; }
        else sub_w[b] = s;  /* root block: the whole component */
    }

    /* ---- packing: blocks in emission order, ids contiguous per block */
    r->np = (int32_t)bverts.n;
    int64_t ne = bedges_u.n;
    r->nnzp = 2 * ne;
    if (r->nnzp > INT32_MAX) { free(sub_w); free(hang_w); st = CENT_E_INVAL; goto fail; }
    const size_t NP = (size_t)(r->np > 0 ? r->np : 1);
    r->prp = (int32_t*)calloc(NP + 1, sizeof(int32_t));
    r->pcol = (int32_t*)malloc((size_t)(r->nnzp > 0 ? r->nnzp : 1) * sizeof(int32_t));
    r->pw = (int64_t*)malloc(NP * sizeof(int64_t));
    r->porig = (int32_t*)malloc(NP * sizeof(int32_t));
    r->pblock = (int32_t*)malloc(NP * sizeof(int32_t));
    int32_t* local = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* fill = (int32_t*)malloc(NP * sizeof(int32_t));
    if (!r->prp || !r->pcol || !r->pw || !r->porig || !r->pblock || !local || !fill) {
        free(sub_w); free(hang_w); free(local); free(fill);
        st = CENT_E_OOM;
        goto fail;
    }
    /* Inside each block, vertices in order of descending block degree (ties
     * in emission order), and each row sorted by id below: the highest-degree
     * vertices of a block share a small id range, and a row lists its
     * highest-degree neighbours — the earliest reached — first. */
    {  /* counting sort per block: a block vertex's degree is below the block size */
        int32_t* bdeg = (int32_t*)calloc(N, sizeof(int32_t));
        int64_t maxbs = 1;
        for (int32_t b = 0; b < nblk; ++b) maxbs = bstart.a[b + 1] - bstart.a[b] > maxbs ? bstart.a[b + 1] - bstart.a[b] : maxbs;
        int32_t* tmp = (int32_t*)malloc((size_t)maxbs * sizeof(int32_t));
        int32_t* cnt = (int32_t*)calloc((size_t)maxbs + 1, sizeof(int32_t));
        if (!bdeg || !tmp || !cnt) {
            free(bdeg); free(tmp); free(cnt); free(sub_w); free(hang_w); free(local); free(fill);
            st = CENT_E_OOM;
            goto fail;
        }
        for (int32_t b = 0; b < nblk; ++b) {
            for (int32_t e = bestart.a[b]; e < bestart.a[b + 1]; ++e) { bdeg[bedges_u.a[e]]++; bdeg[bedges_v.a[e]]++; }
            const int32_t b0 = bstart.a[b], bn = bstart.a[b + 1] - b0;
            int32_t dmax = 0;
            for (int32_t i = 0; i < bn; ++i) {
                const int32_t d = bdeg[bverts.a[b0 + i]];
                cnt[d]++;
                if (d > dmax) dmax = d;
            }
            /* slots for descending degree: degree d starts after all larger ones */
            int32_t at = 0;
            for (int32_t d = dmax; d >= 0; --d) { const int32_t c = cnt[d]; cnt[d] = at; at += c; }
            for (int32_t i = 0; i < bn; ++i) { const int32_t x = bverts.a[b0 + i]; tmp[cnt[bdeg[x]]++] = x; }
            for (int32_t i = 0; i < bn; ++i) { const int32_t x = tmp[i]; bverts.a[b0 + i] = x; bdeg[x] = 0; }
            for (int32_t d = 0; d <= dmax; ++d) cnt[d] = 0;
        }
        free(bdeg);
        free(tmp);
        free(cnt);
    }
    for (int32_t b = 0; b < nblk; ++b) {
        const int32_t h = bheadv.a[b];
        r->boff[b] = bstart.a[b];
        r->bhead[b] = -1;
        const int64_t wc = r->comp_w[r->vcomp[h]];
        for (int32_t i = bstart.a[b]; i < bstart.a[b + 1]; ++i) {
            const int32_t x = bverts.a[i];
            local[x] = i;
            r->porig[i] = x;
            r->pblock[i] = b;
            if (x == h && r->bparent[b] >= 0) {
                r->pw[i] = wc - sub_w[b];
                r->bhead[b] = i;
            } else {
                r->pw[i] = hang_w[x];
            }
            if (r->rep[x] < 0 && (x != h || r->bparent[b] < 0)) r->rep[x] = i;
        }
        for (int32_t e = bestart.a[b]; e < bestart.a[b + 1]; ++e) {
            r->prp[local[bedges_u.a[e]] + 1]++;
            r->prp[local[bedges_v.a[e]] + 1]++;
        }
    }
    r->boff[nblk] = r->np;
    for (int32_t i = 0; i < r->np; ++i) r->prp[i + 1] += r->prp[i];
    for (int32_t i = 0; i < r->np; ++i) fill[i] = r->prp[i];
    for (int32_t b = 0; b < nblk; ++b) {
        for (int32_t i = bstart.a[b]; i < bstart.a[b + 1]; ++i) local[bverts.a[i]] = i;
        for (int32_t e = bestart.a[b]; e < bestart.a[b + 1]; ++e) {
            const int32_t a = local[bedges_u.a[e]], c = local[bedges_v.a[e]];
            r->pcol[fill[a]++] = c;
            r->pcol[fill[c]++] = a;
        }
    }
    /* sorted rows by one transpose: appending a to each neighbour's row for
     * a = 0, 1, ... writes every row in ascending order (the graph is
     * symmetric, so the transpose has the same rows) */
    {
        int32_t* sorted = (int32_t*)malloc((size_t)(r->nnzp > 0 ? r->nnzp : 1) * sizeof(int32_t));
        if (!sorted) { free(local); free(fill); free(sub_w); free(hang_w); st = CENT_E_OOM; goto fail; }
        for (int32_t i = 0; i < r->np; ++i) fill[i] = r->prp[i];
        for (int32_t a = 0; a < r->np; ++a)
            for (int32_t j = r->prp[a]; j < r->prp[a + 1]; ++j) sorted[fill[r->pcol[j]]++] = a;
        free(r->pcol);
        r->pcol = sorted;
    }
    free(local); free(fill); free(sub_w); free(hang_w);

    /* ---- articulation vertices: copies in more than one block */
    {
        int32_t* cnt = (int32_t*)calloc(N, sizeof(int32_t));
        if (!cnt) { st = CENT_E_OOM; goto fail; }
        for (int32_t i = 0; i < r->np; ++i) cnt[r->porig[i]]++;
        r->n_ap = 0;
        for (int32_t v = 0; v < n; ++v) r->n_ap += cnt[v] > 1;
        free(cnt);
    }

    /* ---- work items, finish tables */
    if ((st = cent_reduced_items(r))) goto fail;
    if ((st = cent_reduced_tables(r))) goto fail;
    st = CENT_OK;
fail:
    free(disc); free(low); free(dpar); free(it); free(fstack); free(estack); free(stamp); free(pblock_of);
    cent_ivec_free(&bverts); cent_ivec_free(&bstart); cent_ivec_free(&bheadv);
    cent_ivec_free(&bedges_u); cent_ivec_free(&bedges_v); cent_ivec_free(&bestart);
    if (st) { cent_reduced_free(r); return st; }
    *out = r;
#undef ALIVE
    return CENT_OK;
}

/*
 * Block constants.  For a copy x in block B, F(x) is the far-field of x seen
 * from B: the sum over the vertices hanging at x outside B (x included) of
 * ff[y] + d(x, y) * w[y].  Then K_B = sum_{x in B} F(x) and
 *   child-side copy x (not B's head):  F(x) = ff[x] + sum_{C child block at x} down(C)
 *   down(C) = D(head copy of C) + sum_{z in C, z != head} F(z)
 *   head copy h of B:                  F(h) = farness(v_h) - down(B),
 *   farness(v_h) = D(copy of v_h in its parent block P) + K_P.
 * Blocks are stored children-first, so one forward pass gives down(), one
 * backward pass gives K.
 */
/* Block constants K_B from the packed distance sums D, which need to be
 * valid only at each non-root block's head copy and at the head vertex's
 * copy in the parent block. */
static int reduced_block_constants(const struct cent_reduced* r, const int64_t* D, int64_t* K) {
    /* Over the block-cut tree, blocks in index order (children before
     * parents).  The copies of block b other than its head hang off
     *   s_b = sum of ff over them (fin_ffsum) + down(C) of every child block C,
     * down(C) = D at C's head + s_C being the weighted distance sum from C's
     * head vertex to everything below it through C. */
    const int32_t nb = r->nblocks;
    int64_t* down = (int64_t*)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t* child = (int64_t*)calloc((size_t)nb + 1, sizeof(int64_t));
    if (!down || !child) { free(down); free(child); return CENT_E_OOM; }
    /* bottom-up */
    for (int32_t b = 0; b < nb; ++b) {
        if (r->bparent[b] < 0) continue;
        down[b] = D[r->bhead[b]] + r->fin_ffsum[b] + child[b];
        child[r->bparent[b]] += down[b];
    }
    /* top-down (parents before children) */
    for (int32_t b = nb - 1; b >= 0; --b) {
        const int64_t s = r->fin_ffsum[b] + child[b];
        if (r->bparent[b] < 0) {
            K[b] = s;
        } else {
            /* farness of the head vertex from its copy in the parent block */
            const int32_t P = r->bparent[b];
            const int32_t xp = r->rep[r->porig[r->bhead[b]]];  /* the head vertex's copy in P */
            if (xp < 0 || r->pblock[xp] != P) { free(down); free(child); return CENT_E_INVAL; }
            const int64_t far_h = D[xp] + K[P];
            K[b] = (far_h - down[b]) + s;
        }
    }
    free(down); free(child);
    return CENT_OK;
}

int cent_reduced_finish(const struct cent_reduced* r, const int64_t* D, int64_t* farness) {
    if (!r || (!D && r->np > 0) || !farness) return CENT_E_INVAL;
    const int32_t n = r->n;
    int64_t* K = (int64_t*)calloc((size_t)r->nblocks + 1, sizeof(int64_t));
    if (!K) return CENT_E_OOM;
    const int st = reduced_block_constants(r, D, K);
    if (st) { free(K); return st; }
    /* farness(v) = D_B(x) + K_B at the copy x answering for v's unfolded
     * ancestor, plus the fold chain's terms (tables: cent_reduced_tables) */
    for (int32_t v = 0; v < n; ++v) {
        const int32_t x = r->fin_x[v];
        farness[v] = r->fin_base[v] + (x >= 0 ? D[x] + K[r->fin_b[v]] : 0);
    }
    free(K);
    return CENT_OK;
}

/* farness of the given source vertices only (assigned, one per source).  D
 * needs to be valid at every copy reduced_block_constants reads and at the
 * copy answering for each source's unfolded ancestor (a folded vertex is
 * recovered from its parent: farness(d) = farness(p) + N - 2 w[d]). */
int cent_reduced_farness_sources(const cent_reduced* r, const int64_t* D, const int32_t* src, int32_t k,
                                 int64_t* far) {
    if (!r || k < 0 || (k > 0 && (!src || !far)) || (!D && r->np > 0)) return CENT_E_INVAL;
    for (int32_t i = 0; i < k; ++i)
        if (src[i] < 0 || src[i] >= r->n) return CENT_E_INVAL;
    int64_t* K = (int64_t*)calloc((size_t)r->nblocks + 1, sizeof(int64_t));
    if (!K) return CENT_E_OOM;
    const int st = reduced_block_constants(r, D, K);
    if (st) { free(K); return st; }
    for (int32_t i = 0; i < k; ++i) {
        const int32_t v = src[i], x = r->fin_x[v];
        far[i] = r->fin_base[v] + (x >= 0 ? D[x] + K[r->fin_b[v]] : 0);
    }
    free(K);
    return CENT_OK;
}

/* Source-count weights of a source multiset S (sampled betweenness through
 * the reduction): the counterpart of the vertex weights W with every vertex
 * counted by its multiplicity in S instead of 1.  wsf[v] (optional, [n]) =
 * sources among v and the vertices folded into it; WS[x] ([np]) = sources
 * hanging off copy x on its side of x's block (all of the component's
 * sources minus those below the block for a block's head copy); Scomp
 * (optional, [ncomp]) = sources per component.  With S = every vertex once,
 * WS equals the copy weights pw. */
static int source_weights(const cent_reduced* r, const int32_t* src, int32_t k, int64_t* WS, int64_t* wsf_out,
                          int64_t* Scomp_out) {
    if (!r || k < 0 || (k > 0 && !src) || (!WS && r->np > 0)) return CENT_E_INVAL;
    const int32_t n = r->n, nb = r->nblocks;
    const size_t N = (size_t)(n > 0 ? n : 1);
    int64_t* ws = (int64_t*)calloc(N, sizeof(int64_t));
    int64_t* down = (int64_t*)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t* Sc = (int64_t*)calloc((size_t)r->ncomp + 1, sizeof(int64_t));
    if (!ws || !down || !Sc) { free(ws); free(down); free(Sc); return CENT_E_OOM; }
    for (int32_t i = 0; i < k; ++i) {
        if (src[i] < 0 || src[i] >= n) { free(ws); free(down); free(Sc); return CENT_E_INVAL; }
        ws[src[i]]++;
        Sc[r->vcomp[src[i]]]++;
    }
    /* fold subtrees (children fold before their parents) */
    for (int32_t i = 0; i < r->d1.nfold; ++i) {
        const int32_t d = r->d1.order[i];
        ws[r->d1.parent[d]] += ws[d];
    }
    if (wsf_out) memcpy(wsf_out, ws, N * sizeof(int64_t));
    /* bottom-up over the block-cut tree: ws[v] gathers the child blocks headed at v */
    for (int32_t b = 0; b < nb; ++b) {
        if (r->bparent[b] < 0) continue;
        const int32_t h = r->bhead[b];
        int64_t t = 0;
        for (int32_t i = r->boff[b]; i < r->boff[b + 1]; ++i)
            if (i != h) t += ws[r->porig[i]];
        down[b] = t;
        ws[r->porig[h]] += t;
    }
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t h = r->bparent[b] >= 0 ? r->bhead[b] : -1;
        for (int32_t i = r->boff[b]; i < r->boff[b + 1]; ++i) {
            const int32_t v = r->porig[i];
            WS[i] = (i == h) ? Sc[r->vcomp[v]] - down[b] : ws[v];
        }
    }
    if (Scomp_out) memcpy(Scomp_out, Sc, (size_t)(r->ncomp > 0 ? r->ncomp : 1) * sizeof(int64_t));
    free(ws); free(down); free(Sc);
    return CENT_OK;
}

/* ---- C-ABI wrappers (include/cent.h) */
int cent_reduce(const struct CSR* csr, cent_reduced** out) {
    if (!csr || !out) return CENT_E_INVAL;
    const int multi = cent_csr_is_multigraph(csr);
    if (multi < 0) return CENT_E_OOM;
    if (!multi) return cent_reduced_build(csr, out);
    /* distances (hence farness) ignore parallel edges and self-loops */
    struct CSR* simple = NULL;
    int st = cent_csr_simple(csr, &simple);
    if (st) return st;
    st = cent_reduced_build(simple, out);
    cent_csr_free(simple);
    if (!st) (*out)->multi = 1;
    return st;
}

int cent_reduced_get_view(const cent_reduced* r, cent_reduced_view* v) {
    if (!r || !v) return CENT_E_INVAL;
    v->n = r->n; v->np = r->np; v->nnzp = r->nnzp; v->nblocks = r->nblocks; v->nitems = r->nitems;
    v->nfold = r->d1.nfold; v->n_ap = r->n_ap; v->ncomp = r->ncomp;
    v->prp = r->prp; v->pcol = r->pcol; v->pw = r->pw; v->porig = r->porig; v->boff = r->boff;
    return CENT_OK;
}

int cent_reduced_farness(const cent_reduced* r, const int64_t* D, int64_t* farness) {
    return cent_reduced_finish(r, D, farness);
}

void cent_reduced_drop_device(cent_reduced* r) {
    if (!r) return;
    if (r->dev && r->dev_free) r->dev_free(r->dev);
    r->dev = NULL;
    r->dev_free = NULL;
}

/*
 * Betweenness through the reduction (no reference counterpart: the reference
 * reduces the graph for closeness only).  For a vertex v of a component with
 * N vertices, let the pieces of v be the connected parts of the component
 * once v is removed.  Every shortest path between two different pieces runs
 * through v; a shortest path between two vertices of the same piece can only
 * run through v inside a biconnected block that holds v, entering and leaving
 * that block at its copies' weighted vertices.  Hence
 *   BC(v) = [(N-1)^2 - sum of squared piece sizes]
 *         + sum over the blocks B that hold v of the weighted in-block term
 *           sum_{a != b, both != v} W(a) W(b) sigma_ab(v) / sigma_ab,
 * the second part being what the GPU accumulates at v's copy in B (weighted
 * Brandes from every copy, source and target weights W).  The pieces of v are
 * the trees folded into v (sizes w[c] of its fold children c) and, for every
 * block B holding v, the rest of the component behind B: N - W(v's copy in B)
 * (for a folded vertex: N - w[v]).  Ordered pairs, like computeBC summed over
 * all sources (Seq_Brandes/Brandes.c:249).
 */
int cent_reduced_finish_bc(const struct cent_reduced* r, const double* bcp, double* bc) {
    if (!r || (!bcp && r->np > 0) || !bc) return CENT_E_INVAL;
    const int32_t n = r->n;
    const size_t N = (size_t)(n > 0 ? n : 1);
    double* sumsq = (double*)calloc(N, sizeof(double));  /* sum of squared piece sizes */
    double* inblock = (double*)calloc(N, sizeof(double));
    if (!sumsq || !inblock) { free(sumsq); free(inblock); return CENT_E_OOM; }
    for (int32_t i = 0; i < r->d1.nfold; ++i) {
        const int32_t c = r->d1.order[i];
        const double wc = (double)r->d1.w[c];
        sumsq[r->d1.parent[c]] += wc * wc;
        const double rest = (double)(r->comp_w[r->vcomp[c]] - r->d1.w[c]);
        sumsq[c] += rest * rest;
    }
    for (int32_t x = 0; x < r->np; ++x) {
        const int32_t v = r->porig[x];
        const double rest = (double)(r->comp_w[r->vcomp[v]] - r->pw[x]);
        sumsq[v] += rest * rest;
        inblock[v] += bcp[x];
    }
    for (int32_t v = 0; v < n; ++v) {
        const double s = (double)(r->comp_w[r->vcomp[v]] - 1);
        bc[v] += (s * s - sumsq[v]) + inblock[v];
    }
    free(sumsq);
    free(inblock);
    return CENT_OK;
}

int cent_reduced_source_weights(const cent_reduced* r, const int32_t* src, int32_t k, int64_t* WS) {
    return source_weights(r, src, k, WS, NULL, NULL);
}

/* Betweenness from a source multiset S through the reduction (sampled
 * sources; the all-source case is cent_reduced_finish_bc).  Counting the
 * ordered pairs (s, t), s in S, t any other vertex, whose shortest paths run
 * through v:
 *   BC_S(v) = sum over the pieces P of v of S(P) (N - 1 - |P|)
 *           + sum over the blocks B holding v of the in-block term
 *             sum_{a != b, both != v} WS(a) W(b) sigma_ab(v) / sigma_ab,
 * the first part being the pairs whose endpoints lie in different pieces
 * (S(P) = sources in P: wsf of a fold child's subtree, S_comp - wsf[c] for a
 * folded vertex's rest, S_comp - WS(x) behind the block of copy x), the
 * second what the GPU accumulates at v's copies (weighted Brandes from
 * every copy with WS > 0, source weight WS, target weight W).  With every
 * vertex once in S this is cent_reduced_finish_bc's formula. */
int cent_reduced_source_weights_all(const cent_reduced* r, const int32_t* src, int32_t k, int64_t* WS, int64_t* wsf,
                                    int64_t* Sc) {
    return source_weights(r, src, k, WS, wsf, Sc);
}

int cent_reduced_bc_sources(const cent_reduced* r, const int32_t* src, int32_t k, const double* bcp, double* bc) {
    if (!r || (!bcp && r->np > 0) || !bc || k < 0 || (k > 0 && !src)) return CENT_E_INVAL;
    if (r->multi) return CENT_E_UNSUPPORTED;
    const size_t N = (size_t)(r->n > 0 ? r->n : 1);
    int64_t* WS = (int64_t*)malloc((size_t)(r->np > 0 ? r->np : 1) * sizeof(int64_t));
    int64_t* wsf = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* Sc = (int64_t*)malloc((size_t)(r->ncomp > 0 ? r->ncomp : 1) * sizeof(int64_t));
    if (!WS || !wsf || !Sc) { free(WS); free(wsf); free(Sc); return CENT_E_OOM; }
    int st = source_weights(r, src, k, WS, wsf, Sc);
    if (!st) st = cent_reduced_bc_sources_from(r, WS, wsf, Sc, bcp, bc);
    free(WS); free(wsf); free(Sc);
    return st;
}

int cent_reduced_bc_sources_from(const cent_reduced* r, const int64_t* WS, const int64_t* wsf, const int64_t* Sc,
                                 const double* bcp, double* bc) {
    if (!r || (!bcp && r->np > 0) || !bc || (r->np > 0 && !WS) || !wsf || !Sc) return CENT_E_INVAL;
    if (r->multi) return CENT_E_UNSUPPORTED;
    const int32_t n = r->n;
    const size_t N = (size_t)(n > 0 ? n : 1);
    double* acc = (double*)calloc(N, sizeof(double));
    if (!acc) return CENT_E_OOM;
    int st = CENT_OK;
    {
        for (int32_t i = 0; i < r->d1.nfold; ++i) {
            const int32_t c = r->d1.order[i];
            const int32_t cc = r->vcomp[c];
            const double Nc = (double)r->comp_w[cc], wc = (double)r->d1.w[c];
            acc[r->d1.parent[c]] += (double)wsf[c] * (Nc - 1.0 - wc);      /* c's subtree, a piece of its parent */
            acc[c] += (double)(Sc[cc] - wsf[c]) * (wc - 1.0);             /* the rest, a piece of c */
        }
        for (int32_t x = 0; x < r->np; ++x) {
            const int32_t v = r->porig[x];
            acc[v] += (double)(Sc[r->vcomp[v]] - WS[x]) * ((double)r->pw[x] - 1.0) + bcp[x];
        }
        for (int32_t v = 0; v < n; ++v) bc[v] += acc[v];
    }
    free(acc);
    return st;
}

int cent_reduced_bc(const cent_reduced* r, const double* bcp, double* bc) {
    if (r && r->multi) return CENT_E_UNSUPPORTED;
    return cent_reduced_finish_bc(r, bcp, bc);
}
