/*
 * Degree-1 folding — the role of D1Folding (Lib/D1Process/D1Process.c:5-154).
 *
 * Same arithmetic as the reference: a leaf d hanging from p is folded into p
 * with w[p] += w[d] and ff[p] += w[d] + ff[d] (D1Process.c:43-46), leaves
 * created by a fold are folded in turn (D1Process.c:66-69), and the fold
 * order is kept so closeness can be recovered afterwards as
 *     farness(d) = farness(p) + N - 2 * w[d]
 * (Seq_CC/CC.cpp:2037-2049; N = vertices of d's connected component — moving
 * the source across the edge (d, p) brings the w[d] vertices on d's side one
 * step closer and the other N - w[d] one step further).
 *
 * Unlike the reference it does not mutate the input CSR: the remaining
 * degree is tracked in rdeg[] and the parent of a popped leaf is the one
 * neighbour that is still unfolded.  A component that is a tree folds down to
 * a single vertex with rdeg == 0 (which then holds the whole tree).
 */
#include <stdio.h>
#include "../common.h"
#include "../reduce.h"

void cent_d1_free(struct cent_d1* d) {
    if (!d) return;
    free(d->order); free(d->parent); free(d->rdeg); free(d->w); free(d->ff);
    memset(d, 0, sizeof(*d));
}

int cent_d1_fold(const struct CSR* g, struct cent_d1* out) {
    if (!g || !out) return CENT_E_INVAL;
    const int32_t n = g->csrVSize - 1;
    const size_t N = (size_t)(n > 0 ? n : 1);
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->order = (int32_t*)malloc(N * sizeof(int32_t));
    out->parent = (int32_t*)malloc(N * sizeof(int32_t));
    out->rdeg = (int32_t*)malloc(N * sizeof(int32_t));
    out->w = (int64_t*)malloc(N * sizeof(int64_t));
    out->ff = (int64_t*)calloc(N, sizeof(int64_t));
    int32_t* q = (int32_t*)malloc(N * sizeof(int32_t));
    if (!out->order || !out->parent || !out->rdeg || !out->w || !out->ff || !q) {
        free(q);
        cent_d1_free(out);
        return CENT_E_OOM;
    }
    int64_t qh = 0, qt = 0;
    for (int32_t v = 0; v < n; ++v) {
        out->parent[v] = -1;
        out->w[v] = 1;
        out->rdeg[v] = g->csrV[v + 1] - g->csrV[v];
        if (out->rdeg[v] == 1) q[qt++] = v;
    }
    int32_t nf = 0;
    while (qh < qt) {
        const int32_t d = q[qh++];
        if (out->rdeg[d] != 1) continue;  /* its partner was folded into it */
        int32_t p = -1;
        for (int32_t j = g->csrV[d]; j < g->csrV[d + 1]; ++j) {
            const int32_t y = g->csrE[j];
            if (out->parent[y] < 0 && out->rdeg[y] > 0) { p = y; break; }
        }
        if (p < 0) continue;  /* cannot happen for a consistent rdeg */
        out->parent[d] = p;
        out->rdeg[d] = 0;
        out->w[p] += out->w[d];
        out->ff[p] += out->w[d] + out->ff[d];
        out->order[nf++] = d;
        if (--out->rdeg[p] == 1) q[qt++] = p;
    }
    out->nfold = nf;
    free(q);
    return CENT_OK;
}
