/* Graph reduction for closeness: degree-1 folding + biconnected-block split.
 * Internal header shared by Lib/D1Process, Lib/AP_Process and the engine.
 *
 * The reduced form is a "packed block graph": every biconnected block of the
 * folded graph becomes its own connected piece with private vertex ids
 * (an articulation vertex gets one copy per block), laid out block after
 * block.  A BFS from a packed vertex never leaves its block, so blocks are
 * independent work for the GPU.  With per-copy weights W, the farness of a
 * vertex v whose copy x lies in block B is
 *     farness(v) = sum_{y in B} d_B(x, y) * W(y)  +  K_B
 * (the weighted distance sum is computed on the GPU, K_B by a linear pass
 * over the block-cut tree on the host, see cent_reduced_finish).
 */
#ifndef CENT_LIB_REDUCE_H
#define CENT_LIB_REDUCE_H

#include <stdint.h>
#include "../../include/cent.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Degree-1 folding (the role of D1Folding, Lib/D1Process/D1Process.c:5-154). */
struct cent_d1 {
    int32_t  n;
    int32_t  nfold;        /* number of folded vertices                       */
    int32_t* order;        /* [nfold] folded vertices in fold order            */
    int32_t* parent;       /* [n] vertex a folded vertex was folded into, or -1 */
    int32_t* rdeg;         /* [n] degree after folding (0 for folded vertices) */
    int64_t* w;            /* [n] 1 + vertices folded into v (transitively)    */
    int64_t* ff;           /* [n] sum of distances from v to those vertices    */
};

int  cent_d1_fold(const struct CSR* g, struct cent_d1* out);

/* 1 if a row lists a neighbour twice or lists its own vertex (a graph
 * loaded with CENT_CSR_RAW), 0 if not, -1 when out of memory.  Both
 * reductions work on the simple graph of such an input (Lib/CSR/CSR.c). */
int cent_csr_is_multigraph(const struct CSR* g);
/* The simple graph under g (self-loops dropped, repeated neighbours once),
 * same ids and id base. */
int cent_csr_simple(const struct CSR* g, struct CSR** out);
void cent_d1_free(struct cent_d1* d);

/* One schedulable unit of GPU work: sources s0 .. s0+ns-1 (packed ids, ns <= 64)
 * whose BFS stays inside the packed id range [lo, hi). */
struct cent_item {
    int32_t lo, hi, s0, ns;
};

struct cent_reduced {
    int32_t n;              /* vertices of the input graph                     */
    int64_t m;              /* undirected edges of the input graph             */
    struct cent_d1 d1;
    /* packed block graph */
    int32_t np;             /* packed vertices (sum of block sizes)            */
    int64_t nnzp;           /* directed packed edges                           */
    int32_t* prp;           /* [np + 1]                                         */
    int32_t* pcol;          /* [nnzp]                                           */
    int64_t* pw;            /* [np] weight W of each copy                       */
    int32_t* porig;         /* [np] input vertex of each copy                   */
    int32_t* pblock;        /* [np] block of each copy                          */
    int32_t  nblocks;
    int32_t* boff;          /* [nblocks + 1] packed range of each block        */
    int32_t* bhead;         /* [nblocks] packed id of the parent-AP copy, -1 for a root block */
    int32_t* bparent;       /* [nblocks] parent block (-1 for a root block)    */
    int32_t* rep;           /* [n] packed copy that answers for v (-1: folded or lone) */
    int32_t* vcomp;         /* [n] connected component of v                     */
    int64_t* comp_w;        /* [ncomp] vertices in the component               */
    int32_t  ncomp;
    int32_t  n_ap;          /* articulation vertices of the folded graph       */
    int32_t  multi;         /* built from a multigraph (CENT_CSR_RAW input): the
                               reduction is of its simple graph, which has the
                               same distances, so closeness holds and
                               betweenness (path counts) does not            */
    /* work items over the packed graph */
    int32_t  nitems;
    struct cent_item* items;
    /* farness of every input vertex from D in one pass (cent_reduced_tables):
     * farness[v] = fin_base[v] + (fin_x[v] >= 0 ? D[fin_x[v]] + K[fin_b[v]] : 0),
     * fin_x the copy answering for v's unfolded ancestor a (-1: a lone or
     * fully folded vertex), fin_b its block, fin_base the fold-chain terms
     * sum (N - 2 w[d]) from v up to a, plus ff[a] when fin_x is -1 */
    int32_t* fin_x;
    int32_t* fin_b;
    int64_t* fin_base;
    int64_t* fin_ffsum;     /* [nblocks] sum of ff over the block's copies but its head */
    /* device-side cache owned by the engine, released by cent_reduced_free */
    void*    dev;
    void   (*dev_free)(void*);
};

/* Fold, split, pack.  Linear time in n + m. */
int  cent_reduced_build(const struct CSR* g, struct cent_reduced** out);
void cent_reduced_free(struct cent_reduced* r);

/* Fills r->items / r->nitems from r->boff (both builds share it). */
int  cent_reduced_items(struct cent_reduced* r);
/* Fills the fin_* tables (both builds share it). */
int  cent_reduced_tables(struct cent_reduced* r);

/* cent_reduced_source_weights with the per-vertex fold-subtree counts wsf
 * ([n]) and the sources per component Sc ([ncomp]) as well, and
 * cent_reduced_bc_sources from those three arrays (so a caller that needs
 * the weights computes them once). */
int  cent_reduced_source_weights_all(const struct cent_reduced* r, const int32_t* src, int32_t k, int64_t* WS,
                                     int64_t* wsf, int64_t* Sc);
int  cent_reduced_bc_sources_from(const struct cent_reduced* r, const int64_t* WS, const int64_t* wsf,
                                  const int64_t* Sc, const double* bcp, double* bc);

/* Given D[x] = sum_y d_B(x, y) W(y) for every packed vertex x, fills the
 * farness of every input vertex (assigned, int64), including folded
 * vertices (recovered from their parents) and isolated / fully folded ones. */
int  cent_reduced_finish(const struct cent_reduced* r, const int64_t* D, int64_t* farness);

/* Given bcp[x] = sum over sources a in x's block of W(a) * delta^W_a(x)
 * (weighted Brandes on the packed graph), adds the betweenness of every
 * input vertex into bc (ordered pairs, like computeBC over all sources). */
int  cent_reduced_finish_bc(const struct cent_reduced* r, const double* bcp, double* bc);

#ifdef __cplusplus
}
#endif
#endif
