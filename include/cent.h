/*
 * cent.h — C-ABI of the MI355X centrality engine (betweenness + closeness on
 * undirected, unweighted graphs).
 *
 * Every entry point returns an int status (CENT_OK == 0) and never exits the
 * process.  Graph ids are 0-based int32; edge offsets are int32 (the engine
 * rejects graphs with more than INT32_MAX directed edges).  Host buffers are
 * owned by the caller; device buffers are owned by the engine.  Inputs are
 * never mutated.
 *
 * Reference interfaces replaced (paths relative to the XPPGX/BC tree):
 *   struct CSR / createCSR          Lib/CSR/CSR.h:34-97, Lib/CSR/CSR.c:9-108
 *   buildGraph (edge-list loader)   Lib/AdjList/AdjList.h:53, AdjList.c:13-150
 *   D1Folding                       Lib/D1Process/D1Process.c:5-154
 *   AP_detection / AP_Copy_And_Split / rebuildGraph
 *                                   Lib/AP_Process/AP_Process.c:6-225,468-1065,1141-1321
 *   computeBC                       Seq_Brandes/Brandes.c:224-372
 *   computeCC                       Seq_CC/CC.cpp:37-117
 *   compute_D1_AP_CC_shareBased_DegreeOrder
 *                                   Seq_CC/CC.cpp:1728-2052
 */
#ifndef CENT_H
#define CENT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    CENT_OK          = 0,
    CENT_E_INVAL     = 1,   /* bad argument (null pointer, id out of range, ...) */
    CENT_E_OOM       = 2,   /* host or device allocation failed / over budget   */
    CENT_E_DEVICE    = 3,   /* HIP runtime error, or no usable gfx950 device     */
    CENT_E_OVERFLOW  = 4,   /* shortest-path counts of one BFS level span more
                               than FP64's range even after per-level scaling  */
    CENT_E_IO        = 5,   /* file could not be opened or parsed               */
    CENT_E_BUSY      = 6,   /* server queue full                                */
    CENT_E_CANCELLED = 7,   /* request cancelled                                */
    CENT_E_NOTFOUND  = 8,   /* unknown request / graph id                       */
    CENT_E_UNSUPPORTED = 9  /* valid request the engine does not handle         */
};

const char* cent_strerror(int status);

/* ------------------------------------------------------------------- graph */
/*
 * Own CSR.  Field names follow the reference's struct CSR (Lib/CSR/CSR.h:34-95)
 * where they carry the same meaning; the layout is 0-based and immutable:
 * vertex v's neighbours are csrE[csrV[v] .. csrV[v+1]).  Self-loops and
 * duplicate edges are removed on construction (unless CENT_CSR_RAW keeps
 * them), both directions are stored.
 * For a 1-based input file vertex 0 exists with degree 0 (startNodeID = 1),
 * so outputs stay indexed by the file's own ids.
 */
struct CSR {
    int32_t* csrV;            /* csrVSize = n + 1 row offsets                  */
    int32_t* csrE;            /* csrESize column ids                           */
    int32_t* csrNodesDegree;  /* n degrees                                     */
    int64_t  csrESize;        /* number of directed edges (2 x undirected)     */
    int32_t  csrVSize;        /* n + 1                                         */
    int32_t  startNodeID;     /* first id present in the input (0 or 1)        */
    int32_t  endNodeID;       /* last id (n - 1)                               */
    int32_t  totalNodeNumber; /* endNodeID - startNodeID + 1                   */
    int32_t  maxDegree;
    int32_t  startAtZero;     /* 1 if the input ids start at 0                 */
    int32_t  kind;            /* CENT_CSR_KIND_*, set by the constructors      */
};
/* struct CSR.kind: whether rows may repeat a neighbour or list their own
 * vertex (a CENT_CSR_RAW input).  UNKNOWN (0, a CSR built by the caller)
 * makes the engine look. */
#define CENT_CSR_KIND_UNKNOWN 0
#define CENT_CSR_KIND_SIMPLE  1
#define CENT_CSR_KIND_MULTI   2

/* Build from an undirected edge list (u[i], v[i]), ids in [0, n).
 * flags: CENT_CSR_LCC keeps only the largest connected component and
 * relabels it to 0..n'-1 preserving id order; old_id (optional, length n')
 * receives the original id of every kept vertex when non-null after the
 * call (allocated by the library, free with cent_free).  Replaces createCSR
 * (Lib/CSR/CSR.c:9-108), which also did no cleaning. */
#define CENT_CSR_LCC 1
/* flags: CENT_CSR_RAW keeps the input as the reference's loader does
 * (Lib/AdjList/AdjList.c:76-92, no cleaning): a repeated edge stays a
 * parallel edge (it adds a shortest path wherever it lies on one, like the
 * reference's per-edge path-count sums), and a self-loop u-u lists u twice
 * in its own row.  Rows stay sorted by neighbour id.  Not with CENT_CSR_LCC.
 * Every engine call accepts such a CSR; the reduced (D1/AP) form works on
 * its simple graph, which has the same distances, so reduced closeness is
 * unchanged and reduced betweenness returns CENT_E_UNSUPPORTED for it. */
#define CENT_CSR_RAW 2
int  cent_csr_from_edges(int32_t n, const int32_t* u, const int32_t* v, int64_t m,
                         int32_t flags, struct CSR** out, int32_t** old_id);
/* Copy of existing CSR arrays (validated: offsets monotone, ids in range). */
int  cent_csr_from_arrays(int32_t n, const int32_t* row_ptr, const int32_t* col,
                          struct CSR** out);
void cent_csr_free(struct CSR* csr);
void cent_free(void* p);

/* Loader for the reference's text edge-list format ("n m" header, then m
 * lines "u v"; 0- or 1-based ids detected as in Lib/AdjList/AdjList.c:39-86).
 * Replaces buildGraph + createCSR. */
int  cent_load_edgelist(const char* path, struct CSR** out);
/* The same with CSR flags: CENT_CSR_RAW keeps repeated edges and self-loops
 * exactly as the file lists them (the reference's buildGraph does no
 * cleaning). */
int  cent_load_edgelist_ex(const char* path, int32_t flags, struct CSR** out);

/* Seeded synthetic generators (host C, deterministic for a given seed). */
/* R-MAT, 2^scale ids, n_samples edge draws with quadrant probabilities
 * a, b, c (d = 1-a-b-c), random id permutation, cleaned and symmetrised. */
int  cent_gen_rmat(int32_t scale, int64_t n_samples, double a, double b, double c,
                   uint64_t seed, int32_t flags, struct CSR** out);
/* rows x cols 4-neighbour grid, del_frac of the edges removed uniformly at
 * random, largest component kept, then n_shortcuts random extra edges. */
int  cent_gen_grid(int32_t rows, int32_t cols, double del_frac, int32_t n_shortcuts,
                   uint64_t seed, struct CSR** out);
/* Block tree: n_blocks biconnected blocks (size Pareto(alpha) clipped to
 * [smin, smax], ER of average degree avg_deg plus a Hamiltonian cycle),
 * joined in a random tree through shared articulation vertices, then
 * pendant_frac x (block vertices) extra vertices as pendant chains whose
 * lengths are Geometric(p_geo) on {1, 2, ...}. */
int  cent_gen_blocktree(int32_t n_blocks, double alpha, int32_t smin, int32_t smax,
                        double avg_deg, double pendant_frac, double p_geo,
                        uint64_t seed, struct CSR** out);

/* Source sampling: k ids drawn uniformly without replacement from [0, n)
 * (k <= n), in draw order. */
int  cent_sample_sources(int32_t n, int32_t k, uint64_t seed, int32_t* out);

/* ---------------------------------------------------------- engine (GPU) */
typedef struct cent_ctx   cent_ctx;    /* one HIP device + stream + workspace  */
typedef struct cent_graph cent_graph;  /* device-resident CSR                  */

/* BC accumulation formula. TEXTBOOK: BC[w] += delta_s[w] once per source
 * (Brandes 2001).  REF_COMPAT: BC[w] += delta_s[w] once per shortest-path
 * predecessor of w, the formula of Seq_Brandes/Brandes.c:317-327. */
#define CENT_BC_TEXTBOOK   0
#define CENT_BC_REF_COMPAT 1

/* opts.flags */
#define CENT_OPT_TIMING    1   /* record per-kernel HIP events (stats)         */
#define CENT_OPT_SOURCE_ORDER 2 /* 64-lane steps: batch the sources in the given
                                   order (default: grouped by BFS distance to
                                   the highest-degree vertex, then degree, so a
                                   batch's lanes share levels; results are the
                                   same up to FP64 summation order)          */
#define CENT_OPT_ONE_STREAM 4 /* dense 64-lane step: keep every launch group
                                   on the context's stream (default: groups
                                   alternate between it and a second stream
                                   of a peer workspace on the same GPU, whose
                                   BC / distance sums are added at the end;
                                   results equal up to FP64 summation order) */

/* opts.frontier: how a level finds its work.  DENSE scans every vertex
 * (pull; best for low-diameter graphs), SPARSE keeps compacted frontier
 * lists and replays hipGraph-captured runs of levels (best for hundreds of
 * levels with small frontiers), AUTO picks from the graph's BFS-depth probe. */
#define CENT_FRONTIER_AUTO   0
#define CENT_FRONTIER_DENSE  1   /* 64-source bit-parallel pull over all vertices */
#define CENT_FRONTIER_SPARSE 2   /* 64-source bit-parallel, frontier queues       */
#define CENT_FRONTIER_SOURCE 3   /* one workgroup per source (deep graphs)         */

typedef struct cent_opts {
    int32_t bc_mode;      /* CENT_BC_TEXTBOOK or CENT_BC_REF_COMPAT          */
    int32_t flags;        /* CENT_OPT_*                                      */
    int32_t block_size;   /* reserved, 0                                     */
    int32_t grid_waves;   /* 0 = default cap on waves per dense level launch */
    int32_t frontier;     /* CENT_FRONTIER_*                                 */
    int32_t max_slots;    /* SOURCE mode: cap on concurrent per-source slots
                             (0 = sized from free device memory)          */
    int32_t push_pct;     /* DENSE betweenness: a batch's level 2 is pushed
                             from its level-1 frontier when that frontier's
                             rows hold under push_pct % of the traversed
                             edges, else pulled (0 = the graph's tuned value
                             or the engine default, -1 = always pulled)     */
} cent_opts;

typedef struct cent_stats {
    double  ms_fwd;           /* summed duration of forward level kernels    */
    double  ms_bwd;           /* summed duration of backward level kernels   */
    double  ms_total;         /* wall time of the last call (host clock)     */
    int64_t launches_fwd;
    int64_t launches_bwd;
    int64_t batches;          /* 64-lane source batches processed            */
    int64_t levels;           /* summed BFS depth over batches               */
    int64_t dag_edges;        /* sum over sources of |shortest-path DAG|     */
    int64_t reached;          /* sum over sources of reached vertices        */
    int64_t edges_reached;    /* sum over sources of directed edges of the
                                 reached component (2 x m_cc(s))             */
} cent_stats;

int  cent_ctx_create(int32_t device, cent_ctx** out);
void cent_ctx_destroy(cent_ctx* ctx);
int  cent_ctx_stats(const cent_ctx* ctx, cent_stats* out);
int  cent_ctx_reset_stats(cent_ctx* ctx);
/* Device bytes the context's workspace currently holds, and releasing them
 * (the next call re-allocates what it needs). */
int64_t cent_ctx_workspace_bytes(const cent_ctx* ctx);
int  cent_ctx_trim(cent_ctx* ctx);
/* Text of the last HIP error seen by this context (empty if none). */
const char* cent_ctx_last_error(const cent_ctx* ctx);
/* Measurement helper (no reference counterpart): times `iters` launches of a
 * plain device copy of `bytes` bytes and returns read+write GB/s.  Serves as
 * the calibration kernel for HBM counter readings. */
int  cent_probe_copy_bandwidth(cent_ctx* ctx, int64_t bytes, int32_t iters, double* gbs);
/* Measurement helper (no reference counterpart): times `iters` launches of
 * random gathers from a zeroed buffer of `bytes` bytes, each gather reading
 * one `piece`-byte aligned piece (32, 64, 128, 256 or 512), four in flight
 * per lane, and returns the useful bytes read per second in GB/s: the
 * ceiling for gathers of row slices of that size. */
int  cent_probe_gather_bandwidth(cent_ctx* ctx, int64_t bytes, int32_t piece, int32_t iters, double* gbs);
/* Measurement helper (no reference counterpart): times `iters` launches in
 * which every wave load instruction reads `lines` (1, 2 or 4) random whole
 * 128-byte lines of a zeroed buffer of `bytes` bytes, and returns the line
 * bytes read per second in GB/s: the ceiling for gathers whose loads touch
 * that many lines each (the per-instruction rate limits 1 and 2). */
int  cent_probe_gather_lines(cent_ctx* ctx, int64_t bytes, int32_t lines, int32_t iters, double* gbs);

/* Uploads a CSR (validated; ids in range, offsets monotone).  By default the
 * device copy is relabelled for cache locality: by descending degree when
 * the BFS from the highest-degree vertex is at most 48 levels deep, else in
 * BFS order from that vertex (each further component in BFS order from its
 * smallest id).  Every input and output of the engine stays in the caller's
 * ids either way.
 * flags: CENT_GRAPH_KEEP_ORDER keeps the caller's ids on the device too;
 * CENT_GRAPH_DEGREE_ORDER / CENT_GRAPH_BFS_ORDER force either relabel.
 * The device copy lives on ctx's GPU and does not depend on ctx itself: any
 * context of that GPU may use it, after ctx is destroyed too.  The same
 * holds for the device copy of a reduced graph (made by the first reduced
 * call on it). */
#define CENT_GRAPH_KEEP_ORDER 1
#define CENT_GRAPH_DEGREE_ORDER 4
#define CENT_GRAPH_BFS_ORDER 8
int  cent_graph_create(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col,
                       int32_t n, int64_t nnz, int32_t flags, cent_graph** out);
int  cent_graph_upload(cent_ctx* ctx, const struct CSR* csr, cent_graph** out);
int  cent_graph_from_csr(cent_ctx* ctx, const int32_t* row_ptr, const int32_t* col,
                         int32_t n, int64_t nnz, cent_graph** out);
void cent_graph_free(cent_graph* g);

/*
 * Brandes betweenness from the k sources src[0..k) (k == -1: every vertex),
 * computed 64 sources per wavefront in FP64.  bc (length n) is ACCUMULATED
 * (+=) like computeBC (Seq_Brandes/Brandes.c:224, caller-zeroed output).
 * Optional outputs, may be null:
 *   farness[i]  = sum of BFS distances from src[i] over its component
 *                 (assigned, int64)                 — Seq_CC/CC.cpp:94
 *   dist_sum[v] += sum over the sources of d(src, v) (int64, push form)
 * Path counts are kept in FP64 with a power-of-two scale per BFS level and
 * source, so counts beyond 2^1024 (the reference's FP32 counts overflow
 * silently past 2^128, Seq_Brandes/Brandes.c:226-227) still give Brandes'
 * values; CENT_E_OVERFLOW remains only for counts within one level that
 * differ by more than a factor 2^1278.
 */
int  cent_bc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
             double* bc, int64_t* farness, int64_t* dist_sum, const cent_opts* opts);

/*
 * Closeness by bit-parallel multi-source BFS (64 sources per uint64 word):
 * farness[i] assigned for every source (k == -1: all vertices, indexed by
 * vertex), dist_sum accumulated, both int64.  Replaces computeCC
 * (Seq_CC/CC.cpp:37-117).
 */
int  cent_cc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
             int64_t* farness, int64_t* dist_sum, const cent_opts* opts);

/*
 * The same two calls with the outputs in device memory of the context's GPU
 * (same meaning, vertex order and accumulate/assign rules as above; src
 * stays a host array).  No counterpart in the reference, whose results are
 * host arrays: for callers that keep the scores on the GPU, so a call costs
 * no PCIe copy of n scores.  Each output must lie inside one device
 * allocation made by this process's HIP runtime (hipMalloc, or
 * cent_device_alloc below) on that GPU (checked: CENT_E_INVAL otherwise).
 * Returns once the outputs are written.
 */
int  cent_bc_device(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                    double* d_bc, int64_t* d_farness, int64_t* d_dist_sum, const cent_opts* opts);
int  cent_cc_device(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                    int64_t* d_farness, int64_t* d_dist_sum, const cent_opts* opts);

/* Allocation-free closeness (no reference counterpart).  For callers that
 * cannot allocate device or pinned memory during the call — code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation.
 * Same arguments, meaning and output rules as cent_cc_device (k == -1:
 * every vertex), run on the
 * source-parallel closeness step; all its device buffers live in one fixed
 * device array of 64 MiB per GPU (held from the first call until the
 * process ends), so a call allocates nothing.  A graph whose n vertices and
 * k sources do not fit (about 4.1 n bytes per traversal slot, 12 n with
 * distance sums on a relabelled graph, plus 12 k) returns CENT_E_INVAL with
 * the sizes in cent_ctx_last_error.  All calls on one GPU share its array:
 * calls that overlap on one GPU race, so serialise them.  The context's
 * statistics are not updated.  Returns once the outputs are written. */
int  cent_cc_device_noalloc(cent_ctx* ctx, const cent_graph* g, const int32_t* src, int32_t k,
                            int64_t* d_farness, int64_t* d_dist_sum);

/* Device arrays on the context's GPU, for callers with no HIP allocator of
 * their own (FFI and Python callers; no reference counterpart).  Copies and
 * fills run on the context's stream and return when done. */
enum { CENT_COPY_TO_DEVICE = 1, CENT_COPY_TO_HOST = 2, CENT_COPY_ON_DEVICE = 3 };
int  cent_device_alloc(cent_ctx* ctx, int64_t bytes, void** out);
int  cent_device_free(cent_ctx* ctx, void* p);
int  cent_device_copy(cent_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t kind);
int  cent_device_fill(cent_ctx* ctx, void* p, int32_t byte, int64_t bytes);

/* ------------------------------------------------ result files */
/* The reference's result formats.
 * BC scores (Seq_Brandes/Brandes.c:2313-2322): "start end" then one "%f"
 * line per vertex start..end.  read fills out[start..end] (cap >= end + 1).
 * Closeness answers (Seq_CC/CC.cpp:2621-2684): "<dataset>_ans.txt" next to a
 * ".mtx" dataset, one "id cc" line per vertex; read assigns ans[id]. */
int  cent_write_bc_scores(const char* path, const struct CSR* csr, const double* bc);
/* The input side of the same formats: csr as a reference edge-list file
 * ("n m", then "u v" once per undirected edge). */
int  cent_write_edgelist(const char* path, const struct CSR* csr);
int  cent_read_bc_scores(const char* path, int32_t* start, int32_t* end, double* out, int32_t cap);
int  cent_ans_path(const char* dataset_path, char* out, size_t cap);
int  cent_read_cc_answers(const char* path, const struct CSR* csr, int64_t* ans);
int  cent_write_cc_answers(const char* path, const struct CSR* csr, const int64_t* cc);

/* ------------------------------------------------ autotuner */
/*
 * Launch settings of the plain form, per graph (no reference counterpart:
 * the reference runs one fixed schedule).  The level kernels split rows
 * longer than hub_deg into hub_deg-edge segments, group the other rows into
 * chunks of at most 64 vertices and chunk_edges edges, and cap a dense level
 * launch at grid_waves wavefronts; frontier picks the dense (pull over every
 * vertex) or the frontier-queue (expand from the frontier) level step.
 */
typedef struct cent_tune_params {
    int32_t hub_deg;      /* 32 .. 2^20                                      */
    int32_t chunk_edges;  /* 64 .. 2^22                                      */
    int32_t grid_waves;   /* 0 = engine default (2048 workgroups of 4 waves);
                             source-parallel step: grid_waves / 8 slots   */
    int32_t frontier;     /* CENT_FRONTIER_*; AUTO = upload-time heuristic   */
    int32_t push_pct;     /* DENSE betweenness: level-2 push threshold, as
                             cent_opts.push_pct (0 = engine default)        */
    double  ms;           /* measured trial time (cent_tune), else 0         */
} cent_tune_params;

/* Graph signature the tuning table is keyed on: log2 vertices, 2 x log2
 * average degree, log2 (max degree / average degree), log2 (BFS depth + 1),
 * each rounded. */
typedef struct cent_tune_sig {
    int32_t lg_n, deg_x2, skew, lg_ecc;
} cent_tune_sig;

#define CENT_TUNE_BC 1   /* trials time Brandes BC (+ farness)            */
#define CENT_TUNE_CC 2   /* trials time closeness only                    */

/* Current settings of an uploaded graph for the call type `what`
 * (CENT_TUNE_BC or CENT_TUNE_CC) / new settings.  hub_deg and chunk_edges
 * are shared by both call types (the chunk and hub tables are rebuilt on
 * the device when they change); frontier and grid_waves are set for each
 * call type named in `what` (CENT_TUNE_BC | CENT_TUNE_CC: both).  Not to be
 * called while calls on g run. */
int  cent_graph_get_layout(const cent_graph* g, int32_t what, cent_tune_params* out);
int  cent_graph_set_layout(cent_graph* g, int32_t what, const cent_tune_params* p);
int  cent_graph_signature(const cent_graph* g, cent_tune_sig* out);

/* Times the plain call `what` on the sample sources src[0..k) (k <= 256
 * for betweenness, k <= 1024 for closeness; use the maximum where the graph
 * allows, so the level steps are compared at full width: four 64-lane
 * batches, or two launch groups of eight for closeness) for candidate settings by coordinate descent (level step, then
 * hub_deg, chunk_edges and grid_waves for the dense step or the slot count
 * for the source-parallel step; at most max_trials timed runs, 0 = no
 * limit), applies the fastest to g for `what`, records it in the
 * process-wide table under (g's signature, what) and returns it in *best.  Results never depend on the
 * settings beyond floating-point summation order. */
int  cent_tune(cent_ctx* ctx, cent_graph* g, const int32_t* src, int32_t k, int32_t what,
               int32_t max_trials, cent_tune_params* best);

/* Process-wide table (thread-safe), one entry per (signature, call type).
 * cent_graph_create applies, for each call type, the entry nearest to the
 * new graph's signature (exact, else within 2 buckets; the chunk/hub layout
 * comes from the betweenness entry when there is one) unless
 * CENT_GRAPH_NO_TUNE is given.  load merges a text file written by save; a
 * malformed file returns CENT_E_IO and changes nothing. */
#define CENT_GRAPH_NO_TUNE 2
int  cent_tune_signature(const int32_t* row_ptr, int32_t n, int64_t nnz, int32_t ecc, cent_tune_sig* out);
int  cent_tune_params_check(const cent_tune_params* p);
int  cent_tune_table_put(const cent_tune_sig* sig, int32_t what, const cent_tune_params* p);
int  cent_tune_table_get(const cent_tune_sig* sig, int32_t what, cent_tune_params* out);
int32_t cent_tune_table_size(void);
void cent_tune_table_clear(void);
int  cent_tune_table_load(const char* path);
int  cent_tune_table_save(const char* path);

/* ------------------------------------------------ reduced (D1 + AP) form */
/*
 * Closeness through degree-1 folding and articulation-point splitting, the
 * reduction pipeline of compute_D1_AP_CC_shareBased_DegreeOrder
 * (Seq_CC/CC.cpp:1728-2052: D1Folding, AP_detection, AP_Copy_And_Split,
 * rebuildGraph, then per-component traversal and D1 recovery).  The host
 * folds leaves, splits the folded graph into biconnected blocks (linear
 * time) and packs every block as its own id range with a copy of each
 * articulation vertex; the GPU computes weighted BFS distance sums for every
 * packed vertex, and the host turns them into the farness of every input
 * vertex.
 */
typedef struct cent_reduced cent_reduced;

typedef struct cent_reduced_view {
    int32_t n;             /* input vertices                                   */
    int32_t np;            /* packed vertices (block copies)                   */
    int64_t nnzp;          /* packed directed edges                            */
    int32_t nblocks;
    int32_t nitems;        /* (block range, <= 64 sources) work items          */
    int32_t nfold;         /* folded degree-1 vertices                         */
    int32_t n_ap;          /* articulation vertices of the folded graph        */
    int32_t ncomp;         /* connected components of the input                */
    const int32_t* prp;    /* [np + 1] packed CSR                              */
    const int32_t* pcol;   /* [nnzp]                                           */
    const int64_t* pw;     /* [np] copy weights                                */
    const int32_t* porig;  /* [np] input vertex of each copy                   */
    const int32_t* boff;   /* [nblocks + 1] packed id range of each block      */
} cent_reduced_view;

int  cent_reduce(const struct CSR* csr, cent_reduced** out);
/* The same reduction on the GPU of ctx: degree-1 folding by rounds, then
 * the biconnected split by Tarjan-Vishkin over a BFS spanning forest, then
 * the packing, all as device kernels; only the result comes back to host
 * memory.  Replaces the same host preprocessing as cent_reduce (D1Folding,
 * Lib/D1Process/D1Process.c:5-154; AP_detection / AP_Copy_And_Split /
 * rebuildGraph, Lib/AP_Process/AP_Process.c:6-1065, called from
 * Seq_CC/CC.cpp:1742-1776).  The result is a cent_reduced that every
 * reduced entry point below takes.  Blocks, copies, weights, fold parents
 * and all results equal cent_reduce's; the order of blocks inside the
 * packed graph, of copies inside a block, and which vertex of a tree
 * component survives folding may differ. */
int  cent_reduce_gpu(cent_ctx* ctx, const struct CSR* csr, cent_reduced** out);
int  cent_reduced_get_view(const cent_reduced* r, cent_reduced_view* out);
/* farness[v] (assigned, length n) from D[x] = sum_y d(x, y) * pw[y] over each
 * packed vertex x's block. */
int  cent_reduced_farness(const cent_reduced* r, const int64_t* D, int64_t* farness);
void cent_reduced_free(cent_reduced* r);
/* bc[v] += betweenness of every input vertex (all sources, ordered pairs,
 * textbook formula) from bcp[x] = weighted Brandes dependencies at each packed
 * vertex x (source and target weights pw).  No reference counterpart: the
 * reference reduces the graph for closeness only. */
int  cent_reduced_bc(const cent_reduced* r, const double* bcp, double* bc);
/* Sampled sources through the reduction (no reference counterpart): for a
 * source multiset src[0..k) (repeats count), WS[x] (length np) = sources
 * hanging off packed copy x on its side of x's block; with every vertex
 * once it equals pw.  cent_reduced_bc_sources adds into bc[v] the
 * betweenness from those sources (the sum over them of Brandes'
 * dependencies, like computeBC over a source subset) given bcp[x] = weighted
 * Brandes dependencies at x from every copy a with source weight WS[a] and
 * target weights pw.  cent_reduced_farness_sources assigns far[i] = farness
 * of src[i] given D[x] = sum_y d(x, y) pw[y] at the copies that answer for
 * the sources (and at every non-root block's head copy and its parent-side
 * copy).  Multigraph reductions: CENT_E_UNSUPPORTED. */
int  cent_reduced_source_weights(const cent_reduced* r, const int32_t* src, int32_t k, int64_t* WS);
int  cent_reduced_bc_sources(const cent_reduced* r, const int32_t* src, int32_t k, const double* bcp, double* bc);
int  cent_reduced_farness_sources(const cent_reduced* r, const int64_t* D, const int32_t* src, int32_t k,
                                  int64_t* far);
/* Frees the device copy the engine caches on r (re-created on next use). */
void cent_reduced_drop_device(cent_reduced* r);

/* Schedules for cent_cc_reduced. */
#define CENT_SCHED_PERSISTENT 0  /* one persistent launch over the work-item queue */
#define CENT_SCHED_PER_ITEM   1  /* level-synchronous launches per work item       */
/* neighbour sharing (closeness only): one BFS from a source also answers up
 * to 32 of its neighbours by the SI/R rule of
 * compute_D1_AP_CC_shareBased_DegreeOrder (Seq_CC/CC.cpp:1818-2022), one
 * workgroup per source in a persistent queue over the blocks of at most
 * 2048 vertices; larger blocks as in CENT_SCHED_PERSISTENT */
#define CENT_SCHED_SHARED     2

/* farness (length n, assigned) of every input vertex through the reduced
 * form; equals cent_cc over all sources on every graph. */
int  cent_cc_reduced(cent_ctx* ctx, const cent_reduced* r, int64_t* farness, int32_t sched,
                     const cent_opts* opts);

/* Betweenness of every vertex over all sources (textbook formula, ordered
 * pairs; bc accumulated like computeBC) through the reduced form, plus the
 * farness of every vertex (assigned, may be null) from the same pass.
 * Equals cent_bc + cent_cc over all sources on every graph.  No reference
 * counterpart (the reference reduces for closeness only). */
int  cent_bc_reduced(cent_ctx* ctx, const cent_reduced* r, double* bc, int64_t* farness,
                     int32_t sched, const cent_opts* opts);

/* Betweenness from the sources src[0..k) (repeats count; textbook formula,
 * ordered pairs; bc accumulated like computeBC) through the reduced form,
 * plus farness[i] of src[i] (assigned, may be null).  Equals cent_bc on the
 * same sources.  Inside each block only the copies with sources behind
 * them are traversed, so on graphs of many blocks (block trees, road-like
 * graphs) a sampled call costs a few traversals per block; on a graph with
 * one giant block it costs about what cent_bc does.  No reference
 * counterpart.  Multigraph reductions: CENT_E_UNSUPPORTED. */
int  cent_bc_reduced_sources(cent_ctx* ctx, const cent_reduced* r, const int32_t* src, int32_t k, double* bc,
                             int64_t* farness, int32_t sched, const cent_opts* opts);

/* ------------------------------------------------ workspace estimates */
/*
 * Device memory a call will allocate in a context that holds nothing yet
 * (no reference counterpart: for callers that budget device memory, such
 * as the request server).  *bytes is the most the context holds at once
 * during the call (cent_ctx_workspace_bytes afterwards is at most that),
 * replayed from the engine's own allocation rules: buffer headroom, the
 * per-level stores of the dense step and their growth, the peer context of
 * the two-stream form, the per-source slots of the source-parallel step.
 * what = CENT_TUNE_BC (cent_bc and cent_bc_device) or CENT_TUNE_CC
 * (cent_cc, cent_cc_device); k sources, -1 = all; opts as given to the
 * call.  Graph storage is not included.
 *
 * cent_workspace_estimate works from host rows before the graph is
 * uploaded: depth is the BFS depth from the graph's highest-degree vertex
 * (the upload probes the same), -1 if not known (the estimate then covers
 * the dense step up to the depth it is chosen for by default, and the
 * source-parallel step), gflags the cent_graph_create flags (tuning-table
 * entries apply as they will at upload unless CENT_GRAPH_NO_TUNE; with
 * depth -1 the engine defaults are assumed).  cent_graph_workspace_estimate
 * uses an uploaded graph's own layout.  cent_reduced_workspace_estimate
 * covers cent_cc_reduced (what = CC), cent_bc_reduced (BC, k = -1) and
 * cent_bc_reduced_sources (BC, k sources) under schedule sched; for blocks
 * deeper than 112 levels the dense level store may grow beyond it. */
int  cent_workspace_estimate(const int32_t* row_ptr, int32_t n, int64_t nnz, int32_t depth, int32_t gflags,
                             int32_t what, int32_t k, const cent_opts* opts, int64_t* bytes);
int  cent_graph_workspace_estimate(const cent_graph* g, int32_t what, int32_t k, const cent_opts* opts,
                                   int64_t* bytes);
int  cent_reduced_workspace_estimate(const cent_reduced* r, int32_t what, int32_t k, int32_t sched,
                                     const cent_opts* opts, int64_t* bytes);
/* Device scratch cent_reduce_gpu takes for a graph of n vertices and nnz
 * directed edges (kept by the context afterwards, until cent_ctx_trim). */
int  cent_reduce_gpu_workspace_estimate(int32_t n, int64_t nnz, int64_t* bytes);

/* ---------------------------------------------------------------- serving */
/*
 * Query server: BC / closeness requests on registered graphs, continuously
 * batched on one GPU.  `slots` requests run at once, each on its own context
 * and HIP stream (the HIP runtime spreads a process's streams over
 * GPU_MAX_HW_QUEUES hardware queues, 4 unless the environment sets more, and
 * streams sharing a queue run one after another: set it to at least
 * `slots` before the process's first HIP call); the FIFO queue admits the
 * next request as soon as a slot is
 * free and its estimated device footprint fits the HBM budget next to the
 * admitted requests and cached graphs (least-recently-used unreferenced
 * graphs are evicted to make room).  A request's result does not depend on
 * what runs beside it.  (No reference counterpart: the reference has no
 * serving path.)
 */
typedef struct cent_server cent_server;

#define CENT_SERVER_DRYRUN 1   /* admission/scheduling only: requests sleep for a
                                  modelled time, nothing runs on a device */

typedef struct cent_server_opts {
    int32_t device;
    int32_t slots;        /* concurrent requests (1..64)                      */
    int64_t hbm_budget;   /* bytes for request workspaces + cached graphs     */
    int32_t max_queue;    /* queued requests before submit returns CENT_E_BUSY */
    int32_t flags;        /* CENT_SERVER_*                                    */
} cent_server_opts;

#define CENT_Q_BC 0       /* Brandes BC (+ farness of the sources)            */
#define CENT_Q_CC 1       /* closeness (farness); k == -1: every vertex,
                             through the D1/AP-reduced path                   */

#define CENT_REQ_QUEUED    0
#define CENT_REQ_RUNNING   1
#define CENT_REQ_DONE      2
#define CENT_REQ_FAILED    3
#define CENT_REQ_CANCELLED 4

typedef struct cent_request_info {
    int32_t state;        /* CENT_REQ_*                                       */
    int32_t status;       /* engine status of a finished request              */
    int32_t slot;
    int32_t reserved;
    double  t_submit_ms, t_start_ms, t_end_ms;   /* server clock               */
    int64_t footprint;    /* admitted workspace bytes                         */
} cent_request_info;

typedef struct cent_server_stats {
    int64_t submitted, completed, failed, cancelled, rejected;
    int64_t queued, running, max_queued, max_running;
    int64_t graph_uploads, graph_hits, graph_evictions;
    int64_t used_workspace, used_graph;
} cent_server_stats;

int  cent_server_create(const cent_server_opts* opts, cent_server** out);
void cent_server_destroy(cent_server* s);
int  cent_server_add_graph(cent_server* s, const struct CSR* csr, int32_t* gid);
/* k > 0: sources src[0..k); k == -1: every vertex. */
int  cent_server_submit(cent_server* s, int32_t gid, int32_t kind, const int32_t* src, int32_t k,
                        int64_t* rid);
/* Waits up to timeout_ms (< 0: no limit) for the request to finish. */
int  cent_server_wait(cent_server* s, int64_t rid, int32_t timeout_ms, cent_request_info* info);
/* Copies the result out (bc: n doubles for CENT_Q_BC; farness: one per source,
 * or n for k == -1) and forgets the request.  Returns the request's status. */
int  cent_server_result(cent_server* s, int64_t rid, double* bc, int64_t* farness,
                        cent_request_info* info);
int  cent_server_cancel(cent_server* s, int64_t rid);
int  cent_server_get_stats(cent_server* s, cent_server_stats* out);

/* ------------------------------------------------------------ test hooks */
/* Failure injection for tests of the engine's error paths (no reference
 * counterpart): the nth host-to-device upload of a graph structure from now
 * on (device CSR, reduced-graph arrays, schedule tables) fails with
 * CENT_E_OOM before it allocates; 0 switches injection off.  Process-wide;
 * not for production use. */
int  cent_debug_fail_upload(int32_t nth);

/* ------------------------------------------- reference-shaped entry points */
/* Drop-in shapes of the reference functions over the own struct CSR, running
 * on a process-wide default context (device 0).  Ids startNodeID..endNodeID
 * are the sources.  BCs is accumulated, CCs is assigned. */
int  computeBC(const struct CSR* csr, double* BCs);                      /* Brandes.c:224 */
int  computeBC_compat(const struct CSR* csr, double* BCs);               /* Brandes.c:317-327 formula */
int  computeCC(const struct CSR* csr, int64_t* CCs);                     /* CC.cpp:37 */
int  compute_D1_AP_CC_shareBased_DegreeOrder(const struct CSR* csr, int64_t* CCs); /* CC.cpp:1728 */

#ifdef __cplusplus
}
#endif
#endif /* CENT_H */
