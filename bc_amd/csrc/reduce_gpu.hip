// Graph reduction on the GPU: degree-1 folding and the biconnected-block
// split, producing the same packed block graph as the host build
// (cent_reduced_build: the role of D1Folding, Lib/D1Process/D1Process.c:5-154,
// and AP_detection / AP_Copy_And_Split / rebuildGraph,
// Lib/AP_Process/AP_Process.c:6-225, 468-1065, 1141-1321).
//
// Route, every step linear work in n + m:
//  1. Folding by rounds.  Round r folds every vertex whose remaining degree
//     is 1 into its one unfolded neighbour, with the reference's arithmetic
//     (w[p] += w[d], ff[p] += w[d] + ff[d]); a vertex whose degree drops to 1
//     joins round r + 1.  Two leaves that hold only each other (what is left
//     of a tree) fold the smaller id into the larger.  The fold order is the
//     round order, so recovering farness in reverse order meets every parent
//     before its children.
//  2. Connected components by union-find (roots hooked under smaller roots,
//     so a component's root is its smallest vertex and components are
//     numbered by smallest vertex, as the host numbers them).
//  3. A BFS spanning forest of the folded graph, one root per component
//     (its unfolded vertex of highest remaining degree), level by level.
//  4. Tarjan–Vishkin over that forest: subtree sizes and weights bottom-up,
//     preorder numbers top-down, low/high (extreme preorder reached by a
//     non-tree edge from the subtree) bottom-up, then a union-find over the
//     tree edges: a non-tree edge joins its endpoints' tree edges (in a BFS
//     tree no non-tree edge joins an ancestor to a descendant), and the tree
//     edge into v joins the one into its parent p when the subtree of v
//     reaches outside the subtree of p.  A class of tree edges is a block;
//     its label is the smallest child vertex in it.
//  5. Packing: the vertices of a block form a subtree of the forest, so a
//     block is its members (the child ends of its tree edges) plus its head
//     (the parent of its topmost member).  Blocks are ordered by decreasing
//     head depth, a component's root block after the other blocks at the
//     root (children before parents, which the block-constant passes of
//     cent_reduced_finish rely on); copies inside a block by decreasing block
//     degree, then id; rows sorted.  Radix sorts (hipCUB) order blocks,
//     copies and row entries.
//
// The result lands in a host cent_reduced like the host build's, so every
// consumer (the reduced closeness / BC kernels, cent_reduced_finish,
// cent_reduced_finish_bc, the server) takes either.
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <thread>

#include "engine.h"

namespace {

constexpr int kUnfolded = 0x7fffffff;  // fold round of a vertex that is never folded
constexpr int kG = 16;                 // lanes per vertex in the row-scanning kernels
constexpr int kTB = 256;               // threads per workgroup
constexpr int kLevelBlocks = 512;      // grid of the level kernels whose width the host does not know
constexpr int kRunLevels = 16;         // BFS levels / fold rounds enqueued between host checks
constexpr int kChunk = 256;            // row entries per work item of the row-scanning kernels
constexpr int kHubRow = 1024;          // BFS: rows longer than this are scanned by a whole workgroup

__device__ __forceinline__ int ld_rel(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add64(int64_t* p, int64_t v) {
    atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

// Union-find whose every parent is smaller than its child: a set's root is
// its smallest element, and halving stores only ancestors, so no cycle forms.
__device__ int uf_find(int* p, int x) {
    for (;;) {
        const int px = ld_rel(p + x);
        if (px == x) return x;
        const int gp = ld_rel(p + px);
        if (gp == px) return px;
        st_rel(p + x, gp);
        x = gp;
    }
}

__device__ void uf_union(int* p, int a, int b) {
    for (;;) {
        a = uf_find(p, a);
        b = uf_find(p, b);
        if (a == b) return;
        if (a > b) { const int t = a; a = b; b = t; }
        if (atomicCAS(p + b, b, a) == b) return;
    }
}

__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int gstride() { return gridDim.x * blockDim.x; }

// Per-key aggregation of atomics over a wavefront: the lanes with `ok` are
// peeled one key at a time (the key of the lowest pending lane), and one
// atomic per key carries the whole wave's contribution.  Keys repeat within
// a wave almost always here (one giant component / block, a hub's children),
// where per-lane atomics on one address serialise.  Every lane of the wave
// must reach the call (ok = false where there is nothing to add).
template <class T>
__device__ __forceinline__ void wave_count(T* p, int key, bool ok) {
    uint64_t pend = ballot(ok);
    while (pend) {
        const int lead = __builtin_ctzll(pend);
        const int k0 = __shfl(key, lead);
        const uint64_t m = ballot(ok && key == k0) & pend;
        if (lane_id() == lead) {
            if constexpr (sizeof(T) == 8) atomicAdd((unsigned long long*)(p + k0), (unsigned long long)__builtin_popcountll(m));
            else atomicAdd(p + k0, (T)__builtin_popcountll(m));
        }
        pend &= ~m;
    }
}

template <class T>
__device__ __forceinline__ void wave_add(T* p, int key, T val, bool ok) {
    uint64_t pend = ballot(ok);
    while (pend) {
        const int lead = __builtin_ctzll(pend);
        const int k0 = __shfl(key, lead);
        const bool mine = ok && key == k0 && ((pend >> lane_id()) & 1);
        const uint64_t m = ballot(mine);
        T sum = mine ? val : T(0);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane_id() == lead) {
            if constexpr (sizeof(T) == 8) atomicAdd((unsigned long long*)(p + k0), (unsigned long long)sum);
            else atomicAdd(p + k0, sum);
        }
        pend &= ~m;
    }
}

template <bool kMin>
__device__ __forceinline__ void wave_minmax(int* p, int key, int val, bool ok) {
    uint64_t pend = ballot(ok);
    while (pend) {
        const int lead = __builtin_ctzll(pend);
        const int k0 = __shfl(key, lead);
        const bool mine = ok && key == k0 && ((pend >> lane_id()) & 1);
        const uint64_t m = ballot(mine);
        int r = mine ? val : (kMin ? INT_MAX : INT_MIN);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) r = kMin ? min(r, __shfl_xor(r, o)) : max(r, __shfl_xor(r, o));
        if (lane_id() == lead) {
            if (kMin) atomicMin(p + k0, r);
            else atomicMax(p + k0, r);
        }
        pend &= ~m;
    }
}

// Appends val for every lane with ok at Q[base + (*cnt)++], one atomic per
// wave.  Lanes that do not reach the call are simply not counted.
__device__ __forceinline__ void wave_append(int* Q, int base, int* cnt, int val, bool ok) {
    const uint64_t m = ballot(ok);
    if (!m) return;
    const int lead = __builtin_ctzll(m);
    int pos = 0;
    if (lane_id() == lead) pos = atomicAdd(cnt, (int)__builtin_popcountll(m));
    pos = __shfl(pos, lead);
    if (ok) Q[base + pos + (int)__builtin_popcountll(m & ((1ull << lane_id()) - 1ull))] = val;
}

// Rows split into work items of at most kChunk entries (every vertex has at
// least one item), so a hub row is spread over many lane groups instead of
// holding one group for its whole length.  ioff[i] = first item of list
// entry i (list == nullptr: entry i is vertex i).
__global__ void k_chunk_count(int cnt, const int* list, const int* rp, int* nck) {
    const int i = gtid();
    if (i > cnt) return;
    if (i == cnt) { nck[i] = 0; return; }
    const int v = list ? list[i] : i;
    const int d = rp[v + 1] - rp[v];
    nck[i] = d > 0 ? (d + kChunk - 1) / kChunk : 1;
}

struct RowItem {
    int v, j0, j1;
    bool first;  // the item that starts the row
};

__device__ __forceinline__ RowItem row_item(int g, const int* ioff, int cnt, const int* list, const int* rp) {
    int lo = 0, hi = cnt;  // ioff[lo] <= g < ioff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ioff[mid] <= g) lo = mid;
        else hi = mid;
    }
    RowItem it;
    it.v = list ? list[lo] : lo;
    const int k = g - ioff[lo];
    it.j0 = rp[it.v] + k * kChunk;
    it.j1 = min(rp[it.v + 1], it.j0 + kChunk);
    it.first = (k == 0);
    return it;
}

// ------------------------------------------------------------ folding
__global__ void k_d1_init(const int* rp, int n, int* rdeg, int* fr, int* par, int64_t* w, int64_t* ff,
                          int* inq, int* Q, int* qc) {
    const int v = gtid();
    int d = 0;
    if (v < n) {
        d = rp[v + 1] - rp[v];
        rdeg[v] = d;
        fr[v] = kUnfolded;
        par[v] = -1;
        w[v] = 1;
        ff[v] = 0;
        inq[v] = d == 1 ? 0 : -1;
    }
    wave_append(Q, 0, qc, v, v < n && d == 1);
}

// Round r over the leaves Qin[0 .. qc[r % 3]); new leaves go to Qout with
// count qc[(r + 1) % 3]; qc[(r + 2) % 3] (the previous round's input count,
// read by nobody now) is cleared for the next round's output.
__global__ void k_d1_round(int r, const int* rp, const int* col, const int* Qin, int* Qout, int* qc, int* rdeg,
                           int* fr, int* par, int64_t* w, int64_t* ff, int* inq, int* order, int* ordc) {
    const int cnt = qc[r % 3];
    if (gtid() == 0) qc[(r + 2) % 3] = 0;
    for (int i = gtid(); i < cnt; i += gstride()) {
        const int d = Qin[i];
        if (ld_rel(rdeg + d) != 1) continue;  // lost its last neighbour: a tree's survivor
        // the one neighbour not folded before this round
        int p = -1;
        for (int j = rp[d]; j < rp[d + 1]; ++j) {
            const int y = col[j];
            if (ld_rel(fr + y) >= r) { p = y; break; }
        }
        if (p < 0) continue;
        if (inq[p] == r && p < d) continue;  // p is a leaf of this round too and folds into d
        st_rel(fr + d, r);
        par[d] = p;
        st_rel(rdeg + d, 0);
        const int64_t wd = w[d];
        add64(w + p, wd);
        add64(ff + p, wd + ff[d]);
        order[atomicAdd(ordc, 1)] = d;
        if (atomicSub(rdeg + p, 1) == 2) {
            inq[p] = r + 1;
            Qout[atomicAdd(qc + (r + 1) % 3, 1)] = p;
        }
    }
}

// ------------------------------------------------------------ components
__global__ void k_iota(int n, int* a) {
    const int v = gtid();
    if (v < n) a[v] = v;
}

__global__ void k_fill_i32(int n, int* a, int val) {
    const int v = gtid();
    if (v < n) a[v] = val;
}

// one kG-lane group per row item; each undirected edge once (y > v)
__global__ void k_cc_union(int ni, const int* ioff, const int* rp, const int* col, int n, int* uf) {
    const int g = gtid() / kG, lane = threadIdx.x & (kG - 1);
    if (g >= ni) return;
    const RowItem it = row_item(g, ioff, n, nullptr, rp);
    for (int j = it.j0 + lane; j < it.j1; j += kG) {
        const int y = col[j];
        if (y > it.v) uf_union(uf, it.v, y);
    }
}

// root of every vertex into root[] (not back into uf: another thread's
// halving store may still land on uf[v] after v's own find, leaving a
// non-root ancestor there)
__global__ void k_uf_flat(int n, int* uf, int* flag, int* root) {
    const int v = gtid();
    if (v >= n) return;
    const int r = uf_find(uf, v);
    flag[v] = (r == v);
    root[v] = r;
}

// vcomp holds each vertex's root on entry, its component number on exit
__global__ void k_comp_assign(int n, const int* cid, int* vcomp, int64_t* comp_w) {
    const int v = gtid();
    const bool ok = v < n;
    int c = 0;
    if (ok) {
        c = cid[vcomp[v]];
        vcomp[v] = c;
    }
    wave_count(comp_w, c, ok);
}

// ------------------------------------------------------------ BFS forest
// BFS root per component: its unfolded vertex of highest remaining degree
// (smallest id among equals), as key (rdeg << 32 | ~id) under atomicMax.  A
// hub sits near the middle of a power-law graph and a shortcut end near the
// middle of a grid, so the forest is shallow and the per-level passes few;
// the blocks do not depend on the root.
__global__ void k_croot(int n, const int* fr, const int* rdeg, const int* vcomp, unsigned long long* croot) {
    const int v = gtid();
    const bool ok = v < n && fr[v] == kUnfolded && rdeg[v] > 0;
    const int c = ok ? vcomp[v] : 0;
    const unsigned long long key = ok ? ((unsigned long long)(uint32_t)rdeg[v] << 32) | (0xffffffffu - (uint32_t)v) : 0ull;
    uint64_t pend = ballot(ok);
    while (pend) {
        const int lead = __builtin_ctzll(pend);
        const int k0 = __shfl(c, lead);
        const bool mine = ok && c == k0 && ((pend >> lane_id()) & 1);
        const uint64_t m = ballot(mine);
        unsigned long long mx = mine ? key : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long t = __shfl_xor(mx, o);
            mx = t > mx ? t : mx;
        }
        if (lane_id() == lead) atomicMax(croot + k0, mx);
        pend &= ~m;
    }
}

__global__ void k_bfs_seed(int ncomp, const unsigned long long* croot, int* depth, int* tpar, int* pre, int* cursor,
                           int* Qb, int* lvcnt) {
    const int c = gtid();
    const unsigned long long key = c < ncomp ? croot[c] : 0ull;
    const bool ok = key != 0ull;  // 0: no unfolded vertex of nonzero degree
    const int r = ok ? (int)(0xffffffffu - (uint32_t)key) : 0;
    if (ok) {
        depth[r] = 0;
        tpar[r] = -1;
        pre[r] = 0;
        cursor[r] = 1;
    }
    wave_append(Qb, 0, lvcnt, r, ok);
}

// Level L: the frontier is Qb[loff .. loff + lvcnt[L]), loff = sum of the
// earlier levels' counts (kept in loff[L], written by level L - 1).  Level
// L + 1 is appended after it.  Rows longer than kHubRow are left to
// k_bfs_hubs.
__global__ void k_bfs_level(int L, const int* rp, const int* col, const int* fr, int* Qb, int* loff, int* lvcnt,
                            int* depth, int* tpar) {
    const int lo = loff[L], cnt = lvcnt[L];
    if (gtid() == 0) loff[L + 1] = lo + cnt;
    const int base = lo + cnt;
    const int lane = threadIdx.x & (kG - 1);
    for (int i = gtid() / kG; i < cnt; i += gstride() / kG) {
        const int u = Qb[lo + i];
        const int end = rp[u + 1];
        if (end - rp[u] > kHubRow) continue;
        for (int j = rp[u] + lane; j < end; j += kG) {
            const int y = col[j];
            const bool won = fr[y] == kUnfolded && ld_rel(depth + y) < 0 && atomicCAS(depth + y, -1, L + 1) == -1;
            if (won) tpar[y] = u;
            wave_append(Qb, base, lvcnt + L + 1, y, won);
        }
    }
}

// the long rows of level L's frontier: one workgroup per hub of the graph,
// idle unless the hub is on level L
__global__ void k_bfs_hubs(int L, const int* hubs, const int* rp, const int* col, const int* fr, int* Qb,
                           const int* loff, int* lvcnt, int* depth, int* tpar) {
    const int u = hubs[blockIdx.x];
    if (depth[u] != L) return;
    const int base = loff[L] + lvcnt[L];
    const int end = rp[u + 1];
    for (int j = rp[u] + (int)threadIdx.x; j < end; j += (int)blockDim.x) {
        const int y = col[j];
        const bool won = fr[y] == kUnfolded && ld_rel(depth + y) < 0 && atomicCAS(depth + y, -1, L + 1) == -1;
        if (won) tpar[y] = u;
        wave_append(Qb, base, lvcnt + L + 1, y, won);
    }
}

__global__ void k_hub_flag(int n, const int* rp, int* flag) {
    const int v = gtid();
    if (v < n) flag[v] = rp[v + 1] - rp[v] > kHubRow;
    else if (v == n) flag[v] = 0;
}

__global__ void k_hub_list(int n, const int* flag, const int* off, int* hubs) {
    const int v = gtid();
    if (v < n && flag[v]) hubs[off[v]] = v;
}

// ------------------------------------------------------------ Tarjan–Vishkin
__global__ void k_tv_init(int total, const int* Qb, const int64_t* w, int* size, int64_t* wsub) {
    const int i = gtid();
    if (i >= total) return;
    const int v = Qb[i];
    size[v] = 1;
    wsub[v] = w[v];
}

// children of level L + 1 are final: add them into their parents
__global__ void k_up_size(int lo, int cnt, const int* Qb, const int* tpar, int* size, int64_t* wsub) {
    const int i = gtid();
    const bool ok = i < cnt;
    int p = 0, sz = 0;
    int64_t ws = 0;
    if (ok) {
        const int v = Qb[lo + i];
        p = tpar[v];
        sz = size[v];
        ws = wsub[v];
    }
    wave_add(size, p, sz, ok);
    wave_add(wsub, p, ws, ok);
}

// preorder of level L (>= 1): each vertex takes the next free range of its
// parent's, through the parent's cursor (one atomic per parent per wave), so
// every subtree is one interval [pre, pre + size); children order is free
__global__ void k_pre_level(int lo, int cnt, const int* Qb, const int* tpar, const int* size, int* pre,
                            int* cursor) {
    const int i = gtid();
    const bool ok = i < cnt;
    int v = 0, p = 0, sz = 0;
    if (ok) {
        v = Qb[lo + i];
        p = tpar[v];
        sz = size[v];
    }
    uint64_t pend = ballot(ok);
    while (pend) {
        const int lead = __builtin_ctzll(pend);
        const int k0 = __shfl(p, lead);
        const bool mine = ok && p == k0 && ((pend >> lane_id()) & 1);
        const uint64_t m = ballot(mine);
        const int val = mine ? sz : 0;
        int incl = val;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (lane_id() >= o) incl += t;
        }
        const int tot = __shfl(incl, kWave - 1);
        int base = 0;
        if (lane_id() == lead) base = atomicAdd(cursor + k0, tot);
        base = __shfl(base, lead);
        if (mine) {
            const int pv = base + incl - val;
            pre[v] = pv;
            cursor[v] = pv + 1;
        }
        pend &= ~m;
    }
}

__global__ void k_lohi_init(int total, const int* Qb, const int* pre, int* lo, int* hi) {
    const int i = gtid();
    if (i >= total) return;
    const int x = Qb[i];
    lo[x] = pre[x];
    hi[x] = pre[x];
}

// low/high of a vertex's own non-tree edges, one row item per group
__global__ void k_lohi_local(int ni, const int* ioff, int total, const int* Qb, const int* rp, const int* col,
                             const int* depth, const int* tpar, const int* pre, int* lo, int* hi) {
    const int g = gtid() / kG, lane = threadIdx.x & (kG - 1);
    if (g >= ni) return;
    const RowItem it = row_item(g, ioff, total, Qb, rp);
    const int x = it.v, px = tpar[x];
    int mn = pre[x], mx = pre[x];
    for (int j = it.j0 + lane; j < it.j1; j += kG) {
        const int y = col[j];
        if (depth[y] < 0 || y == px || tpar[y] == x) continue;
        const int py = pre[y];
        mn = min(mn, py);
        mx = max(mx, py);
    }
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o, kG));
        mx = max(mx, __shfl_xor(mx, o, kG));
    }
    if (lane == 0) {
        if (mn < pre[x]) atomicMin(lo + x, mn);
        if (mx > pre[x]) atomicMax(hi + x, mx);
    }
}

__global__ void k_up_lohi(int lo0, int cnt, const int* Qb, const int* tpar, int* lo, int* hi) {
    const int i = gtid();
    const bool ok = i < cnt;
    int p = 0, l = 0, h = 0;
    if (ok) {
        const int v = Qb[lo0 + i];
        p = tpar[v];
        l = lo[v];
        h = hi[v];
    }
    wave_minmax<true>(lo, p, l, ok);
    wave_minmax<false>(hi, p, h, ok);
}

// union-find over tree edges (the edge into x is named x), one row item per group
__global__ void k_tv_union(int ni, const int* ioff, int total, const int* Qb, const int* rp, const int* col,
                           const int* depth, const int* tpar, const int* pre, const int* size, const int* lo,
                           const int* hi, int* au) {
    const int g = gtid() / kG, lane = threadIdx.x & (kG - 1);
    if (g >= ni) return;
    const RowItem it = row_item(g, ioff, total, Qb, rp);
    const int x = it.v, px = tpar[x];
    if (px < 0) return;  // a root: every edge of it is a tree edge
    for (int j = it.j0 + lane; j < it.j1; j += kG) {
        const int y = col[j];
        if (y <= x || depth[y] < 0 || y == px || tpar[y] == x) continue;
        uf_union(au, x, y);
    }
    if (it.first && lane == 0 && tpar[px] >= 0 && (lo[x] < pre[px] || hi[x] >= pre[px] + size[px]))
        uf_union(au, x, px);
}

// blk[x] = block label of the tree edge into x; per block: member count,
// head, and the weight below the head (sub_w); per vertex: the weight of
// the blocks it heads (extra)
__global__ void k_blk_info(int total, const int* Qb, const int* tpar, int* au, int* blk, int* memb,
                           int* head, const int64_t* wsub, int64_t* extra, int64_t* subw) {
    const int i = gtid();
    bool mem = false, top = false;
    int p = 0, B = 0;
    int64_t ws = 0;
    if (i < total) {
        const int x = Qb[i];
        p = tpar[x];
        if (p >= 0) {
            mem = true;
            B = uf_find(au, x);
            blk[x] = B;
            top = tpar[p] < 0 || uf_find(au, p) != B;
            if (top) {
                head[B] = p;
                ws = wsub[x];
            }
        }
    }
    wave_count(memb, B, mem);
    wave_add(subw, B, ws, top);
    wave_add(extra, p, ws, top);
}

__global__ void k_rootblk(int total, const int* Qb, const int* tpar, const int* blk, const int* head,
                          const int* vcomp, int* rootblk, int* nblk) {
    const int i = gtid();
    if (i >= total) return;
    const int x = Qb[i];
    if (tpar[x] < 0 || blk[x] != x) return;
    atomicAdd(nblk, 1);
    const int h = head[x];
    if (tpar[h] < 0) atomicMin(rootblk + vcomp[h], x);
}

// sort key of block B (label): deeper heads first; at a component's root,
// its root block after the root's other blocks; then label
__global__ void k_blk_keys(int n, int maxL, const int* depth, const int* tpar, const int* blk, const int* head,
                           const int* vcomp, const int* rootblk, uint64_t* keys) {
    const int v = gtid();
    if (v >= n) return;
    uint64_t key = ~0ull;
    if (depth[v] > 0 && blk[v] == v) {
        const int h = head[v];
        const int isroot = (tpar[h] < 0 && rootblk[vcomp[h]] == v) ? 1 : 0;
        key = ((uint64_t)(uint32_t)((maxL - depth[h]) * 2 + isroot) << 32) | (uint32_t)v;
    }
    keys[v] = key;
}

__global__ void k_blk_rank(int nblk, const uint64_t* skeys, int* rank, int* blabel, const int* memb, int* bsz) {
    const int b = gtid();
    if (b > nblk) return;
    if (b == nblk) { bsz[b] = 0; return; }
    const int B = (int)(uint32_t)skeys[b];
    rank[B] = b;
    blabel[b] = B;
    bsz[b] = memb[B] + 1;
}

// block of the edge x - y (both unfolded, y a neighbour of x)
__device__ __forceinline__ int edge_block(int x, int y, const int* tpar, const int* blk) {
    return tpar[y] == x ? blk[y] : blk[x];
}

// block degree of every copy: x's member copy (bdm[x]) or the head copy of
// block E (bdh[E]); lanes keep a running count per block to batch atomics
__global__ void k_bdeg(int ni, const int* ioff, int total, const int* Qb, const int* rp, const int* col,
                       const int* depth, const int* tpar, const int* blk, int* bdm, int* bdh) {
    const int g = gtid() / kG, lane = threadIdx.x & (kG - 1);
    if (g >= ni) return;
    const RowItem it = row_item(g, ioff, total, Qb, rp);
    const int x = it.v;
    const int bx = tpar[x] >= 0 ? blk[x] : -1;
    int nm = 0, curE = -1, curc = 0;
    for (int j = it.j0 + lane; j < it.j1; j += kG) {
        const int y = col[j];
        if (depth[y] < 0) continue;
        const int E = edge_block(x, y, tpar, blk);
        if (E == bx) { ++nm; continue; }
        if (E != curE) {
            if (curc) atomicAdd(bdh + curE, curc);
            curE = E;
            curc = 0;
        }
        ++curc;
    }
    if (curc) atomicAdd(bdh + curE, curc);
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) nm += __shfl_xor(nm, o, kG);
    if (lane == 0 && nm) atomicAdd(bdm + x, nm);
}

// copies: members (value x < n) and head copies (value n + block rank);
// key = block rank, then decreasing block degree
__global__ void k_copy_keys(int n, int nblk, const int* depth, const int* blk, const int* rank, const int* bdm,
                            const int* blabel, const int* bdh, uint64_t* keys, int* vals) {
    const int i = gtid();
    if (i >= n + nblk) return;
    uint64_t key = ~0ull;
    int val = -1;
    if (i < n) {
        if (depth[i] > 0) {
            key = ((uint64_t)(uint32_t)rank[blk[i]] << 32) | (uint32_t)(0xffffffffu - (uint32_t)bdm[i]);
            val = i;
        }
    } else {
        const int b = i - n;
        key = ((uint64_t)(uint32_t)b << 32) | (uint32_t)(0xffffffffu - (uint32_t)bdh[blabel[b]]);
        val = n + b;
    }
    keys[i] = key;
    vals[i] = val;
}

__global__ void k_slots(int np, int n, const uint64_t* skeys, const int* svals, const int* blabel, const int* head,
                        const int64_t* w, const int64_t* extra, const int64_t* subw, const int* vcomp,
                        const int64_t* comp_w, int* mslot, int* hslot, int* porig, int* pblock, int* rowlen,
                        int64_t* pw) {
    const int i = gtid();
    if (i > np) return;
    if (i == np) { rowlen[i] = 0; return; }
    const int code = svals[i];
    const int b = (int)(skeys[i] >> 32);
    pblock[i] = b;
    rowlen[i] = (int)(0xffffffffu - (uint32_t)skeys[i]);
    if (code < n) {
        porig[i] = code;
        mslot[code] = i;
        pw[i] = w[code] + extra[code];
    } else {
        const int B = blabel[b], h = head[B];
        porig[i] = h;
        hslot[b] = i;
        pw[i] = comp_w[vcomp[h]] - subw[B];
    }
}

// row entries (a << 32 | c) of every packed edge, a's row at prp[a] + k.
// Row a (a copy of x) gets entries from x's own row only, so only x's group
// writes it; within the group, lanes with the same row take consecutive
// places under one atomic.
__global__ void k_fill_rows(int ni, const int* ioff, int total, const int* Qb, const int* rp, const int* col,
                            const int* depth, const int* tpar, const int* blk, const int* rank, const int* mslot,
                            const int* hslot, const int* prp, int* cur, uint64_t* ent) {
    const int g = gtid() / kG, lane = threadIdx.x & (kG - 1);
    if (g >= ni) return;  // whole groups
    const int gbase = lane_id() & ~(kG - 1);
    const RowItem it = row_item(g, ioff, total, Qb, rp);
    const int x = it.v;
    const int bx = tpar[x] >= 0 ? blk[x] : -1;
    const int end = it.j1;
    for (int j0 = it.j0; j0 < end; j0 += kG) {
        const int j = j0 + lane;
        bool ok = false;
        int a = 0, c = 0;
        if (j < end) {
            const int y = col[j];
            if (depth[y] >= 0) {
                ok = true;
                const int E = edge_block(x, y, tpar, blk);
                const int hs = hslot[rank[E]];
                a = (E == bx) ? mslot[x] : hs;
                c = (tpar[y] >= 0 && blk[y] == E) ? mslot[y] : hs;
            }
        }
        uint32_t pend = (uint32_t)(ballot(ok) >> gbase) & 0xffffu;
        while (pend) {
            const int lead = __builtin_ctz(pend);
            const int a0 = __shfl(a, lead, kG);
            const uint32_t m = (uint32_t)(ballot(ok && a == a0) >> gbase) & pend;
            int base = 0;
            if (lane == lead) base = atomicAdd(cur + a0, __builtin_popcount(m));
            base = __shfl(base, lead, kG);
            if ((m >> lane) & 1u)
                ent[prp[a0] + base + __builtin_popcount(m & ((1u << lane) - 1u))] =
                    ((uint64_t)(uint32_t)a0 << 32) | (uint32_t)c;
            pend &= ~m;
        }
    }
}

__global__ void k_low32(int64_t m, const uint64_t* ent, int* pcol) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) pcol[j] = (int)(uint32_t)ent[j];
}

__global__ void k_blk_meta(int nblk, const int* blabel, const int* head, const int* tpar, const int* blk,
                           const int* rank, const int* rootblk, const int* vcomp, const int* hslot, int* bhead,
                           int* bparent, int* nheads) {
    const int b = gtid();
    if (b >= nblk) return;
    const int B = blabel[b], h = head[B];
    atomicAdd(nheads + h, 1);
    if (tpar[h] >= 0) {
        bparent[b] = rank[blk[h]];
        bhead[b] = hslot[b];
    } else if (rootblk[vcomp[h]] == B) {
        bparent[b] = -1;
        bhead[b] = -1;
    } else {
        bparent[b] = rank[rootblk[vcomp[h]]];
        bhead[b] = hslot[b];
    }
}

__global__ void k_rep(int n, const int* depth, const int* mslot, const int* hslot, const int* rank,
                      const int* rootblk, const int* vcomp, const int* nheads, int* rep, int* nap) {
    const int v = gtid();
    if (v >= n) return;
    const int d = depth[v];
    rep[v] = d > 0 ? mslot[v] : d == 0 ? hslot[rank[rootblk[vcomp[v]]]] : -1;
    const int copies = (d > 0 ? 1 : 0) + (d >= 0 ? nheads[v] : 0);
    if (copies >= 2) atomicAdd(nap, 1);
}

// ------------------------------------------------------------ host side
inline unsigned blocks_for(int64_t threads) {
    return (unsigned)std::max<int64_t>(1, (threads + kTB - 1) / kTB);
}

// Device buffers of one reduction: bump-allocated from the context's arena;
// what does not fit is allocated on its own and freed at the end, and the
// arena grows to the whole request for the next call.
struct Scratch {
    cent_ctx* ctx;
    size_t used = 0, want = 0;
    std::vector<void*> spill;
    explicit Scratch(cent_ctx* c) : ctx(c) {}
    ~Scratch() {
        (void)hipStreamSynchronize(ctx->stream);
        for (void* p : spill) (void)hipFree(p);
        if (want > ctx->red_arena.cap) {
            release(ctx->red_arena);
            (void)ensure_exact(ctx, ctx->red_arena, want);
        }
    }
    template <class T>
    int get(T** out, size_t n) {
        const size_t bytes = ((n > 0 ? n : 1) * sizeof(T) + 255) & ~size_t(255);
        want += bytes;
        if (used + bytes <= ctx->red_arena.cap) {
            *out = (T*)(ctx->red_arena.p + used);
            used += bytes;
            return CENT_OK;
        }
        void* p = nullptr;
        HIPTRY(hipMalloc(&p, bytes));
        spill.push_back(p);
        *out = (T*)p;
        return CENT_OK;
    }
    // sort / scan temporary storage, grown on demand
    char* tmp = nullptr;
    size_t tmp_bytes = 0;
    int tmp_at_least(size_t need) {
        if (need <= tmp_bytes) return CENT_OK;
        TRY(get(&tmp, need));  // a larger block; the smaller one stays unused until the end
        tmp_bytes = need;
        return CENT_OK;
    }
    int scan(const int* in, int* out, int cnt) {
        size_t need = 0;
        HIPTRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, cnt, ctx->stream));
        TRY(tmp_at_least(need));
        HIPTRY(hipcub::DeviceScan::ExclusiveSum(tmp, need, in, out, cnt, ctx->stream));
        return CENT_OK;
    }
    int sort_keys(const uint64_t* in, uint64_t* out, int64_t cnt, int end_bit) {
        size_t need = 0;
        HIPTRY(hipcub::DeviceRadixSort::SortKeys(nullptr, need, in, out, cnt, 0, end_bit, ctx->stream));
        TRY(tmp_at_least(need));
        HIPTRY(hipcub::DeviceRadixSort::SortKeys(tmp, need, in, out, cnt, 0, end_bit, ctx->stream));
        return CENT_OK;
    }
    int sort_pairs(const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, int64_t cnt) {
        size_t need = 0;
        HIPTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, cnt, 0, 64, ctx->stream));
        TRY(tmp_at_least(need));
        HIPTRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, kin, kout, vin, vout, cnt, 0, 64, ctx->stream));
        return CENT_OK;
    }
};

// Results to host memory: every array is copied into the context's pinned
// staging buffer at full link rate, then into its malloc'd destination by
// several host threads (a device-to-pageable copy runs through the
// runtime's small staging chunks at a fraction of the link rate).
struct CopyBack {
    cent_ctx* ctx;
    struct Item {
        void* dst;
        const void* src;
        size_t bytes;
    };
    std::vector<Item> items;
    template <class T>
    int add(T** dst, const T* src, size_t n) {
        *dst = (T*)malloc((n > 0 ? n : 1) * sizeof(T));
        if (!*dst) return CENT_E_OOM;
        if (n) items.push_back({*dst, src, n * sizeof(T)});
        return CENT_OK;
    }
    int flush() {
        size_t total = 0;
        for (const Item& it : items) total += (it.bytes + 255) & ~size_t(255);
        if (total == 0) return CENT_OK;
        TRY(ensure_stage(ctx, total));
        char* stage = (char*)ctx->h_stage;
        size_t off = 0;
        std::vector<size_t> at;
        for (const Item& it : items) {
            HIPTRY(hipMemcpyAsync(stage + off, it.src, it.bytes, hipMemcpyDeviceToHost, ctx->stream));
            at.push_back(off);
            off += (it.bytes + 255) & ~size_t(255);
        }
        HIPTRY(hipStreamSynchronize(ctx->stream));
        // 8 MB pieces shared out to the threads
        constexpr size_t kPiece = size_t(8) << 20;
        struct Piece {
            char* dst;
            const char* src;
            size_t bytes;
        };
        std::vector<Piece> pieces;
        for (size_t i = 0; i < items.size(); ++i)
            for (size_t o = 0; o < items[i].bytes; o += kPiece)
                pieces.push_back({(char*)items[i].dst + o, stage + at[i] + o, std::min(kPiece, items[i].bytes - o)});
        const unsigned nt = (unsigned)std::min<size_t>(pieces.size(), 8);
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k = next++; k < pieces.size(); k = next++) memcpy(pieces[k].dst, pieces[k].src, pieces[k].bytes);
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
        items.clear();
        return CENT_OK;
    }
};

int read_int(cent_ctx* ctx, const int* p, int* out) {
    HIPTRY(hipMemcpyAsync(ctx->h_flags, p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_flags[0];
    return CENT_OK;
}

int bits_for(int64_t v) {
    int b = 1;
    while (b < 63 && (int64_t(1) << b) <= v) ++b;
    return b;
}

int reduce_gpu(cent_ctx* ctx, const struct CSR* g, cent_reduced* r) {
    hipStream_t s = ctx->stream;
    const int n = g->csrVSize - 1;
    const int64_t nnz = g->csrESize;
    Scratch S(ctx);
    CopyBack back{ctx, {}};
    int *rp, *col;
    TRY(S.get(&rp, (size_t)n + 1));
    TRY(S.get(&col, (size_t)nnz));
    HIPTRY(hipMemcpyAsync(rp, g->csrV, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    if (nnz) HIPTRY(hipMemcpyAsync(col, g->csrE, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, s));
    const unsigned gn = blocks_for(n);

    // ---- 1. folding
    int *rdeg, *fr, *par, *inq, *Q0, *Q1, *qc, *order, *ordc;
    int64_t *w, *ff;
    TRY(S.get(&rdeg, n)); TRY(S.get(&fr, n)); TRY(S.get(&par, n)); TRY(S.get(&inq, n));
    TRY(S.get(&Q0, n)); TRY(S.get(&Q1, n)); TRY(S.get(&qc, 4)); TRY(S.get(&order, n)); TRY(S.get(&ordc, 1));
    TRY(S.get(&w, n)); TRY(S.get(&ff, n));
    HIPTRY(hipMemsetAsync(qc, 0, 4 * sizeof(int), s));
    HIPTRY(hipMemsetAsync(ordc, 0, sizeof(int), s));
    k_d1_init<<<gn, kTB, 0, s>>>(rp, n, rdeg, fr, par, w, ff, inq, Q0, qc);
    {
        const unsigned grid = std::min<unsigned>(gn, 1024);
        int rd = 0;
        for (;;) {
            for (int k = 0; k < kRunLevels; ++k, ++rd)
                k_d1_round<<<grid, kTB, 0, s>>>(rd, rp, col, (rd & 1) ? Q1 : Q0, (rd & 1) ? Q0 : Q1, qc, rdeg, fr,
                                                 par, w, ff, inq, order, ordc);
            int left = 0;
            TRY(read_int(ctx, qc + rd % 3, &left));
            if (left == 0) break;
        }
    }
    int nfold = 0;
    TRY(read_int(ctx, ordc, &nfold));

    // ---- 2. components of the input graph
    int *uf, *flag, *cid, *vcomp;
    int64_t* comp_w;
    TRY(S.get(&uf, n)); TRY(S.get(&flag, (size_t)n + 1)); TRY(S.get(&cid, (size_t)n + 1)); TRY(S.get(&vcomp, n));
    TRY(S.get(&comp_w, n));
    // row items over all vertices
    int *nck, *ioff;
    TRY(S.get(&nck, (size_t)n + 1));
    TRY(S.get(&ioff, (size_t)n + 1));
    k_chunk_count<<<blocks_for((int64_t)n + 1), kTB, 0, s>>>(n, nullptr, rp, nck);
    TRY(S.scan(nck, ioff, n + 1));
    int ni_all = 0;
    TRY(read_int(ctx, ioff + n, &ni_all));
    if ((int64_t)ni_all * kG > (int64_t)INT32_MAX) return CENT_E_INVAL;  // int thread ids
    k_iota<<<gn, kTB, 0, s>>>(n, uf);
    k_cc_union<<<blocks_for((int64_t)ni_all * kG), kTB, 0, s>>>(ni_all, ioff, rp, col, n, uf);
    k_uf_flat<<<gn, kTB, 0, s>>>(n, uf, flag, vcomp);
    HIPTRY(hipMemsetAsync(flag + n, 0, sizeof(int), s));
    TRY(S.scan(flag, cid, n + 1));
    int ncomp = 0;
    TRY(read_int(ctx, cid + n, &ncomp));
    HIPTRY(hipMemsetAsync(comp_w, 0, (size_t)n * sizeof(int64_t), s));
    k_comp_assign<<<gn, kTB, 0, s>>>(n, cid, vcomp, comp_w);

    // ---- 3. BFS forest of the folded graph
    unsigned long long* croot;
    int *depth, *tpar, *pre, *cursor, *Qb, *loff, *lvcnt;
    const size_t nlv = (size_t)n + 2 + kRunLevels;
    TRY(S.get(&croot, (size_t)ncomp)); TRY(S.get(&depth, n)); TRY(S.get(&tpar, n)); TRY(S.get(&pre, n));
    TRY(S.get(&cursor, n));
    TRY(S.get(&Qb, n)); TRY(S.get(&loff, nlv)); TRY(S.get(&lvcnt, nlv));
    HIPTRY(hipMemsetAsync(croot, 0, (size_t)ncomp * sizeof(unsigned long long), s));
    k_fill_i32<<<gn, kTB, 0, s>>>(n, depth, -1);
    k_fill_i32<<<gn, kTB, 0, s>>>(n, tpar, -1);
    HIPTRY(hipMemsetAsync(loff, 0, nlv * sizeof(int), s));
    HIPTRY(hipMemsetAsync(lvcnt, 0, nlv * sizeof(int), s));
    k_croot<<<gn, kTB, 0, s>>>(n, fr, rdeg, vcomp, croot);
    k_bfs_seed<<<blocks_for(ncomp), kTB, 0, s>>>(ncomp, croot, depth, tpar, pre, cursor, Qb, lvcnt);
    // hub rows of the BFS (nck / ioff reused: the all-vertex items are done)
    k_hub_flag<<<blocks_for((int64_t)n + 1), kTB, 0, s>>>(n, rp, nck);
    TRY(S.scan(nck, ioff, n + 1));
    int nhub = 0;
    TRY(read_int(ctx, ioff + n, &nhub));
    int* hubs;
    TRY(S.get(&hubs, (size_t)nhub));
    k_hub_list<<<gn, kTB, 0, s>>>(n, nck, ioff, hubs);
    std::vector<int> h_cnt;  // vertices per level
    {
        std::vector<int> chunk(kRunLevels + 1);
        bool done = false;
        for (int L = 0; !done && L <= n; L += kRunLevels) {
            for (int k = 0; k < kRunLevels; ++k) {
                k_bfs_level<<<kLevelBlocks, kTB, 0, s>>>(L + k, rp, col, fr, Qb, loff, lvcnt, depth, tpar);
                if (nhub > 0)
                    k_bfs_hubs<<<nhub, kTB, 0, s>>>(L + k, hubs, rp, col, fr, Qb, loff, lvcnt, depth, tpar);
            }
            HIPTRY(hipMemcpyAsync(chunk.data(), lvcnt + L, (kRunLevels + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
            HIPTRY(hipStreamSynchronize(s));
            // chunk[k] counts level L + k; chunk[kRunLevels] is read again as
            // the next run's first level
            for (int k = 0; k <= kRunLevels && !done; ++k) {
                if (chunk[k] == 0) done = true;
                else if (k < kRunLevels) h_cnt.push_back(chunk[k]);
            }
        }
    }
    const int nlev = (int)h_cnt.size();
    std::vector<int> h_off((size_t)nlev + 1, 0);
    for (int L = 0; L < nlev; ++L) h_off[L + 1] = h_off[L] + h_cnt[L];
    const int total = h_off[nlev];  // vertices in the forest
    const int maxL = nlev > 0 ? nlev - 1 : 0;

    r->n = n;
    r->m = nnz / 2;
    r->ncomp = ncomp;
    r->d1.n = n;
    r->d1.nfold = nfold;
    TRY(back.add(&r->d1.order, order, (size_t)nfold));
    TRY(back.add(&r->d1.parent, par, n));
    TRY(back.add(&r->d1.rdeg, rdeg, n));
    TRY(back.add(&r->d1.w, w, n));
    TRY(back.add(&r->d1.ff, ff, n));
    TRY(back.add(&r->vcomp, vcomp, n));
    TRY(back.add(&r->comp_w, comp_w, (size_t)ncomp));

    if (total == 0) {  // a forest: every component folded to one vertex
        TRY(back.flush());
        r->np = 0;
        r->nnzp = 0;
        r->nblocks = 0;
        r->n_ap = 0;
        r->prp = (int32_t*)calloc(2, sizeof(int32_t));
        r->pcol = (int32_t*)malloc(sizeof(int32_t));
        r->pw = (int64_t*)malloc(sizeof(int64_t));
        r->porig = (int32_t*)malloc(sizeof(int32_t));
        r->pblock = (int32_t*)malloc(sizeof(int32_t));
        r->boff = (int32_t*)calloc(1, sizeof(int32_t));
        r->bhead = (int32_t*)malloc(sizeof(int32_t));
        r->bparent = (int32_t*)malloc(sizeof(int32_t));
        r->rep = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
        if (!r->prp || !r->pcol || !r->pw || !r->porig || !r->pblock || !r->boff || !r->bhead || !r->bparent ||
            !r->rep)
            return CENT_E_OOM;
        for (int v = 0; v < n; ++v) r->rep[v] = -1;
        return CENT_OK;
    }

    // ---- 4. Tarjan–Vishkin
    int *size, *lo, *hi, *au, *blk, *memb, *head, *rootblk, *nblkc;
    int64_t *wsub, *extra, *subw;
    TRY(S.get(&size, n)); TRY(S.get(&lo, n)); TRY(S.get(&hi, n)); TRY(S.get(&au, n)); TRY(S.get(&blk, n));
    TRY(S.get(&memb, n)); TRY(S.get(&head, n)); TRY(S.get(&rootblk, (size_t)ncomp)); TRY(S.get(&nblkc, 1));
    TRY(S.get(&wsub, n)); TRY(S.get(&extra, n)); TRY(S.get(&subw, n));
    const unsigned gt = blocks_for(total);
    // row items over the forest's vertices, in Qb order
    k_chunk_count<<<blocks_for((int64_t)total + 1), kTB, 0, s>>>(total, Qb, rp, nck);
    TRY(S.scan(nck, ioff, total + 1));
    int ni = 0;
    TRY(read_int(ctx, ioff + total, &ni));
    if ((int64_t)ni * kG > (int64_t)INT32_MAX) return CENT_E_INVAL;
    const unsigned gi = blocks_for((int64_t)ni * kG);
    k_tv_init<<<gt, kTB, 0, s>>>(total, Qb, w, size, wsub);
    for (int L = maxL; L >= 1; --L)
        k_up_size<<<blocks_for(h_cnt[L]), kTB, 0, s>>>(h_off[L], h_cnt[L], Qb, tpar, size, wsub);
    for (int L = 1; L <= maxL; ++L)
        k_pre_level<<<blocks_for(h_cnt[L]), kTB, 0, s>>>(h_off[L], h_cnt[L], Qb, tpar, size, pre, cursor);
    k_lohi_init<<<gt, kTB, 0, s>>>(total, Qb, pre, lo, hi);
    k_lohi_local<<<gi, kTB, 0, s>>>(ni, ioff, total, Qb, rp, col, depth, tpar, pre, lo, hi);
    for (int L = maxL; L >= 1; --L)
        k_up_lohi<<<blocks_for(h_cnt[L]), kTB, 0, s>>>(h_off[L], h_cnt[L], Qb, tpar, lo, hi);
    k_iota<<<gn, kTB, 0, s>>>(n, au);
    k_tv_union<<<gi, kTB, 0, s>>>(ni, ioff, total, Qb, rp, col, depth, tpar, pre, size, lo, hi, au);

    // ---- 5. blocks, in children-first order
    HIPTRY(hipMemsetAsync(memb, 0, (size_t)n * sizeof(int), s));
    HIPTRY(hipMemsetAsync(extra, 0, (size_t)n * sizeof(int64_t), s));
    HIPTRY(hipMemsetAsync(subw, 0, (size_t)n * sizeof(int64_t), s));
    HIPTRY(hipMemsetAsync(nblkc, 0, sizeof(int), s));
    k_fill_i32<<<blocks_for(ncomp), kTB, 0, s>>>(ncomp, rootblk, kUnfolded);
    k_blk_info<<<gt, kTB, 0, s>>>(total, Qb, tpar, au, blk, memb, head, wsub, extra, subw);
    k_rootblk<<<gt, kTB, 0, s>>>(total, Qb, tpar, blk, head, vcomp, rootblk, nblkc);
    int nblk = 0;
    TRY(read_int(ctx, nblkc, &nblk));
    uint64_t *bk_in, *bk_out;
    int *rank, *blabel, *bsz, *boff;
    TRY(S.get(&bk_in, n)); TRY(S.get(&bk_out, n)); TRY(S.get(&rank, n));
    TRY(S.get(&blabel, (size_t)nblk + 1)); TRY(S.get(&bsz, (size_t)nblk + 1)); TRY(S.get(&boff, (size_t)nblk + 1));
    k_blk_keys<<<gn, kTB, 0, s>>>(n, maxL, depth, tpar, blk, head, vcomp, rootblk, bk_in);
    TRY(S.sort_keys(bk_in, bk_out, n, 64));
    k_blk_rank<<<blocks_for((int64_t)nblk + 1), kTB, 0, s>>>(nblk, bk_out, rank, blabel, memb, bsz);
    TRY(S.scan(bsz, boff, nblk + 1));

    // ---- copies: members, then head copies; per block by decreasing block degree
    const int np = total - h_cnt[0] + nblk;
    const int64_t ncand = (int64_t)n + nblk;
    int *bdm, *bdh, *cv_in, *cv_out;
    uint64_t *ck_in, *ck_out;
    TRY(S.get(&bdm, n)); TRY(S.get(&bdh, n));
    TRY(S.get(&ck_in, (size_t)ncand)); TRY(S.get(&ck_out, (size_t)ncand));
    TRY(S.get(&cv_in, (size_t)ncand)); TRY(S.get(&cv_out, (size_t)ncand));
    HIPTRY(hipMemsetAsync(bdh, 0, (size_t)n * sizeof(int), s));
    HIPTRY(hipMemsetAsync(bdm, 0, (size_t)n * sizeof(int), s));
    k_bdeg<<<gi, kTB, 0, s>>>(ni, ioff, total, Qb, rp, col, depth, tpar, blk, bdm, bdh);
    k_copy_keys<<<blocks_for(ncand), kTB, 0, s>>>(n, nblk, depth, blk, rank, bdm, blabel, bdh, ck_in, cv_in);
    TRY(S.sort_pairs(ck_in, ck_out, cv_in, cv_out, ncand));
    int *mslot, *hslot, *porig, *pblock, *rowlen, *prp;
    int64_t* pw;
    TRY(S.get(&mslot, n)); TRY(S.get(&hslot, (size_t)nblk)); TRY(S.get(&porig, (size_t)np));
    TRY(S.get(&pblock, (size_t)np)); TRY(S.get(&rowlen, (size_t)np + 1)); TRY(S.get(&prp, (size_t)np + 1));
    TRY(S.get(&pw, (size_t)np));
    k_slots<<<blocks_for((int64_t)np + 1), kTB, 0, s>>>(np, n, ck_out, cv_out, blabel, head, w, extra, subw, vcomp,
                                                         comp_w, mslot, hslot, porig, pblock, rowlen, pw);
    TRY(S.scan(rowlen, prp, np + 1));
    int nnzp = 0;
    TRY(read_int(ctx, prp + np, &nnzp));

    // ---- rows, sorted
    int *cur, *pcol;
    uint64_t *ent, *ent_s;
    TRY(S.get(&cur, (size_t)np)); TRY(S.get(&ent, (size_t)nnzp)); TRY(S.get(&ent_s, (size_t)nnzp));
    TRY(S.get(&pcol, (size_t)nnzp));
    HIPTRY(hipMemsetAsync(cur, 0, (size_t)np * sizeof(int), s));
    k_fill_rows<<<gi, kTB, 0, s>>>(ni, ioff, total, Qb, rp, col, depth, tpar, blk, rank, mslot, hslot, prp, cur,
                                   ent);
    if (nnzp > 0) {
        TRY(S.sort_keys(ent, ent_s, nnzp, 32 + bits_for(np)));
        k_low32<<<blocks_for(nnzp), kTB, 0, s>>>(nnzp, ent_s, pcol);
    }

    // ---- block-cut tree links, representatives, articulation count
    int *bhead, *bparent, *nheads, *rep, *nap;
    TRY(S.get(&bhead, (size_t)nblk)); TRY(S.get(&bparent, (size_t)nblk)); TRY(S.get(&nheads, n));
    TRY(S.get(&rep, n)); TRY(S.get(&nap, 1));
    HIPTRY(hipMemsetAsync(nheads, 0, (size_t)n * sizeof(int), s));
    HIPTRY(hipMemsetAsync(nap, 0, sizeof(int), s));
    k_blk_meta<<<blocks_for(nblk), kTB, 0, s>>>(nblk, blabel, head, tpar, blk, rank, rootblk, vcomp, hslot, bhead,
                                                bparent, nheads);
    k_rep<<<gn, kTB, 0, s>>>(n, depth, mslot, hslot, rank, rootblk, vcomp, nheads, rep, nap);
    int n_ap = 0;
    TRY(read_int(ctx, nap, &n_ap));

    r->np = np;
    r->nnzp = nnzp;
    r->nblocks = nblk;
    r->n_ap = n_ap;
    TRY(back.add(&r->prp, prp, (size_t)np + 1));
    TRY(back.add(&r->pcol, pcol, (size_t)nnzp));
    TRY(back.add(&r->pw, pw, (size_t)np));
    TRY(back.add(&r->porig, porig, (size_t)np));
    TRY(back.add(&r->pblock, pblock, (size_t)np));
    TRY(back.add(&r->boff, boff, (size_t)nblk + 1));
    TRY(back.add(&r->bhead, bhead, (size_t)nblk));
    TRY(back.add(&r->bparent, bparent, (size_t)nblk));
    TRY(back.add(&r->rep, rep, n));
    TRY(back.flush());
    return CENT_OK;
}

}  // namespace

extern "C" int cent_reduce_gpu(cent_ctx* ctx, const struct CSR* csr, cent_reduced** out) {
    if (!ctx || !csr || !out) return CENT_E_INVAL;
    *out = nullptr;
    if (csr->csrVSize < 1 || !csr->csrV || csr->csrESize < 0 || (csr->csrESize > 0 && !csr->csrE) ||
        csr->csrESize > (int64_t)INT32_MAX || (int64_t)csr->csrVSize * kG > (int64_t)INT32_MAX)
        return CENT_E_INVAL;  // the row-scanning kernels index kG-lane groups with int thread ids
    TRY(check_device(ctx));
    const int multi = cent_csr_is_multigraph(csr);
    if (multi < 0) return CENT_E_OOM;
    struct CSR* simple = nullptr;  // distances (hence farness) ignore parallel edges and self-loops
    if (multi) TRY(cent_csr_simple(csr, &simple));
    cent_reduced* r = (cent_reduced*)calloc(1, sizeof(cent_reduced));
    if (!r) {
        cent_csr_free(simple);
        return CENT_E_OOM;
    }
    int st = reduce_gpu(ctx, simple ? simple : csr, r);
    cent_csr_free(simple);
    if (!st) st = cent_reduced_items(r);
    if (!st) st = cent_reduced_tables(r);
    if (st) {
        cent_reduced_free(r);
        return st;
    }
    r->multi = multi;
    *out = r;
    return CENT_OK;
}

// The reduction's scratch (the context's arena afterwards): about 45
// per-vertex arrays of at most 8 bytes, the copies' and candidates' arrays
// (at most 2n entries each), the packed rows' sort keys (two 8-byte keys an
// edge), the rows themselves and the sorts' temporary storage.
extern "C" int cent_reduce_gpu_workspace_estimate(int32_t n, int64_t nnz, int64_t* bytes) {
    if (!bytes || n < 0 || nnz < 0) return CENT_E_INVAL;
    *bytes = 640 * (int64_t)n + 48 * nnz + ((int64_t)32 << 20);
    return CENT_OK;
}
