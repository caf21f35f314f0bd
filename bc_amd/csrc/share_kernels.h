// Neighbour-sharing closeness: the SI/R rule of
// compute_D1_AP_CC_shareBased_DegreeOrder (Seq_CC/CC.cpp:1818-2022) as a
// persistent work-queue kernel over the small blocks of the packed block
// graph.
//
// A unit is a center s and up to 32 of its neighbours u_0 .. u_{k-1} that
// have no answer yet (k = 0: s alone).  One BFS from s gives, per vertex v:
//   dist[v]  = d(s, v);
//   SI[v]    = bits i with a shortest s-v path through u_i, i.e.
//              d(u_i, v) = dist[v] - 1 (u_i starts with bit i, and a vertex
//              takes the bits of every neighbour one level up);
//   R[v]     = bits i, not in SI[v], with d(u_i, v) = dist[v]: set where a
//              same-level neighbour carries bit i (CC.cpp:1917-1923), then
//              passed down every shortest-path edge (CC.cpp:1935-1949).
// Every other vertex is one step further from u_i than from s, so
//   d(u_i, v) = dist[v] + 1 - 2 [i in SI[v]] - [i in R[v]]      (CC.cpp:1971-1995)
// and with weights W the weighted distance sum of u_i is
//   D(u_i) = sum_v W(v) (dist[v] + 1) - 2 A_i - B_i,
//   A_i = sum of W(v) over v with i in SI[v],  B_i = same over R[v] \ SI[v].
// A workgroup keeps dist, SI, R and the BFS queue of its block in LDS and
// accumulates A and B with LDS atomics, so k + 1 weighted distance sums
// come out of one traversal.
#pragma once
#include "wave.h"

namespace cent {

constexpr int kShareLds = 2048;  // largest block (packed vertices) a unit may lie in
constexpr int kShareBlock = 256;
constexpr int kShareMax = 32;    // neighbours answered per unit

struct ShareArgs {
    const int* __restrict__ rp;
    const int* __restrict__ col;
    const int64_t* __restrict__ weight;  // W per packed vertex
    const int4* __restrict__ units;      // (center, block lo, block hi, first neighbour in nbr)
    const int* __restrict__ ucnt;        // neighbours per unit (0 .. kShareMax)
    const int* __restrict__ nbr;         // neighbour lists, packed ids
    int nunits;
    int* __restrict__ head;              // queue counter (zeroed before launch)
    int64_t* __restrict__ D;             // [np] weighted distance sums (assigned)
};

__global__ __launch_bounds__(kShareBlock) void k_cc_share(ShareArgs a) {
    __shared__ int s_dist[kShareLds];
    __shared__ uint32_t s_si[kShareLds];
    __shared__ uint32_t s_r[kShareLds];
    __shared__ int s_q[kShareLds];      // BFS order, level after level
    __shared__ int s_loff[kShareLds + 1];
    __shared__ unsigned long long s_acc[2 * kShareMax + 2];  // A_i, B_i, base, D(s)
    __shared__ int s_unit, s_tail;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    for (;;) {
        if (tid == 0) s_unit = atomicAdd(a.head, 1);
        __syncthreads();
        const int u = s_unit;
        if (u >= a.nunits) break;
        const int4 un = a.units[u];
        const int lo = un.y, R = un.z - un.y, c0 = un.x - lo, nb = a.ucnt[u];
        for (int i = tid; i < R; i += kShareBlock) {
            s_dist[i] = -1;
            s_si[i] = 0;
            s_r[i] = 0;
        }
        for (int i = tid; i < 2 * kShareMax + 2; i += kShareBlock) s_acc[i] = 0;
        __syncthreads();
        if (tid == 0) {
            s_dist[c0] = 0;
            s_q[0] = c0;
            s_tail = 1;
            s_loff[0] = 0;
        }
        if (tid < nb) s_si[a.nbr[un.w + tid] - lo] = 1u << tid;
        __syncthreads();
        // first traversal: dist, SI, the same-level part of R, D(s)
        int64_t far = 0;
        int qb = 0, qe = 1, d = 0;
        while (qb < qe) {
            for (int k = qb + tid; k < qe; k += kShareBlock) {
                const int v = s_q[k];
                const uint32_t siv = s_si[v];
                uint32_t rv = 0;
                const int je = a.rp[lo + v + 1];
                for (int j = a.rp[lo + v]; j < je; ++j) {
                    const int y = a.col[j] - lo;
                    int dy = s_dist[y];
                    if (dy == -1) {
                        dy = atomicCAS(&s_dist[y], -1, d + 1);
                        if (dy == -1) {
                            s_q[atomicAdd(&s_tail, 1)] = y;
                            far += (int64_t)(d + 1) * a.weight[lo + y];
                            dy = d + 1;
                        }
                    }
                    if (dy == d + 1) atomicOr(&s_si[y], siv);
                    else if (dy == d) rv |= s_si[y] & ~siv;  // SI of level d is final
                }
                if (rv) s_r[v] |= rv;  // v's own thread only, at v's level
            }
            __syncthreads();
            qb = qe;
            qe = s_tail;
            ++d;
            if (tid == 0) s_loff[d] = qb;
            __syncthreads();
        }
        const int nlev = d;  // levels 0 .. nlev-1, level L = s_q[s_loff[L] .. s_loff[L+1])
        if (tid == 0) s_loff[nlev] = qe;
        __syncthreads();
        // second traversal: R down every shortest-path edge, level by level
        for (int L = 0; L + 1 < nlev; ++L) {
            for (int k = s_loff[L] + tid; k < s_loff[L + 1]; k += kShareBlock) {
                const int v = s_q[k];
                const uint32_t rv = s_r[v];
                if (!rv) continue;
                const int je = a.rp[lo + v + 1];
                for (int j = a.rp[lo + v]; j < je; ++j) {
                    const int y = a.col[j] - lo;
                    if (s_dist[y] == L + 1) atomicOr(&s_r[y], rv);
                }
            }
            __syncthreads();
        }
        // sums over the block
        int64_t base = 0;
        for (int v = tid; v < R; v += kShareBlock) {
            const int64_t w = a.weight[lo + v];
            base += w * (int64_t)(s_dist[v] + 1);
            uint32_t si = s_si[v], rr = s_r[v] & ~si;
            while (si) {
                const int i = __builtin_ctz(si);
                si &= si - 1;
                atomicAdd(&s_acc[i], (unsigned long long)w);
            }
            while (rr) {
                const int i = __builtin_ctz(rr);
                rr &= rr - 1;
                atomicAdd(&s_acc[kShareMax + i], (unsigned long long)w);
            }
        }
        base = wave_sum_i64(base);
        far = wave_sum_i64(far);
        if (lane == 0) {
            atomicAdd(&s_acc[2 * kShareMax], (unsigned long long)base);
            atomicAdd(&s_acc[2 * kShareMax + 1], (unsigned long long)far);
        }
        __syncthreads();
        if (tid == 0) a.D[lo + c0] = (int64_t)s_acc[2 * kShareMax + 1];
        if (tid < nb)
            a.D[a.nbr[un.w + tid]] =
                (int64_t)s_acc[2 * kShareMax] - 2 * (int64_t)s_acc[tid] - (int64_t)s_acc[kShareMax + tid];
        __syncthreads();
    }
}

}  // namespace cent
