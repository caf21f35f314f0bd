/* Shared helpers of the host library: seeded RNG, 64-bit key radix sort,
 * growable buffers.  Internal header (not part of the C-ABI). */
#ifndef CENT_LIB_COMMON_H
#define CENT_LIB_COMMON_H

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cent.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: splitmix64 seeding + xoshiro256** stream (Blackman & Vigna, 2018,
 *      "Scrambled linear pseudorandom number generators"). */
typedef struct { uint64_t s[4]; } cent_rng;

static inline uint64_t cent_splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static inline void cent_rng_seed(cent_rng* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = cent_splitmix64(&x);
}
static inline uint64_t cent_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t cent_rng_next(cent_rng* r) {
    uint64_t* s = r->s;
    uint64_t out = cent_rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = cent_rotl(s[3], 45);
    return out;
}
/* uniform double in [0, 1) with 53 random bits */
static inline double cent_rng_uniform(cent_rng* r) {
    return (double)(cent_rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}
/* uniform integer in [0, bound), bound > 0 (multiply-high mapping) */
static inline uint64_t cent_rng_below(cent_rng* r, uint64_t bound) {
    return (uint64_t)(((unsigned __int128)cent_rng_next(r) * bound) >> 64);
}

/* ---- sort + unique of 64-bit keys (LSD radix, 8-bit digits, passes whose
 *      digit is constant are skipped).  Returns the number of unique keys. */
int64_t cent_sort_unique_u64(uint64_t* keys, int64_t n);
/* The same sort keeping repeats; returns n (-1: out of memory). */
int64_t cent_sort_u64(uint64_t* keys, int64_t n);

/* ---- growable int32 vector */
typedef struct { int32_t* a; int64_t n, cap; } cent_ivec;
int  cent_ivec_push(cent_ivec* v, int32_t x);
/* cent_ivec_push with the in-capacity case inlined (hot loops). */
static inline int cent_ivec_put(cent_ivec* v, int32_t x) {
    if (v->n < v->cap) { v->a[v->n++] = x; return CENT_OK; }
    return cent_ivec_push(v, x);
}
void cent_ivec_free(cent_ivec* v);

/* Connected-component labels (BFS), returns number of components; comp[v]
 * in [0, ncomp), components numbered in order of their smallest id. */
int32_t cent_components(const struct CSR* g, int32_t* comp, int64_t* comp_size);

/* CSR from sorted unique directed keys (u<<32|v), n vertices. */
int cent_csr_from_sorted_keys(int32_t n, const uint64_t* keys, int64_t nk, struct CSR** out);

/* Keep only vertices with keep[v] != 0, relabelled in id order. */
int cent_csr_induced(const struct CSR* g, const uint8_t* keep, struct CSR** out, int32_t** old_id);

struct CSR* cent_csr_alloc(int32_t n, int64_t nnz);

#ifdef __cplusplus
}
#endif
#endif
