// Wave64 helpers for gfx950 (CDNA4): lane broadcast, ballots and
// butterfly reductions.  All reductions use a fixed butterfly order, so a
// reduction over the same 64 values always returns the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cent {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int rl32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

__device__ __forceinline__ uint64_t rl64(uint64_t x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= (uint64_t)__shfl_xor((unsigned long long)x, o);
    return x;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += (int64_t)__shfl_xor((long long)x, o);
    return x;
}

// Path-count scaling (betweenness steps): counts at or above 2^kSigBig
// mark their level, whose counts the next level brings back to about
// 2^kSigTo by a power-of-two factor.  A level's sums then stay below
// 2^(kSigBig + 31) (at most 2^31 terms), well inside FP64.
constexpr int kSigBig = 512;
constexpr int kSigTo = 256;

// Unbiased exponent of a positive normal double.
__device__ __forceinline__ int dexp(double x) {
    return (int)((__double_as_longlong(x) >> 52) & 0x7ff) - 1023;
}

}  // namespace cent
