// wave.hpp -- wave64 primitives for gfx950 (CDNA4).
//
// Every helper here is called by all 64 lanes of a wavefront with uniform
// control flow; results that are the same in every lane are moved to SGPRs
// with readfirstlane so loop bounds and branches stay scalar.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plpk {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// 64-bit lane mask of lanes where p holds.
__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p ? 1 : 0); }

// Number of set bits of m below this lane (mbcnt).
__device__ __forceinline__ int rank_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t uni(uint64_t v) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t uni(int64_t v) { return (int64_t)uni((uint64_t)v); }

// Value of v in lane src (src uniform).
__device__ __forceinline__ int read_lane(int v, int src) {
    return __builtin_amdgcn_readlane(v, src);
}
__device__ __forceinline__ uint32_t read_lane(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}
__device__ __forceinline__ uint64_t read_lane(uint64_t v, int src) {
    return ((uint64_t)read_lane((uint32_t)(v >> 32), src) << 32) | read_lane((uint32_t)v, src);
}

// Wave-wide integer sums (exact; order-independent).
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return uni(v);
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return uni(v);
}

// LDS ordering point between phases of a block (one or a few waves).
__device__ __forceinline__ void block_sync() { __syncthreads(); }

// LDS ordering point between lanes of one wave.  A wave's LDS instructions
// execute in issue order, so a store seen before a load in program order is
// visible to that load; the fences and the wave barrier stop the compiler
// from moving LDS accesses across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of the NW per-wave counts c[0..NW) (16-byte aligned), read in 16-byte
// pieces: one LDS round trip.
template <int NW>
__device__ __forceinline__ int sum_counts(const int* c) {
    if constexpr (NW == 1) {
        return c[0];
    } else if constexpr (NW == 2) {
        const int2 v = *(const int2*)c;
        return v.x + v.y;
    } else {
        int t = 0;
#pragma unroll
        for (int i = 0; i < NW; i += 4) {
            const int4 v = *(const int4*)(c + i);
            t += v.x + v.y + v.z + v.w;
        }
        return t;
    }
}

// Wave-wide min / max through DPP lane moves (VALU latency, no LDS round
// trip): quad swaps, half-row and row mirrors reduce each 16-lane row, then
// row_bcast:15 and row_bcast:31 fold the rows into lane 63.  Lanes a DPP step
// does not write take `old`, the operation's identity (INT_MAX for min,
// INT_MIN for max, 0 for or), so each step folds into one DPP-source VALU op.
template <bool kMax>
__device__ __forceinline__ int wave_reduce_dpp(int v) {
    constexpr int id = kMax ? INT32_MIN : INT32_MAX;
    auto op = [](int a, int b) { return kMax ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min(int v) { return wave_reduce_dpp<false>(v); }
__device__ __forceinline__ int wave_max(int v) { return wave_reduce_dpp<true>(v); }

// Wave-wide bitwise OR, the same DPP pattern (identity 0).
__device__ __forceinline__ uint32_t wave_or(uint32_t u) {
    int v = (int)u;
    v |= __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
    return (uint32_t)__builtin_amdgcn_readlane(v, 63);
}

}  // namespace plpk
