// mt19937.hpp -- CPython random.Random(int) on one wavefront.
//
// Contract (gpu/plp_common.hpp:164-250, CPython Modules/_randommodule.c):
//   seeding   init_genrand(19650218) then init_by_array(key), key = the
//             little-endian 32-bit chunks of the seed, mt[0] = 0x80000000,
//             index 624 (plp_common.hpp:170-232);
//   next()    regenerate the 624-word block when exhausted, then temper
//             (plp_common.hpp:234-249);
//   random()  (a * 2^26 + b) * 2^-53 with a = next() >> 5, b = next() >> 6
//             (plp_common.hpp:182-186);
//   getrandbits(k) = next() >> (32 - k), 1 <= k <= 32;
//   randrange(n)   k = bit_length(n), redraw getrandbits(k) while >= n
//             (plp_common.hpp:189-194).
//
// The 624-word state lives in LDS.  The block regeneration is split into the
// four ranges where the recurrence has no intra-range dependence:
//   [0,227)    reads old mt[i], mt[i+1], mt[i+397]
//   [227,454)  reads new mt[i-227]
//   [454,623)  reads new mt[i-227]
//   623        reads new mt[0] and mt[396]
// so 64 lanes regenerate a block in 4 load/store rounds instead of 624 serial
// steps.  Every lane computes identical draws (broadcast LDS reads), so the
// RNG's data-dependent branches (randrange rejection) stay wave-uniform.
#pragma once

#include "wave.hpp"

namespace plpk {

constexpr uint32_t kMtUpper = 0x80000000u;
constexpr uint32_t kMtLower = 0x7fffffffu;
constexpr uint32_t kMtMatrix = 0x9908b0dfu;

__device__ __forceinline__ uint32_t mt_mix(uint32_t cur, uint32_t nxt, uint32_t far) {
    uint32_t y = (cur & kMtUpper) | (nxt & kMtLower);
    return far ^ (y >> 1) ^ ((y & 1u) ? kMtMatrix : 0u);
}

// init_genrand(19650218): the state every seeding starts from, a constant.
struct MtInitTable {
    uint32_t w[640];   // 624 words, padded to ten 64-lane rows
};

constexpr MtInitTable mt_init_table() {
    MtInitTable t{};
    uint32_t prev = 19650218u;
    t.w[0] = prev;
    for (int i = 1; i < 624; i++) {
        prev = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
        t.w[i] = prev;
    }
    return t;
}

static __constant__ const MtInitTable kMtInit = mt_init_table();

// CPython seeding, all lanes of one wave.  init_by_array's two loops are a
// serial recurrence through the previous word; the state words sit in ten
// registers (lane l of row r holds word 64 r + l), so each step reads its
// word and writes the result with a lane move, not an LDS round trip.  The
// words visited are 1..623, 1 (first loop) and 2..623, 1 (second loop).
// Caller syncs the wave afterwards.
__device__ inline void mt_seed_wave(uint32_t* mt, uint64_t seed) {
    const int lane = lane_id();
    uint32_t v[10];
#pragma unroll
    for (int r = 0; r < 10; r++) v[r] = kMtInit.w[64 * r + lane];
    // key = the 32-bit chunks of the seed (one chunk below 2^32, seed 0 gives
    // [0]); the first loop adds key[j] + j with j cycling over the chunks
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t kodd = k1 ? k1 + 1u : k0;
    uint32_t prev = 19650218u;   // word 0 of the initial state
    int n = 0;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const int hi = r == 9 ? 624 - 576 : kWave;
        for (int q = r == 0 ? 1 : 0; q < hi; q++) {
            const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)v[r], q);
            prev = (x ^ ((prev ^ (prev >> 30)) * 1664525u)) + ((n & 1) ? kodd : k0);
            n++;
            v[r] = lane == q ? prev : v[r];
        }
    }
    {   // 624th step: word 1 again (n = 623, odd)
        const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)v[0], 1);
        prev = (x ^ ((prev ^ (prev >> 30)) * 1664525u)) + kodd;
        v[0] = lane == 1 ? prev : v[0];
    }
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const int hi = r == 9 ? 624 - 576 : kWave;
        for (int q = r == 0 ? 2 : 0; q < hi; q++) {
            const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)v[r], q);
            prev = (x ^ ((prev ^ (prev >> 30)) * 1566083941u)) - (uint32_t)(kWave * r + q);
            v[r] = lane == q ? prev : v[r];
        }
    }
    {   // last step: word 1
        const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)v[0], 1);
        prev = (x ^ ((prev ^ (prev >> 30)) * 1566083941u)) - 1u;
        v[0] = lane == 1 ? prev : v[0];
    }
    v[0] = lane == 0 ? 0x80000000u : v[0];
#pragma unroll
    for (int r = 0; r < 10; r++)
        if (kWave * r + lane < 624) mt[kWave * r + lane] = v[r];
}

// Regenerate the block: all lanes of one wave.
__device__ inline void mt_twist(uint32_t* mt) {
    const int lane = lane_id();
    uint32_t v[4];
    // [0, 227): every input is an old word.
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = lane + r * kWave;
        if (i < 227) v[r] = mt_mix(mt[i], mt[i + 1], mt[i + 397]);
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = lane + r * kWave;
        if (i < 227) mt[i] = v[r];
    }
    wave_sync();
    // [227, 454): mt[i-227] is new.
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = 227 + lane + r * kWave;
        if (i < 454) v[r] = mt_mix(mt[i], mt[i + 1], mt[i - 227]);
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = 227 + lane + r * kWave;
        if (i < 454) mt[i] = v[r];
    }
    wave_sync();
    // [454, 623): mt[i-227] is new.
#pragma unroll
    for (int r = 0; r < 3; r++) {
        int i = 454 + lane + r * kWave;
        if (i < 623) v[r] = mt_mix(mt[i], mt[i + 1], mt[i - 227]);
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < 3; r++) {
        int i = 454 + lane + r * kWave;
        if (i < 623) mt[i] = v[r];
    }
    wave_sync();
    // 623: wraps to the new mt[0].
    uint32_t last = mt_mix(mt[623], mt[0], mt[396]);
    wave_sync();
    if (lane == 0) mt[623] = last;
    wave_sync();
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// One 32-bit output; mti is wave-uniform state held by the caller.
__device__ __forceinline__ uint32_t mt_next(uint32_t* mt, int& mti) {
    if (mti >= 624) {
        mt_twist(mt);
        mti = 0;
    }
    uint32_t y = uni(mt[mti]);
    mti++;
    return mt_temper(y);
}

// The two words the next random() takes when no regeneration falls inside
// it (mti <= 622; otherwise two in-bounds words that go unused), so that a
// caller can read them together with other LDS data.
struct MtPeek {
    uint32_t a, b;
};

__device__ __forceinline__ MtPeek mt_peek2(const uint32_t* mt, int mti) {
    const int i = mti <= 622 ? mti : 622;
    return MtPeek{mt[i], mt[i + 1]};
}

__device__ __forceinline__ double mt_random(uint32_t* mt, int& mti, MtPeek pk) {
    uint32_t a, b;
    if (mti + 2 <= 624) {   // both words without a regeneration
        mti += 2;
        a = mt_temper(uni(pk.a)) >> 5;
        b = mt_temper(uni(pk.b)) >> 6;
    } else {
        a = mt_next(mt, mti) >> 5;
        b = mt_next(mt, mti) >> 6;
    }
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double mt_random(uint32_t* mt, int& mti) {
    return mt_random(mt, mti, mt_peek2(mt, mti));
}

__device__ __forceinline__ uint32_t mt_getrandbits(uint32_t* mt, int& mti, int k) {
    return mt_next(mt, mti) >> (32 - k);
}

__device__ __forceinline__ int mt_randrange(uint32_t* mt, int& mti, int n) {
    int k = 32 - __clz((unsigned)n);
    int r = (int)mt_getrandbits(mt, mti, k);
    while (r >= n) r = (int)mt_getrandbits(mt, mti, k);
    return r;
}

}  // namespace plpk
