// decoder_dev.hpp -- device building blocks shared by the item decoder
// (kernels.hip) and the block-building decoder (block_kernel.hip): the LDS
// box record, packed extreme-point keys, the feasibility scan, FSR, the
// extreme-point merge and packing output.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.hpp"
#include "wave.hpp"

namespace plpk {

// Placed box in LDS: AABB + item type (32 bytes; the hot loop reads 24).
struct Box {
    int32_t x, y, z, x2, y2, z2, type, pad;
};

// Widths of the per-decoder state.  An extreme point is a packed key
// z << (xb + yb) | y << xb | x (xb, yb the bit lengths of L and W), whose
// numeric order is the (z, y, x) order of the reference's candidate list; a
// pair's state is its support sum.  kNarrow and kWide hold keys and sums in
// 32 bits (bitlen(L) + bitlen(W) + bitlen(H) <= 31, so L * W < 2^30) and the
// pair-state cache in 16 or 32 bits; kBig holds all three in 64 bits, for
// every container side up to 2^21 - 1 (the reference device form's range,
// gpu/plp_gpu.cu:1263-1268).
template <int M>
struct ModeTypes {
    using Key = uint32_t;
    using Sum = uint32_t;
};
template <>
struct ModeTypes<kBig> {
    using Key = uint64_t;
    using Sum = uint64_t;
};

template <class Key>
__device__ __forceinline__ void key_xyz(Key key, int xb, int yb, int& x, int& y, int& z) {
    x = (int)(key & ((Key(1) << xb) - 1));
    y = (int)((key >> xb) & ((Key(1) << yb) - 1));
    z = (int)(key >> (xb + yb));
}

template <class Key>
__device__ __forceinline__ Key xyz_key(int x, int y, int z, int xb, int yb) {
    return ((Key)(uint32_t)z << (xb + yb)) | ((Key)(uint32_t)y << xb) | (Key)(uint32_t)x;
}

// check_feasible_one (plp_common.hpp:297-321) for this lane's pair:
//   C2  x+w <= L, y+d <= W, z+h <= H
//   C3  z > 0: s = sum of top-face overlaps of boxes whose top is at z;
//       infeasible iff (double)s / (double)(w*d) < 0.5
//   C1  no strict-interior overlap with any box
// pair_scan takes lanes that passed C2 (alive) and returns C1 together with
// the support sum s, which is complete for every lane that passes C1; the
// caller applies C3 (pair_verdict).  The result does not depend on which
// boxes are visited in what order as long as every box that can overlap or
// support the pair is visited.  The wave walks the placed boxes newest first
// in windows of 64 (one box per lane), keeps the ones that intersect the
// chunk's region [xlo,xhi) x [ylo,yhi) x [zlo,zhi) or whose top lies in
// [zlo, zhi) over that footprint (a box outside it can neither overlap nor
// support any pair of the chunk), copies them into the wave's cbox[] window,
// pads it with sentinels to a multiple of kBoxGroup and scans it in groups
// whose LDS reads issue together, branch-free per box (overlap and support
// both come from the three overlap extents).  The wave leaves as soon as
// every lane has failed.  The support sum is a Sum: 32 bits when L*W < 2^30
// (ox, oy < 2^21, so the 24-bit multiply is exact), else 64 bits.
constexpr int kBoxGroup = 4;

// a * b + c for a, b < 2^24 in one VALU op (the compiler otherwise pairs two
// products into one three-way add, three ops per two boxes instead of two).
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// The same for 64-bit sums (v_mad_u64_u32).
__device__ __forceinline__ uint32_t mad_sum(uint32_t a, uint32_t b, uint32_t c) {
    return mad_u24(a, b, c);
}
__device__ __forceinline__ uint64_t mad_sum(uint32_t a, uint32_t b, uint64_t c) {
    return (uint64_t)a * b + c;
}
constexpr int kCullWindow = kWave + kBoxGroup;   // cbox[] entries per wave

__device__ __forceinline__ Box sentinel_box() {
    Box b;
    b.x = b.y = b.z = b.x2 = b.y2 = b.z2 = -1;   // never overlaps, never supports
    b.type = 0;
    b.pad = 0;
    return b;
}

template <class Sum>
__device__ __forceinline__ bool pair_scan(int x, int y, int z, int w, int d, int h, bool alive,
                                          const Box* __restrict__ boxes, int n_box, int xlo,
                                          int xhi, int ylo, int yhi, int zlo, int zhi,
                                          Box* __restrict__ cbox, Sum& s_out) {
    const int lane = lane_id();
    const int x2 = x + w, y2 = y + d, z2 = z + h;
    Sum s = 0;
    uint64_t am = ballot(alive);   // lanes still passing C1, as a scalar mask
    for (int top = n_box - 1; top >= 0; top -= kWave) {
        if (am == 0) break;
        const int b = top - lane;
        Box C = sentinel_box();
        if (b >= 0) C = boxes[b];
        const bool r = (b >= 0) & (C.x < xhi) & (C.x2 > xlo) & (C.y < yhi) & (C.y2 > ylo) &
                       (C.z < zhi) & (C.z2 >= zlo);
        const uint64_t rm = ballot(r);
        const int n_rel = __popcll(rm);
        if (n_rel == 0) continue;
        if (r) cbox[rank_below(rm)] = C;
        const int n_pad = (n_rel + kBoxGroup - 1) / kBoxGroup * kBoxGroup;
        if (lane < n_pad - n_rel) cbox[n_rel + lane] = sentinel_box();
        wave_sync();
        for (int g = 0; g < n_pad; g += kBoxGroup) {
            Box B[kBoxGroup];
#pragma unroll
            for (int u = 0; u < kBoxGroup; u++) B[u] = cbox[g + u];
#pragma unroll
            for (int u = 0; u < kBoxGroup; u++) {
                // overlap extents: the boxes meet strictly iff all three are
                // positive; a box whose top is at z supports over ox * oy
                // when the footprints meet
                const int ox = min(x2, B[u].x2) - max(x, B[u].x);
                const int oy = min(y2, B[u].y2) - max(y, B[u].y);
                const int oz = min(z2, B[u].z2) - max(z, B[u].z);
                const int oxy = min(ox, oy);
                am &= ~ballot(min(oxy, oz) > 0);
                const bool sup = (B[u].z2 == z) & (oxy > 0);
                s = mad_sum((uint32_t)(sup ? ox : 0), (uint32_t)oy, s);
            }
            if (am == 0) break;
        }
        wave_sync();   // the next window overwrites cbox[]
    }
    s_out = s;
    return (am >> lane) & 1ull;
}

// Division by an orientation count no in 1..6 without an integer divide:
// p / no = umulhi(p, ceil(2^32 / no)), exact while p * (no - 2^32 mod no) < 2^32,
// i.e. for every p below 2^32 / 6; no == 1 bypasses the product.
__device__ __forceinline__ uint32_t orient_recip(int no) {
    return no == 2 ? 0x80000000u : no == 3 ? 0x55555556u : no == 4 ? 0x40000000u
         : no == 5 ? 0x33333334u : 0x2aaaaaabu;
}
__device__ __forceinline__ int div_orients(int p, int no, uint32_t recip) {
    return no == 1 ? p : (int)__umulhi((uint32_t)p, recip);
}

// C3 on top of C1 and C2 (plp_common.hpp:304-312).  The reference rejects
// when the FP64 quotient (double)s / (double)(w * d) < 0.5.  For integers
// 0 <= s <= A = w * d with 1 <= A <= 2^53 (both exact in FP64) that verdict
// equals 2 * s < A: when 2s < A the exact quotient is at most 0.5 - 1/(2A),
// more than half an ulp (2^-55) below 0.5, so the correctly rounded quotient
// stays below 0.5; when 2s >= A it is at least 0.5 and rounds to at least
// 0.5.  Every box that passes C2 has A <= L * W < 2^42 (sides below 2^21),
// so the integer form gives the reference's verdict bit for bit.  s <= A
// (boxes whose top is at z do not overlap one another, so their overlaps
// with the footprint are disjoint parts of it; a clamped cache sum is
// smaller still).  With 32-bit sums L * W < 2^30, so 2s and the product of
// the 24-bit operands fit 32 bits; with 64-bit sums both fit 64 bits.
// Branch-free; when C1 or C2 fails neither the product nor the sum matters.
__device__ __forceinline__ bool pair_verdict(bool c12, int z, int w, int d, uint32_t s) {
    return c12 & ((z == 0) | (2u * s >= (uint32_t)__umul24((uint32_t)w, (uint32_t)d)));
}
__device__ __forceinline__ bool pair_verdict(bool c12, int z, int w, int d, uint64_t s) {
    return c12 & ((z == 0) | (2ull * s >= (uint64_t)(uint32_t)w * (uint32_t)d));
}

// Pair-state cache entry: the support sum of a pair that passes C1 and C2,
// or blocked (all ones).  Both only grow as boxes are added, so the state
// after one more box follows from the old state and that box alone.
template <class Sum>
constexpr Sum kBlockedOf = ~Sum(0);
constexpr uint32_t kBlocked = kBlockedOf<uint32_t>;
constexpr uint16_t kNewPoint = 0xffffu;   // list entry with no cached state
constexpr uint16_t kCorner = 0xfff0u;     // kCorner + j: corner j of the last box (state in cstate)

// Entries are the Sum itself (kWide: 32-bit, kBig: 64-bit), or 16-bit
// (kNarrow) with the support sum clamped to 65534 and 0xffff for blocked
// (exact for C3 when every w*d <= kNarrowMaxArea).
template <int M>
__device__ __forceinline__ typename ModeTypes<M>::Sum cache_get(const void* c, int i) {
    using Sum = typename ModeTypes<M>::Sum;
    if constexpr (M == kNarrow) {
        const uint32_t v = ((const uint16_t*)c)[i];
        return v == 0xffffu ? kBlocked : v;
    } else {
        return ((const Sum*)c)[i];
    }
}

template <int M>
__device__ __forceinline__ void cache_put(void* c, int i, typename ModeTypes<M>::Sum v) {
    using Sum = typename ModeTypes<M>::Sum;
    if constexpr (M == kNarrow) ((uint16_t*)c)[i] = (uint16_t)(v == kBlocked ? 0xffffu : min(v, 0xfffeu));
    else ((Sum*)c)[i] = v;
}


// ---------------------------------------------------------------- FSR ----
__device__ __forceinline__ double read_lane_f64(double v, int src) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The six coordinates of a box read from LDS (one 16-byte and one 8-byte
// read; type and padding stay behind).
__device__ __forceinline__ Box box_xyz(const Box& b) {
    const int4 p = *(const int4*)&b.x;
    const int2 q = *(const int2*)&b.y2;
    Box r;
    r.x = p.x; r.y = p.y; r.z = p.z; r.x2 = p.w; r.y2 = q.x; r.z2 = q.y;
    r.type = 0; r.pad = 0;
    return r;
}

// Contact area of box i's bottom face with the top face of box j (j's top at
// i's bottom z > 0 and the footprints overlapping), else 0.
template <class Sum>
__device__ __forceinline__ Sum contact_area(const Box& i, const Box& j) {
    const int ox = min(i.x2, j.x2) - max(i.x, j.x);
    const int oy = min(i.y2, j.y2) - max(i.y, j.y);
    const bool c = (j.z2 == i.z) & (i.z > 0) & (ox > 0) & (oy > 0);
    return c ? (Sum)(uint32_t)ox * (uint32_t)oy : Sum(0);
}

// Box j precedes box i in the reference's visiting order (z desc, index asc).
__device__ __forceinline__ int visits_before(const Box& j, int jdx, const Box& i, int idx) {
    return ((j.z > i.z) | ((j.z == i.z) & (jdx < idx))) ? 1 : 0;
}

// fsr_score for n <= R * kWave boxes with the loads in registers: lane i
// holds boxes i, i + 64, ..., their loads and contact totals (also written
// to tot[]).  The visiting order comes from each box's rank (the number of
// boxes before it in (z desc, index asc) order, scattered through ord[], n
// ints of LDS); the k loop reads box k's coordinates and total from LDS and
// its load from the lane that holds it.  Box j's load receives additions
// only from the boxes whose bottom is at j's top, each through the register
// of the lane that holds j, in the visiting order: the reference's sequence
// of FP64 operations on every load[j].
template <int R>
__device__ __forceinline__ int read_lane_of(const int (&v)[R], int q, int src) {
#pragma unroll
    for (int i = 0; i < R - 1; i++)
        if (q == i) return __builtin_amdgcn_readlane(v[i], src);
    return __builtin_amdgcn_readlane(v[R - 1], src);
}
template <int R>
__device__ __forceinline__ double read_lane_of(const double (&v)[R], int q, int src) {
#pragma unroll
    for (int i = 0; i < R - 1; i++)
        if (q == i) return read_lane_f64(v[i], src);
    return read_lane_f64(v[R - 1], src);
}

template <int R, class Sum>
static __device__ double fsr_score_reg(const Box* __restrict__ box, int n, double mu, int* ord,
                                       Sum* tot, double* fsr_out) {
    const int lane = lane_id();
    // m = RHO * V / 1e6 (plp_common.hpp:360-362); load starts at m
    auto mass = [](const Box& b) {
        const int64_t v = (int64_t)(b.x2 - b.x) * (b.y2 - b.y) * (b.z2 - b.z);
        return 50.0 * (double)v / 1.0e6;
    };
    bool v[R];
    Box b[R];
    double l[R];
    Sum t[R];
    int r[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        v[q] = lane + q * kWave < n;
        b[q] = v[q] ? box[lane + q * kWave] : sentinel_box();
        l[q] = v[q] ? mass(b[q]) : 1.0;
        t[q] = 0;
        r[q] = 0;
    }
    // contact totals (the boxes whose tops touch a box's bottom do not
    // overlap one another, so a total is at most w * d, which fits a Sum) and
    // ranks
    constexpr int kG = 2;   // boxes whose reads issue together
    for (int j = 0; j < n; j += kG) {
        Box B[kG];
#pragma unroll
        for (int u = 0; u < kG; u++) B[u] = j + u < n ? box_xyz(box[j + u]) : sentinel_box();
#pragma unroll
        for (int u = 0; u < kG; u++) {
#pragma unroll
            for (int q = 0; q < R; q++) {
                t[q] += contact_area<Sum>(b[q], B[u]);
                r[q] += visits_before(B[u], j + u, b[q], lane + q * kWave);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < R; q++) {
        if (v[q]) {
            ord[r[q]] = lane + q * kWave;
            tot[lane + q * kWave] = t[q];
        }
    }
    wave_sync();
    int o[R];
#pragma unroll
    for (int q = 0; q < R; q++) o[q] = v[q] ? ord[lane + q * kWave] : 0;
    // compute_fsr's k loop (plp_common.hpp:365-385).  Box k's coordinates
    // and total come from LDS one source ahead; its load from its lane.
    int kn = __builtin_amdgcn_readlane(o[0], 0);
    Box Bn = box_xyz(box[kn]);
    Sum totn = tot[kn];
    for (int tt = 0; tt < n; tt++) {
        const int k = kn;
        const int bkz = uni(Bn.z), bkx = uni(Bn.x), bky = uni(Bn.y), bkx2 = uni(Bn.x2),
                  bky2 = uni(Bn.y2);
        const Sum total = uni(totn);
        if (tt + 1 < n) {
            kn = read_lane_of(o, (tt + 1) >> 6, (tt + 1) & (kWave - 1));
            Bn = box_xyz(box[kn]);
            totn = tot[kn];
        }
        if (bkz == 0) break;   // the rest are floor boxes: nothing supports them
        if (total == 0) continue;
        const double lk = read_lane_of(l, k >> 6, k & (kWave - 1));
        const double dt = (double)total;
        // (double)a of the reference's 64-bit area a = ox * oy equals the FP64
        // product of the two extents: both are exact (a < 2^42 < 2^53)
#pragma unroll
        for (int q = 0; q < R; q++) {
            const int ox = min(bkx2, b[q].x2) - max(bkx, b[q].x);
            const int oy = min(bky2, b[q].y2) - max(bky, b[q].y);
            if ((b[q].z2 == bkz) & (ox > 0) & (oy > 0))
                l[q] = l[q] + lk * ((double)ox * (double)oy) / dt;
        }
    }
    // FSR_i = ((mu / 0.3) * load_i) / m_i; score = #(FSR >= 1) / n
    int c = 0;
#pragma unroll
    for (int q = 0; q < R; q++) {
        const double f = (mu / 0.3) * l[q] / (v[q] ? mass(b[q]) : 1.0);   // m recomputed: same value
        if (fsr_out && v[q]) fsr_out[lane + q * kWave] = f;
        c += __popcll(ballot(v[q] && f >= 1.0));
    }
    wave_sync();   // ord[] is free again
    return (double)c / (double)n;
}

// compute_fsr + stability_score (plp_common.hpp:356-399) for the n boxes in
// LDS, on one wave.  The reference adds load[k] * a / total into load[j]
// while visiting k in (z desc, index asc) order, and the k loop stays serial
// so every load[j] receives its additions in the reference's order
// (bit-exact FP64).  Up to 64 * kRegBoxes boxes (kRegBoxes 2 or 4: the
// registers a kernel can spare at its pass end) the state stays in registers
// (fsr_score_reg); beyond, loads and totals sit in LDS and the order comes
// from a bitonic sort of 64-bit keys ((2^31-1-z) << 32 | index).  Returns
// the score; per-box FSR goes to fsr_out when non-null.  Contact totals are
// Sums (32-bit when every footprint is below 2^30, else 64-bit).
template <int kRegBoxes = 2, class Sum = uint32_t>
static __device__ double fsr_score(const Box* __restrict__ box, int n, double mu, double* load,
                            uint64_t* skey, Sum* tot, double* fsr_out) {
    const int lane = lane_id();
    if (n <= 0) return 0.0;
    if (n <= 2 * kWave) return fsr_score_reg<2, Sum>(box, n, mu, (int*)skey, tot, fsr_out);
    if (kRegBoxes >= 4 && n <= 4 * kWave)
        return fsr_score_reg<4, Sum>(box, n, mu, (int*)skey, tot, fsr_out);
    const int n2 = next_pow2(n);
    for (int i = lane; i < n2; i += kWave)
        skey[i] = i < n ? (((uint64_t)(0x7fffffffu - (uint32_t)box[i].z) << 32) | (uint32_t)i)
                        : ~0ull;
    // load = m = RHO * V / 1e6; tot[k] = contact area of box k with the boxes
    // whose top is at its bottom (lane per k, every j in turn).  Those boxes
    // do not overlap one another, so tot[k] <= w * d fits a Sum.
    for (int k = lane; k < n; k += kWave) {
        const Box Bk = box[k];
        const int64_t v = (int64_t)(Bk.x2 - Bk.x) * (Bk.y2 - Bk.y) * (Bk.z2 - Bk.z);
        load[k] = 50.0 * (double)v / 1.0e6;
        Sum t = 0;
        if (Bk.z > 0) {
            for (int j = 0; j < n; j++) {
                const Box Bj = box[j];
                const int ox = min(Bk.x2, Bj.x2) - max(Bk.x, Bj.x);
                const int oy = min(Bk.y2, Bj.y2) - max(Bk.y, Bj.y);
                if (j != k && Bj.z2 == Bk.z && ox > 0 && oy > 0)
                    t += (Sum)(uint32_t)ox * (uint32_t)oy;
            }
        }
        tot[k] = t;
    }
    wave_sync();
    // Bitonic sort, ascending.
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < n2; i += kWave) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t p = skey[i], q = skey[ixj];
                    const bool up = (i & k) == 0;
                    if ((p > q) == up) {
                        skey[i] = q;
                        skey[ixj] = p;
                    }
                }
            }
            wave_sync();
        }
    }
    for (int t = 0; t < n; t++) {
        const int k = (int)(uint32_t)uni(skey[t]);
        const Box Bk = box[k];
        const int bkz = uni(Bk.z);
        if (bkz == 0) break;   // the rest are floor boxes: nothing supports them
        const int bkx = uni(Bk.x), bky = uni(Bk.y), bkx2 = uni(Bk.x2), bky2 = uni(Bk.y2);
        const Sum total = uni(tot[k]);
        if (total > 0) {
            const double lk = load[k];
            for (int j = lane; j < n; j += kWave) {
                const Box Bj = box[j];
                if (j != k && Bj.z2 == bkz) {
                    const int ox = min(bkx2, Bj.x2) - max(bkx, Bj.x);
                    const int oy = min(bky2, Bj.y2) - max(bky, Bj.y);
                    if (ox > 0 && oy > 0) {
                        const uint64_t a = (uint64_t)ox * (uint64_t)oy;
                        load[j] = load[j] + lk * (double)a / (double)total;
                    }
                }
            }
        }
        wave_sync();
    }
    uint32_t cnt = 0;
    for (int i = lane; i < n; i += kWave) {
        const int64_t v = (int64_t)(box[i].x2 - box[i].x) * (box[i].y2 - box[i].y) *
                          (box[i].z2 - box[i].z);
        const double m = 50.0 * (double)v / 1.0e6;
        const double f = (mu / 0.3) * load[i] / m;
        if (fsr_out) fsr_out[i] = f;
        cnt += (f >= 1.0) ? 1u : 0u;
    }
    const uint64_t c = wave_sum_u64(cnt);
    wave_sync();
    return (double)c / (double)n;
}

// Bit of Box::pad set on every box that another box rests on (its top at
// that box's bottom, the footprints overlapping).  The item decoder's pad is
// otherwise 0; the block decoder keeps a, b and k in bits 0..23.
constexpr int kSupports = 1 << 30;

// An upper bound on the number of boxes whose FSR (plp_common.hpp:356-390)
// is >= 1.  compute_fsr adds to load[j] only when box j supports the box
// being visited, so a box that supports none keeps load = m, and its FSR
// ((mu / 0.3) * m) / m is computed here in fsr_score's operation order; a
// supporting box counts as one.  One wave.
__device__ __forceinline__ int fsr_count_bound(const Box* __restrict__ box, int n, double mu) {
    int c = 0;
    for (int i0 = 0; i0 < n; i0 += kWave) {
        const int i = i0 + lane_id();
        bool hit = false;
        if (i < n) {
            const Box B = box[i];
            if (B.pad & kSupports) {
                hit = true;
            } else {
                const int64_t v = (int64_t)(B.x2 - B.x) * (B.y2 - B.y) * (B.z2 - B.z);
                const double m = 50.0 * (double)v / 1.0e6;
                hit = (mu / 0.3) * m / m >= 1.0;
            }
        }
        c += __popcll(ballot(hit));
    }
    return c;
}

// Corners j of the box [x,X2) x [y,Y2) x [z,Z2) (bit 0 of j -> x side, bit 1
// -> y, bit 2 -> z) whose smallest later box [jx, jx+wm) x [jy, jy+dm) x
// [jz, jz+hm) meets one of boxes[0..n_box): bit j of the result (wave-uniform).
// On the way, the boxes the new box rests on get their kSupports mark.
__device__ __forceinline__ uint32_t corner_kill(Box* __restrict__ box, int n_box, int x, int y,
                                                int z, int X2, int Y2, int Z2, int wm, int dm,
                                                int hm) {
    // per lane: the corners its boxes kill, as the AND of per-axis masks
    // (x side 0: corners 0x55, side 1: 0xaa; y: 0x33 / 0xcc; z: 0x0f / 0xf0)
    uint32_t kill = 0;
    for (int b0 = 0; b0 < n_box; b0 += kWave) {
        const int bi = b0 + lane_id();
        const bool ok = bi < n_box;
        const Box B = box[ok ? bi : b0];
        if (ok & (B.z2 == z) & (B.x < X2) & (x < B.x2) & (B.y < Y2) & (y < B.y2))
            box[bi].pad = B.pad | kSupports;
        const uint32_t mx = (((x < B.x2) & (B.x < x + wm)) ? 0x55u : 0u) |
                            (((X2 < B.x2) & (B.x < X2 + wm)) ? 0xaau : 0u);
        const uint32_t my = (((y < B.y2) & (B.y < y + dm)) ? 0x33u : 0u) |
                            (((Y2 < B.y2) & (B.y < Y2 + dm)) ? 0xccu : 0u);
        const uint32_t mz = (((z < B.z2) & (B.z < z + hm)) ? 0x0fu : 0u) |
                            (((Z2 < B.z2) & (B.z < Z2 + hm)) ? 0xf0u : 0u);
        kill |= ok ? (mx & my & mz) : 0u;
    }
    return wave_or(kill);
}

// Corners whose key is in src[0..n): bit j of the result (wave-uniform).
template <class Key>
__device__ __forceinline__ uint32_t corner_present(const Key* src, int n, const Key ck[8]) {
    uint32_t present = 0;
    for (int i0 = 0; i0 < n; i0 += kWave) {
        const int i = i0 + lane_id();
        const Key v = src[i < n ? i : i0];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (ballot((i < n) & (v == ck[j]))) present |= 1u << j;
    }
    return present;
}

// Extreme-point update after placing the box [x, X2) x [y, Y2) x [z, Z2)
// (plp_common.hpp:439-453): points strictly inside the box leave the list,
// the box's 8 corners join it unless already present, and the list stays
// sorted and unique.  The live list keeps only points where a box of the
// later steps' smallest dimensions (wm, dm, hm, componentwise over every
// orientation) fits inside the walls without overlapping a placed box: any
// later box contains that one, so elsewhere nothing can ever be placed.  On
// the live list src -> dst: the points whose smallest box meets the new box
// or a wall leave (the strictly inside ones among them leave the reference's
// list too), and the useful corners are inserted, a point of src equal to
// one giving way to it; oidx[] records where each dst entry was in src
// (kCorner + j for inserted corner j).  Usefulness only shrinks as boxes are
// added and the later steps' smallest box grows, so a useful corner never
// equals a point that left, and the list stays a set.  With count_work
// the other points of the reference's list are kept (unsorted) in dead[]: it
// drops its strictly inside points and receives the points that leave the
// live list without being strictly inside and the corners that are neither
// useful nor already listed.  n_act <= cap on entry keeps every write within
// the cap + 8 slots of a list.  One wave; returns the larger new list size.
//
// With point states (SlotMerge given) each list entry carries the slot of
// its state instead of oidx: a kept point keeps its slot and an inserted
// corner j takes the slot lane j holds in sm->cslot.  (The slots of the
// points that leave are freed by the wave that updates the states, which
// applies the same rule.)  The inserted corners (bit j: corner j) go to
// *ins_out when given, and to the LDS word *ins_post as 0x100 | ins as soon
// as they are known, before the list is merged.
struct SlotMerge {
    const uint16_t* src;   // slot of each src entry
    uint16_t* dst;         // slot of each dst entry
    uint32_t cslot;        // lane j < 8: the slot of corner j
};

template <class Key>
static __device__ int merge_points(const Key* src, Key* dst, uint16_t* oidx, Key* dead,
                            int& n_act, int& n_dead, Box* box, int n_box, int x, int y, int z,
                            int X2, int Y2, int Z2, int L, int W, int H, int xb, int yb, int cap,
                            bool count_work, int wm, int dm, int hm, SlotMerge* sm = nullptr,
                            uint32_t* ins_out = nullptr, int* ins_post = nullptr) {
    const int lane = lane_id();
    // Corner j: bit 0 -> x side, bit 1 -> y, bit 2 -> z; with the key's field
    // order (z, then y, then x) the keys increase with j.  Lane j < 8 holds
    // corner j's key in myk.
    Key ck[8];
    uint32_t cdead = 0;
    Key myk = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int jx = (j & 1) ? X2 : x, jy = (j & 2) ? Y2 : y, jz = (j & 4) ? Z2 : z;
        ck[j] = xyz_key<Key>(jx, jy, jz, xb, yb);
        if (lane == j) myk = ck[j];
        if ((jx + wm > L) | (jy + dm > W) | (jz + hm > H)) cdead |= 1u << j;
    }
    // Corners whose smallest box meets a placed box (the new one included)
    // are not useful; the useful ones are inserted, and a listed point equal
    // to one gives way to it (same point: the list as a set and its order
    // are the reference's; the point's pair states are then the corner's,
    // computed against every box, instead of its old ones plus the new box).
    cdead |= corner_kill(box, n_box, x, y, z, X2, Y2, Z2, wm, dm, hm);
    const uint32_t ins = ~cdead & 0xffu;
    if (ins_out) *ins_out = ins;
    // posted for another wave of the block as soon as known: 0x100 | ins
    if (ins_post != nullptr && lane == 0)
        __hip_atomic_store(ins_post, (int)(0x100u | ins), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);

    // Dead list: drop the strictly inside points in place (a point moves to
    // an index no larger than its own, below the next 64-point chunk).
    uint32_t eq_dead = 0, present = 0;
    int nd = n_dead;
    if (count_work) {
        present = corner_present(src, n_act, ck);
        int db = 0;
        for (int i0 = 0; i0 < n_dead; i0 += kWave) {
            const int i = i0 + lane;
            const bool ok = i < n_dead;
            const Key v = ok ? dead[i] : Key(0);
            int px, py, pz;
            key_xyz(v, xb, yb, px, py, pz);
            const bool strict = (x < px) & (px < X2) & (y < py) & (py < Y2) & (z < pz) & (pz < Z2);
            const bool keep = ok & !strict;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (ballot(keep & (v == ck[j]))) eq_dead |= 1u << j;
            const uint64_t km = ballot(keep);
            if (keep) dead[db + rank_below(km)] = v;
            db += __popcll(km);
        }
        nd = uni(db);
        wave_sync();
    }

    // Live list: survivors go to dst after the inserted corners below them;
    // lane j < 8 counts the survivors below corner j (klt).
    int base = 0, klt = 0;
    for (int i0 = 0; i0 < n_act; i0 += kWave) {
        const int i = i0 + lane;
        const bool ok = i < n_act;
        const Key v = ok ? src[i] : Key(0);
        int px, py, pz;
        key_xyz(v, xb, yb, px, py, pz);
        const bool gone = ok & ((px + wm > L) | (py + dm > W) | (pz + hm > H) |
                                ((px < X2) & (x < px + wm) & (py < Y2) & (y < py + dm) &
                                 (pz < Z2) & (z < pz + hm)));
        int shift = 0, eq = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int on = (int)((ins >> j) & 1u);
            shift += on & (int)(ck[j] < v);
            eq |= on & (int)(ck[j] == v);
        }
        const bool keep = ok & !gone & (eq == 0);
        const uint64_t km = ballot(keep);
        if (sm) {
            if (keep) {
                const int pos = base + rank_below(km) + shift;
                dst[pos] = v;
                sm->dst[pos] = sm->src[i];
            }
        } else if (keep) {
            const int pos = base + rank_below(km) + shift;
            dst[pos] = v;
            oidx[pos] = (uint16_t)i;
        }
        base += __popcll(km);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = __popcll(km & ballot(v < ck[j]));
            if (lane == j) klt += c;
        }
        if (count_work) {
            const bool strict = (x < px) & (px < X2) & (y < py) & (py < Y2) & (z < pz) & (pz < Z2);
            const bool mv = gone & !strict;
            const uint64_t mm = ballot(mv);
            const int pos = nd + rank_below(mm);
            if (mv && pos < cap + 8) dead[pos] = v;
            nd += __popcll(mm);
        }
    }
    base = uni(base);
    if (lane < 8 && ((ins >> lane) & 1u)) {
        const int pos = klt + __popc(ins & ((1u << lane) - 1u));
        dst[pos] = myk;
        if (sm) sm->dst[pos] = (uint16_t)sm->cslot;
        else oidx[pos] = (uint16_t)(kCorner + lane);
    }
    if (count_work) {
        const uint32_t dadd = cdead & ~present & ~eq_dead;
        if (lane < 8 && ((dadd >> lane) & 1u)) {
            const int pos = nd + __popc(dadd & ((1u << lane) - 1u));
            if (pos < cap + 8) dead[pos] = myk;
        }
        nd += __popc(dadd);
    }
    wave_sync();
    n_act = base + __popc(ins);
    n_dead = count_work ? uni(nd) : 0;
    return max(n_act, n_dead);
}

// The pair states (cache entries) of the 8 corners of the box nb just
// placed, for the orientations in ors (no of them), against all n_box boxes:
// lane j * no + o.  One wave.
template <class Sum>
static __device__ inline void corner_states(const int* nb, Sum* cstate, const int* ors,
                                     int no, const Box* box, int n_box, int L, int W, int H,
                                     Box* cbox) {
    const int lane = lane_id();
    const int n = 8 * no;
    const int j = div_orients(lane, no, orient_recip(no)), o = lane - __mul24(j, no);
    // lanes n..63 compute a corner too (j & 7, o < no: in bounds) and are
    // masked out by c2
    const int x = (j & 1) ? nb[3] : nb[0];
    const int y = (j & 2) ? nb[4] : nb[1];
    const int z = (j & 4) ? nb[5] : nb[2];
    const int w = ors[3 * o], d = ors[3 * o + 1], h = ors[3 * o + 2];
    const bool c2 = (lane < n) & (x + w <= L) & (y + d <= W) & (z + h <= H);
    bool c12 = false;
    Sum s = 0;
    if (ballot(c2) != 0) {
        const int xlo = wave_min(c2 ? x : 0x7fffffff);
        const int xhi = wave_max(c2 ? x + w : -1);
        const int ylo = wave_min(c2 ? y : 0x7fffffff);
        const int yhi = wave_max(c2 ? y + d : -1);
        const int zlo = wave_min(c2 ? z : 0x7fffffff);
        const int zhi = wave_max(c2 ? z + h : -1);
        c12 = pair_scan(x, y, z, w, d, h, c2, box, n_box, xlo, xhi, ylo, yhi, zlo, zhi, cbox, s);
    }
    if (lane < n) cstate[lane] = c12 ? s : kBlockedOf<Sum>;
}

__device__ __forceinline__ void write_rows(int32_t* dst, const Box* box, int n) {
    for (int t = lane_id(); t < 7 * n; t += kWave) {
        const int i = t / 7, c = t - (t / 7) * 7;
        const Box& B = box[i];
        int v;
        switch (c) {
            case 0: v = B.type; break;
            case 1: v = B.x; break;
            case 2: v = B.y; break;
            case 3: v = B.z; break;
            case 4: v = B.x2 - B.x; break;
            case 5: v = B.y2 - B.y; break;
            default: v = B.z2 - B.z; break;
        }
        dst[t] = v;
    }
}

// The record of a decoder's selected packing (D4', plp_common.hpp:493-511:
// the best accepted packing, else the best of all with fallback set).
__device__ __forceinline__ RowRecord row_record(const DecState& s, uint64_t seed) {
    const bool have_acc = s.n_acc >= 0;
    RowRecord r;
    r.n_boxes = have_acc ? s.n_acc : max(0, s.n_all);
    r.fallback = have_acc ? 0 : 1;
    r.n_passes = s.passes;
    r.err = s.err;
    r.volume = have_acc ? s.v_acc : s.v_all;
    r.score = have_acc ? s.s_acc : s.s_all;
    r.seed = seed;
    r.work_pair_box = s.work_pair_box;
    r.work_live = s.work_live;
    r.max_ep = (int32_t)s.max_ep;
    r.ep_need = (int32_t)s.ep_need;
    return r;
}

}  // namespace plpk
