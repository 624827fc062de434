// point_state.hpp -- orientation-free feasibility states of the item
// decoder's live extreme points (VerdictKind kPointState).
//
// check_feasible_one (plp_common.hpp:297-321) of a box w x d x h at a live
// point p against the placed boxes depends on them only through:
//   C1  a placed box B meets [p, p + (w, d, h)) iff B is ahead of p
//       (B.x2 > p.x, B.y2 > p.y, B.z2 > p.z) and q < (w, d, h) componentwise,
//       q = (max(0, B.x - p.x), max(0, B.y - p.y), max(0, B.z - p.z));
//   C3  the support sum over the boxes whose top is at p.z, each adding the
//       overlap of its footprint with [p.x, p.x + w) x [p.y, p.y + d).
// So a point keeps, instead of per-orientation results:
//   * the componentwise-minimal q of the boxes ahead of it whose q is below
//     the largest later box on every axis (WM, DM, HM: a larger q can block no
//     later box, and the largest later box only shrinks as steps pass); a q
//     with some kept q' <= q blocks nothing q' does not, and one that is <=
//     kept entries makes them redundant;
//   * the footprints, relative to p, of the boxes whose top is at p.z that
//     reach into [p.x, p.x + WM) x [p.y, p.y + DM) (the others add nothing to
//     any later box's support).
// Both follow from the previous state and the one new box when a box is
// placed, and a new corner's state from one pass over the boxes.  The verdict
// of any orientation is then a test against at most kPtsStair + kPtsSup
// entries, with no box scan.  A point whose entries do not fit is marked
// (kPtsOverflow) and its pairs scan every box instead (exact either way).
//
// Entries are 16-bit in kNarrow, where every orientation side is at most
// 65535: q < WM <= 65535 and a, b < WM fit, and a2, b2 are clamped at 65535,
// which leaves min(w, a2) unchanged for every w <= 65535.  Else 32-bit, where
// no value reaches 2^22.
#pragma once

#include "decoder_dev.hpp"

namespace plpk {

template <int M>
struct PtsEntOf {
    using T = uint32_t;
};
template <>
struct PtsEntOf<kNarrow> {
    using T = uint16_t;
};

template <class Ent>
constexpr uint32_t kEntMax = (uint32_t)(Ent)~Ent(0);

// Entry indices (engine_types.hpp): staircase k at 4k, flags at 3, support k
// at 16 + 4k.
__host__ __device__ constexpr int kPtsQx(int k) { return 4 * k; }
__host__ __device__ constexpr int kPtsSa(int k) { return 16 + 4 * k; }
constexpr int kPtsFlags = 3;

// The entries held in 32-bit words (two 16-bit entries a word in kNarrow),
// read and written at compile-time indices only, so a state copied into a
// lane stays in registers.
template <class Ent>
struct alignas(16) Pts {
    static constexpr int kPerWord = 4 / (int)sizeof(Ent);
    uint32_t w[kPtsWords / kPerWord];
    __device__ __forceinline__ uint32_t get(int i) const {
        if constexpr (kPerWord == 1) return w[i];
        else return (w[i >> 1] >> (16 * (i & 1))) & 0xffffu;
    }
    __device__ __forceinline__ void set(int i, uint32_t v) {
        if constexpr (kPerWord == 1) {
            w[i] = v;
        } else {
            const int sh = 16 * (i & 1);
            w[i >> 1] = (w[i >> 1] & ~(0xffffu << sh)) | ((v & 0xffffu) << sh);
        }
    }
};

// No boxes: neutral staircase entries (max, max, max) block nothing, neutral
// support entries (0, 0, 0, 0) add nothing.
template <class Ent>
__device__ __forceinline__ Pts<Ent> pts_empty() {
    Pts<Ent> s;
#pragma unroll
    for (int i = 0; i < kPtsWords / Pts<Ent>::kPerWord; i++) s.w[i] = 0;
#pragma unroll
    for (int k = 0; k < kPtsStair; k++) {
        s.set(kPtsQx(k), kEntMax<Ent>);
        s.set(kPtsQx(k) + 1, kEntMax<Ent>);
        s.set(kPtsQx(k) + 2, kEntMax<Ent>);
    }
    return s;
}

// C1 and the support sum of a w x d x h box at the state's point.  Returns
// true when no placed box meets it.
template <class Ent, class Sum>
__device__ __forceinline__ bool pts_eval(const Pts<Ent>& s, int w, int d, int h, Sum& sup) {
    bool blocked = false;
#pragma unroll
    for (int k = 0; k < kPtsStair; k++) {
        const uint32_t ex = s.get(kPtsQx(k)), ey = s.get(kPtsQx(k) + 1), ez = s.get(kPtsQx(k) + 2);
        blocked |= (ex < (uint32_t)w) & (ey < (uint32_t)d) & (ez < (uint32_t)h);
    }
    Sum t = 0;
#pragma unroll
    for (int k = 0; k < kPtsSup; k++) {
        const int ox = min(w, (int)s.get(kPtsSa(k) + 1)) - (int)s.get(kPtsSa(k));
        const int oy = min(d, (int)s.get(kPtsSa(k) + 3)) - (int)s.get(kPtsSa(k) + 2);
        t = mad_sum((uint32_t)(((ox > 0) & (oy > 0)) ? ox : 0), (uint32_t)oy, t);
    }
    sup = t;
    return !blocked;
}

// Whether box B changes the state of the point (px, py, pz), with WM, DM, HM
// the largest later box's sides.
__device__ __forceinline__ bool pts_touches(int px, int py, int pz, const Box& B, int WM, int DM,
                                            int HM) {
    const bool ahead = (B.x2 > px) & (B.y2 > py) & (B.z2 > pz) & (B.x - px < WM) &
                       (B.y - py < DM) & (B.z - pz < HM);
    const bool sup = (B.z2 == pz) & (B.x2 > px) & (B.y2 > py) & (B.x - px < WM) & (B.y - py < DM);
    return ahead | sup;
}

// Adds box B to the state of the point (px, py, pz).  The first staircase
// entry that q makes redundant (or an empty one) takes q, the others it makes
// redundant are emptied; a support box takes the first empty support entry.
template <class Ent>
__device__ __forceinline__ void pts_absorb(Pts<Ent>& s, int px, int py, int pz, const Box& B,
                                           int WM, int DM, int HM) {
    const int qx = max(B.x - px, 0), qy = max(B.y - py, 0), qz = max(B.z - pz, 0);
    const bool ahead = (B.x2 > px) & (B.y2 > py) & (B.z2 > pz) & (qx < WM) & (qy < DM) & (qz < HM);
    const bool sup = (B.z2 == pz) & (B.x2 > px) & (B.y2 > py) & (qx < WM) & (qy < DM);
    uint32_t flags = s.get(kPtsFlags);
    if (ahead) {
        bool dom = false;
#pragma unroll
        for (int k = 0; k < kPtsStair; k++) {
            const uint32_t ex = s.get(kPtsQx(k)), ey = s.get(kPtsQx(k) + 1),
                           ez = s.get(kPtsQx(k) + 2);
            dom |= (ex <= (uint32_t)qx) & (ey <= (uint32_t)qy) & (ez <= (uint32_t)qz);
        }
        if (!dom) {
            bool placed = false;
#pragma unroll
            for (int k = 0; k < kPtsStair; k++) {
                const uint32_t ex = s.get(kPtsQx(k)), ey = s.get(kPtsQx(k) + 1),
                               ez = s.get(kPtsQx(k) + 2);
                const bool red = ((uint32_t)qx <= ex) & ((uint32_t)qy <= ey) & ((uint32_t)qz <= ez);
                const bool take = red & !placed;
                const uint32_t nx = take ? (uint32_t)qx : kEntMax<Ent>;
                const uint32_t ny = take ? (uint32_t)qy : kEntMax<Ent>;
                const uint32_t nz = take ? (uint32_t)qz : kEntMax<Ent>;
                if (red) {
                    s.set(kPtsQx(k), nx);
                    s.set(kPtsQx(k) + 1, ny);
                    s.set(kPtsQx(k) + 2, nz);
                }
                placed |= take;
            }
            if (!placed) flags |= kPtsOverflow;
        }
    }
    if (sup) {
        const uint32_t a = (uint32_t)qx, b = (uint32_t)qy;
        const uint32_t a2 = min((uint32_t)(B.x2 - px), kEntMax<Ent>);
        const uint32_t b2 = min((uint32_t)(B.y2 - py), kEntMax<Ent>);
        bool placed = false;
#pragma unroll
        for (int k = 0; k < kPtsSup; k++) {
            const bool take = (s.get(kPtsSa(k) + 1) == 0) & !placed;
            if (take) {
                s.set(kPtsSa(k), a);
                s.set(kPtsSa(k) + 1, a2);
                s.set(kPtsSa(k) + 2, b);
                s.set(kPtsSa(k) + 3, b2);
            }
            placed |= take;
        }
        if (!placed) flags |= kPtsOverflow;
    }
    s.set(kPtsFlags, flags);
}

// C1 and the support sum of this lane's pair against every placed box, for
// lanes whose point state overflowed (alive: the lane's pair passed C2).
// Uniform box reads in groups of kBoxGroup; boxes[n_box ..] are sentinels up
// to n_box + kBoxGroup - 1 (the box array keeps 8 past its capacity).
template <class Sum>
__device__ __forceinline__ bool scan_all(int x, int y, int z, int w, int d, int h, bool alive,
                                         const Box* __restrict__ boxes, int n_box, Sum& s_out) {
    const int x2 = x + w, y2 = y + d, z2 = z + h;
    Sum s = 0;
    uint64_t am = ballot(alive);
    for (int g = 0; g < n_box && am != 0; g += kBoxGroup) {
        Box B[kBoxGroup];
#pragma unroll
        for (int u = 0; u < kBoxGroup; u++) B[u] = boxes[g + u];
#pragma unroll
        for (int u = 0; u < kBoxGroup; u++) {
            const int ox = min(x2, B[u].x2) - max(x, B[u].x);
            const int oy = min(y2, B[u].y2) - max(y, B[u].y);
            const int oz = min(z2, B[u].z2) - max(z, B[u].z);
            const int oxy = min(ox, oy);
            am &= ~ballot(min(oxy, oz) > 0);
            const bool sup = (B[u].z2 == z) & (oxy > 0);
            s = mad_sum((uint32_t)(sup ? ox : 0), (uint32_t)oy, s);
        }
    }
    s_out = s;
    return (am >> lane_id()) & 1ull;
}

// The verdicts of this lane's pair -- a w x d x h box at the point (x, y, z)
// whose state is ps -- for the lanes in act: C2, C1 and the support from the
// state (or, for a point whose state overflowed, from a scan of every placed
// box), C3 on top (pair_verdict).  Returns the wave's ballot of feasible
// pairs.
template <class Sum, class Ent>
__device__ __forceinline__ uint64_t pts_verdicts(const Pts<Ent>& ps, bool act, int x, int y, int z,
                                                 int w, int d, int h, int L, int W, int H,
                                                 const Box* __restrict__ box, int n_box) {
    const bool c2 = act & (x + w <= L) & (y + d <= W) & (z + h <= H);
    Sum sup = 0;
    const bool c1 = pts_eval(ps, w, d, h, sup);
    const bool ovf = (ps.get(kPtsFlags) & kPtsOverflow) != 0;
    uint64_t m = ballot(pair_verdict(c2 & c1 & !ovf, z, w, d, sup));
    if (ballot(c2 & ovf) != 0) {
        Sum s2 = 0;
        const bool a2 = scan_all(x, y, z, w, d, h, c2 & ovf, box, n_box, s2);
        m |= ballot((c2 & ovf) & pair_verdict(a2, z, w, d, s2));
    }
    return m;
}

// Bits b0 .. b0 + n - 1 (n <= 8) of the verdict masks, as the low bits of
// the result: the verdicts of one point's pairs.
__device__ __forceinline__ uint32_t mask_bits(const uint64_t* mask, int b0, int n) {
    const int wi = b0 >> 6, sh = b0 & 63;
    uint64_t v = mask[wi] >> sh;
    if (sh + n > 64) v |= mask[wi + 1] << (64 - sh);
    return (uint32_t)v & ((1u << n) - 1u);
}

// Per-axis corner masks (corner j: bit 0 of j -> x side, bit 1 -> y, bit 2
// -> z; corners on the low side of an axis are 0x55, 0x33 or 0x0f, on the
// high side 0xaa, 0xcc or 0xf0): the corners c whose axis value c_a has
// c_a < hi and lo - c_a < m, for a box [lo, hi) on that axis.
__device__ __forceinline__ uint32_t axis_mask(int lo, int hi, int c0, int c1, int m, uint32_t m0,
                                              uint32_t m1) {
    return (((c0 < hi) & (lo - c0 < m)) ? m0 : 0u) | (((c1 < hi) & (lo - c1 < m)) ? m1 : 0u);
}

// The corners j of the box nb just placed (x, y, z, x2, y2, z2) that the
// extreme-point merge inserts: those where the later steps' smallest box
// (wm, dm, hm) stays inside the walls and meets none of boxes[0..n_box) (the
// new one included) -- merge_points' rule.  Bit j of the (wave-uniform)
// result.
__device__ __forceinline__ uint32_t corner_live(const int* nb, const Box* box, int n_box, int L,
                                                int W, int H, int wm, int dm, int hm) {
    const int lane = lane_id();
    const int x0 = nb[0], y0 = nb[1], z0 = nb[2], x1 = nb[3], y1 = nb[4], z1 = nb[5];
    uint32_t live = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = (c & 1) ? x1 : x0, cy = (c & 2) ? y1 : y0, cz = (c & 4) ? z1 : z0;
        if ((cx + wm <= L) & (cy + dm <= W) & (cz + hm <= H)) live |= 1u << c;
    }
    uint32_t kill = 0;
    for (int g = 0; g < n_box; g += kWave) {
        const int b = g + lane;
        const Box B = b < n_box ? box[b] : sentinel_box();
        kill |= axis_mask(B.x, B.x2, x0, x1, wm, 0x55u, 0xaau) &
                axis_mask(B.y, B.y2, y0, y1, dm, 0x33u, 0xccu) &
                axis_mask(B.z, B.z2, z0, z1, hm, 0x0fu, 0xf0u);
    }
    return live & ~wave_or(kill);
}

// The states of the corners in live (corner_live) of the box nb just placed,
// against all n_box boxes (the new one included), into pts[cslot[j]] for
// corner j.  One wave, lane = box, 64 boxes at a time; per live corner, the
// boxes ahead of it that a later box could meet join the staircase unless
// another such box or a kept entry makes them redundant (then the kept
// entries they make redundant leave), and the support boxes join the support
// list.  Which corners a box is ahead of or supports comes per axis as an
// 8-corner mask.
template <class Ent>
__device__ __forceinline__ void pts_corners(Pts<Ent>* pts, const uint16_t* cslot, const int* nb,
                                            const Box* box, int n_box, uint32_t live, int WM,
                                            int DM, int HM) {
    const int lane = lane_id();
    const int x0 = nb[0], y0 = nb[1], z0 = nb[2], x1 = nb[3], y1 = nb[4], z1 = nb[5];
    if (live == 0) return;
    const int my_slot = cslot[lane & 7];
    if (lane < 8 && ((live >> lane) & 1u)) pts[my_slot] = pts_empty<Ent>();
    wave_sync();
    Ent* const ent = (Ent*)pts;
    for (int g = 0; g < n_box; g += kWave) {
        const int b = g + lane;
        const Box B = b < n_box ? box[b] : sentinel_box();
        const uint32_t ax = axis_mask(B.x, B.x2, x0, x1, WM, 0x55u, 0xaau);
        const uint32_t ay = axis_mask(B.y, B.y2, y0, y1, DM, 0x33u, 0xccu);
        const uint32_t az = axis_mask(B.z, B.z2, z0, z1, HM, 0x0fu, 0xf0u);
        const uint32_t zm = (B.z2 == z0 ? 0x0fu : 0u) | (B.z2 == z1 ? 0xf0u : 0u);
        const uint32_t rel8 = ax & ay & az & live;   // ahead, and a later box could meet it
        const uint32_t sup8 = ax & ay & zm & live;   // top at the corner's height, in reach
        const uint32_t any8 = wave_or(rel8 | sup8);
        for (uint32_t lv = any8; lv != 0; lv &= lv - 1) {
            const int c = __ffs(lv) - 1;
            const int cx = (c & 1) ? x1 : x0, cy = (c & 2) ? y1 : y0, cz = (c & 4) ? z1 : z0;
            Ent* const st = ent + (size_t)read_lane(my_slot, c) * kPtsWords;
            const bool ahead = (rel8 >> c) & 1u;
            const bool sup = (sup8 >> c) & 1u;
            const int qx = max(B.x - cx, 0), qy = max(B.y - cy, 0), qz = max(B.z - cz, 0);
            const uint64_t am = ballot(ahead);
            if (am != 0) {
                uint32_t e[kPtsStair][3];
                bool dom = false;
                if (g == 0) {   // the state is still empty
#pragma unroll
                    for (int k = 0; k < kPtsStair; k++)
#pragma unroll
                        for (int a = 0; a < 3; a++) e[k][a] = kEntMax<Ent>;
                } else {
#pragma unroll
                    for (int k = 0; k < kPtsStair; k++) {
#pragma unroll
                        for (int a = 0; a < 3; a++) e[k][a] = st[kPtsQx(k) + a];
                        dom |= (e[k][0] <= (uint32_t)qx) & (e[k][1] <= (uint32_t)qy) &
                               (e[k][2] <= (uint32_t)qz);
                    }
                }
                // redundant next to another box of this window: its q is <=
                // componentwise (equal ones: the lowest box index stays)
                for (uint64_t m = am & (am - 1) ? am : 0; m != 0; m &= m - 1) {
                    const int bp = __ffsll((unsigned long long)m) - 1;
                    const int px = read_lane(qx, bp), py = read_lane(qy, bp),
                              pz = read_lane(qz, bp);
                    const bool le = (px <= qx) & (py <= qy) & (pz <= qz);
                    const bool same = (px == qx) & (py == qy) & (pz == qz);
                    dom |= le & (!same | (bp < lane));
                }
                const bool surv = ahead & !dom;
                const uint64_t sm = ballot(surv);
                uint32_t keep = 0;
                if (g != 0) {
#pragma unroll
                    for (int k = 0; k < kPtsStair; k++) {
                        const bool red = ballot(surv & ((uint32_t)qx <= e[k][0]) &
                                                ((uint32_t)qy <= e[k][1]) &
                                                ((uint32_t)qz <= e[k][2])) != 0;
                        if ((e[k][0] != kEntMax<Ent>) & !red) keep |= 1u << k;
                    }
                }
                const int nk = __popc(keep), nn = __popcll(sm);
                if (nk + nn > kPtsStair) {
                    if (lane == 0) st[kPtsFlags] = (Ent)(st[kPtsFlags] | kPtsOverflow);
                } else if (nn != 0) {
                    // kept entries first (each moves down or stays), then the
                    // new ones, then empty entries
                    if (lane == 0 && g != 0) {
#pragma unroll
                        for (int k = 0; k < kPtsStair; k++) {
                            const int pos = __popc(keep & ((1u << k) - 1u));
                            if (((keep >> k) & 1u) && pos != k) {
#pragma unroll
                                for (int a = 0; a < 3; a++) st[kPtsQx(pos) + a] = (Ent)e[k][a];
                            }
                        }
#pragma unroll
                        for (int k = 0; k < kPtsStair; k++) {
                            if (k >= nk + nn) {
#pragma unroll
                                for (int a = 0; a < 3; a++) st[kPtsQx(k) + a] = (Ent)kEntMax<Ent>;
                            }
                        }
                    }
                    if (surv) {
                        const int pos = nk + rank_below(sm);
                        st[kPtsQx(pos)] = (Ent)qx;
                        st[kPtsQx(pos) + 1] = (Ent)qy;
                        st[kPtsQx(pos) + 2] = (Ent)qz;
                    }
                }
            }
            const uint64_t spm = ballot(sup);
            if (spm != 0) {
                int ns = 0;
                if (g != 0) {
#pragma unroll
                    for (int k = 0; k < kPtsSup; k++) ns += st[kPtsSa(k) + 1] != 0;
                }
                if (ns + __popcll(spm) > kPtsSup) {
                    if (lane == 0) st[kPtsFlags] = (Ent)(st[kPtsFlags] | kPtsOverflow);
                } else if (sup) {
                    const int pos = ns + rank_below(spm);
                    st[kPtsSa(pos)] = (Ent)qx;
                    st[kPtsSa(pos) + 1] = (Ent)min((uint32_t)(B.x2 - cx), kEntMax<Ent>);
                    st[kPtsSa(pos) + 2] = (Ent)qy;
                    st[kPtsSa(pos) + 3] = (Ent)min((uint32_t)(B.y2 - cy), kEntMax<Ent>);
                }
            }
            wave_sync();
        }
    }
}

}  // namespace plpk
