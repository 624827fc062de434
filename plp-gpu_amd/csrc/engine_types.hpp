// engine_types.hpp -- device-resident layouts shared by host and kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace plpk {

// Per-instance header in device memory (16 ints).
struct InstHdr {
    int32_t L, W, H, n_steps;
    int32_t step_off;   // first entry of this instance in the step-type table
    int32_t ors_off;    // first (instance, type) row of the orientation table
    int32_t n_types;
    int32_t xbits;      // bit length of L: packed key = z << (xbits+ybits) | y << xbits | x
    int32_t ybits;      // bit length of W
    // block-building engine only: first block shape, type record and entry of
    // the volume-ordered type list of this instance
    int32_t shape_off, tinfo_off, vorder_off;
    int32_t gtab_off;   // block engine: first greedy-plan entry of this instance
    int32_t pad[3];
};
static_assert(sizeof(InstHdr) == 64, "InstHdr layout");

// Widths of a decoder's extreme-point keys, support sums and pair-state
// cache entries (the kernels' StateMode template parameter): kNarrow 32-bit
// keys and sums with 16-bit cache entries, kWide the same with 32-bit
// entries, kBig 64 bits for all three.
enum StateMode { kNarrow = 0, kWide = 1, kBig = 2 };

__host__ __device__ constexpr int cache_entry_bytes(int mode) {
    return mode == kNarrow ? 2 : mode == kWide ? 4 : 8;
}
__host__ __device__ constexpr int key_bytes(int mode) { return mode == kBig ? 8 : 4; }

// Orientation table row: n, then up to 6 (w, d, h) triples, padded to 20 ints.
constexpr int kOrsRow = 20;

// Per-decoder state kept in device memory between launches.
struct DecState {
    int32_t passes;          // passes completed
    int32_t err;             // 0 or PLP_E_EPCAP
    int32_t n_all, n_acc;    // D4' best overall / best accepted packing size; -1 = none
    int64_t v_all, v_acc;    // their volumes
    double s_all, s_acc;     // their stability scores
    uint64_t work_pair_box;  // brute-force AABB tests so far (sum of n_cand * n_ors * n_box);
                             // counted only with count_work
    uint64_t work_live;      // the same over the steps and live points this kernel evaluates
                             // (count_work only)
    uint32_t work_steps;     // steps that evaluated feasibility
    uint32_t max_ep;         // largest reference extreme-point list (count_work only)
    uint32_t ep_need;        // largest list this decoder held (the capacity it needs)
    uint32_t box_need;       // most boxes it placed in a pass (box_cap + 1: it ran out)
};
static_assert(sizeof(DecState) == 80, "DecState layout");

constexpr int kMtWords = 625;      // 624 state words + index

// One decoder's result in the layout of plp_row_result (include/plp_hip.h):
// the selected packing's size, volume and score (best accepted, or best of
// all when no packing reached theta), written by the decoder kernels into
// pinned host memory at the end of every run.
struct RowRecord {
    int32_t n_boxes, fallback, n_passes, err;
    int64_t volume;
    double score;
    uint64_t seed;
    uint64_t work_pair_box, work_live;
    int32_t max_ep, ep_need;
};
static_assert(sizeof(RowRecord) == 64, "RowRecord layout");


// Shared prefixes.  Every decoder of an instance makes the same decisions as
// the instance's trajectory -- its pass with feas[0] picked at every step --
// until its first eps pick with an index other than 0: the verdicts depend
// only on the boxes placed, and along the way each placing step draws one
// random() whatever the seed.  A run first decodes each instance's
// trajectory once and records, per placing step j (box j is placed there):
// the step, the feasible-pair count, the live list it started from and the
// largest list held so far; then every decoder replays its own RNG along the
// trajectory and resumes from the list and boxes of the step where it
// diverges, or takes the trajectory's packing when it never does.
struct TrajHdr {
    int32_t valid;    // 1: recorded without a capacity stop
    int32_t n_box;    // boxes placed (= placing steps)
    int32_t need;     // largest live list held during the pass
    int32_t pad;
    int64_t volume;
    double score;     // stability_score of the packing (theta-independent)
};
static_assert(sizeof(TrajHdr) == 32, "TrajHdr layout");

struct TrajStep {
    int32_t s;        // index in the item sequence
    int32_t cnt;      // feasible pairs
    int32_t n_act;    // live points the step starts from (its list: key_off ..)
    int32_t need;     // largest live list held before this step's merge
    int32_t key_off;  // first key of the list in the instance's key area (-1: not kept,
                      // the list exceeds the decoders' capacity)
    int32_t pad[3];
};
static_assert(sizeof(TrajStep) == 32, "TrajStep layout");

struct DecodeArgs {
    const InstHdr* inst;
    const int32_t* step_type;
    const int32_t* step_min;     // 8 ints per step: smallest w, d, h over the later steps, 0,
                                 // largest w, d, h over the later steps, 0
    const int32_t* ors_tab;
    const int32_t* dec_inst;
    const uint64_t* dec_seed;
    const int32_t* work_order;   // queue position -> decoder index
    int32_t n_work;              // queue length
    int32_t n_passes;
    uint32_t* mt;            // n_dec * kMtWords
    const uint32_t* mt0;     // n_dec * kMtWords: each decoder's freshly seeded state
    DecState* state;         // n_dec
    int32_t* out_rows;       // pinned host: n_dec * max_items * 7, each decoder's selected
                             // packing (JSON row order), rewritten whenever the selection changes
    RowRecord* out_rec;      // pinned host: n_dec result records, written at the end of a run
    int32_t max_items;       // row stride of out_rows (the largest item count)
    int32_t ep_cap;
    int32_t box_cap;         // placed boxes a decoder holds in LDS; one more stops it (escalation)
    int32_t step_cap;        // step-table entries in LDS (>= every instance's step count)
    int32_t use_rng;
    int32_t use_theta;
    int32_t reset;           // 1: seed RNG and clear selection state first
    int32_t count_work;      // 1: also keep the dead points, to count the reference's work
    int32_t mode;            // StateMode of the loaded batch
    int32_t verdicts;        // VerdictKind of the kernel
    double eps, theta, mu;
    uint32_t* work_counter;  // [0] persistent-grid work queue head, [1] decoders ending in error
    // Shared prefixes: 0 off; 1 the queue items are instances whose
    // trajectories are recorded; 2 decoders resume from the recorded ones.
    int32_t traj;
    int32_t traj_key_cap;    // keys per instance in traj_keys (placing steps x (list cap + 8))
    int32_t traj_list_cap;   // record: lists longer than this are not kept (the decoders'
                             // capacity; a decoder reaching such a list stops anyway)
    int32_t pad_traj;
    uint32_t* ticks;         // per decoder, or null: the time its work took in the run
                             // (10 ns ticks of the constant-rate counter), for ordering the
                             // queue by measured length
    TrajHdr* traj_hdr;       // per instance
    TrajStep* traj_step;     // per instance: step_cap entries, by placing step
    void* traj_keys;         // per instance: traj_key_cap keys (Key of the mode)
    int32_t* traj_box;       // per instance: step_cap boxes of 8 ints (x, y, z, x2, y2, z2,
                             // type, index of the first box resting on it or 2^30)
    uint32_t* done;          // per decoder, or null: set to 1 (in host memory) once its row
                             // record and packing are there, for a host copy that overlaps
                             // the decode
};

// Dynamic LDS layout of one decoder block.  The pass-time region (two
// live extreme-point lists used alternately, the previous-list index of each
// entry of the current one, two pair-state caches used alternately, the
// dead-point list when work is counted, pair masks, per-wave culled box
// copies) is reused by FSR after the pass (loads, sort keys, contact
// totals).  Each list has ep_cap + 8 slots: a merge writes at most 8 entries
// past the capacity before the overflow is detected.  The box array has
// box_cap slots plus sentinels; the step table (step_cap entries) follows
// the MT state.
struct LdsLayout {
    int32_t off_ep, off_ep2, off_oidx, off_cache, n_cache, off_dead, off_mask, n_mask, off_rel;
    int32_t off_load, off_key, n_key, off_tot;   // FSR view of the first region
    int32_t off_box, off_mt, off_steps, total;
};

__host__ __device__ inline int32_t next_pow2(int32_t v) {
    int32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// A 16-bit pair-state cache entry holds min(s, 65534) exactly enough for C3
// when every orientation has w*d <= 2 * 65534: a clamped s is still >= w*d/2.
constexpr int64_t kNarrowMaxArea = 2 * 65534;

__host__ __device__ inline LdsLayout lds_layout(int32_t ep_cap, int32_t box_cap, int32_t use_rng,
                                                int32_t n_waves, int32_t count_work,
                                                int32_t mode, int32_t step_cap) {
    LdsLayout l;
    const int32_t list = (ep_cap + 8) * key_bytes(mode);
    l.off_ep = 0;
    l.off_ep2 = list;
    l.off_oidx = 2 * list;
    l.off_cache = (l.off_oidx + (ep_cap + 8) * 2 + 15) & ~15;
    l.n_cache = ep_cap * 6;
    l.off_dead = (l.off_cache + 2 * l.n_cache * cache_entry_bytes(mode) + 15) & ~15;
    l.off_mask = (l.off_dead + (count_work ? list : 0) + 7) & ~7;
    l.n_mask = (ep_cap * 6 + 63) / 64;
    l.off_rel = (l.off_mask + l.n_mask * 8 + 15) & ~15;
    int32_t pass_end = l.off_rel + n_waves * (64 + 8) * 32;   // per-wave culled window
    l.n_key = next_pow2(box_cap > 0 ? box_cap : 1);
    l.off_load = 0;
    l.off_key = box_cap * 8;
    l.off_tot = l.off_key + l.n_key * 8;
    int32_t fsr_end = l.off_tot + box_cap * (mode == kBig ? 8 : 4);   // contact totals
    int32_t region = pass_end > fsr_end ? pass_end : fsr_end;
    l.off_box = (region + 15) & ~15;
    l.off_mt = l.off_box + (box_cap + 8) * 32;
    l.off_steps = l.off_mt + (use_rng ? 624 * 4 : 0);   // per step: type | n_ors << 24
    l.total = l.off_steps + step_cap * 4;
    return l;
}

// ----------------------------------------------- per-point states ----
//
// The decoder's two ways to reach a step's pair verdicts (DecodeArgs.verdicts):
//   kPairCache   pair states (support sum or blocked) of the last two steps,
//                reused when the type repeats; other steps scan the boxes
//   kPointState  one orientation-free state per live point, kept current as
//                boxes are placed, from which every pair's verdict follows
enum VerdictKind { kPairCache = 0, kPointState = 1 };

// A point state: Ent v[kPtsWords] (Ent: 16 bits in kNarrow, where every
// orientation side is at most 65535, else 32 bits), in units of 4 entries:
//   v[4k .. 4k + 2]     staircase entry k < kPtsStair: q = (qx, qy, qz), the
//                       offsets max(0, B.lo - p) of a placed box ahead of p
//                       (B.hi > p on every axis) that a later box could still
//                       meet; only the componentwise-minimal ones are kept
//   v[3]                flags: kPtsOverflow when an entry did not fit
//   v[16 + 4k ..]       support entry k < kPtsSup: (a, a2, b, b2) = the
//                       x extent [a, a2) and y extent [b, b2) of a placed box
//                       whose top is at p.z, relative to p (a, b clamped at 0)
// Unused entries are neutral: staircase (max, max, max), support (0, 0, 0,
// 0), so they block and support nothing.
constexpr int kPtsStair = 4, kPtsSup = 4, kPtsWords = 32;
constexpr uint32_t kPtsOverflow = 1u;

__host__ __device__ constexpr int pts_entry_bytes(int mode) { return mode == kNarrow ? 2 : 4; }

// Dynamic LDS layout of one decoder block with per-point states: two live
// lists (keys and slot numbers, used alternately), the stack of free slots,
// ep_cap + 8 point states, per slot the verdict bits of the point's pairs
// for the next step when it repeats the type (bit o: orientation o), the dead
// list when work is counted, pair masks; FSR reuses the region after the
// pass; then boxes, MT state and the step table as in lds_layout.
struct PtsLayout {
    int32_t off_ep, off_ep2, off_sl, off_sl2, off_free, off_pts, off_bits, off_dead, off_mask,
        n_mask;
    int32_t off_load, off_key, n_key, off_tot;
    int32_t off_box, off_mt, off_steps, total;
};

__host__ __device__ inline PtsLayout pts_layout(int32_t ep_cap, int32_t box_cap, int32_t use_rng,
                                                int32_t count_work, int32_t mode, int32_t step_cap) {
    PtsLayout l;
    const int32_t slots = ep_cap + 8;
    const int32_t list = slots * key_bytes(mode);
    l.off_ep = 0;
    l.off_ep2 = list;
    l.off_sl = 2 * list;
    l.off_sl2 = l.off_sl + slots * 2;
    l.off_free = l.off_sl2 + slots * 2;
    l.off_pts = (l.off_free + slots * 2 + 15) & ~15;
    l.off_bits = l.off_pts + slots * kPtsWords * pts_entry_bytes(mode);
    l.off_dead = (l.off_bits + slots + 15) & ~15;
    l.off_mask = (l.off_dead + (count_work ? list : 0) + 7) & ~7;
    l.n_mask = (ep_cap * 6 + 63) / 64;
    const int32_t pass_end = l.off_mask + l.n_mask * 8;
    l.n_key = next_pow2(box_cap > 0 ? box_cap : 1);
    l.off_load = 0;
    l.off_key = box_cap * 8;
    l.off_tot = l.off_key + l.n_key * 8;
    const int32_t fsr_end = l.off_tot + box_cap * (mode == kBig ? 8 : 4);
    const int32_t region = pass_end > fsr_end ? pass_end : fsr_end;
    l.off_box = (region + 15) & ~15;
    l.off_mt = l.off_box + (box_cap + 8) * 32;
    l.off_steps = l.off_mt + (use_rng ? 624 * 4 : 0);
    l.total = l.off_steps + step_cap * 4;
    return l;
}

// ---------------------------------------------- block-building engine ----

// One block shape (blockproof/plp_block.cpp:27-33): cnt = a*b*k copies of
// item ti in one orientation as a bw x bd x bh stack.  An instance's shapes
// are stored in placement order (plp_block.cpp:156-165), so a shape's index
// is its sort key when a sweep's plan is ordered.
struct BlockShapeRec {
    int32_t ti, a, b, k, cnt, bw, bd, bh;
};

// Per (instance, type): copies, whether any orientation is allowed, the
// type's shapes in shape_precedence order (by_type[shp_off .. +shp_cnt)),
// the componentwise smallest orientation, and where the type's greedy-plan
// entries start (instance-local entry index of r = 1 copies left).
struct BlockTypeRec {
    int32_t qty, has_ors, shp_off, shp_cnt, wmin, dmin, hmin, gt_off;
};

// Greedy plan entry of (type, r copies left): the first shape in
// shape_precedence order whose copy count fits r is the start-th of the
// type's list (start == shp_cnt: none fits); s0 is that shape (-1 if none)
// and cnt0 its copy count.
struct GreedyEntry {
    int32_t start, s0, cnt0;
};

struct BlockArgs {
    DecodeArgs d;                  // queue, decoder state, outputs, capacities, eps / theta / mu
    const BlockShapeRec* shape;    // all instances' shapes (InstHdr::shape_off)
    const BlockTypeRec* tinfo;     // all instances' types (InstHdr::tinfo_off)
    const int32_t* by_type;        // shape indices per type, precedence order
    const int32_t* vorder;         // per instance: types by item volume desc, index asc
    const GreedyEntry* gtab;       // per instance and type, r = 1 .. qty (InstHdr::gtab_off)
    int32_t topk;                  // eps-greedy choice among the topk largest fitting shapes
    int32_t pick;                  // 0 first, 1 support, 2 contact (plp_block.cpp:56)
    int32_t max_types;             // largest type count of the batch
};

// LDS of one block-decoder block: two live lists + their previous-list
// index, verdict masks, remaining copies and the type list, the sweep plan
// (sorted in place), per-wave culled box windows; after the pass the first
// region holds FSR's arrays and the decomposed items.  Then the placed
// blocks, the MT state and the instance's type tables.
struct BlockLds {
    int32_t off_ep, off_ep2, off_oidx, off_mask, off_rem, off_tlist, off_plan, n_plan, off_rel;
    int32_t off_load, off_key, n_key, off_tot, off_items;
    int32_t off_box, off_mt, off_types, off_vord, off_gtab, total;
};

// box_cap: items the placed blocks may hold (the blocks, their decomposed
// items and FSR's arrays); full_items: every item of the largest instance
// (the sweep plan and the greedy-plan entries).
__host__ __device__ inline BlockLds block_lds_layout(int32_t ep_cap, int32_t box_cap,
                                                     int32_t full_items, int32_t max_types,
                                                     int32_t use_rng, int32_t n_waves,
                                                     int32_t big = 0) {
    BlockLds l;
    const int32_t list = (ep_cap + 8) * (big ? 8 : 4);
    l.off_ep = 0;
    l.off_ep2 = list;
    l.off_oidx = 2 * list;
    l.off_mask = (l.off_oidx + (ep_cap + 8) * 2 + 7) & ~7;
    l.off_rem = l.off_mask + ((ep_cap + 63) / 64) * 8;
    l.off_tlist = l.off_rem + max_types * 4;
    l.off_plan = l.off_tlist + max_types * 4;
    l.n_plan = next_pow2(full_items > 0 ? full_items : 1);
    l.off_rel = (l.off_plan + l.n_plan * 4 + 15) & ~15;
    const int32_t pass_end = l.off_rel + n_waves * (64 + 8) * 32;
    l.n_key = next_pow2(box_cap > 0 ? box_cap : 1);
    l.off_load = 0;
    l.off_key = box_cap * 8;
    l.off_tot = l.off_key + l.n_key * 8;
    l.off_items = (l.off_tot + box_cap * (big ? 8 : 4) + 15) & ~15;
    const int32_t fsr_end = l.off_items + box_cap * 32;
    const int32_t region = pass_end > fsr_end ? pass_end : fsr_end;
    l.off_box = (region + 15) & ~15;
    l.off_mt = l.off_box + (box_cap + 8) * 32;
    // the instance's type records, volume order and greedy-plan entries
    // (one per item: box_cap >= the item count), staged per decoder
    l.off_types = l.off_mt + (use_rng ? 624 * 4 : 0);
    l.off_vord = l.off_types + max_types * (int32_t)sizeof(BlockTypeRec);
    l.off_gtab = l.off_vord + max_types * 4;
    l.total = l.off_gtab + full_items * (int32_t)sizeof(GreedyEntry);
    return l;
}

}  // namespace plpk
