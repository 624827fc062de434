// kernels.hip -- gfx950 kernels of the pallet-loading decoder.
//
// Arithmetic contract (gpu/Makefile:6-10, gpu/plp_common.hpp:9-15): integer
// geometry stays integer; the only floating point is IEEE-754 binary64 in the
// reference's operand order (C3 ratio, random(), FSR), built with
// -ffp-contract=off and no fast-math, so no multiply-add is fused.
//
// Kernels:
//   feas_kernel        one lane per (candidate, orientation) pair of one step
//                      (plain form, plp_gpu.cu:74-88)
//   feas_batch_kernel  the same for K rows (plp_gpu.cu:123-144)
//   decode_kernel      persistent: one wavefront decodes one (instance, seed)
//                      for whole passes, then takes the next from a work queue
//                      (plp_gpu.cu:701-1095 + stability_kernel 1121-1221 +
//                      the D4' selection of run_device 1438-1473)
//   fsr_kernel         FSR + stability score of given packings
//   rng_script_kernel  device MT19937 probe (plp_gpu.cu:642-661)
#include "kernels.hpp"

#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <utility>

#include "decoder_dev.hpp"
#include "engine_types.hpp"
#include "mt19937.hpp"
#include "point_state.hpp"
#include "wave.hpp"

namespace plpk {

// ------------------------------------------------------------ decoder ----

// One evaluated step as wave 0 publishes it: the step index (-1: the pass is
// over), its type | n_ors << 24, whether the previous evaluated step had the
// same type, and the live points and boxes it starts from.
struct alignas(16) StepDesc {
    int s, sv, inc, n_act, n_box, pad0, pad1, pad2;
};

// Block-wide scalars.  Wave 0 writes step before the barrier that opens the
// step (the one ending the previous step); every wave reads it before the
// step's verdict barrier, and the next write comes after that barrier.
template <class Sum>
struct StepShared {
    alignas(16) int cnt[2][8];   // feasible pairs found by each wave, by evaluated-step
                                 // parity: a wave may start step k+1 while others
                                 // still read step k
    StepDesc step;  // the step being evaluated, published by wave 0
    int item;       // queue item of the block
    int quit;       // the decoder stopped (capacity error): leave the pass loop
    int nb[6];      // the box just placed: x, y, z, x2, y2, z2
    Sum cstate[48];        // pair states of its 8 corners x this type's orientations
    uint16_t cslot[2][8];  // point states: slots for the corners of the box placed
                           // with n_box even / odd
    int nfree;             // point states: free slots on the stack
    int ins_post;          // point states: 0x100 | the corners wave 0's merge inserts,
                           // posted by it, taken and reset to 0 by the slot wave
    int rs_n, rs_box, rs_s;   // a pass resumed from a trajectory: its list length, boxes
                              // and step (rs_n 0: the pass starts from the origin)
};

// ------------------------------------------------------ shared prefixes --

// The largest w, d and h over the steps from s on (step s's type and every
// later step's): the reach of the point states that serve step s.
__device__ __forceinline__ void largest_from(const DecodeArgs& a, const InstHdr& I, int s, int& WM,
                                             int& DM, int& HM) {
    const int32_t* mn = a.step_min + 8 * (size_t)(I.step_off + s);
    WM = mn[4];
    DM = mn[5];
    HM = mn[6];
    const int t = a.step_type[I.step_off + s];
    const int32_t* orow = a.ors_tab + (size_t)(I.ors_off + t) * kOrsRow;
    const int no = orow[0];
    for (int o = 0; o < no; o++) {
        WM = max(WM, orow[1 + 3 * o]);
        DM = max(DM, orow[2 + 3 * o]);
        HM = max(HM, orow[3 + 3 * o]);
    }
}

// Wave 0 at the start of a pass of a decoder whose instance has a recorded
// trajectory (engine_types.hpp, TrajHdr).  Replays the decoder's RNG along
// the trajectory's placing steps: step j draws random() -- from MT words
// mti + 2j and mti + 2j + 1 while no regeneration falls between, so up to 64
// steps are tested at once, lane k taking step j + k -- and, when it is at
// least 1 - eps, randrange(cnt_j); the first index other than 0 is where the
// decoder leaves the trajectory.  Then either
//   * it never does: the pass is the trajectory's (traj_pass), or
//   * it does at placing step j: step j's list goes to ep_a, boxes 0 .. j-1
//     to box[] (with their kSupports marks as of step j), n_act / n_box are
//     set, the step's pick is forced to the drawn index, and the step is
//     returned (loaded set);
// unless a capacity of this decoder would have stopped the pass on the way
// (its list or its boxes outgrow the LDS), which sets quit (and box_over) as
// the decode itself would.  Returns the step to resume at, or -1 when the
// pass does not decode.  An instance without a valid trajectory decodes from
// the origin (step 0, nothing loaded).
template <class Key>
__device__ int resume_from_trajectory(const DecodeArgs& a, int inst_i, uint32_t* mt, int& mti,
                                      Box* box, Key* ep_a, TrajHdr& th, bool& traj_pass,
                                      int& forced, bool& loaded, int& n_act, int& n_box,
                                      uint32_t& ep_need, bool& box_over, int& quit) {
    const int lane = lane_id();
    wave_sync();   // after the pass's list and box initialisation
    loaded = false;
    const TrajHdr h = a.traj_hdr[inst_i];
    th.valid = uni(h.valid);
    th.n_box = uni(h.n_box);
    th.need = uni(h.need);
    th.volume = uni(h.volume);
    th.score = __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(h.score)));
    if (!th.valid) return 0;
    const TrajStep* ts = a.traj_step + (size_t)inst_i * a.step_cap;
    const int nd = th.n_box;
    int j = 0, idx = 0;
    bool div = false;
    if (a.use_rng) {
        while (j < nd) {
            const int room = (624 - mti) >> 1;   // random() calls before a regeneration
            if (room == 0) {
                if (mt_random(mt, mti) >= 1.0 - a.eps) {
                    idx = mt_randrange(mt, mti, uni(ts[j].cnt));
                    if (idx != 0) {
                        div = true;
                        break;
                    }
                }
                j++;
                continue;
            }
            const int w = min(min(room, kWave), nd - j);
            bool hit = false;
            if (lane < w) {
                const uint32_t hi = mt_temper(mt[mti + 2 * lane]) >> 5;
                const uint32_t lo = mt_temper(mt[mti + 2 * lane + 1]) >> 6;
                hit = ((double)hi * 67108864.0 + (double)lo) * (1.0 / 9007199254740992.0) >=
                      1.0 - a.eps;
            }
            const uint64_t hm = ballot(hit);
            if (hm == 0) {
                mti += 2 * w;
                j += w;
                continue;
            }
            const int k = __ffsll((unsigned long long)hm) - 1;
            mti += 2 * (k + 1);
            j += k;
            idx = mt_randrange(mt, mti, uni(ts[j].cnt));
            if (idx != 0) {
                div = true;
                break;
            }
            j++;
        }
    }
    if (!div) {
        ep_need = max(ep_need, (uint32_t)th.need);
        n_box = 0;
        if (th.need > a.ep_cap) {
            quit = 1;
        } else if (nd > a.box_cap) {
            box_over = true;
            quit = 1;
        } else {
            traj_pass = true;
        }
        return -1;
    }
    const int need_j = uni(ts[j].need), na = uni(ts[j].n_act), koff = uni(ts[j].key_off);
    const int s_j = uni(ts[j].s);
    ep_need = max(ep_need, (uint32_t)need_j);
    if (need_j > a.ep_cap || koff < 0) {
        quit = 1;
        return -1;
    }
    if (j >= a.box_cap) {
        box_over = true;
        quit = 1;
        return -1;
    }
    const Key* tk = (const Key*)a.traj_keys + (size_t)inst_i * a.traj_key_cap + koff;
    for (int i = lane; i < na; i += kWave) ep_a[i] = tk[i];
    const int32_t* tb = a.traj_box + (size_t)inst_i * a.step_cap * 8;
    for (int i = lane; i < j; i += kWave) {
        const int4 p = *(const int4*)(tb + 8 * i);
        const int4 q = *(const int4*)(tb + 8 * i + 4);
        Box B;
        B.x = p.x; B.y = p.y; B.z = p.z; B.x2 = p.w; B.y2 = q.x; B.z2 = q.y;
        B.type = q.z;
        B.pad = q.w < j ? kSupports : 0;   // a box of the first j rests on it
        box[i] = B;
    }
    n_act = na;
    n_box = j;
    forced = idx;
    loaded = true;
    wave_sync();
    return s_j;
}

// The rows (JSON layout) of n boxes of a recorded trajectory.
__device__ __forceinline__ void write_rows_traj(int32_t* dst, const int32_t* tb, int n) {
    for (int t = lane_id(); t < 7 * n; t += kWave) {
        const int i = t / 7, c = t - (t / 7) * 7;
        const int32_t* b = tb + 8 * i;
        int v;
        switch (c) {
            case 0: v = b[6]; break;
            case 1: v = b[0]; break;
            case 2: v = b[1]; break;
            case 3: v = b[2]; break;
            case 4: v = b[3] - b[0]; break;
            case 5: v = b[4] - b[1]; break;
            default: v = b[5] - b[2]; break;
        }
        dst[t] = v;
    }
}

// Persistent decoder: one block of NW waves decodes one (instance, seed)
// for n_passes passes (plp_common.hpp:404-456 per pass, FSR and the D4'
// selection after each), then takes the next queue item.
//
// Extreme points are packed (z, y, x) keys.  Only the live ones are kept,
// sorted and unique: a point beyond a far wall or inside the half-open extent
// of a placed box fails C2 or C1 for every box and orientation from then on
// (boxes never move), so dropping it removes only infeasible pairs and pair
// p = e * n_ors + o still enumerates the reference's candidate-major feas
// order (plp_common.hpp:419-435) restricted to the pairs that can be
// feasible.  The live list alternates between two LDS buffers, one merge
// reading one and writing the other.  With count_work the dead points are
// kept too (unsorted) so that the reference's list size, and from it its
// brute-force work, can be reported.
// Every wave evaluates 64-pair chunks c = wave, wave + NW, ...; each chunk's
// verdicts become one ballot mask in LDS and the pick is the idx-th set bit.
// Wave 0 alone walks the step table (memo included) and publishes each
// evaluated step, and runs the serial part of a step (RNG, pick, place,
// extreme-point merge), FSR and the selection; the other waves follow the
// published steps and wait at the block barriers.
// The bound asks for at most 128 VGPRs (four waves per SIMD, i.e. four
// 4-wave decoders per CU).  M (StateMode, a.mode) sets the widths of the
// extreme-point keys, support sums and pair-state cache entries.
//
// V (VerdictKind) chooses how a step's pair verdicts are reached.
// kPairCache: the pair states of the last two evaluated steps are kept; a
// step of the same type as the previous one combines them with the new box,
// and any other step scans the boxes near each 64-pair chunk.  kPointState:
// every live point carries an orientation-free state (point_state.hpp) in a
// slot of its own, which the helper wave brings up to date with each new box
// while wave 0 merges the lists, and builds for the new corners; a step of a
// new type reads its verdicts from the states, and a step that repeats the
// type of the step before it (which placed a box) gathers the verdict bits
// each slot was given during that merge: the previous bits for the points
// the new box does not touch, fresh ones for the touched points and the
// corners.
template <int NW, int M, int V, bool F>
__global__ void __launch_bounds__(64 * NW, 4) decode_kernel(DecodeArgs a) {
    // F: the full kernel, with the work count (count_work) and the shared
    // prefixes' recording and resumption; without F they are compiled out
    const bool cw = F && a.count_work != 0;
    const int traj = F ? a.traj : 0;
    using Key = typename ModeTypes<M>::Key;
    using Sum = typename ModeTypes<M>::Sum;
    using Ent = typename PtsEntOf<M>::T;
    constexpr bool PS = V == kPointState;
    constexpr Sum kBlk = kBlockedOf<Sum>;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_ors[18 * NW];
    __shared__ StepShared<Sum> sh;
    __shared__ DecState sst;
    const int lane = lane_id();
    const int wid = (int)(threadIdx.x >> 6);
    Key *ep_a, *ep_b, *dead;
    uint16_t* oidx = nullptr;
    unsigned char* cache = nullptr;
    int cache_bytes = 0;
    Box* cbox = nullptr;
    uint16_t *sl_a = nullptr, *sl_b = nullptr, *fstack = nullptr;
    Pts<Ent>* pts = nullptr;
    uint8_t* vbits = nullptr;
    uint64_t* mask;
    double* load;
    uint64_t* skey;
    Sum* stot;
    Box* box;
    uint32_t* mt;
    int32_t* s_step;
    if constexpr (PS) {
        const PtsLayout lay = pts_layout(a.ep_cap, a.box_cap, a.use_rng, cw, M,
                                         a.step_cap);
        ep_a = (Key*)(lds + lay.off_ep);
        ep_b = (Key*)(lds + lay.off_ep2);
        dead = (Key*)(lds + lay.off_dead);
        sl_a = (uint16_t*)(lds + lay.off_sl);
        sl_b = (uint16_t*)(lds + lay.off_sl2);
        fstack = (uint16_t*)(lds + lay.off_free);
        pts = (Pts<Ent>*)(lds + lay.off_pts);
        vbits = lds + lay.off_bits;
        mask = (uint64_t*)(lds + lay.off_mask);
        load = (double*)(lds + lay.off_load);
        skey = (uint64_t*)(lds + lay.off_key);
        stot = (Sum*)(lds + lay.off_tot);
        box = (Box*)(lds + lay.off_box);
        mt = (uint32_t*)(lds + lay.off_mt);
        s_step = (int32_t*)(lds + lay.off_steps);
    } else {
        const LdsLayout lay = lds_layout(a.ep_cap, a.box_cap, a.use_rng, NW, cw, M,
                                         a.step_cap);
        ep_a = (Key*)(lds + lay.off_ep);
        ep_b = (Key*)(lds + lay.off_ep2);
        dead = (Key*)(lds + lay.off_dead);
        oidx = (uint16_t*)(lds + lay.off_oidx);
        cache = lds + lay.off_cache;
        cache_bytes = lay.n_cache * cache_entry_bytes(M);
        cbox = (Box*)(lds + lay.off_rel) + wid * kCullWindow;   // this wave's culled window
        mask = (uint64_t*)(lds + lay.off_mask);
        load = (double*)(lds + lay.off_load);
        skey = (uint64_t*)(lds + lay.off_key);
        stot = (Sum*)(lds + lay.off_tot);
        box = (Box*)(lds + lay.off_box);
        mt = (uint32_t*)(lds + lay.off_mt);
        s_step = (int32_t*)(lds + lay.off_steps);
    }
    const int orw = 18 * wid;   // this wave's orientations: s_ors[orw ..]

    if (threadIdx.x == 0) sh.ins_post = 0;
    for (;;) {
        if (threadIdx.x == 0) sh.item = (int)atomicAdd(a.work_counter, 1u);
        block_sync();
        const int item = uni(sh.item);
        block_sync();
        if (item >= a.n_work) break;
        const int d = uni(a.work_order[item]);
        const uint64_t t_item = __builtin_amdgcn_s_memrealtime();
        // recording trajectories, the queue items are instances
        const int inst_i = traj == 1 ? d : uni(a.dec_inst[d]);
        const InstHdr I = a.inst[inst_i];
        const int L = I.L, W = I.W, H = I.H, xb = I.xbits, yb = I.ybits;
        const int32_t* steps = a.step_type + I.step_off;
        // Step table in LDS: type | n_ors << 24 (types < 2^24 at load time).
        for (int i = threadIdx.x; i < I.n_steps; i += 64 * NW) {
            const int t = steps[i];
            s_step[i] = t | (a.ors_tab[(size_t)(I.ors_off + t) * kOrsRow] << 24);
        }
        int cur_ti = -1;   // type whose orientations this wave holds in s_ors[orw ..]

        // ---- decoder state: in LDS (sst) between pass ends; wave 0
        int mti = 624;
        uint32_t* gmt = a.mt + (size_t)d * kMtWords;
        if (wid == 0) {
            if (lane == 0) {
                DecState st0;
                if (a.reset) {
                    st0.passes = 0; st0.err = 0; st0.n_all = -1; st0.n_acc = -1;
                    st0.v_all = 0; st0.v_acc = 0; st0.s_all = 0.0; st0.s_acc = 0.0;
                    st0.work_pair_box = 0; st0.work_live = 0; st0.work_steps = 0; st0.max_ep = 0;
                    st0.ep_need = 0; st0.box_need = 0;
                } else {
                    st0 = a.state[d];
                }
                sst = st0;
                sh.quit = st0.err != 0;
            }
            wave_sync();
            if (uni(sh.quit) == 0 && a.use_rng) {
                if (a.reset) {
                    // the state seeded once at load (mt_seed_kernel)
                    const uint32_t* g0 = a.mt0 + (size_t)d * kMtWords;
                    for (int i = lane; i < 624; i += kWave) mt[i] = g0[i];
                } else {
                    for (int i = lane; i < 624; i += kWave) mt[i] = gmt[i];
                    mti = uni((int)gmt[624]);
                }
                wave_sync();
            }
        }
        block_sync();
        const bool skip_decoder = uni(sh.quit) != 0;
        int32_t* out_sel = a.out_rows + (size_t)d * a.max_items * 7;

        for (int pass = 0; pass < a.n_passes && !skip_decoder; pass++) {
            // ---- pass start (wave 0): origin only, no boxes
            int n_act = 1, n_dead = 0, n_box = 0;
            int fail_ti = -1;   // type whose last copy found nothing since the last placement
            uint64_t work = 0, live = 0;
            uint32_t n_steps_eval = 0, max_ep = 1, ep_need = 1;
            bool box_over = false;   // wave 0: the pass ran out of box capacity
            if (wid == 0) {
                if (lane == 0) {
                    ep_a[0] = Key(0);   // origin; L, W, H > 0 so it is live
                    if constexpr (PS) {
                        sl_a[0] = 0;
                        pts[0] = pts_empty<Ent>();
                    } else {
                        oidx[0] = kNewPoint;
                    }
                }
                if constexpr (PS) {
                    // slot 0 holds the origin, 1 .. 8 wait for the first box's
                    // corners, 9 .. ep_cap + 7 are free
                    for (int i = lane; i < a.ep_cap - 1; i += kWave) fstack[i] = (uint16_t)(i + 9);
                    if (lane < 8) sh.cslot[0][lane] = (uint16_t)(lane + 1);
                    if (lane == 0) sh.nfree = a.ep_cap - 1;
                }
                for (int i = lane; i < a.box_cap + 8; i += kWave) box[i] = sentinel_box();
            }
            // Wave 0 walks the step sequence (the memo decision and the step
            // table need no other wave) and, before the barrier that opens
            // an evaluated step, publishes it in sh.step; the other waves
            // only follow the published steps.
            int ev = 0;     // evaluated steps so far (same count in every wave)
            int cur = 0;    // live list in use: 0 -> ep_a, 1 -> ep_b (every wave)
            int s = 0;          // wave 0: the step being evaluated
            int prev_ti = -1;   // wave 0: type of the previous evaluated step
            // Wave 0: the next step from s0 on that is evaluated (a type with
            // orientations, not skipped by the memo), published with the list
            // and box counts; -1 ends the pass.  The memo: same type, no
            // placement since its last copy found nothing -- the verdicts
            // repeat and nothing is drawn, for every further copy of the type
            // that follows, skipped together.
            auto advance = [&](int s0, bool stop) {
                s = s0;
                int sv = 0;
                while (!stop && s < I.n_steps) {
                    sv = uni(s_step[s]);
                    const int ti = sv & 0xffffff, no = sv >> 24;
                    if (no == 0) {      // type with no allowed orientation
                        s++;
                        continue;
                    }
                    if (cw)
                        work += (uint64_t)(n_act + n_dead) * (uint64_t)no * (uint64_t)n_box;
                    n_steps_eval++;
                    if (ti != fail_ti) break;
                    int run = 1;
                    for (;;) {
                        const int q = s + run + lane;
                        const bool same = (q < I.n_steps) && (s_step[q < I.n_steps ? q : s] == sv);
                        const uint64_t other = ballot(!same);
                        if (other) {
                            run += __ffsll((unsigned long long)other) - 1;
                            break;
                        }
                        run += kWave;
                    }
                    if (cw)
                        work += (uint64_t)(run - 1) * (uint64_t)(n_act + n_dead) *
                                (uint64_t)no * (uint64_t)n_box;
                    n_steps_eval += (uint32_t)(run - 1);
                    s += run;
                }
                const bool more = !stop && s < I.n_steps;
                const int ti = sv & 0xffffff;
                if (lane == 0) {
                    StepDesc sd;
                    sd.s = more ? s : -1;
                    sd.sv = sv;
                    sd.inc = more && ti == prev_ti;
                    sd.n_act = n_act;
                    sd.n_box = n_box;
                    sh.step = sd;
                }
                if (more) prev_ti = ti;
            };
            // ---- shared prefix (decoders of a recorded trajectory)
            int forced = -1;          // wave 0: the pick of the step the pass resumes at
            bool traj_pass = false;   // wave 0: the pass is the trajectory's, whole
            int koff = 0;             // wave 0, recording: keys kept so far
            if (traj == 2) {
                if (wid == 0) {
                    int quit0 = 0;
                    bool loaded = false;
                    TrajHdr th;
                    const int rs_s = resume_from_trajectory<Key>(
                        a, inst_i, mt, mti, box, ep_a, th, traj_pass, forced, loaded, n_act,
                        n_box, ep_need, box_over, quit0);
                    if (lane == 0) {
                        sh.rs_n = loaded ? n_act : 0;
                        sh.rs_box = n_box;
                        sh.rs_s = rs_s;
                        sh.quit = quit0;
                    }
                }
                block_sync();
                if constexpr (PS) {
                    // the resumed list's states, built from its boxes: slot
                    // i for entry i, this box's corners next, the rest free
                    const int rn = uni(sh.rs_n);
                    if (rn > 0) {
                        const int rb = uni(sh.rs_box), rss = uni(sh.rs_s);
                        int WM, DM, HM;
                        largest_from(a, I, rss, WM, DM, HM);
                        for (int i = threadIdx.x; i < rn; i += 64 * NW) {
                            int px, py, pz;
                            key_xyz(ep_a[i], xb, yb, px, py, pz);
                            Pts<Ent> ps = pts_empty<Ent>();
                            for (int b = 0; b < rb; b++) {
                                const Box B = box_xyz(box[b]);
                                if (pts_touches(px, py, pz, B, WM, DM, HM))
                                    pts_absorb(ps, px, py, pz, B, WM, DM, HM);
                            }
                            pts[i] = ps;
                            sl_a[i] = (uint16_t)i;
                        }
                        if (wid == 0) {
                            for (int i = lane; i < a.ep_cap - rn; i += kWave)
                                fstack[i] = (uint16_t)(rn + 8 + i);
                            if (lane < 8) sh.cslot[rb & 1][lane] = (uint16_t)(rn + lane);
                            if (lane == 0) sh.nfree = a.ep_cap - rn;
                        }
                    }
                }
                if (wid == 0) {
                    const int rss = uni(sh.rs_s);
                    if (rss >= 0) advance(rss, false);
                    else advance(I.n_steps, true);   // the trajectory's pass, or a capacity stop
                }
            } else if (wid == 0) {
                advance(0, false);
            }
            block_sync();
            for (;;) {
                const StepDesc sd = sh.step;
                const int s_cur = uni(sd.s);
                if (s_cur < 0) break;
                const int sv = uni(sd.sv);
                const int ti = sv & 0xffffff, no = sv >> 24;
                const bool inc = uni(sd.inc) != 0;
                const int na_w = uni(sd.n_act), nb_w = uni(sd.n_box);
                int wm = 0, dm = 0, hm = 0;
                if (wid == 0) {
                    if (cw) live += (uint64_t)n_act * (uint64_t)no * (uint64_t)n_box;
                    const int32_t* mn = a.step_min + 8 * (size_t)(I.step_off + s_cur);
                    wm = mn[0];
                    dm = mn[1];
                    hm = mn[2];
                }
                if (ti != cur_ti) {
                    const int32_t* orow = a.ors_tab + (size_t)(I.ors_off + ti) * kOrsRow;
                    if (lane < 3 * no) s_ors[orw + lane] = orow[1 + lane];
                    cur_ti = ti;
                }
                wave_sync();
                const Key* act = cur ? ep_b : ep_a;
                const uint16_t* sl_act = cur ? sl_b : sl_a;

                // ---- pair verdicts (every wave).  After a placement of
                // the same type, a pair of a point that was in the previous
                // list takes its cached state and tests the new box alone;
                // the other pairs scan the boxes.
                const unsigned char* c_rd = cache + ((ev + 1) & 1) * cache_bytes;
                unsigned char* c_wr = cache + (ev & 1) * cache_bytes;
                int nx = 0, ny = 0, nz = 0, nx2 = -1, ny2 = -1, nz2 = -1;
                if (inc) {
                    const Box NB = box[nb_w - 1];
                    nx = uni(NB.x); ny = uni(NB.y); nz = uni(NB.z);
                    nx2 = uni(NB.x2); ny2 = uni(NB.y2); nz2 = uni(NB.z2);
                }
                const int n_pairs = na_w * no;
                const uint32_t magic = orient_recip(no);
                const int n_chunks = (n_pairs + kWave - 1) / kWave;
                int cnt_w = 0;
                if constexpr (PS) {
                    if (inc) {
                        // The previous step had this type and placed a box:
                        // every point's verdict bits for this step were left
                        // in vbits[slot] while the lists were merged.
                        for (int c = wid; c < n_chunks; c += NW) {
                            const int p = c * kWave + lane;
                            const bool valid = p < n_pairs;
                            const int pv = valid ? p : 0;
                            const int e = div_orients(pv, no, magic);
                            const int o = pv - __mul24(e, no);
                            const uint32_t bits = vbits[sl_act[e]];
                            const uint64_t m = ballot(valid & (((bits >> o) & 1u) != 0));
                            mask[c] = m;
                            cnt_w += __popcll(m);
                        }
                    } else {
                        for (int c = wid; c < n_chunks; c += NW) {
                            const int p = c * kWave + lane;
                            const bool valid = p < n_pairs;
                            const int pv = valid ? p : 0;
                            const int e = div_orients(pv, no, magic);
                            const int o = pv - __mul24(e, no);
                            int x, y, z;
                            key_xyz(act[e], xb, yb, x, y, z);
                            const Pts<Ent> ps = pts[sl_act[e]];
                            const int w = s_ors[orw + 3 * o], dd = s_ors[orw + 3 * o + 1],
                                      h = s_ors[orw + 3 * o + 2];
                            const uint64_t m = pts_verdicts<Sum>(ps, valid, x, y, z, w, dd, h, L, W,
                                                                 H, box, nb_w);
                            mask[c] = m;
                            cnt_w += __popcll(m);
                        }
                    }
                } else
                for (int c = wid; c < n_chunks; c += NW) {
                    const int p = c * kWave + lane;
                    const bool valid = p < n_pairs;
                    // lanes past the last pair read pair 0 (in bounds) and
                    // are masked out of every result
                    const int pv = valid ? p : 0;
                    const int e = div_orients(pv, no, magic);
                    const int o = pv - __mul24(e, no);
                    int x, y, z;
                    key_xyz(act[e], xb, yb, x, y, z);
                    const int w = s_ors[orw + 3 * o], dd = s_ors[orw + 3 * o + 1],
                              h = s_ors[orw + 3 * o + 2];
                    const int oi = oidx[e];   // read with the point (used by same-type steps)
                    asm volatile("" ::"v"(oi));   // keeps the read beside the point's, not in the branch
                    const bool c12 = valid & (x + w <= L) & (y + dd <= W) & (z + h <= H);
                    // settled: the pair's state st (support sum or kBlocked) is
                    // known without a box scan
                    Sum st = kBlk;
                    bool settled = false;
                    if (inc) {
                        // a corner of the new box takes its precomputed state; a
                        // point of the previous list its old state combined with
                        // the new box; both read at in-bounds indices, selected
                        const bool corner = (oi >= kCorner) & (oi != kNewPoint);
                        const bool cached = oi < kCorner;
                        const Sum cs = sh.cstate[corner ? (oi - kCorner) * no + o : 0];
                        const Sum old = cache_get<M>(c_rd, cached ? oi * no + o : 0);
                        const int ox = min(x + w, nx2) - max(x, nx);
                        const int oy = min(y + dd, ny2) - max(y, ny);
                        const int oz = min(z + h, nz2) - max(z, nz);
                        const int oxy = min(ox, oy);
                        const bool ov = min(oxy, oz) > 0;
                        const Sum add = mad_sum((uint32_t)(((nz2 == z) & (oxy > 0)) ? ox : 0),
                                                (uint32_t)oy, Sum(0));
                        st = corner ? cs : (((old == kBlk) | ov) ? kBlk : old + add);
                        settled = c12 & (corner | cached);
                    }
                    const bool scan = c12 & !settled;
                    // Settled lanes first, so that nothing of theirs stays
                    // live across the scan; a lane failing C2 is blocked.
                    uint64_t m = ballot(settled & pair_verdict(st != kBlk, z, w, dd, st));
                    if (valid && !scan) cache_put<M>(c_wr, p, settled ? st : kBlk);
                    if (ballot(scan) != 0) {
                        const int xlo = wave_min(scan ? x : 0x7fffffff);
                        const int xhi = wave_max(scan ? x + w : -1);
                        const int ylo = wave_min(scan ? y : 0x7fffffff);
                        const int yhi = wave_max(scan ? y + dd : -1);
                        const int zlo = wave_min(scan ? z : 0x7fffffff);
                        const int zhi = wave_max(scan ? z + h : -1);
                        Sum s2 = 0;
                        const bool a2 = pair_scan(x, y, z, w, dd, h, scan, box, nb_w, xlo, xhi, ylo,
                                                  yhi, zlo, zhi, cbox, s2);
                        if (scan) cache_put<M>(c_wr, p, a2 ? s2 : kBlk);
                        m |= ballot(scan & pair_verdict(a2, z, w, dd, s2));
                    }
                    mask[c] = m;   // every lane writes the same (uniform) mask
                    cnt_w += __popcll(m);
                }
                const int par = ev & 1;
                ev++;
                sh.cnt[par][wid] = cnt_w;   // every lane: the same value, no exec mask change
                block_sync();
                // wave 0 reads its next random() words and the first mask
                // with the counts: one LDS round trip for the three
                // (lane l of wave 0 holds mask word l of the first 64 chunks)
                MtPeek pk{0u, 0u};
                uint64_t mw = 0;
                if (wid == 0) {
                    if (a.use_rng) pk = mt_peek2(mt, mti);
                    if (lane < n_chunks) mw = mask[lane];
                }
                const int cnt = uni(sum_counts<NW>(sh.cnt[par]));
                if (cnt == 0) {
                    // this copy is skipped; no RNG use, nothing changed
                    if (wid == 0) {
                        fail_ti = ti;
                        advance(s_cur + 1, false);
                    }
                    block_sync();
                    continue;
                }
                fail_ti = -1;
                if (nb_w >= a.box_cap) {
                    // one box past the LDS box capacity: this decoder stops and
                    // the host reruns it with room for every item
                    if (wid == 0) {
                        box_over = true;
                        if (lane == 0) sh.quit = 1;
                        advance(s_cur + 1, true);
                    }
                    block_sync();
                    continue;
                }

                // ---- pick and place (wave 0)
                int x = 0, y = 0, z = 0, X2 = 0, Y2 = 0, Z2 = 0;
                if (wid == 0) {
                    // feas[0], or feas[randrange(|feas|)] when random() >=
                    // 1 - eps (plp_common.hpp:427-435).
                    int idx = 0;
                    if (forced >= 0) {
                        // drawn while replaying the trajectory (this step's count
                        // equals the trajectory's there; the bound only keeps the
                        // walk below within the masks)
                        idx = min(forced, cnt - 1);
                        forced = -1;
                    } else if (a.use_rng) {
                        if (mt_random(mt, mti, pk) >= 1.0 - a.eps) idx = mt_randrange(mt, mti, cnt);
                    }
                    // The idx-th feasible pair: for idx 0 (every step that
                    // draws no eps pick) the first set bit of the first
                    // non-zero word the lanes hold, else a walk over the words.
                    const uint64_t nz = ballot(mw != 0);
                    int c = 0, bit;
                    uint64_t m;
                    if (idx == 0 && nz != 0) {
                        c = __ffsll((unsigned long long)nz) - 1;
                        m = read_lane(mw, c);
                        bit = __ffsll((unsigned long long)m) - 1;
                    } else {
                        m = read_lane(mw, 0);
                        for (;;) {
                            const int pc = __popcll(m);
                            if (idx < pc) break;
                            idx -= pc;
                            c++;
                            m = c < kWave ? read_lane(mw, c) : uni(mask[c]);
                        }
                        const bool hit = ((m >> lane) & 1ull) && rank_below(m) == idx;
                        bit = __ffsll((unsigned long long)ballot(hit)) - 1;
                    }
                    const int p = c * kWave + bit;
                    const int e = div_orients(p, no, magic);
                    const int o = p - __mul24(e, no);
                    key_xyz(uni(act[e]), xb, yb, x, y, z);
                    x = uni(x);
                    y = uni(y);
                    z = uni(z);
                    X2 = x + uni(s_ors[orw + 3 * o]);
                    Y2 = y + uni(s_ors[orw + 3 * o + 1]);
                    Z2 = z + uni(s_ors[orw + 3 * o + 2]);
                    if (traj == 1) {
                        // recording: placing step n_box -- its step, pair count,
                        // list and the largest list held before it (a list the
                        // decoders' capacity cannot hold is not kept)
                        TrajStep* ts = a.traj_step + (size_t)inst_i * a.step_cap + n_box;
                        const bool keep =
                            na_w <= a.traj_list_cap && koff + na_w <= a.traj_key_cap;
                        if (lane == 0) {
                            ts->s = s_cur;
                            ts->cnt = cnt;
                            ts->n_act = na_w;
                            ts->need = (int32_t)ep_need;
                            ts->key_off = keep ? koff : -1;
                        }
                        if (keep) {
                            Key* tk = (Key*)a.traj_keys + (size_t)inst_i * a.traj_key_cap + koff;
                            for (int i = lane; i < na_w; i += kWave) tk[i] = act[i];
                            koff += na_w;
                        }
                    }
                    if (lane == 0) {
                        Box nbx;
                        nbx.x = x; nbx.y = y; nbx.z = z; nbx.x2 = X2; nbx.y2 = Y2; nbx.z2 = Z2;
                        nbx.type = ti; nbx.pad = 0;
                        box[n_box] = nbx;
                        sh.nb[0] = x; sh.nb[1] = y; sh.nb[2] = z;
                        sh.nb[3] = X2; sh.nb[4] = Y2; sh.nb[5] = Z2;
                    }
                    n_box++;
                    wave_sync();
                }
                if (NW > 1) block_sync();   // every wave sees the new box
                bool quit = false;
                uint32_t ins_m = 0;   // wave 0: the corners its merge inserted
                bool late_post = false;   // the slot wave stopped waiting for wave 0's post
                // ---- extreme-point merge (wave 0) and, meanwhile, the pair
                // states of the new box's corners for this type (last wave),
                // which the next step takes when it has the same type.
                if (wid == 0) {
                    SlotMerge smg;
                    SlotMerge* smp = nullptr;
                    if constexpr (PS) {
                        smg.src = sl_act;
                        smg.dst = cur ? sl_a : sl_b;
                        smg.cslot = sh.cslot[nb_w & 1][lane & 7];
                        smp = &smg;
                    }
                    const int need = merge_points(act, cur ? ep_a : ep_b, oidx, dead, n_act, n_dead,
                                                  box, n_box, x, y, z, X2, Y2, Z2, L, W, H, xb, yb,
                                                  a.ep_cap, cw, uni(wm), uni(dm),
                                                  uni(hm), smp, &ins_m,
                                                  (PS && NW > 1) ? &sh.ins_post : nullptr);
                    max_ep = max(max_ep, (uint32_t)(n_act + n_dead));
                    ep_need = max(ep_need, (uint32_t)need);
                    quit = need > a.ep_cap;   // PLP_E_EPCAP
                    if (lane == 0) sh.quit = quit;
                }
                if constexpr (PS) {
                    // One helper wave frees the slots of the points the merge
                    // drops (the merge's own rule: their smallest later box
                    // meets the new box or a wall, or they equal an inserted
                    // corner), brings the other points' states up to date with
                    // the new box, and takes slots for the next box's corners.
                    // The inserted corners' states are built in two halves:
                    // the corners on the box's bottom face (j < 4) by wave 0
                    // after its merge (by wave 1 when NW > 2), those on its
                    // top face by the last wave (which is the helper wave too
                    // when NW <= 2).
                    constexpr int kFreeWave = NW > 2 ? NW - 2 : NW - 1;
                    constexpr int kLoWave = NW > 2 ? 1 : 0;
                    const int32_t* mn = a.step_min + 8 * (size_t)(I.step_off + s_cur);
                    const int wm_s = mn[0], dm_s = mn[1], hm_s = mn[2];
                    const int par = nb_w & 1;
                    // The next step repeats this type (after a placement the
                    // memo skips nothing): the new box leaves the verdicts of
                    // every point it does not touch as they were, so the
                    // next step's bits are this step's for those points and
                    // are evaluated afresh only for the touched points and
                    // the inserted corners, into vbits[slot].
                    const bool same_next = s_cur + 1 < I.n_steps && uni(s_step[s_cur + 1]) == sv;
                    uint32_t ins = ins_m;
                    // the slot wave (when it is not wave 0) takes the inserted
                    // corners from wave 0's merge, which posts them before
                    // merging the list, after its own pass over the list
                    constexpr bool kDefer = NW > 1;
                    if ((wid == NW - 1 && !(kDefer && wid == kFreeWave)) || (NW > 2 && wid == kLoWave))
                        ins = corner_live(sh.nb, box, nb_w + 1, L, W, H, wm_s, dm_s, hm_s);
                    // the states of the inserted corners in sub, and their bits
                    // for the next step when it repeats the type (lane 8j + o:
                    // orientation o at corner j)
                    auto build_corners = [&](uint32_t sub) {
                        pts_corners(pts, sh.cslot[par], sh.nb, box, nb_w + 1, sub, mn[4], mn[5],
                                    mn[6]);
                        if (same_next && sub != 0) {
                            const int j = lane >> 3, o = lane & 7;
                            const bool on = (((sub >> j) & 1u) != 0) & (o < no);
                            const int oo = on ? o : 0;
                            const int csl = sh.cslot[par][j];
                            const Pts<Ent> ps = pts[csl];
                            const uint64_t m = pts_verdicts<Sum>(
                                ps, on, (j & 1) ? sh.nb[3] : sh.nb[0], (j & 2) ? sh.nb[4] : sh.nb[1],
                                (j & 4) ? sh.nb[5] : sh.nb[2], s_ors[orw + 3 * oo],
                                s_ors[orw + 3 * oo + 1], s_ors[orw + 3 * oo + 2], L, W, H, box,
                                nb_w + 1);
                            if (on & (o == 0)) vbits[csl] = (uint8_t)(m >> (lane & 56));
                        }
                    };
                    if (wid == kLoWave) build_corners(ins & 0x0fu);
                    if (wid == kFreeWave) {
                        const int WM = mn[4], DM = mn[5], HM = mn[6];
                        const Box NB = box[nb_w];
                        Key ck = 0;   // lane j < 8: corner j's key
                        if (lane < 8)
                            ck = xyz_key<Key>((lane & 1) ? NB.x2 : NB.x, (lane & 2) ? NB.y2 : NB.y,
                                              (lane & 4) ? NB.z2 : NB.z, xb, yb);
                        Key cks[8];
#pragma unroll
                        for (int j = 0; j < 8; j++) cks[j] = read_lane(ck, j);
                        int nf = uni(sh.nfree);
                        // The touched points' list indices go to the far end of
                        // the free-slot array, downwards: the old list's points
                        // and this box's corner slots hold all but nf of the
                        // ep_cap + 8 slots, so the stack and these never meet.
                        int nt = 0;
                        uint16_t* const touched = fstack + a.ep_cap + 7;
                        // deferred: lane j < 8 keeps the list index of the point
                        // equal to corner j (the corners are distinct, so a point
                        // equals at most one); such a point gives way to the corner
                        // only if it is inserted, known after this loop
                        int eq_at = -1;
                        for (int i0 = 0; i0 < na_w; i0 += kWave) {
                            const int i = i0 + lane;
                            const bool ok = i < na_w;
                            const Key v = act[ok ? i : i0];
                            const int sl = sl_act[ok ? i : i0];
                            int px, py, pz;
                            key_xyz(v, xb, yb, px, py, pz);
                            bool eq = false;
                            if constexpr (kDefer) {
                                uint32_t em = 0;
#pragma unroll
                                for (int j = 0; j < 8; j++) em |= (cks[j] == v) ? (1u << j) : 0u;
                                uint64_t hit = ballot(ok & (em != 0));
                                while (hit != 0) {
                                    const int hl = __ffsll((unsigned long long)hit) - 1;
                                    hit &= hit - 1;
                                    const int hj = __ffs(read_lane(em, hl)) - 1;
                                    if (lane == hj) eq_at = i0 + hl;
                                }
                            } else {
#pragma unroll
                                for (int j = 0; j < 8; j++)
                                    eq |= (((ins >> j) & 1u) != 0) & (cks[j] == v);
                            }
                            const bool gone =
                                (px + wm_s > L) | (py + dm_s > W) | (pz + hm_s > H) |
                                ((px < NB.x2) & (NB.x < px + wm_s) & (py < NB.y2) &
                                 (NB.y < py + dm_s) & (pz < NB.z2) & (NB.z < pz + hm_s));
                            const bool drop = ok & (gone | eq);
                            const uint64_t fm = ballot(drop);
                            if (drop) fstack[nf + rank_below(fm)] = (uint16_t)sl;
                            nf += __popcll(fm);
                            const bool kept = ok & !drop;
                            // a touched point's state takes the new box after the
                            // loop, 64 touched points a pass
                            const bool touch = kept && pts_touches(px, py, pz, NB, WM, DM, HM);
                            if (same_next & kept & !touch)
                                vbits[sl] = (uint8_t)mask_bits(mask, i * no, no);
                            const uint64_t tm = ballot(touch);
                            if (touch) touched[-(nt + rank_below(tm))] = (uint16_t)i;
                            nt += __popcll(tm);
                        }
                        if (nt != 0) {
                            wave_sync();   // the touched indices are in LDS
                            for (int k0 = 0; k0 < nt; k0 += kWave) {
                                const int k = k0 + lane;
                                const bool on = k < nt;
                                const int ti = touched[-(on ? k : k0)];
                                const int tsl = sl_act[ti];
                                int tx, ty, tz;
                                key_xyz(act[ti], xb, yb, tx, ty, tz);
                                if (on) {
                                    Pts<Ent> ps = pts[tsl];
                                    pts_absorb(ps, tx, ty, tz, NB, WM, DM, HM);
                                    pts[tsl] = ps;
                                }
                            }
                        }
                        if constexpr (kDefer) {
                            // the inserted corners from wave 0's post; past a bound
                            // (never reached: wave 0 posts early in its merge, which
                            // it always runs) they are computed here, and the late
                            // post is cleared after the step's barrier
                            int pv = 0;
                            for (int k = 0; k < (1 << 20); k++) {
                                pv = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                                    &sh.ins_post, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
                                if (pv != 0) break;
                                __builtin_amdgcn_s_sleep(1);
                            }
                            if (pv != 0) {
                                ins = (uint32_t)pv & 0xffu;
                                if (lane == 0)
                                    __hip_atomic_store(&sh.ins_post, 0, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            } else {
                                ins = corner_live(sh.nb, box, nb_w + 1, L, W, H, wm_s, dm_s, hm_s);
                                late_post = true;
                            }
                            // a listed point equal to an inserted corner gives way
                            // to it: its slot joins the free stack (the stack, these
                            // at most 8 slots included, stays below the touched list)
                            const bool leave =
                                (lane < 8) && (((ins >> lane) & 1u) != 0) && (eq_at >= 0);
                            const uint64_t lm = ballot(leave);
                            if (leave) fstack[nf + rank_below(lm)] = sl_act[eq_at];
                            nf += __popcll(lm);
                        }
                        // the inserted corners keep their slots; the next box's
                        // corners take new ones in their place (at most ep_cap
                        // points hold a slot unless the list outgrew the
                        // capacity, and then the decoder stops)
                        const int n_ins = __popc(ins);
                        if (nf >= n_ins) {
                            if (lane < 8) {
                                const bool taken = (ins >> lane) & 1u;
                                const int r = __popc(ins & ((1u << lane) - 1u));
                                sh.cslot[par ^ 1][lane] =
                                    taken ? fstack[nf - 1 - r] : sh.cslot[par][lane];
                            }
                            nf -= n_ins;
                        }
                        if (lane == 0) sh.nfree = nf;
                        if (same_next && nt != 0) {
                            // the touched points' bits, eight points at a time:
                            // lane 8k + o evaluates orientation o of the k-th
                            wave_sync();   // their states and indices are in LDS
                            for (int k0 = 0; k0 < nt; k0 += 8) {
                                const int k = k0 + (lane >> 3), o = lane & 7;
                                const bool on = (k < nt) & (o < no);
                                const int ti = touched[-(k < nt ? k : 0)];
                                const int tsl = sl_act[ti];
                                int tx, ty, tz;
                                key_xyz(act[ti], xb, yb, tx, ty, tz);
                                const Pts<Ent> ps = pts[tsl];
                                const int oo = on ? o : 0;
                                const uint64_t m = pts_verdicts<Sum>(
                                    ps, on, tx, ty, tz, s_ors[orw + 3 * oo], s_ors[orw + 3 * oo + 1],
                                    s_ors[orw + 3 * oo + 2], L, W, H, box, nb_w + 1);
                                if (on & (o == 0)) vbits[tsl] = (uint8_t)(m >> (lane & 56));
                            }
                        }
                    }
                    if (wid == NW - 1) build_corners(ins & 0xf0u);
                } else if (wid == NW - 1) {
                    corner_states(sh.nb, sh.cstate, s_ors + orw, no, box, nb_w + 1, L, W, H, cbox);
                }
                if (wid == 0) advance(s_cur + 1, quit);
                block_sync();
                if (late_post && lane == 0) sh.ins_post = 0;
                cur ^= 1;
            }

            // ---- pass end (wave 0): stability score and D4' selection
            if (wid == 0 && traj == 1) {
                // recording: the trajectory's packing, its score, and per box
                // the first box that rests on it (from which a decoder resuming
                // at placing step j knows the boxes marked kSupports then)
                const bool overflow = uni(sh.quit) != 0;
                const int n = n_box;
                int64_t vpart = 0;
                for (int i = lane; i < n; i += kWave)
                    vpart += (int64_t)(box[i].x2 - box[i].x) * (box[i].y2 - box[i].y) *
                             (box[i].z2 - box[i].z);
                const int64_t vol = wave_sum_i64(vpart);
                const double score =
                    (!overflow && n > 0) ? fsr_score(box, n, a.mu, load, skey, stot, nullptr) : 0.0;
                constexpr int32_t kNone = 1 << 30;
                int32_t* tb = a.traj_box + (size_t)inst_i * a.step_cap * 8;
                for (int i0 = 0; i0 < n; i0 += kWave) {
                    const int i = i0 + lane;
                    const bool ok = i < n;
                    const Box Bi = box[ok ? i : 0];
                    int fr = kNone;
                    for (int k = i0 + 1; k < n; k++) {
                        const Box Bk = box_xyz(box[k]);
                        if (ok & (k > i) & (fr == kNone) & (Bi.z2 == Bk.z) & (Bi.x < Bk.x2) &
                            (Bk.x < Bi.x2) & (Bi.y < Bk.y2) & (Bk.y < Bi.y2))
                            fr = k;
                        if ((k & 15) == 0 && ballot(ok & (fr == kNone)) == 0) break;
                    }
                    if (ok) {
                        *(int4*)(tb + 8 * i) = make_int4(Bi.x, Bi.y, Bi.z, Bi.x2);
                        *(int4*)(tb + 8 * i + 4) = make_int4(Bi.y2, Bi.z2, Bi.type, fr);
                    }
                }
                if (lane == 0) {
                    TrajHdr h;
                    h.valid = overflow ? 0 : 1;
                    h.n_box = n;
                    h.need = (int32_t)ep_need;
                    h.pad = 0;
                    h.volume = vol;
                    h.score = score;
                    a.traj_hdr[inst_i] = h;
                }
            } else if (wid == 0) {
                const bool overflow = uni(sh.quit) != 0;
                const bool scored = !overflow && uni(sst.err) == 0;
                // a pass that followed the trajectory to its end has its packing
                int n = n_box;
                int64_t vol;
                double traj_score = 0.0;
                if (traj_pass) {
                    const TrajHdr h = a.traj_hdr[inst_i];
                    n = uni(h.n_box);
                    vol = uni(h.volume);
                    traj_score = __longlong_as_double(
                        (long long)uni((uint64_t)__double_as_longlong(h.score)));
                } else {
                    int64_t vpart = 0;
                    for (int i = lane; i < n; i += kWave)
                        vpart += (int64_t)(box[i].x2 - box[i].x) * (box[i].y2 - box[i].y) *
                                 (box[i].z2 - box[i].z);
                    vol = wave_sum_i64(vpart);
                }
                // D4' (plp_common.hpp:493-511), kept incrementally: the best of
                // all packings and the best accepted one, strict > on (volume,
                // n_boxes) so the earliest winner stays.  A packing that beats
                // neither cannot be selected, so its score is never needed.
                const bool better_all = uni(sst.n_all) < 0 || vol > uni(sst.v_all) ||
                                        (vol == uni(sst.v_all) && n > uni(sst.n_all));
                const bool better_acc = uni(sst.n_acc) < 0 || vol > uni(sst.v_acc) ||
                                        (vol == uni(sst.v_acc) && n > uni(sst.n_acc));
                // Only a packing that would enter the accepted pool can change
                // the selection unless it beats the best of all; when the
                // count of boxes that can reach FSR >= 1 keeps its score below
                // theta, it cannot, and its score is never needed.
                bool need = scored && (better_all || better_acc);
                if (need && !better_all && a.use_theta && n > 0 && !traj_pass &&
                    (double)fsr_count_bound(box, n, a.mu) / (double)n < a.theta)
                    need = false;
                double score = 0.0;
                if (need)
                    score = traj_pass ? traj_score : fsr_score(box, n, a.mu, load, skey, stot, nullptr);
                DecState st = sst;
                if (overflow) st.err = 4;   // PLP_E_EPCAP
                st.work_pair_box += work;
                st.work_live += live;
                st.work_steps += n_steps_eval;
                if (cw) st.max_ep = max(st.max_ep, max_ep);
                st.ep_need = max(st.ep_need, ep_need);
                st.box_need = max(st.box_need, box_over ? (uint32_t)a.box_cap + 1u : (uint32_t)n);
                if (st.err == 0) {
                    const bool accept = better_acc && (!a.use_theta || score >= a.theta);
                    if (better_all) {
                        st.n_all = n; st.v_all = vol; st.s_all = score;
                    }
                    if (accept) {
                        st.n_acc = n; st.v_acc = vol; st.s_acc = score;
                    }
                    // The selection only ever moves to the packing just decoded
                    // (now the best accepted, or the best of all while nothing is
                    // accepted), so the host copy is rewritten from LDS then.
                    if (accept || (st.n_acc < 0 && better_all)) {
                        if (traj_pass)
                            write_rows_traj(out_sel, a.traj_box + (size_t)inst_i * a.step_cap * 8, n);
                        else write_rows(out_sel, box, n);
                    }
                    st.passes++;
                }
                wave_sync();
                if (lane == 0) sst = st;
            }
            block_sync();
            if (uni(sh.quit) != 0) break;
        }
        if (wid == 0 && !skip_decoder && traj != 1) {
            if (a.use_rng) {
                for (int i = lane; i < 624; i += kWave) gmt[i] = mt[i];
                if (lane == 0) gmt[624] = (uint32_t)mti;
            }
            if (lane == 0) a.state[d] = sst;
        }
        if (wid == 0 && lane == 0 && traj != 1) {
            a.out_rec[d] = row_record(sst, a.dec_seed[d]);
            if (a.ticks) a.ticks[d] = (uint32_t)(__builtin_amdgcn_s_memrealtime() - t_item);
        }
        if (wid == 0 && traj != 1 && a.done != nullptr) {
            // wave 0 wrote the record and every packing row: once they have
            // reached host memory, the decoder's completion word
            __threadfence_system();
            if (lane == 0)
                __hip_atomic_store(a.done + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // work_counter[1] counts the decoders that end with an error, so the
        // host reads one word, not every state, when none did
        if (wid == 0 && lane == 0 && sst.err != 0 && traj != 1) atomicAdd(a.work_counter + 1, 1u);
        block_sync();
    }
}

// -------------------------------------------------------- plain forms ----

// One lane per (candidate, orientation) pair of one row; the row's boxes are
// staged in LDS in tiles of kFeasTile.  bits[t] with t = ci * n_ors + oi
// (plp_gpu.cpp:23-29 order).
constexpr int kFeasTile = 1024;
constexpr int kFeasBlock = 256;

__global__ void __launch_bounds__(kFeasBlock) feas_kernel(FeasRowArgs r) {
    __shared__ Box sbox[kFeasTile];
    const int row = r.block_row[blockIdx.x];
    const int first_block = r.row_block0[row];
    const int n_cand = r.n_cand[row];
    const int n_box = r.n_box[row];
    const int n_ors = r.n_ors;
    const int n_pairs = n_cand * n_ors;
    const int p = (blockIdx.x - first_block) * kFeasBlock + threadIdx.x;
    const bool valid = p < n_pairs;
    const int ci = valid ? p / n_ors : 0;
    const int oi = valid ? p - ci * n_ors : 0;
    const int32_t* cand = r.cand + (size_t)r.cand_off[row] * 3;
    const int32_t* aabb = r.aabb + (size_t)r.box_off[row] * 6;
    int x = 0, y = 0, z = 0, w = 1, d = 1, h = 1;
    if (valid) {
        x = cand[3 * ci]; y = cand[3 * ci + 1]; z = cand[3 * ci + 2];
        w = r.ors[3 * oi]; d = r.ors[3 * oi + 1]; h = r.ors[3 * oi + 2];
    }
    const int x2 = x + w, y2 = y + d, z2 = z + h;
    bool alive = valid && x2 <= r.L && y2 <= r.W && z2 <= r.H;
    uint64_t s = 0;   // support area: up to w * d, any container
    for (int t0 = 0; t0 < n_box; t0 += kFeasTile) {
        const int nt = min(kFeasTile, n_box - t0);
        __syncthreads();
        for (int i = threadIdx.x; i < nt; i += kFeasBlock) {
            const int32_t* q = aabb + (size_t)(t0 + i) * 6;
            Box B;
            B.x = q[0]; B.y = q[1]; B.z = q[2]; B.x2 = q[3]; B.y2 = q[4]; B.z2 = q[5];
            B.type = 0; B.pad = 0;
            sbox[i] = B;
        }
        __syncthreads();
        for (int b = 0; b < nt; b++) {
            if ((b & 7) == 0 && ballot(alive) == 0) break;
            const Box B = sbox[b];
            const bool ov = (x < B.x2) & (B.x < x2) & (y < B.y2) & (B.y < y2) & (z < B.z2) & (B.z < z2);
            alive = alive & !ov;
            const int ox = min(x2, B.x2) - max(x, B.x);
            const int oy = min(y2, B.y2) - max(y, B.y);
            const bool sup = (B.z2 == z) & (ox > 0) & (oy > 0);
            s += sup ? (uint64_t)(uint32_t)ox * (uint32_t)oy : 0ull;
        }
    }
    if (alive && z > 0) alive = !((double)s / (double)((int64_t)w * d) < 0.5);
    if (valid) r.bits[(size_t)r.bits_off[row] + p] = alive ? 1 : 0;
}

// FSR of given packings: one wave per packing.
__global__ void __launch_bounds__(64) fsr_kernel(FsrArgs f) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int pk = blockIdx.x;
    const int r0 = f.pack_off[pk], n = f.pack_off[pk + 1] - r0;
    const LdsLayout lay = lds_layout(0, f.box_cap, 0, 1, 0, kBig, 0);   // 64-bit totals
    double* load = (double*)(lds + lay.off_load);
    uint64_t* skey = (uint64_t*)(lds + lay.off_key);
    uint64_t* stot = (uint64_t*)(lds + lay.off_tot);
    Box* box = (Box*)(lds + lay.off_box);
    for (int i = lane_id(); i < n; i += kWave) {
        const int32_t* q = f.rows7 + (size_t)(r0 + i) * 7;
        Box B;
        B.type = q[0]; B.x = q[1]; B.y = q[2]; B.z = q[3];
        B.x2 = q[1] + q[4]; B.y2 = q[2] + q[5]; B.z2 = q[3] + q[6]; B.pad = 0;
        box[i] = B;
    }
    block_sync();
    const double sc = fsr_score<4>(box, n, f.mu, load, skey, stot,
                                f.fsr_out ? f.fsr_out + r0 : nullptr);
    if (lane_id() == 0) f.score_out[pk] = sc;
}

// Device MT19937 probe: one wave per seed runs the op script.
// CPython's seeding of each decoder's MT19937 (init_by_array of the seed's
// 32-bit words: a serial chain of 1247 steps), once per engine load; every
// reset copies the result instead of seeding again.
__global__ void __launch_bounds__(64) mt_seed_kernel(const uint64_t* seeds, uint32_t* out, int n) {
    __shared__ uint32_t mt[624];
    const int d = blockIdx.x;
    if (d >= n) return;
    mt_seed_wave(mt, seeds[d]);
    wave_sync();
    for (int i = lane_id(); i < 624; i += kWave) out[(size_t)d * kMtWords + i] = mt[i];
    if (lane_id() == 0) out[(size_t)d * kMtWords + 624] = 624u;
}

__global__ void __launch_bounds__(64) rng_script_kernel(RngArgs g) {
    __shared__ uint32_t mt[624];
    const int sd = blockIdx.x;
    mt_seed_wave(mt, g.seeds[sd]);
    block_sync();
    int mti = 624;
    for (int i = 0; i < g.n_ops; i++) {
        const int op = g.ops[i], arg = g.args[i];
        double f = 0.0;
        uint32_t u = 0;
        if (op == 0) f = mt_random(mt, mti);
        else if (op == 1) u = mt_getrandbits(mt, mti, arg);
        else u = (uint32_t)mt_randrange(mt, mti, arg);
        if (lane_id() == 0) {
            g.out_f64[(size_t)sd * g.n_ops + i] = f;
            g.out_u32[(size_t)sd * g.n_ops + i] = u;
        }
    }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ launchers --

// A kernel's own __shared__ bytes, remembered per kernel.
hipError_t cached_static_lds(const void* kernel, int* bytes) {
    static std::mutex mu;
    static std::map<const void*, int> memo;
    std::lock_guard<std::mutex> lk(mu);
    const auto it = memo.find(kernel);
    if (it != memo.end()) {
        *bytes = it->second;
        return hipSuccess;
    }
    hipFuncAttributes at;
    const hipError_t e = hipFuncGetAttributes(&at, kernel);
    if (e == hipSuccess) *bytes = memo[kernel] = (int)at.sharedSizeBytes;
    return e;
}

// Resident blocks per CU of a kernel for a dynamic LDS size, remembered per
// (kernel, size): engine loads ask for many sizes, a serving loop for the
// same ones again and again.  The first query of a kernel also allows it the
// device's whole per-block LDS, so that any size the occupancy admits can be
// launched.
hipError_t cached_occupancy(const void* kernel, int block_threads, int lds_bytes,
                            int* blocks_per_cu) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int>, int> memo;   // (device, kernel, lds)
    static std::set<std::pair<int, const void*>> opened;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(dev, kernel, lds_bytes);
    const auto it = memo.find(key);
    if (it != memo.end()) {
        *blocks_per_cu = it->second;
        return hipSuccess;
    }
    if (!opened.count({dev, kernel})) {
        int max_lds = 0;
        hipFuncAttributes at;
        if ((e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) !=
            hipSuccess)
            return e;
        if ((e = hipFuncGetAttributes(&at, kernel)) != hipSuccess) return e;
        e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                max_lds - (int)at.sharedSizeBytes);
        if (e != hipSuccess) return e;
        opened.insert({dev, kernel});
    }
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kernel, block_threads,
                                                     (size_t)lds_bytes);
    if (e == hipSuccess) memo[key] = *blocks_per_cu;
    return e;
}

using DecodeFn = void (*)(DecodeArgs);

template <int M, int V, bool F>
static DecodeFn decode_fn_w(int waves) {
    switch (waves) {
        case 1: return decode_kernel<1, M, V, F>;
        case 2: return decode_kernel<2, M, V, F>;
        case 4: return decode_kernel<4, M, V, F>;
        case 8: return decode_kernel<8, M, V, F>;
        default: return nullptr;
    }
}

template <int V, bool F>
static DecodeFn decode_fn_v(int waves, int mode) {
    switch (mode) {
        case kNarrow: return decode_fn_w<kNarrow, V, F>(waves);
        case kWide: return decode_fn_w<kWide, V, F>(waves);
        case kBig: return decode_fn_w<kBig, V, F>(waves);
        default: return nullptr;
    }
}

// The decode_kernel instance for a launch shape, key width and verdict kind,
// full (work count, shared prefixes) or not.
static DecodeFn decode_fn(int waves, int mode, int verdicts, bool full = true) {
    if (verdicts == kPointState)
        return full ? decode_fn_v<kPointState, true>(waves, mode)
                    : decode_fn_v<kPointState, false>(waves, mode);
    if (verdicts == kPairCache)
        return full ? decode_fn_v<kPairCache, true>(waves, mode)
                    : decode_fn_v<kPairCache, false>(waves, mode);
    return nullptr;
}

hipError_t launch_decode(const DecodeArgs& a, int waves, int grid, int lds_bytes,
                         hipStream_t st) {
    // the instance without the work count and the shared prefixes when the
    // launch uses neither (it keeps fewer values live: fewer spills)
    const DecodeFn f = decode_fn(waves, a.mode, a.verdicts, a.count_work != 0 || a.traj != 0);
    if (!f) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, a.work_counter, 0u, 2);
    DecodeArgs arg = a;
    void* args[] = {&arg};
    return hipLaunchKernel((const void*)f, dim3(grid), dim3(64 * waves), args, (size_t)lds_bytes,
                           st);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int n, hipStream_t st) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_feas(const FeasRowArgs& r, int n_blocks, hipStream_t st) {
    hipLaunchKernelGGL(feas_kernel, dim3(n_blocks), dim3(kFeasBlock), 0, st, r);
    return hipGetLastError();
}

hipError_t launch_fsr(const FsrArgs& f, int n_packs, int lds_bytes, hipStream_t st) {
    hipLaunchKernelGGL(fsr_kernel, dim3(n_packs), dim3(64), (size_t)lds_bytes, st, f);
    return hipGetLastError();
}

hipError_t launch_mt_seed(const uint64_t* seeds, uint32_t* out, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mt_seed_kernel, dim3(n), dim3(64), 0, st, seeds, out, n);
    return hipGetLastError();
}

hipError_t launch_rng_script(const RngArgs& g, int n_seeds, hipStream_t st) {
    hipLaunchKernelGGL(rng_script_kernel, dim3(n_seeds), dim3(64), 0, st, g);
    return hipGetLastError();
}

hipError_t decode_static_lds(int waves, int mode, int verdicts, int* bytes) {
    const DecodeFn f = decode_fn(waves, mode, verdicts);
    if (!f) return hipErrorInvalidValue;
    return cached_static_lds((const void*)f, bytes);
}

hipError_t decode_occupancy(int waves, int mode, int verdicts, int lds_bytes, int* blocks_per_cu) {
    const DecodeFn f = decode_fn(waves, mode, verdicts);
    if (!f) return hipErrorInvalidValue;
    return cached_occupancy((const void*)f, 64 * waves, lds_bytes, blocks_per_cu);
}

hipError_t fsr_prepare(int lds_bytes) {
    return hipFuncSetAttribute((const void*)fsr_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

int feas_block_size() { return kFeasBlock; }

}  // namespace plpk
