// block_kernel.hip -- gfx950 block-building decoder (the reference's second
// search unit, blockproof/plp_block.cpp:95-278).
//
// A pass repeats sweeps of plan -> placement until nothing is left or a sweep
// places nothing.  Plan (plp_block.cpp:129-153): the types with copies left
// in item-volume order (eps: Fisher-Yates shuffle), each decomposed into
// blocks (a x b base of one orientation, k high) by taking, while copies
// remain, the first shape in shape_precedence order whose copy count fits
// (eps: uniformly among the topk largest that fit), read from per-type
// greedy tables staged in LDS.  The plan is sorted into placement order
// (plp_block.cpp:156-165) on the shapes' placement ranks: up to 64 blocks by
// counting ranks in registers, beyond by a bitonic network in LDS.
// Placement (plp_block.cpp:168-183): every
// planned block still affordable is tested at every extreme point, in (z, y,
// x) order, with the decoder's feasibility predicate; the pick is the first
// feasible point, a random one (eps), or the one with the most bottom support
// (ties: side contact) / the most total contact (ties: bottom), earliest on
// full ties.  After the pass the blocks are decomposed into item boxes
// (plp_block.cpp:188-201), which are scored with FSR and kept by the D4' rule
// (plp_block.cpp:228-275).
//
// One block of NW waves per (instance, seed), persistent over a work queue
// like decode_kernel.  Every wave tests 64-point chunks c = wave, wave + NW,
// ...; with one wave and at most 32 live points, the next 2 or 4 tests of the
// plan share one pass over the list instead.  Wave 0 runs the plan, the
// picks, the extreme-point merge and the pass end.  The extreme-point list keeps only points where a box of the
// remaining types' smallest dimensions fits (every later block contains one
// such box), so the feasible points, and so every pick, are the reference's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "decoder_dev.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"
#include "mt19937.hpp"
#include "wave.hpp"

namespace plpk {

namespace {

struct BlockShared {
    alignas(16) int cnt[2][8];   // feasible points found by each wave, by evaluated-test parity
    int item;       // queue item of the block
    int quit;       // the decoder stopped (capacity error)
    int n_act;      // live extreme points
    int n_box;      // placed blocks
    int n_plan;     // blocks planned by the current sweep
};

// Componentwise smallest orientation over the types with copies left.
__device__ void remaining_min(const int* rem, const BlockTypeRec* tr, int nt, int& wm, int& dm,
                              int& hm) {
    const int lane = lane_id();
    int w = 0x3fffffff, d = 0x3fffffff, h = 0x3fffffff;
    for (int t = lane; t < nt; t += kWave)
        if (rem[t] > 0 && tr[t].has_ors) {
            w = min(w, tr[t].wmin);
            d = min(d, tr[t].dmin);
            h = min(h, tr[t].hmin);
        }
    wm = wave_min(w);
    dm = wave_min(d);
    hm = wave_min(h);
}

// contact_stats (plp_block.cpp:64-81) of a bw x bd x bh block at (x, y, z):
// bottom support (the floor counts in full) and flush side contact.
__device__ void contact_stats(int x, int y, int z, int bw, int bd, int bh, const Box* box,
                              int n_box, int64_t& bottom, int64_t& side) {
    int64_t bt = z == 0 ? (int64_t)bw * bd : 0, sd = 0;
    for (int i = 0; i < n_box; i++) {
        const Box B = box[i];
        const int64_t ox = (int64_t)min(x + bw, B.x2) - max(x, B.x);
        const int64_t oy = (int64_t)min(y + bd, B.y2) - max(y, B.y);
        const int64_t oz = (int64_t)min(z + bh, B.z2) - max(z, B.z);
        if (B.z2 == z && ox > 0 && oy > 0) bt += ox * oy;
        if (oy > 0 && oz > 0 && (B.x2 == x || B.x == x + bw)) sd += oy * oz;
        if (ox > 0 && oz > 0 && (B.y2 == y || B.y == y + bd)) sd += ox * oz;
    }
    bottom = bt;
    side = sd;
}

// Ascending bitonic sort of n2 (a power of two) keys in LDS, one wave.
__device__ void bitonic_u32(uint32_t* key, int n2) {
    const int lane = lane_id();
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < n2; i += kWave) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t p = key[i], q = key[ixj];
                    if ((p > q) == ((i & k) == 0)) {
                        key[i] = q;
                        key[ixj] = p;
                    }
                }
            }
            wave_sync();
        }
}

}  // namespace

// BIG: 64-bit extreme-point keys and support sums (StateMode kBig), else
// 32-bit ones.  The bound asks for at most 168 VGPRs (three waves per SIMD):
// the 32-bit form then needs no scratch memory.  A block decoder's time is
// the latency of its serial chain (one wave per decoder runs as fast as
// four), so residency beyond that buys nothing.
template <int NW, bool BIG>
__global__ void __launch_bounds__(64 * NW, 3) block_kernel(BlockArgs ba) {
    using Key = typename ModeTypes<BIG ? kBig : kNarrow>::Key;
    using Sum = typename ModeTypes<BIG ? kBig : kNarrow>::Sum;
    const DecodeArgs& a = ba.d;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ BlockShared sh;
    __shared__ DecState sst;
    const BlockLds lay =
        block_lds_layout(a.ep_cap, a.box_cap, a.step_cap, ba.max_types, a.use_rng, NW, BIG);
    Key* const ep_a = (Key*)(lds + lay.off_ep);
    Key* const ep_b = (Key*)(lds + lay.off_ep2);
    uint16_t* const oidx = (uint16_t*)(lds + lay.off_oidx);
    uint64_t* const mask = (uint64_t*)(lds + lay.off_mask);
    int* const rem = (int*)(lds + lay.off_rem);
    int* const tlist = (int*)(lds + lay.off_tlist);
    uint32_t* const plan = (uint32_t*)(lds + lay.off_plan);
    double* const load = (double*)(lds + lay.off_load);
    uint64_t* const skey = (uint64_t*)(lds + lay.off_key);
    Sum* const stot = (Sum*)(lds + lay.off_tot);
    Box* const items = (Box*)(lds + lay.off_items);
    Box* const box = (Box*)(lds + lay.off_box);
    uint32_t* const mt = (uint32_t*)(lds + lay.off_mt);
    BlockTypeRec* const ltr = (BlockTypeRec*)(lds + lay.off_types);
    int32_t* const lvord = (int32_t*)(lds + lay.off_vord);
    GreedyEntry* const lgt = (GreedyEntry*)(lds + lay.off_gtab);
    const int lane = lane_id();
    const int wid = (int)(threadIdx.x >> 6);
    Box* const cbox = (Box*)(lds + lay.off_rel) + wid * kCullWindow;
    const bool use_eps = a.eps > 0.0;

    for (;;) {
        if (threadIdx.x == 0) sh.item = (int)atomicAdd(a.work_counter, 1u);
        block_sync();
        const int item = uni(sh.item);
        block_sync();
        if (item >= a.n_work) break;
        const int d = uni(a.work_order[item]);
        const InstHdr I = a.inst[uni(a.dec_inst[d])];
        const int L = I.L, W = I.W, H = I.H, xb = I.xbits, yb = I.ybits, nt = I.n_types;
        const BlockShapeRec* const shapes = ba.shape + I.shape_off;
        // the instance's type records, type order and greedy-plan entries go
        // to LDS once per decoder: the plan reads them in serial chains
        for (int t = threadIdx.x; t < nt; t += 64 * NW) {
            ltr[t] = ba.tinfo[I.tinfo_off + t];
            lvord[t] = ba.vorder[I.vorder_off + t];
        }
        for (int e = threadIdx.x; e < I.n_steps; e += 64 * NW) lgt[e] = ba.gtab[I.gtab_off + e];
        const BlockTypeRec* const tr = ltr;
        const int32_t* const vord = lvord;

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
            // ---- pass start (wave 0): all copies left, origin only, no blocks
            int n_act = 1, n_dead = 0, n_box = 0;
            uint32_t ep_need = 1;
            // wave 0: the items the placed blocks hold, and whether a block
            // would have passed the item capacity (box_cap)
            int n_it = 0;
            bool item_over = false;
            int wm = 0, dm = 0, hm = 0;
            if (wid == 0) {
                for (int t = lane; t < nt; t += kWave) rem[t] = tr[t].qty;
                if (lane == 0) {
                    ep_a[0] = Key(0);   // origin; L, W, H > 0
                    oidx[0] = kNewPoint;
                }
                for (int i = lane; i < a.box_cap + 8; i += kWave) box[i] = sentinel_box();
                wave_sync();
                remaining_min(rem, tr, nt, wm, dm, hm);
            }
            block_sync();
            int na_w = 1, nb_w = 0, cur = 0, ev = 0;
            bool quit = false;
            for (;;) {   // sweeps
                // ---- plan (wave 0)
                if (wid == 0) {
                    int n_types = 0;
                    for (int q0 = 0; q0 < nt; q0 += kWave) {
                        const int q = q0 + lane;
                        const int t = q < nt ? vord[q] : 0;
                        const bool in = q < nt && rem[t] > 0 && tr[t].has_ors;
                        const uint64_t m = ballot(in);
                        if (in) tlist[n_types + rank_below(m)] = t;
                        n_types += __popcll(m);
                    }
                    n_types = uni(n_types);
                    wave_sync();
                    int n_plan = 0;
                    if (n_types > 0) {
                        if (use_eps && mt_random(mt, mti) >= 1.0 - a.eps) {
                            for (int i = n_types - 1; i > 0; i--) {
                                const int j = mt_randrange(mt, mti, i + 1);
                                if (lane == 0) {
                                    const int tmp = tlist[i];
                                    tlist[i] = tlist[j];
                                    tlist[j] = tmp;
                                }
                                wave_sync();
                            }
                        }
                        // Blocks of each type in turn (plp_block.cpp:129-153).
                        // While no decision makes an eps pick, every type's
                        // blocks are its greedy chain (the first shape in
                        // precedence order whose copy count fits the copies
                        // left), so the chains are walked lane-parallel (lane
                        // = type) and placed by a prefix sum of their lengths,
                        // and decision k's random() comes from MT words mti +
                        // 2k and mti + 2k + 1, all computed at once.  From the
                        // first decision that makes a pick on -- and for the
                        // whole plan when those words would cross a
                        // regeneration or there are more than 64 types -- the
                        // decisions run one by one, drawing as they go.
                        int q_ser = 0, r_ser = -1;   // one by one from type q_ser, r_ser copies
                                                     // left there (-1: all of them)
                        bool drawn = false;          // that first decision's random() made a pick
                        if (n_types <= kWave) {
                            const bool okq = lane < n_types;
                            const int t = okq ? tlist[lane] : 0;
                            const int gt = tr[t].gt_off, size = tr[t].shp_cnt;
                            const int r0 = okq ? rem[t] : 0;
                            int len = 0;
                            for (int rr = r0; ballot(rr > 0) != 0;) {
                                if (rr > 0) {
                                    const GreedyEntry g = lgt[gt + rr - 1];
                                    if (g.start == size) {
                                        rr = 0;
                                    } else {
                                        len++;
                                        rr -= g.cnt0;
                                    }
                                }
                            }
                            int inc = len;
#pragma unroll
                            for (int dl = 1; dl < kWave; dl <<= 1) {
                                const int up = __shfl_up(inc, dl);
                                if (lane >= dl) inc += up;
                            }
                            const int off = inc - len;
                            const int n_dec = __shfl(inc, kWave - 1);
                            int kstar = n_dec;   // first decision with a pick
                            bool par = true;
                            if (use_eps) {
                                if (mti + 2 * n_dec <= 624) {
                                    for (int k0 = 0; k0 < n_dec; k0 += kWave) {
                                        const int k = k0 + lane;
                                        bool trig = false;
                                        if (k < n_dec) {
                                            const uint32_t wa = mt_temper(mt[mti + 2 * k]) >> 5;
                                            const uint32_t wb = mt_temper(mt[mti + 2 * k + 1]) >> 6;
                                            const double rv = ((double)wa * 67108864.0 + (double)wb) *
                                                              (1.0 / 9007199254740992.0);
                                            trig = rv >= 1.0 - a.eps;
                                        }
                                        const uint64_t m = ballot(trig);
                                        if (m != 0) {
                                            kstar = k0 + __ffsll((unsigned long long)m) - 1;
                                            break;
                                        }
                                    }
                                    mti += 2 * (kstar < n_dec ? kstar + 1 : n_dec);
                                } else {
                                    par = false;
                                }
                            }
                            if (par) {
                                // the greedy decisions before kstar
                                int rr = r0, p = 0;
                                for (;;) {
                                    const bool go = (p < len) & (off + p < kstar);
                                    if (ballot(go) == 0) break;
                                    if (go) {
                                        const GreedyEntry g = lgt[gt + rr - 1];
                                        plan[off + p] = (uint32_t)g.s0;
                                        p++;
                                        rr -= g.cnt0;
                                    }
                                }
                                n_plan = min(kstar, n_dec);
                                q_ser = n_types;
                                if (kstar < n_dec) {
                                    // decision kstar: its type, and the copies left
                                    // there (where that lane's walk stopped)
                                    const bool at = okq & (off <= kstar) & (kstar < off + len);
                                    const int qs = __ffsll((unsigned long long)ballot(at)) - 1;
                                    q_ser = qs;
                                    r_ser = read_lane(rr, qs);
                                    drawn = true;
                                }
                            }
                            wave_sync();
                        }
                        for (int q = q_ser; q < n_types; q++) {
                            const int t = uni(tlist[q]);
                            int r = (q == q_ser && r_ser >= 0) ? r_ser : uni(rem[t]);
                            const int off = uni(tr[t].shp_off), size = uni(tr[t].shp_cnt);
                            const int gt = uni(tr[t].gt_off);
                            while (r > 0) {
                                // the first shape (precedence order) whose copy
                                // count fits r, from the type's greedy entries
                                const GreedyEntry g = lgt[gt + r - 1];
                                const int start = uni(g.start);
                                if (start == size) break;
                                const int tk = min(ba.topk, size - start);
                                int pk = 0;
                                bool pick = drawn;
                                drawn = false;
                                if (!pick) pick = use_eps && mt_random(mt, mti) >= 1.0 - a.eps;
                                if (pick) pk = mt_randrange(mt, mti, tk);
                                int s = uni(g.s0), cnt = uni(g.cnt0);
                                if (pk != 0) {   // an eps pick among the topk largest that fit
                                    s = uni(ba.by_type[off + start + pk]);
                                    cnt = uni(shapes[s].cnt);
                                }
                                if (lane == 0) plan[n_plan] = (uint32_t)s;
                                n_plan++;
                                r -= cnt;
                            }
                        }
                        if (n_plan <= kWave) {
                            // placement order in registers: lane i's key goes to
                            // its rank (equal keys are the same shape, any order)
                            wave_sync();
                            const uint32_t key = lane < n_plan ? plan[lane] : 0xffffffffu;
                            int rank = 0;
                            for (int j = 0; j < n_plan; j++) {
                                const uint32_t kj = read_lane(key, j);
                                rank += (kj < key) | ((kj == key) & (j < lane));
                            }
                            wave_sync();
                            if (lane < n_plan) plan[rank] = key;
                            wave_sync();
                        } else {
                            const int n2 = next_pow2(n_plan);
                            for (int i = n_plan + lane; i < n2; i += kWave) plan[i] = 0xffffffffu;
                            wave_sync();
                            bitonic_u32(plan, n2);
                        }
                    }
                    if (lane == 0) sh.n_plan = n_plan;
                }
                block_sync();
                const int n_plan = uni(sh.n_plan);
                if (n_plan == 0) break;

                // ---- placement
                bool any = false;
                int fail_s = -1;   // shape whose last test found nothing, no block placed since
                // each test's shape record is read one test ahead
                BlockShapeRec S_next = shapes[uni((int)plan[0])];
                for (int q = 0; q < n_plan; q++) {
                    const Key* act = cur ? ep_b : ep_a;
                    const int n_chunks = (na_w + kWave - 1) / kWave;
                    BlockShapeRec S;   // the test that places
                    int cnt;           // its feasible points
                    if (NW == 1 && na_w <= kWave / 2) {
                        // Short list, one wave: the tests of plan[q .. q+G) run
                        // together, test g on lanes [g*seg, (g+1)*seg), lane =
                        // point.  The tests see the same list (nothing is placed
                        // before the first that finds a point), so the first
                        // group with a feasible point is the serial loop's next
                        // placement, and the groups before it found nothing.
                        const int lg = na_w <= kWave / 4 ? 4 : 5;   // log2(seg)
                        const int g = lane >> lg;
                        const int qg = q + g;
                        const bool inr = qg < n_plan;
                        const int sg = inr ? (int)plan[qg] : 0;
                        const int sp = g == 0 ? fail_s : (inr ? (int)plan[qg - 1] : -1);
                        const BlockShapeRec Sg = shapes[sg];
                        // stale tests and repeats of the test before (the same
                        // answer) test nothing
                        const bool tv = inr && rem[Sg.ti] >= Sg.cnt && sg != sp;
                        const int e = lane & ((1 << lg) - 1);
                        const bool ok = tv & (e < na_w);
                        int x, y, z;
                        key_xyz(act[ok ? e : 0], xb, yb, x, y, z);
                        const bool c2 = ok & (x + Sg.bw <= L) & (y + Sg.bd <= W) & (z + Sg.bh <= H);
                        uint64_t m = 0;
                        if (ballot(c2) != 0) {
                            // the region of all groups' pairs (a box outside it
                            // meets none of them)
                            const int xlo = wave_min(c2 ? x : 0x7fffffff);
                            const int xhi = wave_max(c2 ? x + Sg.bw : -1);
                            const int ylo = wave_min(c2 ? y : 0x7fffffff);
                            const int yhi = wave_max(c2 ? y + Sg.bd : -1);
                            const int zlo = wave_min(c2 ? z : 0x7fffffff);
                            const int zhi = wave_max(c2 ? z + Sg.bh : -1);
                            Sum sup = 0;
                            const bool c12 = pair_scan(x, y, z, Sg.bw, Sg.bd, Sg.bh, c2, box, nb_w,
                                                       xlo, xhi, ylo, yhi, zlo, zhi, cbox, sup);
                            m = ballot(pair_verdict(c12, z, Sg.bw, Sg.bd, sup));
                        }
                        const int q_end = min(q + (kWave >> lg), n_plan);
                        if (m == 0) {
                            // every test of the round found nothing
                            fail_s = uni((int)plan[q_end - 1]);
                            q = q_end - 1;
                            if (q + 1 < n_plan) S_next = shapes[uni((int)plan[q + 1])];
                            continue;
                        }
                        const int gw = (__ffsll((unsigned long long)m) - 1) >> lg;
                        const uint64_t mw = (m >> (gw << lg)) & ((1ull << (1 << lg)) - 1);
                        q += gw;
                        if (lane == 0) mask[0] = mw;
                        S = shapes[uni((int)plan[q])];
                        if (q + 1 < n_plan) S_next = shapes[uni((int)plan[q + 1])];
                        cnt = __popcll(mw);
                        wave_sync();
                    } else {
                        const int s = uni((int)plan[q]);
                        S = S_next;
                        if (q + 1 < n_plan) S_next = shapes[uni((int)plan[q + 1])];
                        const int ti = uni(S.ti), cnt_s = uni(S.cnt);
                        const int bw = uni(S.bw), bd = uni(S.bd), bh = uni(S.bh);
                        if (uni(rem[ti]) < cnt_s) continue;   // planned before, stale
                        if (s == fail_s) continue;            // same test, same (empty) answer
                        int cnt_w = 0;
                        for (int c = wid; c < n_chunks; c += NW) {
                            const int e = c * kWave + lane;
                            const bool ok = e < na_w;
                            int x, y, z;
                            // in bounds; ok lanes count
                            key_xyz(act[ok ? e : 0], xb, yb, x, y, z);
                            const bool c2 = ok & (x + bw <= L) & (y + bd <= W) & (z + bh <= H);
                            uint64_t m = 0;
                            if (ballot(c2) != 0) {
                                const int xlo = wave_min(c2 ? x : 0x7fffffff);
                                const int xhi = wave_max(c2 ? x + bw : -1);
                                const int ylo = wave_min(c2 ? y : 0x7fffffff);
                                const int yhi = wave_max(c2 ? y + bd : -1);
                                const int zlo = wave_min(c2 ? z : 0x7fffffff);
                                const int zhi = wave_max(c2 ? z + bh : -1);
                                Sum sup = 0;
                                const bool c12 = pair_scan(x, y, z, bw, bd, bh, c2, box, nb_w, xlo,
                                                           xhi, ylo, yhi, zlo, zhi, cbox, sup);
                                m = ballot(pair_verdict(c12, z, bw, bd, sup));
                            }
                            mask[c] = m;   // every lane writes the same (uniform) mask
                            cnt_w += __popcll(m);
                        }
                        const int par = ev & 1;
                        ev++;
                        sh.cnt[par][wid] = cnt_w;   // every lane: the same value
                        block_sync();
                        cnt = uni(sum_counts<NW>(sh.cnt[par]));
                        if (cnt == 0) {
                            fail_s = s;
                            continue;
                        }
                    }
                    fail_s = -1;
                    const int ti = uni(S.ti), cnt_s = uni(S.cnt);
                    const int bw = uni(S.bw), bd = uni(S.bd), bh = uni(S.bh);
                    // ---- pick, place, merge (wave 0)
                    if (wid == 0 && n_it + cnt_s > a.box_cap) {
                        // one block past the item capacity: this decoder stops
                        // and the host reruns it with room for every item
                        item_over = true;
                        if (lane == 0) sh.quit = 1;
                    } else if (wid == 0) {
                        n_it += cnt_s;
                        int pe = -1;   // picked extreme point
                        if (use_eps && mt_random(mt, mti) >= 1.0 - a.eps) {
                            int idx = mt_randrange(mt, mti, cnt);
                            for (int c = 0;; c++) {
                                const uint64_t m = uni(mask[c]);
                                const int pc = __popcll(m);
                                if (idx < pc) {
                                    const bool hit = ((m >> lane) & 1ull) && rank_below(m) == idx;
                                    pe = c * kWave + __ffsll((unsigned long long)ballot(hit)) - 1;
                                    break;
                                }
                                idx -= pc;
                            }
                        } else if (ba.pick == 0) {
                            for (int c = 0;; c++) {
                                const uint64_t m = uni(mask[c]);
                                if (m) {
                                    pe = c * kWave + __ffsll((unsigned long long)m) - 1;
                                    break;
                                }
                            }
                        } else {
                            // lane-parallel contact statistics; the wave keeps
                            // the largest (primary, secondary), earliest point
                            int64_t k0 = -1, k1 = -1;
                            int bi = 0x7fffffff;
                            for (int c = 0; c < n_chunks; c++) {
                                const uint64_t m = uni(mask[c]);
                                if (!m) continue;
                                const int e = c * kWave + lane;
                                if ((m >> lane) & 1ull) {
                                    int x, y, z;
                                    key_xyz(act[e], xb, yb, x, y, z);
                                    int64_t sup, side;
                                    contact_stats(x, y, z, bw, bd, bh, box, n_box, sup, side);
                                    const int64_t p0 = ba.pick == 1 ? sup : sup + side;
                                    const int64_t p1 = ba.pick == 1 ? side : sup;
                                    if (p0 > k0 || (p0 == k0 && p1 > k1)) {
                                        k0 = p0;
                                        k1 = p1;
                                        bi = e;
                                    }
                                }
                            }
                            // wave argmax: (k0, k1) largest, then smallest index
                            for (int off = 32; off >= 1; off >>= 1) {
                                const int64_t o0 = __shfl_xor(k0, off, kWave);
                                const int64_t o1 = __shfl_xor(k1, off, kWave);
                                const int ob = __shfl_xor(bi, off, kWave);
                                const bool take = o0 > k0 || (o0 == k0 && (o1 > k1 ||
                                                                          (o1 == k1 && ob < bi)));
                                if (take) {
                                    k0 = o0;
                                    k1 = o1;
                                    bi = ob;
                                }
                            }
                            pe = uni(bi);
                        }
                        int x, y, z;
                        key_xyz(uni(act[pe]), xb, yb, x, y, z);
                        x = uni(x);
                        y = uni(y);
                        z = uni(z);
                        const int X2 = x + bw, Y2 = y + bd, Z2 = z + bh;
                        if (lane == 0) {
                            Box nbx;
                            nbx.x = x; nbx.y = y; nbx.z = z; nbx.x2 = X2; nbx.y2 = Y2; nbx.z2 = Z2;
                            nbx.type = ti;
                            nbx.pad = uni(S.a) | (uni(S.b) << 8) | (uni(S.k) << 16);
                            box[n_box] = nbx;
                            rem[ti] -= cnt_s;
                        }
                        n_box++;
                        wave_sync();
                        if (uni(rem[ti]) <= 0) remaining_min(rem, tr, nt, wm, dm, hm);
                        const int need = merge_points(act, cur ? ep_a : ep_b, oidx, (Key*)nullptr, n_act,
                                                      n_dead, box, n_box, x, y, z, X2, Y2, Z2, L, W,
                                                      H, xb, yb, a.ep_cap, false, wm, dm, hm);
                        ep_need = max(ep_need, (uint32_t)need);
                        if (lane == 0) {
                            sh.quit = need > a.ep_cap;   // PLP_E_EPCAP
                            sh.n_act = n_act;
                            sh.n_box = n_box;
                        }
                    }
                    block_sync();
                    if (uni(sh.quit) != 0) {
                        quit = true;
                        break;
                    }
                    na_w = uni(sh.n_act);
                    nb_w = uni(sh.n_box);
                    cur ^= 1;
                    any = true;
                }
                if (quit || !any) break;
            }

            // ---- pass end (wave 0): decomposed items, FSR, D4'.  The state
            // in LDS (sst) is updated field by field (a whole-struct copy in
            // registers would partly live in scratch memory).
            if (wid == 0) {
                const bool overflow = uni(sh.quit) != 0;
                const int err = overflow ? 4 : uni(sst.err);   // 4: PLP_E_EPCAP
                const uint32_t need_ep = max(uni(sst.ep_need), ep_need);
                wave_sync();
                if (lane == 0) {
                    sst.err = err;
                    sst.ep_need = need_ep;
                    sst.box_need = max(uni(sst.box_need),
                                       item_over ? (uint32_t)a.box_cap + 1u : (uint32_t)n_it);
                }
                if (err == 0) {
                    // blocks -> items in block order, i (x) outer, j (y), t (z)
                    // inner.  Lane = block: its item sides (the block's sides
                    // are exact multiples, so the float quotient of these
                    // integers below 2^24 is exact), its first item from a
                    // wave prefix sum of the counts, then its items with
                    // (i, j, t) stepped, no division.
                    int n = 0;
                    int64_t vol = 0;
                    for (int b0 = 0; b0 < n_box; b0 += kWave) {
                        const int b = b0 + lane;
                        const bool okb = b < n_box;
                        const Box B = box[okb ? b : b0];
                        const int ba_ = max(B.pad & 0xff, 1), bb = max((B.pad >> 8) & 0xff, 1),
                                  bk = max((B.pad >> 16) & 0xff, 1);
                        const int iw = (int)((float)(B.x2 - B.x) / (float)ba_);
                        const int id = (int)((float)(B.y2 - B.y) / (float)bb);
                        const int ih = (int)((float)(B.z2 - B.z) / (float)bk);
                        const int cnt_b = okb ? ba_ * bb * bk : 0;
                        int inc = cnt_b;
#pragma unroll
                        for (int dl = 1; dl < kWave; dl <<= 1) {
                            const int up = __shfl_up(inc, dl);
                            if (lane >= dl) inc += up;
                        }
                        const int off = n + inc - cnt_b;
                        // an item below the block's top layer supports the one
                        // above it; a top-layer item may support something
                        // only if some block rests on the block
                        const bool top_sup = (B.pad & kSupports) != 0;
                        const int m_end = wave_max(cnt_b);
                        int i = 0, j = 0, t = 0;
                        for (int m = 0; m < m_end; m++) {
                            if (m < cnt_b) {
                                Box it;
                                it.x = B.x + i * iw; it.y = B.y + j * id; it.z = B.z + t * ih;
                                it.x2 = it.x + iw; it.y2 = it.y + id; it.z2 = it.z + ih;
                                it.type = B.type;
                                it.pad = ((t < bk - 1) | top_sup) ? kSupports : 0;
                                items[off + m] = it;
                            }
                            if (++t == bk) {
                                t = 0;
                                if (++j == bb) {
                                    j = 0;
                                    ++i;
                                }
                            }
                        }
                        vol += wave_sum_i64(okb ? (int64_t)cnt_b * iw * id * ih : 0);
                        n += __shfl(inc, kWave - 1);
                    }
                    wave_sync();
                    // D4' (plp_block.cpp:228-275): a packing that beats neither
                    // best cannot be selected, so its score is never needed;
                    // nor is the score of one that beats only the best
                    // accepted when it provably stays below theta
                    const int n_all = uni(sst.n_all), n_acc = uni(sst.n_acc);
                    const int64_t v_all = uni(sst.v_all), v_acc = uni(sst.v_acc);
                    const bool better_all =
                        n_all < 0 || vol > v_all || (vol == v_all && n > n_all);
                    const bool better_acc =
                        n_acc < 0 || vol > v_acc || (vol == v_acc && n > n_acc);
                    bool need = better_all || better_acc;
                    if (need && !better_all && a.use_theta && n > 0 &&
                        (double)fsr_count_bound(items, n, a.mu) / (double)n < a.theta)
                        need = false;
                    const double score =
                        need ? fsr_score<4>(items, n, a.mu, load, skey, stot, nullptr) : 0.0;
                    const bool accept = better_acc && (!a.use_theta || score >= a.theta);
                    // the selection only moves to the packing just decoded
                    if (accept || (n_acc < 0 && better_all)) write_rows(out_sel, items, n);
                    wave_sync();
                    if (lane == 0) {
                        if (better_all) {
                            sst.n_all = n; sst.v_all = vol; sst.s_all = score;
                        }
                        if (accept) {
                            sst.n_acc = n; sst.v_acc = vol; sst.s_acc = score;
                        }
                        sst.passes = sst.passes + 1;
                    }
                }
                wave_sync();
            }
            block_sync();
            if (uni(sh.quit) != 0) break;
        }
        if (wid == 0 && !skip_decoder) {
            if (a.use_rng) {
                for (int i = lane; i < 624; i += kWave) gmt[i] = mt[i];
                if (lane == 0) gmt[624] = (uint32_t)mti;
            }
            if (lane == 0) a.state[d] = sst;
        }
        if (wid == 0 && lane == 0) a.out_rec[d] = row_record(sst, a.dec_seed[d]);
        if (wid == 0 && a.done != nullptr) {
            // wave 0 wrote the record and every packing row: once they have
            // reached host memory, the decoder's completion word
            __threadfence_system();
            if (lane == 0)
                __hip_atomic_store(a.done + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // work_counter[1] counts the decoders that end with an error, so the
        // host reads one word, not every state, when none did
        if (wid == 0 && lane == 0 && sst.err != 0) atomicAdd(a.work_counter + 1, 1u);
        block_sync();
    }
}

template <bool BIG>
static hipError_t launch_block_w(const BlockArgs& b, int waves, int grid, int lds_bytes,
                                 hipStream_t st) {
    switch (waves) {
        case 1: hipLaunchKernelGGL((block_kernel<1, BIG>), dim3(grid), dim3(64), (size_t)lds_bytes, st, b); break;
        case 2: hipLaunchKernelGGL((block_kernel<2, BIG>), dim3(grid), dim3(128), (size_t)lds_bytes, st, b); break;
        case 4: hipLaunchKernelGGL((block_kernel<4, BIG>), dim3(grid), dim3(256), (size_t)lds_bytes, st, b); break;
        case 8: hipLaunchKernelGGL((block_kernel<8, BIG>), dim3(grid), dim3(512), (size_t)lds_bytes, st, b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_block(const BlockArgs& b, int waves, int grid, int lds_bytes, hipStream_t st) {
    hipError_t e = launch_fill_u32(b.d.work_counter, 0u, 2, st);
    if (e != hipSuccess) return e;
    return b.d.mode == kBig ? launch_block_w<true>(b, waves, grid, lds_bytes, st)
                            : launch_block_w<false>(b, waves, grid, lds_bytes, st);
}

template <int NW, bool BIG>
static hipError_t block_occupancy_of(int lds_bytes, int* blocks_per_cu) {
    return cached_occupancy((const void*)block_kernel<NW, BIG>, 64 * NW, lds_bytes, blocks_per_cu);
}

template <int NW, bool BIG>
static hipError_t block_static_of(int* bytes) {
    return cached_static_lds((const void*)block_kernel<NW, BIG>, bytes);
}

template <bool BIG>
static hipError_t block_occupancy_w(int waves, int lds_bytes, int* blocks_per_cu) {
    switch (waves) {
        case 1: return block_occupancy_of<1, BIG>(lds_bytes, blocks_per_cu);
        case 2: return block_occupancy_of<2, BIG>(lds_bytes, blocks_per_cu);
        case 4: return block_occupancy_of<4, BIG>(lds_bytes, blocks_per_cu);
        case 8: return block_occupancy_of<8, BIG>(lds_bytes, blocks_per_cu);
        default: return hipErrorInvalidValue;
    }
}

hipError_t block_occupancy(int waves, int mode, int lds_bytes, int* blocks_per_cu) {
    return mode == kBig ? block_occupancy_w<true>(waves, lds_bytes, blocks_per_cu)
                        : block_occupancy_w<false>(waves, lds_bytes, blocks_per_cu);
}

template <bool BIG>
static hipError_t block_static_w(int waves, int* bytes) {
    switch (waves) {
        case 1: return block_static_of<1, BIG>(bytes);
        case 2: return block_static_of<2, BIG>(bytes);
        case 4: return block_static_of<4, BIG>(bytes);
        case 8: return block_static_of<8, BIG>(bytes);
        default: return hipErrorInvalidValue;
    }
}

hipError_t block_static_lds(int waves, int mode, int* bytes) {
    return mode == kBig ? block_static_w<true>(waves, bytes) : block_static_w<false>(waves, bytes);
}

}  // namespace plpk
