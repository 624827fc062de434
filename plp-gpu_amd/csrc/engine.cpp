// engine.cpp -- device-resident lockstep decoders (C ABI seam 3).
//
// Replaces CudaDeviceBackend (gpu/plp_gpu.cu:1225-1390) and run_device's
// D4' selection (plp_gpu.cu:1395-1473) for many instances at once.  Layout in
// HBM, all allocated once per load:
//   instance headers (64 B each), step-type table, orientation table (20 ints
//   per instance type), decoder -> instance map, decoder seeds, work order,
//   per decoder: MT19937 state (625 words) and DecState (80 B).
// Results go to pinned host memory that the kernels write directly: per
// decoder one 64-byte record (the plp_row_result layout) at the end of a run
// and the selected packing (max_items rows x 7 ints) whenever the selection
// moves to the packing just decoded, so a fetch copies nothing over PCIe
// after the kernel.  A run is one fill kernel (queue head) + one persistent
// decode launch; it can be captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace plpi;
using plpk::DecState;
using plpk::InstHdr;

struct plp_engine {
    plp_ctx* ctx = nullptr;
    plp_engine_config cfg{};
    bool loaded = false;
    int n_inst = 0, n_dec = 0, max_items = 1;
    int box_cap = 8;    // placed boxes a decoder holds in LDS (multiple of 8)
    int full_box = 8;   // every item of the largest instance, rounded up to 8
    int step_cap = 8;   // step-table entries in LDS
    int ep_cap = 0, lds = 0, grid = 0, occ = 0, wpd = 1;
    bool wpd_tuned = false;   // wpd chosen by plp_engine_autotune (kept across loads)
    int mode = plpk::kNarrow;   // widths of keys, support sums and cache entries (StateMode)
    int verd = plpk::kPointState;   // how the item decoder reaches its verdicts (VerdictKind)
    bool blocks = false;   // block-building decoder (plp_engine_load_blocks)
    plp_block_params bp{};
    int max_types = 1;
    // Read-only tables of a load, in one device buffer uploaded from one
    // pinned staging buffer (one copy per load).
    enum Tab { kInst, kStep, kStepMin, kShape, kTinfo, kByType, kVorder, kGtab, kOrs, kDecInst,
               kDecSeed, kOrder, kTrajOrder, kNumTabs };
    DevBuf d_tab;
    HostBuf h_tab;
    size_t tab_off[kNumTabs] = {};
    plp_params params{};
    int use_rng = 0;
    int passes_done = 0;
    bool need_reset = true;
    bool states_fresh = false;
    DevBuf d_retry, d_mt, d_mt0, d_state, d_counter, d_ticks;
    HostBuf h_rows, h_rec;   // results, written by the kernels (fine-grained pinned memory)
    // per decoder: 1 once its results are in h_rows / h_rec (item decoder
    // launches; cleared by every sync), and the runs enqueued since the last sync
    HostBuf h_done;
    int runs_unsynced = 0;
    std::vector<int32_t> h_order;
    std::vector<int32_t> h_dec_inst;
    std::vector<uint64_t> h_seed;
    std::vector<DecState> h_state;
    uint32_t* h_failed = nullptr;   // pinned: decoders of the last run that ended in error
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_passes = -1, g_reset = -1, g_grid = -1, g_cap = -1, g_wpd = -1, g_verd = -1,
        g_share = -1;
    const void* g_tbuf = nullptr;
    int g_tshape = -1;
    // Shared prefixes (plpk::TrajHdr): the mode (plp_engine_config
    // share_prefix; the autotuner sets 1 or 2), whether the automatic mode
    // expects decoders to follow their trajectories for long (no RNG, or eps
    // small against the item count), whether a block can hold a trajectory,
    // the recording launch's shape (waves, list capacity, LDS bytes, grid),
    // and one device buffer holding per instance the header, the placing
    // steps, the boxes and the kept lists.
    int share_mode = 0;
    bool share_eps = false, share_fit = true;
    // runs since the load, and whether the queue was ordered by measured times
    int runs = 0;
    bool ordered = false;
    int t_wpd = 4, t_cap = 0, t_lds = 0, t_grid = 0;
    int t_wpd_req = 0;   // recording waves, 0: at least 4 and the decoders'
    int t_verd = -1;     // recording verdict kind, -1: the decoders'
    size_t t_key_cap = 0;   // keys per instance
    DevBuf d_traj;
};

// The kernels write plp_row_result records as plpk::RowRecord.
static_assert(sizeof(plp_row_result) == sizeof(plpk::RowRecord), "record size");
static_assert(offsetof(plp_row_result, volume) == offsetof(plpk::RowRecord, volume), "record");
static_assert(offsetof(plp_row_result, score) == offsetof(plpk::RowRecord, score), "record");
static_assert(offsetof(plp_row_result, seed) == offsetof(plpk::RowRecord, seed), "record");
static_assert(offsetof(plp_row_result, work_live_pair_box) == offsetof(plpk::RowRecord, work_live),
              "record");
static_assert(offsetof(plp_row_result, ep_need) == offsetof(plpk::RowRecord, ep_need), "record");

namespace {

// A few persistent host threads that, with the calling thread, run the parts
// of one job: the copy of the selected packings out of the pinned result
// memory in plp_engine_fetch, which one thread does at about 12 GB/s.  One
// pool per process; jobs from several threads run one after another.
class CopyPool {
public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; i++) th_.emplace_back([this] { work(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : th_) t.join();
    }
    // Threads that take parts: the workers and the caller.
    int threads() const { return (int)th_.size() + 1; }
    // fn(i) for every i in [0, parts); returns when all have run.
    void run(int parts, const std::function<void(int)>& fn) {
        std::lock_guard<std::mutex> job(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            parts_ = parts;
            next_.store(0);
            active_ = (int)th_.size();
            gen_++;
        }
        cv_.notify_all();
        for (int i; (i = next_.fetch_add(1)) < parts;) fn(i);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return active_ == 0; });
    }

private:
    void work() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(int)>* fn = fn_;
            const int parts = parts_;
            lk.unlock();
            for (int i; (i = next_.fetch_add(1)) < parts;) (*fn)(i);
            lk.lock();
            if (--active_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_, job_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int parts_ = 0, active_ = 0;
    std::atomic<int> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

CopyPool& copy_pool() {
    static CopyPool pool(std::max(0, std::min(7, (int)std::thread::hardware_concurrency() - 1)));
    return pool;
}

// "No later box": larger than any container side (< 2^21), small enough that
// a coordinate plus it stays in int32.
constexpr int32_t kNoBox = 1 << 30;

int round_up(int v, int m) { return (v + m - 1) / m * m; }

const void* table(const plp_engine* g, int t) { return (const char*)g->d_tab.p + g->tab_off[t]; }

void drop_graph(plp_engine* g) {
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->gexec = nullptr;
    g->graph = nullptr;
    g->g_passes = g->g_reset = g->g_grid = g->g_cap = g->g_wpd = g->g_verd = g->g_share = -1;
}

// Dynamic LDS bytes of one decoder block for a capacity, for the loaded
// decoder kind.
int dyn_lds(plp_engine* g, int ep_cap) {
    if (g->blocks)
        return plpk::block_lds_layout(ep_cap, g->box_cap, g->full_box, g->max_types, g->use_rng,
                                      g->wpd,
                                      g->mode == plpk::kBig)
            .total;
    if (g->verd == plpk::kPointState)
        return plpk::pts_layout(ep_cap, g->box_cap, g->use_rng, g->cfg.count_work, g->mode,
                                g->step_cap)
            .total;
    return plpk::lds_layout(ep_cap, g->box_cap, g->use_rng, g->wpd, g->cfg.count_work, g->mode,
                            g->step_cap)
        .total;
}

// The decoder kernel's own static LDS bytes.
int static_lds(plp_engine* g) {
    int st = 0;
    hipError_t e = g->blocks ? plpk::block_static_lds(g->wpd, g->mode, &st)
                             : plpk::decode_static_lds(g->wpd, g->mode, g->verd, &st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        st = 4096;
    }
    return st;
}

// Whether one block's dynamic plus static LDS fits for a capacity.
bool lds_fits(plp_engine* g, int ep_cap) {
    return dyn_lds(g, ep_cap) + static_lds(g) <= g->ctx->max_lds;
}

// Launch shape for an EP capacity: LDS bytes, resident blocks per CU, grid.
int shape(plp_engine* g, int ep_cap, int blocks_per_cu_req, int n_work) {
    const int lds = dyn_lds(g, ep_cap);
    if (!lds_fits(g, ep_cap))
        return fail(PLP_E_EPCAP, "decoder state (" + std::to_string(lds) + " B of LDS for " +
                                     std::to_string(ep_cap) +
                                     " extreme points) exceeds the block LDS limit");
    int occ = 0;
    hipError_t e = g->blocks ? plpk::block_occupancy(g->wpd, g->mode, lds, &occ)
                             : plpk::decode_occupancy(g->wpd, g->mode, g->verd, lds, &occ);
    if (e != hipSuccess) return hip_fail(e, "decoder occupancy");
    if (occ < 1) return fail(PLP_E_EPCAP, "the decoder cannot be resident with this LDS size");
    int bpc = blocks_per_cu_req > 0 ? std::min(blocks_per_cu_req, occ) : occ;
    g->ep_cap = ep_cap;
    g->lds = lds;
    g->occ = occ;
    g->grid = std::max(1, std::min(n_work, bpc * g->ctx->n_cu));
    return PLP_OK;
}

// The largest extreme-point capacity whose LDS one block can hold.
int largest_cap(plp_engine* g) {
    int lo = 64, hi = 1 << 15;
    if (!lds_fits(g, lo)) return 0;
    while (hi - lo > 32) {
        const int mid = (lo + hi) / 2 / 32 * 32;
        if (lds_fits(g, mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

plpk::DecodeArgs make_args(plp_engine* g, int n_passes, int reset, const int32_t* order,
                           int n_work) {
    plpk::DecodeArgs a;
    a.inst = (const InstHdr*)table(g, plp_engine::kInst);
    a.step_type = (const int32_t*)table(g, plp_engine::kStep);
    a.step_min = (const int32_t*)table(g, plp_engine::kStepMin);
    a.ors_tab = (const int32_t*)table(g, plp_engine::kOrs);
    a.dec_inst = (const int32_t*)table(g, plp_engine::kDecInst);
    a.dec_seed = (const uint64_t*)table(g, plp_engine::kDecSeed);
    a.work_order = order;
    a.n_work = n_work;
    a.n_passes = n_passes;
    a.mt = (uint32_t*)g->d_mt.p;
    a.mt0 = (const uint32_t*)g->d_mt0.p;
    a.state = (DecState*)g->d_state.p;
    a.out_rows = (int32_t*)g->h_rows.dev;
    a.out_rec = (plpk::RowRecord*)g->h_rec.dev;
    a.max_items = g->max_items;
    a.ep_cap = g->ep_cap;
    a.box_cap = g->box_cap;
    a.step_cap = g->step_cap;
    a.use_rng = g->use_rng;
    a.use_theta = g->params.use_theta;
    a.reset = reset;
    a.count_work = g->cfg.count_work ? 1 : 0;
    a.mode = g->mode;
    a.verdicts = g->verd;
    a.eps = g->params.eps;
    a.theta = g->params.theta;
    a.mu = g->params.mu;
    a.work_counter = (uint32_t*)g->d_counter.p;
    a.ticks = (uint32_t*)g->d_ticks.p;
    a.traj = 0;
    a.traj_key_cap = 0;
    a.traj_list_cap = 0;
    a.pad_traj = 0;
    a.traj_hdr = nullptr;
    a.traj_step = nullptr;
    a.traj_keys = nullptr;
    a.traj_box = nullptr;
    a.done = (uint32_t*)g->h_done.dev;
    return a;
}

// Shared prefixes: the recording launch's shape for the current decoder
// capacity -- the decoders' waves and verdict kind with at least four
// waves, room for every item, and a list capacity of at least the
// decoders' (twice it, up to what one block's LDS holds) -- and the buffer
// of the recorded trajectories.  Clears g->share_fit when no block can hold
// a trajectory at the decoders' capacity.
int traj_prepare(plp_engine* g) {
    g->t_wpd = g->t_wpd_req > 0 ? g->t_wpd_req : std::max(4, g->wpd);
    const int tv = g->t_verd >= 0 ? g->t_verd : g->verd;
    const int wide_cap = round_up(std::max(2 * g->ep_cap, 256), 32);
    auto lds_of = [&](int cap) {
        return tv == plpk::kPointState
                   ? plpk::pts_layout(cap, g->full_box, 0, 0, g->mode, g->step_cap).total
                   : plpk::lds_layout(cap, g->full_box, 0, g->t_wpd, 0, g->mode, g->step_cap)
                         .total;
    };
    int st = 0;
    if (plpk::decode_static_lds(g->t_wpd, g->mode, tv, &st) != hipSuccess) {
        (void)hipGetLastError();
        st = 4096;
    }
    int cap = wide_cap;
    while (cap > g->ep_cap && lds_of(cap) + st > g->ctx->max_lds) cap -= 32;
    if (cap < g->ep_cap || lds_of(cap) + st > g->ctx->max_lds) {
        g->share_fit = false;
        return PLP_OK;
    }
    int occ = 0;
    hipError_t e = plpk::decode_occupancy(g->t_wpd, g->mode, tv, lds_of(cap), &occ);
    if (e != hipSuccess) return hip_fail(e, "trajectory occupancy");
    if (occ < 1) {
        g->share_fit = false;
        return PLP_OK;
    }
    g->t_cap = cap;
    g->t_lds = lds_of(cap);
    g->t_grid = std::max(1, std::min(g->n_inst, occ * g->ctx->n_cu));
    g->t_key_cap = (size_t)g->step_cap * (size_t)(g->ep_cap + 8);
    const size_t kb = g->mode == plpk::kBig ? 8 : 4;
    const size_t per = sizeof(plpk::TrajHdr) + (size_t)g->step_cap * sizeof(plpk::TrajStep) +
                       (size_t)g->step_cap * 32 + g->t_key_cap * kb;
    return g->d_traj.reserve(per * (size_t)g->n_inst, "recorded trajectories");
}

// Queue order from the run just made: instances by the mean time their
// decoders took (longest first; the cost estimate of the load only orders
// by item count), decoders of an instance by index.  The persistent grid's
// tail then holds the instances that really are short.  (Ordering decoders
// by their own times instead, which splits an instance's seeds across the
// queue, measured 10 % slower on the 256-instance batch.)
int order_by_run_time(plp_engine* g) {
    std::vector<uint32_t> t((size_t)g->n_dec);
    hipError_t e = hipMemcpyAsync(t.data(), g->d_ticks.p, 4 * (size_t)g->n_dec,
                                  hipMemcpyDeviceToHost, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H decoder run times");
    std::vector<double> sum((size_t)g->n_inst, 0.0);
    std::vector<int> cnt((size_t)g->n_inst, 0);
    for (int d = 0; d < g->n_dec; d++) {
        sum[(size_t)g->h_dec_inst[(size_t)d]] += t[(size_t)d];
        cnt[(size_t)g->h_dec_inst[(size_t)d]]++;
    }
    std::vector<double> mean((size_t)g->n_inst, 0.0);
    for (int i = 0; i < g->n_inst; i++)
        mean[(size_t)i] = cnt[(size_t)i] ? sum[(size_t)i] / cnt[(size_t)i] : 0.0;
    for (int d = 0; d < g->n_dec; d++) g->h_order[(size_t)d] = d;
    if (g->n_dec <= g->grid) {
        // one round of the grid: every decoder starts at once, and the first
        // blocks to take work sit on different CUs, so longest first gives
        // each CU one decoder of each length class
        std::stable_sort(g->h_order.begin(), g->h_order.end(),
                         [&](int a, int b) { return t[(size_t)a] > t[(size_t)b]; });
    } else {
        std::stable_sort(g->h_order.begin(), g->h_order.end(), [&](int a, int b) {
            return mean[(size_t)g->h_dec_inst[(size_t)a]] > mean[(size_t)g->h_dec_inst[(size_t)b]];
        });
    }
    // the table sits in the pinned staging copy of the load's tables too
    char* stage = (char*)g->h_tab.p + g->tab_off[plp_engine::kOrder];
    std::memcpy(stage, g->h_order.data(), 4 * (size_t)g->n_dec);
    e = hipMemcpyAsync((char*)g->d_tab.p + g->tab_off[plp_engine::kOrder], stage,
                       4 * (size_t)g->n_dec, hipMemcpyHostToDevice, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "H2D work order");
    g->ordered = true;
    return PLP_OK;
}

// Whether this run records the trajectories first (before traj_prepare).
bool share_wanted(const plp_engine* g) {
    if (g->blocks || g->cfg.count_work || g->share_mode == 1) return false;
    if (g->share_mode == 2) return true;
    return g->share_eps && g->n_dec >= 3 * g->grid;
}

// The trajectory fields of a launch's arguments (traj: 1 record, 2 resume).
void traj_args(plp_engine* g, plpk::DecodeArgs& a, int traj) {
    const size_t n = (size_t)g->n_inst;
    char* p = (char*)g->d_traj.p;
    a.traj = traj;
    a.traj_key_cap = (int32_t)g->t_key_cap;
    a.traj_list_cap = g->ep_cap;
    a.traj_hdr = (plpk::TrajHdr*)p;
    a.traj_step = (plpk::TrajStep*)(p + n * sizeof(plpk::TrajHdr));
    a.traj_box = (int32_t*)(p + n * (sizeof(plpk::TrajHdr) +
                                     (size_t)g->step_cap * sizeof(plpk::TrajStep)));
    a.traj_keys = p + n * (sizeof(plpk::TrajHdr) + (size_t)g->step_cap * sizeof(plpk::TrajStep) +
                           (size_t)g->step_cap * 32);
}

// One launch of the loaded decoder kind over `grid` blocks.
hipError_t launch(plp_engine* g, const plpk::DecodeArgs& a, int grid, hipStream_t st) {
    if (!g->blocks) return plpk::launch_decode(a, g->wpd, grid, g->lds, st);
    plpk::BlockArgs b;
    b.d = a;
    b.shape = (const plpk::BlockShapeRec*)table(g, plp_engine::kShape);
    b.tinfo = (const plpk::BlockTypeRec*)table(g, plp_engine::kTinfo);
    b.by_type = (const int32_t*)table(g, plp_engine::kByType);
    b.vorder = (const int32_t*)table(g, plp_engine::kVorder);
    b.gtab = (const plpk::GreedyEntry*)table(g, plp_engine::kGtab);
    b.topk = g->bp.topk;
    b.pick = g->bp.pick;
    b.max_types = g->max_types;
    return plpk::launch_block(b, g->wpd, grid, g->lds, st);
}

int launch_run(plp_engine* g, int n_passes, int reset) {
    plp_ctx* c = g->ctx;
    plpk::DecodeArgs a = make_args(g, n_passes, reset,
                                   (const int32_t*)table(g, plp_engine::kOrder), g->n_dec);
    hipError_t e;
    int rc;
    // Shared prefixes: one recording launch over the instances, then the
    // decoders, which resume from the recorded trajectories.  Automatic mode:
    // when the decoders are expected to follow their trajectories for long and
    // the queue holds at least three rounds of the persistent grid (with
    // fewer, the launch lasts as long as its slowest decoders, whose early
    // divergence sharing does not shorten, and the recording launch only adds
    // to it).
    plpk::DecodeArgs t{};
    const bool share = share_wanted(g);
    if (share) {
        if ((rc = traj_prepare(g)) != PLP_OK) return rc;
    }
    const bool rec = share && g->share_fit;
    if (rec) {
        t = make_args(g, 1, 1, (const int32_t*)table(g, plp_engine::kTrajOrder), g->n_inst);
        t.ep_cap = g->t_cap;
        t.box_cap = g->full_box;
        t.use_rng = 0;
        t.verdicts = g->t_verd >= 0 ? g->t_verd : g->verd;
        traj_args(g, t, 1);
        traj_args(g, a, 2);
    }
    auto launch_all = [&]() -> hipError_t {
        if (rec) {
            const hipError_t te = plpk::launch_decode(t, g->t_wpd, g->t_grid, g->t_lds, c->stream);
            if (te != hipSuccess) return te;
        }
        return launch(g, a, g->grid, c->stream);
    };
    if (g->cfg.use_graph) {
        if (!g->gexec || g->g_passes != n_passes || g->g_reset != reset || g->g_grid != g->grid ||
            g->g_cap != g->ep_cap || g->g_wpd != g->wpd || g->g_verd != g->verd ||
            g->g_share != (rec ? 1 : 0) || g->g_tbuf != g->d_traj.p ||
            g->g_tshape != g->t_wpd * 4 + t.verdicts) {
            drop_graph(g);
            e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
            hipError_t le = launch_all();
            e = hipStreamEndCapture(c->stream, &g->graph);
            if (le != hipSuccess) return hip_fail(le, "decode capture");
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            g->g_passes = n_passes;
            g->g_reset = reset;
            g->g_grid = g->grid;
            g->g_cap = g->ep_cap;
            g->g_wpd = g->wpd;
            g->g_verd = g->verd;
            g->g_share = rec ? 1 : 0;
            g->g_tbuf = g->d_traj.p;
            g->g_tshape = g->t_wpd * 4 + t.verdicts;
        }
        e = hipGraphLaunch(g->gexec, c->stream);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        return PLP_OK;
    }
    e = launch_all();
    if (e != hipSuccess) return hip_fail(e, "decoder launch");
    return PLP_OK;
}

int read_states(plp_engine* g) {
    g->h_state.resize((size_t)g->n_dec);
    hipError_t e = hipMemcpyAsync(g->h_state.data(), g->d_state.p,
                                  sizeof(DecState) * (size_t)g->n_dec, hipMemcpyDeviceToHost,
                                  g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H decoder states");
    g->states_fresh = true;
    return PLP_OK;
}

}  // namespace

static bool valid_waves(int w) { return w == 0 || w == 1 || w == 2 || w == 4 || w == 8; }

// The recording launch's requested waves and verdict kind from the config.
static void trace_shape_from_cfg(plp_engine* g) {
    g->t_wpd_req = g->cfg.trace_waves;
    g->t_verd = g->cfg.trace_verdicts == 0 ? -1
              : g->cfg.trace_verdicts == 1 ? plpk::kPairCache : plpk::kPointState;
}

extern "C" int plp_engine_create(plp_ctx* ctx, const plp_engine_config* cfg, plp_engine** out) {
    if (!ctx || !out) return fail(PLP_E_ARG, "NULL argument");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    plp_engine* g = new plp_engine();
    g->ctx = ctx;
    if (cfg) g->cfg = *cfg;
    const int wq = g->cfg.waves_per_decoder;
    if (!valid_waves(wq) || g->cfg.ep_cap < 0 ||
        g->cfg.box_cap < 0 || g->cfg.blocks_per_cu < 0 || g->cfg.verdicts < 0 ||
        g->cfg.verdicts > 2 || g->cfg.share_prefix < 0 || g->cfg.share_prefix > 2 ||
        (g->cfg.trace_waves != 0 && g->cfg.trace_waves != 4 && g->cfg.trace_waves != 8) ||
        g->cfg.trace_verdicts < 0 ||
        g->cfg.trace_verdicts > 2) {
        delete g;
        return fail(PLP_E_ARG, "config: waves_per_decoder must be 0, 1, 2, 4 or 8, "
                               "trace_waves 0, 4 or 8; caps >= 0; verdicts, share_prefix and "
                               "trace_verdicts 0, 1 or 2");
    }
    g->wpd = wq > 0 ? wq : 4;
    g->verd = g->cfg.verdicts == 1 ? plpk::kPairCache : plpk::kPointState;
    // fine-grained: the kernels' stores reach host memory while they run
    g->h_rows.flags = g->h_rec.flags = g->h_done.flags = hipHostMallocMapped | hipHostMallocCoherent;
    if ((e = hipEventCreate(&g->ev0)) != hipSuccess || (e = hipEventCreate(&g->ev1)) != hipSuccess) {
        delete g;
        return hip_fail(e, "hipEventCreate");
    }
    int rc = g->d_counter.reserve(64, "queue head");
    if (rc != PLP_OK) {
        plp_engine_destroy(g);
        return rc;
    }
    if ((e = hipHostMalloc((void**)&g->h_failed, 64, hipHostMallocDefault)) != hipSuccess) {
        g->h_failed = nullptr;
        plp_engine_destroy(g);
        return hip_fail(e, "pinned error count");
    }
    *g->h_failed = 0;
    *out = g;
    return PLP_OK;
}

extern "C" void plp_engine_destroy(plp_engine* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    drop_graph(g);
    for (DevBuf* b : {&g->d_tab, &g->d_retry, &g->d_mt, &g->d_mt0, &g->d_state, &g->d_counter,
                      &g->d_traj, &g->d_ticks})
        b->release();
    g->h_tab.release();
    g->h_rows.release();
    g->h_rec.release();
    g->h_done.release();
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    if (g->h_failed) (void)hipHostFree(g->h_failed);
    delete g;
}

namespace {

// Block catalogue of one instance (blockproof/plp_block.cpp:193-218): shapes
// stored in placement order (plp_block.cpp:156-165), so a shape's index is
// its plan sort key; each type's shape list in shape_precedence order
// (plp_block.cpp:38-45); the types in item-volume order (plp_block.cpp:134-138).
struct BlockTables {
    std::vector<plpk::BlockShapeRec> shape;
    std::vector<plpk::BlockTypeRec> tinfo;
    std::vector<int32_t> by_type;   // instance-local shape indices
    std::vector<int32_t> vorder;
    std::vector<plpk::GreedyEntry> gtab;   // per type, r = 1 .. qty
};

void block_tables(const plp::Instance& in, const plp_block_params& bp, int32_t by_type_base,
                  BlockTables* out) {
    struct Gen {
        int ti, oi, a, b, k, cnt, bw, bd, bh;
    };
    std::vector<Gen> cat;
    const int nt = (int)in.items.size();
    std::vector<plpk::BlockTypeRec> tinfo((size_t)nt);
    for (int ti = 0; ti < nt; ti++) {
        const plp::Item& it = in.items[(size_t)ti];
        const auto ors = plp::orientations(it);
        plpk::BlockTypeRec& r = tinfo[(size_t)ti];
        r = plpk::BlockTypeRec{it.qty, ors.empty() ? 0 : 1, 0, 0, kNoBox, kNoBox, kNoBox, 0};
        for (size_t oi = 0; oi < ors.size(); oi++) {
            const int w = ors[oi][0], d = ors[oi][1], h = ors[oi][2];
            r.wmin = std::min(r.wmin, w);
            r.dmin = std::min(r.dmin, d);
            r.hmin = std::min(r.hmin, h);
            for (int a = 1; a <= bp.amax && a * w <= in.L; a++)
                for (int b = 1; b <= bp.bmax && b * d <= in.W; b++)
                    for (int k = 1; k <= bp.kmax && k * h <= in.H; k++) {
                        const long long cnt = (long long)a * b * k;
                        if (cnt > it.qty) break;
                        cat.push_back({ti, (int)oi, a, b, k, (int)cnt, a * w, b * d, k * h});
                    }
        }
    }
    std::vector<int> by_rank(cat.size());
    for (size_t i = 0; i < cat.size(); i++) by_rank[i] = (int)i;
    std::sort(by_rank.begin(), by_rank.end(), [&](int x, int y) {
        const Gen& s = cat[(size_t)x];
        const Gen& t = cat[(size_t)y];
        const long long vs = (long long)s.bw * s.bd * s.bh, vt = (long long)t.bw * t.bd * t.bh;
        if (vs != vt) return vs > vt;
        if (s.ti != t.ti) return s.ti < t.ti;
        if (s.oi != t.oi) return s.oi < t.oi;
        if (s.k != t.k) return s.k > t.k;
        if (s.a != t.a) return s.a > t.a;
        return s.b > t.b;
    });
    std::vector<int> rank(cat.size());
    out->shape.clear();
    for (size_t r = 0; r < by_rank.size(); r++) {
        const Gen& g = cat[(size_t)by_rank[r]];
        rank[(size_t)by_rank[r]] = (int)r;
        out->shape.push_back(plpk::BlockShapeRec{g.ti, g.a, g.b, g.k, g.cnt, g.bw, g.bd, g.bh});
    }
    out->by_type.clear();
    for (int ti = 0; ti < nt; ti++) {
        std::vector<int> mine;
        for (size_t i = 0; i < cat.size(); i++)
            if (cat[i].ti == ti) mine.push_back((int)i);
        std::sort(mine.begin(), mine.end(), [&](int x, int y) {
            const Gen& p = cat[(size_t)x];
            const Gen& q = cat[(size_t)y];
            if (p.cnt != q.cnt) return p.cnt > q.cnt;
            if (p.bh != q.bh) return p.bh > q.bh;
            if (p.a != q.a) return p.a > q.a;
            if (p.b != q.b) return p.b > q.b;
            return p.oi < q.oi;
        });
        tinfo[(size_t)ti].shp_off = by_type_base + (int32_t)out->by_type.size();
        tinfo[(size_t)ti].shp_cnt = (int32_t)mine.size();
        for (int i : mine) out->by_type.push_back(rank[(size_t)i]);
        // the precedence list is cnt-descending, so the shapes that fit r
        // copies are the suffix after every cnt > r
        tinfo[(size_t)ti].gt_off = (int32_t)out->gtab.size();
        std::vector<plpk::GreedyEntry> col((size_t)std::max(0, in.items[(size_t)ti].qty));
        for (int r = (int)col.size(); r >= 1; r--) {
            size_t start = 0;
            while (start < mine.size() && cat[(size_t)mine[start]].cnt > r) start++;
            plpk::GreedyEntry& e = col[(size_t)r - 1];
            e.start = (int32_t)start;
            e.s0 = start < mine.size() ? rank[(size_t)mine[start]] : -1;
            e.cnt0 = start < mine.size() ? cat[(size_t)mine[start]].cnt : 0;
        }
        out->gtab.insert(out->gtab.end(), col.begin(), col.end());
    }
    out->tinfo = tinfo;
    out->vorder.resize((size_t)nt);
    for (int t = 0; t < nt; t++) out->vorder[(size_t)t] = t;
    std::sort(out->vorder.begin(), out->vorder.end(), [&](int x, int y) {
        const plp::Item& a = in.items[(size_t)x];
        const plp::Item& b = in.items[(size_t)y];
        const long long vx = (long long)a.l * a.w * a.h, vy = (long long)b.l * b.w * b.h;
        if (vx != vy) return vx > vy;
        return x < y;
    });
}

int load_impl(plp_engine* g, const plp_instance_desc* insts, int32_t n_inst,
              const int32_t* dec_inst, const uint64_t* dec_seed, int32_t n_dec,
              const plp_params* params, const plp_block_params* bp);

}  // namespace

extern "C" int plp_engine_load(plp_engine* g, const plp_instance_desc* insts, int32_t n_inst,
                               const int32_t* dec_inst, const uint64_t* dec_seed, int32_t n_dec,
                               const plp_params* params) {
    return load_impl(g, insts, n_inst, dec_inst, dec_seed, n_dec, params, nullptr);
}

extern "C" int plp_engine_load_blocks(plp_engine* g, const plp_instance_desc* insts,
                                      int32_t n_inst, const int32_t* dec_inst,
                                      const uint64_t* dec_seed, int32_t n_dec,
                                      const plp_params* params, const plp_block_params* blocks) {
    if (!blocks) return fail(PLP_E_ARG, "NULL block parameters");
    plp_block_params bp = *blocks;
    // the reference CLI clamps these to >= 1 (plp_block.cpp:325-327, 315)
    bp.amax = std::max(1, bp.amax);
    bp.bmax = std::max(1, bp.bmax);
    bp.kmax = std::max(1, bp.kmax);
    bp.topk = std::max(1, bp.topk);
    if (bp.amax > 255 || bp.bmax > 255 || bp.kmax > 255)
        return fail(PLP_E_ARG, "amax, bmax and kmax must be <= 255");
    if (bp.pick < 0 || bp.pick > 2) return fail(PLP_E_ARG, "pick must be 0, 1 or 2");
    if (g && g->cfg.count_work) return fail(PLP_E_ARG, "count_work is not supported for blocks");
    return load_impl(g, insts, n_inst, dec_inst, dec_seed, n_dec, params, &bp);
}

namespace {

int load_impl(plp_engine* g, const plp_instance_desc* insts, int32_t n_inst,
              const int32_t* dec_inst, const uint64_t* dec_seed, int32_t n_dec,
              const plp_params* params, const plp_block_params* bp) {
    if (!g || !insts || !dec_inst || !dec_seed || !params) return fail(PLP_E_ARG, "NULL argument");
    if (n_inst < 1 || n_dec < 1) return fail(PLP_E_ARG, "need at least one instance and decoder");
    if (params->order != 0 && params->order != 1) return fail(PLP_E_ARG, "order must be 0 or 1");
    if (!(params->eps >= 0.0) || params->eps > 1.0) return fail(PLP_E_ARG, "eps must be in [0, 1]");
    hipError_t e = hipSetDevice(g->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "engine sync");
    drop_graph(g);
    g->loaded = false;
    std::vector<InstHdr> hdr((size_t)n_inst);
    std::vector<int32_t> steps;
    std::vector<int32_t> step_min;   // per step: min w, d, h, 0, max w, d, h, 0 over the later steps
    std::vector<int32_t> ors;
    BlockTables bt_all;               // block engine: every instance's tables
    int max_types = 1;
    std::vector<int64_t> cost((size_t)n_inst, 0);
    int max_items = 1;
    int wide = 0, big = 0;
    for (int i = 0; i < n_inst; i++) {
        plp::Instance in;
        int rc = to_instance(&insts[i], &in);
        if (rc != PLP_OK) return fail(rc, "instance " + std::to_string(i) + ": " + plp_last_error());
        std::vector<int> seq = plp::item_sequence(in, params->order == 0 ? "volume" : "input");
        InstHdr& h = hdr[(size_t)i];
        std::memset(&h, 0, sizeof h);
        h.L = in.L;
        h.W = in.W;
        h.H = in.H;
        h.n_steps = (int32_t)seq.size();
        h.step_off = (int32_t)steps.size();
        h.ors_off = (int32_t)(ors.size() / plpk::kOrsRow);
        h.n_types = (int32_t)in.items.size();
        if (h.n_types >= (1 << 24))   // the decoder's step table packs the type into 24 bits
            return fail(PLP_E_RANGE, "instance " + std::to_string(i) + ": more than 2^24 - 1 types");
        h.xbits = bitlen(in.L);
        h.ybits = bitlen(in.W);
        // 32-bit keys and sums need bitlen(L) + bitlen(W) + bitlen(H) <= 31
        // (then L * W < 2^30); beyond, the batch runs with 64-bit ones
        if (h.xbits + h.ybits + bitlen(in.H) > 31) big = 1;
        max_types = std::max(max_types, h.n_types);
        if (bp) {
            BlockTables bt;
            block_tables(in, *bp, (int32_t)bt_all.by_type.size(), &bt);
            h.shape_off = (int32_t)bt_all.shape.size();
            h.tinfo_off = (int32_t)bt_all.tinfo.size();
            h.vorder_off = (int32_t)bt_all.vorder.size();
            h.gtab_off = (int32_t)bt_all.gtab.size();
            bt_all.shape.insert(bt_all.shape.end(), bt.shape.begin(), bt.shape.end());
            bt_all.tinfo.insert(bt_all.tinfo.end(), bt.tinfo.begin(), bt.tinfo.end());
            bt_all.by_type.insert(bt_all.by_type.end(), bt.by_type.begin(), bt.by_type.end());
            bt_all.vorder.insert(bt_all.vorder.end(), bt.vorder.begin(), bt.vorder.end());
            bt_all.gtab.insert(bt_all.gtab.end(), bt.gtab.begin(), bt.gtab.end());
        }
        steps.insert(steps.end(), seq.begin(), seq.end());
        // Componentwise smallest orientation of every type; a point where a
        // box of the smallest dimensions of all later steps cannot go can
        // host none of them.
        // The componentwise largest of the later steps bounds which placed
        // boxes can still meet or support a later box at a point.
        std::vector<std::array<int32_t, 3>> tmin(in.items.size(), {kNoBox, kNoBox, kNoBox});
        std::vector<std::array<int32_t, 3>> tmax(in.items.size(), {0, 0, 0});
        for (size_t t = 0; t < in.items.size(); t++)
            for (const auto& o : plp::orientations(in.items[t]))
                for (int c = 0; c < 3; c++) {
                    tmin[t][(size_t)c] = std::min(tmin[t][(size_t)c], (int32_t)o[(size_t)c]);
                    tmax[t][(size_t)c] = std::max(tmax[t][(size_t)c], (int32_t)o[(size_t)c]);
                    // 16-bit point-state entries hold sides up to 65535
                    if (o[(size_t)c] > 65535) wide = 1;
                }
        {
            std::vector<int32_t> mins(8 * seq.size(), 0);
            std::array<int32_t, 3> m = {kNoBox, kNoBox, kNoBox}, mx = {0, 0, 0};
            for (size_t k = seq.size(); k-- > 0;) {
                for (int c = 0; c < 3; c++) {
                    mins[8 * k + (size_t)c] = m[(size_t)c];
                    mins[8 * k + 4 + (size_t)c] = mx[(size_t)c];
                }
                for (int c = 0; c < 3; c++) {
                    m[(size_t)c] = std::min(m[(size_t)c], tmin[(size_t)seq[k]][(size_t)c]);
                    mx[(size_t)c] = std::max(mx[(size_t)c], tmax[(size_t)seq[k]][(size_t)c]);
                }
            }
            step_min.insert(step_min.end(), mins.begin(), mins.end());
        }
        for (const plp::Item& it : in.items) {
            auto o = plp::orientations(it);
            int32_t row[plpk::kOrsRow] = {0};
            row[0] = (int32_t)o.size();
            for (size_t k = 0; k < o.size(); k++) {
                row[1 + 3 * k] = o[k][0];
                row[2 + 3 * k] = o[k][1];
                row[3 + 3 * k] = o[k][2];
                if ((int64_t)o[k][0] * o[k][1] > plpk::kNarrowMaxArea) wide = 1;
            }
            ors.insert(ors.end(), row, row + plpk::kOrsRow);
        }
        max_items = std::max(max_items, (int)seq.size());
        cost[(size_t)i] = (int64_t)seq.size() * (int64_t)seq.size();
    }
    if (ors.empty()) ors.assign(plpk::kOrsRow, 0);
    if (steps.empty()) steps.push_back(0);
    if (step_min.empty()) step_min = {kNoBox, kNoBox, kNoBox, 0, 0, 0, 0, 0};
    for (int d = 0; d < n_dec; d++)
        if (dec_inst[d] < 0 || dec_inst[d] >= n_inst)
            return fail(PLP_E_ARG, "dec_inst out of range");
    g->n_inst = n_inst;
    g->n_dec = n_dec;
    g->max_items = max_items;
    g->full_box = g->step_cap = round_up(max_items, 8);   // sentinel-padded groups of 8
    // the block engine keeps every item (its blocks decompose into items)
    g->box_cap = g->cfg.box_cap <= 0 ? g->full_box
                                     : std::min(g->full_box, round_up(g->cfg.box_cap, 8));
    g->mode = big ? plpk::kBig : wide ? plpk::kWide : plpk::kNarrow;
    g->params = *params;
    g->blocks = bp != nullptr;
    if (bp) g->bp = *bp;
    // automatic waves per decoder: one for the block engine (its chain is
    // the latency of one wave, and one wave tests several plan entries at
    // once on short lists), four for the item decoder
    if (g->cfg.waves_per_decoder == 0 && !g->wpd_tuned) g->wpd = g->blocks ? 1 : 4;
    g->max_types = max_types;
    g->use_rng = bp ? (params->eps > 0.0 ? 1 : 0)
                    : ((params->first_fit || params->eps == 0.0) ? 0 : 1);
    if (bt_all.shape.empty()) bt_all.shape.push_back(plpk::BlockShapeRec{});
    if (bt_all.tinfo.empty()) bt_all.tinfo.push_back(plpk::BlockTypeRec{});
    if (bt_all.by_type.empty()) bt_all.by_type.push_back(0);
    if (bt_all.vorder.empty()) bt_all.vorder.push_back(0);
    if (bt_all.gtab.empty()) bt_all.gtab.push_back(plpk::GreedyEntry{});
    g->h_dec_inst.assign(dec_inst, dec_inst + n_dec);
    g->h_seed.assign(dec_seed, dec_seed + n_dec);
    // Work order: most expensive instances first (longest-processing-time
    // first keeps the persistent grid's tail short); stable by decoder index.
    g->h_order.resize((size_t)n_dec);
    for (int d = 0; d < n_dec; d++) g->h_order[(size_t)d] = d;
    std::stable_sort(g->h_order.begin(), g->h_order.end(), [&](int a, int b) {
        return cost[(size_t)dec_inst[a]] > cost[(size_t)dec_inst[b]];
    });
    // Shared prefixes: the instances' trajectories are recorded in the same
    // order; the automatic mode expects a decoder to follow its trajectory
    // for a good part of a pass when the RNG is unused or 1/eps placing steps
    // (before its first eps pick) are at least an eighth of the item count
    std::vector<int32_t> traj_order((size_t)n_inst);
    for (int i = 0; i < n_inst; i++) traj_order[(size_t)i] = i;
    std::stable_sort(traj_order.begin(), traj_order.end(),
                     [&](int a, int b) { return cost[(size_t)a] > cost[(size_t)b]; });
    g->share_mode = g->cfg.share_prefix;
    g->share_eps = g->use_rng == 0 || params->eps * (double)max_items <= 8.0;
    g->share_fit = true;
    const size_t rows_bytes = (size_t)n_dec * (size_t)max_items * 7 * sizeof(int32_t);
    struct Up {
        const void* src;
        size_t n;
    } ups[plp_engine::kNumTabs] = {
        {hdr.data(), sizeof(InstHdr) * hdr.size()},
        {steps.data(), 4 * steps.size()},
        {step_min.data(), 4 * step_min.size()},
        {bt_all.shape.data(), sizeof(plpk::BlockShapeRec) * bt_all.shape.size()},
        {bt_all.tinfo.data(), sizeof(plpk::BlockTypeRec) * bt_all.tinfo.size()},
        {bt_all.by_type.data(), 4 * bt_all.by_type.size()},
        {bt_all.vorder.data(), 4 * bt_all.vorder.size()},
        {bt_all.gtab.data(), sizeof(plpk::GreedyEntry) * bt_all.gtab.size()},
        {ors.data(), 4 * ors.size()},
        {dec_inst, 4 * (size_t)n_dec},
        {dec_seed, 8 * (size_t)n_dec},
        {g->h_order.data(), 4 * (size_t)n_dec},
        {traj_order.data(), 4 * (size_t)n_inst},
    };
    size_t tab_bytes = 0;
    for (int t = 0; t < plp_engine::kNumTabs; t++) {
        g->tab_off[t] = tab_bytes;
        tab_bytes += (ups[t].n + 255) & ~(size_t)255;
    }
    int rc;
    if ((rc = g->d_tab.reserve(tab_bytes, "engine tables")) ||
        (rc = g->h_tab.reserve(tab_bytes, "pinned table staging")) ||
        (rc = g->d_retry.reserve(4 * (size_t)n_dec, "retry order")) ||
        (rc = g->d_mt.reserve(4 * plpk::kMtWords * (size_t)n_dec, "RNG states")) ||
        (rc = g->d_mt0.reserve(4 * plpk::kMtWords * (size_t)(g->use_rng ? n_dec : 1),
                               "seeded RNG states")) ||
        (rc = g->d_state.reserve(sizeof(DecState) * (size_t)n_dec, "decoder states")) ||
        (rc = g->d_ticks.reserve(4 * (size_t)n_dec, "decoder run times")) ||
        (rc = g->h_rows.reserve(rows_bytes, "pinned result packings")) ||
        (rc = g->h_rec.reserve(sizeof(plpk::RowRecord) * (size_t)n_dec, "pinned result records")) ||
        (rc = g->h_done.reserve(4 * (size_t)n_dec, "pinned completion words")))
        return rc;
    std::memset(g->h_done.p, 0, 4 * (size_t)n_dec);
    g->runs_unsynced = 0;
    // One copy: the stream has drained (above), so the staging buffer is free;
    // the decode launches that read the tables follow the copy on the stream.
    for (int t = 0; t < plp_engine::kNumTabs; t++)
        if (ups[t].n) std::memcpy((char*)g->h_tab.p + g->tab_off[t], ups[t].src, ups[t].n);
    e = hipMemcpyAsync(g->d_tab.p, g->h_tab.p, tab_bytes, hipMemcpyHostToDevice, g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "H2D engine tables");
    // every decoder's seeded MT19937 state, once per load: a reset copies it
    if (g->use_rng) {
        e = plpk::launch_mt_seed((const uint64_t*)table(g, plp_engine::kDecSeed),
                                 (uint32_t*)g->d_mt0.p, n_dec, g->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "RNG seeding");
    }
    int cap = g->cfg.ep_cap;
    if (cap <= 0) {
        // Automatic capacity: the largest (up to 4 points per item) that
        // keeps the residency the registers allow at the smallest capacity;
        // a decoder that needs more is rerun with more (sync()).
        cap = 64;
        if ((rc = shape(g, cap, g->cfg.blocks_per_cu, n_dec)) != PLP_OK) return rc;
        const int occ_min = g->occ;
        const int want = g->cfg.blocks_per_cu > 0 ? std::min(g->cfg.blocks_per_cu, occ_min) : occ_min;
        for (int c = round_up(4 * max_items + 64, 32); c > 64; c -= 32)
            if (shape(g, c, g->cfg.blocks_per_cu, n_dec) == PLP_OK && g->occ >= want) {
                cap = c;
                break;
            }
    }
    rc = shape(g, cap, g->cfg.blocks_per_cu, n_dec);
    if (rc != PLP_OK) return rc;
    g->passes_done = 0;
    g->need_reset = true;
    g->states_fresh = false;
    g->runs = 0;
    g->ordered = false;
    trace_shape_from_cfg(g);
    g->loaded = true;
    return PLP_OK;
}

}  // namespace

extern "C" int plp_engine_reset(plp_engine* g) {
    if (!g || !g->loaded) return fail(PLP_E_ARG, "engine not loaded");
    g->need_reset = true;
    g->passes_done = 0;
    g->states_fresh = false;
    return PLP_OK;
}

extern "C" int plp_engine_run(plp_engine* g, int32_t n_passes) {
    if (!g || !g->loaded) return fail(PLP_E_ARG, "engine not loaded");
    if (n_passes < 1) return fail(PLP_E_ARG, "n_passes must be >= 1");
    hipError_t e = hipSetDevice(g->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int reset = g->need_reset ? 1 : 0;
    // the second run of a load takes its instances in the order of the times
    // the first run measured (a batch decoded once, like a served request,
    // never pays for it)
    if (g->runs == 1 && !g->ordered && !g->blocks && (g->n_inst > 1 || g->n_dec <= g->grid)) {
        const int rc0 = order_by_run_time(g);
        if (rc0 != PLP_OK) return rc0;
    }
    g->runs++;
    g->runs_unsynced++;
    if ((e = hipEventRecord(g->ev0, g->ctx->stream)) != hipSuccess) return hip_fail(e, "event");
    int rc = launch_run(g, n_passes, reset);
    if (rc != PLP_OK) return rc;
    if ((e = hipEventRecord(g->ev1, g->ctx->stream)) != hipSuccess) return hip_fail(e, "event");
    // the error count rides the stream to pinned memory; sync reads it
    e = hipMemcpyAsync(g->h_failed, (const uint32_t*)g->d_counter.p + 1, sizeof(uint32_t),
                       hipMemcpyDeviceToHost, g->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H error count");
    g->passes_done = reset ? n_passes : g->passes_done + n_passes;
    g->need_reset = false;
    g->states_fresh = false;
    return PLP_OK;
}

static int sync_impl(plp_engine* g) {
    plp_ctx* c = g->ctx;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "decode sync");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, g->ev0, g->ev1) != hipSuccess) {
        (void)hipGetLastError();
        ms = 0.0f;
    }
    g->last_ms = ms;
    if (g->passes_done == 0) return PLP_OK;
    if (*g->h_failed == 0) return PLP_OK;   // no decoder failed: states stay on the device
    int rc = read_states(g);
    if (rc != PLP_OK) return rc;
    // Capacity escalation: a decoder whose extreme-point list or placed
    // boxes outgrew its LDS capacity stopped with err set; rerun exactly those
    // decoders from their seed with room for every item and the largest
    // extreme-point capacity one block can then hold.  Results do not depend
    // on the capacities, so the rerun reproduces what an unlimited run would
    // give.  Afterwards the capacities are the most any decoder needed, which
    // these decoders' later runs fit with the best residency.
    bool escalated = false;
    for (;;) {
        std::vector<int32_t> failed;
        int need = 0;
        for (int d = 0; d < g->n_dec; d++)
            if (g->h_state[(size_t)d].err != 0) {
                failed.push_back(d);
                need = std::max(need, (int)g->h_state[(size_t)d].ep_need);
            }
        if (failed.empty()) break;
        const bool box_grows = g->box_cap < g->full_box;
        g->box_cap = g->full_box;
        const int cap = largest_cap(g);
        if ((cap <= g->ep_cap && !box_grows) || cap < need || shape(g, cap, 1, g->n_dec) != PLP_OK)
            return fail(PLP_E_EPCAP, std::to_string(failed.size()) +
                                         " decoder(s) need more than " + std::to_string(g->ep_cap) +
                                         " extreme points, beyond the block LDS limit");
        drop_graph(g);
        e = hipMemcpyAsync(g->d_retry.p, failed.data(), 4 * failed.size(), hipMemcpyHostToDevice,
                           c->stream);
        if (e != hipSuccess) return hip_fail(e, "H2D retry order");
        g->grid = std::max(1, std::min((int)failed.size(), g->grid));
        plpk::DecodeArgs a = make_args(g, g->passes_done, 1, (const int32_t*)g->d_retry.p,
                                       (int)failed.size());
        if ((e = hipEventRecord(g->ev0, c->stream)) != hipSuccess) return hip_fail(e, "event");
        e = launch(g, a, g->grid, c->stream);
        if (e != hipSuccess) return hip_fail(e, "decode retry launch");
        if ((e = hipEventRecord(g->ev1, c->stream)) != hipSuccess) return hip_fail(e, "event");
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(e, "retry sync");
        if (hipEventElapsedTime(&ms, g->ev0, g->ev1) == hipSuccess) g->last_ms += ms;
        else (void)hipGetLastError();
        if ((rc = read_states(g)) != PLP_OK) return rc;
        escalated = true;
    }
    if (escalated) {
        int need_all = 64, box_all = 1;
        for (const DecState& st : g->h_state) {
            need_all = std::max(need_all, (int)st.ep_need);
            box_all = std::max(box_all, (int)st.box_need);
        }
        g->box_cap = std::min(g->full_box, round_up(box_all, 8));
        if ((rc = shape(g, round_up(need_all, 32), g->cfg.blocks_per_cu, g->n_dec)) != PLP_OK)
            return rc;
        drop_graph(g);
    }
    return PLP_OK;
}

extern "C" int plp_engine_sync(plp_engine* g) {
    if (!g || !g->loaded) return fail(PLP_E_ARG, "engine not loaded");
    const int rc = sync_impl(g);
    // the launches of the runs just synchronised (reruns included) have
    // ended: their completion words start the next run at 0
    std::memset(g->h_done.p, 0, 4 * (size_t)g->n_dec);
    g->runs_unsynced = 0;
    return rc;
}

namespace {

// One decoder's row and packing from the pinned result memory.
void copy_result(const plp_engine* g, int d, plp_row_result* rows, int32_t* placements) {
    const plpk::RowRecord* rec = (const plpk::RowRecord*)g->h_rec.p;
    std::memcpy(rows + d, rec + d, sizeof(plp_row_result));
    if (placements) {
        const size_t stride = (size_t)g->max_items * 7;
        std::memcpy(placements + stride * (size_t)d,
                    (const int32_t*)g->h_rows.p + stride * (size_t)d,
                    4 * 7 * (size_t)std::max(0, rows[d].n_boxes));
    }
}

}  // namespace

extern "C" int plp_engine_sync_fetch(plp_engine* g, plp_row_result* rows, int32_t* placements) {
    if (!g || !g->loaded || !rows) return fail(PLP_E_ARG, "engine not loaded");
    if (g->passes_done == 0) return fail(PLP_E_ARG, "nothing decoded since the last load or reset");
    hipError_t e = hipSetDevice(g->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // Overlapped only for one outstanding run: the completion words then
    // tell exactly which decoders of this run have finished.
    if (g->cfg.count_work || g->runs_unsynced != 1) {
        const int rc = plp_engine_sync(g);
        return rc != PLP_OK ? rc : plp_engine_fetch(g, rows, placements);
    }
    const uint32_t* done = (const uint32_t*)g->h_done.p;
    std::atomic<bool> missed{false};
    // one part per pool thread (the caller included), each polling its
    // decoders' completion words and copying a decoder's results as soon as
    // its word is set; after the run's end event has completed, one more scan
    // finds every remaining word set
    const int parts = copy_pool().threads();
    const int per = (g->n_dec + parts - 1) / parts;
    copy_pool().run(parts, [&](int i) {
        const int d0 = i * per, d1 = std::min(g->n_dec, d0 + per);
        std::vector<int> left;
        for (int d = d0; d < d1; d++) left.push_back(d);
        bool ended = false;
        while (!left.empty()) {
            size_t k = 0;
            for (const int d : left) {
                if (__atomic_load_n(done + d, __ATOMIC_ACQUIRE) != 0u) copy_result(g, d, rows, placements);
                else left[k++] = d;
            }
            const bool progress = k < left.size();
            left.resize(k);
            if (left.empty() || progress) continue;
            if (ended) {
                missed.store(true);
                break;
            }
            // the run has ended, or failed (any answer but "not ready"):
            // one more scan, then the sync reports what happened
            if (hipEventQuery(g->ev1) != hipErrorNotReady) ended = true;
            else
                for (int s = 0; s < 64; s++) __builtin_ia32_pause();
        }
    });
    const int rc = plp_engine_sync(g);
    if (rc != PLP_OK) return rc;
    // decoders rerun at a larger capacity (or a word not seen): copy all again
    if (missed.load() || *g->h_failed != 0) return plp_engine_fetch(g, rows, placements);
    return PLP_OK;
}

extern "C" int plp_engine_max_items(plp_engine* g, int32_t* max_items) {
    if (!g || !g->loaded || !max_items) return fail(PLP_E_ARG, "engine not loaded");
    *max_items = g->max_items;
    return PLP_OK;
}

extern "C" int plp_engine_fetch(plp_engine* g, plp_row_result* rows, int32_t* placements) {
    if (!g || !g->loaded || !rows) return fail(PLP_E_ARG, "engine not loaded");
    if (g->passes_done == 0) return fail(PLP_E_ARG, "nothing decoded since the last load or reset");
    hipError_t e = hipSetDevice(g->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(g->ctx->stream)) != hipSuccess) return hip_fail(e, "fetch sync");
    // The kernels wrote the records and packings into pinned host memory.
    const plpk::RowRecord* rec = (const plpk::RowRecord*)g->h_rec.p;
    std::memcpy(rows, rec, sizeof(plp_row_result) * (size_t)g->n_dec);
    if (placements) {
        const size_t stride = (size_t)g->max_items * 7;
        const int32_t* src = (const int32_t*)g->h_rows.p;
        auto copy = [&](int d0, int d1) {
            for (int d = d0; d < d1; d++)
                std::memcpy(placements + stride * (size_t)d, src + stride * (size_t)d,
                            4 * 7 * (size_t)std::max(0, rec[d].n_boxes));
        };
        // a result slab past 4 MiB is copied by the pool's threads in 32 parts
        if ((size_t)g->n_dec * g->max_items * 28 < (4u << 20)) {
            copy(0, g->n_dec);
        } else {
            const int parts = 32, per = (g->n_dec + parts - 1) / parts;
            copy_pool().run(parts, [&](int i) {
                copy(i * per, std::min(g->n_dec, (i + 1) * per));
            });
        }
    }
    return PLP_OK;
}

extern "C" int plp_engine_results(plp_engine* g, const plp_row_result** rows,
                                  const int32_t** placements, int32_t* row_stride) {
    if (!g || !g->loaded || !rows) return fail(PLP_E_ARG, "engine not loaded");
    if (g->passes_done == 0) return fail(PLP_E_ARG, "nothing decoded since the last load or reset");
    hipError_t e = hipSetDevice(g->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = hipStreamSynchronize(g->ctx->stream)) != hipSuccess) return hip_fail(e, "results sync");
    *rows = (const plp_row_result*)g->h_rec.p;
    if (placements) *placements = (const int32_t*)g->h_rows.p;
    if (row_stride) *row_stride = g->max_items * 7;
    return PLP_OK;
}

extern "C" int plp_engine_last_ms(plp_engine* g, float* ms) {
    if (!g || !ms) return fail(PLP_E_ARG, "NULL argument");
    *ms = g->last_ms;
    return PLP_OK;
}

extern "C" int plp_engine_describe(plp_engine* g, int32_t* wpd, int32_t* ep_cap, int32_t* lds,
                                   int32_t* grid) {
    if (!g || !g->loaded) return fail(PLP_E_ARG, "engine not loaded");
    if (wpd) *wpd = g->wpd;
    if (ep_cap) *ep_cap = g->ep_cap;
    if (lds) *lds = g->lds;
    if (grid) *grid = g->grid;
    return PLP_OK;
}

extern "C" int plp_engine_autotune(plp_engine* g, int32_t n_passes, plp_engine_config* best,
                                   float* best_ms) {
    if (!g || !g->loaded) return fail(PLP_E_ARG, "engine not loaded");
    if (n_passes < 1) return fail(PLP_E_ARG, "n_passes must be >= 1");
    const plp_engine_config saved = g->cfg;
    // Probe run: learn the largest extreme-point list and the most placed
    // boxes of this batch; every candidate shape then holds exactly those
    // boxes (less LDS per decoder, more decoders per CU).
    g->need_reset = true;
    int rc = plp_engine_run(g, n_passes);
    if (rc == PLP_OK) rc = plp_engine_sync(g);
    if (rc == PLP_OK && !g->states_fresh) rc = read_states(g);
    if (rc != PLP_OK) return rc;
    int need = 1, box_need = 1;
    for (const DecState& s : g->h_state) {
        need = std::max(need, (int)s.ep_need);
        box_need = std::max(box_need, (int)s.box_need);
    }
    if (!g->blocks && (rc = order_by_run_time(g)) != PLP_OK) return rc;
    g->box_cap = std::min(g->full_box, round_up(box_need, 8));
    std::vector<int> caps;
    for (int num : {4, 5, 6})   // 1.0x, 1.25x, 1.5x the observed peak
        caps.push_back(std::max(64, round_up(need * num / 4, 32)));
    caps.push_back(g->ep_cap);
    std::sort(caps.begin(), caps.end());
    caps.erase(std::unique(caps.begin(), caps.end()), caps.end());
    float bms = -1.0f;
    plp_engine_config bcfg = saved;
    const int saved_wpd = g->wpd;
    const int saved_verd = g->verd;
    int best_verd = saved_verd;
    // verdict kinds to try: the configured one, or both for the item decoder
    std::vector<int> verds = {saved_verd};
    if (!g->blocks && saved.verdicts == 0) verds = {plpk::kPointState, plpk::kPairCache};
    // every candidate once, then the three fastest again (the fastest of
    // three runs each): a single run varies by a few per cent
    struct Cand {
        float ms;
        plp_engine_config cfg;
        int verd;
    };
    std::vector<Cand> cands;
    auto apply = [&](const Cand& c) {
        g->wpd = c.cfg.waves_per_decoder;
        g->verd = c.verd;
        g->cfg.ep_cap = c.cfg.ep_cap;
        g->cfg.blocks_per_cu = c.cfg.blocks_per_cu;
        g->box_cap = c.cfg.box_cap;
        return shape(g, c.cfg.ep_cap, c.cfg.blocks_per_cu, g->n_dec);
    };
    auto time_runs = [&](int runs, float* ms) {
        drop_graph(g);
        *ms = -1.0f;
        for (int r = 0; r < runs; r++) {
            g->need_reset = true;
            int rc2 = plp_engine_run(g, n_passes);
            if (rc2 == PLP_OK) rc2 = plp_engine_sync(g);
            if (rc2 != PLP_OK) return rc2;
            if (*ms < 0.0f || g->last_ms < *ms) *ms = g->last_ms;
        }
        return PLP_OK;
    };
    const int box_probe = g->box_cap;
    for (int verd : verds)
    for (int wpd : {1, 2, 4, 8}) {
        if (saved.waves_per_decoder != 0 && wpd != saved.waves_per_decoder) continue;
        g->wpd = wpd;
        g->verd = verd;
        for (int cap : caps) {
            g->cfg.ep_cap = cap;
            g->cfg.blocks_per_cu = 0;
            g->box_cap = box_probe;
            if (shape(g, cap, 0, g->n_dec) != PLP_OK) continue;
            const int occ = g->occ;
            std::vector<int> bpcs = {occ, std::max(1, occ * 3 / 4), std::max(1, occ / 2)};
            bpcs.erase(std::unique(bpcs.begin(), bpcs.end()), bpcs.end());
            for (int bpc : bpcs) {
                g->cfg.blocks_per_cu = bpc;
                g->box_cap = box_probe;
                if (shape(g, cap, bpc, g->n_dec) != PLP_OK) continue;
                float ms = 0.0f;
                if ((rc = time_runs(1, &ms)) != PLP_OK) return rc;
                Cand c;
                c.ms = ms;
                c.cfg = g->cfg;
                c.cfg.ep_cap = g->ep_cap;   // escalation may have raised it
                c.cfg.box_cap = g->box_cap;
                c.cfg.waves_per_decoder = wpd;
                c.verd = verd;
                cands.push_back(c);
            }
        }
    }
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
    if (cands.size() > 3) cands.resize(3);
    for (Cand& c : cands) {
        if (apply(c) != PLP_OK) {
            c.ms = 1e30f;
            continue;
        }
        if ((rc = time_runs(3, &c.ms)) != PLP_OK) return rc;
        c.cfg.ep_cap = g->ep_cap;
        c.cfg.box_cap = g->box_cap;
    }
    for (const Cand& c : cands)
        if (bms < 0.0f || c.ms < bms) {
            bms = c.ms;
            bcfg = c.cfg;
            best_verd = c.verd;
        }
    g->cfg = saved;
    g->cfg.ep_cap = bcfg.ep_cap;
    g->cfg.blocks_per_cu = bcfg.blocks_per_cu;
    g->cfg.box_cap = bcfg.box_cap;
    g->box_cap = bcfg.box_cap > 0 ? bcfg.box_cap : g->box_cap;
    g->wpd = bcfg.waves_per_decoder > 0 ? bcfg.waves_per_decoder : saved_wpd;
    g->wpd_tuned = true;
    g->verd = best_verd;
    rc = shape(g, bcfg.ep_cap, bcfg.blocks_per_cu, g->n_dec);
    if (rc != PLP_OK) return rc;
    drop_graph(g);
    if (!g->blocks && !g->cfg.count_work && saved.share_prefix == 0) {
        // shared prefixes on and off at the chosen shape, the fastest of five
        // runs each; on only when it wins by more than the run-to-run noise
        float t[2] = {-1.0f, -1.0f};
        for (int k = 0; k < 2; k++) {
            g->share_mode = k == 0 ? 1 : 2;
            drop_graph(g);
            for (int r = 0; r < 5; r++) {
                g->need_reset = true;
                rc = plp_engine_run(g, n_passes);
                if (rc == PLP_OK) rc = plp_engine_sync(g);
                if (rc != PLP_OK) return rc;
                if (t[k] < 0.0f || g->last_ms < t[k]) t[k] = g->last_ms;
            }
        }
        g->share_mode = t[1] < 0.97f * t[0] ? 2 : 1;
        bms = g->share_mode == 2 ? t[1] : t[0];
        drop_graph(g);
        if (g->share_mode == 2 && saved.trace_waves == 0 && saved.trace_verdicts == 0) {
            // the recording launch's waves and verdict kind (its own, apart
            // from the decoders'), the fastest of three runs each
            float best_t = t[1];
            int bw = g->t_wpd_req, bv = g->t_verd;
            for (int tw : {4, 8})
                for (int tv : {plpk::kPointState, plpk::kPairCache}) {
                    g->t_wpd_req = tw;
                    g->t_verd = tv;
                    float ms = -1.0f;
                    drop_graph(g);
                    for (int r = 0; r < 3; r++) {
                        g->need_reset = true;
                        rc = plp_engine_run(g, n_passes);
                        if (rc == PLP_OK) rc = plp_engine_sync(g);
                        if (rc != PLP_OK) return rc;
                        if (ms < 0.0f || g->last_ms < ms) ms = g->last_ms;
                    }
                    if (ms < best_t) {
                        best_t = ms;
                        bw = tw;
                        bv = tv;
                    }
                }
            g->t_wpd_req = bw;
            g->t_verd = bv;
            bms = best_t;
            drop_graph(g);
        }
    }
    if (!g->blocks && g->n_dec <= g->grid) {
        // one round of the grid at the chosen shape: the queue takes the
        // decoders by their own times measured at that shape (the probe ran at
        // another one, often with a smaller grid)
        g->need_reset = true;
        rc = plp_engine_run(g, n_passes);
        if (rc == PLP_OK) rc = plp_engine_sync(g);
        if (rc == PLP_OK) rc = order_by_run_time(g);
        if (rc != PLP_OK) return rc;
    }
    g->need_reset = true;
    g->passes_done = 0;
    g->states_fresh = false;
    if (best) {
        *best = g->cfg;
        best->waves_per_decoder = g->wpd;
        best->verdicts = g->blocks ? 0 : (g->verd == plpk::kPairCache ? 1 : 2);
        best->share_prefix = g->blocks ? 0 : g->share_mode;
        const bool tr = !g->blocks && g->share_mode == 2;
        best->trace_waves = tr ? (g->t_wpd_req > 0 ? g->t_wpd_req : std::max(4, g->wpd)) : 0;
        const int tv = g->t_verd >= 0 ? g->t_verd : g->verd;
        best->trace_verdicts = tr ? (tv == plpk::kPairCache ? 1 : 2) : 0;
    }
    if (best_ms) *best_ms = bms;
    return PLP_OK;
}
