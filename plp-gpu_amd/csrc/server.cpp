// server.cpp -- continuous-batching request loop (no reference counterpart;
// the reference decodes one instance per process).
//
// Each admitted request owns a slot: a context (its own non-blocking HIP
// stream) and an engine whose device buffers are grow-only and reused by
// later requests, so steady-state admission allocates nothing and never
// calls hipFree (which would wait for the whole device).  A request is
// admitted when a slot is free and its estimated device footprint fits both
// the server's budget and hipMemGetInfo's free memory; it is launched at once
// and overlaps on the GPU with requests already running on other slots.  A
// request's results depend only on its own instance, seeds and parameters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace plpi;

namespace {

struct Slot {
    plp_ctx* ctx = nullptr;
    plp_engine* eng = nullptr;
    hipEvent_t done = nullptr;
    int64_t req = -1;       // request occupying the slot, -1 if free
    int64_t footprint = 0;  // bytes charged to the budget
    int n_dec = 0;
};

// Device bytes a request needs (tables + per-decoder state + two packings).
int64_t footprint(const plp_instance_desc* inst, int32_t n_dec) {
    int64_t items = 0;
    for (int t = 0; t < inst->n_types; t++) items += inst->types[7 * t + 6];
    items = std::max<int64_t>(items, 1);
    const int64_t per_dec = 4 * 625 + 64 + 2 * items * 7 * 4 + 4 * 4 + 8;
    return 65536 + (int64_t)n_dec * per_dec + items * 8 + (int64_t)inst->n_types * 80;
}

}  // namespace

struct plp_server {
    int device = 0;
    int max_inflight = 1;
    int64_t budget = 0;
    int64_t in_use = 0;
    int64_t next_id = 1;
    std::vector<Slot> slots;
    std::map<int64_t, int> req_slot;
    std::mutex mu;
};

extern "C" int plp_server_create(int32_t device, int32_t max_inflight, int64_t mem_budget_bytes,
                                 plp_server** out) {
    if (!out || max_inflight < 1 || max_inflight > 64 || mem_budget_bytes < 0)
        return fail(PLP_E_ARG, "bad server arguments (1 <= max_inflight <= 64)");
    *out = nullptr;
    plp_server* s = new plp_server();
    s->device = device;
    s->max_inflight = max_inflight;
    s->budget = mem_budget_bytes;
    s->slots.resize((size_t)max_inflight);
    for (Slot& sl : s->slots) {
        int rc = plp_ctx_create(device, &sl.ctx);
        if (rc == PLP_OK) {
            plp_engine_config cfg{};
            rc = plp_engine_create(sl.ctx, &cfg, &sl.eng);
        }
        if (rc == PLP_OK && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
            rc = fail(PLP_E_HIP, "hipEventCreate");
        if (rc != PLP_OK) {
            plp_server_destroy(s);
            return rc;
        }
    }
    *out = s;
    return PLP_OK;
}

extern "C" void plp_server_destroy(plp_server* s) {
    if (!s) return;
    for (Slot& sl : s->slots) {
        if (sl.eng) plp_engine_destroy(sl.eng);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.ctx) plp_ctx_destroy(sl.ctx);
    }
    delete s;
}

extern "C" int plp_server_submit(plp_server* s, const plp_instance_desc* inst, int32_t n_dec,
                                 uint64_t seed0, int32_t n_passes, const plp_params* params,
                                 int64_t* req_id) {
    if (!s || !inst || !params || !req_id || n_dec < 1 || n_passes < 1)
        return fail(PLP_E_ARG, "bad request");
    std::lock_guard<std::mutex> lock(s->mu);
    Slot* slot = nullptr;
    for (Slot& sl : s->slots)
        if (sl.req < 0) {
            slot = &sl;
            break;
        }
    if (!slot) return fail(PLP_E_BUSY, "all slots busy");
    const int64_t fp = footprint(inst, n_dec);
    if (s->budget > 0 && s->in_use + fp > s->budget) return fail(PLP_E_BUSY, "memory budget full");
    if (hipSetDevice(s->device) != hipSuccess) return fail(PLP_E_HIP, "hipSetDevice");
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (int64_t)free_b < fp + (64ll << 20))
        return fail(PLP_E_BUSY, "device memory low");
    std::vector<int32_t> dinst((size_t)n_dec, 0);
    std::vector<uint64_t> seeds((size_t)n_dec);
    for (int d = 0; d < n_dec; d++) seeds[(size_t)d] = seed0 + (uint64_t)d;
    int rc = plp_engine_load(slot->eng, inst, 1, dinst.data(), seeds.data(), n_dec, params);
    if (rc != PLP_OK) return rc;
    rc = plp_engine_run(slot->eng, n_passes);
    if (rc != PLP_OK) return rc;
    if (hipEventRecord(slot->done, slot->ctx->stream) != hipSuccess)
        return fail(PLP_E_HIP, "hipEventRecord");
    const int64_t id = s->next_id++;
    slot->req = id;
    slot->footprint = fp;
    slot->n_dec = n_dec;
    s->in_use += fp;
    s->req_slot[id] = (int)(slot - s->slots.data());
    *req_id = id;
    return PLP_OK;
}

extern "C" int plp_server_poll(plp_server* s, int64_t req_id, int32_t* done, plp_row_result* rows,
                               int32_t* placements, int32_t placements_cap_rows,
                               int32_t* max_items) {
    if (!s || !done) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->req_slot.find(req_id);
    if (it == s->req_slot.end()) return fail(PLP_E_ARG, "unknown request id");
    Slot& sl = s->slots[(size_t)it->second];
    *done = 0;
    hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) return PLP_OK;
    if (q != hipSuccess) return hip_fail(q, "request event");
    int rc = plp_engine_sync(sl.eng);
    if (rc != PLP_OK && rc != PLP_E_EPCAP) return rc;
    int32_t mi = 0;
    plp_engine_max_items(sl.eng, &mi);
    if (max_items) *max_items = mi;
    if (rows) {
        if (placements && placements_cap_rows < sl.n_dec * mi)
            return fail(PLP_E_ARG, "placements buffer too small");
        int frc = plp_engine_fetch(sl.eng, rows, placements);
        if (frc != PLP_OK) return frc;
    }
    s->in_use -= sl.footprint;
    sl.req = -1;
    sl.footprint = 0;
    s->req_slot.erase(it);
    *done = 1;
    return rc;
}

extern "C" int plp_server_query(plp_server* s, int64_t req_id, int32_t* done) {
    if (!s || !done) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->req_slot.find(req_id);
    if (it == s->req_slot.end()) return fail(PLP_E_ARG, "unknown request id");
    const hipError_t q = hipEventQuery(s->slots[(size_t)it->second].done);
    if (q != hipSuccess && q != hipErrorNotReady) return hip_fail(q, "request event");
    *done = q == hipSuccess ? 1 : 0;
    return PLP_OK;
}

extern "C" int plp_server_inflight(plp_server* s, int32_t* n) {
    if (!s || !n) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    *n = (int32_t)s->req_slot.size();
    return PLP_OK;
}
