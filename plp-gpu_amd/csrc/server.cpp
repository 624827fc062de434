// server.cpp -- continuous-batching request loop (no reference counterpart;
// the reference decodes one instance per process).
//
// A request is one instance decoded by n_decoders seeds.  Submitted requests
// wait in a queue.  Every poll or query first collects the launches that
// have finished, then, while a slot is free, launches every queued request
// that shares the queue head's pass count and parameters (up to
// kBatchDecoders decoders) together as ONE multi-instance engine run on that
// slot: one persistent decode launch in which each request's decoders are a
// range of the work queue.  So the requests that arrived since the last poll
// share a launch, and a lone request starts at the next poll.  Slots have
// their own contexts (non-blocking HIP streams) and engines, so launches on
// different slots overlap on the device.  Admission checks
// the memory of everything admitted and not yet returned against the
// server's budget and hipMemGetInfo's free memory, and a queue bound.
// Engine buffers are grow-only and reused by later batches, so steady-state
// admission allocates nothing and never calls hipFree.  A request's results
// depend only on its own instance, seeds and parameters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace plpi;

namespace {

constexpr int kBatchDecoders = 1 << 16;   // decoders of one coalesced launch, at most
constexpr int kBatchRequests = 512;       // requests of one coalesced launch, at most
constexpr int kMaxQueued = 4096;          // requests waiting for a slot, at most

struct Request {
    int64_t id = 0;
    plp_instance_desc desc{};
    std::vector<int32_t> types;   // owned copy of desc.types
    int32_t n_dec = 0, n_passes = 1, n_items = 1;
    uint64_t seed0 = 0;
    plp_params params{};
    int64_t footprint = 0;
    enum { kQueued, kRunning, kDone } state = kQueued;
    int status = PLP_OK;          // of a finished request
    std::vector<plp_row_result> rows;
    std::vector<int32_t> placements;   // n_dec * n_items * 7
};

struct Slot {
    plp_ctx* ctx = nullptr;
    plp_engine* eng = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Request*> batch;   // requests of the running launch; empty if free
    std::vector<int32_t> dec_off;  // first decoder of each
};

// Device bytes a request needs (tables + per-decoder state).
int64_t footprint(const plp_instance_desc* inst, int32_t n_dec) {
    int64_t items = 0;
    for (int t = 0; t < inst->n_types; t++) items += inst->types[7 * t + 6];
    items = std::max<int64_t>(items, 1);
    const int64_t per_dec = 4 * 625 + 80 + 4 * 4 + 8;
    return 65536 + (int64_t)n_dec * per_dec + items * 24 + (int64_t)inst->n_types * 80;
}

bool same_params(const Request& a, const Request& b) {
    return a.n_passes == b.n_passes && a.params.eps == b.params.eps &&
           a.params.theta == b.params.theta && a.params.mu == b.params.mu &&
           a.params.use_theta == b.params.use_theta && a.params.first_fit == b.params.first_fit &&
           a.params.order == b.params.order;
}

}  // namespace

struct plp_server {
    int device = 0;
    int max_inflight = 1;
    int64_t budget = 0;
    int64_t admitted_bytes = 0;   // footprints of requests admitted and not yet returned
    int64_t next_id = 1;
    int64_t launches = 0, launched_requests = 0, launched_decoders = 0;
    std::vector<Slot> slots;
    std::deque<Request*> queue;
    std::map<int64_t, std::unique_ptr<Request>> reqs;   // admitted, not yet returned
    std::mutex mu;
};

namespace {

// Launch the queued requests that fit one batch on a free slot, while
// there are both.
int dispatch(plp_server* s) {
    for (Slot& sl : s->slots) {
        if (s->queue.empty()) break;
        if (!sl.batch.empty()) continue;
        // the queue head always goes (whatever its size), then every later
        // request with the same passes and parameters that still fits
        Request* const head_req = s->queue.front();
        const Request& head = *head_req;
        s->queue.pop_front();
        std::vector<Request*> batch{head_req};
        int64_t n_dec = head.n_dec;
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            Request* r = *it;
            if (same_params(*r, head) && n_dec + r->n_dec <= kBatchDecoders &&
                (int)batch.size() < kBatchRequests) {
                batch.push_back(r);
                n_dec += r->n_dec;
                it = s->queue.erase(it);
            } else {
                ++it;
            }
        }
        std::vector<plp_instance_desc> insts;
        std::vector<int32_t> dinst;
        std::vector<uint64_t> seeds;
        sl.dec_off.clear();
        for (size_t i = 0; i < batch.size(); i++) {
            Request* r = batch[i];
            insts.push_back(r->desc);
            sl.dec_off.push_back((int32_t)dinst.size());
            for (int k = 0; k < r->n_dec; k++) {
                dinst.push_back((int32_t)i);
                seeds.push_back(r->seed0 + (uint64_t)k);
            }
        }
        int rc = plp_engine_load(sl.eng, insts.data(), (int32_t)insts.size(), dinst.data(),
                                 seeds.data(), (int32_t)seeds.size(), &head.params);
        if (rc == PLP_OK) rc = plp_engine_run(sl.eng, head.n_passes);
        if (rc == PLP_OK && hipEventRecord(sl.done, sl.ctx->stream) != hipSuccess)
            rc = fail(PLP_E_HIP, "hipEventRecord");
        if (rc != PLP_OK) {
            const std::string msg = plp_last_error();
            for (Request* r : batch) {
                r->state = Request::kDone;
                r->status = rc;
            }
            return fail(rc, msg);
        }
        for (Request* r : batch) r->state = Request::kRunning;
        sl.batch = std::move(batch);
        s->launches++;
        s->launched_requests += (int64_t)sl.batch.size();
        s->launched_decoders += n_dec;
    }
    return PLP_OK;
}

// Collect the results of every finished launch into its requests, free
// those slots and launch what is queued.
int progress(plp_server* s) {
    for (Slot& sl : s->slots) {
        if (sl.batch.empty()) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        int rc = q == hipSuccess ? PLP_OK : hip_fail(q, "request event");
        if (rc == PLP_OK) rc = plp_engine_sync(sl.eng);   // capacity escalation, if any
        const plp_row_result* rows = nullptr;
        const int32_t* pl = nullptr;
        int32_t stride = 0;
        if (rc == PLP_OK || rc == PLP_E_EPCAP) {
            const int rr = plp_engine_results(sl.eng, &rows, &pl, &stride);
            if (rr != PLP_OK) rc = rr;
        }
        for (size_t i = 0; i < sl.batch.size(); i++) {
            Request* r = sl.batch[i];
            r->state = Request::kDone;
            r->status = rc;
            if (!rows) continue;
            const int d0 = sl.dec_off[i];
            r->rows.assign(rows + d0, rows + d0 + r->n_dec);
            r->placements.assign((size_t)r->n_dec * r->n_items * 7, 0);
            for (int k = 0; k < r->n_dec; k++)
                std::memcpy(r->placements.data() + (size_t)k * r->n_items * 7,
                            pl + (size_t)(d0 + k) * stride,
                            4 * 7 * (size_t)std::max(0, r->rows[(size_t)k].n_boxes));
        }
        sl.batch.clear();
    }
    return dispatch(s);
}

}  // namespace

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
    plp::Instance model;
    int rc = to_instance(inst, &model);   // same checks as a load, before admission
    if (rc != PLP_OK) return rc;
    if (params->order != 0 && params->order != 1) return fail(PLP_E_ARG, "order must be 0 or 1");
    if (!(params->eps >= 0.0) || params->eps > 1.0) return fail(PLP_E_ARG, "eps must be in [0, 1]");
    std::lock_guard<std::mutex> lock(s->mu);
    if ((int)s->queue.size() >= kMaxQueued) return fail(PLP_E_BUSY, "request queue full");
    const int64_t fp = footprint(inst, n_dec);
    if (s->budget > 0 && s->admitted_bytes + fp > s->budget)
        return fail(PLP_E_BUSY, "memory budget full");
    if (hipSetDevice(s->device) != hipSuccess) return fail(PLP_E_HIP, "hipSetDevice");
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (int64_t)free_b < fp + (64ll << 20))
        return fail(PLP_E_BUSY, "device memory low");
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->types.assign(inst->types, inst->types + 7 * (size_t)inst->n_types);
    r->desc = *inst;
    r->desc.types = r->types.empty() ? nullptr : r->types.data();
    r->n_dec = n_dec;
    r->n_passes = n_passes;
    r->seed0 = seed0;
    r->params = *params;
    r->footprint = fp;
    int64_t items = 0;
    for (const plp::Item& it : model.items) items += it.qty;
    r->n_items = (int32_t)std::max<int64_t>(1, items);
    s->admitted_bytes += fp;
    s->queue.push_back(r.get());
    *req_id = r->id;
    s->reqs[r->id] = std::move(r);
    return PLP_OK;   // launched by the next poll or query, with the others queued by then
}

extern "C" int plp_server_poll(plp_server* s, int64_t req_id, int32_t* done, plp_row_result* rows,
                               int32_t* placements, int32_t placements_cap_rows,
                               int32_t* max_items) {
    if (!s || !done) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    *done = 0;
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return fail(PLP_E_ARG, "unknown request id");
    Request& r = *it->second;
    if (r.state != Request::kDone) {
        const int rc = progress(s);
        if (rc != PLP_OK) return rc;
        if (r.state != Request::kDone) return PLP_OK;
    }
    if (max_items) *max_items = r.n_items;
    if (rows && !r.rows.empty()) {
        if (placements && placements_cap_rows < (int64_t)r.n_dec * r.n_items)
            return fail(PLP_E_ARG, "placements buffer too small");
        std::memcpy(rows, r.rows.data(), sizeof(plp_row_result) * (size_t)r.n_dec);
        if (placements)
            std::memcpy(placements, r.placements.data(), 4 * r.placements.size());
    }
    const int status = r.status;
    s->admitted_bytes -= r.footprint;
    s->reqs.erase(it);
    *done = 1;
    return status;
}

extern "C" int plp_server_query(plp_server* s, int64_t req_id, int32_t* done) {
    if (!s || !done) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return fail(PLP_E_ARG, "unknown request id");
    if (it->second->state != Request::kDone) {
        const int rc = progress(s);
        if (rc != PLP_OK) return rc;
    }
    *done = it->second->state == Request::kDone ? 1 : 0;
    return PLP_OK;
}

extern "C" int plp_server_inflight(plp_server* s, int32_t* n) {
    if (!s || !n) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    *n = (int32_t)s->reqs.size();
    return PLP_OK;
}

extern "C" int plp_server_stats(plp_server* s, int64_t* launches, int64_t* requests,
                                int64_t* decoders) {
    if (!s) return fail(PLP_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lock(s->mu);
    if (launches) *launches = s->launches;
    if (requests) *requests = s->launched_requests;
    if (decoders) *decoders = s->launched_decoders;
    return PLP_OK;
}
