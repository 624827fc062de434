// capi.cpp -- the C ABI (include/plp_hip.h) over the gfx950 kernels.
//
// Host runtime only: argument checks, table building, one pinned staging
// buffer per direction, launches on the context's stream, HIP events for
// device time.  Nothing here computes a decode result on the CPU.
#include "plp_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "host_core.hpp"
#include "kernels.hpp"

using plpk::DecState;
using plpk::InstHdr;

namespace plpi {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? PLP_E_NOMEM : PLP_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

int DevBuf::reserve(size_t n, const char* what) {
    if (n <= cap) return PLP_OK;
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    size_t want = n < 4096 ? 4096 : n + n / 4;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_fail(e, what);
    }
    cap = want;
    return PLP_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

int HostBuf::reserve(size_t n, const char* what) {
    if (n <= cap) return PLP_OK;
    if (p) {
        (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    size_t want = n < 4096 ? 4096 : n + n / 4;
    hipError_t e = hipHostMalloc(&p, want, flags);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_fail(e, what);
    }
    e = hipHostGetDevicePointer(&dev, p, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(p);
        p = dev = nullptr;
        return hip_fail(e, what);
    }
    cap = want;
    return PLP_OK;
}

void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    cap = 0;
}

int bitlen(int v) {
    int b = 0;
    while (v > 0) {
        b++;
        v >>= 1;
    }
    return b;
}

// Validates an instance descriptor and returns the host model.
int to_instance(const plp_instance_desc* d, plp::Instance* out, int32_t max_side) {
    if (!d) return fail(PLP_E_ARG, "instance descriptor is NULL");
    if (d->L <= 0 || d->W <= 0 || d->H <= 0)
        return fail(PLP_E_ARG, "container dimensions must be positive");
    if (d->n_types < 0 || (d->n_types > 0 && !d->types))
        return fail(PLP_E_ARG, "bad item type table");
    if (d->L > max_side || d->W > max_side || d->H > max_side)
        return fail(PLP_E_RANGE, "container " + std::to_string(d->L) + "x" + std::to_string(d->W) +
                                     "x" + std::to_string(d->H) + " has a side above " +
                                     std::to_string(max_side) +
                                     (max_side == kMaxSide ? " (21-bit packed extreme-point keys)"
                                                           : ""));
    out->L = d->L;
    out->W = d->W;
    out->H = d->H;
    out->items.clear();
    for (int t = 0; t < d->n_types; t++) {
        const int32_t* q = d->types + 7 * t;
        if (q[0] <= 0 || q[2] <= 0 || q[4] <= 0)
            return fail(PLP_E_ARG, "item dimensions must be positive");
        if ((q[1] != 0 && q[1] != 1) || (q[3] != 0 && q[3] != 1) || (q[5] != 0 && q[5] != 1))
            return fail(PLP_E_ARG, "orientation flags must be 0/1");
        if (q[6] < 0) return fail(PLP_E_ARG, "item qty must be >= 0");
        out->items.push_back({q[0], q[2], q[4], q[1] == 1, q[3] == 1, q[5] == 1, q[6]});
    }
    return PLP_OK;
}

}  // namespace plpi

using namespace plpi;

// ------------------------------------------------------------------ misc --

extern "C" const char* plp_status_string(int s) {
    switch (s) {
        case PLP_OK: return "ok";
        case PLP_E_ARG: return "invalid argument";
        case PLP_E_HIP: return "HIP runtime error";
        case PLP_E_NOMEM: return "device memory exhausted";
        case PLP_E_EPCAP: return "extreme-point capacity exceeded";
        case PLP_E_RANGE: return "container exceeds key range";
        case PLP_E_FORMAT: return "instance format error";
        case PLP_E_BUSY: return "no free slot";
        default: return "unknown status";
    }
}

extern "C" const char* plp_last_error(void) { return g_err.c_str(); }

extern "C" int plp_device_count(int32_t* count) {
    if (!count) return fail(PLP_E_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return PLP_OK;
}

extern "C" int plp_parse_instance(const char* text, int32_t* L, int32_t* W, int32_t* H,
                                  int32_t* n_types, int32_t* types_out, int32_t types_cap,
                                  char* err, int32_t errlen) {
    if (!text || !L || !W || !H || !n_types) return fail(PLP_E_ARG, "NULL argument");
    std::vector<plp::Instance> v;
    try {
        v = plp::parse(text);
    } catch (const plp::InstanceFormatError& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return fail(PLP_E_FORMAT, e.what());
    }
    if (v.size() != 1) {
        std::string m = "expected exactly 1 instance, got " + std::to_string(v.size());
        if (err && errlen > 0) {
            std::strncpy(err, m.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return fail(PLP_E_FORMAT, m);
    }
    const plp::Instance& in = v[0];
    *L = in.L;
    *W = in.W;
    *H = in.H;
    *n_types = (int32_t)in.items.size();
    if ((int64_t)in.items.size() > types_cap || (!types_out && !in.items.empty()))
        return fail(PLP_E_ARG, "types_out too small");
    for (size_t t = 0; t < in.items.size(); t++) {
        const plp::Item& it = in.items[t];
        int32_t* q = types_out + 7 * t;
        q[0] = it.l; q[1] = it.lf; q[2] = it.w; q[3] = it.wf; q[4] = it.h; q[5] = it.hf;
        q[6] = it.qty;
    }
    return PLP_OK;
}

// --------------------------------------------------------------- context --

extern "C" int plp_ctx_create(int32_t device, plp_ctx** out) {
    if (!out) return fail(PLP_E_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PLP_E_HIP, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(PLP_E_ARG, "device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    noalloc_setup(device);
    plp_ctx* c = new plp_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    (void)hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    (void)hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (c->n_cu <= 0) c->n_cu = 1;
    if (c->max_lds <= 0) c->max_lds = 65536;
    *out = c;
    return PLP_OK;
}

extern "C" void plp_ctx_destroy(plp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->d_stage.release();
    c->d_out.release();
    c->h_stage.release();
    c->h_out.release();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

// ----------------------------------------------------------- feasibility --

namespace {

size_t align8(size_t v) { return (v + 7) & ~(size_t)7; }

}  // namespace

extern "C" int plp_feas_check_batch(plp_ctx* c, int32_t K, const int32_t* const* cand_xyz,
                                    const int32_t* n_cand, const int32_t* ors_wdh,
                                    int32_t n_ors, const int32_t* const* aabb6,
                                    const int32_t* n_aabb, int32_t L, int32_t W, int32_t H,
                                    uint8_t* const* bits_out) {
    if (!c) return fail(PLP_E_ARG, "context is NULL");
    if (K < 0 || n_ors < 0 || n_ors > 6) return fail(PLP_E_ARG, "K >= 0 and 0 <= n_ors <= 6 required");
    if (K == 0) return PLP_OK;
    if (!n_cand || !n_aabb || !cand_xyz || !aabb6 || !bits_out || (n_ors > 0 && !ors_wdh))
        return fail(PLP_E_ARG, "NULL argument");
    if (L <= 0 || W <= 0 || H <= 0) return fail(PLP_E_ARG, "container dimensions must be positive");
    if (L > kMaxPlainSide || W > kMaxPlainSide || H > kMaxPlainSide)
        return fail(PLP_E_RANGE, "container side above 2^30 - 1 (coordinate sums leave int32)");
    const int bs = plpk::feas_block_size();
    int64_t tot_cand = 0, tot_box = 0, tot_bits = 0, tot_blocks = 0;
    for (int k = 0; k < K; k++) {
        if (n_cand[k] < 0 || n_aabb[k] < 0) return fail(PLP_E_ARG, "negative count");
        if ((n_cand[k] > 0 && !cand_xyz[k]) || (n_aabb[k] > 0 && !aabb6[k]) ||
            (n_cand[k] > 0 && n_ors > 0 && !bits_out[k]))
            return fail(PLP_E_ARG, "NULL row buffer");
        tot_cand += n_cand[k];
        tot_box += n_aabb[k];
        const int64_t np = (int64_t)n_cand[k] * n_ors;
        tot_bits += np;
        tot_blocks += (np + bs - 1) / bs;
    }
    if (tot_blocks == 0) return PLP_OK;
    if (tot_blocks > (1ll << 30)) return fail(PLP_E_ARG, "too many pairs");
    // Staging layout (one H2D): cand | aabb | ors | block_row | row_block0 |
    // cand_off | box_off | n_cand | n_box | bits_off(int64).
    size_t o_cand = 0;
    size_t o_box = align8(o_cand + 12 * (size_t)tot_cand);
    size_t o_ors = align8(o_box + 24 * (size_t)tot_box);
    size_t o_brow = align8(o_ors + 12 * (size_t)n_ors);
    size_t o_rb0 = align8(o_brow + 4 * (size_t)tot_blocks);
    size_t o_coff = align8(o_rb0 + 4 * (size_t)K);
    size_t o_boff = align8(o_coff + 4 * (size_t)K);
    size_t o_nc = align8(o_boff + 4 * (size_t)K);
    size_t o_nb = align8(o_nc + 4 * (size_t)K);
    size_t o_bits = align8(o_nb + 4 * (size_t)K);
    size_t total = align8(o_bits + 8 * (size_t)K);
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc;
    if ((rc = c->h_stage.reserve(total, "pinned staging")) != PLP_OK) return rc;
    if ((rc = c->d_stage.reserve(total, "device staging")) != PLP_OK) return rc;
    if ((rc = c->h_out.reserve((size_t)tot_bits, "pinned bits")) != PLP_OK) return rc;
    if ((rc = c->d_out.reserve((size_t)tot_bits, "device bits")) != PLP_OK) return rc;
    char* hs = (char*)c->h_stage.p;
    int32_t* brow = (int32_t*)(hs + o_brow);
    int32_t* rb0 = (int32_t*)(hs + o_rb0);
    int32_t* coff = (int32_t*)(hs + o_coff);
    int32_t* boff = (int32_t*)(hs + o_boff);
    int32_t* nc = (int32_t*)(hs + o_nc);
    int32_t* nb = (int32_t*)(hs + o_nb);
    int64_t* bo = (int64_t*)(hs + o_bits);
    int64_t ccur = 0, bcur = 0, bitcur = 0, blk = 0;
    for (int k = 0; k < K; k++) {
        if (n_cand[k] > 0) std::memcpy(hs + o_cand + 12 * ccur, cand_xyz[k], 12 * (size_t)n_cand[k]);
        if (n_aabb[k] > 0) std::memcpy(hs + o_box + 24 * bcur, aabb6[k], 24 * (size_t)n_aabb[k]);
        coff[k] = (int32_t)ccur;
        boff[k] = (int32_t)bcur;
        nc[k] = n_cand[k];
        nb[k] = n_aabb[k];
        bo[k] = bitcur;
        rb0[k] = (int32_t)blk;
        const int64_t np = (int64_t)n_cand[k] * n_ors;
        const int64_t nblk = (np + bs - 1) / bs;
        for (int64_t b = 0; b < nblk; b++) brow[blk + b] = k;
        blk += nblk;
        ccur += n_cand[k];
        bcur += n_aabb[k];
        bitcur += np;
    }
    if (n_ors > 0) std::memcpy(hs + o_ors, ors_wdh, 12 * (size_t)n_ors);
    e = hipMemcpyAsync(c->d_stage.p, hs, total, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(e, "H2D staging");
    char* ds = (char*)c->d_stage.p;
    plpk::FeasRowArgs r;
    r.cand = (const int32_t*)(ds + o_cand);
    r.aabb = (const int32_t*)(ds + o_box);
    r.ors = (const int32_t*)(ds + o_ors);
    r.block_row = (const int32_t*)(ds + o_brow);
    r.row_block0 = (const int32_t*)(ds + o_rb0);
    r.cand_off = (const int32_t*)(ds + o_coff);
    r.box_off = (const int32_t*)(ds + o_boff);
    r.n_cand = (const int32_t*)(ds + o_nc);
    r.n_box = (const int32_t*)(ds + o_nb);
    r.bits_off = (const int64_t*)(ds + o_bits);
    r.bits = (uint8_t*)c->d_out.p;
    r.n_ors = n_ors;
    r.L = L;
    r.W = W;
    r.H = H;
    e = plpk::launch_feas(r, (int)tot_blocks, c->stream);
    if (e != hipSuccess) return hip_fail(e, "feas_kernel launch");
    e = hipMemcpyAsync(c->h_out.p, c->d_out.p, (size_t)tot_bits, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H bits");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "feas sync");
    const uint8_t* hb = (const uint8_t*)c->h_out.p;
    for (int k = 0; k < K; k++) {
        const int64_t np = (int64_t)n_cand[k] * n_ors;
        if (np > 0) std::memcpy(bits_out[k], hb + bo[k], (size_t)np);
    }
    return PLP_OK;
}

extern "C" int plp_feas_check_all(plp_ctx* c, const int32_t* cand_xyz, int32_t n_cand,
                                  const int32_t* ors_wdh, int32_t n_ors, const int32_t* aabb6,
                                  int32_t n_aabb, int32_t L, int32_t W, int32_t H,
                                  uint8_t* bits_out) {
    const int32_t* cp[1] = {cand_xyz};
    const int32_t* bp[1] = {aabb6};
    uint8_t* op[1] = {bits_out};
    return plp_feas_check_batch(c, 1, cp, &n_cand, ors_wdh, n_ors, bp, &n_aabb, L, W, H, op);
}

// ------------------------------------------------------------------- FSR --

extern "C" int plp_fsr_eval(plp_ctx* c, const int32_t* rows7, const int32_t* pack_off,
                            int32_t n_packs, double mu, double* fsr_out, double* score_out) {
    if (!c) return fail(PLP_E_ARG, "context is NULL");
    if (n_packs < 0 || !pack_off || (n_packs > 0 && !score_out))
        return fail(PLP_E_ARG, "bad packing list");
    if (n_packs == 0) return PLP_OK;
    if (pack_off[0] != 0) return fail(PLP_E_ARG, "pack_off[0] must be 0");
    int32_t box_cap = 1;
    for (int p = 0; p < n_packs; p++) {
        const int32_t n = pack_off[p + 1] - pack_off[p];
        if (n < 0) return fail(PLP_E_ARG, "pack_off must be non-decreasing");
        box_cap = std::max(box_cap, n);
    }
    const int64_t n_rows = pack_off[n_packs];
    if (n_rows > 0 && !rows7) return fail(PLP_E_ARG, "rows7 is NULL");
    // the layout fsr_kernel uses: 64-bit contact totals
    const plpk::LdsLayout lay = plpk::lds_layout(0, box_cap, 0, 1, 0, plpk::kBig, 0);
    if (lay.total > c->max_lds)
        return fail(PLP_E_ARG, "packing too large for one block's LDS (" + std::to_string(box_cap) + " boxes)");
    size_t o_rows = 0;
    size_t o_off = align8(28 * (size_t)n_rows);
    size_t total = align8(o_off + 4 * (size_t)(n_packs + 1));
    size_t o_fsr = 0;
    size_t o_score = align8(8 * (size_t)n_rows);
    size_t out_total = o_score + 8 * (size_t)n_packs;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc;
    if ((rc = c->h_stage.reserve(total, "pinned staging")) != PLP_OK) return rc;
    if ((rc = c->d_stage.reserve(total, "device staging")) != PLP_OK) return rc;
    if ((rc = c->h_out.reserve(out_total, "pinned out")) != PLP_OK) return rc;
    if ((rc = c->d_out.reserve(out_total, "device out")) != PLP_OK) return rc;
    char* hs = (char*)c->h_stage.p;
    if (n_rows > 0) std::memcpy(hs + o_rows, rows7, 28 * (size_t)n_rows);
    std::memcpy(hs + o_off, pack_off, 4 * (size_t)(n_packs + 1));
    e = hipMemcpyAsync(c->d_stage.p, hs, total, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(e, "H2D rows");
    char* ds = (char*)c->d_stage.p;
    char* dout = (char*)c->d_out.p;
    plpk::FsrArgs f;
    f.rows7 = (const int32_t*)(ds + o_rows);
    f.pack_off = (const int32_t*)(ds + o_off);
    f.mu = mu;
    f.fsr_out = (double*)(dout + o_fsr);
    f.score_out = (double*)(dout + o_score);
    f.box_cap = box_cap;
    if ((e = plpk::fsr_prepare(lay.total)) != hipSuccess) return hip_fail(e, "fsr_kernel attribute");
    e = plpk::launch_fsr(f, n_packs, lay.total, c->stream);
    if (e != hipSuccess) return hip_fail(e, "fsr_kernel launch");
    e = hipMemcpyAsync(c->h_out.p, c->d_out.p, out_total, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H fsr");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "fsr sync");
    const char* ho = (const char*)c->h_out.p;
    if (fsr_out && n_rows > 0) std::memcpy(fsr_out, ho + o_fsr, 8 * (size_t)n_rows);
    std::memcpy(score_out, ho + o_score, 8 * (size_t)n_packs);
    return PLP_OK;
}

// ------------------------------------------------------------------- RNG --

extern "C" int plp_rng_script(plp_ctx* c, const uint64_t* seeds, int32_t n_seeds,
                              const int32_t* ops, const int32_t* args, int32_t n_ops,
                              double* out_f64, uint32_t* out_u32) {
    if (!c) return fail(PLP_E_ARG, "context is NULL");
    if (n_seeds < 0 || n_ops < 0) return fail(PLP_E_ARG, "negative count");
    if (n_seeds == 0 || n_ops == 0) return PLP_OK;
    if (!seeds || !ops || !args || !out_f64 || !out_u32) return fail(PLP_E_ARG, "NULL argument");
    for (int i = 0; i < n_ops; i++) {
        if (ops[i] < 0 || ops[i] > 2) return fail(PLP_E_ARG, "op must be 0, 1 or 2");
        if (ops[i] == 1 && (args[i] < 1 || args[i] > 32))
            return fail(PLP_E_ARG, "getrandbits needs 1 <= k <= 32");
        if (ops[i] == 2 && args[i] < 1) return fail(PLP_E_ARG, "randrange needs n >= 1");
    }
    size_t o_seed = 0;
    size_t o_ops = align8(8 * (size_t)n_seeds);
    size_t o_args = align8(o_ops + 4 * (size_t)n_ops);
    size_t total = align8(o_args + 4 * (size_t)n_ops);
    size_t o_f = 0;
    size_t o_u = align8(8 * (size_t)n_seeds * n_ops);
    size_t out_total = o_u + 4 * (size_t)n_seeds * n_ops;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc;
    if ((rc = c->h_stage.reserve(total, "pinned staging")) != PLP_OK) return rc;
    if ((rc = c->d_stage.reserve(total, "device staging")) != PLP_OK) return rc;
    if ((rc = c->h_out.reserve(out_total, "pinned out")) != PLP_OK) return rc;
    if ((rc = c->d_out.reserve(out_total, "device out")) != PLP_OK) return rc;
    char* hs = (char*)c->h_stage.p;
    std::memcpy(hs + o_seed, seeds, 8 * (size_t)n_seeds);
    std::memcpy(hs + o_ops, ops, 4 * (size_t)n_ops);
    std::memcpy(hs + o_args, args, 4 * (size_t)n_ops);
    e = hipMemcpyAsync(c->d_stage.p, hs, total, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(e, "H2D rng script");
    char* ds = (char*)c->d_stage.p;
    char* dout = (char*)c->d_out.p;
    plpk::RngArgs g;
    g.seeds = (const uint64_t*)(ds + o_seed);
    g.ops = (const int32_t*)(ds + o_ops);
    g.args = (const int32_t*)(ds + o_args);
    g.n_ops = n_ops;
    g.out_f64 = (double*)(dout + o_f);
    g.out_u32 = (uint32_t*)(dout + o_u);
    e = plpk::launch_rng_script(g, n_seeds, c->stream);
    if (e != hipSuccess) return hip_fail(e, "rng_script launch");
    e = hipMemcpyAsync(c->h_out.p, c->d_out.p, out_total, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H rng");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "rng sync");
    const char* ho = (const char*)c->h_out.p;
    std::memcpy(out_f64, ho + o_f, 8 * (size_t)n_seeds * n_ops);
    std::memcpy(out_u32, ho + o_u, 4 * (size_t)n_seeds * n_ops);
    return PLP_OK;
}
