// noalloc.hip -- plp_feas_check_all_noalloc: the plain form's feasibility
// step (plp_feas_check_all) for callers that cannot allocate memory.
//
// plp_feas_check_all stages its inputs, row tables and output bits in
// grow-only device and pinned host buffers that the context allocates on
// first use.  This variant stages them in one file-scope device array,
// `scratch`, and copies straight from and to the caller's buffers, so a
// call makes no allocation of any kind.  Each device holds its own copy of
// the array for the life of the process; its address on a device is taken
// once (hipGetSymbolAddress, with that device current) when the first
// context on the device is created, and only that address is used by host
// code.  Calls on one device share the array: two calls that overlap in
// time on one device race (the context is single-threaded anyway).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "kernels.hpp"

namespace {

constexpr size_t kScratchBytes = 16u << 20;   // the most one call can stage
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices];        // device address of scratch, per device
hipError_t g_err[kMaxDevices];   // result of taking it

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

namespace plpi {

// Called by plp_ctx_create with `device` current.
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    void* p = nullptr;
    g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_err[device] == hipSuccess) g_ptr[device] = p;
    else (void)hipGetLastError();
}

}  // namespace plpi

using namespace plpi;

extern "C" int plp_feas_check_all_noalloc(plp_ctx* c, const int32_t* cand_xyz, int32_t n_cand,
                                          const int32_t* ors_wdh, int32_t n_ors,
                                          const int32_t* aabb6, int32_t n_aabb, int32_t L,
                                          int32_t W, int32_t H, uint8_t* bits_out) {
    if (!c) return fail(PLP_E_ARG, "context is NULL");
    if (n_cand < 0 || n_aabb < 0 || n_ors < 0 || n_ors > 6)
        return fail(PLP_E_ARG, "counts must be >= 0 and n_ors <= 6");
    if ((n_cand > 0 && !cand_xyz) || (n_aabb > 0 && !aabb6) || (n_ors > 0 && !ors_wdh) ||
        ((int64_t)n_cand * n_ors > 0 && !bits_out))
        return fail(PLP_E_ARG, "NULL argument");
    if (L <= 0 || W <= 0 || H <= 0) return fail(PLP_E_ARG, "container dimensions must be positive");
    if (L > kMaxPlainSide || W > kMaxPlainSide || H > kMaxPlainSide)
        return fail(PLP_E_RANGE, "container side above 2^30 - 1 (coordinate sums leave int32)");
    const int64_t n_pairs = (int64_t)n_cand * n_ors;
    if (n_pairs == 0) return PLP_OK;
    const int bs = plpk::feas_block_size();
    const int64_t n_blocks = (n_pairs + bs - 1) / bs;
    // Scratch layout: cand | aabb | ors | per-row table (one row) | block_row | bits.
    const size_t o_cand = 0;
    const size_t o_box = align256(o_cand + 12 * (size_t)n_cand);
    const size_t o_ors = align256(o_box + 24 * (size_t)n_aabb);
    const size_t o_row = align256(o_ors + 12 * (size_t)n_ors);
    const size_t o_brow = align256(o_row + 64);
    const size_t o_bits = align256(o_brow + 4 * (size_t)n_blocks);
    const size_t total = o_bits + (size_t)n_pairs;
    if (total > kScratchBytes)
        return fail(PLP_E_ARG, "inputs and outputs need " + std::to_string(total) +
                                   " bytes, more than the " + std::to_string(kScratchBytes) +
                                   "-byte scratch of the allocation-free entry point");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = g_ptr[c->device];
        if (!base)
            return fail(PLP_E_HIP, std::string("scratch address of this device: ") +
                                       hipGetErrorString(c->device >= 0 && c->device < kMaxDevices
                                                             ? g_err[c->device]
                                                             : hipErrorInvalidDevice));
    }
    char* ds = (char*)base;
    // One row: its first block, first candidate, first box, counts, and
    // output offset (int64 at byte 24), as feas_kernel reads them.
    int32_t row[16] = {0};
    row[3] = n_cand;
    row[4] = n_aabb;
    const int64_t bits_off = 0;
    std::memcpy(&row[6], &bits_off, sizeof bits_off);
    struct Up {
        size_t off;
        const void* src;
        size_t n;
    } ups[] = {{o_cand, cand_xyz, 12 * (size_t)n_cand},
               {o_box, aabb6, 24 * (size_t)n_aabb},
               {o_ors, ors_wdh, 12 * (size_t)n_ors},
               {o_row, row, sizeof row}};
    for (const Up& u : ups) {
        if (u.n == 0) continue;
        e = hipMemcpyAsync(ds + u.off, u.src, u.n, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(e, "H2D to scratch");
    }
    // every block belongs to row 0: a fill kernel, not a memset, so that the
    // sequence stays valid inside a captured graph
    e = plpk::launch_fill_u32((uint32_t*)(ds + o_brow), 0u, (int)n_blocks, c->stream);
    if (e != hipSuccess) return hip_fail(e, "block table fill");
    plpk::FeasRowArgs r;
    r.cand = (const int32_t*)(ds + o_cand);
    r.aabb = (const int32_t*)(ds + o_box);
    r.ors = (const int32_t*)(ds + o_ors);
    r.block_row = (const int32_t*)(ds + o_brow);
    r.row_block0 = (const int32_t*)(ds + o_row);
    r.cand_off = (const int32_t*)(ds + o_row + 4);
    r.box_off = (const int32_t*)(ds + o_row + 8);
    r.n_cand = (const int32_t*)(ds + o_row + 12);
    r.n_box = (const int32_t*)(ds + o_row + 16);
    r.bits_off = (const int64_t*)(ds + o_row + 24);
    r.bits = (uint8_t*)(ds + o_bits);
    r.n_ors = n_ors;
    r.L = L;
    r.W = W;
    r.H = H;
    e = plpk::launch_feas(r, (int)n_blocks, c->stream);
    if (e != hipSuccess) return hip_fail(e, "feas_kernel launch");
    e = hipMemcpyAsync(bits_out, ds + o_bits, (size_t)n_pairs, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(e, "D2H bits");
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "feas sync");
    return PLP_OK;
}
