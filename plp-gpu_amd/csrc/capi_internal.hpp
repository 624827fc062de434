// capi_internal.hpp -- shared pieces of the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "host_core.hpp"
#include "plp_hip.h"

namespace plpi {

// Largest container side of the device-resident engines: 21 bits per packed
// key field, as the reference device form (gpu/plp_gpu.cu:1263-1268).
constexpr int32_t kMaxSide = (1 << 21) - 1;
// Largest side of the plain form's feasibility kernel (x + w stays in int32).
constexpr int32_t kMaxPlainSide = (1 << 30) - 1;

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int bitlen(int v);
// Takes this device's address of the allocation-free entry point's scratch
// array (noalloc.hip); called with the device current.
void noalloc_setup(int device);

// Validates an instance descriptor (sides up to max_side) and returns the
// host model.
int to_instance(const plp_instance_desc* d, plp::Instance* out, int32_t max_side = kMaxSide);

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int reserve(size_t n, const char* what);
    void release();
};

// Grow-only pinned host buffer (hipHostMalloc with `flags`).  dev is the
// address kernels use for it.
struct HostBuf {
    void* p = nullptr;
    void* dev = nullptr;
    size_t cap = 0;
    unsigned flags = hipHostMallocDefault;
    int reserve(size_t n, const char* what);
    void release();
};

}  // namespace plpi

struct plp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 1;
    int max_lds = 65536;
    plpi::DevBuf d_stage, d_out;
    plpi::HostBuf h_stage, h_out;
};
