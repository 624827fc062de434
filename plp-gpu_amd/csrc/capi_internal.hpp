// capi_internal.hpp -- shared pieces of the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "host_core.hpp"
#include "plp_hip.h"

namespace plpi {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int bitlen(int v);
int to_instance(const plp_instance_desc* d, plp::Instance* out);

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int reserve(size_t n, const char* what);
    void release();
};

// Grow-only pinned host buffer (hipHostMalloc).
struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
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
