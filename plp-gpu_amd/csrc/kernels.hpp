// kernels.hpp -- launch interface between the host runtime and kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.hpp"

namespace plpk {

// Plain / batched feasibility: K rows sharing one orientation table.  Block
// b works on row block_row[b], pairs [(b - row_block0[row]) * 256, +256).
struct FeasRowArgs {
    const int32_t* cand;       // all rows' candidates, 3 ints each
    const int32_t* aabb;       // all rows' boxes, 6 ints each
    const int32_t* ors;        // n_ors * 3
    const int32_t* block_row;  // per block
    const int32_t* row_block0; // per row: first block
    const int32_t* cand_off;   // per row: first candidate
    const int32_t* box_off;    // per row: first box
    const int32_t* n_cand;
    const int32_t* n_box;
    const int64_t* bits_off;   // per row: first output byte
    uint8_t* bits;
    int32_t n_ors, L, W, H;
};

struct FsrArgs {
    const int32_t* rows7;      // placement rows (type, x, y, z, w, d, h)
    const int32_t* pack_off;   // n_packs + 1 row offsets
    double mu;
    double* fsr_out;           // per row, may be null
    double* score_out;         // per packing
    int32_t box_cap;
};

struct RngArgs {
    const uint64_t* seeds;
    const int32_t* ops;
    const int32_t* args;
    int32_t n_ops;
    double* out_f64;
    uint32_t* out_u32;
};

hipError_t launch_decode(const DecodeArgs& a, int waves, int grid, int lds_bytes,
                         hipStream_t st);
hipError_t launch_block(const BlockArgs& b, int waves, int grid, int lds_bytes, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int n, hipStream_t st);
hipError_t launch_feas(const FeasRowArgs& r, int n_blocks, hipStream_t st);
hipError_t launch_fsr(const FsrArgs& f, int n_packs, int lds_bytes, hipStream_t st);
hipError_t launch_rng_script(const RngArgs& g, int n_seeds, hipStream_t st);
// Every decoder's seeded MT19937 state (kMtWords words each) into out.
hipError_t launch_mt_seed(const uint64_t* seeds, uint32_t* out, int n, hipStream_t st);
hipError_t decode_occupancy(int waves, int mode, int verdicts, int lds_bytes, int* blocks_per_cu);
// the kernel's own __shared__ bytes
hipError_t decode_static_lds(int waves, int mode, int verdicts, int* bytes);
hipError_t block_occupancy(int waves, int mode, int lds_bytes, int* blocks_per_cu);
hipError_t block_static_lds(int waves, int mode, int* bytes);
hipError_t fsr_prepare(int lds_bytes);
// Occupancy and static LDS of a kernel, memoised (kernels.hip).
hipError_t cached_occupancy(const void* kernel, int block_threads, int lds_bytes,
                            int* blocks_per_cu);
hipError_t cached_static_lds(const void* kernel, int* bytes);
int feas_block_size();

}  // namespace plpk
