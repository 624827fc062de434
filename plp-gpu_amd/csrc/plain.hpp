// plain.hpp -- host-driven GPU forms (see plain.cpp).
#pragma once

#include <array>
#include <vector>

#include "capi_internal.hpp"

namespace plpi {

// K lockstep rows over one instance; row k uses random.Random(seed0 + k)
// (plp_gpu.cu:350).  RNG and D4' state persist across run_pass calls, as in
// the reference's run loop.
class PlainRunner {
public:
    PlainRunner(plp_ctx* ctx, const plp::Instance& inst, uint64_t seed0, int K,
                const plp_params& p);
    int run_pass();
    int rows() const { return (int)rows_.size(); }
    void result(int k, plp_row_result* out, std::vector<std::array<int32_t, 7>>* rows) const;

private:
    struct Pack {
        long long vol = 0;
        int n = -1;
        double score = 0.0;
        std::vector<std::array<int32_t, 7>> rows;
    };
    struct Row {
        explicit Row(uint64_t s) : rng(s), seed(s) {}
        plp::PyRng rng;
        uint64_t seed = 0;
        std::vector<std::array<int32_t, 3>> cand;
        std::vector<std::array<int32_t, 6>> aabb;
        std::vector<std::array<int32_t, 7>> placed;   // (type, x, y, z, w, d, h)
        Pack all, acc;
        int passes = 0;
        int fail_ti = -1;   // type whose last copy failed since the last placement
        int max_ep = 1;
        uint64_t work = 0;
    };
    plp_ctx* ctx_;
    plp::Instance inst_;
    plp_params p_;
    std::vector<int> seq_;
    std::vector<std::vector<std::array<int, 3>>> orients_;
    bool use_rng_ = false;
    std::vector<Row> rows_;
};

}  // namespace plpi
