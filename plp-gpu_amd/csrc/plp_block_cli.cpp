// plp_block_cli.cpp -- drop-in `plp_block` command line for MI355X.
//
// Flags, JSON and exit codes follow the reference block engine's CLI
// (blockproof/plp_block.cpp:281-386):
//   plp_block <instance> --seed N [--time S] [--eps F] [--theta F | --no-theta]
//             [--amax A] [--bmax B] [--kmax K] [--topk N]
//             [--pick first|support|contact] [--single-pass] [--json]
// plus additions that leave every default unchanged:
//   --passes P   exactly P passes (the reference can only stop on --time)
//   --device K   K decoders seeded N..N+K-1 in one launch, printed as the
//                {"K", "n_passes", "time_s", "rows"} envelope of plp_gpu
//   --mu F       FSR friction (default PLP_MU or 0.2, as plp_gpu)
// The single-run JSON goes to stdout and fill_pct to stderr, as the
// reference prints them.  Exit codes: 0 ok, 1 instance / device error, 2 usage.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "host_core.hpp"
#include "plp_hip.h"

namespace {

void usage(const char* prog) {
    std::fprintf(stderr,
                 "usage: %s <instance> --seed N [--time S] [--eps F] [--theta F | --no-theta]\n"
                 "                 [--amax A] [--bmax B] [--kmax K] [--topk N]\n"
                 "                 [--pick first|support|contact]\n"
                 "                 [--single-pass] [--passes P] [--device K] [--mu F] [--json]\n",
                 prog);
}

int die_device(const char* what, int rc) {
    std::fprintf(stderr, "hip %s: %s (%s)\n", what, plp_last_error(), plp_status_string(rc));
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    std::string instance_path;
    long long seed = 0;
    double time_s = 300.0, eps = 0.3, theta = 0.9, mu = 0.2;
    bool no_theta = false, single_pass = false, as_json = false;
    int passes = 0, device = 0;
    plp_block_params bp{};
    bp.amax = 3;
    bp.bmax = 3;
    bp.kmax = 4;
    bp.topk = 8;
    bp.pick = 0;
    if (const char* e = std::getenv("PLP_MU")) mu = std::strtod(e, nullptr);
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--seed") seed = std::atoi(next(a.c_str()).c_str());   // an int, as the reference
        else if (a == "--time") time_s = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--eps") eps = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--theta") theta = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--no-theta") no_theta = true;
        else if (a == "--amax") bp.amax = std::atoi(next(a.c_str()).c_str());
        else if (a == "--bmax") bp.bmax = std::atoi(next(a.c_str()).c_str());
        else if (a == "--kmax") bp.kmax = std::atoi(next(a.c_str()).c_str());
        else if (a == "--topk") bp.topk = std::max(1, std::atoi(next(a.c_str()).c_str()));
        else if (a == "--pick") {
            const std::string v = next(a.c_str());
            if (v == "first") bp.pick = 0;
            else if (v == "support") bp.pick = 1;
            else if (v == "contact") bp.pick = 2;
            else {
                std::fprintf(stderr, "bad --pick: %s (first|support|contact)\n", v.c_str());
                return 2;
            }
        }
        else if (a == "--single-pass") single_pass = true;
        else if (a == "--passes") passes = std::atoi(next(a.c_str()).c_str());
        else if (a == "--device") device = std::atoi(next(a.c_str()).c_str());
        else if (a == "--mu") mu = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--json") as_json = true;
        else if (a == "-h" || a == "--help") {
            usage(argv[0]);
            return 0;
        } else if (!a.empty() && a[0] != '-') {
            if (!instance_path.empty()) {
                usage(argv[0]);
                return 2;
            }
            instance_path = a;
        } else {
            std::fprintf(stderr, "unknown flag: %s\n", a.c_str());
            usage(argv[0]);
            return 2;
        }
    }
    if (instance_path.empty()) {
        usage(argv[0]);
        return 2;
    }
    if (passes < 0 || (passes > 0 && single_pass) || device < 0) {
        std::fprintf(stderr, "--passes P needs P >= 1 and excludes --single-pass; --device K >= 1\n");
        return 2;
    }
    if (device > 0 && !as_json) {
        std::fprintf(stderr, "--device requires --json\n");
        return 2;
    }
    // the reference clamps the block bounds to >= 1 (plp_block.cpp:327)
    bp.amax = std::max(1, bp.amax);
    bp.bmax = std::max(1, bp.bmax);
    bp.kmax = std::max(1, bp.kmax);
    plp::Instance inst;
    try {
        inst = plp::load_instance(instance_path);
    } catch (const plp::InstanceFormatError& e) {
        std::fprintf(stderr, "instance error: %s\n", e.what());
        return 1;
    }
    if (single_pass) passes = 1;
    plp_params p{};
    p.eps = eps;
    p.theta = theta;
    p.mu = mu;
    p.use_theta = no_theta ? 0 : 1;
    std::vector<int32_t> types;
    for (const plp::Item& it : inst.items) {
        const int32_t q[7] = {it.l, it.lf, it.w, it.wf, it.h, it.hf, it.qty};
        types.insert(types.end(), q, q + 7);
    }
    plp_instance_desc desc{inst.L, inst.W, inst.H, (int32_t)inst.items.size(),
                           types.empty() ? nullptr : types.data()};
    const int K = device > 0 ? device : 1;

    plp_ctx* ctx = nullptr;
    int rc = plp_ctx_create(0, &ctx);
    if (rc != PLP_OK) return die_device("init", rc);
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    plp_engine* eng = nullptr;
    plp_engine_config cfg{};
    rc = plp_engine_create(ctx, &cfg, &eng);
    std::vector<int32_t> dinst((size_t)K, 0);
    std::vector<uint64_t> seeds((size_t)K);
    for (int k = 0; k < K; k++) seeds[(size_t)k] = (uint64_t)(seed + k);   // PyRng((uint64_t)seed)
    if (rc == PLP_OK) rc = plp_engine_load_blocks(eng, &desc, 1, dinst.data(), seeds.data(), K, &p, &bp);
    if (rc == PLP_E_RANGE || rc == PLP_E_ARG) {
        std::fprintf(stderr, "%s\n", plp_last_error());
        return 2;
    }
    if (rc != PLP_OK) return die_device("load", rc);
    if (passes > 0) {
        rc = plp_engine_run(eng, passes);
        if (rc == PLP_OK) rc = plp_engine_sync(eng);
    } else {
        do {
            rc = plp_engine_run(eng, 1);
            if (rc == PLP_OK) rc = plp_engine_sync(eng);
        } while (rc == PLP_OK && elapsed() < time_s);
    }
    if (rc != PLP_OK) return die_device("decode", rc);
    int32_t mi = 1;
    plp_engine_max_items(eng, &mi);
    std::vector<plp_row_result> rows((size_t)K);
    std::vector<int32_t> pl((size_t)K * (size_t)mi * 7);
    rc = plp_engine_fetch(eng, rows.data(), pl.data());
    if (rc != PLP_OK) return die_device("fetch", rc);
    const double t = elapsed();
    plp_engine_destroy(eng);
    plp_ctx_destroy(ctx);
    std::vector<plp::RunResult> res;
    for (int k = 0; k < K; k++) {
        const plp_row_result& r = rows[(size_t)k];
        plp::RunResult o;
        o.n_boxes = r.n_boxes;
        o.volume = r.volume;
        o.score = r.score;
        o.fallback = r.fallback != 0;
        o.n_passes = r.n_passes;
        o.time_s = t;
        o.seed = (long long)r.seed;
        const int32_t* q = pl.data() + (size_t)k * (size_t)mi * 7;
        for (int i = 0; i < r.n_boxes; i++) {
            std::array<int, 7> b;
            for (int c = 0; c < 7; c++) b[(size_t)c] = q[7 * i + c];
            o.placements.push_back(b);
        }
        res.push_back(o);
    }
    if (device > 0) {
        std::cout << "{\"K\": " << K << ", \"n_passes\": " << res[0].n_passes
                  << ", \"time_s\": " << plp::py_repr(t) << ", \"rows\": [";
        for (size_t r = 0; r < res.size(); r++) std::cout << (r ? ", " : "") << plp::to_json(res[r]);
        std::cout << "]}\n";
        return 0;
    }
    const plp::RunResult& r = res[0];
    const double fill = 100.0 * (double)r.volume / (double)((long long)inst.L * inst.W * inst.H);
    if (as_json) {
        std::cout << plp::to_json(r) << "\n";
        std::fprintf(stderr, "fill_pct=%.4f\n", fill);
    } else {
        std::printf("n_boxes=%d volume_cm3=%lld fill_pct=%.4f stability_score=%.4f "
                    "fallback=%s n_passes=%d time_s=%.3f\n",
                    r.n_boxes, (long long)r.volume, fill, r.score, r.fallback ? "true" : "false",
                    r.n_passes, r.time_s);
    }
    return 0;
}
