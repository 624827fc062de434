// plp_gpu_cli.cpp -- drop-in `plp_gpu` command line for MI355X.
//
// Flags, JSON and exit codes follow the reference CLIs (gpu/plp_gpu.cpp:80-143
// and gpu/plp_gpu.cu:1575-1696):
//   plp_gpu <instance> [--seed N] [--time S] [--eps F] [--theta F | --no-theta]
//           [--order volume|input] [--first-fit] [--single-pass] [--mu F]
//           [--batch K | --device K] [--json]
//   plp_gpu --selftest | --rngdump N
// plus additions that leave every default unchanged:
//   --passes P   exactly P passes (the reference can only stop on --time)
//   --plain      the one-step-per-launch form for the default single run
//   --serve [--max-inflight N]   JSON-lines request/response server on
//                stdin/stdout (serve_cli.cpp)
// Default single runs use the device-resident decoder; --batch K drives K
// lockstep rows from the host with one batched feasibility launch per step;
// --device K runs K device-resident decoders (no K <= 16 limit).
// Exit codes: 0 ok, 1 instance / device error, 2 usage error.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "host_core.hpp"
#include "plain.hpp"
#include "plp_hip.h"
#include "serve_cli.hpp"

namespace {

void print_hexfloat(double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, sizeof bits);
    std::printf("%016llx\n", (unsigned long long)bits);
}

// RNG + repr blocks of the reference --selftest (plp_gpu.cpp:43-66).
void selftest() {
    for (uint32_t s : {0u, 1u, 7u, 123456789u}) {
        plp::PyRng rng(s);
        std::printf("SEED %u\n", s);
        for (int i = 0; i < 16; i++) print_hexfloat(rng.random());
        for (int i = 0; i < 16; i++) std::printf("%u\n", rng.getrandbits(13));
    }
    const double vals[] = {
        0.0, 1.0, 0.5, 0.1, 2.0 / 3.0, 0.5785123966942148,
        0.6132075471698113, 0.6595744680851063, 1e-05, 1e-04,
        1e15, 1e16, 0.001, 123456789.5, 1.0 / 121.0, 1.0 / 106.0,
    };
    for (double v : vals) std::printf("REPR %s\n", plp::py_repr(v).c_str());
}

int die_device(const char* what, int rc) {
    std::fprintf(stderr, "device %s: %s (%s)\n", what, plp_status_string(rc), plp_last_error());
    return 1;
}

// --rngdump N (plp_gpu.cu:1479-1536): device MT19937 of seeds 0..N-1 against
// the host PyRng, 16 random() + 16 getrandbits(13) each.
int rngdump(int N) {
    if (N < 1) {
        std::fprintf(stderr, "--rngdump N must be >= 1\n");
        return 2;
    }
    plp_ctx* ctx = nullptr;
    int rc = plp_ctx_create(0, &ctx);
    if (rc != PLP_OK) return die_device("init", rc);
    std::vector<uint64_t> seeds((size_t)N);
    for (int k = 0; k < N; k++) seeds[(size_t)k] = (uint64_t)k;
    std::vector<int32_t> ops, args;
    for (int i = 0; i < 16; i++) { ops.push_back(0); args.push_back(0); }
    for (int i = 0; i < 16; i++) { ops.push_back(1); args.push_back(13); }
    std::vector<double> f((size_t)N * 32);
    std::vector<uint32_t> u((size_t)N * 32);
    rc = plp_rng_script(ctx, seeds.data(), N, ops.data(), args.data(), 32, f.data(), u.data());
    if (rc != PLP_OK) return die_device("rngdump", rc);
    bool ok = true;
    for (int k = 0; k < N; k++) {
        plp::PyRng rng((uint64_t)k);
        bool row_ok = true;
        for (int i = 0; i < 16; i++) {
            const double h = rng.random();
            if (f[(size_t)k * 32 + i] != h) {
                row_ok = false;
                std::fprintf(stderr, "row %d random[%d]: host %s dev %s\n", k, i,
                             plp::py_repr(h).c_str(), plp::py_repr(f[(size_t)k * 32 + i]).c_str());
            }
        }
        for (int i = 0; i < 16; i++) {
            const uint32_t h = rng.getrandbits(13);
            if (u[(size_t)k * 32 + 16 + i] != h) {
                row_ok = false;
                std::fprintf(stderr, "row %d getrandbits[%d]: host %u dev %u\n", k, i, h,
                             u[(size_t)k * 32 + 16 + i]);
            }
        }
        std::printf("SEED %u: %s\n", (unsigned)k, row_ok ? "PASS" : "FAIL");
        ok = ok && row_ok;
    }
    plp_ctx_destroy(ctx);
    std::printf("RNGDUMP: %s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}

void usage(const char* prog) {
    std::fprintf(stderr,
                 "usage: %s <instance> [--seed N] [--time S] [--eps F] [--theta F | --no-theta]\n"
                 "                 [--order volume|input] [--first-fit] [--single-pass] "
                 "[--passes P]\n"
                 "                 [--batch K | --device K] [--plain] [--mu F] [--json]\n"
                 "       %s --selftest | --rngdump N | --serve [--max-inflight N] [--mu F]\n",
                 prog, prog);
}

plp::RunResult to_result(const plp_row_result& r, const int32_t* rows, double time_s) {
    plp::RunResult o;
    o.n_boxes = r.n_boxes;
    o.volume = r.volume;
    o.score = r.score;
    o.fallback = r.fallback != 0;
    o.n_passes = r.n_passes;
    o.time_s = time_s;
    o.seed = (long long)r.seed;
    for (int i = 0; i < r.n_boxes; i++) {
        std::array<int, 7> p;
        for (int c = 0; c < 7; c++) p[(size_t)c] = rows[7 * i + c];
        o.placements.push_back(p);
    }
    return o;
}

}  // namespace

// Largest container side of the device-resident engine (2^21 - 1).
constexpr int plpi_max_side = (1 << 21) - 1;

int main(int argc, char** argv) {
    std::string instance_path, order = "volume";
    long long seed = 0;
    double time_s = 300.0, eps = 0.3, theta = 0.9, mu = 0.2;
    bool no_theta = false, first_fit = false, single_pass = false, as_json = false, plain = false;
    int batch = 0, device = 0, passes = 0, max_inflight = 8;
    bool serve = false;
    if (const char* e = std::getenv("PLP_MU")) mu = std::strtod(e, nullptr);
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--selftest") { selftest(); return 0; }
        else if (a == "--rngdump") return rngdump(std::atoi(next(a.c_str()).c_str()));
        else if (a == "--seed") seed = std::atoll(next(a.c_str()).c_str());
        else if (a == "--time") time_s = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--eps") eps = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--mu") mu = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--theta") theta = std::strtod(next(a.c_str()).c_str(), nullptr);
        else if (a == "--no-theta") no_theta = true;
        else if (a == "--order") order = next(a.c_str());
        else if (a == "--first-fit") first_fit = true;
        else if (a == "--single-pass") single_pass = true;
        else if (a == "--passes") passes = std::atoi(next(a.c_str()).c_str());
        else if (a == "--batch") batch = std::atoi(next(a.c_str()).c_str());
        else if (a == "--device") device = std::atoi(next(a.c_str()).c_str());
        else if (a == "--plain") plain = true;
        else if (a == "--json") as_json = true;
        else if (a == "--serve") serve = true;
        else if (a == "--max-inflight") max_inflight = std::atoi(next(a.c_str()).c_str());
        else if (a == "-h" || a == "--help") { usage(argv[0]); return 0; }
        else if (!a.empty() && a[0] != '-') {
            if (!instance_path.empty()) { usage(argv[0]); return 2; }
            instance_path = a;
        } else {
            std::fprintf(stderr, "unknown flag: %s\n", a.c_str());
            usage(argv[0]);
            return 2;
        }
    }
    if (serve) {
        if (!instance_path.empty() || max_inflight < 1 || max_inflight > 64) {
            usage(argv[0]);
            return 2;
        }
        return serve_main(max_inflight, mu);
    }
    if (instance_path.empty()) { usage(argv[0]); return 2; }
    plp::Instance inst;
    try {
        inst = plp::load_instance(instance_path);
    } catch (const plp::InstanceFormatError& e) {
        std::fprintf(stderr, "instance error: %s\n", e.what());
        return 1;
    }
    if (order != "volume" && order != "input") {
        std::fprintf(stderr, "unknown order %s\n", order.c_str());
        return 2;
    }
    if (passes < 0 || (passes > 0 && single_pass)) {
        std::fprintf(stderr, "--passes P needs P >= 1 and excludes --single-pass\n");
        return 2;
    }
    if (batch < 0 || device < 0) {
        std::fprintf(stderr, "--batch/--device K must be >= 1\n");
        return 2;
    }
    if (batch > 0 && device > 0) {
        std::fprintf(stderr, "--device and --batch are mutually exclusive\n");
        return 2;
    }
    if ((batch > 0 || device > 0) && !as_json) {
        std::fprintf(stderr, "%s requires --json\n", batch > 0 ? "--batch" : "--device");
        return 2;
    }
    if (single_pass) passes = 1;
    plp_params p{};
    p.eps = eps;
    p.theta = theta;
    p.mu = mu;
    p.use_theta = no_theta ? 0 : 1;
    p.first_fit = first_fit ? 1 : 0;
    p.order = order == "volume" ? 0 : 1;
    std::vector<int32_t> types;
    for (const plp::Item& it : inst.items) {
        int32_t q[7] = {it.l, it.lf, it.w, it.wf, it.h, it.hf, it.qty};
        types.insert(types.end(), q, q + 7);
    }
    plp_instance_desc desc{inst.L, inst.W, inst.H, (int32_t)inst.items.size(),
                           types.empty() ? nullptr : types.data()};
    const int K = batch > 0 ? batch : (device > 0 ? device : 1);

    plp_ctx* ctx = nullptr;
    int rc = plp_ctx_create(0, &ctx);
    if (rc != PLP_OK) return die_device("init", rc);
    auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    std::vector<plp::RunResult> results;
    // The device-resident engine packs extreme points into keys of 21 bits
    // per side.  The reference's --device exits 2 beyond that
    // (gpu/plp_gpu.cu:1263-1268); its default host run has no such limit
    // (gpu/plp_gpu.cpp:80-143), so a single run of a larger container takes
    // the plain form, whose host decode loop keeps the points unpacked.
    const bool beyond_keys = inst.L > plpi_max_side || inst.W > plpi_max_side || inst.H > plpi_max_side;
    if (device > 0 && beyond_keys) {
        std::fprintf(stderr, "--device: container dims exceed the 21-bit pack (L=%d W=%d H=%d)\n",
                     inst.L, inst.W, inst.H);
        return 2;
    }
    if (batch > 0 || plain || beyond_keys) {
        plpi::PlainRunner runner(ctx, inst, (uint64_t)seed, K, p);
        for (int done = 0;;) {
            rc = runner.run_pass();
            if (rc != PLP_OK) return die_device("pass", rc);
            done++;
            if (passes > 0 ? done >= passes : elapsed() >= time_s) break;
        }
        const double t = elapsed();
        for (int k = 0; k < K; k++) {
            plp_row_result r;
            std::vector<std::array<int32_t, 7>> rows;
            runner.result(k, &r, &rows);
            results.push_back(to_result(r, rows.empty() ? nullptr : rows[0].data(), t));
        }
    } else {
        plp_engine* eng = nullptr;
        plp_engine_config cfg{};
        rc = plp_engine_create(ctx, &cfg, &eng);
        std::vector<int32_t> dinst((size_t)K, 0);
        std::vector<uint64_t> seeds((size_t)K);
        for (int k = 0; k < K; k++) seeds[(size_t)k] = (uint64_t)(seed + k);
        if (rc == PLP_OK) rc = plp_engine_load(eng, &desc, 1, dinst.data(), seeds.data(), K, &p);
        if (rc == PLP_E_RANGE) {
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
        for (int k = 0; k < K; k++)
            results.push_back(to_result(rows[(size_t)k], pl.data() + (size_t)k * mi * 7, t));
        plp_engine_destroy(eng);
    }
    plp_ctx_destroy(ctx);
    if (batch > 0 || device > 0) {
        std::cout << "{\"K\": " << K << ", \"n_passes\": " << results[0].n_passes
                  << ", \"time_s\": " << plp::py_repr(results[0].time_s) << ", \"rows\": [";
        for (size_t r = 0; r < results.size(); r++) {
            if (r) std::cout << ", ";
            std::cout << plp::to_json(results[r]);
        }
        std::cout << "]}\n";
        return 0;
    }
    const plp::RunResult& r = results[0];
    if (as_json) {
        std::cout << plp::to_json(r) << "\n";
    } else {
        std::printf("n_boxes=%d volume_cm3=%lld stability_score=%.4f fallback=%s "
                    "n_passes=%d time_s=%.3f\n",
                    r.n_boxes, (long long)r.volume, r.score, r.fallback ? "true" : "false",
                    r.n_passes, r.time_s);
    }
    return 0;
}
