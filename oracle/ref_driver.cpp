// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY (never shipped, never linked
// into the product).
//
// A small harness around the reference's own decoder header
// (gpu/plp_common.hpp, included from the reference checkout at build time by
// oracle/Makefile; nothing from it is copied here).  It exists because the
// reference CLI (gpu/plp_gpu.cpp:80-143) can only stop multi-pass runs on a
// wall-clock budget (plp_common.hpp:490-491) and never prints per-box FSR.
// This driver calls the reference functions directly:
//
//   run    fixed pass count: plp::decode_pass (plp_common.hpp:404) P times with
//          one plp::PyRng (plp_common.hpp:164), plp::stability_score
//          (plp_common.hpp:393) per pass, then the D4' rule of plp::run
//          (plp_common.hpp:493-511), restated here for a fixed pass count.
//          Prints plp::to_json (plp_common.hpp:592) with time_s forced to 0.0
//          plus per-box FSR bits of the selected packing (plp::compute_fsr,
//          plp_common.hpp:356) and the brute-force work count of every pass.
//   rows   K decoders seeded seed..seed+K-1 (plp_gpu.cu:350), one JSON line
//          each, as the reference's --batch/--device envelopes do.
//   bench  the same rows timed over T std::threads (CPU baseline); several
//          instance files may be given, each decoded by K seeds.
//
// Build: see oracle/Makefile (g++ -O3 -std=c++17 -ffp-contract=off, the flags
// of gpu/Makefile:18).
#include "plp_common.hpp"

#include <atomic>
#include <thread>

namespace {

// Feasibility seam (plp_common.hpp:326): evaluates plp::check_feasible_one for
// every (candidate, orientation) pair and counts the brute-force pair x box
// work of the step.
struct CountingBackend : plp::FeasibilityBackend {
    unsigned long long pair_box = 0;   // sum over steps of n_cand*n_ors*n_fast
    unsigned long long pairs = 0;      // sum over steps of n_cand*n_ors
    unsigned long long steps = 0;
    size_t max_cand = 0;
    std::vector<uint8_t> check_all(const std::vector<plp::Pt>& cand,
                                   const std::vector<plp::Pt>& ors,
                                   const std::vector<plp::FastBox>& fast,
                                   int L, int W, int H) override {
        std::vector<uint8_t> bits(cand.size() * ors.size(), 0);
        for (size_t ci = 0; ci < cand.size(); ci++)
            for (size_t oi = 0; oi < ors.size(); oi++)
                bits[ci * ors.size() + oi] = plp::check_feasible_one(
                    cand[ci][0], cand[ci][1], cand[ci][2], ors[oi], fast, L, W, H) ? 1 : 0;
        pair_box += (unsigned long long)cand.size() * ors.size() * fast.size();
        pairs += (unsigned long long)cand.size() * ors.size();
        steps++;
        if (cand.size() > max_cand) max_cand = cand.size();
        return bits;
    }
};

struct Opts {
    std::string inst_path, order = "volume";
    std::vector<std::string> more_paths;   // bench: further instances
    long long seed = 0;
    int passes = 1, K = 1, threads = 1;
    double eps = plp::EPS_QUALITY, theta = plp::THETA;
    bool no_theta = false, first_fit = false, fsr = false, work = false;
};

struct Pack {
    long long volume;
    int n;
    double score;
    std::vector<plp::PlacedBox> placed;
};

struct Row {
    plp::RunResult res;
    std::vector<double> fsr;            // per-box FSR of the selected packing
    std::vector<unsigned long long> pass_pair_box;
    std::vector<long long> pass_volume;
    std::vector<double> pass_score;
    size_t max_cand = 0;
};

Row run_fixed(const plp::Instance& inst, long long seed, const Opts& o) {
    plp::PyRng rng((uint64_t)seed);
    CountingBackend be;
    std::vector<Pack> packs;
    Row row;
    for (int p = 0; p < o.passes; p++) {
        unsigned long long before = be.pair_box;
        auto placed = plp::decode_pass(inst, o.order, o.first_fit, o.eps, rng, be);
        long long vol = 0;
        for (const auto& b : placed) vol += 1LL * b.w * b.d * b.h;
        double sc = plp::stability_score(placed);
        row.pass_pair_box.push_back(be.pair_box - before);
        row.pass_volume.push_back(vol);
        row.pass_score.push_back(sc);
        packs.push_back(Pack{vol, (int)placed.size(), sc, std::move(placed)});
    }
    row.max_cand = be.max_cand;
    // D4' selection over the fixed set of packings (same rule as
    // plp_common.hpp:493-511: accepted pool if non-empty, else all packings;
    // strict > on (volume, n) so the earliest winner is kept).
    std::vector<const Pack*> pool;
    for (const auto& p : packs)
        if (o.no_theta || p.score >= o.theta) pool.push_back(&p);
    bool fallback = pool.empty();
    if (fallback)
        for (const auto& p : packs) pool.push_back(&p);
    const Pack* best = pool.front();
    for (size_t i = 1; i < pool.size(); i++)
        if (pool[i]->volume > best->volume ||
            (pool[i]->volume == best->volume && pool[i]->n > best->n))
            best = pool[i];
    plp::RunResult& r = row.res;
    for (const auto& b : best->placed)
        r.placements.push_back({b.item_type, b.x, b.y, b.z, b.w, b.d, b.h});
    r.n_boxes = best->n;
    r.volume = best->volume;
    r.score = best->score;
    r.fallback = fallback;
    r.n_passes = o.passes;
    r.seed = (int)seed;
    r.time_s = 0.0;
    if (o.fsr) row.fsr = plp::compute_fsr(best->placed);
    return row;
}

void print_row(const Row& row, const Opts& o, long long seed) {
    std::string js = plp::to_json(row.res);
    // RunResult::seed is an int (plp_common.hpp:466); print the full 64-bit
    // seed the PyRng was built from, as the Python solver would.
    const std::string key = "\"seed\": " + std::to_string(row.res.seed) + ",";
    const size_t at = js.find(key);
    if (at != std::string::npos)
        js.replace(at, key.size(), "\"seed\": " + std::to_string(seed) + ",");
    js.pop_back();   // reopen the object to append the harness fields
    std::cout << js;
    if (o.fsr) {
        std::cout << ", \"fsr_hex\": [";
        for (size_t i = 0; i < row.fsr.size(); i++) {
            uint64_t bits;
            std::memcpy(&bits, &row.fsr[i], 8);
            char buf[32];
            std::snprintf(buf, sizeof buf, "\"%016llx\"", (unsigned long long)bits);
            if (i) std::cout << ", ";
            std::cout << buf;
        }
        std::cout << "]";
    }
    if (o.work) {
        std::cout << ", \"max_cand\": " << row.max_cand << ", \"pass_pair_box\": [";
        for (size_t i = 0; i < row.pass_pair_box.size(); i++)
            std::cout << (i ? ", " : "") << row.pass_pair_box[i];
        std::cout << "], \"pass_volume\": [";
        for (size_t i = 0; i < row.pass_volume.size(); i++)
            std::cout << (i ? ", " : "") << row.pass_volume[i];
        std::cout << "], \"pass_score\": [";
        for (size_t i = 0; i < row.pass_score.size(); i++)
            std::cout << (i ? ", " : "") << plp::py_repr(row.pass_score[i]);
        std::cout << "]";
    }
    std::cout << "}\n";
}

int usage() {
    std::fprintf(stderr,
                 "usage: ref_driver run|rows|bench <instance> [<instance> ... (bench)] "
                 "[--seed N] [--passes P] "
                 "[--K K] [--threads T] [--eps F] [--theta F|--no-theta] "
                 "[--order volume|input] [--first-fit] [--mu F] [--fsr] [--work]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    std::string mode = argv[1];
    Opts o;
    o.inst_path = argv[2];
    int first_opt = 3;
    while (first_opt < argc && argv[first_opt][0] != '-') o.more_paths.push_back(argv[first_opt++]);
    if (!o.more_paths.empty() && mode != "bench") return usage();
    for (int i = first_opt; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { std::exit(usage()); }
            return argv[++i];
        };
        if (a == "--seed") o.seed = std::atoll(next().c_str());
        else if (a == "--passes") o.passes = std::atoi(next().c_str());
        else if (a == "--K") o.K = std::atoi(next().c_str());
        else if (a == "--threads") o.threads = std::atoi(next().c_str());
        else if (a == "--eps") o.eps = std::strtod(next().c_str(), nullptr);
        else if (a == "--theta") o.theta = std::strtod(next().c_str(), nullptr);
        else if (a == "--no-theta") o.no_theta = true;
        else if (a == "--order") o.order = next();
        else if (a == "--first-fit") o.first_fit = true;
        else if (a == "--mu") plp::MU = std::strtod(next().c_str(), nullptr);
        else if (a == "--fsr") o.fsr = true;
        else if (a == "--work") o.work = true;
        else return usage();
    }
    if (o.passes < 1 || o.K < 1 || o.threads < 1) return usage();
    std::vector<plp::Instance> insts;
    try {
        insts.push_back(plp::load_instance(o.inst_path));
        for (const auto& pth : o.more_paths) insts.push_back(plp::load_instance(pth));
    } catch (const plp::InstanceFormatError& e) {
        std::fprintf(stderr, "instance error: %s\n", e.what());
        return 1;
    }
    const plp::Instance& inst = insts[0];
    if (mode == "run") {
        print_row(run_fixed(inst, o.seed, o), o, o.seed);
        return 0;
    }
    if (mode == "rows") {
        for (int k = 0; k < o.K; k++) print_row(run_fixed(inst, o.seed + k, o), o, o.seed + k);
        return 0;
    }
    if (mode == "bench") {
        std::atomic<int> next_row{0};
        std::atomic<long long> checksum{0};
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int t = 0; t < o.threads; t++)
            pool.emplace_back([&]() {
                for (;;) {
                    // decoder k: instance k / K with seed o.seed + k % K
                    int k = next_row.fetch_add(1);
                    if (k >= o.K * (int)insts.size()) break;
                    Row r = run_fixed(insts[(size_t)(k / o.K)], o.seed + k % o.K, o);
                    checksum.fetch_add(r.res.volume);
                }
            });
        for (auto& th : pool) th.join();
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double packings = (double)o.K * (double)insts.size() * o.passes;
        std::printf("{\"packings\": %.0f, \"seconds\": %.6f, \"packings_per_s\": %.3f, "
                    "\"threads\": %d, \"checksum_volume\": %lld}\n",
                    packings, el, packings / el, o.threads, checksum.load());
        return 0;
    }
    return usage();
}
