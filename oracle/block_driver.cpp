// block_driver.cpp -- TEST INFRASTRUCTURE ONLY (never shipped, never linked
// into the product).
//
// A harness around the reference's block-building engine
// (blockproof/plp_block.cpp, compiled into this translation unit from the
// reference checkout by oracle/Makefile with its main() renamed; nothing from
// it is copied here).  The reference CLI stops only on a wall-clock budget or
// after one pass (plp_block.cpp:238-279), so this driver calls its block_pass
// (plp_block.cpp:95) a fixed number of times, with the catalogue of
// run_block (plp_block.cpp:193-218) and its online D4' selection
// (plp_block.cpp:228-275) restated for a fixed pass count:
//
//   run    one decoder: prints plp::to_json (plp_common.hpp:592) with time_s
//          0.0 plus the volume and score of every pass
//   rows   K decoders seeded seed..seed+K-1, one JSON line each
//   bench  the same rows over T std::threads, timed (CPU baseline)
#define main plp_block_reference_main
#include "plp_block.cpp"
#undef main

#include <atomic>
#include <chrono>
#include <iostream>
#include <thread>

namespace {

struct Opts {
    std::string inst_path;
    long long seed = 0;
    int passes = 1, K = 1, threads = 1;
    int amax = 3, bmax = 3, kmax = 4, topk = 8;
    double eps = EPS_QUALITY, theta = THETA;
    bool no_theta = false;
    PickRule rule = PICK_FIRST;
};

struct Catalogue {
    std::vector<std::vector<Pt>> orients;
    std::vector<BlockShape> cat;
    std::vector<std::vector<int>> by_type;
};

// The block catalogue of run_block (plp_block.cpp:193-218).
Catalogue catalogue(const Instance& inst, const Opts& o) {
    Catalogue c;
    for (const auto& it : inst.items) c.orients.push_back(orientations(it));
    for (size_t ti = 0; ti < inst.items.size(); ti++)
        for (size_t oi = 0; oi < c.orients[ti].size(); oi++) {
            const int w = c.orients[ti][oi][0], d = c.orients[ti][oi][1], h = c.orients[ti][oi][2];
            for (int a = 1; a <= o.amax && a * w <= inst.L; a++)
                for (int b = 1; b <= o.bmax && b * d <= inst.W; b++)
                    for (int k = 1; k <= o.kmax && k * h <= inst.H; k++) {
                        const long long cnt = 1LL * a * b * k;
                        if (cnt > inst.items[ti].qty) break;
                        c.cat.push_back({(int)ti, (int)oi, a, b, k, cnt, a * w, b * d, k * h});
                    }
        }
    c.by_type.resize(inst.items.size());
    for (size_t i = 0; i < c.cat.size(); i++) c.by_type[(size_t)c.cat[i].ti].push_back((int)i);
    for (auto& v : c.by_type)
        std::sort(v.begin(), v.end(),
                  [&](int x, int y) { return shape_precedence(c.cat[(size_t)x], c.cat[(size_t)y]); });
    return c;
}

struct Row {
    RunResult r;
    std::vector<long long> pass_volume;
    std::vector<double> pass_score;
};

// P passes of block_pass with one PyRng, D4' kept online as run_block does.
Row run_one(const Instance& inst, const Catalogue& c, const Opts& o, uint64_t seed) {
    PyRng rng(seed);
    Row row;
    long long v_acc = -1, v_all = -1;
    int n_acc = 0, n_all = 0;
    double s_acc = 0.0, s_all = 0.0;
    bool has_acc = false;
    std::vector<PlacedBox> p_acc, p_all;
    for (int p = 0; p < o.passes; p++) {
        BlockPass bp = block_pass(inst, c.orients, c.cat, c.by_type, o.eps, o.rule, o.topk, rng);
        const double sc = stability_score(bp.items);
        row.pass_volume.push_back(bp.volume);
        row.pass_score.push_back(sc);
        const bool upd_all = bp.volume > v_all || (bp.volume == v_all && bp.n > n_all);
        const bool upd_acc = (o.no_theta || sc >= o.theta) &&
                             (!has_acc || bp.volume > v_acc || (bp.volume == v_acc && bp.n > n_acc));
        if (upd_acc) {
            v_acc = bp.volume; n_acc = bp.n; s_acc = sc; p_acc = bp.items; has_acc = true;
        }
        if (upd_all) {
            v_all = bp.volume; n_all = bp.n; s_all = sc; p_all = bp.items;
        }
    }
    const std::vector<PlacedBox>& best = has_acc ? p_acc : p_all;
    for (const auto& b : best) row.r.placements.push_back({b.item_type, b.x, b.y, b.z, b.w, b.d, b.h});
    row.r.n_boxes = has_acc ? n_acc : n_all;
    row.r.volume = has_acc ? v_acc : v_all;
    row.r.score = has_acc ? s_acc : s_all;
    row.r.fallback = !has_acc;
    row.r.n_passes = o.passes;
    row.r.seed = (long long)seed;
    row.r.time_s = 0.0;
    return row;
}

std::string row_json(const Row& row, uint64_t seed) {
    std::string j = to_json(row.r);
    // full 64-bit seed (RunResult::seed is a long long)
    const std::string key = "\"seed\": ";
    const size_t at = j.find(key);
    const size_t end = j.find(',', at);
    j = j.substr(0, at + key.size()) + std::to_string(seed) + j.substr(end);
    j.pop_back();
    j += ", \"pass_volume\": [";
    for (size_t i = 0; i < row.pass_volume.size(); i++)
        j += (i ? ", " : "") + std::to_string(row.pass_volume[i]);
    j += "], \"pass_score\": [";
    for (size_t i = 0; i < row.pass_score.size(); i++)
        j += (i ? ", " : "") + py_repr(row.pass_score[i]);
    j += "]}";
    return j;
}

int usage() {
    std::fprintf(stderr,
                 "usage: block_driver run|rows|bench <instance> [--seed S] [--passes P] [--K K]\n"
                 "       [--threads T] [--eps F] [--theta F | --no-theta] [--amax A] [--bmax B]\n"
                 "       [--kmax K] [--topk N] [--pick first|support|contact]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    const std::string mode = argv[1];
    Opts o;
    o.inst_path = argv[2];
    for (int i = 3; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) std::exit(usage());
            return argv[++i];
        };
        if (a == "--seed") o.seed = std::stoll(next());
        else if (a == "--passes") o.passes = std::stoi(next());
        else if (a == "--K") o.K = std::stoi(next());
        else if (a == "--threads") o.threads = std::stoi(next());
        else if (a == "--eps") o.eps = std::stod(next());
        else if (a == "--theta") o.theta = std::stod(next());
        else if (a == "--no-theta") o.no_theta = true;
        else if (a == "--amax") o.amax = std::max(1, std::stoi(next()));
        else if (a == "--bmax") o.bmax = std::max(1, std::stoi(next()));
        else if (a == "--kmax") o.kmax = std::max(1, std::stoi(next()));
        else if (a == "--topk") o.topk = std::max(1, std::stoi(next()));
        else if (a == "--pick") {
            const std::string v = next();
            o.rule = v == "support" ? PICK_SUPPORT : v == "contact" ? PICK_CONTACT : PICK_FIRST;
        } else return usage();
    }
    const Instance inst = load_instance(o.inst_path);
    const Catalogue c = catalogue(inst, o);
    if (mode == "run") {
        std::cout << row_json(run_one(inst, c, o, (uint64_t)o.seed), (uint64_t)o.seed) << "\n";
        return 0;
    }
    if (mode == "rows") {
        for (int k = 0; k < o.K; k++) {
            const uint64_t s = (uint64_t)o.seed + (uint64_t)k;
            std::cout << row_json(run_one(inst, c, o, s), s) << "\n";
        }
        return 0;
    }
    if (mode == "bench") {
        std::atomic<int> next_k{0};
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (int t = 0; t < std::max(1, o.threads); t++)
            th.emplace_back([&]() {
                for (int k; (k = next_k++) < o.K;)
                    (void)run_one(inst, c, o, (uint64_t)o.seed + (uint64_t)k);
            });
        for (auto& x : th) x.join();
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("{\"packings\": %lld, \"seconds\": %.6f, \"packings_per_s\": %.3f}\n",
                    (long long)o.K * o.passes, s, (double)o.K * o.passes / s);
        return 0;
    }
    return usage();
}
