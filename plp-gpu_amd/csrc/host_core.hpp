// host_core.hpp -- host side of the decoder contract (C++17).
//
// The pieces of gpu/plp_common.hpp that stay on the host in every GPU form:
// the instance model and thpack parser, orientation lists, the item
// sequence, CPython's MT19937 for the host-driven plain form, D4' selection
// and the CPython-compatible JSON writer.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace plp {

struct Item {
    int l, w, h;
    bool lf, wf, hf;
    int qty;
    long long volume() const { return 1LL * l * w * h; }
};

struct Instance {
    int L = 0, W = 0, H = 0;
    std::vector<Item> items;
};

class InstanceFormatError : public std::runtime_error {
public:
    explicit InstanceFormatError(const std::string& m) : std::runtime_error(m) {}
};

// thpack "original" format (plp_common.hpp:68-141, parser.py:67-139).
std::vector<Instance> parse(const std::string& text);
// Exactly one instance (plp_common.hpp:143-151).
Instance load_instance(const std::string& path);

using Dims = std::array<int, 3>;
// Unique (w, d, h) in parser order (plp_common.hpp:269-280).
std::vector<std::array<int, 3>> orientations(const Item& it);
// Item type of every (type, copy) step: (-volume, type, copy) or file order
// (plp_common.hpp:335-351).  Throws std::invalid_argument on other orders.
std::vector<int> item_sequence(const Instance& inst, const std::string& order);

// CPython random.Random(int) (plp_common.hpp:164-250).
class PyRng {
public:
    explicit PyRng(uint64_t seed);
    uint32_t next();
    double random();
    uint32_t getrandbits(int k) { return next() >> (32 - k); }
    int randrange(int n);

private:
    uint32_t mt_[624];
    int mti_;
};

struct RunResult {
    std::vector<std::array<int, 7>> placements;   // (type, x, y, z, w, d, h)
    int n_boxes = 0;
    long long volume = 0;
    double score = 0.0;
    bool fallback = false;
    int n_passes = 0;
    double time_s = 0.0;
    long long seed = 0;
};

// CPython float repr (plp_common.hpp:535-590): shortest round-trip digits,
// fixed notation iff -4 < decimal exponent <= 16.
std::string py_repr(double x);
// json.dumps layout of solver.py:298-307 (plp_common.hpp:592-610).
std::string to_json(const RunResult& r);

}  // namespace plp
