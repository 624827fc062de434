// host_core.cpp -- see host_core.hpp.
#include "host_core.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace plp {

namespace {

std::vector<std::vector<std::string>> tokenize_lines(const std::string& text) {
    std::vector<std::vector<std::string>> out;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<std::string> t;
        std::string w;
        while (ls >> w) t.push_back(w);
        if (!t.empty()) out.push_back(std::move(t));
    }
    return out;
}

long long to_int(const std::string& s, const char* what) {
    char* end = nullptr;
    long long v = std::strtoll(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0')
        throw InstanceFormatError(std::string("not an integer: ") + s + " (" + what + ")");
    return v;
}

}  // namespace

// Format and errors of parser.py:67-139; token rules of plp_common.hpp:68-141.
std::vector<Instance> parse(const std::string& text) {
    auto toks = tokenize_lines(text);
    if (toks.empty()) throw InstanceFormatError("empty file");
    long long P = to_int(toks[0][0], "instance count");
    if (P <= 0 || toks.size() < 2) throw InstanceFormatError("instance count must be >= 1");
    std::vector<Instance> out;
    size_t pos = 1;
    for (long long p = 0; p < P; p++) {
        if (pos >= toks.size() || toks[pos].size() > 2)
            throw InstanceFormatError("malformed problem line");
        pos++;
        if (pos >= toks.size() || toks[pos].size() != 3)
            throw InstanceFormatError("container line must have 3 dimensions");
        long long L = to_int(toks[pos][0], "container dimension");
        long long W = to_int(toks[pos][1], "container dimension");
        long long H = to_int(toks[pos][2], "container dimension");
        if (L <= 0 || W <= 0 || H <= 0)
            throw InstanceFormatError("container dimensions must be positive");
        pos++;
        if (pos >= toks.size() || toks[pos].size() != 1)
            throw InstanceFormatError("type-count line must be a single integer");
        long long n = to_int(toks[pos][0], "type count");
        pos++;
        Instance inst;
        inst.L = (int)L;
        inst.W = (int)W;
        inst.H = (int)H;
        for (long long k = 1; k <= n; k++) {
            if (pos >= toks.size()) throw InstanceFormatError("instance truncated");
            const auto& t = toks[pos++];
            if (t.size() != 8) throw InstanceFormatError("item line must have 8 tokens");
            if (to_int(t[0], "item index") != k)
                throw InstanceFormatError("item index must be sequential");
            long long dl = to_int(t[1], "item dimension");
            long long dw = to_int(t[3], "item dimension");
            long long dh = to_int(t[5], "item dimension");
            if (dl <= 0 || dw <= 0 || dh <= 0)
                throw InstanceFormatError("item dimensions must be positive");
            for (int f : {2, 4, 6})
                if (t[f] != "0" && t[f] != "1")
                    throw InstanceFormatError("orientation flags must be 0/1");
            long long qty = to_int(t[7], "item qty");
            if (qty < 0) throw InstanceFormatError("item qty must be >= 0");
            inst.items.push_back({(int)dl, (int)dw, (int)dh, t[2] == "1", t[4] == "1",
                                  t[6] == "1", (int)qty});
        }
        out.push_back(std::move(inst));
    }
    if (pos != toks.size()) throw InstanceFormatError("trailing data after last instance");
    return out;
}

Instance load_instance(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw InstanceFormatError("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    auto v = parse(ss.str());
    if (v.size() != 1)
        throw InstanceFormatError("expected exactly 1 instance, got " + std::to_string(v.size()));
    return v[0];
}

std::vector<std::array<int, 3>> orientations(const Item& it) {
    std::vector<std::array<int, 3>> out;
    auto add = [&](int w, int d, int h) {
        std::array<int, 3> o{w, d, h};
        if (std::find(out.begin(), out.end(), o) == out.end()) out.push_back(o);
    };
    if (it.lf) { add(it.w, it.h, it.l); add(it.h, it.w, it.l); }
    if (it.wf) { add(it.l, it.h, it.w); add(it.h, it.l, it.w); }
    if (it.hf) { add(it.l, it.w, it.h); add(it.w, it.l, it.h); }
    return out;
}

std::vector<int> item_sequence(const Instance& inst, const std::string& order) {
    struct Step {
        long long vol;
        int type, copy;
    };
    std::vector<Step> seq;
    for (size_t t = 0; t < inst.items.size(); t++)
        for (int c = 0; c < inst.items[t].qty; c++)
            seq.push_back({inst.items[t].volume(), (int)t, c});
    if (order == "volume") {
        std::sort(seq.begin(), seq.end(), [](const Step& a, const Step& b) {
            if (a.vol != b.vol) return a.vol > b.vol;
            if (a.type != b.type) return a.type < b.type;
            return a.copy < b.copy;
        });
    } else if (order != "input") {
        throw std::invalid_argument("unknown order " + order);
    }
    std::vector<int> types(seq.size());
    for (size_t i = 0; i < seq.size(); i++) types[i] = seq[i].type;
    return types;
}

// ------------------------------------------------------------------ PyRng

PyRng::PyRng(uint64_t seed) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int keyused = key[1] ? 2 : 1;
    mt_[0] = 19650218u;
    for (int i = 1; i < 624; i++)
        mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
    int i = 1, j = 0;
    for (int k = 624; k; k--) {
        mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++;
        j++;
        if (i >= 624) { mt_[0] = mt_[623]; i = 1; }
        if (j >= keyused) j = 0;
    }
    for (int k = 623; k; k--) {
        mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= 624) { mt_[0] = mt_[623]; i = 1; }
    }
    mt_[0] = 0x80000000u;
    mti_ = 624;
}

uint32_t PyRng::next() {
    if (mti_ >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % 624] & 0x7fffffffu);
            mt_[i] = mt_[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        mti_ = 0;
    }
    uint32_t y = mt_[mti_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double PyRng::random() {
    uint32_t a = next() >> 5;
    uint32_t b = next() >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

int PyRng::randrange(int n) {
    int k = 0;
    for (unsigned t = (unsigned)n; t; t >>= 1) k++;
    int r = (int)getrandbits(k);
    while (r >= n) r = (int)getrandbits(k);
    return r;
}

// ------------------------------------------------------------------- JSON

std::string py_repr(double x) {
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    if (std::isnan(x)) return "nan";
    if (std::isinf(x)) return x < 0 ? "-inf" : "inf";
    // std::to_chars(scientific) without a precision gives the shortest digit
    // string that reads back to x, nearest to x on ties -- the digits CPython's
    // repr uses.
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof buf, std::fabs(x), std::chars_format::scientific);
    std::string sci(buf, res.ptr);
    size_t epos = sci.find('e');
    int e10 = std::atoi(sci.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; i++)
        if (sci[i] != '.') digits.push_back(sci[i]);
    const int nd = (int)digits.size();
    const int decpt = e10 + 1;   // |x| = 0.digits * 10^decpt
    std::string out = x < 0 ? "-" : "";
    if (decpt > -4 && decpt <= 16) {
        if (decpt <= 0) {
            out += "0." + std::string((size_t)(-decpt), '0') + digits;
        } else if (decpt >= nd) {
            out += digits + std::string((size_t)(decpt - nd), '0') + ".0";
        } else {
            out += digits.substr(0, (size_t)decpt) + "." + digits.substr((size_t)decpt);
        }
    } else {
        out += digits.substr(0, 1);
        if (nd > 1) out += "." + digits.substr(1);
        int ex = decpt - 1;
        char eb[16];
        std::snprintf(eb, sizeof eb, "e%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        out += eb;
    }
    return out;
}

namespace {

template <class T>
void append_int(std::string& o, T v) {
    char buf[24];
    const auto res = std::to_chars(buf, buf + sizeof buf, v);
    o.append(buf, res.ptr);
}

}  // namespace

std::string to_json(const RunResult& r) {
    std::string o;
    o.reserve(160 + r.placements.size() * 36);
    o += "{\"n_boxes\": ";
    append_int(o, r.n_boxes);
    o += ", \"volume_cm3\": ";
    append_int(o, r.volume);
    o += ", \"stability_score\": ";
    o += py_repr(r.score);
    o += r.fallback ? ", \"fallback\": true" : ", \"fallback\": false";
    o += ", \"n_passes\": ";
    append_int(o, r.n_passes);
    o += ", \"time_s\": ";
    o += py_repr(r.time_s);
    o += ", \"seed\": ";
    append_int(o, r.seed);
    o += ", \"placements\": [";
    for (size_t i = 0; i < r.placements.size(); i++) {
        const auto& p = r.placements[i];
        o += i ? ", [" : "[";
        for (int c = 0; c < 7; c++) {
            if (c) o += ", ";
            append_int(o, p[(size_t)c]);
        }
        o += ']';
    }
    o += "]}";
    return o;
}

}  // namespace plp
