// host_core.cpp -- see host_core.hpp.
#include "host_core.hpp"

#include <algorithm>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string_view>

namespace plp {

namespace {

// The thpack text read as records: each record is one line that holds at
// least one field, its fields being the runs of non-whitespace characters
// (the blank and whitespace-only lines carry nothing and are passed over).
// One record of look-ahead, so the grammar can ask whether more follow.
class Records {
public:
    explicit Records(std::string_view text) : rest_(text) { advance(); }
    bool more() const { return have_; }
    const std::vector<std::string_view>& peek() const { return cur_; }
    // Current record, then move past it.
    std::vector<std::string_view> take() {
        std::vector<std::string_view> r;
        r.swap(cur_);
        advance();
        return r;
    }

private:
    static bool blank(char c) {
        return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
    }
    void advance() {
        have_ = false;
        cur_.clear();
        while (!have_ && !rest_.empty()) {
            const size_t eol = rest_.find('\n');
            std::string_view line = rest_.substr(0, eol);
            rest_ = eol == std::string_view::npos ? std::string_view() : rest_.substr(eol + 1);
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && blank(line[i])) i++;
                const size_t start = i;
                while (i < line.size() && !blank(line[i])) i++;
                if (i > start) cur_.push_back(line.substr(start, i - start));
            }
            have_ = !cur_.empty();
        }
    }
    std::string_view rest_;
    std::vector<std::string_view> cur_;
    bool have_ = false;
};

[[noreturn]] void reject(const std::string& why) { throw InstanceFormatError(why); }

// A decimal integer field, read the way C's strtoll(s, &end, 10) reads it
// (optional sign, digits, saturation at the long long range); anything else
// in the field is an error naming the field and its role.
long long field_value(std::string_view f, const char* role) {
    f = f.substr(0, f.find('\0'));   // strtoll stops at a NUL byte as at the field's end
    size_t i = 0;
    const bool neg = i < f.size() && f[i] == '-';
    if (i < f.size() && (f[i] == '-' || f[i] == '+')) i++;
    const size_t first_digit = i;
    unsigned long long mag = 0;
    bool saturated = false;
    const unsigned long long limit =
        neg ? (unsigned long long)LLONG_MAX + 1ull : (unsigned long long)LLONG_MAX;
    for (; i < f.size() && f[i] >= '0' && f[i] <= '9'; i++) {
        const unsigned digit = (unsigned)(f[i] - '0');
        if (saturated || mag > (limit - digit) / 10) saturated = true;
        else mag = mag * 10 + digit;
    }
    if (i == first_digit || i != f.size())
        reject("not an integer: " + std::string(f) + " (" + role + ")");
    if (saturated) mag = limit;
    return neg ? (long long)(0ull - mag) : (long long)mag;
}

bool is_flag(std::string_view f) { return f == "0" || f == "1"; }

// The record the grammar expects next: present and with a field count that
// the rule accepts, else the rule's message.
template <class Fits>
std::vector<std::string_view> expect(Records& rec, Fits fits, const char* absent,
                                     const char* wrong_shape) {
    if (!rec.more()) reject(absent);
    if (!fits(rec.peek().size())) reject(wrong_shape);
    return rec.take();
}

// One item line: "k l lf w wf h hf qty", k counting from 1.
Item item_record(Records& rec, long long k) {
    const auto f = expect(rec, [](size_t n) { return n == 8; }, "instance truncated",
                          "item line must have 8 tokens");
    if (field_value(f[0], "item index") != k) reject("item index must be sequential");
    long long dim[3];
    for (int a = 0; a < 3; a++) dim[a] = field_value(f[1 + 2 * a], "item dimension");
    if (dim[0] <= 0 || dim[1] <= 0 || dim[2] <= 0) reject("item dimensions must be positive");
    if (!is_flag(f[2]) || !is_flag(f[4]) || !is_flag(f[6]))
        reject("orientation flags must be 0/1");
    const long long qty = field_value(f[7], "item qty");
    if (qty < 0) reject("item qty must be >= 0");
    return Item{(int)dim[0], (int)dim[1], (int)dim[2], f[2] == "1", f[4] == "1", f[6] == "1",
                (int)qty};
}

// One problem: "id [seed]", "L W H", "n", then n item lines.
Instance problem_records(Records& rec) {
    expect(rec, [](size_t n) { return n <= 2; }, "malformed problem line", "malformed problem line");
    const auto box = expect(rec, [](size_t n) { return n == 3; },
                            "container line must have 3 dimensions",
                            "container line must have 3 dimensions");
    long long side[3];
    for (int a = 0; a < 3; a++) side[a] = field_value(box[(size_t)a], "container dimension");
    if (side[0] <= 0 || side[1] <= 0 || side[2] <= 0)
        reject("container dimensions must be positive");
    const auto count = expect(rec, [](size_t n) { return n == 1; },
                              "type-count line must be a single integer",
                              "type-count line must be a single integer");
    const long long n_types = field_value(count[0], "type count");
    Instance inst;
    inst.L = (int)side[0];
    inst.W = (int)side[1];
    inst.H = (int)side[2];
    for (long long k = 1; k <= n_types; k++) inst.items.push_back(item_record(rec, k));
    return inst;
}

}  // namespace

// Grammar of the thpack "original" format with the reference's error
// messages (plp/parser.py:67-139; messages and check order as
// gpu/plp_common.hpp:68-141 reports them): a count line whose first field is
// the number of problems P >= 1, P problems, nothing after them.
std::vector<Instance> parse(const std::string& text) {
    Records rec{std::string_view(text)};
    if (!rec.more()) reject("empty file");
    const long long n_problems = field_value(rec.take()[0], "instance count");
    if (n_problems <= 0 || !rec.more()) reject("instance count must be >= 1");
    std::vector<Instance> out;
    for (long long p = 0; p < n_problems; p++) out.push_back(problem_records(rec));
    if (rec.more()) reject("trailing data after last instance");
    return out;
}

Instance load_instance(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) reject("cannot open " + path);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<Instance> all = parse(text);
    if (all.size() != 1)
        reject("expected exactly 1 instance, got " + std::to_string(all.size()));
    return std::move(all.front());
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
