// serve_cli.cpp -- `plp_gpu --serve`: the continuous-batching server over a
// JSON-lines wire format (no reference counterpart: the reference decodes one
// instance per process).
//
// Request, one JSON object per line on stdin:
//   {"id": 7, "instance": "<thpack file>" | "instance_text": "<thpack text>",
//    "decoders": K, "seed": S, "passes": P, "eps": F, "theta": F | null,
//    "mu": F, "order": "volume" | "input", "first_fit": false}
// Defaults: decoders 1, seed 0, passes 1, eps 0.3, theta 0.9 (null = the
// reference's --no-theta), mu from PLP_MU or 0.2, order volume.  Decoder k
// uses seed S + k, as the reference's --device rows (plp_gpu.cu:350).
// Response, one line on stdout per request in completion order:
//   {"id": 7, "K": K, "n_passes": P, "time_s": arrival to completion,
//    "format_ms": host time to write the rows, "rows": [<row>, ...]}
// where each row is the reference's single-run JSON (plp_common.hpp:592-610;
// its time_s is the request's).  A request that cannot run gets
// {"id": ..., "error": "<message>"}.  The server admits requests while a
// slot and device memory are free (plp_server_submit) and exits when stdin
// ends and every admitted request has been answered.
#include <algorithm>
#include <array>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <iostream>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_core.hpp"
#include "plp_hip.h"
#include "serve_cli.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// A JSON scalar of a flat request object; raw keeps the literal text (the
// id is echoed back verbatim).
struct JVal {
    enum Kind { STR, NUM, BOOL, NUL } kind = NUL;
    std::string str, raw;
    double num = 0.0;
    bool b = false;
};

struct ParseError {
    std::string what;
};

class FlatJson {
public:
    explicit FlatJson(const std::string& s) : s_(s) {}

    std::map<std::string, JVal> object() {
        std::map<std::string, JVal> out;
        ws();
        expect('{');
        ws();
        if (peek() == '}') {
            i_++;
            return out;
        }
        for (;;) {
            ws();
            const std::string key = string_lit();
            ws();
            expect(':');
            ws();
            out[key] = value();
            ws();
            if (peek() == ',') {
                i_++;
                continue;
            }
            expect('}');
            break;
        }
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return out;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& m) {
        throw ParseError{m + " at offset " + std::to_string(i_)};
    }
    char peek() const { return i_ < s_.size() ? s_[i_] : '\0'; }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        i_++;
    }
    static void put_utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string string_lit() {
        expect('"');
        std::string o;
        for (;;) {
            if (i_ >= s_.size()) fail("unterminated string");
            const char c = s_[i_++];
            if (c == '"') break;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (i_ >= s_.size()) fail("unterminated escape");
            const char e = s_[i_++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    if (i_ + 4 > s_.size()) fail("short \\u escape");
                    put_utf8(o, (unsigned)std::stoul(s_.substr(i_, 4), nullptr, 16));
                    i_ += 4;
                    break;
                }
                default: fail("bad escape");
            }
        }
        return o;
    }
    JVal value() {
        JVal v;
        const size_t start = i_;
        const char c = peek();
        if (c == '"') {
            v.kind = JVal::STR;
            v.str = string_lit();
        } else if (s_.compare(i_, 4, "true") == 0) {
            v.kind = JVal::BOOL;
            v.b = true;
            i_ += 4;
        } else if (s_.compare(i_, 5, "false") == 0) {
            v.kind = JVal::BOOL;
            i_ += 5;
        } else if (s_.compare(i_, 4, "null") == 0) {
            v.kind = JVal::NUL;
            i_ += 4;
        } else {
            char* end = nullptr;
            v.num = std::strtod(s_.c_str() + i_, &end);
            if (end == s_.c_str() + i_) fail("expected a value");
            v.kind = JVal::NUM;
            i_ = (size_t)(end - s_.c_str());
        }
        v.raw = s_.substr(start, i_ - start);
        return v;
    }
};

struct Request {
    std::string id_raw = "null";
    Clock::time_point arrival;
    plp::Instance inst;
    std::vector<int32_t> types;
    plp_instance_desc desc{};
    int32_t n_dec = 1, passes = 1, n_items = 1;
    uint64_t seed = 0;
    plp_params params{};
    int64_t rid = -1;
};

std::string error_line(const std::string& id_raw, const std::string& msg) {
    std::string esc;
    for (char c : msg) {
        if (c == '"' || c == '\\') esc += '\\';
        if ((unsigned char)c < 0x20) {
            esc += ' ';
            continue;
        }
        esc += c;
    }
    return "{\"id\": " + id_raw + ", \"error\": \"" + esc + "\"}";
}

// Builds a request from one line; throws ParseError with the reason.
Request make_request(const std::string& line, double mu_default) {
    Request r;
    r.arrival = Clock::now();
    std::map<std::string, JVal> o = FlatJson(line).object();
    auto num = [&](const char* k, double dflt) {
        auto it = o.find(k);
        if (it == o.end()) return dflt;
        if (it->second.kind != JVal::NUM) throw ParseError{std::string(k) + " must be a number"};
        return it->second.num;
    };
    auto integer = [&](const char* k, double dflt, double lo, double hi) {
        const double v = num(k, dflt);
        if (v != std::floor(v) || v < lo || v > hi)
            throw ParseError{std::string(k) + " must be an integer in range"};
        return v;
    };
    if (o.count("id")) r.id_raw = o["id"].raw;
    if (o.count("instance") && o["instance"].kind == JVal::STR) {
        r.inst = plp::load_instance(o["instance"].str);
    } else if (o.count("instance_text") && o["instance_text"].kind == JVal::STR) {
        std::vector<plp::Instance> v = plp::parse(o["instance_text"].str);
        if (v.empty()) throw ParseError{"instance_text holds no instance"};
        r.inst = v[0];
    } else {
        throw ParseError{"request needs \"instance\" (a path) or \"instance_text\""};
    }
    r.n_dec = (int32_t)integer("decoders", 1, 1, 1 << 20);
    r.passes = (int32_t)integer("passes", 1, 1, 1 << 20);
    r.seed = (uint64_t)integer("seed", 0, 0, 9007199254740991.0);
    r.params.eps = num("eps", 0.3);
    r.params.mu = num("mu", mu_default);
    r.params.use_theta = 1;
    r.params.theta = 0.9;
    if (o.count("theta")) {
        if (o["theta"].kind == JVal::NUL) r.params.use_theta = 0;
        else r.params.theta = num("theta", 0.9);
    }
    r.params.first_fit = o.count("first_fit") && o["first_fit"].kind == JVal::BOOL && o["first_fit"].b;
    r.params.order = 0;
    if (o.count("order")) {
        if (o["order"].kind != JVal::STR || (o["order"].str != "volume" && o["order"].str != "input"))
            throw ParseError{"order must be \"volume\" or \"input\""};
        r.params.order = o["order"].str == "input" ? 1 : 0;
    }
    int64_t items = 0;
    for (const plp::Item& it : r.inst.items) {
        int32_t q[7] = {it.l, it.lf, it.w, it.wf, it.h, it.hf, it.qty};
        r.types.insert(r.types.end(), q, q + 7);
        items += it.qty;
    }
    r.n_items = (int32_t)std::max<int64_t>(1, items);
    r.desc = plp_instance_desc{r.inst.L, r.inst.W, r.inst.H, (int32_t)r.inst.items.size(),
                               r.types.empty() ? nullptr : r.types.data()};
    return r;
}

// The rows' JSON, written by up to 16 host threads (one contiguous range of
// rows each) and joined in row order.
std::string rows_json(const Request& r, const std::vector<plp_row_result>& rows,
                      const std::vector<int32_t>& pl, double t) {
    const int K = r.n_dec;
    std::vector<std::string> part;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = std::max(1, std::min({K / 256 + 1, 16, (int)hw}));
    part.resize((size_t)T);
    auto work = [&](int w) {
        std::string& o = part[(size_t)w];
        for (int k = (int)((int64_t)K * w / T); k < (int)((int64_t)K * (w + 1) / T); k++) {
            const plp_row_result& x = rows[(size_t)k];
            plp::RunResult res;
            res.n_boxes = x.n_boxes;
            res.volume = x.volume;
            res.score = x.score;
            res.fallback = x.fallback != 0;
            res.n_passes = x.n_passes;
            res.time_s = t;
            res.seed = (long long)x.seed;
            const int32_t* p = pl.data() + (size_t)k * (size_t)r.n_items * 7;
            res.placements.resize((size_t)x.n_boxes);
            for (int i = 0; i < x.n_boxes; i++)
                for (int c = 0; c < 7; c++) res.placements[(size_t)i][(size_t)c] = p[7 * i + c];
            if (k) o += ", ";
            o += plp::to_json(res);
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < T; w++) th.emplace_back(work, w);
    work(0);
    for (std::thread& x : th) x.join();
    size_t n = 0;
    for (const std::string& x : part) n += x.size();
    std::string body;
    body.reserve(n);
    for (const std::string& x : part) body += x;
    return body;
}

std::string response_line(const Request& r, const std::vector<plp_row_result>& rows,
                          const std::vector<int32_t>& pl, double* format_ms) {
    const auto t0 = Clock::now();
    const double t = std::chrono::duration<double>(t0 - r.arrival).count();
    const std::string body = rows_json(r, rows, pl, t);
    *format_ms = 1e3 * std::chrono::duration<double>(Clock::now() - t0).count();
    std::string line = "{\"id\": " + r.id_raw + ", \"K\": " + std::to_string(r.n_dec) +
                       ", \"n_passes\": " + std::to_string(r.n_dec ? rows[0].n_passes : 0) +
                       ", \"time_s\": " + plp::py_repr(t) + ", \"format_ms\": " +
                       plp::py_repr(*format_ms) + ", \"rows\": [";
    line.reserve(line.size() + body.size() + 2);
    line += body;
    line += "]}";
    return line;
}

}  // namespace

int serve_main(int max_inflight, double mu_default) {
    plp_server* srv = nullptr;
    int rc = plp_server_create(0, max_inflight, 0, &srv);
    if (rc != PLP_OK) {
        std::fprintf(stderr, "server init: %s\n", plp_last_error());
        return 1;
    }
    std::mutex mu;
    std::deque<std::string> lines;
    bool eof = false;
    std::thread reader([&]() {
        std::string l;
        while (std::getline(std::cin, l)) {
            std::lock_guard<std::mutex> g(mu);
            lines.push_back(l);
        }
        std::lock_guard<std::mutex> g(mu);
        eof = true;
    });
    std::deque<Request> pending;
    std::vector<Request> inflight;
    auto emit = [](const std::string& s) {
        std::fwrite(s.data(), 1, s.size(), stdout);
        std::fputc('\n', stdout);
        std::fflush(stdout);
    };
    int status = 0;
    for (;;) {
        bool input_done;
        std::deque<std::string> got;
        {
            std::lock_guard<std::mutex> g(mu);
            got.swap(lines);
            input_done = eof;
        }
        bool progress = !got.empty();
        for (const std::string& l : got) {
            if (l.find_first_not_of(" \t\r") == std::string::npos) continue;
            try {
                pending.push_back(make_request(l, mu_default));
            } catch (const ParseError& e) {
                std::string id = "null";
                try {
                    auto o = FlatJson(l).object();
                    if (o.count("id")) id = o["id"].raw;
                } catch (const ParseError&) {
                }
                emit(error_line(id, e.what));
            } catch (const std::exception& e) {
                std::string id = "null";
                try {
                    auto o = FlatJson(l).object();
                    if (o.count("id")) id = o["id"].raw;
                } catch (const ParseError&) {
                }
                emit(error_line(id, e.what()));
            }
        }
        while (!pending.empty()) {
            Request& r = pending.front();
            int64_t rid = -1;
            rc = plp_server_submit(srv, &r.desc, r.n_dec, r.seed, r.passes, &r.params, &rid);
            if (rc == PLP_E_BUSY) break;
            progress = true;
            if (rc != PLP_OK) {
                emit(error_line(r.id_raw, plp_last_error()));
                pending.pop_front();
                continue;
            }
            r.rid = rid;
            inflight.push_back(std::move(r));
            pending.pop_front();
        }
        for (size_t i = 0; i < inflight.size();) {
            Request& r = inflight[i];
            int32_t done = 0, mi = 0;
            rc = plp_server_query(srv, r.rid, &done);   // result buffers only once done
            if (!done && rc == PLP_OK) {
                i++;
                continue;
            }
            std::vector<plp_row_result> rows((size_t)r.n_dec);
            std::vector<int32_t> pl((size_t)r.n_dec * (size_t)r.n_items * 7);
            if (rc == PLP_OK)
                rc = plp_server_poll(srv, r.rid, &done, rows.data(), pl.data(),
                                     r.n_dec * r.n_items, &mi);
            progress = true;
            if (rc != PLP_OK) {
                emit(error_line(r.id_raw, plp_last_error()));
                if (!done) status = 1;   // the device failed: nothing more will finish
            } else {
                double fms = 0.0;
                emit(response_line(r, rows, pl, &fms));
            }
            inflight.erase(inflight.begin() + (std::ptrdiff_t)i);
            if (status) break;
        }
        if (status) break;
        if (input_done && pending.empty() && inflight.empty()) break;
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    if (status) {
        reader.detach();   // stdin may stay open; the process exits anyway
        return status;
    }
    reader.join();
    plp_server_destroy(srv);
    return status;
}
