// plain.cpp -- the host-driven GPU forms (plain and batched per step).
//
// Mirrors the reference's run() / decode_pass() (gpu/plp_common.hpp:404-523)
// and its K-row lockstep variant (decode_pass_batch / run_batch,
// plp_gpu.cu:279-412): the decode loop, RNG and extreme-point update run on
// the host; every item step makes one plp_feas_check_batch call (one H2D +
// one launch + one D2H + sync for all K rows; K = 1 is the CudaBackend
// seam, plp_gpu.cu:414-502) and each pass's stability scores come from one
// plp_fsr_eval call.  This is the per-step-round-trip baseline the
// device-resident engine is measured against.
#include "plain.hpp"

#include <algorithm>
#include <cstring>

using namespace plpi;

namespace plpi {

PlainRunner::PlainRunner(plp_ctx* ctx, const plp::Instance& inst, uint64_t seed0, int K,
                         const plp_params& p)
    : ctx_(ctx), inst_(inst), p_(p) {
    seq_ = plp::item_sequence(inst_, p.order == 0 ? "volume" : "input");
    for (const plp::Item& it : inst_.items) orients_.push_back(plp::orientations(it));
    use_rng_ = !(p.first_fit || p.eps == 0.0);
    for (int k = 0; k < K; k++) rows_.emplace_back(seed0 + (uint64_t)k);
}

int PlainRunner::run_pass() {
    const int K = (int)rows_.size();
    for (Row& r : rows_) {
        r.cand.assign(1, {0, 0, 0});
        r.aabb.clear();
        r.placed.clear();
        r.fail_ti = -1;
    }
    std::vector<const int32_t*> cp((size_t)K), bp((size_t)K);
    std::vector<int32_t> nc((size_t)K), nb((size_t)K);
    std::vector<uint8_t*> op((size_t)K);
    std::vector<std::vector<uint8_t>> bits((size_t)K);
    for (int ti : seq_) {
        const auto& ors = orients_[(size_t)ti];
        if (ors.empty()) continue;
        // A row whose previous copy of this type found nothing, with no
        // placement since, would get the same all-zero verdicts: it sits out.
        int n_live = 0;
        std::vector<int> live((size_t)K);
        for (int k = 0; k < K; k++) {
            Row& r = rows_[(size_t)k];
            r.work += (uint64_t)r.cand.size() * ors.size() * r.aabb.size();
            if (r.fail_ti == ti) continue;
            const size_t j = (size_t)n_live++;
            live[j] = k;
            bits[(size_t)k].assign(r.cand.size() * ors.size(), 0);
            cp[j] = &r.cand[0][0];
            nc[j] = (int32_t)r.cand.size();
            bp[j] = r.aabb.empty() ? nullptr : &r.aabb[0][0];
            nb[j] = (int32_t)r.aabb.size();
            op[j] = bits[(size_t)k].data();
        }
        if (n_live == 0) continue;
        int rc = plp_feas_check_batch(ctx_, n_live, cp.data(), nc.data(), &ors[0][0],
                                      (int32_t)ors.size(), bp.data(), nb.data(), inst_.L, inst_.W,
                                      inst_.H, op.data());
        if (rc != PLP_OK) return rc;
        for (int j = 0; j < n_live; j++) {
            Row& r = rows_[(size_t)live[(size_t)j]];
            const std::vector<uint8_t>& b = bits[(size_t)live[(size_t)j]];
            std::vector<int> feas;   // candidate-major pair indices
            for (size_t t = 0; t < b.size(); t++)
                if (b[t]) feas.push_back((int)t);
            if (feas.empty()) {
                r.fail_ti = ti;
                continue;
            }
            r.fail_ti = -1;
            int pick = feas[0];
            if (use_rng_ && r.rng.random() >= 1.0 - p_.eps)
                pick = feas[(size_t)r.rng.randrange((int)feas.size())];
            const int ci = pick / (int)ors.size(), oi = pick % (int)ors.size();
            const int x = r.cand[(size_t)ci][0], y = r.cand[(size_t)ci][1], z = r.cand[(size_t)ci][2];
            const int w = ors[(size_t)oi][0], d = ors[(size_t)oi][1], h = ors[(size_t)oi][2];
            r.placed.push_back({ti, x, y, z, w, d, h});
            r.aabb.push_back({x, y, z, x + w, y + d, z + h});
            std::vector<std::array<int32_t, 3>> n2;
            n2.reserve(r.cand.size() + 8);
            for (const auto& q : r.cand)
                if (!(x < q[0] && q[0] < x + w && y < q[1] && q[1] < y + d && z < q[2] &&
                      q[2] < z + h))
                    n2.push_back(q);
            for (int cx : {x, x + w})
                for (int cy : {y, y + d})
                    for (int cz : {z, z + h}) n2.push_back({cx, cy, cz});
            std::sort(n2.begin(), n2.end(),
                      [](const std::array<int32_t, 3>& a, const std::array<int32_t, 3>& b2) {
                          if (a[2] != b2[2]) return a[2] < b2[2];
                          if (a[1] != b2[1]) return a[1] < b2[1];
                          return a[0] < b2[0];
                      });
            n2.erase(std::unique(n2.begin(), n2.end()), n2.end());
            r.cand.swap(n2);
            r.max_ep = std::max(r.max_ep, (int)r.cand.size());
        }
    }
    // Stability of every row's packing in one launch.
    std::vector<int32_t> rows7, off(1, 0);
    for (const Row& r : rows_) {
        for (const auto& q : r.placed) rows7.insert(rows7.end(), q.begin(), q.end());
        off.push_back(off.back() + (int32_t)r.placed.size());
    }
    std::vector<double> score((size_t)K, 0.0);
    int rc = plp_fsr_eval(ctx_, rows7.empty() ? nullptr : rows7.data(), off.data(), K, p_.mu,
                          nullptr, score.data());
    if (rc != PLP_OK) return rc;
    for (int k = 0; k < K; k++) {
        Row& r = rows_[(size_t)k];
        long long vol = 0;
        for (const auto& q : r.placed) vol += 1LL * q[4] * q[5] * q[6];
        const int n = (int)r.placed.size();
        const double sc = score[(size_t)k];
        if (r.all.n < 0 || vol > r.all.vol || (vol == r.all.vol && n > r.all.n))
            r.all = Pack{vol, n, sc, r.placed};
        if ((!p_.use_theta || sc >= p_.theta) &&
            (r.acc.n < 0 || vol > r.acc.vol || (vol == r.acc.vol && n > r.acc.n)))
            r.acc = Pack{vol, n, sc, r.placed};
        r.passes++;
    }
    return PLP_OK;
}

void PlainRunner::result(int k, plp_row_result* out, std::vector<std::array<int32_t, 7>>* rows) const {
    const Row& r = rows_[(size_t)k];
    const bool have_acc = r.acc.n >= 0;
    const Pack& b = have_acc ? r.acc : r.all;
    std::memset(out, 0, sizeof *out);
    out->n_boxes = std::max(0, b.n);
    out->fallback = have_acc ? 0 : 1;
    out->n_passes = r.passes;
    out->volume = b.vol;
    out->score = b.score;
    out->seed = r.seed;
    out->work_pair_box = r.work;
    out->max_ep = r.max_ep;
    if (rows) *rows = b.rows;
}

}  // namespace plpi

extern "C" int plp_plain_run(plp_ctx* ctx, const plp_instance_desc* desc, uint64_t seed,
                             int32_t n_passes, const plp_params* p, plp_row_result* row,
                             int32_t* placements) {
    if (!ctx || !desc || !p || !row) return fail(PLP_E_ARG, "NULL argument");
    if (n_passes < 1) return fail(PLP_E_ARG, "n_passes must be >= 1");
    if (p->order != 0 && p->order != 1) return fail(PLP_E_ARG, "order must be 0 or 1");
    plp::Instance in;
    int rc = to_instance(desc, &in, kMaxPlainSide);
    if (rc != PLP_OK) return rc;
    PlainRunner runner(ctx, in, seed, 1, *p);
    for (int i = 0; i < n_passes; i++)
        if ((rc = runner.run_pass()) != PLP_OK) return rc;
    std::vector<std::array<int32_t, 7>> rows;
    runner.result(0, row, &rows);
    if (placements)
        for (size_t i = 0; i < rows.size(); i++)
            std::memcpy(placements + 7 * i, rows[i].data(), 7 * sizeof(int32_t));
    return PLP_OK;
}
