#include "search.h"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <functional>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_set>

#include <immintrin.h>

#include "policy.h"

namespace tfim {

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One wave of count rows in CSR form, handed to tfim_count_nodes.
class WaveBatch {
   public:
    void clear() {
        item_ids_.clear();
        col_ids_.clear();
        item_off_.assign(1, 0);
        col_off_.assign(1, 0);
        compl_rows_.clear();
        compl_of_.clear();
        excl_.clear();
        compl_cols_ = 0;
        pm_.clear();
        cm_.clear();
        mask_off_.assign(1, 0);
        ord_kind_.clear();
        ord_at_.clear();
        ord_.clear();
        mask_ = false;
    }
    // A row as bit masks (databases of at most 64 dense ids; a wave holds
    // rows of one form): prefix ids pm, columns cm.  The library returns the
    // counts in ascending column order; result() gives them in the row's own
    // order: `lead` (a column of cm) first and then the others ascending, or
    // the order of the ids in `order` (n_order of them, exactly the ids of
    // cm), or ascending when neither is given.
    int add_mask(uint64_t pm, uint64_t cm, int lead = -1, const int32_t *order = nullptr, int n_order = 0) {
        mask_ = true;
        pm_.push_back(pm);
        cm_.push_back(cm);
        mask_off_.push_back(mask_off_.back() + __builtin_popcountll(cm));
        if (order) {
            ord_kind_.push_back(2);
            ord_at_.push_back((int32_t)ord_.size());
            ord_.insert(ord_.end(), order, order + n_order);
        } else if (lead >= 0) {
            ord_kind_.push_back(1);
            ord_at_.push_back(lead);
        } else {
            ord_kind_.push_back(0);
            ord_at_.push_back(0);
        }
        return rows() - 1;
    }
    int add(const int32_t *prefix, int k, const int32_t *cols, int nc) {
        item_ids_.insert(item_ids_.end(), prefix, prefix + k);
        item_off_.push_back((int32_t)item_ids_.size());
        col_ids_.insert(col_ids_.end(), cols, cols + nc);
        col_off_.push_back((int32_t)col_ids_.size());
        compl_of_.push_back(-1);
        return rows() - 1;
    }
    // A complement row (tfim_count_nodes_ex_submit): counted on the device
    // against every survivor whose bit in `excl` is clear; its result is
    // (support of the prefix, number of those survivors other than prefix[0]
    // reaching the threshold).  `live` = columns it counts, for statistics.
    int add_complement(const int32_t *prefix, int k, const uint64_t *excl, int words, int64_t live) {
        item_ids_.insert(item_ids_.end(), prefix, prefix + k);
        item_off_.push_back((int32_t)item_ids_.size());
        col_off_.push_back((int32_t)col_ids_.size());
        compl_of_.push_back((int32_t)compl_rows_.size());
        compl_rows_.push_back(rows() - 1);
        excl_.insert(excl_.end(), excl, excl + words);
        compl_cols_ += live;
        return rows() - 1;
    }
    int rows() const { return mask_ ? (int)pm_.size() : (int)item_off_.size() - 1; }
    int64_t cols_total() const { return mask_ ? (int64_t)mask_off_.back() : (int64_t)col_ids_.size() + compl_cols_; }
    // Launches the wave asynchronously (tfim_count_nodes_submit, or the _ex
    // form when it holds complement rows); wait() makes its counts
    // available.  A wave with no columns submits nothing.
    int submit(tfim_ctx *ctx, SearchStats &st, int64_t n_bits, int32_t min_count, std::string &err) {
        out_.resize(std::max<size_t>(1, mask_ ? (size_t)mask_off_.back() : col_ids_.size() + 2 * compl_rows_.size()));
        st.nodes += rows();
        ticket_ = -1;
        if (mask_ ? mask_off_.back() == 0 : col_ids_.empty() && compl_rows_.empty()) return TFIM_OK;   // nothing to count
        const double t0 = now();
        const int32_t one = 0;
        const int rc = mask_ ? tfim_count_masks_submit(ctx, pm_.data(), cm_.data(), rows(), TFIM_PATH_AUTO, &ticket_)
                       : compl_rows_.empty()
                           ? tfim_count_nodes_submit(ctx, item_ids_.data(), item_off_.data(), rows(), col_ids_.data(),
                                                     col_off_.data(), TFIM_PATH_AUTO, &ticket_)
                           : tfim_count_nodes_ex_submit(ctx, item_ids_.data(), item_off_.data(), rows(),
                                                        col_ids_.empty() ? &one : col_ids_.data(), col_off_.data(),
                                                        compl_rows_.data(), (int32_t)compl_rows_.size(),
                                                        excl_.data(), min_count, TFIM_PATH_AUTO, &ticket_);
        st.count_seconds += now() - t0;
        st.submit_seconds += now() - t0;
        if (rc != TFIM_OK) {
            err = std::string("tfim_count_nodes_submit: ") + tfim_last_error(ctx);
            return rc;
        }
        st.waves++;
        st.counted_bitops += (double)cols_total() * (double)n_bits;
        return TFIM_OK;
    }
    int wait(tfim_ctx *ctx, SearchStats &st, std::string &err) {
        const double t0 = now();
        const int rc = tfim_count_nodes_wait(ctx, ticket_, out_.data());
        st.count_seconds += now() - t0;
        ticket_ = -1;
        if (rc != TFIM_OK) {
            err = std::string("tfim_count_nodes_wait: ") + tfim_last_error(ctx);
            return rc;
        }
        if (mask_) order_masks();
        return rc;
    }
    // Counts of row r (CSR-aligned with its columns), or for a complement
    // row its two results.
    const int32_t *result(int r) const {
        if (mask_) return ord_kind_[r] ? perm_.data() + perm_at_[r] : out_.data() + mask_off_[r];
        const int32_t j = compl_of_[r];
        return j < 0 ? out_.data() + col_off_[r] : out_.data() + col_ids_.size() + 2 * (size_t)j;
    }
    int ncols(int r) const { return mask_ ? mask_off_[r + 1] - mask_off_[r] : col_off_[r + 1] - col_off_[r]; }

   private:
    // The counts of mask rows with an order of their own, in that order (the
    // position of column t among a row's ascending columns is the number of
    // its columns below t).
    void order_masks() {
        perm_.clear();
        perm_at_.assign(pm_.size(), 0);
        for (size_t r = 0; r < pm_.size(); r++) {
            if (!ord_kind_[r]) continue;
            const uint64_t cm = cm_[r];
            const int32_t *asc = out_.data() + mask_off_[r];
            const int n = mask_off_[r + 1] - mask_off_[r];
            perm_at_[r] = (int32_t)perm_.size();
            auto pos = [&](int32_t t) { return __builtin_popcountll(cm & ((1ull << t) - 1)); };
            if (ord_kind_[r] == 1) {
                const int lp = pos(ord_at_[r]);
                perm_.push_back(asc[lp]);
                for (int j = 0; j < n; j++)
                    if (j != lp) perm_.push_back(asc[j]);
            } else {
                const int32_t *o = ord_.data() + ord_at_[r];
                for (int j = 0; j < n; j++) perm_.push_back(asc[pos(o[j])]);
            }
        }
    }
    bool mask_ = false;
    std::vector<uint64_t> pm_, cm_;
    std::vector<int32_t> mask_off_{0}, ord_kind_, ord_at_, ord_, perm_, perm_at_;

    std::vector<int32_t> item_ids_, item_off_{0}, col_ids_, col_off_{0}, out_, compl_rows_, compl_of_;
    std::vector<uint64_t> excl_;
    int64_t compl_cols_ = 0;
    int64_t ticket_ = -1;
};

// Smallest integer count that passes the reference's `sup >= thrAbs`
// (double compare, kernel_bmma.cu:864): for integer s, s >= t <=> s >= ceil(t).
int32_t min_count_of(const Database &db) {
    if (!(db.thr_abs > 0)) return 0;
    return (int32_t)std::min<double>(std::ceil(db.thr_abs), (double)INT32_MAX);
}

bool frequent(const Database &db, int32_t s) { return !((double)s < db.thr_abs); }

// ---------------------------------------------------------------------------
// Reference enumeration (plain and wave engines).  A node is (items in
// insertion order, frequency, progress); counting covers every survivor;
// maximal iff #{e : sup(X+e) >= thr} == |X|; children X+i for i < progress
// with sup(X+i) >= thr, i not in X (kernel_bmma.cu:874-890).
struct RefNode {
    std::vector<int32_t> items;
    int32_t freq = 0;
    int32_t progress = 0;
};

void ref_expand(const Database &db, RefNode &nd, const int32_t *res, std::vector<RefNode> &stack,
                std::vector<Mfi> &out) {
    int32_t cnt = 0;
    for (int32_t e = 0; e < db.n_surv; e++) cnt += frequent(db, res[e]);
    if (cnt == (int32_t)nd.items.size()) {
        out.push_back({nd.items, nd.freq});
        return;
    }
    for (int32_t i = 0; i <= nd.progress - 1 && i < db.n_surv; i++) {
        if (!frequent(db, res[i])) continue;
        if (std::find(nd.items.begin(), nd.items.end(), i) != nd.items.end()) continue;
        RefNode c;
        c.items = nd.items;
        c.items.push_back(i);
        c.freq = res[i];
        c.progress = i;
        stack.push_back(std::move(c));
    }
}

std::vector<RefNode> ref_roots(const Database &db, bool only_item0) {
    std::vector<RefNode> st;
    for (int32_t i = 0; i < db.n_roots; i++) {
        const int32_t it = db.sorted_items[i];
        if (only_item0 && it != db.n_surv) continue;
        RefNode r;
        r.items = {it};
        r.freq = db.freq_dense[it];
        r.progress = it;
        st.push_back(std::move(r));
    }
    return st;
}

// Worker pool over one LIFO stack (kernel_bitmap.cu:853-894): each worker has
// its own context (stream + scratch) sharing the resident bitmaps, pops a
// node, counts it with one launch and pushes its children.
int run_plain(tfim_ctx *ctx, const Database &db, int workers, std::vector<Mfi> &out, SearchStats &st,
              std::string &err, std::vector<std::vector<int32_t>> *node_log) {
    std::vector<RefNode> stack = ref_roots(db, false);
    if (db.n_surv == 0) return TFIM_OK;
    workers = std::max(1, workers);
    std::vector<tfim_ctx *> ctxs(workers, nullptr);
    ctxs[0] = ctx;
    for (int w = 1; w < workers; w++)
        if (tfim_open_shared(ctx, &ctxs[w]) != TFIM_OK) {
            err = std::string("tfim_open_shared: ") + tfim_last_error(ctx);
            for (int v = 1; v < w; v++) tfim_close(ctxs[v]);
            return TFIM_E_HIP;
        }
    std::mutex mu;
    std::condition_variable cv;
    int active = 0, rc_all = TFIM_OK;
    std::vector<SearchStats> wst(workers);
    std::vector<std::vector<std::vector<int32_t>>> wlog(node_log ? workers : 0);
    auto worker = [&](int w) {
        std::vector<int32_t> res((size_t)db.n_surv + 1, 0);
        std::vector<RefNode> kids;
        std::vector<Mfi> found;
        for (;;) {
            RefNode nd;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !stack.empty() || active == 0 || rc_all != TFIM_OK; });
                if (stack.empty() || rc_all != TFIM_OK) break;
                nd = std::move(stack.back());
                stack.pop_back();
                active++;
            }
            const double t0 = now();
            const int rc = tfim_count_node(ctxs[w], nd.items.data(), (int32_t)nd.items.size(), res.data());
            wst[w].count_seconds += now() - t0;
            kids.clear();
            found.clear();
            if (rc == TFIM_OK) ref_expand(db, nd, res.data(), kids, found);
            {
                std::lock_guard<std::mutex> lk(mu);
                if (rc != TFIM_OK && rc_all == TFIM_OK) {
                    rc_all = rc;
                    err = std::string("tfim_count_node: ") + tfim_last_error(ctxs[w]);
                }
                for (auto &k : kids) stack.push_back(std::move(k));
                for (auto &m : found) out.push_back(std::move(m));
                active--;
            }
            cv.notify_all();
            wst[w].nodes++;
            if (node_log) wlog[w].push_back(nd.items);
        }
        cv.notify_all();
    };
    std::vector<std::thread> th;
    for (int w = 0; w < workers; w++) th.emplace_back(worker, w);
    for (auto &t : th) t.join();
    for (int w = 1; w < workers; w++) tfim_close(ctxs[w]);
    if (node_log)
        for (auto &l : wlog)
            for (auto &v : l) node_log->push_back(std::move(v));
    for (auto &x : wst) {
        st.nodes += x.nodes;
        st.count_seconds += x.count_seconds;
    }
    st.waves = st.nodes;
    st.ref_rows = st.nodes;
    st.counted_bitops = (double)st.nodes * (double)db.n_surv * (double)db.n_bits();
    return rc_all;
}

// Waves whose rows touch few distinct items fit the LDS-staged kernel when
// their total output stays within its accumulator budget; such waves are
// capped by outputs.  Otherwise a wave's column traffic is kept near 4 GiB of
// bitmap words so it amortises the launch while the DFS frontier stays small,
// with at least 128 rows (the matrix-core path counts 64-, 128- or 256-row
// tiles; a sweep on the clickstream workload put its best cap at 128).  The
// MAFIA engine tunes around this base (WaveSizeTuner).
constexpr int64_t kLdsOutBudget = 8192;   // = the library's LDS accumulator limit
constexpr int32_t kLdsItemBudget = 90;    // survivors (+ item 0) that fit its slots

bool lds_shape(const Database &db) { return db.d_count <= kLdsItemBudget; }

int32_t auto_wave_cap(const Database &db) {
    if (lds_shape(db)) return 2048;
    const double row_bytes = (double)db.n_bits() / 8.0 * std::max(1, db.n_surv);
    const double cap = 4.0 * (1ull << 30) / std::max(1.0, row_bytes);
    return (int32_t)std::min(2048.0, std::max(128.0, cap));
}

// Online choice of the MAFIA engine's wave size among base/2, base and
// 2*base rows: each arm runs a window of kWindow waves in turn (after the
// first kSkip waves, while the stack fills), then the arm with the lowest
// wall time per counted row runs, another one being tried every fourth
// window.  The time of a wave is the interval between the ends of two
// consecutive wave processings, so it covers the device and the host sides
// of the pipeline.  Wave size changes which rows are counted together and
// how much pruning later rows see, never the output.
class WaveSizeTuner {
   public:
    explicit WaveSizeTuner(int32_t base) {
        caps_[0] = std::max(1, base / 2);
        caps_[1] = base;
        caps_[2] = base * 2;
    }
    // the arm for the next wave to pack
    int next() {
        if (skip_ < kSkip) return 1;
        if (in_window_ >= kWindow) {
            in_window_ = 0;
            windows_++;
            if (windows_ < kArms) {
                arm_ = windows_;
            } else {
                int best = 0;
                for (int a = 1; a < kArms; a++)
                    if (cost_[a] < cost_[best]) best = a;
                arm_ = windows_ % 4 == 0 ? (best + 1 + (windows_ / 4) % (kArms - 1)) % kArms : best;
            }
        }
        in_window_++;
        return arm_;
    }
    int32_t cap(int arm) const { return caps_[arm]; }
    // a processed wave of `rows` rows packed with `arm`, `dt` seconds since
    // the previous wave's processing ended
    void record(int arm, int64_t rows, double dt) {
        if (skip_ < kSkip) {
            skip_++;
            return;
        }
        if (arm < 0 || rows <= 0) return;
        const double per = dt / (double)rows;
        cost_[arm] = n_[arm] == 0 ? per : 0.7 * cost_[arm] + 0.3 * per;
        n_[arm]++;
    }
    int32_t chosen() const {
        int best = 1;
        for (int a = 0; a < kArms; a++)
            if (n_[a] && cost_[a] < cost_[best]) best = a;
        return caps_[best];
    }

   private:
    static constexpr int kArms = 3, kWindow = 4, kSkip = 6;
    int32_t caps_[kArms];
    double cost_[kArms] = {1e30, 1e30, 1e30};
    int64_t n_[kArms] = {};
    int arm_ = 0, in_window_ = 0, windows_ = 0, skip_ = 0;
};

int64_t out_budget(const Database &db) { return lds_shape(db) ? kLdsOutBudget : INT64_MAX; }

int run_wave(tfim_ctx *ctx, const Database &db, int32_t cap, std::vector<Mfi> &out, SearchStats &st,
             std::string &err) {
    std::vector<RefNode> stack = ref_roots(db, false);
    std::vector<int32_t> all(db.n_surv);
    std::iota(all.begin(), all.end(), 0);
    const int64_t budget = out_budget(db);
    // two waves in flight: the next one is counted while this one expands
    struct Flight {
        std::vector<RefNode> wave;
        WaveBatch wb;
        bool live = false;
    } fl[2];
    auto launch = [&](Flight &f) -> int {
        f.wave.clear();
        int64_t outs = 0;
        while ((int)f.wave.size() < cap && !stack.empty() && (f.wave.empty() || outs + db.n_surv <= budget)) {
            f.wave.push_back(std::move(stack.back()));
            stack.pop_back();
            outs += db.n_surv;
        }
        f.live = !f.wave.empty();
        if (!f.live) return TFIM_OK;
        f.wb.clear();
        for (auto &nd : f.wave) f.wb.add(nd.items.data(), (int)nd.items.size(), all.data(), db.n_surv);
        st.ref_rows += (int64_t)f.wave.size();
        return f.wb.submit(ctx, st, db.n_bits(), 0, err);
    };
    int cur = 0;
    if (int rc = launch(fl[cur])) return rc;
    while (fl[cur].live) {
        if (int rc = launch(fl[1 - cur])) return rc;
        Flight &f = fl[cur];
        if (int rc = f.wb.wait(ctx, st, err)) return rc;
        for (size_t i = 0; i < f.wave.size(); i++) ref_expand(db, f.wave[i], f.wb.result((int)i), stack, out);
        f.live = false;
        cur = 1 - cur;
        if (!fl[cur].live)
            if (int rc = launch(fl[cur])) return rc;
    }
    return TFIM_OK;
}

// ---------------------------------------------------------------------------
// MAFIA-style engine.
//
// Node kinds, each one count row (prefix items, column list):
//   EXPAND (head, tail)          cols = tail      -> PEP, children or lookahead
//   HUT    (head' + tail'')      cols = {h0} + (U \ HUT)
//                                 -> sup(HUT) in col 0; HUT maximal iff the
//                                    other columns are all below threshold
//   LEAF   (H)                   cols = U \ H     -> H maximal iff all below
//   SPEC   (head, tail)          cols = U \ H     -> an EXPAND node predicted to
//                                 be a maximal leaf (its tail expected to die):
//                                 its maximality is checked at once; the tail's
//                                 counts are among the columns, so the node is
//                                 expanded from the same row when it lives on
//                                 (complement rows carry no per-column counts:
//                                 a SPEC row that fails is counted again as
//                                 EXPAND)
//   REF    (reference node)      cols = U         -> only in the subtree of a
//                                                    frequent item 0 (quirk A.3)
// U is the survivor set [0, F).  The reference reports the MFIs of U whose
// largest dense id is a root (strict-threshold roots, quirk A.2), plus its
// own enumeration under a frequent item-0 root; both are reproduced.
struct MNode {
    enum Kind : uint8_t { EXPAND, HUT, LEAF, REF, SPEC } kind = EXPAND;
    std::vector<int32_t> head;       // EXPAND/HUT/LEAF head; REF items
    int32_t sup = 0;                 // sup(head) (REF: node frequency)
    std::vector<int32_t> tail;       // EXPAND: candidates; HUT: ordered tail''
    std::vector<int32_t> tail_sup;   // HUT: sup(head'+t) for the fallback children
    int32_t progress = 0;            // REF
    int32_t depth = 0;
    int32_t decision = -1;           // HUT: policy record index for its outcome
    int32_t item_record = -1;        // EXPAND: ranker record of the item that made it
    // survivors e already known to have sup(head + e) < thr (bitset): they
    // stay infrequent for every superset of head, so HUT and LEAF rows skip
    // them (a pruning of columns only; decisions are unchanged)
    std::vector<uint64_t> known;
    // SPEC rows of wide databases: the tail is (*tail_ref)[tail_off ..),
    // copied into `tail` only if the node lives on (materialize_tail)
    std::shared_ptr<const std::vector<int32_t>> tail_ref;
    int32_t tail_off = 0;

    void materialize_tail() {
        if (!tail_ref) return;
        tail.assign(tail_ref->begin() + tail_off, tail_ref->end());
        tail_ref.reset();
    }
    // back to a fresh node, keeping the vectors' storage (node recycling)
    void reset() {
        kind = EXPAND;
        head.clear();
        tail.clear();
        tail_sup.clear();
        known.clear();
        tail_ref.reset();
        tail_off = 0;
        sup = progress = depth = 0;
        decision = item_record = -1;
    }
};

// Whether the AND of the k >= 2 slices sl[] is non-zero in some word of
// [w0, w1): portable form, one word at a time with an early exit per word.
bool any_and_scalar(const uint64_t *const *sl, size_t k, size_t w0, size_t w1) {
    for (size_t w = w0; w < w1; w++) {
        uint64_t acc = sl[0][w] & sl[1][w];
        for (size_t j = 2; j < k && acc; j++) acc &= sl[j][w];
        if (acc) return true;
    }
    return false;
}

// The same eight words at a time with 512-bit ANDs (the slices are read in
// rarest-first order, and a block of eight words stops once it is all zero).
__attribute__((target("avx512f"))) bool any_and_avx512(const uint64_t *const *sl, size_t k, size_t w0,
                                                        size_t w1) {
    size_t w = w0;
    for (; w + 8 <= w1; w += 8) {
        __m512i a = _mm512_and_si512(_mm512_loadu_si512(sl[0] + w), _mm512_loadu_si512(sl[1] + w));
        size_t j = 2;
        for (; j + 1 < k && _mm512_test_epi64_mask(a, a); j += 2)
            a = _mm512_and_si512(a, _mm512_and_si512(_mm512_loadu_si512(sl[j] + w), _mm512_loadu_si512(sl[j + 1] + w)));
        if (j < k) a = _mm512_and_si512(a, _mm512_loadu_si512(sl[j] + w));
        if (_mm512_test_epi64_mask(a, a)) return true;
    }
    return any_and_scalar(sl, k, w, w1);
}

using AnyAndFn = bool (*)(const uint64_t *const *, size_t, size_t, size_t);
AnyAndFn pick_any_and() { return __builtin_cpu_supports("avx512f") ? any_and_avx512 : any_and_scalar; }
const AnyAndFn any_and = pick_any_and();

// Found MFIs, sliced by item: bits_[i] has bit m set when MFI m contains
// item i.  X is inside some found MFI iff the AND of its items' slices is
// non-zero in some word.
class MfiIndex {
   public:
    explicit MfiIndex(int32_t n) : bits_((size_t)std::max(1, n)), cnt_((size_t)std::max(1, n), 0) {}
    bool subsumed(const std::vector<int32_t> &x) const { return subsumed_from(x, 0); }
    // Only the MFIs numbered >= n0 (and those sharing their first word).
    bool subsumed_from(const std::vector<int32_t> &x, int64_t n0) const {
        if (x.empty()) return n_ > n0;
        // rarest items first, so the AND of a word empties sooner
        thread_local std::vector<const uint64_t *> sl;
        thread_local std::vector<int32_t> ord;
        ord.assign(x.begin(), x.end());
        std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return cnt_[a] < cnt_[b]; });
        if (cnt_[ord[0]] == 0) return false;
        sl.clear();
        for (int32_t i : ord) sl.push_back(bits_[i].data());
        if (sl.size() == 1) sl.push_back(sl[0]);
        return any_and(sl.data(), sl.size(), (size_t)(n0 >> 6), words_);
    }
    void add(const std::vector<int32_t> &x) {
        grow();
        for (int32_t i : x) {
            bits_[i][n_ >> 6] |= 1ull << (n_ & 63);
            cnt_[i]++;
        }
        n_++;
    }
    // The sets keys[0 .. n) of items < 64 given as masks, numbered in that
    // order.  The bits of the MFIs that share an index word are gathered per
    // item first, so each item's word is written once per 64 MFIs.
    void add_masks(const uint64_t *keys, size_t n) {
        uint64_t col[64];
        for (size_t j = 0; j < n;) {
            grow();
            const size_t w = (size_t)(n_ >> 6);
            const size_t take = std::min(n - j, (size_t)(64 - (n_ & 63)));
            uint64_t any = 0;
            for (size_t q = 0; q < take; q++) any |= keys[j + q];
            for (uint64_t m = any; m; m &= m - 1) col[__builtin_ctzll(m)] = 0;
            for (size_t q = 0; q < take; q++) {
                const uint64_t bit = 1ull << ((n_ + (int64_t)q) & 63);
                for (uint64_t m = keys[j + q]; m; m &= m - 1) col[__builtin_ctzll(m)] |= bit;
            }
            for (uint64_t m = any; m; m &= m - 1) {
                const int i = __builtin_ctzll(m);
                bits_[(size_t)i][w] |= col[i];
                cnt_[(size_t)i] += __builtin_popcountll(col[i]);
            }
            n_ += (int64_t)take;
            j += take;
        }
    }
    int64_t size() const { return n_; }

   private:
    void grow() {
        if (n_ % 64 == 0) {
            words_++;
            for (auto &b : bits_) b.push_back(0ull);
        }
    }
    std::vector<std::vector<uint64_t>> bits_;
    std::vector<int64_t> cnt_;   // MFIs containing each item
    size_t words_ = 0;
    int64_t n_ = 0;
};

// Fork-join helper for the engine's independent per-row work: run(n, f)
// calls f(0..n-1) on the calling thread and the pool's threads and returns
// when all calls are done.  Between runs the threads spin on the job counter
// for a short while (a run per wave arrives every few hundred microseconds)
// before they sleep.
class HostPool {
   public:
    explicit HostPool(int threads) {
        for (int i = 1; i < threads; i++) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_.store(true);
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    template <class F>
    void run(int64_t n, F &&f) {
        if (th_.empty() || n < 2 * kGrain) {
            for (int64_t i = 0; i < n; i++) f(i);
            return;
        }
        fn_ = std::function<void(int64_t)>(std::forward<F>(f));
        n_ = n;
        next_.store(0);
        pending_.store((int)th_.size());
        {
            std::lock_guard<std::mutex> g(mu_);
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        work();
        while (pending_.load() != 0) std::this_thread::yield();
    }

   private:
    static constexpr int64_t kGrain = 8;
    void work() {
        for (int64_t i; (i = next_.fetch_add(kGrain)) < n_;)
            for (int64_t j = i, e = std::min(n_, i + kGrain); j < e; j++) fn_(j);
    }
    void loop() {
        uint64_t seen = 0;
        while (true) {
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load() == seen && std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(500))
                std::this_thread::yield();
            if (gen_.load() == seen) {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return gen_.load() != seen; });
            }
            seen = gen_.load();
            if (stop_.load()) return;
            work();
            pending_.fetch_sub(1);
        }
    }
    std::vector<std::thread> th_;
    std::function<void(int64_t)> fn_;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
    std::atomic<int> pending_{0};
    std::atomic<uint64_t> gen_{0};
    std::atomic<bool> stop_{false};
    std::mutex mu_;
    std::condition_variable cv_;
};

// One persistent helper thread running one job at a time: post(f) starts f,
// join() returns when it is done.  It spins briefly between jobs (one per
// wave) before it sleeps.
class AsyncWorker {
   public:
    AsyncWorker() : th_([this] { loop(); }) {}
    ~AsyncWorker() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_.store(true);
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        th_.join();
    }
    void post(std::function<void()> f) {
        job_ = std::move(f);
        busy_.store(true);
        {
            std::lock_guard<std::mutex> g(mu_);
            gen_.fetch_add(1);
        }
        cv_.notify_one();
    }
    void join() {
        while (busy_.load()) std::this_thread::yield();
    }

   private:
    void loop() {
        uint64_t seen = 0;
        while (true) {
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load() == seen && std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(2000))
                std::this_thread::yield();
            if (gen_.load() == seen) {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return gen_.load() != seen; });
            }
            seen = gen_.load();
            if (stop_.load()) return;
            job_();
            busy_.store(false);
        }
    }
    std::function<void()> job_;
    std::atomic<bool> busy_{false}, stop_{false};
    std::atomic<uint64_t> gen_{0};
    std::mutex mu_;
    std::condition_variable cv_;
    std::thread th_;
};

// Set of 64-bit keys (itemsets of at most 64 survivors as bitmasks): open
// addressing with linear probing, doubled when half full.  Key 0 (the empty
// set) is never inserted.
class KeySet64 {
   public:
    // true if the key was not there yet
    bool insert(uint64_t key) {
        if ((n_ + 1) * 2 > slots_.size()) grow();
        size_t i = mix(key) & (slots_.size() - 1);
        while (slots_[i]) {
            if (slots_[i] == key) return false;
            i = (i + 1) & (slots_.size() - 1);
        }
        slots_[i] = key;
        n_++;
        return true;
    }
    // Room for `more` keys beyond those inserted (at most half full then),
    // so that insert_shared never has to grow the table.
    void reserve(size_t more) {
        while ((n_ + more) * 2 > slots_.size()) grow();
    }
    // insert() for several threads at once, after reserve(): an empty slot
    // is claimed by compare-and-swap.  The caller reports how many keys it
    // inserted with added() once the threads are done.
    bool insert_shared(uint64_t key) {
        const size_t mask = slots_.size() - 1;
        for (size_t i = mix(key) & mask;; i = (i + 1) & mask) {
            uint64_t cur = __atomic_load_n(&slots_[i], __ATOMIC_ACQUIRE);
            if (cur == 0 &&
                __atomic_compare_exchange_n(&slots_[i], &cur, key, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE))
                return true;
            if (cur == key) return false;   // there, or just put there by another thread
        }
    }
    void added(size_t k) { n_ += k; }
    size_t size() const { return n_; }

   private:
    static uint64_t mix(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    void grow() {
        std::vector<uint64_t> old(std::max<size_t>(1024, slots_.size() * 2), 0ull);
        old.swap(slots_);
        n_ = 0;
        for (uint64_t k : old)
            if (k) insert(k);
    }
    std::vector<uint64_t> slots_;
    size_t n_ = 0;
};

struct VecHash {
    size_t operator()(const std::vector<int32_t> &v) const {
        uint64_t h = 1469598103934665603ull;
        for (int32_t x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
        return (size_t)h;
    }
};

class Mafia {
   public:
    Mafia(tfim_ctx *ctx, const Database &db, const SearchOptions &opt, Policy &pol, SearchStats &st)
        : ctx_(ctx), db_(db), opt_(opt), pol_(pol), st_(st), index_(db.n_surv),
          pool_(opt.host_threads > 0 ? opt.host_threads : default_pool_threads()) {}

    // The host pool's default size: the CPUs this process may run on, less
    // four (the main thread, the packing helper and the runtime's own
    // threads), at most 12 (measured on the dense replicated workload with a
    // 16-CPU share: 12 threads 0.111-0.114 s, 8 threads 0.120-0.124 s, 15
    // threads 0.117-0.123 s).
    static int default_pool_threads() {
        cpu_set_t set;
        int n = (int)std::thread::hardware_concurrency();
        if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
        return std::max(1, std::min(12, n - 4));
    }

    int run(std::vector<Mfi> &out, std::string &err) {
        const int32_t F = db_.n_surv;
        pol_.set_parallel([this](int64_t n, const std::function<void(int64_t)> &f) { pool_.run(n, f); });
        cap_ = opt_.wave_cap > 0 ? opt_.wave_cap : auto_wave_cap(db_);
        flights_ = std::max(2, std::min(kFlights, opt_.flights));
        if (opt_.wave_cap <= 0) sizer_.reset(new WaveSizeTuner(cap_));
        // maximality rows (HUT, LEAF) as complement rows reduced on the
        // device when the survivors are too many for listing them per row
        // to be cheap (the same bound as the LDS kernel's)
        compl_ = opt_.compl_rows == "on" || (opt_.compl_rows == "auto" && !lds_shape(db_));
        // rows as bit masks where every dense id fits one word
        mask_rows_ = !compl_ && db_.d_count <= 64 && opt_.rows != "lists";
        keys64_ = F <= 64;
        min_count_ = min_count_of(db_);
        spec_ = opt_.spec != "off";
        excl_words_ = (F + 63) / 64;
        // item-0 subtree: the reference enumeration, reproduced node for node
        if (db_.d_count > F && db_.is_root[F]) {
            for (auto &r : ref_roots(db_, true)) {
                MNode n;
                n.kind = MNode::REF;
                n.head = r.items;
                n.sup = r.freq;
                n.progress = r.progress;
                stack_.push_back(std::move(n));
            }
        }
        // root: empty head, tail = U; children are the single items
        if (F > 0) {
            Decision root;
            root.head = {};
            root.sup = db_.trans_num_reduced;
            root.depth = 0;
            for (int32_t i = 0; i < F; i++) {
                root.tail.push_back(i);
                root.tail_sup.push_back(db_.freq_dense[i]);
            }
            root.lookahead = false;
            std::vector<Decision *> ds{&root};
            pol_.decide(ds, db_);
            root.lookahead = false;   // the empty head has no HUT row
            push_children(root);
        }
        budget_ = out_budget(db_);
        // flights_ waves in flight (four by default): while this thread
        // processes wave k, the device counts k+1 and the next ones; the
        // helper thread packs and issues the newest meanwhile (the library
        // allows one submit beside one wait).  Those waves are taken off the
        // stack before k's children are on it; the answers do not depend on
        // the order of the search.
        int cur = 0;
        const int nf = flights_;
        if (int rc = launch(fl_[cur], false, err)) return rc;
        while (fl_[cur].live) {
            // the waves between this one and the newest are in flight (or
            // launched here if the stack ran dry before); the newest is
            // packed and issued by the helper
            for (int j = 1; j + 1 < nf; j++) {
                Flight &nj = fl_[(cur + j) % nf];
                if (!nj.live)
                    if (int rc = launch(nj, false, err)) return rc;
            }
            Flight &n1 = fl_[(cur + nf - 2) % nf];
            Flight &n2 = fl_[(cur + nf - 1) % nf];
            if (int rc = launch(n2, n1.live, err)) return rc;
            struct Join {
                AsyncWorker &w;
                bool on;
                ~Join() {
                    if (on) w.join();
                }
            } join{helper_, n2.live && n1.live};
            Flight &f = fl_[cur];
            if (int rc = f.wb.wait(ctx_, st_, err)) return rc;
            const double tp = now();
            if (int rc = process_wave(f, out)) return rc;
            const double te = now();
            st_.process_seconds += te - tp;
            if (sizer_ && last_end_ > 0) sizer_->record(f.cap_arm, (int64_t)f.wave.size(), te - last_end_);
            last_end_ = te;
            if (join.on) {
                helper_.join();
                join.on = false;
                merge_build(n2);
                if (n2.rc) {
                    err = n2.err;
                    return n2.rc;
                }
            }
            f.live = false;
            cur = (cur + 1) % nf;
            if (!fl_[cur].live)
                if (int rc = launch(fl_[cur], false, err)) return rc;
        }
        st_.wave_size = sizer_ ? sizer_->chosen() : cap_;
        // the MFIs kept as masks, as id lists (largest id first), in parallel
        // (chunks of 1,024 sets: one id list costs about one allocation)
        constexpr int64_t kChunk = 1024;
        const size_t o0 = out.size();
        const int64_t n_masks = (int64_t)out_masks_.size();
        out.resize(o0 + (size_t)n_masks);
        pool_.run((n_masks + kChunk - 1) / kChunk, [&](int64_t c) {
            for (int64_t i = c * kChunk; i < std::min(n_masks, (c + 1) * kChunk); i++) {
                Mfi &m = out[o0 + (size_t)i];
                const uint64_t key = out_masks_[(size_t)i].first;
                m.items.resize((size_t)__builtin_popcountll(key));
                size_t k = 0;
                for (uint64_t b = key; b;) {
                    const int32_t top = 63 - __builtin_clzll(b);
                    m.items[k++] = top;
                    b &= ~(1ull << top);
                }
                m.frequency = out_masks_[(size_t)i].second;
            }
        });
        out_masks_.clear();
        return TFIM_OK;
    }

   private:
    struct Flight {
        std::vector<MNode> wave;
        WaveBatch wb;
        bool live = false;
        // packing state of the helper thread: its own scratch, statistics
        // (added to the run's after it finishes) and status
        std::vector<uint8_t> mark;
        std::vector<int32_t> cols;
        std::vector<uint64_t> excl;
        int cap_arm = -1;   // wave-size tuner arm it was packed with
        SearchStats st;
        std::string err;
        int rc = 0;
    };

    // Pops the next wave off the stack (this thread) and packs and issues it,
    // on the helper thread when `async`.
    int launch(Flight &f, bool async, std::string &err) {
        if (stack_.empty()) settle_undecided();   // nothing else left to count
        const double t0 = now();
        if (free_.size() < kMaxFree)
            for (auto &n : f.wave) free_.push_back(std::move(n));
        f.wave.clear();
        f.cap_arm = sizer_ ? sizer_->next() : -1;
        const int32_t cap = sizer_ ? sizer_->cap(f.cap_arm) : cap_;
        int64_t outs = 0;
        while ((int)f.wave.size() < cap && !stack_.empty()) {
            const int64_t o = row_outputs(stack_.back());
            if (!f.wave.empty() && outs + o > budget_) break;
            outs += o;
            f.wave.push_back(std::move(stack_.back()));
            stack_.pop_back();
        }
        f.live = !f.wave.empty();
        st_.build_seconds += now() - t0;
        if (!f.live) return TFIM_OK;
        f.rc = 0;
        f.st = SearchStats();
        if (async) {
            helper_.post([this, &f] {
                const double t1 = now();
                f.rc = count_wave(f, f.err);
                f.st.build_seconds += now() - t1;
            });
            return TFIM_OK;
        }
        const double t1 = now();
        f.rc = count_wave(f, f.err);
        f.st.build_seconds += now() - t1;
        merge_build(f);
        if (f.rc) err = f.err;
        return f.rc;
    }

    // Adds the statistics of a wave's packing to the run's.
    void merge_build(Flight &f) {
        st_.nodes += f.st.nodes;
        st_.waves += f.st.waves;
        st_.expand_rows += f.st.expand_rows;
        st_.hut_rows += f.st.hut_rows;
        st_.leaf_rows += f.st.leaf_rows;
        st_.spec_rows += f.st.spec_rows;
        st_.ref_rows += f.st.ref_rows;
        st_.counted_bitops += f.st.counted_bitops;
        st_.count_seconds += f.st.count_seconds;
        st_.submit_seconds += f.st.submit_seconds;
        st_.build_seconds += f.st.build_seconds;
        f.st = SearchStats();
    }

    // Columns a row lists (what it costs of the wave's output budget): the
    // complement rows skip the survivors known infrequent with the head.
    int64_t row_outputs(const MNode &n) const {
        int64_t known = 0;
        if (n.kind != MNode::EXPAND && n.kind != MNode::REF)
            for (uint64_t w : n.known) known += __builtin_popcountll(w);
        switch (n.kind) {
            case MNode::EXPAND: return (int64_t)n.tail.size();
            case MNode::HUT: return 1 + db_.n_surv - (int64_t)(n.head.size() + n.tail.size()) - known;
            case MNode::LEAF:
            case MNode::SPEC: return db_.n_surv - (int64_t)n.head.size() - known;
            default: return db_.n_surv;
        }
    }

    // Survivors outside x that are not known infrequent with x's head; with
    // `lead`, x[0] first (the HUT row's own support).
    void complement(const std::vector<int32_t> &x, const std::vector<uint64_t> &known, std::vector<int32_t> &cols,
                    bool lead, std::vector<uint8_t> &mark) const {
        mark.assign(db_.n_surv, 0);
        for (int32_t i : x) mark[i] = 1;
        for (size_t w = 0; w < known.size(); w++)
            for (uint64_t b = known[w]; b; b &= b - 1) mark[(int32_t)(w * 64 + __builtin_ctzll(b))] = 1;
        cols.clear();
        if (lead) cols.push_back(x[0]);
        for (int32_t i = 0; i < db_.n_surv; i++)
            if (!mark[i]) cols.push_back(i);
    }

    // A SPEC row counted against explicit complement columns (narrow
    // databases): reads the tail's counts out of the complement's into
    // spec_tres_[r], counts the frequent columns outside the tail into
    // spec_other_[r] and adds the infrequent ones to the node's known set.
    // Touches only row r's data (runs on the host pool).
    void spec_prepare(int r, MNode &n, const int32_t *res) {
        const int W = (db_.n_surv + 63) / 64;
        thread_local std::vector<uint64_t> live;
        thread_local std::vector<int32_t> cum;
        live.assign((size_t)W, ~0ull);
        cum.assign((size_t)W + 1, 0);
        if (db_.n_surv % 64) live[(size_t)W - 1] = (1ull << (db_.n_surv % 64)) - 1;
        for (size_t w = 0; w < n.known.size() && w < (size_t)W; w++) live[w] &= ~n.known[w];
        for (int32_t i : n.head) live[(size_t)i >> 6] &= ~(1ull << (i & 63));
        for (int w = 0; w < W; w++) cum[(size_t)w + 1] = cum[(size_t)w] + __builtin_popcountll(live[(size_t)w]);
        auto &tres = spec_tres_[(size_t)r];
        tres.resize(n.tail.size());
        int freq_tail = 0;
        for (size_t j = 0; j < n.tail.size(); j++) {
            const int32_t t = n.tail[j];
            const uint64_t below = live[(size_t)t >> 6] & ((1ull << (t & 63)) - 1);
            tres[j] = res[cum[(size_t)t >> 6] + __builtin_popcountll(below)];
            freq_tail += ok_(tres[j]);
        }
        if (n.known.empty()) n.known.assign((size_t)W, 0ull);
        int freq_all = 0, j = 0;
        for (int w = 0; w < W; w++)
            for (uint64_t b = live[(size_t)w]; b; b &= b - 1, j++) {
                if (ok_(res[j])) freq_all++;
                else n.known[(size_t)w] |= b & (~b + 1);
            }
        spec_other_[(size_t)r] = freq_all - freq_tail;
    }

    // Exclusion bitmap of a complement row: x and the survivors known
    // infrequent with its head; returns how many survivors stay to count.
    int64_t exclusion(const std::vector<int32_t> &x, const std::vector<uint64_t> &known,
                      std::vector<uint64_t> &ex) const {
        ex.assign((size_t)excl_words_, 0ull);
        for (size_t w = 0; w < known.size() && w < ex.size(); w++) ex[w] = known[w];
        for (int32_t i : x) ex[(size_t)i >> 6] |= 1ull << (i & 63);
        int64_t n = 0;
        for (uint64_t w : ex) n += __builtin_popcountll(w);
        return db_.n_surv - n;
    }

    // Packs a wave and issues it; touches only the flight, the database and
    // the library (so it may run beside process_wave).
    int count_wave(Flight &f, std::string &err) {
        WaveBatch &wb_ = f.wb;
        wb_.clear();
        if (mask_rows_) return count_wave_masks(f, err);
        std::vector<int32_t> &cols = f.cols;
        for (auto &n : f.wave) {
            switch (n.kind) {
                case MNode::EXPAND:
                    wb_.add(n.head.data(), (int)n.head.size(), n.tail.data(), (int)n.tail.size());
                    f.st.expand_rows++;
                    break;
                case MNode::HUT: {
                    std::vector<int32_t> hut = n.head;
                    hut.insert(hut.end(), n.tail.begin(), n.tail.end());
                    if (compl_) {
                        const int64_t live = exclusion(hut, n.known, f.excl);
                        wb_.add_complement(hut.data(), (int)hut.size(), f.excl.data(), excl_words_, live + 1);
                    } else {
                        complement(hut, n.known, cols, true, f.mark);
                        wb_.add(hut.data(), (int)hut.size(), cols.data(), (int)cols.size());
                    }
                    f.st.hut_rows++;
                    break;
                }
                case MNode::LEAF:
                case MNode::SPEC:
                    if (compl_) {
                        const int64_t live = exclusion(n.head, n.known, f.excl);
                        wb_.add_complement(n.head.data(), (int)n.head.size(), f.excl.data(), excl_words_, live + 1);
                    } else {
                        complement(n.head, n.known, cols, false, f.mark);
                        wb_.add(n.head.data(), (int)n.head.size(), cols.data(), (int)cols.size());
                    }
                    if (n.kind == MNode::LEAF) f.st.leaf_rows++;
                    else f.st.spec_rows++;
                    break;
                case MNode::REF:
                    cols.resize(db_.n_surv);
                    std::iota(cols.begin(), cols.end(), 0);
                    wb_.add(n.head.data(), (int)n.head.size(), cols.data(), (int)cols.size());
                    f.st.ref_rows++;
                    break;
            }
        }
        return wb_.submit(ctx_, f.st, db_.n_bits(), min_count_, err);
    }

    // count_wave for databases of at most 64 dense ids: each row as two bit
    // masks (its columns the same as the list form's: a maximality row's are
    // the survivors outside its prefix and its known set, a HUT row's lead
    // column is its first item).
    int count_wave_masks(Flight &f, std::string &err) {
        WaveBatch &wb_ = f.wb;
        const uint64_t all = db_.n_surv >= 64 ? ~0ull : (1ull << db_.n_surv) - 1;
        auto bits = [](const std::vector<int32_t> &v) {
            uint64_t b = 0;
            for (int32_t i : v) b |= 1ull << i;
            return b;
        };
        for (auto &n : f.wave) {
            const uint64_t pm = bits(n.head);
            const uint64_t known = n.known.empty() ? 0ull : n.known[0];
            switch (n.kind) {
                case MNode::EXPAND:
                    wb_.add_mask(pm, bits(n.tail), -1, n.tail.data(), (int)n.tail.size());
                    f.st.expand_rows++;
                    break;
                case MNode::HUT: {
                    const uint64_t hut = pm | bits(n.tail);
                    const int32_t x0 = n.head.empty() ? n.tail[0] : n.head[0];
                    wb_.add_mask(hut, (all & ~(hut | known)) | (1ull << x0), x0);
                    f.st.hut_rows++;
                    break;
                }
                case MNode::LEAF:
                case MNode::SPEC:
                    wb_.add_mask(pm, all & ~(pm | known));
                    if (n.kind == MNode::LEAF) f.st.leaf_rows++;
                    else f.st.spec_rows++;
                    break;
                case MNode::REF:
                    wb_.add_mask(pm, all);
                    f.st.ref_rows++;
                    break;
            }
        }
        return wb_.submit(ctx_, f.st, db_.n_bits(), min_count_, err);
    }

    // An MFI found by wave row r: with at most 64 survivors the set is the
    // 64-bit key the row claimed on the host pool (none when another row
    // found the same set first), kept as is and indexed with the wave's
    // others after the row loop; its id list (largest id first) is made
    // once, after the search.  Else the id list is deduplicated and indexed
    // here.
    void emit(int r, std::vector<int32_t> &&items, int32_t sup, std::vector<Mfi> &out) {
        if (keys64_) {
            const uint64_t key = emit_key_[(size_t)r];
            if (!key) return;
            new_keys_.push_back(key);
            st_.mfi_total++;
            const int32_t top = 63 - __builtin_clzll(key);
            if (db_.is_root[top])
                out_masks_.push_back({key, __builtin_popcountll(key) == 1 ? db_.freq_dense[top] : sup});
            return;
        }
        std::sort(items.begin(), items.end(), std::greater<int32_t>());   // insertion order: root first
        if (!seen_.insert(items).second) return;
        index_.add(items);
        st_.mfi_total++;
        if (db_.is_root[items[0]]) {
            // a single-item MFI carries the root's occurrence count, as the
            // reference reports roots with freqPerItem (kernel_bmma.cu:1227)
            const int32_t f = items.size() == 1 ? db_.freq_dense[items[0]] : sup;
            out.push_back({std::move(items), f});
        }
    }

    // Subsumption of the leaf child (the last one, with no tail) of each
    // decision that will stack children, in parallel; push_children uses the
    // answers (and checks any MFI found after them).
    void check_leaves(std::vector<Decision *> &ds) {
        const int64_t n_now = index_.size();
        pool_.run((int64_t)ds.size(), [&](int64_t i) { check_leaf(*ds[i], n_now); });
    }
    // One decision's leaf check against the n_now MFIs found so far.
    void check_leaf(Decision &d, int64_t n_now) const {
        d.leaf_sub = -1;
        if (d.order.empty() || (d.lookahead && d.tail.size() >= 2)) return;   // a HUT row goes first
        thread_local std::vector<int32_t> leaf;
        leaf.assign(d.head.begin(), d.head.end());
        leaf.push_back(d.tail[d.order.back()]);
        d.leaf_sub = index_.subsumed(leaf) ? 1 : 0;
        d.leaf_n = n_now;
    }

    // The children of d (child i: head + t_i with tail t_{i+1..}), in stacking
    // order (last child first, so the first child, with the largest tail, is
    // counted first), into out[0 ..) -- nodes the caller provides, at least
    // d.order.size() of them.  Returns how many are kept: a leaf child inside
    // a found MFI is dropped and counted in *subsumed.  Reads only d, the MFI
    // index and the database (runs on the host pool).
    size_t build_children(const Decision &d, MNode *out, int64_t *subsumed) const {
        const size_t k = d.order.size();
        // largest sup(head + t) over each suffix of the ordered tail
        thread_local std::vector<int32_t> smax;
        smax.assign(k + 1, 0);
        for (size_t r = k; r-- > 0;) smax[r] = std::max(smax[r + 1], d.tail_sup[d.order[r]]);
        // the ordered tail, shared by the speculative children of a wide
        // database (their complement rows do not list it; a child that lives
        // on copies its suffix then), so a node with a long tail does not copy
        // a quadratic number of items
        std::shared_ptr<std::vector<int32_t>> ordered;
        size_t m = 0;
        for (size_t q = k; q-- > 0;) {
            MNode &c = out[m];
            c.reset();
            c.head.assign(d.head.begin(), d.head.end());
            c.head.push_back(d.tail[d.order[q]]);
            c.sup = d.tail_sup[d.order[q]];
            c.depth = d.depth + 1;
            c.known.assign(d.known.begin(), d.known.end());
            if (!d.item_records.empty()) c.item_record = d.item_records[d.order[q]];
            const bool has_tail = q + 1 < k;
            // Narrow databases: every child with a tail is a SPEC row.  It
            // lists every survivor outside head and known, so its tail's
            // counts are read out of it when it lives on, its maximality is
            // settled at once when the tail dies, and its infrequent columns
            // join the known set of its whole subtree, which shortens every
            // maximality row below it (on the dense replicated workload this
            // halves the columns counted against listing tails).  Wide
            // databases (complement rows, no per-column counts): speculate
            // that the child is a maximal leaf when even its best tail item
            // is expected to drop below the threshold if it were independent
            // of the added item: sup(H+t_q) * sup(H+t) / sup(H) < thr.
            if (spec_ && has_tail &&
                (!compl_ || (double)c.sup * (double)smax[q + 1] / (double)std::max(1, d.sup) < db_.thr_abs))
                c.kind = MNode::SPEC;
            if (c.kind == MNode::SPEC && compl_) {
                if (!ordered) {
                    ordered = std::make_shared<std::vector<int32_t>>(k);
                    for (size_t r = 0; r < k; r++) (*ordered)[r] = d.tail[d.order[r]];
                }
                c.tail_ref = ordered;
                c.tail_off = (int32_t)(q + 1);
            } else {
                for (size_t r = q + 1; r < k; r++) c.tail.push_back(d.tail[d.order[r]]);
            }
            if (!has_tail) {
                // no candidates left: the child is a leaf
                const bool sub = d.leaf_sub >= 0 ? d.leaf_sub == 1 || index_.subsumed_from(c.head, d.leaf_n)
                                                 : index_.subsumed(c.head);
                if (sub) {
                    ++*subsumed;
                    continue;
                }
                c.kind = MNode::LEAF;
            }
            m++;
        }
        return m;
    }

    // At least n nodes in kids_, spent ones refilled from the recycled nodes
    // (their vectors keep their storage).
    void arena(size_t n) {
        if (kids_.size() < n) kids_.resize(n);
        for (size_t i = 0; i < n && !free_.empty(); i++)
            if (kids_[i].head.capacity() == 0) {
                kids_[i] = std::move(free_.back());
                free_.pop_back();
            }
    }

    void push_children(const Decision &d) {
        arena(d.order.size());
        int64_t sub = 0;
        const size_t m = build_children(d, kids_.data(), &sub);
        st_.leaf_subsumed += sub;
        for (size_t i = 0; i < m; i++) stack_.push_back(std::move(kids_[i]));
    }

    // Outcomes of a row that settles itself, found on the host pool from its
    // own counts: it found an MFI; a HUT row is frequent.
    enum : uint8_t { kFound = 1, kHutFreq = 2 };

    int process_wave(Flight &f, std::vector<Mfi> &out) {
        settle_undecided();
        std::vector<MNode> &wave_ = f.wave;
        WaveBatch &wb_ = f.wb;
        const size_t rows = wave_.size();
        // SPEC rows counted against explicit complement columns: the tail's
        // counts read out of the complement's, how many frequent columns lie
        // outside the tail, and the infrequent ones (known for the subtree)
        spec_tres_.resize(rows);
        spec_other_.assign(rows, 0);
        const int64_t n_start = index_.size();
        pre_sub_.assign(rows, 0);
        row_flag_.assign(rows, 0);
        emit_key_.assign(rows, 0ull);
        if (keys64_) seen64_.reserve(rows);
        if (pre_dec_.size() < rows) pre_dec_.resize(rows);
        pre_kept_.assign(rows, 0);
        pre_pep_.assign(rows, 0);
        auto expands = [&](const MNode &n) { return n.kind == MNode::EXPAND || (n.kind == MNode::SPEC && !compl_); };
        auto tail_n = [&](int r) { return wave_[r].kind == MNode::SPEC ? (int)spec_tres_[r].size() : wb_.ncols(r); };
        auto key_of = [](const std::vector<int32_t> &v) {
            uint64_t b = 0;
            for (int32_t i : v) b |= 1ull << i;
            return b;
        };
        // a found set's key, claimed at once so that a set two rows find is
        // reported by one of them
        auto claim = [&](int64_t r, uint64_t key) { emit_key_[(size_t)r] = seen64_.insert_shared(key) ? key : 0ull; };
        // Per row, in parallel, what depends on the row alone:
        //  - a row that settles itself (HUT, LEAF, complement SPEC): whether
        //    it found an MFI (with at most 64 survivors its key is claimed
        //    here) and whether a HUT row is frequent;
        //  - SPEC preparation (the row's columns are the survivors outside
        //    head and known, ascending, so a column's index is a prefix
        //    popcount of that live mask);
        //  - a row that expands: its decision (frequent tail, PEP moves into
        //    the head, infrequent items into known) and its subsumption check
        //    against the MFIs found before the wave.  An EXPAND row checks
        //    head + every frequent tail item (the leaf head after PEP when no
        //    other item stays frequent, else its HUT); the checks against the
        //    MFIs this wave finds follow the row loop, so decisions equal a
        //    row-by-row pass.
        // What depends on the order of the rows (the stack, the policy's
        // observations, the MFI index) is done by the serial loop below.
        pool_.run((int64_t)rows, [&](int64_t r) {
            MNode &n = wave_[r];
            if (!expands(n)) {
                const int32_t *res = wb_.result((int)r);
                const int nc = wb_.ncols((int)r);
                uint8_t fl = 0;
                if (n.kind == MNode::SPEC) {
                    // complement row: res = {sup(head), frequent extensions};
                    // none means its whole tail dropped out
                    if (res[1] == 0) fl = kFound;
                } else if (n.kind == MNode::HUT || n.kind == MNode::LEAF) {
                    // explicit: res = {[sup(HUT),] counts of the complement};
                    // complement row: res = {sup(row), frequent extensions}
                    const bool hut = n.kind == MNode::HUT;
                    if (!hut || ok_(res[0])) {
                        bool maximal = !compl_ || res[1] == 0;
                        for (int j = hut ? 1 : 0; j < nc && maximal; j++) maximal = !ok_(res[j]);
                        fl = (uint8_t)((hut ? kHutFreq : 0) | (maximal ? kFound : 0));
                    }
                }
                row_flag_[(size_t)r] = fl;
                if ((fl & kFound) && keys64_)
                    claim(r, n.kind == MNode::HUT ? key_of(n.head) | key_of(n.tail) : key_of(n.head));
                return;
            }
            const bool spec = n.kind == MNode::SPEC;
            if (spec) spec_prepare((int)r, n, wb_.result((int)r));
            const int32_t *res = spec ? spec_tres_[r].data() : wb_.result((int)r);
            const int nc = spec ? (int)spec_tres_[r].size() : wb_.ncols((int)r);
            Decision &d = pre_dec_[r];
            d.head.assign(n.head.begin(), n.head.end());
            d.sup = n.sup;
            d.depth = n.depth;
            if (n.known.empty()) d.known.assign((size_t)(db_.n_surv + 63) / 64, 0ull);
            else d.known.assign(n.known.begin(), n.known.end());
            d.tail.clear();
            d.tail_sup.clear();
            d.order.clear();
            d.item_records.clear();
            d.lookahead = d.pending = false;
            d.record = -1;
            d.leaf_sub = -1;
            d.leaf_n = 0;
            d.prepared = false;
            int kept = 0, pep = 0;
            for (int j = 0; j < nc; j++) {
                const int32_t t = n.tail[j], sv = res[j];
                if (!ok_(sv)) {
                    d.known[(size_t)t >> 6] |= 1ull << (t & 63);
                    continue;
                }
                kept++;
                if (sv == n.sup) {   // PEP: every MFI above head contains t
                    d.head.push_back(t);
                    pep++;
                } else {
                    d.tail.push_back(t);
                    d.tail_sup.push_back(sv);
                }
            }
            pre_kept_[r] = kept;
            pre_pep_[r] = pep;
            // a SPEC row whose tail died is settled by its own counts: a
            // maximal leaf when no frequent column lies outside the
            // (equivalent) tail items, else nothing
            if (spec && d.tail.empty()) {
                if (spec_other_[r] == 0) {
                    row_flag_[(size_t)r] = kFound;
                    if (keys64_) claim(r, key_of(d.head));
                }
                return;
            }
            // head + frequent tail: the leaf head when the tail is empty,
            // else the node's head-union-tail
            thread_local std::vector<int32_t> q;
            q.assign(d.head.begin(), d.head.end());
            q.insert(q.end(), d.tail.begin(), d.tail.end());
            pre_sub_[r] = index_.subsumed(q);
        });
        // Rows that expand into a leaf or a decision are checked against the
        // MFIs found before the wave (pre_sub_) and against those this wave
        // finds: the row loop below reports the MFIs and notes these rows (a
        // leaf gets its place among the pending nodes), the checks then run
        // on the host pool against every MFI the wave found, and a last pass
        // in row order settles them.  (Pruning by any found MFI is sound, so
        // the checks need not see only the MFIs of earlier rows.)
        chk_.clear();
        new_keys_.clear();
        for (int r = 0; r < (int)rows; r++) {
            MNode &n = wave_[r];
            const uint8_t fl = row_flag_[(size_t)r];
            if (n.kind == MNode::SPEC && compl_) {
                if (fl & kFound) {
                    st_.spec_hits++;
                    if (n.item_record >= 0) pol_.observe_value(n.item_record, 0.0f);
                    emit(r, std::move(n.head), n.sup, out);
                } else {
                    st_.spec_misses++;
                    n.kind = MNode::EXPAND;   // lives on: count its tail
                    n.materialize_tail();
                    pending_.push_back(std::move(n));
                }
            } else if (expands(n)) {
                const bool spec = n.kind == MNode::SPEC;
                const int nc = tail_n(r);
                Decision &d = pre_dec_[r];
                if (n.item_record >= 0) pol_.observe_value(n.item_record, nc > 0 ? (float)pre_kept_[r] / nc : 0.0f);
                st_.pep_moves += pre_pep_[r];
                if (d.tail.empty() && spec) {
                    if (fl & kFound) {
                        st_.spec_hits++;
                        emit(r, std::move(d.head), d.sup, out);
                    }
                    continue;
                }
                if (d.tail.empty()) {
                    // a leaf unless a found MFI holds it: its place among the
                    // pending nodes, settled after the checks
                    chk_.push_back({r, (int64_t)pending_.size(), false});
                    pending_.push_back(take_node());
                    continue;
                }
                chk_.push_back({r, -1, false});   // a decision unless a found MFI holds head + tail
            } else if (n.kind == MNode::HUT) {
                const bool hut_freq = (fl & kHutFreq) != 0;
                pol_.observe(n.decision, hut_freq);
                if (hut_freq) {
                    st_.hut_success++;
                    if (fl & kFound) {
                        std::vector<int32_t> hut;
                        if (!keys64_) {
                            hut = n.head;
                            hut.insert(hut.end(), n.tail.begin(), n.tail.end());
                        }
                        emit(r, std::move(hut), wb_.result(r)[0], out);
                    }
                } else {
                    Decision d;
                    d.head = n.head;
                    d.sup = n.sup;
                    d.depth = n.depth;
                    d.tail = n.tail;
                    d.tail_sup = n.tail_sup;
                    d.known = std::move(n.known);
                    d.order.resize(d.tail.size());
                    std::iota(d.order.begin(), d.order.end(), 0);   // tail kept in decided order
                    hut_children_.push_back(std::move(d));
                }
            } else if (n.kind == MNode::LEAF) {
                if (fl & kFound) emit(r, std::move(n.head), n.sup, out);
            } else {   // REF: reference rule, own output path
                const int32_t *res = wb_.result(r);
                const int nc = wb_.ncols(r);
                int32_t cnt = 0;
                for (int j = 0; j < nc; j++) cnt += ok_(res[j]);
                if (cnt == (int32_t)n.head.size()) {
                    out.push_back({n.head, n.sup});
                } else {
                    for (int32_t i = 0; i <= n.progress - 1 && i < db_.n_surv; i++) {
                        if (!ok_(res[i])) continue;
                        if (std::find(n.head.begin(), n.head.end(), i) != n.head.end()) continue;
                        MNode c;
                        c.kind = MNode::REF;
                        c.head = n.head;
                        c.head.push_back(i);
                        c.sup = res[i];
                        c.progress = i;
                        pending_.push_back(std::move(c));
                    }
                }
            }
        }
        if (keys64_) {
            index_.add_masks(new_keys_.data(), new_keys_.size());
            seen64_.added(new_keys_.size());
        }
        // The checks against the MFIs found so far; a decision that survives
        // them is decided here too when the policy decides each node alone
        // (else the part of its decision that depends on it alone is made),
        // and its leaf child checked (no MFI is found before the children are
        // built, so the leaf check sees every MFI there is then; a policy
        // that decides the wave together must keep the tail order, and any
        // lookahead, made here).
        const bool alone = pol_.decides_alone();
        const bool early_leaf = alone || pol_.keeps_order();
        const int64_t n_now = index_.size();
        pool_.run((int64_t)chk_.size(), [&](int64_t i) {
            Check &c = chk_[i];
            Decision &d = pre_dec_[c.row];
            if (pre_sub_[c.row]) {
                c.sub = true;
            } else if (c.pend >= 0) {
                c.sub = index_.subsumed_from(d.head, n_start);
            } else {
                thread_local std::vector<int32_t> hut;
                hut.assign(d.head.begin(), d.head.end());
                hut.insert(hut.end(), d.tail.begin(), d.tail.end());
                c.sub = index_.subsumed_from(hut, n_start);
            }
            if (c.pend < 0 && !c.sub) {
                if (alone) {
                    thread_local std::vector<Decision *> one(1);
                    one[0] = &d;
                    pol_.decide(one, db_);
                } else {
                    pol_.prepare(d, db_);
                }
                if (early_leaf) check_leaf(d, n_now);
            }
        });
        pend_drop_.assign(pending_.size(), 0);
        decs_.clear();
        for (const Check &c : chk_) {
            Decision &d = pre_dec_[c.row];
            if (c.pend >= 0) {
                if (c.sub) {
                    st_.leaf_subsumed++;
                    pend_drop_[(size_t)c.pend] = 1;
                } else {
                    MNode &leaf = pending_[(size_t)c.pend];
                    leaf.reset();
                    leaf.kind = MNode::LEAF;
                    leaf.head.swap(d.head);
                    leaf.sup = d.sup;
                    leaf.known.swap(d.known);
                }
            } else if (c.sub) {
                st_.hutmfi_pruned++;
            } else {
                decs_.push_back(&d);
            }
        }
        // a policy that decides the wave's nodes together (order tails,
        // choose lookaheads) runs here
        if (!alone && !decs_.empty()) {
            const double t0 = now();
            pol_.decide(decs_, db_);
            st_.learn_seconds += now() - t0;
        }
        // push in reverse so the first decision's work is on top
        for (size_t i = 0; i < pending_.size(); i++)
            if (!pend_drop_[i]) stack_.push_back(std::move(pending_[i]));
        pending_.clear();
        {
            // the leaf checks not made above, pending decisions too: their
            // leaf check then needs no pass of its own when they are settled
            // (push_children checks the MFIs found in between; a lookahead
            // taken after all leaves it unused)
            std::vector<Decision *> leaves;
            for (auto &d : hut_children_) leaves.push_back(&d);
            if (!early_leaf) leaves.insert(leaves.end(), decs_.begin(), decs_.end());
            if (!leaves.empty()) check_leaves(leaves);
        }
        for (auto &d : hut_children_) push_children(d);
        hut_children_.clear();
        // the decisions' work, stacked in reverse order (the first decision's
        // on top): their children are built on the host pool into kids_,
        // then stacked here in that order
        work_.clear();
        size_t total = 0;
        for (size_t q = decs_.size(); q-- > 0;) {
            Decision &d = *decs_[q];
            if (d.pending) {
                undecided_.push_back(std::move(d));   // settled after the next wave
                continue;
            }
            const bool hut = d.lookahead && d.tail.size() >= 2;
            work_.push_back({q, hut, total, 0, 0});
            if (!hut) total += d.order.size();
        }
        arena(total);
        kid_work_.clear();
        for (size_t i = 0; i < work_.size(); i++)
            if (!work_[i].hut) kid_work_.push_back(i);
        pool_.run((int64_t)kid_work_.size(), [&](int64_t i) {
            Work &w = work_[kid_work_[i]];
            w.n = build_children(*decs_[w.dec], kids_.data() + w.off, &w.subsumed);
        });
        for (const Work &w : work_) {
            Decision &d = *decs_[w.dec];
            if (w.hut) {
                apply(d);   // its head-union-tail row
                continue;
            }
            if (d.tail.size() >= 2) st_.hut_skipped++;
            pol_.observe(d.record, -1);
            st_.leaf_subsumed += w.subsumed;
            for (size_t i = 0; i < w.n; i++) stack_.push_back(std::move(kids_[w.off + i]));
        }
        return TFIM_OK;
    }

    // Stacks a decided node's work: its head-union-tail row, or its children.
    void apply(Decision &d) {
        if (d.lookahead && d.tail.size() >= 2) {
            MNode h;
            h.kind = MNode::HUT;
            h.head = d.head;
            h.sup = d.sup;
            h.depth = d.depth;
            for (int32_t o : d.order) {
                h.tail.push_back(d.tail[o]);
                h.tail_sup.push_back(d.tail_sup[o]);
            }
            h.decision = d.record;
            h.known = d.known;
            stack_.push_back(std::move(h));
        } else {
            if (d.tail.size() >= 2) st_.hut_skipped++;
            pol_.observe(d.record, -1);
            push_children(d);
        }
    }

    // Settles the decisions the policy left pending one wave earlier (kept in
    // reverse stacking order) and stacks their work.
    void settle_undecided() {
        if (undecided_.empty()) return;
        std::vector<Decision *> ptrs;
        for (size_t q = undecided_.size(); q-- > 0;) ptrs.push_back(&undecided_[q]);
        const double t0 = now();
        pol_.resolve(ptrs);
        st_.learn_seconds += now() - t0;
        for (auto &d : undecided_) apply(d);
        undecided_.clear();
    }

    tfim_ctx *ctx_;
    const Database &db_;
    const SearchOptions &opt_;
    Policy &pol_;
    SearchStats &st_;
    MfiIndex index_;
    HostPool pool_;
    std::vector<uint8_t> pre_sub_;
    std::vector<Decision> pre_dec_;             // per wave row: its decision, built on the host pool
    std::vector<int32_t> pre_kept_, pre_pep_;   // per wave row: frequent tail items, PEP moves
    std::vector<std::vector<int32_t>> spec_tres_;   // SPEC rows (explicit): the tail's counts
    std::vector<int32_t> spec_other_;
    std::unordered_set<std::vector<int32_t>, VecHash> seen_;
    KeySet64 seen64_;
    std::vector<std::pair<uint64_t, int32_t>> out_masks_;   // reported MFIs of <= 64 survivors (mask, frequency)
    std::vector<uint8_t> row_flag_;   // per wave row: kFound, kHutFreq
    std::vector<uint64_t> emit_key_;  // per wave row: key of the MFI it found (<= 64 survivors), 0 = none
    std::vector<uint64_t> new_keys_;  // keys the wave found, in row order
    std::vector<MNode> stack_, pending_;
    std::vector<Decision> hut_children_, undecided_;
    std::vector<Decision *> decs_;   // decisions of the wave being processed (in pre_dec_)
    // a decision's stacking work: its HUT row, or its children (n kept of
    // those built at kids_[off ..), `subsumed` leaves dropped)
    struct Work {
        size_t dec;
        bool hut;
        size_t off, n;
        int64_t subsumed;
    };
    std::vector<Work> work_;
    // a wave row that expands, to check against the found MFIs: a leaf (its
    // place `pend` among the pending nodes) or a decision (pend < 0)
    struct Check {
        int row;
        int64_t pend;
        bool sub;
    };
    std::vector<Check> chk_;
    std::vector<uint8_t> pend_drop_;   // pending nodes dropped (leaves inside a found MFI)
    std::vector<size_t> kid_work_;
    std::vector<MNode> kids_;   // children under construction
    static constexpr int kFlights = 4;   // at most; flights_ in use
    // spent nodes kept for reuse, at most: a wave's rows and children need a
    // few thousand; more would only be freed at the end of the search
    static constexpr size_t kMaxFree = 4096;
    Flight fl_[kFlights];
    int flights_ = 4;
    AsyncWorker helper_;
    int32_t cap_ = 256;
    int64_t budget_ = INT64_MAX;
    // frequent(db, s) as an integer compare: s >= ceil(thr_abs)
    bool ok_(int32_t s) const { return s >= min_count_; }
    // spent nodes whose vectors are reused for new ones
    std::vector<MNode> free_;
    MNode take_node() {
        if (free_.empty()) return MNode();
        MNode n = std::move(free_.back());
        free_.pop_back();
        n.reset();
        return n;
    }
    std::unique_ptr<WaveSizeTuner> sizer_;   // wave size (unless --wave-cap is given)
    double last_end_ = 0;
    bool compl_ = false;         // HUT, LEAF and SPEC rows are complement rows
    bool mask_rows_ = false;     // rows counted as bit masks (at most 64 dense ids)
    bool keys64_ = false;        // found MFIs kept as 64-bit keys (at most 64 survivors)
    bool spec_ = true;           // predicted maximal leaves are checked at once (SPEC rows)
    int32_t min_count_ = 0;
    int excl_words_ = 0;
};

}  // namespace

int run_search(tfim_ctx *ctx, const Database &db, const SearchOptions &opt, std::vector<Mfi> &out,
               SearchStats &st, std::string &err, Policy *given) {
    if (opt.engine == "plain") return run_plain(ctx, db, opt.workers, out, st, err, opt.node_log);
    if (opt.engine == "wave") {
        const int32_t cap = opt.wave_cap > 0 ? opt.wave_cap : auto_wave_cap(db);
        const int rc = run_wave(ctx, db, cap, out, st, err);
        tfim_wave_stats(ctx, nullptr, &st.lds_waves, nullptr, &st.kernel_ms);
        return rc;
    }
    if (opt.engine == "mafia") {
        if (db.duplicates) {
            // Repeated ids inside one transaction make occurrence counts differ
            // from bitmap supports, which the reference's own rules mix; its
            // enumeration is reproduced as is in that case.
            const int32_t cap = opt.wave_cap > 0 ? opt.wave_cap : auto_wave_cap(db);
            return run_wave(ctx, db, cap, out, st, err);
        }
        std::unique_ptr<Policy> own;
        if (!given) {
            own = make_policy(ctx, opt, err);
            if (!own) return TFIM_E_STATE;
        }
        Policy &pol = given ? *given : *own;
        Mafia m(ctx, db, opt, pol, st);
        const int rc = m.run(out, err);
        pol.finish(st);
        tfim_wave_stats(ctx, nullptr, &st.lds_waves, nullptr, &st.kernel_ms);
        return rc;
    }
    err = "unknown engine " + opt.engine;
    return TFIM_E_INVALID;
}

}  // namespace tfim
