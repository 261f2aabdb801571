#include "policy.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "search.h"

namespace tfim {

void PlainPolicy::decide(std::vector<Decision *> &ds, const Database &db) {
    const double log_thr = std::log(std::max(1e-300, db.thr_abs));
    for (Decision *d : ds) {
        d->order.resize(d->tail.size());
        std::iota(d->order.begin(), d->order.end(), 0);
        std::stable_sort(d->order.begin(), d->order.end(), [&](int32_t a, int32_t b) {
            if (d->tail_sup[a] != d->tail_sup[b]) return d->tail_sup[a] < d->tail_sup[b];
            return d->tail[a] < d->tail[b];
        });
        d->lookahead = d->tail.size() >= 2;
        if (d->lookahead && !always_) {
            const double ls = std::log(std::max(1, d->sup));
            double est = ls;
            for (int32_t t : d->tail_sup) est += std::log(std::max(1, t)) - ls;
            d->lookahead = est >= log_thr;
        }
    }
}

namespace {

constexpr int NF = TFIM_MLP_INPUTS;

// Learned policy.  One MLP, two heads, trained online on the GPU from what the
// exact engine observes:
//   head 0 (gate)   P(head+tail is frequent) for a node; label = the exact
//                   outcome of a lookahead that was taken (1/0);
//   head 1 (ranker) for each tail item t, the fraction of the child
//                   (head+t)'s tail that stays frequent when that child is
//                   counted; the tail is ordered by increasing prediction
//                   (items that cut the tail down go first, so later children's
//                   head-union-tail sets exclude them, MAFIA's rationale for
//                   ordering by support).
// Until the first training round the plain policy decides.  Every decision is
// confirmed by exact counts: a skipped lookahead only forgoes a pruning, and
// any tail order gives the same set enumeration.
class LearnedPolicy : public Policy {
   public:
    LearnedPolicy(tfim_ctx *ctx, const SearchOptions &opt) : ctx_(ctx), opt_(opt), plain_(false) {
        const auto t0 = std::chrono::steady_clock::now();
        if (tfim_mlp_create(ctx, 0x5EEDull, &mlp_) != TFIM_OK) {
            mlp_ = nullptr;
            create_error_ = std::string("the learned policy's network could not be created: ") + tfim_last_error(ctx);
        }
        if (mlp_ && !opt.model_path.empty()) {
            std::vector<float> p(TFIM_MLP_PARAMS);
            FILE *f = fopen(opt.model_path.c_str(), "rb");
            if (f && fread(p.data(), 4, p.size(), f) == p.size()) {
                tfim_mlp_set_params(mlp_, p.data());
                trained_ = true;
            }
            if (f) fclose(f);
        }
        if (!opt.trace_path.empty()) trace_ = fopen(opt.trace_path.c_str(), "wb");
        setup_seconds_ = secs_since(t0);
    }
    ~LearnedPolicy() override {
        if (mlp_) tfim_mlp_destroy(mlp_);
        if (trace_) fclose(trace_);
    }

    void decide(std::vector<Decision *> &ds, const Database &db) override {
        N_ = (double)std::max(1, db.trans_num_reduced);
        thr_ = std::max(1.0, db.thr_abs);
        freq_ = &db.freq_dense;
        // plain order and the independence-estimate rule, node by node (the
        // engine may have made it already: prepare)
        if (!std::all_of(ds.begin(), ds.end(), [](const Decision *d) { return d->prepared; }))
            run_parallel((int64_t)ds.size(), [&](int64_t q) {
                if (!ds[q]->prepared) prepare(*ds[q], db);
            });
        // Which nodes get feature rows: one gate row per node with a tail of
        // two or more (kept as a training record when its lookahead runs);
        // the model is asked only about nodes the rule declines with tails of
        // at most kMaxGateTail, plus the ranker rows of tails it reorders.
        // The draws are made here in node order; the rows are then filled in
        // parallel (each node writes only its own rows).
        feats_.clear();
        infer_.clear();
        node_row_.assign(ds.size(), -1);
        item_row0_.assign(ds.size(), -1);
        infer_row_.assign(ds.size(), -1);
        item_infer0_.assign(ds.size(), -1);
        const bool use_model = trained_ && mlp_;
        // gate-only: the model's verdicts are settled one wave later (resolve)
        const bool deferred = use_model && !opt_.rank;
        int64_t rows = 0, n_infer = 0;
        for (size_t q = 0; q < ds.size(); q++) {
            Decision *d = ds[q];
            if (d->head.empty() || d->tail.size() < 2) continue;
            // features only where they are used: every node during warm-up;
            // afterwards the model's candidates, a sample of the rule's own
            // lookaheads (their outcomes keep the labels balanced) and the
            // ranker's nodes
            // ranker rows: every ranked node's items; without --rank a
            // sample of nodes still records them, so head 1 keeps learning
            // (and small runs collect enough labels for training rounds)
            const bool items = d->tail.size() >= 3 && d->tail.size() <= kMaxRankTail &&
                               (opt_.rank || (rank_recorded_ < kRankRecordMax && rand01() < kRankRecord));
            if (items && !opt_.rank) rank_recorded_ += (int64_t)d->tail.size();
            bool need = !trained_ || items;
            // the model is asked about a share ask_p_ of the declined nodes
            // (all of them while its additions pay off)
            const bool ask = !d->lookahead && d->tail.size() <= kMaxGateTail && (ask_p_ >= 1.0f || rand01() < ask_p_);
            if (!need) need = d->lookahead ? rand01() < kRecordAccepted : ask;
            if (!need) continue;
            node_row_[q] = rows++;
            if (use_model && ask) infer_row_[q] = n_infer++;
            if (items) {
                item_row0_[q] = rows;
                rows += (int64_t)d->tail.size();
                if (use_model && opt_.rank) {
                    item_infer0_[q] = n_infer;
                    n_infer += (int64_t)d->tail.size();
                }
            }
        }
        feats_.resize((size_t)(rows * NF));
        infer_.resize((size_t)(n_infer * NF));
        // the nodes with rows; a few rows are filled here (handing them to
        // the host pool costs more than computing them)
        with_rows_.clear();
        for (size_t q = 0; q < ds.size(); q++)
            if (node_row_[q] >= 0) with_rows_.push_back((int64_t)q);
        const auto fill = [&](int64_t w) {
            const int64_t q = with_rows_[(size_t)w];
            const Decision &d = *ds[q];
            float *nf = &feats_[(size_t)(node_row_[q] * NF)];
            const NodeAgg a = node_features(d, nf);
            if (infer_row_[q] >= 0) std::copy_n(nf, NF, &infer_[(size_t)(infer_row_[q] * NF)]);
            if (item_row0_[q] >= 0) {
                // rank of each item in the plain (increasing support) order
                thread_local std::vector<int32_t> rank;
                rank.assign(d.tail.size(), 0);
                for (size_t j = 0; j < d.order.size(); j++) rank[(size_t)d.order[j]] = (int32_t)j;
                for (size_t i = 0; i < d.tail.size(); i++) {
                    float *f = &feats_[(size_t)((item_row0_[q] + (int64_t)i) * NF)];
                    item_features(d, a, rank[i], i, f);
                    if (item_infer0_[q] >= 0) std::copy_n(f, NF, &infer_[(size_t)((item_infer0_[q] + (int64_t)i) * NF)]);
                }
            }
        };
        if (rows <= kSerialRows) {
            for (size_t w = 0; w < with_rows_.size(); w++) fill((int64_t)w);
        } else {
            run_parallel((int64_t)with_rows_.size(), fill);
        }
        if (feats_.empty()) return;
        const int64_t n = (int64_t)(infer_.size() / NF);
        if (deferred) {
            if (n > 0) {
                if (tfim_mlp_forward_submit(mlp_, infer_.data(), (int32_t)n, 1) != TFIM_OK) {
                    model_failed("inference");   // no model: the rule decides from here on
                } else {
                    pend_feats_ = infer_;
                    pend_n_ = n;
                    forward_calls_++;
                    forward_rows_ += n;
                }
            }
        } else if (use_model && n > 0) {
            const auto tf0 = std::chrono::steady_clock::now();
            logits_.resize((size_t)n * 2);
            if (tfim_mlp_forward(mlp_, infer_.data(), (int32_t)n, logits_.data(), 1) != TFIM_OK) model_failed("inference");
            forward_seconds_ += secs_since(tf0);
            forward_calls_++;
            forward_rows_ += n;
        }
        for (size_t q = 0; q < ds.size(); q++) {
            Decision *d = ds[q];
            if (node_row_[q] < 0) continue;
            const bool rule = d->lookahead;
            if (!trained_) {
                // warm-up: the rule's lookaheads plus a sample of the
                // declined bounded tails, so that both outcomes get labelled;
                // the sample share calibrates itself like the exploration
                // share (it shrinks where hardly any sampled lookahead succeeds)
                const bool warm = !rule && d->tail.size() <= kMaxGateTail && rand01() < warm_p_;
                d->lookahead = rule || warm;
                if (warm) gate_src_ = kSrcWarm;
            } else if (!rule && d->tail.size() <= kMaxGateTail) {
                if (deferred && infer_row_[q] >= 0) {
                    d->pending = true;   // settled by resolve() from the queued inference
                    continue;
                }
                // the model adds the lookaheads it expects to succeed to the
                // rule's; a small exploration share keeps negative labels coming
                const bool model = infer_row_[q] >= 0 && sigmoid(logits_[infer_row_[q] * 2 + 0]) >= gate_p_;
                const bool explore = !model && rand01() < explore_p_;
                d->lookahead = model || explore;
                if (d->lookahead) model_gated_++;
                gate_src_ = model ? kSrcModel : explore ? kSrcExplore : kSrcRule;
            }
            if (trained_ && item_row0_[q] >= 0 && n > 0 && !deferred) {
                // items predicted to keep >= 90% of their child's tail go
                // last (they sit in every earlier child's head+tail)
                const int64_t base = item_infer0_[q];
                std::stable_partition(d->order.begin(), d->order.end(), [&](int32_t a) {
                    return sigmoid(logits_[(base + a) * 2 + 1]) < 0.9f;
                });
            }
            // the gate record is labelled only if the lookahead runs
            d->record = d->lookahead ? add_record(&feats_[node_row_[q] * NF], 0, gate_src_) : -1;
            gate_src_ = kSrcRule;
            if (item_row0_[q] >= 0) {
                d->item_records.resize(d->tail.size());
                for (size_t i = 0; i < d->tail.size(); i++)
                    d->item_records[i] = add_record(&feats_[(item_row0_[q] + i) * NF], 1);
            }
        }
    }

    void set_parallel(ParallelFor pf) override { par_ = std::move(pf); }
    void prepare(Decision &d, const Database &db) override {
        thread_local std::vector<Decision *> one(1);
        one[0] = &d;
        plain_.decide(one, db);
        d.prepared = true;
    }
    // the ranker reorders tails; the gate only adds lookaheads to the rule's
    bool keeps_order() const override { return !opt_.rank; }

    void resolve(std::vector<Decision *> &ds) override {
        // ds: the decisions decide() left pending, in the order it saw them;
        // their feature rows are the rows of the queued inference
        if (ds.empty()) return;
        bool ok = pend_n_ == (int64_t)ds.size();
        if (pend_n_ > 0) {
            const auto t0 = std::chrono::steady_clock::now();
            logits_.resize((size_t)pend_n_ * 2);
            if (tfim_mlp_forward_wait(mlp_, logits_.data()) != TFIM_OK) {
                model_failed("inference");
                ok = false;
            }
            forward_seconds_ += secs_since(t0);
        }
        for (size_t q = 0; q < ds.size(); q++) {
            Decision *d = ds[q];
            const bool model = ok && sigmoid(logits_[q * 2 + 0]) >= gate_p_;
            const bool explore = !model && rand01() < explore_p_;
            d->lookahead = model || explore;
            d->pending = false;
            if (d->lookahead) model_gated_++;
            d->record = d->lookahead && ok ? add_record(&pend_feats_[q * NF], 0, model ? kSrcModel : kSrcExplore) : -1;
        }
        pend_n_ = 0;
    }

    void observe(int32_t record, int outcome) override {
        if (record < 0 || outcome < 0) return;
        label(record, 0, (float)outcome);
    }
    void observe_value(int32_t record, float value) override {
        if (record < 0) return;
        label(record, 1, value);
    }

    void finish(SearchStats &st) override {
        st.mlp_train_steps = steps_;
        st.model_lookaheads = model_gated_;
        st.model_lookahead_hits = model_success_;
        st.gate_threshold = gate_p_;
        st.explore_share = explore_p_;
        st.model_failures = failures_;
        st.learn_seconds += extra_seconds_;   // training (decide() is timed by the engine)
        st.model_setup_seconds = setup_seconds_;
        if (opt_.verbose)
            fprintf(stderr, "learned policy: set-up %.4f s, training %.4f s (%lld steps), forward %.4f s in %lld calls (%lld rows)\n",
                    setup_seconds_, extra_seconds_, (long long)steps_, forward_seconds_, (long long)forward_calls_,
                    (long long)forward_rows_);
        if (mlp_ && !opt_.model_path.empty() && opt_.learned && trained_) {
            // keep what was learned for the next run on the same data
            std::vector<float> p(TFIM_MLP_PARAMS);
            if (tfim_mlp_get_params(mlp_, p.data()) == TFIM_OK) {
                FILE *f = fopen(opt_.model_path.c_str(), "wb");
                if (f) {
                    fwrite(p.data(), 4, p.size(), f);
                    fclose(f);
                }
            }
        }
    }

   private:
    // A failed model call leaves the rule in charge from then on; it is
    // reported on stderr (once) and counted in the run statistics.
    void model_failed(const char *what) {
        if (failures_++ == 0) fprintf(stderr, "learned policy: %s failed (%s); the rule decides from here on\n", what,
                                      tfim_last_error(ctx_));
        trained_ = false;
        disabled_ = true;
    }
    static double secs_since(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    }
    static float sigmoid(float x) { return 1.0f / (1.0f + std::exp(-x)); }
    float rand01() {
        uint64_t z = (rng_ += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((z >> 40) * (1.0 / 16777216.0));
    }

    struct NodeAgg {
        double s, sum_log_r, min_r, max_r, mean_r, frac_hi, mean_freq, log_min_s;
    };
    NodeAgg agg(const Decision &d) const {
        NodeAgg a{};
        a.s = std::max(1, d.sup);
        a.min_r = 1e9;
        a.max_r = 0;
        double min_s = 1e18;
        for (size_t i = 0; i < d.tail.size(); i++) {
            const double r = std::max(1, d.tail_sup[i]) / a.s;
            a.sum_log_r += std::log(r);
            a.min_r = std::min(a.min_r, r);
            a.max_r = std::max(a.max_r, r);
            a.mean_r += r;
            a.frac_hi += r >= 0.98;
            a.mean_freq += (double)(*freq_)[d.tail[i]] / N_;
            min_s = std::min(min_s, (double)d.tail_sup[i]);
        }
        const double k = (double)d.tail.size();
        a.mean_r /= k;
        a.frac_hi /= k;
        a.mean_freq /= k;
        a.log_min_s = std::log(std::max(1.0, min_s) / thr_);
        return a;
    }
    // The gate features of a node into f[0 .. NF); returns the tail
    // aggregates its item rows use.
    NodeAgg node_features(const Decision &d, float *f) const {
        const NodeAgg a = agg(d);
        const double margin = std::log(a.s / thr_);
        const float v[NF] = {(float)(a.s / N_), (float)(thr_ / a.s), (float)(d.head.size() / 16.0),
                             (float)(d.tail.size() / 16.0), (float)(10 * margin), (float)a.sum_log_r,
                             (float)(10 * (margin + a.sum_log_r)), (float)a.min_r, (float)a.mean_r, (float)a.max_r,
                             (float)a.frac_hi, (float)(d.depth / 16.0), (float)a.mean_freq, 0.0f};
        std::copy_n(v, NF, f);
        return a;
    }
    // The ranker features of tail item i (rank: its place in the plain
    // order) into f[0 .. NF).
    void item_features(const Decision &d, const NodeAgg &a, int rank, size_t i, float *f) const {
        const double margin = std::log(a.s / thr_);
        const double r = std::max(1, d.tail_sup[i]) / a.s;
        const float v[NF] = {(float)(a.s / N_), (float)(thr_ / a.s), (float)(d.head.size() / 16.0),
                             (float)(d.tail.size() / 16.0), (float)(10 * margin), (float)(10 * std::log(r)),
                             (float)(10 * std::log(std::max(1, d.tail_sup[i]) / thr_)),
                             (float)(rank / std::max<double>(1, d.tail.size() - 1)),
                             (float)((double)(*freq_)[d.tail[i]] / N_), (float)(r - a.mean_r),
                             (float)(10 * (margin + a.sum_log_r)), (float)(d.depth / 16.0),
                             (float)(10 * (margin + a.sum_log_r - std::log(r))), 1.0f};
        std::copy_n(v, NF, f);
    }
    void run_parallel(int64_t n, const std::function<void(int64_t)> &f) {
        if (par_) {
            par_(n, f);
        } else {
            for (int64_t i = 0; i < n; i++) f(i);
        }
    }

    // Records live in a ring of kRing slots; a label that arrives after its
    // slot was reused is dropped.
    // src: who asked for the lookahead (kSrcModel / kSrcExplore outcomes
    // calibrate the gate and the exploration share).
    int32_t add_record(const float *f, int head, uint8_t src = kSrcRule) {
        (void)head;
        const int64_t id = next_id_++;
        const size_t slot = (size_t)(id % kRing);
        std::copy_n(f, NF, &rec_x_[slot * NF]);
        rec_id_[slot] = id;
        rec_src_[slot] = src;
        rec_y_[slot * 2] = rec_y_[slot * 2 + 1] = 0.0f;
        rec_m_[slot * 2] = rec_m_[slot * 2 + 1] = 0.0f;
        return (int32_t)(id & 0x7FFFFFFF);
    }
    void label(int32_t id32, int head, float v) {
        const size_t slot = (size_t)(id32 % kRing);
        if ((rec_id_[slot] & 0x7FFFFFFF) != id32) return;
        rec_y_[slot * 2 + head] = v;
        rec_m_[slot * 2 + head] = 1.0f;
        if (head == 0 && rec_src_[slot] != kSrcRule) calibrate(rec_src_[slot], v > 0.5f);
        if (trace_) {
            fwrite(&rec_x_[slot * NF], 4, NF, trace_);
            const float yh[2] = {(float)head, v};
            fwrite(yh, 4, 2, trace_);
        }
        // each head keeps its most recent labelled records (half a batch each)
        const int64_t nh = n_head_[head]++;
        const size_t b = (size_t)(head * kHalf + nh % kHalf);
        if (bx_.empty()) {
            bx_.assign((size_t)kBatch * NF, 0.0f);
            by_.assign((size_t)kBatch * 2, 0.0f);
            bm_.assign((size_t)kBatch * 2, 0.0f);
        }
        std::copy_n(&rec_x_[slot * NF], NF, &bx_[b * NF]);
        by_[b * 2] = head == 0 ? v : 0.0f;
        by_[b * 2 + 1] = head == 1 ? v : 0.0f;
        bm_[b * 2] = head == 0 ? 1.0f : 0.0f;
        bm_[b * 2 + 1] = head == 1 ? 1.0f : 0.0f;
        n_labelled_++;
        // training rounds after 256, 1,024 and 4,096 labels (small runs get
        // a trained network), then every kTrainEvery new labels
        const int64_t due = last_train_ == 0 ? kFirstTrain : last_train_ < kTrainEvery / 2 ? last_train_ * 3
                                                                                          : kTrainEvery;
        if (n_labelled_ - last_train_ >= due) train();
    }
    // Self-calibration from the outcomes of the lookaheads the rule declined.
    // Model-added: over each window of kGateWindow, a success share below
    // kGateLow raises the threshold (those failures cost a row each and save
    // nothing), one above kGateHigh lowers it again, never below kGateP.
    // Exploration: where hardly any explored lookahead succeeds (fewer than
    // kExploreMinHits in a window) the share halves, down to kExploreMin;
    // where many do it doubles, up to kExplore.
    void calibrate(uint8_t src, bool success) {
        if (src == kSrcWarm) {
            warm_taken_++;
            warm_success_ += success;
            if (warm_taken_ < kGateWindow) return;
            if (warm_success_ < kExploreMinHits) warm_p_ = std::max(kExploreMin, warm_p_ * 0.5f);
            warm_taken_ = warm_success_ = 0;
            return;
        }
        if (src == kSrcModel) {
            gate_taken_++;
            gate_success_ += success;
            model_success_ += success;
            if (gate_taken_ < kGateWindow) return;
            const double prec = (double)gate_success_ / (double)gate_taken_;
            if (prec < kGateLow) {
                // far below: jump; at the ceiling already: ask the model about
                // fewer nodes (its features and inference cost host time)
                if (gate_p_ >= kGateMax) ask_p_ = std::max(kAskMin, ask_p_ * 0.5f);
                gate_p_ = std::min(kGateMax, gate_p_ + (prec < kGateLow / 4 ? 0.2f : 0.1f));
            } else if (prec > kGateHigh) {
                gate_p_ = std::max(kGateP, gate_p_ - 0.05f);
                ask_p_ = std::min(1.0f, ask_p_ * 2.0f);
            }
            gate_taken_ = gate_success_ = 0;
            return;
        }
        expl_taken_++;
        expl_success_ += success;
        if (expl_taken_ < kGateWindow) return;
        if (expl_success_ < kExploreMinHits) explore_p_ = std::max(kExploreMin, explore_p_ * 0.5f);
        else if (expl_success_ > kGateWindow / 4) explore_p_ = std::min(kExplore, explore_p_ * 2.0f);
        expl_taken_ = expl_success_ = 0;
    }

    void train() {
        if (!mlp_ || disabled_) return;
        const auto t0 = std::chrono::steady_clock::now();
        struct Acc {
            double &s;
            std::chrono::steady_clock::time_point t;
            ~Acc() { s += secs_since(t); }
        } acc{extra_seconds_, t0};
        last_train_ = n_labelled_;
        // batch = filled part of each head's half; snapshot so queued steps
        // read a stable copy
        tx_.clear();
        ty_.clear();
        tm_.clear();
        for (int h = 0; h < 2; h++) {
            const int64_t n = std::min<int64_t>(n_head_[h], kHalf);
            tx_.insert(tx_.end(), bx_.begin() + h * kHalf * NF, bx_.begin() + (h * kHalf + n) * NF);
            ty_.insert(ty_.end(), by_.begin() + h * kHalf * 2, by_.begin() + (h * kHalf + n) * 2);
            tm_.insert(tm_.end(), bm_.begin() + h * kHalf * 2, bm_.begin() + (h * kHalf + n) * 2);
        }
        const int32_t take = (int32_t)(tx_.size() / NF);
        if (take == 0) return;
        float loss = 0;
        // kEpochs steps on the batch, staged once and queued on the model's
        // stream (with --verbose the last one reports its loss and waits)
        if (tfim_mlp_train(mlp_, tx_.data(), ty_.data(), tm_.data(), take, 2e-3f, kEpochs,
                           opt_.verbose ? &loss : nullptr) != TFIM_OK) {
            model_failed("training");
            return;
        }
        steps_ += kEpochs;
        if (opt_.verbose) fprintf(stderr, "mlp train: %d records, loss %.4f\n", take, loss);
        trained_ = !disabled_;
    }

    static constexpr int64_t kTrainEvery = 8192;
    static constexpr int64_t kFirstTrain = 256;
    static constexpr int64_t kBatch = 8192;
    static constexpr int64_t kHalf = kBatch / 2;
    static constexpr int64_t kRing = 1 << 16;   // labels arrive within a few waves
    static constexpr int kEpochs = 4;
    static constexpr size_t kMaxRankTail = 128;   // longer tails keep the plain order
    static constexpr size_t kMaxGateTail = 64;     // the model gates tails of 2..64 items
    static constexpr float kGateP = 0.5f;          // model-added lookahead: P(HUT frequent) >= this (start, floor)
    static constexpr float kGateMax = 0.95f;       // ceiling of the self-calibrated threshold
    static constexpr int64_t kGateWindow = 64;     // model-added outcomes per calibration step
    static constexpr double kGateLow = 0.25, kGateHigh = 0.6;
    static constexpr float kExplore = 0.02f;       // share of rule-declined lookaheads run anyway (start, cap)
    static constexpr float kWarmSample = 0.25f;    // declined tails looked ahead before the first training
    static constexpr float kExploreMin = 0.0025f;  // floor of the self-calibrated exploration share
    static constexpr int64_t kExploreMinHits = 2;  // per window, below which exploration halves
    static constexpr uint8_t kSrcRule = 0, kSrcModel = 1, kSrcExplore = 2, kSrcWarm = 3;
    static constexpr float kAskMin = 1.0f / 16;     // floor of the share of declined nodes the model is asked about
    static constexpr float kRecordAccepted = 0.25f;  // share of the rule's lookaheads kept as training records
    static constexpr float kRankRecord = 0.125f;     // without --rank: share of nodes whose items are recorded
    static constexpr int64_t kRankRecordMax = 8192;  // ... until this many item records exist
    static constexpr int64_t kSerialRows = 256;      // feature rows of one wave filled without the host pool

    tfim_ctx *ctx_;
    const SearchOptions &opt_;
    tfim_mlp *mlp_ = nullptr;

   public:
    std::string create_error_;   // non-empty: no network (the run must not be labelled learned)

   private:
    FILE *trace_ = nullptr;
    PlainPolicy plain_;
    bool trained_ = false;
    bool disabled_ = false;   // a model call failed: no further training or inference
    int64_t failures_ = 0;
    double N_ = 1, thr_ = 1;
    const std::vector<int32_t> *freq_ = nullptr;
    ParallelFor par_;
    std::vector<int64_t> node_row_, item_row0_, infer_row_, item_infer0_, with_rows_;
    std::vector<float> infer_, pend_feats_;
    int64_t pend_n_ = 0;   // rows of the inference queued for pending decisions
    int64_t model_gated_ = 0, model_success_ = 0;
    float gate_p_ = kGateP;
    int64_t gate_taken_ = 0, gate_success_ = 0;
    float explore_p_ = kExplore;
    float warm_p_ = kWarmSample;     // warm-up sample share (self-calibrated)
    int64_t warm_taken_ = 0, warm_success_ = 0;
    float ask_p_ = kAskMin;          // share of declined nodes the model is asked about (grows while it pays)
    int64_t expl_taken_ = 0, expl_success_ = 0;
    uint8_t gate_src_ = kSrcRule;
    double extra_seconds_ = 0, forward_seconds_ = 0, setup_seconds_ = 0;
    int64_t forward_calls_ = 0, forward_rows_ = 0;
    std::vector<float> feats_, logits_;
    std::vector<float> rec_x_ = std::vector<float>((size_t)kRing * NF);
    std::vector<float> rec_y_ = std::vector<float>((size_t)kRing * 2), rec_m_ = std::vector<float>((size_t)kRing * 2);
    std::vector<int64_t> rec_id_ = std::vector<int64_t>((size_t)kRing, -1);
    std::vector<uint8_t> rec_src_ = std::vector<uint8_t>((size_t)kRing, 0);
    std::vector<float> bx_, by_, bm_, tx_, ty_, tm_;
    int64_t next_id_ = 0, n_labelled_ = 0, last_train_ = 0, steps_ = 0;
    int64_t rank_recorded_ = 0;   // item records made without --rank
    int64_t n_head_[2] = {0, 0};
    uint64_t rng_ = 0xC0FFEEull;
};

}  // namespace

std::unique_ptr<Policy> make_policy(tfim_ctx *ctx, const SearchOptions &opt, std::string &err) {
    if (opt.learned) {
        std::unique_ptr<LearnedPolicy> p(new LearnedPolicy(ctx, opt));
        if (!p->create_error_.empty()) {
            err = p->create_error_;
            return nullptr;
        }
        return std::unique_ptr<Policy>(p.release());
    }
    return std::unique_ptr<Policy>(new PlainPolicy(opt.gate == "always"));
}

}  // namespace tfim
