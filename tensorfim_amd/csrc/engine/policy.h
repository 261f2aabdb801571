// Search policy of the MAFIA engine: the order of each node's tail (which
// child gets the largest tail) and whether to spend a head-union-tail
// lookahead row on it.  Neither choice can change the output: the set
// enumeration is complete for any tail order, and a skipped lookahead only
// forgoes a pruning; every pruning that is taken is confirmed by exact counts.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "dataset.h"
#include "tfim.h"

namespace tfim {

struct SearchOptions;
struct SearchStats;

// A node after tail counting and parent-equivalence pruning.
struct Decision {
    std::vector<int32_t> head;      // dense ids
    int32_t sup = 0;                // sup(head)
    std::vector<int32_t> tail;      // frequent, non-equivalent tail items
    std::vector<int32_t> tail_sup;  // sup(head + tail[i])
    int32_t depth = 0;
    std::vector<int32_t> order;     // out: permutation of tail indices
    bool lookahead = false;         // out: count head+tail before expanding
    bool pending = false;           // out: lookahead undecided until resolve()
    int32_t record = -1;            // out: gate record id for observe()
    std::vector<int32_t> item_records;  // out: ranker record id per tail index
    std::vector<uint64_t> known;    // survivors e with sup(head + e) < thr (bitset), engine-owned
    int8_t leaf_sub = -1;           // engine: last child's leaf subsumed (1) or not (0) by the
    int64_t leaf_n = 0;             //   MFIs found before leaf_n; -1 = not yet checked
    bool prepared = false;          // prepare() has run on it (engine resets it per node)
};

class Policy {
   public:
    virtual ~Policy() {}
    virtual void decide(std::vector<Decision *> &ds, const Database &db) = 0;
    // Settles the decisions decide() left pending (lookahead, record).  The
    // engine calls it one wave later, so work queued on the device for them
    // (model inference) overlaps the next wave's count.
    virtual void resolve(std::vector<Decision *> &ds) { (void)ds; }
    // outcome: 1 = HUT frequent, 0 = not frequent, -1 = lookahead not taken
    virtual void observe(int32_t record, int outcome) { (void)record; (void)outcome; }
    // ranker label of a child: fraction of its tail that stayed frequent
    virtual void observe_value(int32_t record, float value) { (void)record; (void)value; }
    virtual void finish(SearchStats &st) { (void)st; }
    // true when decide() of one node depends on that node alone (no state,
    // no batch): the engine may then call it per node from several threads
    virtual bool decides_alone() const { return false; }
    // For a policy that decides a wave's nodes together: the part of
    // decide() that depends on one node alone, which the engine may call per
    // node from several threads before decide() (which then skips it for
    // nodes marked `prepared`).
    virtual void prepare(Decision &d, const Database &db) { (void)d; (void)db; }
    // true when decide() keeps the tail order prepare() made and takes back
    // no lookahead prepare() chose (the engine then checks leaves before it)
    virtual bool keeps_order() const { return false; }
    // A parallel loop the engine lends a batch policy: pf(n, f) calls f(i)
    // for every i < n on the engine's host threads and returns when all are
    // done; each call may touch only its own outputs.
    using ParallelFor = std::function<void(int64_t, const std::function<void(int64_t)> &)>;
    virtual void set_parallel(ParallelFor pf) { (void)pf; }
};

// MAFIA's plain policy: tail by increasing support (ties: ascending id).
// Lookahead rule ("indep", default): take it when the support head+tail would
// have if its items were independent given the head, sup(H) * prod_t
// sup(H+t)/sup(H), reaches the threshold; "always": whenever the tail holds
// two or more items.
class PlainPolicy : public Policy {
   public:
    explicit PlainPolicy(bool always = false) : always_(always) {}
    void decide(std::vector<Decision *> &ds, const Database &db) override;
    bool decides_alone() const override { return true; }

   private:
    bool always_;
};

// Returns nullptr, with the reason in err, when --learned is asked for and the
// network cannot be created (the run then fails rather than mining without it).
std::unique_ptr<Policy> make_policy(tfim_ctx *ctx, const SearchOptions &opt, std::string &err);

}  // namespace tfim
