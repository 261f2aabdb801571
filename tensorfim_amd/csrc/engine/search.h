// Depth-first MFI search engines above the C ABI of libtfim.
//
//   plain  -- the reference baseline engine in GPU-only mode (RUN_MODE 1,
//             kernel_bitmap.cu:853-894): one count launch per DFS node, all
//             survivors counted, reference maximality and extension rule.
//   wave   -- the reference TensorFIM engine (runWaveEngine, kernel_bmma.cu:
//             896-910): the same enumeration, nodes popped in waves and each
//             wave counted in one fused launch.
//   mafia  -- batched MAFIA-style search (Burdick, Calimlim & Gehrke, ICDE
//             2001): tail counting, parent-equivalence pruning, head-union-tail
//             lookahead, HUT-subset pruning against found MFIs, maximality
//             decided by an exact all-survivor count.  Optional learned tail
//             ranking and lookahead gating (every decision confirmed by exact
//             counts).  Output is set-identical to the reference enumeration.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dataset.h"
#include "results.h"
#include "tfim.h"

namespace tfim {

struct SearchOptions {
    std::string engine = "mafia";
    int32_t wave_cap = 0;         // rows per wave; 0 = automatic
    int32_t workers = 1;          // plain engine: host workers (one stream each)
    int32_t host_threads = 0;     // mafia: host pool threads (0 = the CPUs this process may use less four, at most 12)
    int32_t flights = 4;          // mafia: waves in flight (2 to 4)
    bool learned = false;         // mafia: use the MLP ranker/gate
    std::string gate = "indep";   // mafia plain lookahead rule: indep | always
    std::string compl_rows = "auto";   // mafia: HUT/LEAF rows as device-reduced complement rows: auto | on | off
    std::string spec = "on";      // mafia: check nodes predicted to be maximal leaves at once: on | off
    std::string rows = "auto";    // waves of at most 64 dense ids as bit-mask rows (auto) or as id lists (lists)
    bool rank = false;            // mafia learned: reorder tails with the ranker head
    std::string model_path;       // mafia: load MLP weights (else trained online)
    std::string trace_path;       // mafia: write (features, outcome) records
    bool verbose = false;
    // plain: when set, receives every counted node's items (dense ids,
    // insertion order), for timing samples of the reference DFS elsewhere
    std::vector<std::vector<int32_t>> *node_log = nullptr;
};

struct SearchStats {
    int64_t nodes = 0;            // count rows issued (all kinds)
    int64_t waves = 0;            // count launches
    int64_t expand_rows = 0, hut_rows = 0, leaf_rows = 0, ref_rows = 0;
    int64_t spec_rows = 0;        // rows of nodes predicted to be maximal leaves
    int64_t spec_hits = 0;        // of those, found maximal by the same row
    int64_t spec_misses = 0;      // complement SPEC rows that had to be counted again as EXPAND
    int64_t hut_success = 0;      // HUT lookaheads that found HUT frequent
    int64_t hut_skipped = 0;      // lookaheads the gate declined
    int64_t pep_moves = 0;        // items moved into heads by PEP
    int64_t hutmfi_pruned = 0;    // subtrees cut because HUT is inside a found MFI
    int64_t leaf_subsumed = 0;    // leaves discarded by a found superset
    int64_t mfi_total = 0;        // MFIs of the survivor universe found
    double counted_bitops = 0;    // sum over rows of columns x transactions
    double count_seconds = 0;     // host time inside count calls (submit + wait)
    double submit_seconds = 0;    // of which: packing and issuing waves
    double process_seconds = 0;   // host time turning wave results into decisions and new nodes
    double build_seconds = 0;     // host time popping nodes into waves and issuing them (includes submit)
    double kernel_ms = 0;         // device time of count kernels (HIP events)
    int64_t lds_waves = 0;        // waves counted by the LDS-staged kernel
    int32_t wave_size = 0;        // mafia: rows per wave the wave-size tuner settled on (or the fixed cap)
    double learn_seconds = 0;     // host time in the learned part while mining
    double model_setup_seconds = 0;   // creating the learned part's network (before mining starts)
    int64_t mlp_train_steps = 0;
    int64_t model_lookaheads = 0;  // lookaheads run because the model (not the rule) asked
    int64_t model_lookahead_hits = 0;   // of those (exploration excluded), the ones whose HUT was frequent
    double gate_threshold = 0;          // learned gate's final self-calibrated threshold
    double explore_share = 0;           // and its final exploration share
    int64_t model_failures = 0;         // model calls that failed (the rule decided from then on)
};

class Policy;

// Runs the engine; returns MFIs (dense ids, insertion order) with frequencies.
// The MAFIA engine's search policy is `pol` when given (made beforehand with
// make_policy, so that setting up a learned policy's network is not part of
// the search), else one it makes itself.
int run_search(tfim_ctx *ctx, const Database &db, const SearchOptions &opt, std::vector<Mfi> &out,
               SearchStats &stats, std::string &err, Policy *pol = nullptr);

}  // namespace tfim
