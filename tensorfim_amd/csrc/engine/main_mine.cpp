// tfim_mine -- the reference CLI (kernel_bmma.cu:1179-1296) on MI355X.
//
//   tfim_mine <transactionFile> <relativeSupport> [--engine plain|wave|mafia]
//             [--replicate K] [--out PATH] [--wave-cap N] [--device D]
//             [--workers N] [--host-threads N] [--learned] [--rank] [--gate indep|always]
//             [--compl auto|on|off] [--spec on|off]
//             [--wave-kernel 0|1|2|3|4|5] [--mfma-format fp4|int8] [--split-k auto|off]
//             [--graphs 0|1] [--parse-threads N]
//             [--prep gpu|host] [--model PATH] [--trace PATH] [--verbose]
//
// Writes "<transactionFile>-<%f support>=Results.txt" (or PATH) in the
// reference layout and prints run statistics as "key: value" lines.  Exit
// status is 0 on success (the reference always returns 1, kernel_bmma.cu:
// 1295; callers here need a real status) and 1 on any failure, including an
// unreadable input (no Results file is written then, as in the reference).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dataset.h"
#include "results.h"
#include "policy.h"
#include "search.h"
#include "tfim.h"

using namespace tfim;

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
    SearchOptions opt;
    std::string out_path;
    int wave_kernel = 0;
    int graphs = -1;   // -1: library default
    int parse_threads = 0;   // 0: hardware threads for files of 4 MB and more
    std::string prep = "gpu";   // counts, reduction and bitmaps: gpu | host
    int replicate = 0, device = 0;
    std::string dump_nodes;        // plain engine: write every counted node (original ids), one per line
    std::string split_k = "auto";  // matrix-core waves: auto | off (one word range per tile)
    std::string mfma_format = "fp4";   // matrix-core waves: fp4 | int8 matrix cores
    std::vector<std::vector<int32_t>> node_log;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(1); }
            return argv[++i];
        };
        if (a == "--engine") opt.engine = next();
        else if (a == "--replicate") replicate = atoi(next().c_str());
        else if (a == "--out") out_path = next();
        else if (a == "--wave-cap") opt.wave_cap = atoi(next().c_str());
        else if (a == "--device") device = atoi(next().c_str());
        else if (a == "--learned") opt.learned = true;
        else if (a == "--gate") opt.gate = next();
        else if (a == "--compl") opt.compl_rows = next();
        else if (a == "--spec") opt.spec = next();
        else if (a == "--rows") opt.rows = next();
        else if (a == "--workers") opt.workers = atoi(next().c_str());
        else if (a == "--host-threads") opt.host_threads = atoi(next().c_str());
        else if (a == "--flights") opt.flights = atoi(next().c_str());
        else if (a == "--rank") opt.rank = true;
        else if (a == "--wave-kernel") wave_kernel = atoi(next().c_str());
        else if (a == "--graphs") graphs = atoi(next().c_str());
        else if (a == "--parse-threads") parse_threads = atoi(next().c_str());
        else if (a == "--prep") prep = next();
        else if (a == "--model") opt.model_path = next();
        else if (a == "--trace") opt.trace_path = next();
        else if (a == "--verbose") opt.verbose = true;
        else if (a == "--dump-nodes") dump_nodes = next();
        else if (a == "--split-k") split_k = next();
        else if (a == "--mfma-format") mfma_format = next();
        else pos.push_back(a);
    }
    if (pos.size() < 2) {
        fprintf(stderr, "usage: tfim_mine <file> <support> [--engine plain|wave|mafia] [--replicate K] [--out PATH]\n");
        return 1;
    }
    const std::string file = pos[0];
    const double support = atof(pos[1].c_str());
    const double t0 = now();

    tfim_ctx *ctx = nullptr;
    if (tfim_open(device, &ctx) != TFIM_OK) {
        fprintf(stderr, "tfim_open(%d) failed: no usable GPU\n", device);
        return 1;
    }
    if (graphs >= 0 && tfim_config(ctx, "graphs", graphs) != TFIM_OK) {
        fprintf(stderr, "%s\n", tfim_last_error(ctx));
        tfim_close(ctx);
        return 1;
    }
    if (wave_kernel && tfim_config(ctx, "wave_kernel", wave_kernel) != TFIM_OK) {
        fprintf(stderr, "%s\n", tfim_last_error(ctx));
        tfim_close(ctx);
        return 1;
    }
    if ((mfma_format != "fp4" && mfma_format != "int8") ||
        tfim_config(ctx, "mfma_format", mfma_format == "int8" ? 1 : 0) != TFIM_OK) {
        fprintf(stderr, "--mfma-format must be fp4 or int8\n");
        tfim_close(ctx);
        return 1;
    }
    if ((split_k != "auto" && split_k != "off") ||
        (split_k == "off" && tfim_config(ctx, "mfma_split", 1) != TFIM_OK)) {
        fprintf(stderr, "--split-k must be auto or off\n");
        tfim_close(ctx);
        return 1;
    }
    if (!dump_nodes.empty()) {
        if (opt.engine != "plain") {
            fprintf(stderr, "--dump-nodes needs --engine plain\n");
            tfim_close(ctx);
            return 1;
        }
        opt.node_log = &node_log;
    }
    if (opt.spec != "on" && opt.spec != "off") {
        fprintf(stderr, "--spec must be on or off\n");
        tfim_close(ctx);
        return 1;
    }
    if (opt.rows != "auto" && opt.rows != "lists") {
        fprintf(stderr, "--rows must be auto or lists\n");
        tfim_close(ctx);
        return 1;
    }
    if (opt.compl_rows != "auto" && opt.compl_rows != "on" && opt.compl_rows != "off") {
        fprintf(stderr, "--compl must be auto, on or off\n");
        tfim_close(ctx);
        return 1;
    }
    if (prep != "gpu" && prep != "host") {
        fprintf(stderr, "--prep must be gpu or host\n");
        tfim_close(ctx);
        return 1;
    }
    Database db;
    std::string err;
    if (!load_database(file, support, replicate, db, err, parse_threads, prep == "gpu" ? ctx : nullptr)) {
        fprintf(stderr, "%s\n", err.c_str());
        tfim_close(ctx);
        return 1;
    }
    const double t_up0 = now();
    int rc = TFIM_OK;
    if (db.d_count > 0 && db.base_bits > 0 && !db.on_device) {
        if (db.replicate > 1)
            rc = tfim_load_bitmaps_replicated(ctx, db.bitmap.data(), db.d_count, db.base_bits, db.words, db.n_surv,
                                              db.replicate);
        else
            rc = tfim_load_bitmaps(ctx, db.bitmap.data(), db.d_count, db.base_bits, db.words, db.n_surv);
    }
    if (rc != TFIM_OK) {
        fprintf(stderr, "upload failed: %s\n", tfim_last_error(ctx));
        tfim_close(ctx);
        return 1;
    }
    const double t_up1 = now();

    std::vector<Mfi> mfis;
    SearchStats st;
    // the MAFIA engine's policy, made before the search (a learned policy
    // creates its network here: device set-up, like loading the bitmaps)
    std::unique_ptr<Policy> pol;
    if (opt.engine == "mafia" && !db.duplicates && db.d_count > 0 && db.base_bits > 0) {
        pol = make_policy(ctx, opt, err);
        if (!pol) {
            fprintf(stderr, "search failed: %s\n", err.c_str());
            tfim_close(ctx);
            return 1;
        }
    }
    const double c0 = now();
    if (db.d_count > 0 && db.base_bits > 0) rc = run_search(ctx, db, opt, mfis, st, err, pol.get());
    const double c1 = now();
    if (rc != TFIM_OK) {
        fprintf(stderr, "search failed: %s\n", err.c_str());
        tfim_close(ctx);
        return 1;
    }
    RunInfo info;
    char name[256];
    int32_t cus = 0, clk = 0;
    int64_t hbm = 0;
    tfim_device_info(ctx, name, sizeof name, &cus, &hbm, &clk);
    info.device = name;
    info.compute_units = cus;
    info.hbm_bytes = hbm;
    info.engine = opt.engine + (opt.learned ? " (learned ranker/gate)" : "") +
                  (opt.engine == "plain" ? " (" + std::to_string(std::max(1, opt.workers)) + " host workers)" : "");
    info.preprocessing_time = db.preprocessing_seconds;
    info.computation_time = c1 - c0;
    char path[4096];
    if (!out_path.empty()) snprintf(path, sizeof path, "%s", out_path.c_str());
    else snprintf(path, sizeof path, "%s-%f=Results.txt", file.c_str(), support);
    info.wall_time = now() - t0;
    if (!write_results(path, file, db, mfis, info)) {
        fprintf(stderr, "cannot write %s\n", path);
        tfim_close(ctx);
        return 1;
    }
    if (!dump_nodes.empty()) {   // after the timed search
        FILE *f = fopen(dump_nodes.c_str(), "w");
        if (!f) {
            fprintf(stderr, "cannot write %s\n", dump_nodes.c_str());
            tfim_close(ctx);
            return 1;
        }
        for (const auto &v : node_log) {
            for (size_t i = 0; i < v.size(); i++) fprintf(f, i ? " %d" : "%d", db.dense_to_orig[v[i]]);
            fputc('\n', f);
        }
        fclose(f);
    }
    const double wall = now() - t0;
    printf("results: %s\n", path);
    printf("engine: %s\n", info.engine.c_str());
    printf("device: %s\n", name);
    printf("transNum: %d\ntransNumReduced: %d\nsurvivors: %d\nroots: %d\n", db.trans_num, db.trans_num_reduced,
           db.n_surv, db.n_roots);
    printf("MFIsNumber: %zu\n", mfis.size());
    printf("mfi_total: %lld\n", (long long)st.mfi_total);
    printf("parse_time(s): %g\npreprocessing_time(s): %g\nupload_time(s): %g\n", db.parse_seconds,
           db.preprocessing_seconds, t_up1 - t_up0);
    printf("computation_time(s): %g\nwall_time(s): %g\n", info.computation_time, wall);
    printf("count_time(s): %g\nkernel_ms: %g\n", st.count_seconds, st.kernel_ms);
    printf("submit_time(s): %g\nbuild_time(s): %g\nprocess_time(s): %g\n", st.submit_seconds, st.build_seconds,
           st.process_seconds);
    printf("rows: %lld\nwaves: %lld\nlds_waves: %lld\n", (long long)st.nodes, (long long)st.waves,
           (long long)st.lds_waves);
    int64_t n_graphs = 0, graph_launches = 0;
    tfim_wave_graph_stats(ctx, &n_graphs, &graph_launches);
    printf("graphs: %lld\ngraph_launches: %lld\n", (long long)n_graphs, (long long)graph_launches);
    int64_t l2_waves = 0, mfma_waves = 0, fused_waves = 0;
    tfim_wave_path_stats(ctx, &l2_waves, nullptr, &mfma_waves, &fused_waves);
    printf("l2_waves: %lld\nmfma_waves: %lld\nfused_waves: %lld\nwave_size: %d\n", (long long)l2_waves,
           (long long)mfma_waves, (long long)fused_waves, st.wave_size);
    printf("expand_rows: %lld\nhut_rows: %lld\nleaf_rows: %lld\nref_rows: %lld\n", (long long)st.expand_rows,
           (long long)st.hut_rows, (long long)st.leaf_rows, (long long)st.ref_rows);
    printf("spec_rows: %lld\nspec_hits: %lld\nspec_misses: %lld\n", (long long)st.spec_rows, (long long)st.spec_hits,
           (long long)st.spec_misses);
    printf("hut_success: %lld\nhut_skipped: %lld\npep_moves: %lld\nhutmfi_pruned: %lld\nleaf_subsumed: %lld\n",
           (long long)st.hut_success, (long long)st.hut_skipped, (long long)st.pep_moves,
           (long long)st.hutmfi_pruned, (long long)st.leaf_subsumed);
    printf("counted_bitops: %.6g\nlearn_time(s): %g\nmodel_setup_time(s): %g\nmlp_train_steps: %lld\n"
           "model_lookaheads: %lld\n",
           st.counted_bitops, st.learn_seconds, st.model_setup_seconds, (long long)st.mlp_train_steps,
           (long long)st.model_lookaheads);
    printf("model_lookahead_hits: %lld\ngate_threshold: %g\nexplore_share: %g\nmodel_failures: %lld\n",
           (long long)st.model_lookahead_hits, st.gate_threshold, st.explore_share, (long long)st.model_failures);
    tfim_close(ctx);
    return 0;
}
