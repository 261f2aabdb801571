#include "results.h"

#include <algorithm>
#include <charconv>
#include <string>
#include <cstdio>
#include <cstring>
#include <thread>

namespace tfim {

static std::string cpu_model() {
    std::string name = "Unknown CPU";
    FILE *f = fopen("/proc/cpuinfo", "r");
    if (!f) return name;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, "model name", 10) == 0) {
            const char *c = strchr(line, ':');
            if (c) {
                std::string s(c + 1);
                while (!s.empty() && (s.front() == ' ' || s.front() == '\t')) s.erase(0, 1);
                while (!s.empty() && (s.back() == '\n' || s.back() == ' ')) s.pop_back();
                name = s;
            }
            break;
        }
    }
    fclose(f);
    return name;
}

bool write_results(const std::string &path, const std::string &trans_set_file, const Database &db,
                   std::vector<Mfi> &mfis, const RunInfo &info) {
    std::sort(mfis.begin(), mfis.end(), [](const Mfi &a, const Mfi &b) {
        if (a.frequency != b.frequency) return a.frequency > b.frequency;
        return a.items < b.items;
    });
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return false;
    const char *bar = "==================================================================";
    fprintf(f, "%s\n", bar);
    fprintf(f, "GPU 0: %s\n", info.device.c_str());
    fprintf(f, "     Compute Units: %d\n", info.compute_units);
    fprintf(f, "     HBM: %.1f GB\n\n", (double)info.hbm_bytes / 1e9);
    fprintf(f, "CPU: %s\n", cpu_model().c_str());
    fprintf(f, "     Number of Logical Cores: %u\n\n", std::thread::hardware_concurrency());
    fprintf(f, "Engine: %s\n", info.engine.c_str());
    fprintf(f, "%s\n", bar);
    fprintf(f, "transSetFile: %s\n", trans_set_file.c_str());
    fprintf(f, "dimension: %d\n", db.dimension);
    fprintf(f, "transNum: %d\n", db.trans_num);
    fprintf(f, "itemIDmin: %d\n", db.item_id_min);
    fprintf(f, "itemIDmax: %d\n", db.item_id_max_dense());
    fprintf(f, "transLengthMin: %lld\n", (long long)db.len_min);
    fprintf(f, "transLengthMax: %lld\n", (long long)db.len_max);
    fprintf(f, "transLengthMean: %g\n", db.len_mean);
    fprintf(f, "\nsupportThreshold: %g\n\n", db.support);
    fprintf(f, "dimensionReduced: %d\n", db.dimension_reduced);
    fprintf(f, "transNumReduced: %d\n", db.trans_num_reduced);
    fprintf(f, "transLengthReducedMin: %lld\n", (long long)db.red_len_min);
    fprintf(f, "transLengthReducedMax: %lld\n", (long long)db.red_len_max);
    fprintf(f, "transLengthReducedMean: %g\n", db.red_len_mean);
    fprintf(f, "\nGPUtaskPercentage: 100%%\n");
    fprintf(f, "%s\n", bar);
    fprintf(f, "preprocessing_time(s): %g\n", info.preprocessing_time);
    fprintf(f, "computation_time(s): %g\n", info.computation_time);
    fprintf(f, "total_time(s): %g\n", info.preprocessing_time + info.computation_time);
    fprintf(f, "wall_time(s): %g\n", info.wall_time);
    fprintf(f, "\nMFIsNumber: %zu\n", mfis.size());
    fprintf(f, "============================List of MFIs:=============================\n");
    // the list, formatted into one buffer per block of lines (integers by
    // to_chars; the support text changes only with the frequency, which the
    // order groups)
    std::string buf;
    buf.reserve(1 << 22);
    char tail[96] = {0};
    int tail_len = 0;
    int32_t tail_freq = -1;
    auto put = [&](long long v) {
        char t[24];
        const auto r = std::to_chars(t, t + sizeof t, v);
        buf.append(t, r.ptr);
    };
    for (size_t i = 0; i < mfis.size(); i++) {
        const Mfi &m = mfis[i];
        put((long long)(i + 1));
        buf += "th: ";
        put((long long)m.items.size());
        buf += " { ";
        for (int32_t x : m.items) {
            put(db.dense_to_orig[x]);
            buf += ' ';
        }
        if (m.frequency != tail_freq) {
            tail_freq = m.frequency;
            tail_len = snprintf(tail, sizeof tail, "}   support: %g  frequency: %d\n",
                                (double)m.frequency / (double)db.trans_num, m.frequency);
        }
        buf.append(tail, (size_t)tail_len);
        if (buf.size() >= (1 << 22) - 4096) {
            fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    fwrite(buf.data(), 1, buf.size(), f);
    fprintf(f, "%s\n", bar);
    fclose(f);
    return true;
}

}  // namespace tfim
