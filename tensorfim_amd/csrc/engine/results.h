// Results file in the reference layout (kernel_bmma.cu:1034-1073): an
// environment block, dataset statistics, timers and one line per MFI,
//   "<i>th: <len> { <ids in insertion order> }   support: <%g>  frequency: <f>"
// sorted by frequency, descending.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dataset.h"

namespace tfim {

struct Mfi {
    std::vector<int32_t> items;   // dense ids, insertion order (largest first)
    int32_t frequency = 0;
};

struct RunInfo {
    std::string device;           // GPU name / arch
    int32_t compute_units = 0;
    int64_t hbm_bytes = 0;
    std::string engine;           // engine description line
    double preprocessing_time = 0;  // reduce + bitmap build (reference convention)
    double computation_time = 0;    // root push .. DFS end (reference convention)
    double wall_time = 0;           // parse .. file written
};

// Sorts `mfis` (frequency desc, ties by item list) and writes the file.
bool write_results(const std::string &path, const std::string &trans_set_file, const Database &db,
                   std::vector<Mfi> &mfis, const RunInfo &info);

}  // namespace tfim
