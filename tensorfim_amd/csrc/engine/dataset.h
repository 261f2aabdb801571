// Transaction database: reader, reduction, dense compaction, vertical bitmaps.
// Semantics follow the reference engine exactly (kernel_bmma.cu:208-301,
// 388-426, 1075-1086, 1186-1215); see the parity quirks in dataset.cpp.
#pragma once

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

struct tfim_ctx;

namespace tfim {

// Largest item id the reader accepts.  Per-id tables are sized id_max + 1 and
// the device histogram id_max + 2 entries, counted in int32, so two ids below
// INT32_MAX stay free and every loop `x <= id_max` ends.  Larger ids make the
// load fail with a message.
constexpr int32_t kMaxItemId = INT32_MAX - 2;

struct Database {
    // raw statistics (readTransSetFile)
    int32_t trans_num = 0;        // counts empty lines
    int32_t dimension = 0;
    int32_t item_id_min = 0;
    int32_t item_id_max_raw = 0;
    int64_t len_min = 0, len_max = 0;
    double len_mean = 0;

    double support = 0;           // relative threshold (argv[2])
    double thr_abs = 0;           // frequencyThreshold = transNum * support

    int32_t trans_num_reduced = 0;
    int32_t dimension_reduced = 0;
    int64_t red_len_min = 0, red_len_max = 0;
    double red_len_mean = 0;

    // dense namespace
    int32_t n_surv = 0;           // F: survivors with id >= 1 (columns)
    int32_t d_count = 0;          // F + (item 0 frequent)
    std::vector<int32_t> dense_to_orig;
    std::vector<int32_t> freq_dense;      // occurrence counts, dense ids
    std::vector<int32_t> sorted_items;    // items_inFreqSort, d_count entries
    std::vector<uint8_t> is_root;         // [d_count]
    int32_t n_roots = 0;

    // base bitmaps (one copy of the rows when replicate > 1)
    int32_t replicate = 1;
    int64_t base_bits = 0;        // transactions in one copy (after reduction)
    int64_t words = 0;            // u64 words per base row
    std::vector<uint64_t> bitmap; // [d_count][words]
    bool duplicates = false;      // some row repeats an item id
    bool on_device = false;       // bitmaps built on the device (bitmap stays empty)

    double parse_seconds = 0, preprocessing_seconds = 0;

    int32_t item_id_max_dense() const { return d_count - 1; }
    int64_t n_bits() const { return base_bits * replicate; }
};

// Reads `path` (whitespace-separated ids, one transaction per line) and
// prepares it at `support`.  replicate >= 1 gives the database
// data/make_replica.py writes (file + '\n' if missing, concatenated
// `replicate` times); 0 reads the file as is.  parse_threads = 0 uses the
// hardware threads (files under 4 MB are parsed by one).  With a context and
// replicate == 0 the counts, reduction and bitmaps are computed on the device
// and left resident in `ctx` (db.on_device); otherwise on the host into
// db.bitmap.  Returns false if unreadable (or on a device error).
bool load_database(const std::string &path, double support, int replicate, Database &db,
                   std::string &err, int parse_threads = 0, tfim_ctx *ctx = nullptr);

}  // namespace tfim
