// Reader and preprocessing.  Reference semantics (kernel_bmma.cu), kept
// exactly because they decide which itemsets are reported:
//   - every '\n' ends a transaction, empty ones included, so empty lines
//     raise transNum and the absolute threshold (:219, quirk A.1); a final
//     line without '\n' counts only if it holds an id (:233);
//   - a digit run is an id, '-' + digits is dropped, any other byte
//     separates (:220-231);
//   - item frequencies are occurrence counts (:272-276);
//   - survivors: id >= 1 with freq >= transNum*support (double compare), in
//     ascending id order -> dense ids; a frequent item 0 gets dense id F and is
//     never a column (:394-403, quirk A.3);
//   - roots: the first dimensionReduced items of the dense list sorted by
//     frequency (stable, ties ascending id), dimensionReduced counting
//     freq/transNum > support strictly over all raw ids (:1199, :411-414,
//     :1224-1233, quirk A.2).
#include "dataset.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <numeric>
#include <thread>

namespace tfim {

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Parses bytes [p, end) with the reader's rules; appends ids and the length
// of every row ended by '\n'.  The bytes after the last '\n' form a final row
// when `final_row` is 1 (always) or 2 (only if it holds an id); 0 ignores
// them (a piece that ends right after a '\n' has none).
void parse_piece(const char *p, const char *end, int final_row, std::vector<int32_t> &ids,
                 std::vector<int32_t> &lens, int32_t &idmax) {
    size_t row_start = ids.size();
    while (p < end) {
        const char c = *p;
        if (c == '\n') {
            lens.push_back((int32_t)(ids.size() - row_start));
            row_start = ids.size();
            ++p;
            continue;
        }
        if (c >= '0' && c <= '9') {
            long long v = 0;
            while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
            ids.push_back((int32_t)v);
            idmax = std::max(idmax, (int32_t)v);
            continue;
        }
        if (c == '-') {
            ++p;
            while (p < end && *p >= '0' && *p <= '9') ++p;
            continue;
        }
        ++p;
    }
    if (final_row == 1 || (final_row == 2 && ids.size() > row_start)) lens.push_back((int32_t)(ids.size() - row_start));
}

// Parses a whole file image into CSR rows with up to `threads` threads: the
// bytes are cut right after '\n' bytes into one piece per thread, each piece
// is parsed on its own and the pieces are joined in order, so the rows are
// those of one sequential pass.  `newline_at_end` treats the image as if a
// '\n' followed it (the replicated database appends one when missing).
void parse_parallel(const char *p, size_t n, bool newline_at_end, int threads, std::vector<int32_t> &ids,
                    std::vector<int64_t> &off, int32_t &idmax) {
    threads = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), n / 2 + 1));
    std::vector<size_t> cut(threads + 1, n);
    cut[0] = 0;
    for (int i = 1; i < threads; i++) {
        size_t b = std::max(cut[i - 1], n / threads * i);
        while (b < n && p[b - 1] != '\n') b++;
        cut[i] = b;
    }
    std::vector<std::vector<int32_t>> pid(threads), plen(threads);
    std::vector<int32_t> pmax(threads, 0);
    auto work = [&](int i) {
        pid[i].reserve((cut[i + 1] - cut[i]) / 3 + 16);
        // the bytes after the file's last '\n' belong to the piece that ends
        // the image (trailing pieces past it are empty)
        const bool ends = cut[i + 1] == n && (cut[i] < n || i == 0);
        const int final_row = ends ? (newline_at_end ? 1 : 2) : 0;
        parse_piece(p + cut[i], p + cut[i + 1], final_row, pid[i], plen[i], pmax[i]);
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; i++) pool.emplace_back(work, i);
    work(0);
    for (auto &t : pool) t.join();
    size_t n_ids = 0, n_rows = 0;
    for (int i = 0; i < threads; i++) {
        n_ids += pid[i].size();
        n_rows += plen[i].size();
        idmax = std::max(idmax, pmax[i]);
    }
    ids.resize(n_ids);
    off.resize(n_rows + 1);
    off[0] = 0;
    size_t io = 0, ro = 0;
    for (int i = 0; i < threads; i++) {
        std::copy(pid[i].begin(), pid[i].end(), ids.begin() + (ptrdiff_t)io);
        for (int32_t l : plen[i]) {
            off[ro + 1] = off[ro] + l;
            ro++;
        }
        io += pid[i].size();
        std::vector<int32_t>().swap(pid[i]);
    }
}

}  // namespace

bool load_database(const std::string &path, double support, int replicate, Database &db,
                   std::string &err, int parse_threads) {
    const double t0 = now();
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        err = path + " open failed!";
        return false;
    }
    struct stat stt;
    if (fstat(fd, &stt) != 0) {
        close(fd);
        err = path + " open failed!";
        return false;
    }
    const size_t n = (size_t)stt.st_size;
    const char *img = nullptr;
    void *map = MAP_FAILED;
    if (n > 0) {
        map = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            close(fd);
            err = path + " read failed!";
            return false;
        }
        img = static_cast<const char *>(map);
    }
    // the replicated database is the file plus '\n' when it lacks one
    const bool newline_at_end = replicate >= 1 && (n == 0 || img[n - 1] != '\n');
    std::vector<int32_t> ids;
    std::vector<int64_t> off;
    int32_t parsed_max = 0;
    const int threads = parse_threads > 0 ? parse_threads
                        : n < ((size_t)4 << 20) ? 1
                                                : (int)std::min(32u, std::max(1u, std::thread::hardware_concurrency()));
    parse_parallel(img, n, newline_at_end, threads, ids, off, parsed_max);
    if (map != MAP_FAILED) munmap(map, n);
    close(fd);
    const int64_t base_rows = (int64_t)off.size() - 1;
    const int64_t rep = replicate >= 1 ? replicate : 1;
    db.replicate = (int32_t)rep;
    db.support = support;
    db.parse_seconds = now() - t0;
    const double t1 = now();

    // statistics over the (replicated) database: every copy repeats the base
    int32_t idmin = INT_MAX, idmax = 0;
    int64_t lmin = INT_MAX, lmax = 0;
    double lsum = 0;
    bool dup = false;
    for (int64_t t = 0; t < base_rows; t++) {
        for (int64_t j = off[t]; j < off[t + 1]; j++) {
            idmax = std::max(idmax, ids[j]);
            idmin = std::min(idmin, ids[j]);
        }
        const int64_t len = off[t + 1] - off[t];
        lmin = std::min(lmin, len);
        lmax = std::max(lmax, len);
        lsum += (double)len;
    }
    const int64_t rows = base_rows * rep;
    db.trans_num = (int32_t)rows;
    db.item_id_min = idmin;
    db.item_id_max_raw = idmax;
    db.len_min = lmin;
    db.len_max = lmax;
    db.len_mean = rows > 0 ? (lsum * (double)rep) / (double)rows : 0.0;

    std::vector<int32_t> freq((size_t)idmax + 1, 0);
    {
        std::vector<int32_t> seen((size_t)idmax + 1, -1);
        for (int64_t t = 0; t < base_rows; t++)
            for (int64_t j = off[t]; j < off[t + 1]; j++) {
                if (seen[ids[j]] == (int32_t)t) dup = true;
                seen[ids[j]] = (int32_t)t;
                freq[ids[j]]++;
            }
    }
    db.duplicates = dup;
    int32_t dim = 0;
    for (auto f : freq) dim += f > 0;
    db.dimension = dim;
    for (auto &f : freq) f = (int32_t)((int64_t)f * rep);

    const double thr = (double)db.trans_num * support;
    db.thr_abs = thr;
    auto frequent = [&](int32_t x) { return !((double)freq[x] < thr); };

    // reduction: drop infrequent ids, then empty rows (one copy; copies agree)
    std::vector<int64_t> roff(1, 0);
    std::vector<int32_t> rids;
    rids.reserve(ids.size());
    double rsum = 0;
    int64_t rmin = INT_MAX, rmax = 0;
    for (int64_t t = 0; t < base_rows; t++) {
        const size_t m0 = rids.size();
        for (int64_t j = off[t]; j < off[t + 1]; j++)
            if (frequent(ids[j])) rids.push_back(ids[j]);
        if (rids.size() > m0) {
            roff.push_back((int64_t)rids.size());
            const int64_t len = (int64_t)(rids.size() - m0);
            rmin = std::min(rmin, len);
            rmax = std::max(rmax, len);
            rsum += (double)len;
        }
    }
    std::vector<int32_t>().swap(ids);
    const int64_t base_red = (int64_t)roff.size() - 1;
    db.trans_num_reduced = (int32_t)(base_red * rep);
    db.red_len_min = rmin;
    db.red_len_max = rmax;
    db.red_len_mean = (rsum * (double)rep) / (double)(base_red * rep);   // NaN when nothing survives

    int32_t dred = 0;
    for (int32_t x = 0; x <= idmax; x++)
        if ((double)freq[x] / (double)db.trans_num > support) dred++;
    db.dimension_reduced = dred;

    // dense compaction
    std::vector<int32_t> o2d((size_t)idmax + 1, -1);
    db.dense_to_orig.clear();
    db.freq_dense.clear();
    for (int32_t i = 1; i <= idmax; i++)
        if (frequent(i)) {
            o2d[i] = (int32_t)db.dense_to_orig.size();
            db.dense_to_orig.push_back(i);
            db.freq_dense.push_back(freq[i]);
        }
    db.n_surv = (int32_t)db.dense_to_orig.size();
    if (frequent(0)) {
        o2d[0] = db.n_surv;
        db.dense_to_orig.push_back(0);
        db.freq_dense.push_back(freq[0]);
    }
    db.d_count = (int32_t)db.dense_to_orig.size();
    db.sorted_items.resize(db.d_count);
    std::iota(db.sorted_items.begin(), db.sorted_items.end(), 0);
    std::stable_sort(db.sorted_items.begin(), db.sorted_items.end(),
                     [&](int32_t a, int32_t b) { return db.freq_dense[a] > db.freq_dense[b]; });
    // The reference pushes items_inFreqSort[0..dimensionReduced); its sorted
    // list carries one zero-frequency slot past d_count without a bitmap row,
    // so the root count is clamped to real items.
    db.n_roots = std::min(dred, db.d_count);
    db.is_root.assign(db.d_count, 0);
    for (int32_t i = 0; i < db.n_roots; i++) db.is_root[db.sorted_items[i]] = 1;

    // vertical bitmaps of one copy, LSB-first
    db.base_bits = base_red;
    db.words = (base_red + 63) / 64;
    db.bitmap.assign((size_t)std::max(1, db.d_count) * (size_t)std::max<int64_t>(1, db.words), 0ull);
    for (int64_t t = 0; t < base_red; t++)
        for (int64_t j = roff[t]; j < roff[t + 1]; j++) {
            const int32_t x = o2d[rids[j]];
            db.bitmap[(size_t)x * db.words + (t >> 6)] |= 1ull << (t & 63);
        }
    db.preprocessing_seconds = now() - t1;
    return true;
}

}  // namespace tfim
