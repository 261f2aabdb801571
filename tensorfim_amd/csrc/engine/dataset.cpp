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

#include "tfim.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <numeric>
#include <thread>
#include <utility>

namespace tfim {

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// std::allocator whose resize leaves new elements uninitialised: the
// parser's threads write every element of the gigabyte id and offset arrays,
// and their first touch spreads the page faults over those threads instead
// of one serial zero fill.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U> &) {}
    template <class U>
    void construct(U *q) noexcept {
        ::new ((void *)q) U;
    }
    template <class U, class... A>
    void construct(U *q, A &&...a) {
        ::new ((void *)q) U(std::forward<A>(a)...);
    }
};
using IdVec = std::vector<int32_t, UninitAlloc<int32_t>>;
using OffVec = std::vector<int64_t, UninitAlloc<int64_t>>;

inline bool is_digit(unsigned char c) { return (unsigned char)(c - '0') < 10; }

// Ids and rows of bytes [p, end) under the rules of scan_piece below, without
// parsing: an id starts at every digit whose previous byte is neither a digit
// nor '-' (a '-' drops the digit run after it), and every '\n' ends a row.
// The byte before p is a '\n' or the start of the file.  Branch-free over
// blocks so that the compiler vectorises it.
void count_piece(const char *p, const char *end, int final_row, int64_t &n_ids, int64_t &n_rows) {
    const unsigned char *u = reinterpret_cast<const unsigned char *>(p);
    const size_t n = (size_t)(end - p);
    int64_t ni = 0, nr = 0;
    if (n > 0) ni = is_digit(u[0]);
    for (size_t b = 1; b < n; b += 4096) {
        const size_t e = std::min(n, b + 4096);
        uint32_t si = 0, sr = 0;
        for (size_t i = b; i < e; i++) {
            const unsigned char c = u[i], q = u[i - 1];
            si += (uint32_t)(is_digit(c) & !is_digit(q) & (q != '-'));
            sr += (uint32_t)(c == '\n');
        }
        ni += si;
        nr += sr;
    }
    if (n > 0) nr += u[0] == '\n';
    if (final_row == 1) {
        nr++;
    } else if (final_row == 2) {
        // a final row without '\n' counts if it holds an id
        size_t t = n;
        while (t > 0 && u[t - 1] != '\n') t--;
        bool any = false;
        for (size_t i = t; i < n && !any; i++)
            any = is_digit(u[i]) && (i == 0 || (!is_digit(u[i - 1]) && u[i - 1] != '-'));
        nr += any;
    }
    n_ids = ni;
    n_rows = nr;
}

// The reader's rules over bytes [p, end): a digit run is an id, '-' and its
// digits are dropped, any other byte separates, '\n' ends a row.  The bytes
// after the last '\n' form a final row when `final_row` is 1 (always) or 2
// (only if it holds an id); 0 ignores them (a piece that ends right after a
// '\n' has none).  Writes the ids from ids[0] and the end offset of each row,
// id_base plus the ids so far, from ends[0]; returns the largest id through
// idmax, and sets too_big if an id exceeds kMaxItemId.  At most max_ids ids
// and max_rows rows are written (the counts of count_piece); n_ids and n_rows
// report what the bytes hold.
void scan_piece(const char *p, const char *end, int final_row, int32_t *ids, int64_t max_ids, int64_t *ends,
                int64_t max_rows, int64_t id_base, int64_t &n_ids, int64_t &n_rows, int32_t &idmax,
                bool &too_big) {
    int64_t ni = 0, nr = 0, row_start = 0;
    int32_t mx = idmax;
    while (p < end) {
        const char c = *p;
        if (c == '\n') {
            if (nr < max_rows) ends[nr] = id_base + ni;
            nr++;
            row_start = ni;
            ++p;
            continue;
        }
        if (c >= '0' && c <= '9') {
            // ids past kMaxItemId are refused (saturated here, reported
            // through too_big)
            int64_t v = 0;
            while (p < end && *p >= '0' && *p <= '9') {
                v = v * 10 + (*p++ - '0');
                if (v > kMaxItemId) v = (int64_t)kMaxItemId + 1;
            }
            if (v > kMaxItemId) {
                too_big = true;
                v = kMaxItemId;
            }
            if (ni < max_ids) ids[ni] = (int32_t)v;
            mx = std::max(mx, (int32_t)v);
            ni++;
            continue;
        }
        if (c == '-') {
            ++p;
            while (p < end && *p >= '0' && *p <= '9') ++p;
            continue;
        }
        ++p;
    }
    if (final_row == 1 || (final_row == 2 && ni > row_start)) {
        if (nr < max_rows) ends[nr] = id_base + ni;
        nr++;
    }
    n_ids = ni;
    n_rows = nr;
    idmax = mx;
}

// Parses a whole file image into CSR rows with up to `threads` threads: the
// bytes are cut right after '\n' bytes into one piece per thread; a first
// pass counts each piece's ids and rows, a second parses them straight into
// the joined arrays (ids and row offsets), so the rows are those of one
// sequential pass and no thread allocates memory of its own.
// `newline_at_end` treats the image as if a '\n' followed it (the replicated
// database appends one when missing).  Returns false, with the reason in
// err, if an id exceeds kMaxItemId or the two passes disagree (a bug;
// nothing is written beyond the arrays either way).
bool parse_parallel(const char *p, size_t n, bool newline_at_end, int threads, IdVec &ids, OffVec &off,
                    int32_t &idmax, std::string &err) {
    threads = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), n / 2 + 1));
    std::vector<size_t> cut(threads + 1, n);
    cut[0] = 0;
    for (int i = 1; i < threads; i++) {
        size_t b = std::max(cut[i - 1], n / threads * i);
        while (b < n && p[b - 1] != '\n') b++;
        cut[i] = b;
    }
    // the bytes after the file's last '\n' belong to the piece that ends the
    // image (trailing pieces past it are empty)
    auto final_row = [&](int i) {
        const bool ends = cut[i + 1] == n && (cut[i] < n || i == 0);
        return ends ? (newline_at_end ? 1 : 2) : 0;
    };
    std::vector<int64_t> pids(threads + 1, 0), prows(threads + 1, 0);
    std::vector<int32_t> pmax(threads, 0);
    std::vector<uint8_t> agree(threads, 1), big(threads, 0);
    auto run = [&](auto &&fn) {
        std::vector<std::thread> pool;
        for (int i = 1; i < threads; i++) pool.emplace_back(fn, i);
        fn(0);
        for (auto &t : pool) t.join();
    };
    run([&](int i) { count_piece(p + cut[i], p + cut[i + 1], final_row(i), pids[i + 1], prows[i + 1]); });
    for (int i = 0; i < threads; i++) {
        pids[i + 1] += pids[i];
        prows[i + 1] += prows[i];
    }
    // large arrays: ask for transparent huge pages before the first touch
    // (page faults of 4 KB pages dominate filling gigabytes)
    auto huge = [](void *ptr, size_t bytes) {
        const uintptr_t a = ((uintptr_t)ptr + (2u << 20) - 1) & ~(uintptr_t)((2u << 20) - 1);
        const uintptr_t e = ((uintptr_t)ptr + bytes) & ~(uintptr_t)((2u << 20) - 1);
        if (e > a) madvise((void *)a, e - a, MADV_HUGEPAGE);
    };
    ids.reserve((size_t)pids[threads]);
    huge(ids.data(), (size_t)pids[threads] * 4);
    ids.resize((size_t)pids[threads]);
    off.reserve((size_t)prows[threads] + 1);
    huge(off.data(), ((size_t)prows[threads] + 1) * 8);
    off.resize((size_t)prows[threads] + 1);
    off[0] = 0;
    run([&](int i) {
        const int64_t want_ids = pids[i + 1] - pids[i], want_rows = prows[i + 1] - prows[i];
        int64_t a = 0, b = 0;
        bool too_big = false;
        scan_piece(p + cut[i], p + cut[i + 1], final_row(i), ids.data() + pids[i], want_ids,
                   off.data() + prows[i] + 1, want_rows, pids[i], a, b, pmax[i], too_big);
        agree[i] = a == want_ids && b == want_rows;
        big[i] = too_big;
    });
    for (int i = 0; i < threads; i++) {
        idmax = std::max(idmax, pmax[i]);
        if (!agree[i]) {
            err = "parser passes disagree on the id or row count";
            return false;
        }
        if (big[i]) {
            err = "an id exceeds " + std::to_string(kMaxItemId);
            return false;
        }
    }
    return true;
}

// Dense namespace from occurrence counts: survivors (id >= 1, ascending),
// a frequent item 0 last, the frequency order and the roots.  Returns the
// map original id -> dense id (-1 = dropped).
std::vector<int32_t> dense_namespace(Database &db, const std::vector<int32_t> &freq, int32_t idmax, double support) {
    const double thr = db.thr_abs;
    auto frequent = [&](int32_t x) { return !((double)freq[x] < thr); };
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

    return o2d;
}

}  // namespace

bool load_database(const std::string &path, double support, int replicate, Database &db,
                   std::string &err, int parse_threads, tfim_ctx *ctx) {
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
    IdVec ids;
    OffVec off;
    int32_t parsed_max = 0;
    const int threads = parse_threads > 0 ? parse_threads
                        : n < ((size_t)4 << 20) ? 1
                                                : (int)std::min(32u, std::max(1u, std::thread::hardware_concurrency()));
    std::string perr;
    const bool parsed = parse_parallel(img, n, newline_at_end, threads, ids, off, parsed_max, perr);
    if (map != MAP_FAILED) munmap(map, n);
    close(fd);
    if (!parsed) {
        err = path + ": " + perr;
        return false;
    }
    const int64_t base_rows = (int64_t)off.size() - 1;
    const int64_t rep = replicate >= 1 ? replicate : 1;
    db.replicate = (int32_t)rep;
    db.support = support;
    db.parse_seconds = now() - t0;
    const double t1 = now();

    if (ctx && replicate < 1) {
        // occurrence counts, reduction and bitmaps on the device; lengths
        // from the row offsets here
        int64_t lmin = INT_MAX, lmax = 0;
        double lsum = 0;
        for (int64_t t = 0; t < base_rows; t++) {
            const int64_t len = off[t + 1] - off[t];
            lmin = std::min(lmin, len);
            lmax = std::max(lmax, len);
            lsum += (double)len;
        }
        std::vector<int32_t> freq((size_t)parsed_max + 1, 0);
        if (tfim_prep_begin(ctx, ids.data(), (int64_t)ids.size(), off.data(), base_rows, parsed_max, freq.data()) !=
            TFIM_OK) {
            err = std::string("device preprocessing: ") + tfim_last_error(ctx);
            return false;
        }
        IdVec().swap(ids);
        OffVec().swap(off);
        int32_t idmin = INT_MAX, dim = 0;
        for (int32_t x = 0; x <= parsed_max; x++)
            if (freq[x] > 0) {
                idmin = std::min(idmin, x);
                dim++;
            }
        db.trans_num = (int32_t)base_rows;
        db.item_id_min = idmin;
        db.item_id_max_raw = parsed_max;
        db.len_min = lmin;
        db.len_max = lmax;
        db.len_mean = base_rows > 0 ? lsum / (double)base_rows : 0.0;
        db.dimension = dim;
        db.thr_abs = (double)db.trans_num * support;
        std::vector<int32_t> o2d = dense_namespace(db, freq, parsed_max, support);
        int64_t n_red = 0, rmin = 0, rmax = 0, rsum = 0;
        int32_t dupf = 0;
        if (tfim_prep_finish(ctx, o2d.data(), db.d_count, db.n_surv, &n_red, &rmin, &rmax, &rsum, &dupf) != TFIM_OK) {
            err = std::string("device preprocessing: ") + tfim_last_error(ctx);
            return false;
        }
        db.trans_num_reduced = (int32_t)n_red;
        db.red_len_min = n_red > 0 ? rmin : INT_MAX;
        db.red_len_max = rmax;
        db.red_len_mean = (double)rsum / (double)n_red;   // NaN when nothing survives
        db.duplicates = dupf != 0;
        db.base_bits = n_red;
        db.words = (n_red + 63) / 64;
        db.on_device = true;
        db.preprocessing_seconds = now() - t1;
        return true;
    }

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
    IdVec().swap(ids);
    const int64_t base_red = (int64_t)roff.size() - 1;
    db.trans_num_reduced = (int32_t)(base_red * rep);
    db.red_len_min = rmin;
    db.red_len_max = rmax;
    db.red_len_mean = (rsum * (double)rep) / (double)(base_red * rep);   // NaN when nothing survives

    std::vector<int32_t> o2d = dense_namespace(db, freq, idmax, support);

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
