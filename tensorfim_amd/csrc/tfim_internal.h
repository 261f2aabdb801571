// Internal declarations shared by the libtfim HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "tfim.h"

namespace tfim {

// Row pitch granularity: 2048 bits = 32 u64 words (kernel_bmma.cu:389).
constexpr int64_t kPitchWords = 32;

inline int64_t pitch_for_bits(int64_t n_bits) {
    return ((n_bits + 2047) / 2048) * kPitchWords;
}

// One kernel launch, described so that it can be issued directly or set as
// the parameters of a graph kernel node.  Arguments live in `buf`; params()
// returns the pointer array HIP expects (valid while this object lives).  An
// argument list that overflows `buf` marks the launch invalid (`overflow`):
// launch() then returns hipErrorInvalidValue and the graph path refuses it.
struct KernelLaunch {
    const void *func = nullptr;
    dim3 grid{1, 1, 1}, block{1, 1, 1};
    uint32_t shmem = 0;
    static constexpr int kMaxArgs = 12;
    alignas(16) unsigned char buf[256];
    int off[kMaxArgs] = {};
    int n_args = 0, used = 0;
    bool overflow = false;
    void *ptrs[kMaxArgs] = {};

    void reset() {
        n_args = used = 0;
        overflow = false;
    }
    template <class T>
    void arg(const T &v) {
        const int a = (int)alignof(T) < 8 ? 8 : (int)alignof(T);
        const int at = (used + a - 1) / a * a;
        static_assert(sizeof(T) <= sizeof(buf), "argument too large");
        if (n_args >= kMaxArgs || at + (int)sizeof(T) > (int)sizeof(buf)) {
            overflow = true;
            return;
        }
        __builtin_memcpy(buf + at, &v, sizeof(T));
        off[n_args++] = at;
        used = at + (int)sizeof(T);
    }
    void **params() {
        for (int i = 0; i < n_args; i++) ptrs[i] = buf + off[i];
        return ptrs;
    }
    hipError_t launch(hipStream_t s) {
        if (overflow || !func) return hipErrorInvalidValue;
        return hipLaunchKernel(func, grid, block, params(), shmem, s);
    }
};

// ---- launchers (defined in kernels_*.hip) --------------------------------

// Plain per-node count: out[c] += popcount(AND rows[items] & rows[c]) for
// c < n_cols; `out` must be zeroed.  items are device ids.  With
// spread_stride > 0 the counts are spread over 8 copies out[j * spread_stride
// + c], j < 8, which the caller adds up.
void launch_count_node(const uint64_t *bm, int64_t pitch_words, const int32_t *items, int32_t k,
                       int32_t n_cols, int32_t *out, int64_t spread_stride, int n_cu, int shape, hipStream_t s);

// A node's prefix passed by value in the kernel's argument block.
constexpr int kNodeArgItems = TFIM_NOALLOC_MAX_ITEMS;
struct NodeItems {
    int32_t n;
    int32_t id[kNodeArgItems];
};
// launch_count_node with the prefix in the argument block.
void launch_count_node_args(const uint64_t *bm, int64_t pitch_words, const NodeItems &items, int32_t n_cols,
                            int32_t *out, int64_t spread_stride, int n_cu, int shape, hipStream_t s);
// out[e] = sum of the 8 copies copies[j * stride + e], e < n.
void launch_sum_copies8(const int32_t *copies, int64_t stride, int32_t n, int32_t *out, hipStream_t s);

// One tile of a wave: a row's prefix, and a run of its columns.  compl == 0:
// the columns are entries [col_begin, col_end) of the row's column list.
// compl == j + 1: the row is complement row j, the columns are the survivor
// ids [col_begin, col_end) (one 64-aligned word of its exclusion bitmap) and
// only those whose exclusion bit is clear, plus the row's first prefix item,
// are counted.
struct WaveTile {
    int32_t row;        // node index
    int32_t col_begin;
    int32_t col_end;
    int32_t compl_;
};

// Wave descriptor consumed by the fused wave kernel (device pointers).
struct WaveDesc {
    const int32_t *item_ids;   // prefix items, CSR by item_off
    const int32_t *item_off;   // [n_rows+1]
    const int32_t *col_ids;    // columns, CSR by col_off; nullptr = identity over [0,n_cols)
    const int32_t *col_off;    // [n_rows+1] output offsets (also when col_ids is null)
    const WaveTile *tiles;     // [n_tiles]
    int32_t *out;              // counts, indexed by col_off[row] + j (complement rows: + survivor id)
    const uint64_t *excl;      // complement rows' exclusion bitmaps, excl_words each
    int32_t n_tiles;
    int32_t n_chunks;          // word chunks of the wave kernel's chunk size
    int32_t excl_words;
};

// Complement-row reduction after the wave kernel: for complement row j (row
// compl_rows[j], first prefix item x0, full counts at out[col_off[row] + e]):
// res[2j] = count of x0 (= support of the prefix) and res[2j+1] = number of
// survivors e != x0 with a clear exclusion bit and count >= min_count.
void plan_compl_reduce(const WaveDesc &d, const int32_t *compl_rows, int32_t n_compl, int32_t n_cols,
                       int32_t min_count, int32_t *res, KernelLaunch &L);

// Fused wave kernel: prefix AND + column counts + split-k accumulation
// (atomics).  `out` must be zeroed.  grid_blocks must be a multiple of 8.
void launch_count_wave(const uint64_t *bm, int64_t pitch_words, const WaveDesc &d,
                       int grid_blocks, int chunk_uint4_per_lane, hipStream_t s);
void plan_count_wave(const uint64_t *bm, int64_t pitch_words, const WaveDesc &d, int grid_blocks,
                     int chunk_uint4_per_lane, KernelLaunch &L);
int wave_chunk_words(int chunk_uint4_per_lane);

// LDS-staged wave: every item the wave touches gets an LDS slot; prefixes and
// columns are given as slot indices.  Each workgroup stages one word chunk of
// all slots at a time and runs every tile of the wave against it.
// A tile of the LDS kernel, resolved on the host: prefix slots
// item_slots[ib .. ib+k), column slots col_slots[cb .. cb+nc), outputs
// out[cb .. cb+nc).
struct LdsTile {
    int32_t ib, k, cb, nc;
};
struct WaveLdsDesc {
    const int32_t *slot_rows;  // [n_slots] bitmap row of each slot
    const int32_t *item_slots; // prefix slots, CSR by item_off
    const int32_t *item_off;   // [n_rows+1]
    const int32_t *col_slots;  // column slots, CSR by col_off
    const int32_t *col_off;    // [n_rows+1] (also output offsets)
    const LdsTile *tiles;
    int32_t *out;              // [n_out], zeroed
    int32_t n_slots, n_tiles, n_out;
};
// Largest slot count / output count the LDS kernel accepts.
constexpr int kLdsMaxSlots = 96;
// Largest prefix-mask block of one wave on the matrix-core path (each of the
// three wave slots keeps its own; 288 GB of HBM holds them beside the bitmaps).
constexpr int64_t kMfmaMaskBytes = (int64_t)8 << 30;
constexpr int kLdsMaxOut = 8192;
// Returns false if the shape does not fit LDS (caller uses launch_count_wave).
bool launch_count_wave_lds(const uint64_t *bm, int64_t pitch_words, const WaveLdsDesc &d, int n_cu,
                           hipStream_t s);
bool plan_count_wave_lds(const uint64_t *bm, int64_t pitch_words, const WaveLdsDesc &d, int n_cu, bool wide,
                         KernelLaunch &L);

// LDS-staged wave with rows given as bit masks (databases of at most 64 dense
// ids): slot i is dense id i, every id is staged; a tile holds its prefix and
// its columns as masks, so neither the host nor the kernel handles id lists.
// Outputs of a tile: out[ob + j] for its j-th column in ascending id order.
struct MaskTile {
    uint64_t pm;     // prefix ids (at least one)
    uint64_t cm;     // column ids (at least one, all < n_cols)
    int32_t ob;      // output index of the first column
    int32_t k, nc;   // popcount(pm), popcount(cm)
    int32_t pad;
};
static_assert(sizeof(MaskTile) == 32, "mask tiles are 32-byte records");
struct WaveMaskDesc {
    const MaskTile *tiles;
    int32_t *out;        // [n_out], zeroed
    // [2], zeroed: the kernel's span on the device wall clock, as the max of
    // the complement of each workgroup's start and the max of its end
    unsigned long long *stamp;
    int32_t n_items;     // ids staged: 0 .. n_items - 1 (<= 64)
    int32_t n_tiles, n_out;
};
constexpr int kMaskMaxItems = 64;
// Returns false if the shape does not fit LDS with the requested staging
// (wide: one workgroup per CU; narrow: two).
bool plan_count_wave_mask(const uint64_t *bm, int64_t pitch_words, const WaveMaskDesc &d, int n_cu, bool wide,
                          KernelLaunch &L);

// Device preprocessing (kernels_prep.hip).
void launch_prep_hist(const int32_t *ids, int64_t n, int32_t nbins, int32_t *freq, int32_t *bad, int n_cu,
                      hipStream_t s);
int64_t prep_blocks(int64_t n_rows);
void launch_prep_rows(const int32_t *ids, const int64_t *off, int64_t n_rows, const int32_t *dense_of, int32_t *cnt,
                      int32_t *blk_keep, unsigned long long *stats, hipStream_t s);
void launch_prep_scan(int32_t *blk, int64_t n_blk, int64_t *total, hipStream_t s);
void launch_prep_index(const int32_t *cnt, int64_t n_rows, const int32_t *blk_off, int32_t *kept, hipStream_t s);
void launch_prep_build(const int32_t *ids, const int64_t *off, const int32_t *kept, int64_t n_red,
                       const int32_t *dense_of, uint64_t *bm, int64_t pitch, int32_t *dup, hipStream_t s);

// Zero `n` int32 with a fill kernel (graph-capture safe).
void launch_zero_i32(int32_t *p, int64_t n, hipStream_t s);
void plan_zero_i32(int32_t *p, int64_t n, KernelLaunch &L);
// A wave's first and last nodes: its descriptor block from pinned host memory
// to the device, and its counts from the device to pinned host memory with
// the device region [0, max(n_copy, n_zero)) zeroed again.
void plan_stage16(const void *src, void *dst, int64_t bytes, KernelLaunch &L);
void plan_drain_i32(int32_t *dev, int32_t *host, int64_t n_copy, int64_t n_zero, KernelLaunch &L);

// Matrix AND-popcount, packed (VALU) path.  C must be zeroed when split_k>1.
void launch_popmat_packed(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                          int64_t pitch_words, int split_k, hipStream_t s);
// Matrix AND-popcount on the matrix cores, 256x128 tiles with the operands
// expanded in registers from packed LDS: int8 MFMA, or (fp4) e2m1 MFMA at
// twice the int8 rate.  C must be zeroed when split_k>1.
hipError_t launch_popmat_mfma_rp(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                                 int64_t pitch_words, int split_k, hipStream_t s, bool fp4 = false);
// The same as a launch description (graph node), with a row tile of 64, 128
// or 256 rows (rp_rows_per_tile(b)) x 128 columns, or for fp4 with `wide` and
// more than 128 rows and columns 256 x 256 (one workgroup per CU); C must be
// zeroed when the returned number of word splits is above 1.
int64_t plan_popmat_mfma_rp(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                            int64_t pitch_words, int split_k, KernelLaunch &L, bool fp4 = false, bool wide = true);
int rp_rows_per_tile(int64_t b);

// A training round of a learned policy's network was queued on stream s of
// ctx's device: waves issued before it ends share the device with it, and
// their times are kept out of the wave tuners.
void ctx_model_training(tfim_ctx *c, hipStream_t s);

// Matrix-core wave path: the prefix masks of a wave's rows are built in
// `masks` ([n_rows][pitch] u64), C = masks AND-popcount survivors is counted
// by k_popmat_mfma_rp ([n_rows][n_cols] int32), and k_wave_gather writes the
// wave's outputs from C: explicit rows their listed columns
// (out[col_off[r] + j] = C[r][col_ids[col_off[r] + j]]), complement row j its
// two results at res[2j] (as k_compl_reduce).  row_compl[r] = j or -1.
void plan_wave_masks(const uint64_t *bm, int64_t pitch_words, const int32_t *item_ids, const int32_t *item_off,
                     int32_t n_rows, uint64_t *masks, KernelLaunch &L);
void plan_wave_gather(const WaveDesc &d, const int32_t *C, int32_t n_rows, int32_t n_cols, const int32_t *row_compl,
                      int32_t min_count, int32_t *res, int32_t max_row_cols, KernelLaunch &L);

// Fused matrix-core wave (k_wave_mfma): one launch per wave builds each
// row's prefix mask from LDS-staged item words, counts it against survivor
// tiles on the int8 matrix cores with the word axis split over
// gridDim.y, and writes only the wave's outputs.
struct WaveMfmaDesc {
    const int32_t *slot_rows;   // [n_slots] bitmap row of each distinct prefix item
    const int32_t *item_off;    // [n_rows + 1] prefix CSR
    const int32_t *item_slots;  // prefix slots, CSR by item_off
    const int32_t *seg_off;     // [tiles_n][n_rows + 1] listed-output entries per (column tile, row), or null
    const int2 *entries;        // (output index, row << 8 | column within the tile)
    const int32_t *row_compl;   // [n_rows] complement index or -1; null when the wave has none
    const int32_t *compl_x0;    // [n_compl] first prefix item of each complement row
    const uint64_t *excl;       // [n_compl][excl_words] exclusion bitmaps
    int32_t *out;               // listed outputs (dense_out: [n_rows][n_cols])
    int32_t *res;               // two results per complement row
    int32_t *acc;               // [n_compl][n_cols] full counts of complement rows, zeroed
    int32_t *arrive;            // [tiles_m] arrival counters, zeroed
    int64_t k_per_split;        // words per split (a multiple of 8)
    int32_t n_rows, n_cols, n_slots, excl_words, min_count, tiles_n, dense_out;
};
// Dynamic LDS of a launch: the staging buffers (with up to max_tile_list
// prefix entries per row tile) or the epilogue's count tile, the larger.
size_t wave_mfma_lds_bytes(int ti, int tj, int n_slots, int max_tile_list);
// Launch description; false if the shape is outside the kernel's limits
// (more than 256 slots, more LDS than a CU has, an unsupported tile).
bool plan_wave_mfma(const uint64_t *bm, int64_t pitch_words, const WaveMfmaDesc &d, int ti, int tj, int tiles_m,
                    int splits, size_t lds_bytes, KernelLaunch &L, bool int8 = false);
constexpr int kFusedMaxSlots = 256;
// Dynamic LDS limit of k_wave_mfma: a CU's 160 KB less room for its static
// shared variable.
constexpr size_t kFusedMaxLds = 160 * 1024 - 256;

// Matrix AND-popcount, int8 MFMA path, 128x128 tiles with the operand bytes
// expanded into LDS.  C must be zeroed when split_k>1.
void launch_popmat_mfma(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                        int64_t pitch_words, int split_k, hipStream_t s);

void launch_fill_bernoulli(uint64_t *dst, int64_t rows, int64_t n_bits, int64_t pitch_words,
                           uint64_t seed, uint32_t matrix, uint32_t thr32, hipStream_t s);

void launch_replicate_rows(const uint64_t *base, int64_t base_pitch, int64_t base_bits,
                           uint64_t *dst, int64_t dst_pitch, int32_t rows, int32_t replicate,
                           hipStream_t s);

void launch_row_popcount(const uint64_t *bm, int64_t pitch_words, int32_t rows, int32_t *out,
                         hipStream_t s);

// MLP kernels (kernels_mlp.hip)
int mlp_nparam();
void launch_mlp_fwd_bwd(const float *P, const float *X, const float *Y, const float *M, int n, float inv_count,
                        float *H1, float *H2, float *D1, float *D2, float *DY, float *lossv, hipStream_t s);
int mlp_grad_slices(int n);
void launch_mlp_grad(const float *X, const float *H1, const float *H2, const float *D1, const float *D2,
                     const float *DY, int n, float *part, float *G, hipStream_t s);
void launch_mlp_adam(float *P, const float *G, float *Mo, float *Vo, uint16_t *Pb, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t s);
void launch_mlp_to_bf16(const float *P, uint16_t *Pb, hipStream_t s);
void launch_mlp_infer(const float *P, const uint16_t *Pb, const float *X, int n, float *out, hipStream_t s);

// Context accessors for translation units other than the runtime.
hipStream_t ctx_stream(tfim_ctx *c);
int ctx_device(tfim_ctx *c);
int ctx_fail(tfim_ctx *c, int code, const char *msg);

}  // namespace tfim
