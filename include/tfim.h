/*
 * tfim.h -- C ABI of the MI355X support-counting library (libtfim.so).
 *
 * The reference has no library API: its counting path sits behind internal
 * C++ functions reading process globals.  Each entry point below replaces one
 * of them and is cited.  Conventions:
 *   - every call returns an int status (TFIM_OK == 0) and never exits;
 *     tfim_last_error() holds the message of the last failing call;
 *   - no hidden globals: all state lives in a tfim_ctx (one per engine; one
 *     HIP stream per ctx, default priority);
 *   - plain pointers and sizes only.  Host pointers are host memory unless a
 *     parameter is a tfim_buf (device memory owned by the library).
 *   - counts are exact int32 supports: out = popcount(AND of rows).
 *
 * Bitmap layout (identical to the reference's buildBitmap, kernel_bmma.cu:
 * 388-426): one row per dense item id, row-major, LSB-first (transaction t is
 * bit t&63 of u64 word t>>6), zero padding beyond n_bits.
 */
#ifndef TFIM_H
#define TFIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tfim_ctx tfim_ctx;
typedef struct tfim_buf tfim_buf;

enum {
    TFIM_OK = 0,
    TFIM_E_INVALID = -1,   /* bad argument */
    TFIM_E_HIP = -2,       /* HIP runtime error */
    TFIM_E_NOMEM = -3,     /* device allocation failed */
    TFIM_E_STATE = -4,     /* call made before the state it needs */
    TFIM_E_RANGE = -5      /* size beyond a documented limit */
};

/* Counting paths. */
enum {
    TFIM_PATH_AUTO = 0,     /* library chooses (autotuned per shape) */
    TFIM_PATH_PACKED = 1,   /* 128-bit loads, AND, v_bcnt, wave reductions */
    TFIM_PATH_MFMA_I8 = 2,  /* bits unpacked to 0/1 int8 in LDS, v_mfma_i32_*_i8 */
    TFIM_PATH_MFMA_FP4 = 3  /* matrix count only: bits as exact e2m1 values (0, 1.0,
                               2.0), v_mfma_f32_32x32x64_f8f6f4, twice the int8
                               rate; fp32 sums stay exact (at most 2^25 per split) */
};

/* ---- context ---------------------------------------------------------- */

/* Opens device `device` (HIP ordinal).  Replaces the implicit CUDA context and
 * the thread_local streams of kernel_bitmap.cu:447-449. */
int tfim_open(int device, tfim_ctx **out);
void tfim_close(tfim_ctx *ctx);
/* A second context on the same device with its own stream and scratch that
 * shares `parent`'s resident bitmaps (read-only; parent must outlive it).
 * One per host worker replaces the per-thread streams and buffers of the
 * reference's worker pool (kernel_bitmap.cu:447-449,853-894). */
int tfim_open_shared(tfim_ctx *parent, tfim_ctx **out);
/* Message of the last failing call on this ctx ("" if none). */
const char *tfim_last_error(const tfim_ctx *ctx);
/* Device facts for the Results header (replaces displayWorkingEnvironment,
 * kernel_bmma.cu:166-206). */
int tfim_device_info(tfim_ctx *ctx, char *name, int name_len, int32_t *compute_units,
                     int64_t *hbm_bytes, int32_t *clock_mhz);

/* ---- resident item bitmaps -------------------------------------------- */

/* Uploads the vertical bitmaps once; replaces uploadBitmapToGPU
 * (kernel_bmma.cu:428-444, kernel_bitmap.cu:382-388).  `rows` is host memory,
 * n_items rows of words_per_row u64 words.  The first n_cols rows are the
 * survivors counted by every count call (freqItems, kernel_bmma.cu:394-418);
 * rows n_cols..n_items-1 (item 0's row, quirk A.3) may appear in prefixes
 * only.  The device copy is padded to a 2048-bit multiple per row. */
int tfim_load_bitmaps(tfim_ctx *ctx, const uint64_t *rows, int32_t n_items,
                      int64_t n_bits, int64_t words_per_row, int32_t n_cols);

/* Like tfim_load_bitmaps, for a database that is `replicate` concatenated
 * copies of a base of `n_bits` transactions (data/make_replica.py): row bits
 * are repeated on the device, so the host holds only the base. */
int tfim_load_bitmaps_replicated(tfim_ctx *ctx, const uint64_t *rows, int32_t n_items,
                                 int64_t n_bits, int64_t words_per_row, int32_t n_cols,
                                 int32_t replicate);

/* F1 on the device: out[i] = popcount(row i) for i < n_items
 * (batch_proto.cu:57-65 item_count_kernel). */
int tfim_item_supports(tfim_ctx *ctx, int32_t *out);

/* ---- counting ---------------------------------------------------------- */

/* Plain per-node path; replaces getResFrequencyCuda (kernel_bitmap.cu:451-536)
 * and its kernel (:412-442).  out[e] = popcount(AND_i row[items[i]] & row[e])
 * for every survivor e < n_cols; synchronous, `out` is host memory. */
int tfim_count_node(tfim_ctx *ctx, const int32_t *items, int32_t k, int32_t *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tfim_count_node without any allocation.  items is host
 * memory (1 <= k <= TFIM_NOALLOC_MAX_ITEMS ids, checked here and passed to
 * the kernel by value); out is DEVICE memory owned by the caller, n_cols
 * int32, written asynchronously on the ctx stream (tfim_sync waits).  The
 * kernel's partial counts live in a fixed per-device scratch array of the
 * library (8 x TFIM_NOALLOC_MAX_COLS int32 = 256 KiB, kept until the process
 * ends); more than TFIM_NOALLOC_MAX_COLS survivors is TFIM_E_RANGE.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (other
 * contexts or streams) race; calls on one ctx are ordered by its stream. */
#define TFIM_NOALLOC_MAX_COLS 8192
#define TFIM_NOALLOC_MAX_ITEMS 60
int tfim_count_node_noalloc(tfim_ctx *ctx, const int32_t *items, int32_t k, int32_t *out);

/* Times `reps` back-to-back launches of tfim_count_node's kernel for one node
 * (after `warmup` untimed ones) with HIP events on the ctx stream; avg_ms =
 * mean per launch.  Measurement only: no counts are returned.  Stands beside
 * the reference's per-node timing of getResFrequencyCuda (kernel_bitmap.cu:
 * 451-536), which it times without the copies and the synchronisation. */
int tfim_time_count_node(tfim_ctx *ctx, const int32_t *items, int32_t k, int32_t warmup, int32_t reps,
                         double *avg_ms);

/* Batched path; replaces batchCountBMMA (kernel_bmma.cu:772-872), including
 * its mask build (:619-634), count (:636-758) and split-k reduce (:761-768),
 * as one fused launch.  Row r has prefix items item_ids[item_off[r] ..
 * item_off[r+1]) (CSR; uneven lengths allowed, k >= 1).  If col_ids is NULL,
 * row r is counted against all survivors and out is [n_rows][n_cols];
 * otherwise row r is counted against its tail col_ids[col_off[r] ..
 * col_off[r+1]) and out is CSR-aligned with col_ids.  Host memory in and out;
 * synchronous.  path: TFIM_PATH_AUTO lets the library choose the kernel per
 * wave (its tuners); TFIM_PATH_PACKED keeps the wave on the packed (AND +
 * v_bcnt) kernels; TFIM_PATH_MFMA_I8 counts it on the matrix cores (the fused
 * kernel; the three-launch form -- prefix masks, fp4 matrix count with
 * split-k, gather -- for waves outside its limits; both count on the fp4
 * matrix cores, whose sums are the exact integer counts). */
int tfim_count_nodes(tfim_ctx *ctx, const int32_t *item_ids, const int32_t *item_off,
                     int32_t n_rows, const int32_t *col_ids, const int32_t *col_off,
                     int32_t *out, int32_t path);

/* The same count split in two, so the host can prepare or process one wave
 * while the GPU counts another: submit packs and launches the wave
 * asynchronously on the ctx stream (every step a kernel: the descriptor is
 * read from pinned host memory by the wave's first kernel and the counts are
 * written to pinned host memory by its last) and returns a ticket; wait
 * blocks until
 * that wave's counts are done and copies them into `out` (layout as in
 * tfim_count_nodes).  At most four waves may be in flight per ctx; tickets
 * are waited in submission order.  A ticket of -1 (nothing to count) waits
 * trivially.  One thread may call submit while another waits on an earlier
 * ticket of the same ctx; no other calls on a ctx may overlap. */
int tfim_count_nodes_submit(tfim_ctx *ctx, const int32_t *item_ids, const int32_t *item_off,
                            int32_t n_rows, const int32_t *col_ids, const int32_t *col_off,
                            int32_t path, int64_t *ticket);
int tfim_count_nodes_wait(tfim_ctx *ctx, int64_t ticket, int32_t *out);
/* Batch count with rows given as bit masks, for databases of at most 64
 * dense ids (n_items <= 64 in tfim_load_bitmaps): row r counts the prefix of
 * the ids set in prefix_mask[r] (at least one) against the columns set in
 * col_mask[r] (survivor ids < n_cols) in ascending id order; the outputs are
 * CSR by popcount(col_mask[r]) (row r's first output follows every earlier
 * row's columns).  The same counts as tfim_count_nodes[_submit] with item and
 * column lists, but each row is two words: the host neither builds nor
 * checks id lists, and the kernel finds the ids with scalar bit scans.
 * Waited on with tfim_count_nodes_wait (the tickets are shared with
 * tfim_count_nodes_submit); path as there.  Replaces batchCountBMMA
 * (kernel_bmma.cu:772) for narrow databases; TFIM_E_RANGE if more than 64
 * ids are loaded. */
int tfim_count_masks_submit(tfim_ctx *ctx, const uint64_t *prefix_mask, const uint64_t *col_mask,
                            int32_t n_rows, int32_t path, int64_t *ticket);
/* Blocking form of tfim_count_masks_submit + tfim_count_nodes_wait. */
int tfim_count_masks(tfim_ctx *ctx, const uint64_t *prefix_mask, const uint64_t *col_mask, int32_t n_rows,
                     int32_t *out, int32_t path);
/* Complement rows, for maximality checks over many survivors.  As
 * tfim_count_nodes_submit with col_off required (col_ids may be NULL when
 * col_off[n_rows] == 0), plus n_compl complement rows: compl_rows lists their
 * row indices in ascending order, each with an empty column range and a
 * first prefix item x0 < n_cols, and excl holds one exclusion bitmap of
 * ceil(n_cols / 64) u64 words per complement row (bit e of word e >> 6 set =
 * survivor e excluded).  Complement row j is counted on the device against
 * every survivor e whose bit is clear, and the wave returns for it two ints
 * after the explicit counts (out[col_off[n_rows] + 2j ...]):
 *   - popcount(AND of the prefix rows), the support of the prefix;
 *   - #{e != x0 : bit e clear, popcount(AND prefix & row e) >= min_count}.
 * The reference's maximality test counts every survivor and compares the
 * number of frequent ones on the host (kernel_bmma.cu:857-877,
 * kernel_bitmap.cu:861-866); here the threshold step runs on the device and
 * only the two numbers travel.  tfim_count_nodes_wait returns
 * col_off[n_rows] + 2 * n_compl ints.  tfim_count_nodes_ex is the blocking
 * form. */
int tfim_count_nodes_ex_submit(tfim_ctx *ctx, const int32_t *item_ids, const int32_t *item_off,
                               int32_t n_rows, const int32_t *col_ids, const int32_t *col_off,
                               const int32_t *compl_rows, int32_t n_compl, const uint64_t *excl,
                               int32_t min_count, int32_t path, int64_t *ticket);
int tfim_count_nodes_ex(tfim_ctx *ctx, const int32_t *item_ids, const int32_t *item_off,
                        int32_t n_rows, const int32_t *col_ids, const int32_t *col_off,
                        const int32_t *compl_rows, int32_t n_compl, const uint64_t *excl,
                        int32_t min_count, int32_t *out, int32_t path);

/* Tuning knobs (ablation and autotuning; the reference's equivalents are
 * environment variables, kernel_bmma.cu:351-368):
 *   "node_spread"  per-node counts: 0 auto (8 copies for at most 512
 *                  survivors), 1 one copy, 8 spread over 8 copies 4 KB
 *                  apart (the same-address atomics of one column then come
 *                  from 1/8 of the word chunks; the library adds the copies)
 *   "node_shape"   launch shape of the per-node kernel (tfim_count_node):
 *                  0 auto, else R * 100 + cpb with R in {1, 2, 4} 128-bit
 *                  words per lane and cpb in {8, 16, 32, 64} columns per
 *                  block
 *   "wave_kernel"  0 auto, 1 L2-resident kernel, 2 LDS-staged kernel (two
 *                  workgroups per CU), 3 LDS-staged with wide staging (one
 *                  workgroup per CU, up to 512 bits per lane per column),
 *                  4 matrix cores in three launches (prefix masks, fp4 matrix
 *                  count of every (row, survivor) pair, gather; waves of up
 *                  to 8 GiB of masks), 5 fused matrix-core kernel (one
 *                  launch: prefix masks built in LDS from the wave's staged
 *                  item words, fp4 matrix count, split-k reduced in the
 *                  epilogue; waves of at most 256 distinct prefix items).
 *                  In auto mode the wave tuner also tries the fused kernel on
 *                  waves the LDS kernel takes, and waves outside the LDS
 *                  kernel go to the L2 kernel or the three-launch
 *                  matrix-core form by a second tuner that predicts each
 *                  one's time from its measured cost per unit of work (the
 *                  fused kernel there only when forced)
 *   "tile_cols"    columns per work tile of the batched path: 8..64, or 0
 *                  (default) to let the wave tuner choose
 *   "mfma_tile"    workgroup tile of the int8-MFMA matrix count: 0 auto
 *                  (= 256), 256 (256 x 128 outputs, operand bytes expanded
 *                  in registers from packed bits in LDS) or 128 (128 x 128,
 *                  bytes expanded into LDS)
 *   "mfma_format"  matrix-core waves (wave_kernel 4 and 5 and their automatic
 *                  use): 0 (default) the fp4 matrix cores, each bit an exact
 *                  e2m1 value; 1 the int8 matrix cores (bits as 0/1 bytes,
 *                  int32 sums).  Both give the same exact counts.
 *   "mfma_split"   matrix-core waves: 0 (default) split the word axis by an
 *                  occupancy rule scaled by a tuner; 1 no split, one word
 *                  range per tile (the reference's BMMA_SPLITK_OFF ablation,
 *                  kernel_bmma.cu:351-368)
 *   "graphs"       1: each wave's kernels are replayed from an instantiated
 *                  HIP graph (descriptor stage, count kernels, output drain)
 *                  whose node parameters are updated per wave; 0 (default):
 *                  the same kernels launched one by one (on MI355X cheaper
 *                  on the host, 12 against 22 us per wave, and without the
 *                  graph's gaps between nodes on the device)
 * Unknown keys and out-of-range values return TFIM_E_INVALID. */
int tfim_config(tfim_ctx *ctx, const char *key, int64_t value);

/* Statistics of the batched path since tfim_open: launches, how many used
 * the LDS-staged kernel, rows, summed kernel milliseconds (HIP events around
 * a wave's count kernels; a mask wave's kernel stamps its own span, first
 * workgroup start to last workgroup end, on the device wall clock).  In
 * TFIM_PATH_AUTO the library's wave tuner picks, for waves that fit both
 * kernels, the kernel and the columns per work tile (L2 kernel with 32; LDS
 * kernel with 16, 32 or 64, narrow or wide staging): each arm is timed
 * three times, then the cheapest
 * per row-word of work is kept and another arm re-timed every 16 waves. */
int tfim_wave_stats(tfim_ctx *ctx, int64_t *launches, int64_t *lds_launches,
                    int64_t *rows, double *kernel_ms);

/* Waves since tfim_open by kernel: the L2-resident kernel, the LDS-staged
 * kernel, the matrix-core path in three launches (prefix masks, fp4 matrix
 * count of every (row, survivor) pair with split-k over word ranges, gather
 * of the wave's outputs) and the fused matrix-core kernel (the same in one
 * launch, masks built in LDS, split-k reduced in the epilogue) -- the forms
 * of the reference's batchCountBMMA dispatch (naive / smem / split-k,
 * kernel_bmma.cu:772-872) on MI355X. */
int tfim_wave_path_stats(tfim_ctx *ctx, int64_t *l2, int64_t *lds, int64_t *mfma, int64_t *fused);
/* Batched-path graphs instantiated and graph launches since tfim_open.  The
 * reference launches each wave directly (runWaveEngine, kernel_bmma.cu:896-910);
 * no counterpart. */
int tfim_wave_graph_stats(tfim_ctx *ctx, int64_t *graphs, int64_t *launches);

/* ---- device buffers and the isolated count (microbench) ---------------- */

int tfim_buf_alloc(tfim_ctx *ctx, int64_t bytes, tfim_buf **out);
void tfim_buf_free(tfim_ctx *ctx, tfim_buf *buf);
int tfim_buf_write(tfim_ctx *ctx, tfim_buf *buf, int64_t offset, const void *src, int64_t bytes);
int tfim_buf_read(tfim_ctx *ctx, const tfim_buf *buf, int64_t offset, void *dst, int64_t bytes);

/* Row pitch in u64 words the matrix kernels expect for n_bits-bit rows:
 * ceil(n_bits / 2048) * 32 (the reference's 2048-bit padding,
 * kernel_bmma.cu:389). */
int64_t tfim_pitch_words(int64_t n_bits);

/* Fills `rows` rows of `pitch_words` u64 words with i.i.d. Bernoulli bits:
 * bit = (u32 draw < thr32), draws from SplitMix64 keyed by (seed, matrix,
 * row, bit pair); bits >= n_bits are zero. */
int tfim_fill_bernoulli(tfim_ctx *ctx, tfim_buf *dst, int64_t rows, int64_t n_bits,
                        int64_t pitch_words, uint64_t seed, uint32_t matrix, uint32_t thr32);

/* C[i][j] = popcount(A_i AND B_j) over pitch_words words, int32, row-major
 * [b][d] in C.  Replaces bmma_kernel / bmma_smem_kernel / bitmap_kernel of
 * code/microbench/bmma_bench.cu:50-174.  split_k >= 1 splits the word axis;
 * 0 lets the library's tuner choose the split and (int8-MFMA path, unless
 * "mfma_tile" is configured) the tile: the first call for a (path, b, d,
 * pitch) shape times every candidate once on these buffers and blocks; the
 * choice is kept for the shape.  Otherwise asynchronous on the ctx stream. */
int tfim_popcount_matrix(tfim_ctx *ctx, const tfim_buf *A, const tfim_buf *B, tfim_buf *C,
                         int64_t b, int64_t d, int64_t pitch_words, int32_t path,
                         int32_t split_k);

/* The tile (64 for the packed path) and split_k that split_k == 0 selects for
 * this shape, tuning first if needed.  No counterpart in the reference (its
 * split factor is fixed, kernel_bmma.cu:715-758). */
int tfim_popcount_matrix_plan(tfim_ctx *ctx, const tfim_buf *A, const tfim_buf *B, tfim_buf *C,
                              int64_t b, int64_t d, int64_t pitch_words, int32_t path,
                              int32_t *tile, int32_t *split_k);

/* Times `reps` back-to-back tfim_popcount_matrix launches after `warmup`
 * untimed ones with HIP events on the ctx stream; avg_ms = mean per launch. */
int tfim_time_popcount_matrix(tfim_ctx *ctx, const tfim_buf *A, const tfim_buf *B, tfim_buf *C,
                              int64_t b, int64_t d, int64_t pitch_words, int32_t path,
                              int32_t split_k, int32_t warmup, int32_t reps, double *avg_ms);

/* ---- device preprocessing ---------------------------------------------- */
/* The reference's frequency count, reduction and bitmap build
 * (kernel_bmma.cu:272-276, :1075-1086, buildBitmap :388-426) on the device,
 * for a database parsed into CSR rows (ids[row_off[t] .. row_off[t+1]) of
 * transaction t, ids >= 0).
 *   tfim_prep_begin  uploads the rows and returns freq[0..id_max], the
 *                    occurrence count of every id; an id outside [0, id_max]
 *                    is an error.
 *   tfim_prep_finish keeps item x as dense row dense_of[x] (-1 = dropped),
 *                    drops the rows left empty and makes the result the
 *                    context's resident bitmaps, exactly as tfim_load_bitmaps
 *                    of the host-built rows would (n_items rows, the first
 *                    n_cols of them columns).  Returns the kept row count and
 *                    the min, max and sum of kept row lengths, and
 *                    *duplicates = 1 if some transaction repeats a kept id.
 * The uploaded rows are released by tfim_prep_finish. */
int tfim_prep_begin(tfim_ctx *ctx, const int32_t *ids, int64_t n_ids, const int64_t *row_off, int64_t n_rows,
                    int32_t id_max, int32_t *freq);
int tfim_prep_finish(tfim_ctx *ctx, const int32_t *dense_of, int32_t n_items, int32_t n_cols,
                     int64_t *rows_reduced, int64_t *len_min, int64_t *len_max, int64_t *len_sum,
                     int32_t *duplicates);

/* ---- learned search policy (MLP 14 -> 64 -> 64 -> 2) -------------------- */
/* No counterpart in the reference (its search learns nothing).  The MAFIA
 * engine's tail ranker (head 1) and head-union-tail lookahead gate (head 0).
 * Parameters: W1[64][14] b1[64] W2[64][64] b2[64] W3[2][64] b3[2], fp32;
 * inference may read a bf16 copy.  Loss: mean over masked outputs of binary
 * cross-entropy of sigmoid(logit) against a 0/1 label; Adam updates. */
#define TFIM_MLP_INPUTS 14
#define TFIM_MLP_OUTPUTS 2
#define TFIM_MLP_PARAMS 5250
typedef struct tfim_mlp tfim_mlp;

/* He-uniform weights from SplitMix64(seed), zero biases; the device scratch
 * and pinned staging for batches and requests of up to 8,192 rows are
 * allocated here, so training and inference on such batches allocate
 * nothing. */
int tfim_mlp_create(tfim_ctx *ctx, uint64_t seed, tfim_mlp **out);
void tfim_mlp_destroy(tfim_mlp *mlp);
int tfim_mlp_get_params(tfim_mlp *mlp, float *params);
int tfim_mlp_set_params(tfim_mlp *mlp, const float *params);
/* logits[n][2] for x[n][14]; bf16 != 0 uses the bf16 weight copy. */
int tfim_mlp_forward(tfim_mlp *mlp, const float *x, int32_t n, float *logits, int32_t bf16);
/* The same forward split in two: submit copies x (n >= 1) and queues the
 * inference on the model's stream, returning at once; wait blocks until it
 * has finished and writes logits[n][2].  One request may be outstanding;
 * training steps queued before or after it run in submission order. */
int tfim_mlp_forward_submit(tfim_mlp *mlp, const float *x, int32_t n, int32_t bf16);
int tfim_mlp_forward_wait(tfim_mlp *mlp, float *logits);
/* Gradient of the loss w.r.t. all parameters (no update). */
int tfim_mlp_gradients(tfim_mlp *mlp, const float *x, const float *labels, const float *mask, int32_t n,
                       float *grads, float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch.  The MLP
 * works on a stream of its own; with loss == NULL the step is queued and the
 * call returns once the inputs are copied (later MLP calls order after it). */
int tfim_mlp_train_step(tfim_mlp *mlp, const float *x, const float *labels, const float *mask, int32_t n,
                        float lr, float *loss);
/* `steps` such Adam steps on one batch, staged to the device once; with loss
 * != NULL it waits and returns the loss the last step saw. */
int tfim_mlp_train(tfim_mlp *mlp, const float *x, const float *labels, const float *mask, int32_t n, float lr,
                   int32_t steps, float *loss);

/* Waits for all work on the ctx stream. */
int tfim_sync(tfim_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
