/*
 * oracle.h -- CPU oracle for exact maximal-frequent-itemset (MFI) mining.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * bitmap engine's algorithm, used as the checker for the GPU library and
 * engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product (tensorfim_amd/) links, loads or
 * calls anything in this directory.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference checkout).
 */
#ifndef TFIM_ORACLE_H
#define TFIM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One parsed + preprocessed transaction database, in the reference's
 * process-global state laid out as a struct (kernel_bmma.cu:48-72,337-343). */
typedef struct oracle_db {
    /* raw statistics (readTransSetFile, kernel_bmma.cu:208-270) */
    int32_t trans_num;          /* transNum: counts empty lines (quirk A.1) */
    int32_t dimension;          /* distinct item ids present */
    int32_t item_id_min;        /* INT_MAX when the file holds no ids */
    int32_t item_id_max_raw;    /* raw max id before compaction */
    int64_t len_min, len_max;   /* transLengthMin/Max (size_t in the reference) */
    double len_mean;

    /* threshold (kernel_bmma.cu:1189) */
    double support;             /* relative support threshold (atof of argv[2]) */
    double thr_abs;             /* frequencyThreshold = transNum * support (double) */

    /* reduced statistics (kernel_bmma.cu:1193-1209) */
    int32_t trans_num_reduced;
    int32_t dimension_reduced;  /* #{ids : (double)freq/transNum > support} (strict) */
    int64_t red_len_min, red_len_max;
    double red_len_mean;

    /* dense namespace (buildBitmap, kernel_bmma.cu:388-426) */
    int32_t n_surv;             /* F = freqItemsNum: survivors with id >= 1 */
    int32_t d_count;            /* F + (item 0 frequent ? 1 : 0) */
    int32_t item_id_max;        /* d_count - 1, the value the Results header prints */
    int32_t *dense_to_orig;     /* [d_count] */
    int32_t *freq_dense;        /* [d_count] freqPerItem in dense ids */
    int32_t *sorted_items;      /* [d_count] items_inFreqSort (stable, freq desc) */
    int32_t n_roots;            /* min(dimension_reduced, d_count) */

    /* vertical bitmaps, LSB-first, row-major [d_count][words] */
    int64_t words;              /* ceil(trans_num_reduced / 64) */
    uint64_t *bitmap;
} oracle_db;

/* One found MFI: dense ids in insertion order (root first). */
typedef struct oracle_mfi {
    int32_t k;
    int32_t frequency;
    int32_t *items;
} oracle_mfi;

typedef struct oracle_result {
    int64_t n_mfi;
    oracle_mfi *mfi;
    int64_t nodes;              /* DFS nodes counted */
    double computation_time;    /* seconds, steady clock around the DFS */
    double preprocessing_time;  /* seconds, reduce + bitmap build */
} oracle_result;

/* Parse the transaction file (kernel_bmma.cu:208-270) and prepare it at the
 * given support (kernel_bmma.cu:1186-1215).  `replicate` >= 1 gives the
 * database data/make_replica.py would write (file + '\n' if missing,
 * concatenated `replicate` times).  Returns 0 on success, -1 if the file
 * cannot be read. */
int oracle_load(const char *path, double support, int replicate, oracle_db *db);

/* Build a database from an in-memory CSR transaction list (same semantics as
 * parsing a file whose lines hold these ids; negative ids are dropped). */
int oracle_load_rows(const int32_t *ids, const int64_t *row_off, int64_t n_rows,
                     double support, oracle_db *db);

void oracle_free(oracle_db *db);

/* getResFrequency (kernel_bmma.cu:446-464): res[e] = popcount(AND_x bv[x] & bv[e])
 * for every survivor e in [0, n_surv); returns #{e : res[e] >= thr_abs}. */
int32_t oracle_count_node(const oracle_db *db, const int32_t *items, int32_t k,
                          int32_t *res);

/* The reference DFS (kernel_bitmap.cu:841-894 worker pool; kernel_bmma.cu:874-910
 * expansion rule).  `threads` workers share one LIFO stack. */
int oracle_mine(const oracle_db *db, int threads, oracle_result *out);

/* Same DFS, limited to the first `max_nodes` popped nodes (0 = unlimited);
 * returns the number of nodes processed.  Used for bounded CPU timing. */
int64_t oracle_mine_bounded(const oracle_db *db, int threads, int64_t max_nodes,
                            oracle_result *out);

void oracle_result_free(oracle_result *r);

/* The worker's per-node step of the same DFS (kernel_bitmap.cu:853-894:
 * getResFrequency, the maximality test and the scan for children) on a given
 * list of nodes (CSR: node i = items[off[i] .. off[i+1]), dense ids in
 * insertion order, its progress being the last item), `threads` workers
 * taking nodes from a shared index.  Returns the seconds of the whole pass
 * (steady clock); used to time a random sample of a run's nodes. */
double oracle_time_nodes(const oracle_db *db, const int32_t *items, const int64_t *off, int64_t n, int threads);

/* Sort MFIs by frequency desc (ties: item list, for a deterministic file) and
 * write the Results file in the reference layout (kernel_bmma.cu:1034-1073). */
int oracle_write_results(const oracle_db *db, oracle_result *r, const char *path,
                         const char *trans_set_file, const char *engine_line);

/* Plain AND-popcount matrix (bmma_bench.cu:180-191 cpu_ref):
 * C[i*ldc + j] = sum_w popcount(A[i*words + w] & B[j*words + w]). */
void oracle_popcount_matrix(const uint64_t *A, const uint64_t *B, int32_t *C,
                            int64_t b, int64_t d, int64_t words, int64_t ldc,
                            int threads);

/* Counter-based Bernoulli bitmap generator (SplitMix64 keyed by seed, matrix,
 * row, bit-pair).  Fills `words` u64 words of row `row` of matrix `matrix`;
 * bits at positions >= n_bits are zero.  Threshold is compared as u32 < thr. */
void oracle_bernoulli_row(uint64_t seed, uint32_t matrix, uint64_t row,
                          int64_t n_bits, int64_t words, uint32_t thr32,
                          uint64_t *out);

/* batch_proto.cu:193-230 generator restated: B is [512][2^23/64] u64 rows.
 * Returns the planted pattern lists (80 patterns, <= 6 items, sorted) in
 * pat_items[80*6] with lengths in pat_len[80]. */
void oracle_proto_gen(uint64_t *B, int32_t *pat_items, int32_t *pat_len);

/* batch_proto.cu:287-374 level-wise search (canonical j > last extension,
 * minsup = (int)(N * 0.15 / 100)); writes per-level found counts into
 * level_found[0..7] (index = level-1; level 1 = F1) and returns the total. */
int64_t oracle_proto_levelwise(const uint64_t *B, int threads, int64_t *level_found);

/* Search-policy MLP (mlp_oracle.c): loss and gradients in double. */
int oracle_mlp_nparam(void);
double oracle_mlp_loss_grad(const double *P, const float *X, const float *Y, const float *M, int n,
                            double *logits, double *grads);

#ifdef __cplusplus
}
#endif
#endif
