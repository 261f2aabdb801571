/*
 * oracle.c -- CPU oracle (test infrastructure only; see oracle.h).
 *
 * Restates, in plain C, the reference bitmap engine:
 *   reader          kernel_bmma.cu:208-270  (== kernel_bitmap.cu:205-267)
 *   item frequency  kernel_bmma.cu:272-301
 *   threshold/roots kernel_bmma.cu:1189,1199,1224-1233
 *   reduction       kernel_bmma.cu:1075-1086
 *   bitmap build    kernel_bmma.cu:388-426  (word count as kernel_bitmap.cu:343)
 *   counter         kernel_bmma.cu:446-464
 *   DFS             kernel_bitmap.cu:853-894, expandNode kernel_bmma.cu:875-890
 *   output          kernel_bmma.cu:1034-1073,1284-1293
 *   cpu_ref         code/microbench/bmma_bench.cu:180-191
 *   proto generator code/batch_proto/batch_proto.cu:193-230, level loop :287-374
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* small growable int vector                                                 */
typedef struct { int32_t *v; int64_t n, cap; } ivec;
static void iv_push(ivec *a, int32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 64;
        a->v = (int32_t *)realloc(a->v, (size_t)a->cap * sizeof(int32_t));
    }
    a->v[a->n++] = x;
}
typedef struct { int64_t *v; int64_t n, cap; } lvec;
static void lv_push(lvec *a, int64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 64;
        a->v = (int64_t *)realloc(a->v, (size_t)a->cap * sizeof(int64_t));
    }
    a->v[a->n++] = x;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Reader: kernel_bmma.cu:208-233.  '\n' ends a row (empty rows included);
 * a digit run is one id; '-' followed by digits is dropped; every other byte
 * separates.  A final row without '\n' is kept only if non-empty. */
static void parse_blob(const char *p, const char *end, ivec *ids, lvec *off) {
    lv_push(off, 0);
    int64_t row_start = ids->n;
    while (p < end) {
        char c = *p;
        if (c == '\n') { lv_push(off, ids->n); row_start = ids->n; p++; continue; }
        if (c >= '0' && c <= '9') {
            long v = 0;
            while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
            iv_push(ids, (int32_t)v);
            continue;
        }
        if (c == '-') {
            p++;
            while (p < end && *p >= '0' && *p <= '9') p++;
            continue;
        }
        p++;
    }
    if (ids->n > row_start) lv_push(off, ids->n);
}

static int prepare(oracle_db *db, int32_t *ids, int64_t *off, int64_t n_rows, double support);

int oracle_load(const char *path, double support, int replicate, oracle_db *db) {
    memset(db, 0, sizeof(*db));
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    fseeko(fp, 0, SEEK_END);
    long long sz = ftello(fp);
    fseeko(fp, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)sz + 2);
    if (sz > 0 && fread(buf, 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); free(buf); return -1; }
    fclose(fp);
    if (replicate >= 1 && (sz == 0 || buf[sz - 1] != '\n')) buf[sz++] = '\n';  /* make_replica.py:26-27 */
    ivec ids = {0}; lvec off = {0};
    parse_blob(buf, buf + sz, &ids, &off);
    free(buf);
    int64_t n_rows = off.n - 1;
    if (replicate > 1) {
        int64_t base_ids = ids.n;
        for (int r = 1; r < replicate; r++) {
            for (int64_t i = 0; i < base_ids; i++) iv_push(&ids, ids.v[i]);
            for (int64_t t = 1; t <= n_rows; t++) lv_push(&off, off.v[t] + (int64_t)r * base_ids);
        }
        n_rows *= replicate;
    }
    int rc = prepare(db, ids.v, off.v, n_rows, support);
    free(ids.v); free(off.v);
    return rc;
}

int oracle_load_rows(const int32_t *ids_in, const int64_t *row_off, int64_t n_rows,
                     double support, oracle_db *db) {
    memset(db, 0, sizeof(*db));
    /* drop negative ids, as the parser does */
    int64_t total = row_off[n_rows];
    int32_t *ids = (int32_t *)malloc((size_t)(total + 1) * sizeof(int32_t));
    int64_t *off = (int64_t *)malloc((size_t)(n_rows + 1) * sizeof(int64_t));
    int64_t n = 0;
    off[0] = 0;
    for (int64_t t = 0; t < n_rows; t++) {
        for (int64_t j = row_off[t]; j < row_off[t + 1]; j++)
            if (ids_in[j] >= 0) ids[n++] = ids_in[j];
        off[t + 1] = n;
    }
    int rc = prepare(db, ids, off, n_rows, support);
    free(ids); free(off);
    return rc;
}

/* stable sort of ids by freq desc (std::stable_sort, kernel_bmma.cu:411-414) */
static const int32_t *g_sort_freq;
static int cmp_freq_desc(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    int32_t fx = g_sort_freq[x], fy = g_sort_freq[y];
    if (fx != fy) return fx > fy ? -1 : 1;
    return x < y ? -1 : (x > y);   /* ties by ascending id == stability over iota */
}

static int prepare(oracle_db *db, int32_t *ids, int64_t *off, int64_t n_rows, double support) {
    db->support = support;
    /* statistics (kernel_bmma.cu:235-265) */
    int32_t idmin = INT_MAX, idmax = 0;
    int64_t lmin = INT_MAX, lmax = 0;
    double lmean = 0;
    for (int64_t t = 0; t < n_rows; t++) {
        for (int64_t j = off[t]; j < off[t + 1]; j++) {
            if (ids[j] > idmax) idmax = ids[j];
            if (ids[j] < idmin) idmin = ids[j];
        }
        int64_t len = off[t + 1] - off[t];
        if (len < lmin) lmin = len;
        if (len > lmax) lmax = len;
        lmean += (double)len;
    }
    if (n_rows > 0) lmean /= (double)n_rows;
    db->trans_num = (int32_t)n_rows;
    db->item_id_min = idmin;
    db->item_id_max_raw = idmax;
    db->len_min = lmin; db->len_max = lmax; db->len_mean = lmean;

    /* getItemsFrequenc (kernel_bmma.cu:272-276): occurrence counts */
    int32_t *freq = (int32_t *)calloc((size_t)idmax + 1, sizeof(int32_t));
    int32_t dim = 0;
    for (int64_t j = 0; j < off[n_rows]; j++) {
        if (freq[ids[j]] == 0) dim++;
        freq[ids[j]]++;
    }
    db->dimension = dim;

    /* threshold (kernel_bmma.cu:1189) and reduction (:1075-1086) */
    double thr = (double)db->trans_num * support;
    db->thr_abs = thr;
    double t0 = now_s();
    int64_t *roff = (int64_t *)malloc((size_t)(n_rows + 1) * sizeof(int64_t));
    int32_t *rids = (int32_t *)malloc((size_t)(off[n_rows] + 1) * sizeof(int32_t));
    int64_t nr = 0, m = 0;
    roff[0] = 0;
    for (int64_t t = 0; t < n_rows; t++) {
        int64_t m0 = m;
        for (int64_t j = off[t]; j < off[t + 1]; j++)
            if (!((double)freq[ids[j]] < thr)) rids[m++] = ids[j];
        if (m > m0) roff[++nr] = m;   /* empty rows removed after reduction */
    }
    db->trans_num_reduced = (int32_t)nr;

    /* dimensionReduced: strict relative test over every raw id (:1199) */
    int32_t dred = 0;
    for (int32_t x = 0; x <= idmax; x++)
        if ((double)freq[x] / (double)db->trans_num > support) dred++;
    db->dimension_reduced = dred;

    int64_t rmin = INT_MAX, rmax = 0;
    double rmean = 0;
    for (int64_t t = 0; t < nr; t++) {
        int64_t len = roff[t + 1] - roff[t];
        if (len < rmin) rmin = len;
        if (len > rmax) rmax = len;
        rmean += (double)len;
    }
    rmean /= (double)nr;   /* NaN when nothing survives, as in the reference */
    db->red_len_min = rmin; db->red_len_max = rmax; db->red_len_mean = rmean;

    /* buildBitmap (kernel_bmma.cu:388-426) */
    int32_t F = 0;
    for (int32_t i = 1; i <= idmax; i++) if (!((double)freq[i] < thr)) F++;
    int zero_freq = !((double)freq[0] < thr);
    int32_t dcount = F + (zero_freq ? 1 : 0);
    db->n_surv = F;
    db->d_count = dcount;
    db->item_id_max = dcount - 1;
    db->dense_to_orig = (int32_t *)malloc((size_t)(dcount + 1) * sizeof(int32_t));
    db->freq_dense = (int32_t *)calloc((size_t)dcount + 1, sizeof(int32_t));
    int32_t *o2d = (int32_t *)malloc(((size_t)idmax + 1) * sizeof(int32_t));
    for (int32_t i = 0; i <= idmax; i++) o2d[i] = -1;
    int32_t k = 0;
    for (int32_t i = 1; i <= idmax; i++)
        if (!((double)freq[i] < thr)) { db->dense_to_orig[k] = i; db->freq_dense[k] = freq[i]; o2d[i] = k++; }
    if (zero_freq) { o2d[0] = F; db->dense_to_orig[F] = 0; db->freq_dense[F] = freq[0]; }

    /* items_inFreqSort over dCount+1 entries (freqDense carries one zero slot) */
    int32_t *srt = (int32_t *)malloc((size_t)(dcount + 1) * sizeof(int32_t));
    for (int32_t i = 0; i <= dcount; i++) srt[i] = i;
    g_sort_freq = db->freq_dense;
    qsort(srt, (size_t)dcount + 1, sizeof(int32_t), cmp_freq_desc);
    db->sorted_items = srt;
    /* The reference pushes items_inFreqSort[0..dimensionReduced) as roots
     * (:1224-1233); past d_count it would index the zero slot's missing bitmap
     * row, so the count is clamped to real items. */
    db->n_roots = dred < dcount ? dred : dcount;

    int64_t W = ((int64_t)nr + 63) / 64;
    db->words = W;
    db->bitmap = (uint64_t *)calloc((size_t)(dcount > 0 ? dcount : 1) * (size_t)(W > 0 ? W : 1), sizeof(uint64_t));
    for (int64_t t = 0; t < nr; t++)
        for (int64_t j = roff[t]; j < roff[t + 1]; j++) {
            int32_t x = o2d[rids[j]];
            db->bitmap[(size_t)x * W + (t >> 6)] |= 1ull << (t & 63);
        }
    (void)t0;
    free(o2d); free(freq); free(roff); free(rids);
    return 0;
}

void oracle_free(oracle_db *db) {
    free(db->dense_to_orig); free(db->freq_dense); free(db->sorted_items); free(db->bitmap);
    memset(db, 0, sizeof(*db));
}

/* ------------------------------------------------------------------------ */
/* getResFrequency (kernel_bmma.cu:446-464) */
static int32_t count_node_buf(const oracle_db *db, const int32_t *items, int32_t k,
                              int32_t *res, uint64_t *mask) {
    const int64_t W = db->words;
    for (int64_t w = 0; w < W; w++) mask[w] = ~0ull;
    for (int32_t i = 0; i < k; i++) {
        const uint64_t *bv = db->bitmap + (size_t)items[i] * W;
        for (int64_t w = 0; w < W; w++) mask[w] &= bv[w];
    }
    int32_t cnt = 0;
    for (int32_t e = 0; e < db->n_surv; e++) {
        const uint64_t *bv = db->bitmap + (size_t)e * W;
        uint64_t s = 0;
        for (int64_t w = 0; w < W; w++) s += (uint64_t)__builtin_popcountll(mask[w] & bv[w]);
        res[e] = (int32_t)s;
        if (!((double)res[e] < db->thr_abs)) cnt++;
    }
    return cnt;
}

int32_t oracle_count_node(const oracle_db *db, const int32_t *items, int32_t k, int32_t *res) {
    uint64_t *mask = (uint64_t *)malloc((size_t)(db->words > 0 ? db->words : 1) * sizeof(uint64_t));
    int32_t c = count_node_buf(db, items, k, res, mask);
    free(mask);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Worker-pool DFS (kernel_bitmap.cu:841-894).  A node is (items in insertion
 * order, frequency, progress); roots are pushed in sorted order so the LIFO
 * pops the last root first, as the reference does. */
typedef struct node { int32_t k, freq, progress; int32_t *items; } node;

typedef struct {
    const oracle_db *db;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    node *stack; int64_t sn, scap;
    int active;
    int64_t popped, max_nodes;
    oracle_mfi *mfi; int64_t nm, mcap;
} dfs_state;

static void st_push(dfs_state *s, node nd) {
    if (s->sn == s->scap) {
        s->scap = s->scap ? s->scap * 2 : 1024;
        s->stack = (node *)realloc(s->stack, (size_t)s->scap * sizeof(node));
    }
    s->stack[s->sn++] = nd;
}

static void *dfs_worker(void *arg) {
    dfs_state *s = (dfs_state *)arg;
    const oracle_db *db = s->db;
    int32_t *res = (int32_t *)calloc((size_t)db->d_count + 2, sizeof(int32_t));
    uint64_t *mask = (uint64_t *)malloc((size_t)(db->words > 0 ? db->words : 1) * sizeof(uint64_t));
    node *kids = NULL; int64_t nk = 0, kcap = 0;
    for (;;) {
        pthread_mutex_lock(&s->mu);
        while (s->sn == 0 && s->active > 0) pthread_cond_wait(&s->cv, &s->mu);
        if (s->sn == 0 || (s->max_nodes > 0 && s->popped >= s->max_nodes)) {
            pthread_cond_broadcast(&s->cv);
            pthread_mutex_unlock(&s->mu);
            break;
        }
        node nd = s->stack[--s->sn];
        s->popped++;
        s->active++;
        pthread_mutex_unlock(&s->mu);

        memset(res, 0, ((size_t)db->d_count + 2) * sizeof(int32_t));
        int32_t cnt = count_node_buf(db, nd.items, nd.k, res, mask);
        nk = 0;
        int is_mfi = (cnt == nd.k);
        if (!is_mfi) {
            for (int32_t i = 0; i <= nd.progress - 1; i++) {
                if ((double)res[i] < db->thr_abs) continue;
                int outside = 1;
                for (int32_t j = 0; j < nd.k; j++) if (nd.items[j] == i) { outside = 0; break; }
                if (!outside) continue;
                node c;
                c.k = nd.k + 1; c.freq = res[i]; c.progress = i;
                c.items = (int32_t *)malloc((size_t)c.k * sizeof(int32_t));
                memcpy(c.items, nd.items, (size_t)nd.k * sizeof(int32_t));
                c.items[nd.k] = i;
                if (nk == kcap) { kcap = kcap ? kcap * 2 : 64; kids = (node *)realloc(kids, (size_t)kcap * sizeof(node)); }
                kids[nk++] = c;
            }
        }
        pthread_mutex_lock(&s->mu);
        if (is_mfi) {
            if (s->nm == s->mcap) { s->mcap = s->mcap ? s->mcap * 2 : 256; s->mfi = (oracle_mfi *)realloc(s->mfi, (size_t)s->mcap * sizeof(oracle_mfi)); }
            oracle_mfi m = { nd.k, nd.freq, nd.items };
            s->mfi[s->nm++] = m;
        } else {
            free(nd.items);
        }
        for (int64_t i = 0; i < nk; i++) st_push(s, kids[i]);
        s->active--;
        pthread_cond_broadcast(&s->cv);
        pthread_mutex_unlock(&s->mu);
    }
    free(kids); free(res); free(mask);
    return NULL;
}

int64_t oracle_mine_bounded(const oracle_db *db, int threads, int64_t max_nodes, oracle_result *out) {
    memset(out, 0, sizeof(*out));
    if (threads < 1) threads = 1;
    dfs_state s;
    memset(&s, 0, sizeof(s));
    s.db = db;
    s.max_nodes = max_nodes;
    pthread_mutex_init(&s.mu, NULL);
    pthread_cond_init(&s.cv, NULL);
    double t0 = now_s();
    for (int32_t i = 0; i < db->n_roots; i++) {
        node r;
        int32_t it = db->sorted_items[i];
        r.k = 1; r.freq = db->freq_dense[it]; r.progress = it;
        r.items = (int32_t *)malloc(sizeof(int32_t));
        r.items[0] = it;
        st_push(&s, r);
    }
    pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, dfs_worker, &s);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    out->computation_time = now_s() - t0;
    for (int64_t i = 0; i < s.sn; i++) free(s.stack[i].items);
    free(s.stack); free(th);
    out->n_mfi = s.nm;
    out->mfi = s.mfi;
    out->nodes = s.popped;
    pthread_mutex_destroy(&s.mu);
    pthread_cond_destroy(&s.cv);
    return s.popped;
}

typedef struct {
    const oracle_db *db;
    const int32_t *items;
    const int64_t *off;
    int64_t n;
    int64_t next;          /* shared index, taken under mu */
    int64_t kids;          /* children generated (keeps the scan's work live) */
    pthread_mutex_t mu;
} tn_state;

static void *tn_worker(void *arg) {
    tn_state *s = (tn_state *)arg;
    const oracle_db *db = s->db;
    int32_t *res = (int32_t *)calloc((size_t)db->d_count + 2, sizeof(int32_t));
    uint64_t *mask = (uint64_t *)malloc((size_t)(db->words > 0 ? db->words : 1) * sizeof(uint64_t));
    int64_t kids = 0;
    for (;;) {
        pthread_mutex_lock(&s->mu);
        const int64_t i = s->next++;
        pthread_mutex_unlock(&s->mu);
        if (i >= s->n) break;
        const int32_t *it = s->items + s->off[i];
        const int32_t k = (int32_t)(s->off[i + 1] - s->off[i]);
        memset(res, 0, ((size_t)db->d_count + 2) * sizeof(int32_t));
        const int32_t cnt = count_node_buf(db, it, k, res, mask);
        if (cnt == k) continue;   /* maximal: the worker records it */
        const int32_t progress = it[k - 1];
        for (int32_t e = 0; e <= progress - 1; e++) {
            if ((double)res[e] < db->thr_abs) continue;
            int outside = 1;
            for (int32_t j = 0; j < k; j++) if (it[j] == e) { outside = 0; break; }
            if (!outside) continue;
            /* the child's item list, as the worker builds it */
            int32_t *c = (int32_t *)malloc((size_t)(k + 1) * sizeof(int32_t));
            memcpy(c, it, (size_t)k * sizeof(int32_t));
            c[k] = e;
            kids += c[k] >= 0;
            free(c);
        }
    }
    pthread_mutex_lock(&s->mu);
    s->kids += kids;
    pthread_mutex_unlock(&s->mu);
    free(res); free(mask);
    return NULL;
}

double oracle_time_nodes(const oracle_db *db, const int32_t *items, const int64_t *off, int64_t n, int threads) {
    tn_state s;
    memset(&s, 0, sizeof(s));
    s.db = db; s.items = items; s.off = off; s.n = n;
    pthread_mutex_init(&s.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    const double t0 = now_s();
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, tn_worker, &s);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    const double dt = now_s() - t0;
    free(th);
    pthread_mutex_destroy(&s.mu);
    return dt;
}

int oracle_mine(const oracle_db *db, int threads, oracle_result *out) {
    oracle_mine_bounded(db, threads, 0, out);
    return 0;
}

void oracle_result_free(oracle_result *r) {
    for (int64_t i = 0; i < r->n_mfi; i++) free(r->mfi[i].items);
    free(r->mfi);
    memset(r, 0, sizeof(*r));
}

/* ------------------------------------------------------------------------ */
/* Results file (kernel_bmma.cu:1034-1073).  ostream's default double format
 * is printf's %g (6 significant digits). */
static int cmp_mfi(const void *a, const void *b) {
    const oracle_mfi *x = (const oracle_mfi *)a, *y = (const oracle_mfi *)b;
    if (x->frequency != y->frequency) return x->frequency > y->frequency ? -1 : 1;
    int32_t n = x->k < y->k ? x->k : y->k;
    for (int32_t i = 0; i < n; i++)
        if (x->items[i] != y->items[i]) return x->items[i] < y->items[i] ? -1 : 1;
    return x->k < y->k ? -1 : (x->k > y->k);
}

static void cpu_model(char *buf, size_t n) {
    snprintf(buf, n, "Unknown CPU");
    FILE *f = fopen("/proc/cpuinfo", "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, "model name", 10) == 0) {
            char *c = strchr(line, ':');
            if (c) {
                c++;
                while (*c == ' ' || *c == '\t') c++;
                size_t L = strlen(c);
                while (L && (c[L - 1] == '\n' || c[L - 1] == ' ')) c[--L] = 0;
                snprintf(buf, n, "%s", c);
            }
            break;
        }
    }
    fclose(f);
}

int oracle_write_results(const oracle_db *db, oracle_result *r, const char *path,
                         const char *trans_set_file, const char *engine_line) {
    qsort(r->mfi, (size_t)r->n_mfi, sizeof(oracle_mfi), cmp_mfi);
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    char cpu[256];
    cpu_model(cpu, sizeof cpu);
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    fprintf(f, "==================================================================\n");
    fprintf(f, "CPU: %s\n     Number of Logical Cores: %ld\n\n", cpu, ncpu);
    fprintf(f, "Engine: %s\n", engine_line ? engine_line : "cpu-oracle");
    fprintf(f, "==================================================================\n");
    fprintf(f, "transSetFile: %s\n", trans_set_file);
    fprintf(f, "dimension: %d\n", db->dimension);
    fprintf(f, "transNum: %d\n", db->trans_num);
    fprintf(f, "itemIDmin: %d\n", db->item_id_min);
    fprintf(f, "itemIDmax: %d\n", db->item_id_max);
    fprintf(f, "transLengthMin: %lld\n", (long long)db->len_min);
    fprintf(f, "transLengthMax: %lld\n", (long long)db->len_max);
    fprintf(f, "transLengthMean: %g\n", db->len_mean);
    fprintf(f, "\nsupportThreshold: %g\n\n", db->support);
    fprintf(f, "dimensionReduced: %d\n", db->dimension_reduced);
    fprintf(f, "transNumReduced: %d\n", db->trans_num_reduced);
    fprintf(f, "transLengthReducedMin: %lld\n", (long long)db->red_len_min);
    fprintf(f, "transLengthReducedMax: %lld\n", (long long)db->red_len_max);
    fprintf(f, "transLengthReducedMean: %g\n", db->red_len_mean);
    fprintf(f, "\nGPUtaskPercentage: 0%%\n");
    fprintf(f, "==================================================================\n");
    fprintf(f, "preprocessing_time(s): %g\n", r->preprocessing_time);
    fprintf(f, "computation_time(s): %g\n", r->computation_time);
    fprintf(f, "total_time(s): %g\n", r->preprocessing_time + r->computation_time);
    fprintf(f, "\nMFIsNumber: %lld\n", (long long)r->n_mfi);
    fprintf(f, "============================List of MFIs:=============================\n");
    for (int64_t i = 0; i < r->n_mfi; i++) {
        const oracle_mfi *m = &r->mfi[i];
        fprintf(f, "%lldth: %d { ", (long long)(i + 1), m->k);
        for (int32_t j = 0; j < m->k; j++) fprintf(f, "%d ", db->dense_to_orig[m->items[j]]);
        fprintf(f, "}   support: %g  frequency: %d\n", (double)m->frequency / (double)db->trans_num, m->frequency);
    }
    fprintf(f, "==================================================================\n");
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* cpu_ref (bmma_bench.cu:180-191), 64-bit popcount, row-parallel */
typedef struct {
    const uint64_t *A, *B; int32_t *C;
    int64_t b, d, words, ldc;
    int64_t next; pthread_mutex_t mu;
} pm_job;

static void *pm_worker(void *arg) {
    pm_job *j = (pm_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t i = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->b) break;
        const uint64_t *a = j->A + (size_t)i * j->words;
        for (int64_t c = 0; c < j->d; c++) {
            const uint64_t *v = j->B + (size_t)c * j->words;
            uint64_t s = 0;
            for (int64_t w = 0; w < j->words; w++) s += (uint64_t)__builtin_popcountll(a[w] & v[w]);
            j->C[(size_t)i * j->ldc + c] = (int32_t)s;
        }
    }
    return NULL;
}

void oracle_popcount_matrix(const uint64_t *A, const uint64_t *B, int32_t *C,
                            int64_t b, int64_t d, int64_t words, int64_t ldc, int threads) {
    pm_job j = { A, B, C, b, d, words, ldc, 0, PTHREAD_MUTEX_INITIALIZER };
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, pm_worker, &j);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea & Flood 2014): state advances by the golden gamma,
 * output is the 64-bit finaliser of the state. */
static inline uint64_t sm64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
#define SM64_GAMMA 0x9E3779B97F4A7C15ull

void oracle_bernoulli_row(uint64_t seed, uint32_t matrix, uint64_t row,
                          int64_t n_bits, int64_t words, uint32_t thr32, uint64_t *out) {
    uint64_t km = sm64_mix(seed + SM64_GAMMA * ((uint64_t)matrix + 1));
    uint64_t kr = sm64_mix(km + SM64_GAMMA * (row + 1));
    for (int64_t w = 0; w < words; w++) {
        uint64_t word = 0;
        for (int p = 0; p < 32; p++) {
            uint64_t pair = (uint64_t)w * 32 + (uint64_t)p;
            uint64_t x = sm64_mix(kr + SM64_GAMMA * (pair + 1));
            uint64_t b0 = (uint32_t)x < thr32, b1 = (uint32_t)(x >> 32) < thr32;
            word |= b0 << (2 * p);
            word |= b1 << (2 * p + 1);
        }
        int64_t base = w * 64;
        if (base + 64 > n_bits) {
            int64_t keep = n_bits - base;
            word = keep <= 0 ? 0 : (keep >= 64 ? word : (word & ((1ull << keep) - 1)));
        }
        out[w] = word;
    }
}

/* ------------------------------------------------------------------------ */
/* batch_proto.cu:193-230 restated (xorshift64 13/7/17, seed 0xC0FFEE123456789) */
#define PROTO_D 512
#define PROTO_N (1u << 23)
#define PROTO_NPAT 80
#define PROTO_MAXPAT 6
#define PROTO_MAXK 8
static inline uint64_t xs64(uint64_t *s) { *s ^= *s << 13; *s ^= *s >> 7; *s ^= *s << 17; return *s; }

void oracle_proto_gen(uint64_t *B, int32_t *pat_items, int32_t *pat_len) {
    const int64_t W = PROTO_N / 64;
    uint64_t seed = 0xC0FFEE123456789ull;
    double pfreq[PROTO_NPAT];
    for (int p = 0; p < PROTO_NPAT; p++) {
        int sz = 2 + (int)(xs64(&seed) % (PROTO_MAXPAT - 1));
        pfreq[p] = (0.2 + 1.3 * (double)(xs64(&seed) % 1000) / 1000.0) / 100.0;
        int n = 0;
        int32_t *it = pat_items + p * PROTO_MAXPAT;
        while (n < sz) {
            int x = (int)(xs64(&seed) % PROTO_D);
            int dup = 0;
            for (int q = 0; q < n; q++) if (it[q] == x) dup = 1;
            if (!dup) it[n++] = x;
        }
        for (int a = 1; a < n; a++)   /* sort ascending */
            for (int b2 = a; b2 > 0 && it[b2 - 1] > it[b2]; b2--) { int32_t t = it[b2]; it[b2] = it[b2 - 1]; it[b2 - 1] = t; }
        pat_len[p] = n;
    }
    uint32_t thr[PROTO_NPAT];
    for (int p = 0; p < PROTO_NPAT; p++) thr[p] = (uint32_t)(pfreq[p] * 4294967296.0);
    memset(B, 0, (size_t)PROTO_D * W * sizeof(uint64_t));
    for (uint64_t t = 0; t < PROTO_N; t++) {
        for (int p = 0; p < PROTO_NPAT; p++) {
            if ((uint32_t)xs64(&seed) < thr[p]) {
                const int32_t *it = pat_items + p * PROTO_MAXPAT;
                for (int q = 0; q < pat_len[p]; q++) B[(size_t)it[q] * W + (t >> 6)] |= 1ull << (t & 63);
            }
        }
    }
    for (int j = 0; j < PROTO_D; j++) {
        int n = (int)(PROTO_N * 0.0002);
        for (int q = 0; q < n; q++) {
            uint64_t t = xs64(&seed) % PROTO_N;
            B[(size_t)j * W + (t >> 6)] |= 1ull << (t & 63);
        }
    }
}

typedef struct { int32_t items[PROTO_MAXK]; int32_t k; } pset;
typedef struct {
    const uint64_t *B; const pset *cur; int64_t ncur; int level; int32_t minsup;
    const uint8_t *f1; int64_t next_i; pthread_mutex_t mu;
    int64_t *nfound_per_prefix; /* [ncur] */
    int32_t *ext; /* [ncur * PROTO_D] extension flags */
} lw_job;

static void *lw_worker(void *arg) {
    lw_job *j = (lw_job *)arg;
    const int64_t W = PROTO_N / 64;
    uint64_t *mask = (uint64_t *)malloc((size_t)W * sizeof(uint64_t));
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t i = j->next_i++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->ncur) break;
        const pset *s = &j->cur[i];
        for (int64_t w = 0; w < W; w++) mask[w] = ~0ull;
        for (int t = 0; t < s->k; t++) {
            const uint64_t *v = j->B + (size_t)s->items[t] * W;
            for (int64_t w = 0; w < W; w++) mask[w] &= v[w];
        }
        int last = s->items[s->k - 1];
        int64_t nf = 0;
        for (int c = last + 1; c < PROTO_D; c++) {
            /* a column below minsup at level 1 cannot reach minsup later */
            if (!j->f1[c]) { j->ext[i * PROTO_D + c] = 0; continue; }
            const uint64_t *v = j->B + (size_t)c * W;
            uint64_t sup = 0;
            for (int64_t w = 0; w < W; w++) sup += (uint64_t)__builtin_popcountll(mask[w] & v[w]);
            int ok = (int64_t)sup >= j->minsup;
            j->ext[i * PROTO_D + c] = ok;
            nf += ok;
        }
        j->nfound_per_prefix[i] = nf;
    }
    free(mask);
    return NULL;
}

int64_t oracle_proto_levelwise(const uint64_t *B, int threads, int64_t *level_found) {
    const int64_t W = PROTO_N / 64;
    const int32_t minsup = (int32_t)(PROTO_N * 0.15 / 100.0);
    if (threads < 1) threads = 1;
    uint8_t f1[PROTO_D];
    pset *cur = (pset *)malloc(PROTO_D * sizeof(pset));
    int64_t ncur = 0, total = 0;
    for (int j = 0; j < PROTO_D; j++) {
        uint64_t s = 0;
        for (int64_t w = 0; w < W; w++) s += (uint64_t)__builtin_popcountll(B[(size_t)j * W + w]);
        f1[j] = (int64_t)s >= minsup;
        if (f1[j]) { cur[ncur].k = 1; cur[ncur].items[0] = j; ncur++; }
    }
    for (int l = 0; l < PROTO_MAXK; l++) level_found[l] = 0;
    level_found[0] = ncur;
    total = ncur;
    for (int level = 2; level <= PROTO_MAXK && ncur > 0; level++) {
        lw_job j;
        memset(&j, 0, sizeof j);
        j.B = B; j.cur = cur; j.ncur = ncur; j.level = level; j.minsup = minsup; j.f1 = f1;
        pthread_mutex_init(&j.mu, NULL);
        j.nfound_per_prefix = (int64_t *)calloc((size_t)ncur, sizeof(int64_t));
        j.ext = (int32_t *)calloc((size_t)ncur * PROTO_D, sizeof(int32_t));
        pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
        for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, lw_worker, &j);
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        free(th);
        int64_t nn = 0;
        for (int64_t i = 0; i < ncur; i++) nn += j.nfound_per_prefix[i];
        pset *next = (pset *)malloc((size_t)(nn > 0 ? nn : 1) * sizeof(pset));
        int64_t q = 0;
        for (int64_t i = 0; i < ncur; i++)
            for (int c = 0; c < PROTO_D; c++)
                if (j.ext[i * PROTO_D + c]) {
                    next[q] = cur[i];
                    next[q].items[next[q].k++] = c;
                    q++;
                }
        level_found[level - 1] = nn;
        total += nn;
        free(j.nfound_per_prefix); free(j.ext); free(cur);
        pthread_mutex_destroy(&j.mu);
        cur = next; ncur = nn;
    }
    free(cur);
    return total;
}
