/*
 * oracle_main.c -- CLI for the CPU oracle (test infrastructure only).
 *
 *   tfim_oracle <file> <thr> [--threads N] [--replicate K] [--out PATH]
 *                            [--max-nodes M]
 *       The reference CLI contract (kernel_bmma.cu:1179-1181,1291-1293):
 *       mines <file> at relative support <thr> with the reference DFS and
 *       writes "<file>-<%f thr>=Results.txt" (or PATH).  --max-nodes bounds
 *       the DFS for timing samples; the output is then partial.
 *   tfim_oracle <file> <thr> --time-nodes NODES --sample K [--seed S] [--threads N] [--replicate K]
 *       Times the reference worker's per-node step on K nodes drawn at random
 *       (without replacement, SplitMix64(S)) from NODES, one node per line as
 *       original item ids in insertion order (the node list of a run); prints
 *       the time and its extrapolation to every node of the list.
 *   tfim_oracle proto [--threads N]
 *       batch_proto.cu level-wise known-answer run; prints per-level counts.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static uint64_t sm64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Reads the node list, draws the sample and times it (see the usage above). */
static int time_node_sample(const oracle_db *db, const char *path, long long sample, unsigned long long seed,
                            int threads) {
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "%s open failed!\n", path); return 1; }
    int32_t mx = 0;
    for (int32_t i = 0; i < db->d_count; i++) if (db->dense_to_orig[i] > mx) mx = db->dense_to_orig[i];
    int32_t *o2d = (int32_t *)malloc(((size_t)mx + 1) * sizeof(int32_t));
    for (int32_t i = 0; i <= mx; i++) o2d[i] = -1;
    for (int32_t i = 0; i < db->d_count; i++) o2d[db->dense_to_orig[i]] = i;
    /* all nodes, CSR */
    size_t ni = 0, ci = 1 << 20, nn = 0, cn = 1 << 16;
    int32_t *ids = (int32_t *)malloc(ci * sizeof(int32_t));
    int64_t *off = (int64_t *)malloc(cn * sizeof(int64_t));
    off[0] = 0;
    char line[1 << 16];
    while (fgets(line, sizeof line, f)) {
        char *p = line;
        for (;;) {
            while (*p == ' ' || *p == '\t') p++;
            if (*p < '0' || *p > '9') break;
            long v = strtol(p, &p, 10);
            if (v < 0 || v > mx || o2d[v] < 0) { fprintf(stderr, "node list: id %ld is not an item of the database\n", v); return 1; }
            if (ni == ci) { ci *= 2; ids = (int32_t *)realloc(ids, ci * sizeof(int32_t)); }
            ids[ni++] = o2d[v];
        }
        if ((int64_t)ni == off[nn]) continue;   /* empty line */
        if (nn + 2 > cn) { cn *= 2; off = (int64_t *)realloc(off, cn * sizeof(int64_t)); }
        off[++nn] = (int64_t)ni;
    }
    fclose(f);
    if (nn == 0) { fprintf(stderr, "node list %s is empty\n", path); return 1; }
    /* sample without replacement: partial Fisher-Yates over node indices */
    const int64_t k = sample <= 0 || (size_t)sample > nn ? (int64_t)nn : (int64_t)sample;
    int64_t *perm = (int64_t *)malloc(nn * sizeof(int64_t));
    for (size_t i = 0; i < nn; i++) perm[i] = (int64_t)i;
    uint64_t st = seed;
    for (int64_t i = 0; i < k; i++) {
        const int64_t j = i + (int64_t)(sm64(&st) % (uint64_t)(nn - (size_t)i));
        const int64_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    int32_t *sids = (int32_t *)malloc((ni + 1) * sizeof(int32_t));
    int64_t *soff = (int64_t *)malloc(((size_t)k + 1) * sizeof(int64_t));
    soff[0] = 0;
    for (int64_t i = 0; i < k; i++) {
        const int64_t a = off[perm[i]], b = off[perm[i] + 1];
        memcpy(sids + soff[i], ids + a, (size_t)(b - a) * sizeof(int32_t));
        soff[i + 1] = soff[i] + (b - a);
    }
    const double secs = oracle_time_nodes(db, sids, soff, k, threads);
    printf("nodes_total: %lld\nnodes_timed: %lld\nseconds: %g\nextrapolated_s: %g\nthreads: %d\n",
           (long long)nn, (long long)k, secs, secs * (double)nn / (double)k, threads);
    free(o2d); free(ids); free(off); free(perm); free(sids); free(soff);
    return 0;
}

int main(int argc, char **argv) {
    int threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    int replicate = 0;
    long long max_nodes = 0;
    const char *out = NULL;
    const char *time_nodes = NULL;
    long long sample = 0;
    unsigned long long seed = 1;
    const char *pos[2] = {NULL, NULL};
    int npos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--replicate") && i + 1 < argc) replicate = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "--max-nodes") && i + 1 < argc) max_nodes = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--time-nodes") && i + 1 < argc) time_nodes = argv[++i];
        else if (!strcmp(argv[i], "--sample") && i + 1 < argc) sample = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], NULL, 10);
        else if (npos < 2) pos[npos++] = argv[i];
    }
    if (npos >= 1 && !strcmp(pos[0], "proto")) {
        size_t words = (size_t)512 * ((1u << 23) / 64);
        uint64_t *B = (uint64_t *)malloc(words * sizeof(uint64_t));
        int32_t pat[80 * 6], plen[80];
        double t0 = now_s();
        oracle_proto_gen(B, pat, plen);
        double t1 = now_s();
        int64_t lf[8];
        int64_t tot = oracle_proto_levelwise(B, threads, lf);
        double t2 = now_s();
        printf("gen %.2f s, levelwise %.2f s (%d threads)\n", t1 - t0, t2 - t1, threads);
        for (int l = 0; l < 8; l++) printf("L%d %lld\n", l + 1, (long long)lf[l]);
        printf("total %lld\n", (long long)tot);
        free(B);
        return 0;
    }
    if (npos < 2) {
        fprintf(stderr, "usage: tfim_oracle <file> <thr> [--threads N] [--replicate K] [--out PATH] [--max-nodes M]\n");
        return 2;
    }
    double support = atof(pos[1]);
    oracle_db db;
    double t0 = now_s();
    if (oracle_load(pos[0], support, replicate, &db) != 0) {
        fprintf(stderr, "%s open failed!\n", pos[0]);
        return 1;
    }
    double t1 = now_s();
    if (time_nodes) return time_node_sample(&db, time_nodes, sample, seed, threads);
    oracle_result r;
    oracle_mine_bounded(&db, threads, max_nodes, &r);
    r.preprocessing_time = t1 - t0;
    char path[4096];
    if (out) snprintf(path, sizeof path, "%s", out);
    else snprintf(path, sizeof path, "%s-%f=Results.txt", pos[0], support);
    char eng[128];
    snprintf(eng, sizeof eng, "cpu-oracle (reference DFS, %d threads)", threads);
    oracle_write_results(&db, &r, path, pos[0], eng);
    printf("nodes: %lld\nMFIsNumber: %lld\ncomputation_time(s): %g\nload_time(s): %g\nthreads: %d\n",
           (long long)r.nodes, (long long)r.n_mfi, r.computation_time, t1 - t0, threads);
    oracle_result_free(&r);
    oracle_free(&db);
    return 0;
}
