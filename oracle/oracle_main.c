/*
 * oracle_main.c -- CLI for the CPU oracle (test infrastructure only).
 *
 *   tfim_oracle <file> <thr> [--threads N] [--replicate K] [--out PATH]
 *                            [--max-nodes M]
 *       The reference CLI contract (kernel_bmma.cu:1179-1181,1291-1293):
 *       mines <file> at relative support <thr> with the reference DFS and
 *       writes "<file>-<%f thr>=Results.txt" (or PATH).  --max-nodes bounds
 *       the DFS for timing samples; the output is then partial.
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

int main(int argc, char **argv) {
    int threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    int replicate = 0;
    long long max_nodes = 0;
    const char *out = NULL;
    const char *pos[2] = {NULL, NULL};
    int npos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--replicate") && i + 1 < argc) replicate = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "--max-nodes") && i + 1 < argc) max_nodes = atoll(argv[++i]);
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
