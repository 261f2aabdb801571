// tfim_gen -- deterministic synthetic transaction databases for the benchmark
// workloads, written in the reference's input format (one transaction per
// line, ascending item ids >= 1, space separated).
//
//   tfim_gen <dense-small|clickstream|dense-replicated|gene> <out.txt> [--rows N]
//
// Every draw comes from one SplitMix64 stream seeded by the workload's seed,
// so a file is a pure function of (workload, rows).  The isolated count
// (65,536 x 512 bitmaps over 2^22 transactions) has no transaction file: its
// bitmaps are generated on the device (tfim_fill_bernoulli).
// dense-replicated writes the 49,046-row base table; the engine replicates it
// x256 (--replicate 256 = data/make_replica.py).
//
// dense-small       3,200 x items 1..75, exactly 37 distinct items per transaction:
//           6 blocks of 8 items (a seeded permutation of 1..75), visited in a
//           random order, each added with p=0.7 when it still fits; the rest
//           drawn without replacement with weight 1/(i+1)^0.8.  seed 0.
// clickstream       990,000 transactions over 41,270 items: 1+Poisson(7) distinct
//           items by Zipf(1.05) popularity (id = rank), plus 200 planted
//           patterns of 3..8 uniform items, each inserted into exactly 4,950
//           (0.5%) distinct transactions.  seed 2.
// dense-replicated  49,046 transactions over 2,113 items, exactly 74 per transaction
//           (pumsb-like: 74 attribute slots, slot k owns ids k+1, k+75, ...,
//           values Zipf(0.7) in id order).  20 blocks of 12 slots drawn from
//           the first 40 slots fire independently with p=0.8; a core slot
//           takes its first value when at least 60% of its blocks fire or,
//           independently, with p=0.45; otherwise its Zipf draw.  seed 3.
// gene              1,100 transactions over 20,000 items: each item present with
//           p=0.05; 40 co-regulated modules of 50..300 items; transaction t
//           has a latent factor z_t ~ N(0,1) and module m switches on with
//           probability sigmoid(a_m z_t + b_m), a_m ~ U(1,3), b_m ~ U(-3.5,-1.5);
//           an active module's items are present with p=0.6.  seed 4.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
    uint64_t below(uint64_t n) { return next() % n; }
    double normal() {   // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int poisson(double lam) {   // Knuth, fine for small lambda
        const double L = std::exp(-lam);
        int k = 0;
        double p = 1.0;
        do {
            k++;
            p *= uniform();
        } while (p > L);
        return k - 1;
    }
};

// Index drawn from a cumulative weight table.
int draw_cdf(Rng &r, const std::vector<double> &cdf) {
    const double u = r.uniform() * cdf.back();
    return (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
}

void write_row(FILE *f, std::vector<int> &row) {
    std::sort(row.begin(), row.end());
    for (size_t i = 0; i < row.size(); i++) fprintf(f, i ? " %d" : "%d", row[i]);
    fputc('\n', f);
}

void gen1(FILE *f, int rows) {
    Rng r(0);
    const int n = 75, L = 37;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 1);
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
    std::vector<std::vector<int>> blocks(6);
    for (int b = 0; b < 6; b++) blocks[b].assign(perm.begin() + 8 * b, perm.begin() + 8 * b + 8);
    std::vector<double> w(n + 1, 0.0);
    for (int i = 1; i <= n; i++) w[i] = 1.0 / std::pow((double)(i + 1), 0.8);
    std::vector<int> order(6), row;
    std::vector<char> in(n + 1);
    for (int t = 0; t < rows; t++) {
        std::fill(in.begin(), in.end(), 0);
        row.clear();
        std::iota(order.begin(), order.end(), 0);
        for (int i = 5; i > 0; i--) std::swap(order[i], order[r.below(i + 1)]);
        for (int b : order) {
            if (r.uniform() >= 0.7) continue;
            int add = 0;
            for (int x : blocks[b]) add += !in[x];
            if ((int)row.size() + add > L) continue;
            for (int x : blocks[b])
                if (!in[x]) { in[x] = 1; row.push_back(x); }
        }
        while ((int)row.size() < L) {   // weighted draw without replacement
            double tot = 0;
            for (int i = 1; i <= n; i++) if (!in[i]) tot += w[i];
            double u = r.uniform() * tot;
            int pick = -1;
            for (int i = 1; i <= n; i++) {
                if (in[i]) continue;
                pick = i;
                u -= w[i];
                if (u < 0) break;
            }
            in[pick] = 1;
            row.push_back(pick);
        }
        write_row(f, row);
    }
}

void gen3(FILE *f, int rows) {
    Rng r(2);
    const int n = 41270, npat = 200, per = 4950;
    std::vector<double> cdf(n);
    double acc = 0;
    for (int i = 0; i < n; i++) cdf[i] = (acc += 1.0 / std::pow((double)(i + 1), 1.05));
    std::vector<std::vector<int>> pats(npat);
    for (auto &p : pats) {
        const int len = 3 + (int)r.below(6);
        while ((int)p.size() < len) {
            const int x = 1 + (int)r.below(n);
            if (std::find(p.begin(), p.end(), x) == p.end()) p.push_back(x);
        }
    }
    // pattern -> transactions: a partial Fisher-Yates draw of `per` rows each
    std::vector<std::vector<int>> plant((size_t)rows);
    std::vector<int> idx(rows);
    std::iota(idx.begin(), idx.end(), 0);
    const int take = std::min(per, rows);
    for (int p = 0; p < npat; p++) {
        for (int i = 0; i < take; i++) {
            const int j = i + (int)r.below((uint64_t)(rows - i));
            std::swap(idx[i], idx[j]);
            plant[idx[i]].push_back(p);
        }
    }
    std::vector<int> row;
    for (int t = 0; t < rows; t++) {
        row.clear();
        const int len = 1 + r.poisson(7.0);
        int guard = 0;
        while ((int)row.size() < len && guard++ < 100 * len) {
            const int x = 1 + draw_cdf(r, cdf);
            if (std::find(row.begin(), row.end(), x) == row.end()) row.push_back(x);
        }
        for (int p : plant[t])
            for (int x : pats[p])
                if (std::find(row.begin(), row.end(), x) == row.end()) row.push_back(x);
        write_row(f, row);
    }
}

void gen4(FILE *f, int rows) {
    Rng r(3);
    const int S = 74, n = 2113, nblk = 20, bsz = 12, core = 40;
    std::vector<std::vector<int>> slot(S);
    std::vector<std::vector<double>> cdf(S);
    for (int k = 0; k < S; k++) {
        double acc = 0;
        for (int id = k + 1; id <= n; id += S) {
            slot[k].push_back(id);
            cdf[k].push_back(acc += 1.0 / std::pow((double)slot[k].size(), 0.7));
        }
    }
    std::vector<std::vector<int>> member(S);   // slot -> blocks
    for (int b = 0; b < nblk; b++) {
        std::vector<int> p(core);
        std::iota(p.begin(), p.end(), 0);
        for (int i = 0; i < bsz; i++) std::swap(p[i], p[i + r.below((uint64_t)(core - i))]);
        for (int i = 0; i < bsz; i++) member[p[i]].push_back(b);
    }
    std::vector<char> fire(nblk);
    std::vector<int> row(S);
    for (int t = 0; t < rows; t++) {
        for (int b = 0; b < nblk; b++) fire[b] = r.uniform() < 0.8;
        for (int k = 0; k < S; k++) {
            int v = slot[k][draw_cdf(r, cdf[k])];
            if (k < core) {
                const int m = (int)member[k].size();
                int on = 0;
                for (int b : member[k]) on += fire[b];
                const bool boost = r.uniform() < 0.45;
                if ((m > 0 && 10 * on >= 6 * m) || boost) v = slot[k][0];
            }
            row[k] = v;
        }
        std::vector<int> tmp = row;
        write_row(f, tmp);
    }
}

void gen5(FILE *f, int rows) {
    Rng r(4);
    const int n = 20000, nmod = 40;
    std::vector<std::vector<int>> mods(nmod);
    std::vector<double> a(nmod), b(nmod);
    for (int m = 0; m < nmod; m++) {
        const int sz = 50 + (int)r.below(251);
        std::vector<char> used(n + 1, 0);
        while ((int)mods[m].size() < sz) {
            const int x = 1 + (int)r.below(n);
            if (!used[x]) { used[x] = 1; mods[m].push_back(x); }
        }
        a[m] = 1.0 + 2.0 * r.uniform();
        b[m] = -1.5 - 2.0 * r.uniform();
    }
    std::vector<char> in(n + 1);
    std::vector<int> row;
    for (int t = 0; t < rows; t++) {
        std::fill(in.begin(), in.end(), 0);
        for (int i = 1; i <= n; i++) in[i] = r.uniform() < 0.05;
        const double z = r.normal();
        for (int m = 0; m < nmod; m++) {
            const double pon = 1.0 / (1.0 + std::exp(-(a[m] * z + b[m])));
            if (r.uniform() >= pon) continue;
            for (int x : mods[m])
                if (r.uniform() < 0.6) in[x] = 1;
        }
        row.clear();
        for (int i = 1; i <= n; i++)
            if (in[i]) row.push_back(i);
        write_row(f, row);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: tfim_gen <dense-small|clickstream|dense-replicated|gene> <out.txt> [--rows N]\n");
        return 1;
    }
    const std::string w = argv[1];
    const int cfg = w == "dense-small" ? 1 : w == "clickstream" ? 3 : w == "dense-replicated" ? 4 : w == "gene" ? 5 : 0;
    int rows = cfg == 1 ? 3200 : cfg == 3 ? 990000 : cfg == 4 ? 49046 : cfg == 5 ? 1100 : 0;
    if (!cfg) {
        fprintf(stderr, "unknown workload %s\n", argv[1]);
        return 1;
    }
    for (int i = 3; i + 1 < argc; i++)
        if (!strcmp(argv[i], "--rows")) rows = atoi(argv[i + 1]);
    if (rows <= 0) {
        fprintf(stderr, "bad --rows\n");
        return 1;
    }
    FILE *f = fopen(argv[2], "w");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    switch (cfg) {
        case 1: gen1(f, rows); break;
        case 3: gen3(f, rows); break;
        case 4: gen4(f, rows); break;
        case 5: gen5(f, rows); break;
    }
    fclose(f);
    return 0;
}
