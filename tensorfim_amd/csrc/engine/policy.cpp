#include "policy.h"

#include <algorithm>
#include <numeric>

#include "search.h"

namespace tfim {

void PlainPolicy::decide(std::vector<Decision *> &ds, const Database &db) {
    (void)db;
    for (Decision *d : ds) {
        d->order.resize(d->tail.size());
        std::iota(d->order.begin(), d->order.end(), 0);
        std::stable_sort(d->order.begin(), d->order.end(), [&](int32_t a, int32_t b) {
            if (d->tail_sup[a] != d->tail_sup[b]) return d->tail_sup[a] < d->tail_sup[b];
            return d->tail[a] < d->tail[b];
        });
        d->lookahead = d->tail.size() >= 2;
    }
}

std::unique_ptr<Policy> make_policy(tfim_ctx *ctx, const SearchOptions &opt) {
    (void)ctx;
    (void)opt;
    return std::unique_ptr<Policy>(new PlainPolicy());
}

}  // namespace tfim
