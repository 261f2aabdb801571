// Device preprocessing of a parsed transaction database: occurrence counts,
// row reduction (infrequent items dropped, rows left empty removed) and the
// vertical bitmaps, replacing the host passes of the reference
// (kernel_bmma.cu:272-276 frequency count, :1075-1086 reduction, :388-426
// buildBitmap).  Rows arrive as CSR (ids, int64 offsets).
#include <climits>

#include "tfim_internal.h"

namespace tfim {

// Occurrence counts of ids in [0, nbins): a private histogram per workgroup
// in LDS, merged with one global atomic per non-zero bin.  An id outside the
// range is not counted and sets *bad.
__global__ __launch_bounds__(1024) void k_prep_hist_lds(const int32_t *__restrict__ ids, int64_t n, int32_t nbins,
                                                        int32_t *__restrict__ freq, int32_t *__restrict__ bad) {
    extern __shared__ int hist[];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = ids[i];
        if (x >= 0 && x < nbins) atomicAdd(&hist[x], 1);
        else atomicOr(bad, 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (hist[i]) atomicAdd(&freq[i], hist[i]);
}

// The same straight into global memory, for id ranges too wide for LDS.
__global__ __launch_bounds__(256) void k_prep_hist_global(const int32_t *__restrict__ ids, int64_t n, int32_t nbins,
                                                          int32_t *__restrict__ freq, int32_t *__restrict__ bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = ids[i];
        if (x >= 0 && x < nbins) atomicAdd(&freq[x], 1);
        else atomicOr(bad, 1);
    }
}

constexpr int kPrepRowsPerThread = 4;
constexpr int kPrepThreads = 256;
constexpr int kPrepRowsPerBlock = kPrepRowsPerThread * kPrepThreads;

// Per row: the number of surviving items (dense_of[id] >= 0).  Per block:
// how many rows keep at least one item.  Over kept rows: sum, min and max of
// that number (stats[0..2]).
__global__ __launch_bounds__(kPrepThreads) void k_prep_rows(const int32_t *__restrict__ ids,
                                                            const int64_t *__restrict__ off, int64_t n_rows,
                                                            const int32_t *__restrict__ dense_of,
                                                            int32_t *__restrict__ cnt, int32_t *__restrict__ blk_keep,
                                                            unsigned long long *__restrict__ stats) {
    __shared__ int keep_s;
    __shared__ unsigned long long sum_s;
    __shared__ int min_s, max_s;
    if (threadIdx.x == 0) {
        keep_s = 0;
        sum_s = 0;
        min_s = INT_MAX;
        max_s = 0;
    }
    __syncthreads();
    int keep = 0, mn = INT_MAX, mx = 0;
    unsigned long long sum = 0;
    const int64_t r0 = (int64_t)blockIdx.x * kPrepRowsPerBlock + (int64_t)threadIdx.x * kPrepRowsPerThread;
    for (int q = 0; q < kPrepRowsPerThread; q++) {
        const int64_t t = r0 + q;
        if (t >= n_rows) break;
        int c = 0;
        for (int64_t j = off[t]; j < off[t + 1]; j++) c += dense_of[ids[j]] >= 0;
        cnt[t] = c;
        if (c > 0) {
            keep++;
            sum += (unsigned long long)c;
            mn = min(mn, c);
            mx = max(mx, c);
        }
    }
    if (keep) {
        atomicAdd(&keep_s, keep);
        atomicAdd(&sum_s, sum);
        atomicMin(&min_s, mn);
        atomicMax(&max_s, mx);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_keep[blockIdx.x] = keep_s;
        if (keep_s) {
            atomicAdd(&stats[0], sum_s);
            atomicMin(&stats[1], (unsigned long long)min_s);
            atomicMax(&stats[2], (unsigned long long)max_s);
        }
    }
}

// Exclusive scan of the per-block kept-row counts, in place, by one
// workgroup; total[0] receives the number of kept rows.
__global__ __launch_bounds__(1024) void k_prep_scan_blocks(int32_t *__restrict__ v, int64_t n,
                                                           int64_t *__restrict__ total) {
    __shared__ int64_t s[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < n; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int64_t x = i < n ? v[i] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {   // inclusive Hillis-Steele scan
            const int64_t y = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < n) v[i] = (int32_t)(carry + s[threadIdx.x] - x);
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[0] = carry;
}

// kept[r] = the original index of the r-th row that keeps an item.
__global__ __launch_bounds__(kPrepThreads) void k_prep_index(const int32_t *__restrict__ cnt, int64_t n_rows,
                                                             const int32_t *__restrict__ blk_off,
                                                             int32_t *__restrict__ kept) {
    __shared__ int s[kPrepThreads];
    const int64_t r0 = (int64_t)blockIdx.x * kPrepRowsPerBlock + (int64_t)threadIdx.x * kPrepRowsPerThread;
    int local = 0;
    for (int q = 0; q < kPrepRowsPerThread; q++)
        if (r0 + q < n_rows && cnt[r0 + q] > 0) local++;
    s[threadIdx.x] = local;
    __syncthreads();
    for (int d = 1; d < kPrepThreads; d <<= 1) {
        const int y = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    int pos = blk_off[blockIdx.x] + s[threadIdx.x] - local;
    for (int q = 0; q < kPrepRowsPerThread; q++)
        if (r0 + q < n_rows && cnt[r0 + q] > 0) kept[pos++] = (int32_t)(r0 + q);
}

// Vertical bitmaps.  One wave per 64 kept rows (one bitmap word): lane L
// walks row 64w + L's surviving items; each step the wave takes the smallest
// pending item m, and the first lane holding m ORs the ballot of all lanes
// holding it into word w of row m.  A bit that is already set means an item
// repeated within a transaction (*dup = 1).
__global__ __launch_bounds__(256) void k_prep_build(const int32_t *__restrict__ ids, const int64_t *__restrict__ off,
                                                    const int32_t *__restrict__ kept, int64_t n_red,
                                                    const int32_t *__restrict__ dense_of, uint64_t *__restrict__ bm,
                                                    int64_t pitch, int32_t *__restrict__ dup) {
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (w * 64 >= n_red) return;   // whole waves
    const int64_t tr = w * 64 + lane;
    int64_t p = 0, e = 0;
    if (tr < n_red) {
        const int32_t t = kept[tr];
        p = off[t];
        e = off[(int64_t)t + 1];
    }
    auto next = [&]() {
        while (p < e) {
            const int x = dense_of[ids[p]];
            if (x >= 0) return x;
            p++;
        }
        return INT_MAX;
    };
    int x = next();
    while (true) {
        int m = x;
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) m = min(m, __shfl_xor(m, s));
        if (m == INT_MAX) break;
        const unsigned long long mask = __ballot(x == m);
        if (lane == __ffsll((long long)mask) - 1) {
            const unsigned long long old =
                atomicOr(reinterpret_cast<unsigned long long *>(&bm[(int64_t)m * pitch + w]), mask);
            if (old & mask) atomicOr(dup, 1);
        }
        if (x == m) {
            p++;
            x = next();
        }
    }
}

void launch_prep_hist(const int32_t *ids, int64_t n, int32_t nbins, int32_t *freq, int32_t *bad, int n_cu,
                      hipStream_t s) {
    if (n <= 0) return;
    if (nbins <= 16384) {
        const int64_t grid = std::min<int64_t>((n + 1023) / 1024, (int64_t)n_cu * 2);
        hipLaunchKernelGGL(k_prep_hist_lds, dim3((unsigned)std::max<int64_t>(1, grid)), dim3(1024),
                           (size_t)nbins * 4, s, ids, n, nbins, freq, bad);
    } else {
        const int64_t grid = std::min<int64_t>((n + 255) / 256, (int64_t)n_cu * 16);
        hipLaunchKernelGGL(k_prep_hist_global, dim3((unsigned)std::max<int64_t>(1, grid)), dim3(256), 0, s, ids, n,
                           nbins, freq, bad);
    }
}

int64_t prep_blocks(int64_t n_rows) { return (n_rows + kPrepRowsPerBlock - 1) / kPrepRowsPerBlock; }

void launch_prep_rows(const int32_t *ids, const int64_t *off, int64_t n_rows, const int32_t *dense_of, int32_t *cnt,
                      int32_t *blk_keep, unsigned long long *stats, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_rows, dim3((unsigned)prep_blocks(n_rows)), dim3(kPrepThreads), 0, s, ids, off, n_rows,
                       dense_of, cnt, blk_keep, stats);
}

void launch_prep_scan(int32_t *blk, int64_t n_blk, int64_t *total, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_scan_blocks, dim3(1), dim3(1024), 0, s, blk, n_blk, total);
}

void launch_prep_index(const int32_t *cnt, int64_t n_rows, const int32_t *blk_off, int32_t *kept, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_index, dim3((unsigned)prep_blocks(n_rows)), dim3(kPrepThreads), 0, s, cnt, n_rows,
                       blk_off, kept);
}

void launch_prep_build(const int32_t *ids, const int64_t *off, const int32_t *kept, int64_t n_red,
                       const int32_t *dense_of, uint64_t *bm, int64_t pitch, int32_t *dup, hipStream_t s) {
    if (n_red <= 0) return;
    const int64_t waves = (n_red + 63) / 64;
    const int64_t grid = (waves + 3) / 4;
    hipLaunchKernelGGL(k_prep_build, dim3((unsigned)grid), dim3(256), 0, s, ids, off, kept, n_red, dense_of, bm, pitch,
                       dup);
}

}  // namespace tfim
