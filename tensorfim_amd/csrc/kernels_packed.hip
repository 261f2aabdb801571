// Packed-bitmap kernels for gfx950: 128-bit loads, AND, v_bcnt_u32_b32,
// wave-level reductions into int32 supports.
//
//   k_count_node  -- plain per-node path (replaces getResFrequencyCudaKernel,
//                    kernel_bitmap.cu:412-442).  The reference re-ANDs the
//                    prefix rows once per column block; here a block ANDs the
//                    prefix once for its word chunk and reuses it for up to 64
//                    columns, so per node the traffic is (k + n_cols) rows.
//   k_count_wave  -- fused, persistent wave kernel (replaces maskBuildKernelFast
//                    + bmmaKernel* + bmmaSplitReduceKernel, kernel_bmma.cu:
//                    619-768): prefix AND, per-row tail counts and the split
//                    over word chunks accumulated with int32 atomics.
//   k_popmat_packed -- isolated C = A AND-popcount B^T (bitmap_kernel of
//                    code/microbench/bmma_bench.cu:159-174), LDS tiled 64x64,
//                    4x4 register micro-tiles.
#include <mutex>

#include "tfim_internal.h"

namespace tfim {

// acc + popcount(a AND b) as a chain of v_bcnt_u32_b32 with accumulate (one
// AND and one BCNT per dword): the empty asm statements keep the compiler
// from re-associating the sum into separate adds.
__device__ __forceinline__ uint32_t popc_and4_acc(const uint4 a, const uint4 b, uint32_t acc) {
    acc = __popc(a.x & b.x) + acc;
    asm("" : "+v"(acc));
    acc = __popc(a.y & b.y) + acc;
    asm("" : "+v"(acc));
    acc = __popc(a.z & b.z) + acc;
    asm("" : "+v"(acc));
    acc = __popc(a.w & b.w) + acc;
    asm("" : "+v"(acc));
    return acc;
}

// The same for one 64-bit word pair: two ANDs, two chained BCNTs.
__device__ __forceinline__ uint32_t popc_and64_acc(const uint64_t a, const uint64_t b, uint32_t acc) {
    acc = __popc((uint32_t)a & (uint32_t)b) + acc;
    asm("" : "+v"(acc));
    acc = __popc((uint32_t)(a >> 32) & (uint32_t)(b >> 32)) + acc;
    asm("" : "+v"(acc));
    return acc;
}

__device__ __forceinline__ uint4 and4(const uint4 a, const uint4 b) {
    return make_uint4(a.x & b.x, a.y & b.y, a.z & b.z, a.w & b.w);
}

// a AND b AND c in one v_bitop3_b32 per dword (gfx950's three-input logic
// op; truth table 0x80: only the all-ones input gives one), where two
// v_and_b32 would be needed.
__device__ __forceinline__ uint4 and3_4(const uint4 a, const uint4 b, const uint4 c) {
    return make_uint4(__builtin_amdgcn_bitop3_b32(a.x, b.x, c.x, 0x80), __builtin_amdgcn_bitop3_b32(a.y, b.y, c.y, 0x80),
                      __builtin_amdgcn_bitop3_b32(a.z, b.z, c.z, 0x80), __builtin_amdgcn_bitop3_b32(a.w, b.w, c.w, 0x80));
}

// Reduce-scatter of 8 per-lane values over a 64-lane wave in 10 shuffles:
// afterwards lane L with (L & 7) == 0 holds the wave total of v[(L >> 3) & 7].
__device__ __forceinline__ int wave_reduce_scatter8(const int v[8]) {
    const int lane = threadIdx.x & 63;
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    int w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int send = b5 ? v[j] : v[j + 4];
        const int keep = b5 ? v[j + 4] : v[j];
        w[j] = keep + __shfl_xor(send, 32);
    }
    int x[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int send = b4 ? w[j] : w[j + 2];
        const int keep = b4 ? w[j + 2] : w[j];
        x[j] = keep + __shfl_xor(send, 16);
    }
    const int send = b3 ? x[0] : x[1];
    const int keep = b3 ? x[1] : x[0];
    int y = keep + __shfl_xor(send, 8);
    y += __shfl_xor(y, 4);
    y += __shfl_xor(y, 2);
    y += __shfl_xor(y, 1);
    return y;
}

// Sums 8 per-lane counts (each < 2^12) over every 16-lane row of the wave with
// DPP row ops only (no LDS traffic).  Two columns share a 32-bit register (16
// bits each; a row total is < 2^16) through the reduction.  Afterwards the
// lanes with (lane & 3) == 0 hold, for group g = (lane >> 2) & 3, the row
// totals of columns 2g (low half) and 2g + 1 (high half).
//   step 1: partner 15 - l (row mirror); lanes 0-7 keep columns 0-3, 8-15 keep 4-7
//   step 2: partner 7 - (l & 7) (half mirror); (l & 4) keeps the upper pair
//   step 3: sum over the quad (xor 1, then xor 2)
__device__ __forceinline__ uint32_t row_reduce8_packed(const uint32_t v[8]) {
    const int lane = threadIdx.x & 63;
    const bool hi = lane & 8, h2 = lane & 4;
    uint32_t p[4];
#pragma unroll
    for (int i = 0; i < 4; i++) p[i] = v[2 * i] | (v[2 * i + 1] << 16);
    uint32_t q[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const uint32_t keep = hi ? p[2 + i] : p[i], send = hi ? p[i] : p[2 + i];
        q[i] = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x140, 0xF, 0xF, false);
    }
    const uint32_t keep = h2 ? q[1] : q[0], send = h2 ? q[0] : q[1];
    uint32_t r = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x141, 0xF, 0xF, false);
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    return r;
}

// The same for 4 columns: afterwards lanes with (lane & 7) == 0 hold the row
// totals of columns 0 and 1 (lane & 8 clear) or 2 and 3 (lane & 8 set), low
// half first.
//   step 1: partner 15 - l (row mirror); lanes 0-7 keep columns 0-1, 8-15 keep 2-3
//   step 2: partner 7 - (l & 7) (half mirror), both keep the same pair
//   step 3: sum over the quad (xor 1, then xor 2)
__device__ __forceinline__ uint32_t row_reduce4_packed(const uint32_t v[4]) {
    const int lane = threadIdx.x & 63;
    const bool hi = lane & 8;
    const uint32_t p0 = v[0] | (v[1] << 16), p1 = v[2] | (v[3] << 16);
    const uint32_t keep = hi ? p1 : p0, send = hi ? p0 : p1;
    uint32_t r = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x140, 0xF, 0xF, false);
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0x141, 0xF, 0xF, false);
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    return r;
}

// ---------------------------------------------------------------------------
// Plain per-node count.  grid.x = word chunks of 256*R uint4, grid.y = column
// blocks of cpb <= 64 columns.  Every block ANDs the k prefix rows over its
// chunk once (registers), then streams its columns 8*G at a time: the R*8*G
// loads of a batch are independent, so they are all in flight together.
// `item(i)` gives prefix item i (from device memory, or from the kernel's
// argument block).
template <int R, int G, class ItemFn>
__device__ __forceinline__ void count_node_body(const uint4 *__restrict__ bm, int64_t pitch4, ItemFn item, int k,
                                                int n_cols, int cpb, int32_t *__restrict__ out,
                                                int64_t spread_stride) {
    __shared__ int red[4][64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t base = (int64_t)blockIdx.x * (256 * R) + tid;
    const int c0 = blockIdx.y * cpb;
    const int c1 = min(n_cols, c0 + cpb);
    uint4 m[R];
    {
        const uint4 *row = bm + (int64_t)item(0) * pitch4;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t idx = base + r * 256;
            m[r] = idx < pitch4 ? row[idx] : make_uint4(0, 0, 0, 0);
        }
    }
    for (int i = 1; i < k; i++) {
        const uint4 *row = bm + (int64_t)item(i) * pitch4;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t idx = base + r * 256;
            if (idx < pitch4) m[r] = and4(m[r], row[idx]);
        }
    }
    for (int c = c0; c < c1; c += 8 * G) {
        int v[8 * G];
#pragma unroll
        for (int j = 0; j < 8 * G; j++) {
            uint32_t s = 0;
            if (c + j < c1) {
                const uint4 *row = bm + (int64_t)(c + j) * pitch4;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int64_t idx = base + r * 256;
                    if (idx < pitch4) s = popc_and4_acc(m[r], row[idx], s);
                }
            }
            v[j] = (int)s;
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int y = wave_reduce_scatter8(&v[8 * g]);
            if ((lane & 7) == 0) {
                const int col = c + 8 * g + ((lane >> 3) & 7);
                if (col < c1) red[wave][col - c0] = y;
            }
        }
    }
    __syncthreads();
    if (tid < c1 - c0) {
        const int s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
        // with a spread stride, blocks of different word chunks add into 8
        // copies of the counts (4 KB apart), so that the same-address atomics
        // of one column come from 1/8 of the chunks; the caller sums them
        int32_t *o = spread_stride ? out + (int64_t)(blockIdx.x & 7) * spread_stride : out;
        if (s) atomicAdd(&o[c0 + tid], s);
    }
}

template <int R, int G>
__global__ __launch_bounds__(256) void k_count_node(const uint4 *__restrict__ bm, int64_t pitch4,
                                                    const int32_t *__restrict__ items, int k, int n_cols,
                                                    int cpb, int32_t *__restrict__ out, int64_t spread_stride) {
    count_node_body<R, G>(bm, pitch4, [&](int i) { return items[i]; }, k, n_cols, cpb, out, spread_stride);
}

// The same with the prefix passed in the kernel's argument block (nothing to
// copy to device memory first).
template <int R, int G>
__global__ __launch_bounds__(256) void k_count_node_args(const uint4 *__restrict__ bm, int64_t pitch4,
                                                         NodeItems items, int n_cols, int cpb,
                                                         int32_t *__restrict__ out, int64_t spread_stride) {
    count_node_body<R, G>(bm, pitch4, [&](int i) { return items.id[i]; }, items.n, n_cols, cpb, out,
                          spread_stride);
}

// shape = R * 100 + cpb (R in {1, 2, 4}, cpb in {8, 16, 32, 64}).  0 picks
// from measurements on MI355X: (2, 32) once that gives >= 8 blocks per CU,
// else (1, 16); short, narrow nodes are latency-bound, and more blocks with
// fewer columns each keep more loads in flight.
static void node_shape_of(int64_t pitch4, int32_t n_cols, int n_cu, int shape, int &R, int &cpb) {
    auto blocks = [&](int r, int c) { return ((pitch4 + 256 * r - 1) / (256 * r)) * ((n_cols + c - 1) / c); };
    R = shape / 100;
    cpb = shape % 100;
    if (shape == 0) {
        R = 1;
        cpb = 16;
        if (blocks(2, 32) >= 8LL * n_cu) {
            R = 2;
            cpb = 32;
        }
    }
}

void launch_count_node(const uint64_t *bm, int64_t pitch_words, const int32_t *items, int32_t k,
                       int32_t n_cols, int32_t *out, int64_t spread_stride, int n_cu, int shape, hipStream_t s) {
    const int64_t pitch4 = pitch_words / 2;
    int R, cpb;
    node_shape_of(pitch4, n_cols, n_cu, shape, R, cpb);
    const dim3 grid((unsigned)((pitch4 + 256 * R - 1) / (256 * R)), (unsigned)((n_cols + cpb - 1) / cpb));
    const uint4 *b4 = reinterpret_cast<const uint4 *>(bm);
    if (cpb == 8) {
        if (R == 4)
            hipLaunchKernelGGL((k_count_node<4, 1>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
        else if (R == 2)
            hipLaunchKernelGGL((k_count_node<2, 1>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
        else
            hipLaunchKernelGGL((k_count_node<1, 1>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
    } else if (R == 4) {
        hipLaunchKernelGGL((k_count_node<4, 1>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
    } else if (R == 2) {
        hipLaunchKernelGGL((k_count_node<2, 2>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
    } else {
        hipLaunchKernelGGL((k_count_node<1, 2>), grid, dim3(256), 0, s, b4, pitch4, items, k, n_cols, cpb, out, spread_stride);
    }
}

void launch_count_node_args(const uint64_t *bm, int64_t pitch_words, const NodeItems &items, int32_t n_cols,
                            int32_t *out, int64_t spread_stride, int n_cu, int shape, hipStream_t s) {
    const int64_t pitch4 = pitch_words / 2;
    int R, cpb;
    node_shape_of(pitch4, n_cols, n_cu, shape, R, cpb);
    const dim3 grid((unsigned)((pitch4 + 256 * R - 1) / (256 * R)), (unsigned)((n_cols + cpb - 1) / cpb));
    const uint4 *b4 = reinterpret_cast<const uint4 *>(bm);
    if (cpb == 8 || R == 4) {
        if (R == 4)
            hipLaunchKernelGGL((k_count_node_args<4, 1>), grid, dim3(256), 0, s, b4, pitch4, items, n_cols, cpb, out, spread_stride);
        else if (R == 2)
            hipLaunchKernelGGL((k_count_node_args<2, 1>), grid, dim3(256), 0, s, b4, pitch4, items, n_cols, cpb, out, spread_stride);
        else
            hipLaunchKernelGGL((k_count_node_args<1, 1>), grid, dim3(256), 0, s, b4, pitch4, items, n_cols, cpb, out, spread_stride);
    } else if (R == 2) {
        hipLaunchKernelGGL((k_count_node_args<2, 2>), grid, dim3(256), 0, s, b4, pitch4, items, n_cols, cpb, out, spread_stride);
    } else {
        hipLaunchKernelGGL((k_count_node_args<1, 2>), grid, dim3(256), 0, s, b4, pitch4, items, n_cols, cpb, out, spread_stride);
    }
}

// out[e] = sum over j < 8 of copies[j * stride + e], e < n.
__global__ __launch_bounds__(256) void k_sum_copies8(const int32_t *__restrict__ copies, int64_t stride, int32_t n,
                                                     int32_t *__restrict__ out) {
    const int e = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (e >= n) return;
    int32_t t = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) t += copies[(int64_t)j * stride + e];
    out[e] = t;
}

void launch_sum_copies8(const int32_t *copies, int64_t stride, int32_t n, int32_t *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sum_copies8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, copies, stride, n, out);
}

// ---------------------------------------------------------------------------
// Fused persistent wave kernel.  A unit of work is (word chunk s, tile t): one
// 64-lane wave ANDs the prefix of tile t's row over chunk s into registers and
// counts the tile's columns against it.  Units are spread so that every chunk
// s is handled by blocks of one XCD (blockIdx % 8 == s % 8): the chunk of each
// item row is then fetched into that XCD's L2 once and reused by every row of
// the wave.
template <int RW>
__global__ __launch_bounds__(256) void k_count_wave(const uint4 *__restrict__ bm, int64_t pitch4,
                                                    WaveDesc d) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x & 7;
    const int64_t gw = (int64_t)(blockIdx.x >> 3) * 4 + wave;
    const int64_t GW = (int64_t)(gridDim.x >> 3) * 4;
    const int S = d.n_chunks;
    const int nchunks_x = xcd < S ? (S - xcd + 7) / 8 : 0;
    const int64_t U = (int64_t)nchunks_x * d.n_tiles;
    for (int64_t v = gw; v < U; v += GW) {
        const int s = (int)(v / d.n_tiles) * 8 + xcd;
        const int t = (int)(v % d.n_tiles);
        const WaveTile tile = d.tiles[t];
        const int r = tile.row;
        const int64_t base = (int64_t)s * (64 * RW) + lane;
        const int ib = d.item_off[r], ie = d.item_off[r + 1];
        uint4 m[RW];
        {
            const uint4 *row = bm + (int64_t)d.item_ids[ib] * pitch4;
#pragma unroll
            for (int q = 0; q < RW; q++) {
                const int64_t idx = base + q * 64;
                m[q] = idx < pitch4 ? row[idx] : make_uint4(0, 0, 0, 0);
            }
        }
        // prefix rows two at a time, so that both rows' loads are in flight
        // together
        int i = ib + 1;
        for (; i + 1 < ie; i += 2) {
            const uint4 *r0 = bm + (int64_t)d.item_ids[i] * pitch4;
            const uint4 *r1 = bm + (int64_t)d.item_ids[i + 1] * pitch4;
#pragma unroll
            for (int q = 0; q < RW; q++) {
                const int64_t idx = base + q * 64;
                if (idx < pitch4) m[q] = and3_4(m[q], r0[idx], r1[idx]);
            }
        }
        for (; i < ie; i++) {
            const uint4 *row = bm + (int64_t)d.item_ids[i] * pitch4;
#pragma unroll
            for (int q = 0; q < RW; q++) {
                const int64_t idx = base + q * 64;
                if (idx < pitch4) m[q] = and4(m[q], row[idx]);
            }
        }
        const int ob = d.col_off[r];
        // complement tile: the columns are survivor ids; live = exclusion bit
        // clear, or the row's first prefix item (its count is the prefix's
        // support).  Wave-uniform, so dead groups and columns are skipped by
        // scalar branches.
        uint64_t live = ~0ull;
        if (tile.compl_) {
            live = ~d.excl[(int64_t)(tile.compl_ - 1) * d.excl_words + (tile.col_begin >> 6)];
            const int x0 = d.item_ids[ib] - tile.col_begin;
            if (x0 >= 0 && x0 < 64) live |= 1ull << x0;
        }
        for (int j = tile.col_begin; j < tile.col_end; j += 8) {
            const uint32_t live8 = tile.compl_ ? (uint32_t)(live >> (j - tile.col_begin)) & 0xffu : 0xffu;
            if (!live8) continue;
            int v8[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                uint32_t acc = 0;
                if (j + q < tile.col_end && ((live8 >> q) & 1u)) {
                    const int col = (d.col_ids && !tile.compl_) ? d.col_ids[ob + j + q] : j + q;
                    const uint4 *row = bm + (int64_t)col * pitch4;
#pragma unroll
                    for (int qq = 0; qq < RW; qq++) {
                        const int64_t idx = base + qq * 64;
                        if (idx < pitch4) acc = popc_and4_acc(m[qq], row[idx], acc);
                    }
                }
                v8[q] = (int)acc;
            }
            const int y = wave_reduce_scatter8(v8);
            if ((lane & 7) == 0) {
                const int q = (lane >> 3) & 7;
                if (j + q < tile.col_end && y) atomicAdd(&d.out[ob + j + q], y);
            }
        }
    }
}

int wave_chunk_words(int rw) { return 64 * rw * 2; }

// One wave per complement row: lanes walk the survivors 64 at a time (one
// exclusion word per step, read once for the wave), count the live columns
// that reach min_count, and reduce over the wave.
__global__ __launch_bounds__(256) void k_compl_reduce(WaveDesc d, const int32_t *__restrict__ compl_rows,
                                                      int32_t n_compl, int32_t n_cols, int32_t min_count,
                                                      int32_t *__restrict__ res) {
    const int lane = threadIdx.x & 63;
    const int j = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (j >= n_compl) return;
    const int r = compl_rows[j];
    const int x0 = d.item_ids[d.item_off[r]];
    const int32_t *full = d.out + d.col_off[r];
    const uint64_t *ex = d.excl + (int64_t)j * d.excl_words;
    int n = 0;
    for (int w = 0; w < d.excl_words; w++) {
        const int e = w * 64 + lane;
        const bool live = e < n_cols && e != x0 && !((ex[w] >> lane) & 1ull);
        if (live && full[e] >= min_count) n++;
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (lane == 0) {
        res[2 * j] = full[x0];
        res[2 * j + 1] = n;
    }
}

void plan_compl_reduce(const WaveDesc &d, const int32_t *compl_rows, int32_t n_compl, int32_t n_cols,
                       int32_t min_count, int32_t *res, KernelLaunch &L) {
    L.func = (const void *)k_compl_reduce;
    L.grid = dim3((unsigned)std::max(1, (n_compl + 3) / 4));
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(d);
    L.arg(compl_rows);
    L.arg(n_compl);
    L.arg(n_cols);
    L.arg(min_count);
    L.arg(res);
}

void plan_count_wave(const uint64_t *bm, int64_t pitch_words, const WaveDesc &d, int grid_blocks, int rw,
                     KernelLaunch &L) {
    switch (rw) {
        case 1: L.func = (const void *)k_count_wave<1>; break;
        case 2: L.func = (const void *)k_count_wave<2>; break;
        case 8: L.func = (const void *)k_count_wave<8>; break;
        default: L.func = (const void *)k_count_wave<4>; break;
    }
    L.grid = dim3((unsigned)grid_blocks);
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(reinterpret_cast<const uint4 *>(bm));
    L.arg((int64_t)(pitch_words / 2));
    L.arg(d);
}

void launch_count_wave(const uint64_t *bm, int64_t pitch_words, const WaveDesc &d, int grid_blocks,
                       int rw, hipStream_t s) {
    KernelLaunch L;
    plan_count_wave(bm, pitch_words, d, grid_blocks, rw, L);
    L.launch(s);
}

// ---------------------------------------------------------------------------
// LDS-staged fused wave kernel.  Workgroup g (1024 threads = 16 waves) handles
// word chunks g, g+G, ...; per chunk it copies the chunk of every slot (item
// row) the wave uses into LDS with 128-bit loads, four in flight per thread,
// then its waves take tiles from a workgroup counter in LDS (the host orders
// the tiles by falling cost, so the waves reach the end of a chunk together).
// A tile's record is fetched two tiles ahead and its prefix and column slot
// numbers one tile ahead, lane-parallel, then broadcast with readlane, so the
// column loop issues only LDS reads and VALU work: prefix AND into registers,
// then 8 (then 4) columns at a time, each column's words read while the one
// before it is counted, a DPP reduction over each 16-lane row, and one LDS
// atomic per (row, column).  Partial counts leave the workgroup once, as one
// global atomic per non-zero output.  Global traffic per wave is one pass over
// the slots' rows whatever the number of rows.
constexpr int kLdsThreads = 1024;
constexpr int kLdsWaves = kLdsThreads / 64;

// LDS beyond the slot chunks: the partial counts (n_out ints, rounded up to
// 16 bytes), the slots' bitmap rows and the tile counter.
__host__ __device__ constexpr int lds_acc_ints(int n_out) { return (n_out + 3) / 4 * 4; }
constexpr int kLdsTailBytes = kLdsMaxSlots * 4 + 16;

template <int RW>
__device__ __forceinline__ void lds_col_load(uint4 (&x)[RW], const uint4 *lds, int slot, int lane) {
    const uint4 *row = lds + slot * (64 * RW);
#pragma unroll
    for (int q = 0; q < RW; q++) x[q] = row[q * 64 + lane];
}

template <int RW>
__device__ __forceinline__ uint32_t lds_col_count(const uint4 (&m)[RW], const uint4 (&x)[RW]) {
    uint32_t a = 0;
#pragma unroll
    for (int q = 0; q < RW; q++) a = popc_and4_acc(m[q], x[q], a);
    return a;
}

template <int RW>
__global__ __launch_bounds__(kLdsThreads) void k_count_wave_lds(const uint4 *__restrict__ bm, int64_t pitch4,
                                                                WaveLdsDesc d, int64_t n_chunks) {
    extern __shared__ uint4 lds[];
    constexpr int C4 = 64 * RW;   // uint4 per slot per chunk
    int *acc = reinterpret_cast<int *>(lds + (size_t)d.n_slots * C4);
    int *srow = acc + lds_acc_ints(d.n_out);
    int *counter = srow + kLdsMaxSlots;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n_tiles = d.n_tiles, n_stage = d.n_slots * C4;
    for (int i = tid; i < d.n_out; i += kLdsThreads) acc[i] = 0;
    for (int i = tid; i < d.n_slots; i += kLdsThreads) srow[i] = d.slot_rows[i];
    // the next tile of this chunk for the calling wave (wave-uniform)
    auto grab = [&]() {
        int g = 0;
        if (lane == 0) g = atomicAdd(counter, 1);
        return __builtin_amdgcn_readfirstlane(g);
    };
    auto tile_at = [&](int t) {
        LdsTile r = {0, 0, 0, 0};
        if (t < n_tiles) r = d.tiles[t];
        r.ib = __builtin_amdgcn_readfirstlane(r.ib);
        r.k = __builtin_amdgcn_readfirstlane(r.k);
        r.cb = __builtin_amdgcn_readfirstlane(r.cb);
        r.nc = __builtin_amdgcn_readfirstlane(r.nc);
        return r;
    };
    for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        __syncthreads();
        if (tid == 0) *counter = kLdsWaves;   // tiles 0 .. 15 go to waves 0 .. 15
        const int64_t base = c * C4;
        for (int i0 = tid; i0 < n_stage; i0 += 4 * kLdsThreads) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * kLdsThreads;
                v[u] = make_uint4(0, 0, 0, 0);
                if (i < n_stage) {
                    const int64_t g = base + i % C4;
                    if (g < pitch4) v[u] = bm[(int64_t)srow[i / C4] * pitch4 + g];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * kLdsThreads;
                if (i < n_stage) lds[i] = v[u];
            }
        }
        __syncthreads();
        // t: tile in work (record tl, slots pslot / cslot); t1: the next
        // (record n1, slots loaded during this tile); t2: the one after
        // (record loaded during this tile)
        int t = wave;
        LdsTile tl = tile_at(t);
        int pslot = 0, cslot = 0;
        if (t < n_tiles) {
            pslot = lane < tl.k ? d.item_slots[tl.ib + lane] : 0;
            cslot = lane < tl.nc ? d.col_slots[tl.cb + lane] : 0;
        }
        int t1 = grab();
        LdsTile n1 = tile_at(t1);
        while (t < n_tiles) {
            const int t2 = t1 < n_tiles ? grab() : n_tiles;
            const LdsTile n2 = tile_at(t2);
            int npslot = 0, ncslot = 0;
            if (t1 < n_tiles) {
                npslot = lane < n1.k ? d.item_slots[n1.ib + lane] : 0;
                ncslot = lane < n1.nc ? d.col_slots[n1.cb + lane] : 0;
            }
            const int k = tl.k, nc = tl.nc;   // nc <= 64 (tile_cols)
            uint4 m[RW];
            lds_col_load<RW>(m, lds, __builtin_amdgcn_readlane(pslot, 0), lane);
            // prefix rows two at a time, so that both rows' reads are in
            // flight together (the two ANDs fold into one 3-input op)
            int i = 1;
            for (; i + 1 < k && i + 1 < 64; i += 2) {
                uint4 a0[RW], a1[RW];
                lds_col_load<RW>(a0, lds, __builtin_amdgcn_readlane(pslot, i), lane);
                lds_col_load<RW>(a1, lds, __builtin_amdgcn_readlane(pslot, i + 1), lane);
#pragma unroll
                for (int q = 0; q < RW; q++) m[q] = and3_4(m[q], a0[q], a1[q]);
            }
            for (; i < k; i++) {
                const int sl = i < 64 ? __builtin_amdgcn_readlane(pslot, i) : d.item_slots[tl.ib + i];
                uint4 a0[RW];
                lds_col_load<RW>(a0, lds, __builtin_amdgcn_readfirstlane(sl), lane);
#pragma unroll
                for (int q = 0; q < RW; q++) m[q] = and4(m[q], a0[q]);
            }
            // full groups of eight columns, then groups of four for the rest
            // (the last one stops at the tile's end); column j + 1's words
            // are read while column j is counted
            uint4 x[RW];
            if (nc > 0) lds_col_load<RW>(x, lds, __builtin_amdgcn_readlane(cslot, 0), lane);
            int j = 0;
            for (; j + 8 <= nc; j += 8) {
                uint32_t v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    uint4 y[RW];   // the next column (the last one again at the tile's end)
                    lds_col_load<RW>(y, lds, __builtin_amdgcn_readlane(cslot, min(j + q + 1, nc - 1)), lane);
                    v[q] = lds_col_count<RW>(m, x);
#pragma unroll
                    for (int qq = 0; qq < RW; qq++) x[qq] = y[qq];
                }
                const uint32_t r = row_reduce8_packed(v);
                if ((lane & 3) == 0) {
                    const int c0 = j + 2 * ((lane >> 2) & 3);
                    const int lo = (int)(r & 0xFFFFu), hi = (int)(r >> 16);
                    if (lo) atomicAdd(&acc[tl.cb + c0], lo);
                    if (hi) atomicAdd(&acc[tl.cb + c0 + 1], hi);
                }
            }
            for (; j < nc; j += 4) {
                uint32_t v[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (j + q >= nc) break;   // wave-uniform: columns past the tile's end are skipped
                    uint4 y[RW];   // the next column (the last one again at the tile's end)
                    lds_col_load<RW>(y, lds, __builtin_amdgcn_readlane(cslot, min(j + q + 1, nc - 1)), lane);
                    v[q] = lds_col_count<RW>(m, x);
#pragma unroll
                    for (int qq = 0; qq < RW; qq++) x[qq] = y[qq];
                }
                const uint32_t r = row_reduce4_packed(v);
                if ((lane & 7) == 0) {
                    const int c0 = j + ((lane & 8) ? 2 : 0);
                    const int lo = (int)(r & 0xFFFFu), hi = (int)(r >> 16);
                    if (c0 < nc && lo) atomicAdd(&acc[tl.cb + c0], lo);
                    if (c0 + 1 < nc && hi) atomicAdd(&acc[tl.cb + c0 + 1], hi);
                }
            }
            t = t1;
            tl = n1;
            pslot = npslot;
            cslot = ncslot;
            t1 = t2;
            n1 = n2;
        }
    }
    __syncthreads();
    for (int i = tid; i < d.n_out; i += kLdsThreads)
        if (acc[i]) atomicAdd(&d.out[i], acc[i]);
}

bool plan_count_wave_lds(const uint64_t *bm, int64_t pitch_words, const WaveLdsDesc &d, int n_cu, bool wide,
                         KernelLaunch &L) {
    if (d.n_slots < 1 || d.n_slots > kLdsMaxSlots || d.n_out > kLdsMaxOut) return false;
    const int64_t pitch4 = pitch_words / 2;
    const size_t tail_bytes = (size_t)lds_acc_ints(d.n_out) * 4 + kLdsTailBytes;
    auto bytes = [&](int rw) { return (size_t)d.n_slots * 64 * rw * 16 + tail_bytes; };
    // Narrow: two workgroups per CU (<= 80 KB each), the widest chunk that
    // fits.  Wide: the widest chunk (up to 4 x 128 bits per lane) that fits
    // one workgroup per CU, so each lane counts more bits per column between
    // reductions, at half the occupancy.
    constexpr size_t kHalf = 80 * 1024, kFull = 160 * 1024;
    // the kernels may take up to the full 160 KB: set once per device
    // (contexts on several threads or devices may plan at the same time)
    {
        static std::mutex mu;
        static uint64_t done = 0;   // bit per device id below 64
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        std::lock_guard<std::mutex> g(mu);
        if (dev >= 64 || !((done >> dev) & 1)) {
            const void *fn[4] = {(const void *)k_count_wave_lds<1>, (const void *)k_count_wave_lds<2>,
                                 (const void *)k_count_wave_lds<3>, (const void *)k_count_wave_lds<4>};
            for (const void *f : fn)
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFull) != hipSuccess)
                    return false;
            if (dev < 64) done |= 1ull << dev;
        }
    }
    int rw;
    if (wide) {
        rw = 4;
        while (rw > 1 && bytes(rw) > kFull) rw--;
        if (bytes(rw) > kFull) return false;
    } else {
        rw = 4;
        while (rw > 1 && bytes(rw) > kHalf) rw /= 2;
        if (bytes(rw) > kHalf && bytes(1) > kFull) return false;
    }
    const size_t lds_bytes = bytes(rw);
    const int64_t n_chunks = (pitch4 + 64 * rw - 1) / (64 * rw);
    const int per_cu = lds_bytes <= kHalf ? 2 : 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_chunks, (int64_t)n_cu * per_cu));
    L.func = rw == 4   ? (const void *)k_count_wave_lds<4>
             : rw == 3 ? (const void *)k_count_wave_lds<3>
             : rw == 2 ? (const void *)k_count_wave_lds<2>
                       : (const void *)k_count_wave_lds<1>;
    L.grid = dim3((unsigned)grid);
    L.block = dim3(kLdsThreads);
    L.shmem = (uint32_t)lds_bytes;
    L.reset();
    L.arg(reinterpret_cast<const uint4 *>(bm));
    L.arg(pitch4);
    L.arg(d);
    L.arg(n_chunks);
    return true;
}

bool launch_count_wave_lds(const uint64_t *bm, int64_t pitch_words, const WaveLdsDesc &d, int n_cu,
                           hipStream_t s) {
    KernelLaunch L;
    if (!plan_count_wave_lds(bm, pitch_words, d, n_cu, false, L)) return false;
    L.launch(s);
    return true;
}

// ---------------------------------------------------------------------------
// LDS-staged wave kernel for rows given as bit masks (at most 64 dense ids).
// The scheme of k_count_wave_lds -- workgroup g stages word chunks g, g+G, ...
// of every id into LDS (slot i = id i), its 16 waves take tiles from a
// counter in LDS -- with a tile that is (prefix mask, column mask, first
// output): its record is read one tile ahead with one vector load per lane
// and broadcast to scalar registers, the ids come from scalar bit scans, and
// every column group is straight-line code: groups of eight, then one group
// of four or eight for the rest, padded with the tile's last column (its
// surplus counts are not added).

// Counts `n` (<= G) columns of a group against the prefix m: ids from the
// column mask `cm` (consumed), padded with `last`; each column's words are
// read while the one before it is counted.
template <int RW, int G>
__device__ __forceinline__ void mask_group(const uint4 (&m)[RW], const uint4 *lds, uint64_t &cm, int n, int last,
                                           int lane, uint32_t (&v)[G]) {
    int s[G];
#pragma unroll
    for (int q = 0; q < G; q++) {
        if (q < n) {
            s[q] = __builtin_ctzll(cm);
            cm &= cm - 1;
        } else {
            s[q] = last;
        }
    }
    uint4 x[RW];
    lds_col_load<RW>(x, lds, s[0], lane);
#pragma unroll
    for (int q = 0; q < G; q++) {
        uint4 y[RW];
        if (q + 1 < G) lds_col_load<RW>(y, lds, s[q + 1], lane);
        v[q] = lds_col_count<RW>(m, x);
#pragma unroll
        for (int qq = 0; qq < RW; qq++) x[qq] = y[qq];
    }
}

// The last G (1 .. 7) columns of a tile: counted exactly, reduced with the
// 4- or 8-column row reduction over zeros, one LDS atomic per (row of lanes,
// column) into out[0 .. G).
template <int RW, int G>
__device__ __forceinline__ void mask_tail(const uint4 (&m)[RW], const uint4 *lds, uint64_t &cm, int lane,
                                          int *out) {
    uint32_t v[G];
    mask_group<RW, G>(m, lds, cm, G, 0, lane, v);
    if (G <= 4) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < G; q++) w[q] = v[q];
        const uint32_t r = row_reduce4_packed(w);
        if ((lane & 7) == 0) {
            const int q0 = (lane & 8) ? 2 : 0;
            const int lo = (int)(r & 0xFFFFu), hi = (int)(r >> 16);
            if (q0 < G && lo) atomicAdd(&out[q0], lo);
            if (q0 + 1 < G && hi) atomicAdd(&out[q0 + 1], hi);
        }
    } else {
        uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < G; q++) w[q] = v[q];
        const uint32_t r = row_reduce8_packed(w);
        if ((lane & 3) == 0) {
            const int q0 = 2 * ((lane >> 2) & 3);
            const int lo = (int)(r & 0xFFFFu), hi = (int)(r >> 16);
            if (q0 < G && lo) atomicAdd(&out[q0], lo);
            if (q0 + 1 < G && hi) atomicAdd(&out[q0 + 1], hi);
        }
    }
}

template <int RW>
__global__ __launch_bounds__(kLdsThreads) void k_count_wave_mask(const uint4 *__restrict__ bm, int64_t pitch4,
                                                                 WaveMaskDesc d, int64_t n_chunks) {
    extern __shared__ uint4 lds[];
    constexpr int C4 = 64 * RW;   // uint4 per id per chunk
    int *acc = reinterpret_cast<int *>(lds + (size_t)d.n_items * C4);
    int *counter = acc + lds_acc_ints(d.n_out);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n_tiles = d.n_tiles, n_stage = d.n_items * C4;
    if (tid == 0) atomicMax(&d.stamp[0], ~(unsigned long long)wall_clock64());
    for (int i = tid; i < d.n_out; i += kLdsThreads) acc[i] = 0;
    auto grab = [&]() {
        int g = 0;
        if (lane == 0) g = atomicAdd(counter, 1);
        return __builtin_amdgcn_readfirstlane(g);
    };
    // a tile's record, read by every lane (one cache line) and broadcast
    auto tile_at = [&](int t) {
        uint4 a = make_uint4(0, 0, 0, 0), b = make_uint4(0, 0, 0, 0);
        if (t < n_tiles) {
            const uint4 *p = reinterpret_cast<const uint4 *>(d.tiles + t);
            a = p[0];
            b = p[1];
        }
        // (readfirstlane returns a signed int: widen through uint32_t, so a
        // set bit 31 of a low word does not spread into the high word)
        auto u = [](uint32_t x) { return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(x); };
        MaskTile r;
        r.pm = u(a.x) | (u(a.y) << 32);
        r.cm = u(a.z) | (u(a.w) << 32);
        r.ob = __builtin_amdgcn_readfirstlane(b.x);
        r.k = __builtin_amdgcn_readfirstlane(b.y);
        r.nc = __builtin_amdgcn_readfirstlane(b.z);
        r.pad = 0;
        return r;
    };
    for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        __syncthreads();
        if (tid == 0) *counter = kLdsWaves;   // tiles 0 .. 15 go to waves 0 .. 15
        const int64_t base = c * C4;
        for (int i0 = tid; i0 < n_stage; i0 += 4 * kLdsThreads) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * kLdsThreads;
                v[u] = make_uint4(0, 0, 0, 0);
                if (i < n_stage) {
                    const int64_t g = base + i % C4;
                    if (g < pitch4) v[u] = bm[(int64_t)(i / C4) * pitch4 + g];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * kLdsThreads;
                if (i < n_stage) lds[i] = v[u];
            }
        }
        __syncthreads();
        int t = wave;
        MaskTile tl = tile_at(t);
        while (t < n_tiles) {
            const int tn = grab();
            const MaskTile nx = tile_at(tn);   // in flight while this tile is counted
            // prefix AND, two ids per step (both reads in flight, one 3-input AND)
            uint64_t pm = tl.pm;
            uint4 m[RW];
            lds_col_load<RW>(m, lds, __builtin_ctzll(pm), lane);
            pm &= pm - 1;
            while (pm) {
                const int a = __builtin_ctzll(pm);
                pm &= pm - 1;
                if (pm) {
                    const int b = __builtin_ctzll(pm);
                    pm &= pm - 1;
                    uint4 a0[RW], a1[RW];
                    lds_col_load<RW>(a0, lds, a, lane);
                    lds_col_load<RW>(a1, lds, b, lane);
#pragma unroll
                    for (int q = 0; q < RW; q++) m[q] = and3_4(m[q], a0[q], a1[q]);
                } else {
                    uint4 a0[RW];
                    lds_col_load<RW>(a0, lds, a, lane);
#pragma unroll
                    for (int q = 0; q < RW; q++) m[q] = and4(m[q], a0[q]);
                }
            }
            uint64_t cm = tl.cm;
            const int nc = tl.nc, ob = tl.ob;
            const int last = 63 - __builtin_clzll(cm);
            int j = 0;
            for (; j + 8 <= nc; j += 8) {
                uint32_t v[8];
                mask_group<RW, 8>(m, lds, cm, 8, last, lane, v);
                const uint32_t r = row_reduce8_packed(v);
                if ((lane & 3) == 0) {
                    const int c0 = ob + j + 2 * ((lane >> 2) & 3);
                    const int lo = (int)(r & 0xFFFFu), hi = (int)(r >> 16);
                    if (lo) atomicAdd(&acc[c0], lo);
                    if (hi) atomicAdd(&acc[c0 + 1], hi);
                }
            }
            // the last 1 .. 7 columns: a group of exactly that many (no
            // padding columns counted), reduced as a group of 4 or 8
            switch (nc - j) {
                case 1: mask_tail<RW, 1>(m, lds, cm, lane, acc + ob + j); break;
                case 2: mask_tail<RW, 2>(m, lds, cm, lane, acc + ob + j); break;
                case 3: mask_tail<RW, 3>(m, lds, cm, lane, acc + ob + j); break;
                case 4: mask_tail<RW, 4>(m, lds, cm, lane, acc + ob + j); break;
                case 5: mask_tail<RW, 5>(m, lds, cm, lane, acc + ob + j); break;
                case 6: mask_tail<RW, 6>(m, lds, cm, lane, acc + ob + j); break;
                case 7: mask_tail<RW, 7>(m, lds, cm, lane, acc + ob + j); break;
                default: break;
            }
            t = tn;
            tl = nx;
        }
    }
    __syncthreads();
    for (int i = tid; i < d.n_out; i += kLdsThreads)
        if (acc[i]) atomicAdd(&d.out[i], acc[i]);
    __syncthreads();
    if (tid == 0) atomicMax(&d.stamp[1], (unsigned long long)wall_clock64());
}

bool plan_count_wave_mask(const uint64_t *bm, int64_t pitch_words, const WaveMaskDesc &d, int n_cu, bool wide,
                          KernelLaunch &L) {
    if (d.n_items < 1 || d.n_items > kMaskMaxItems || d.n_out < 1 || d.n_out > kLdsMaxOut || d.n_tiles < 1)
        return false;
    const int64_t pitch4 = pitch_words / 2;
    const size_t tail_bytes = (size_t)lds_acc_ints(d.n_out) * 4 + 16;
    auto bytes = [&](int rw) { return (size_t)d.n_items * 64 * rw * 16 + tail_bytes; };
    constexpr size_t kHalf = 80 * 1024, kFull = 160 * 1024;
    {
        static std::mutex mu;
        static uint64_t done = 0;   // bit per device id below 64
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        std::lock_guard<std::mutex> g(mu);
        if (dev >= 64 || !((done >> dev) & 1)) {
            const void *fn[4] = {(const void *)k_count_wave_mask<1>, (const void *)k_count_wave_mask<2>,
                                 (const void *)k_count_wave_mask<3>, (const void *)k_count_wave_mask<4>};
            for (const void *f : fn)
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFull) != hipSuccess)
                    return false;
            if (dev < 64) done |= 1ull << dev;
        }
    }
    // wide: the widest chunk (up to 4 x 128 bits per lane per id) with one
    // workgroup per CU; narrow: the widest with two per CU
    const size_t cap = wide ? kFull : kHalf;
    int rw = 4;
    while (rw > 1 && bytes(rw) > cap) rw--;
    if (bytes(rw) > cap) return false;
    const size_t lds_bytes = bytes(rw);
    const int64_t n_chunks = (pitch4 + 64 * rw - 1) / (64 * rw);
    const int per_cu = lds_bytes <= kHalf ? 2 : 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(n_chunks, (int64_t)n_cu * per_cu));
    L.func = rw == 4   ? (const void *)k_count_wave_mask<4>
             : rw == 3 ? (const void *)k_count_wave_mask<3>
             : rw == 2 ? (const void *)k_count_wave_mask<2>
                       : (const void *)k_count_wave_mask<1>;
    L.grid = dim3((unsigned)grid);
    L.block = dim3(kLdsThreads);
    L.shmem = (uint32_t)lds_bytes;
    L.reset();
    L.arg(reinterpret_cast<const uint4 *>(bm));
    L.arg(pitch4);
    L.arg(d);
    L.arg(n_chunks);
    return true;
}

// ---------------------------------------------------------------------------
__global__ void k_zero_i32(int32_t *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

void plan_zero_i32(int32_t *p, int64_t n, KernelLaunch &L) {
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    L.func = (const void *)k_zero_i32;
    L.grid = dim3((unsigned)blocks);
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(p);
    L.arg(n);
}

// A wave's descriptor block from pinned host memory (mapped, uncached on the
// device) into device memory, 16 bytes per thread: the first node of a wave,
// so no copy call and no second stream are involved.
__global__ void k_stage16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void plan_stage16(const void *src, void *dst, int64_t bytes, KernelLaunch &L) {
    const int64_t n16 = (bytes + 15) / 16;
    L.func = (const void *)k_stage16;
    L.grid = dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, 256)));
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(reinterpret_cast<const uint4 *>(src));
    L.arg(reinterpret_cast<uint4 *>(dst));
    L.arg(n16);
}

// A wave's counts out to pinned host memory (mapped), and its device output
// region back to zero for the slot's next wave: the last node of a wave.
// 16-byte loads and stores (the buffers are 256-byte aligned), so each wave
// of the kernel writes 1 KB runs to the host; every element is read and
// zeroed by the same thread.
__global__ void k_drain_i32(int32_t *__restrict__ dev, int32_t *__restrict__ host, int64_t n_copy, int64_t n_zero) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x, t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint4 *d4 = reinterpret_cast<uint4 *>(dev);
    uint4 *h4 = reinterpret_cast<uint4 *>(host);
    const int64_t c4 = n_copy / 4, z4 = n_zero / 4;   // n_copy <= n_zero
    for (int64_t i = t0; i < z4; i += stride) {
        const uint4 v = d4[i];
        if (i < c4) {
            h4[i] = v;
        } else if (i == c4) {   // the copy's last partial 16 bytes
            const int64_t b = 4 * i;
            if (b < n_copy) host[b] = (int32_t)v.x;
            if (b + 1 < n_copy) host[b + 1] = (int32_t)v.y;
            if (b + 2 < n_copy) host[b + 2] = (int32_t)v.z;
        }
        d4[i] = make_uint4(0, 0, 0, 0);
    }
    for (int64_t i = 4 * z4 + t0; i < n_zero; i += stride) {
        if (i < n_copy) host[i] = dev[i];
        dev[i] = 0;
    }
    __threadfence_system();
}

void plan_drain_i32(int32_t *dev, int32_t *host, int64_t n_copy, int64_t n_zero, KernelLaunch &L) {
    const int64_t nz = std::max(n_copy, n_zero);
    const int64_t n4 = std::max<int64_t>(1, (nz + 3) / 4);
    L.func = (const void *)k_drain_i32;
    L.grid = dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n4 + 255) / 256, 1024)));
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(dev);
    L.arg(host);
    L.arg(n_copy);
    L.arg(nz);
}

void launch_zero_i32(int32_t *p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    KernelLaunch L;
    plan_zero_i32(p, n, L);
    L.launch(s);
}

// ---------------------------------------------------------------------------
// Packed matrix count: 64x64 output tile per 256-thread block; thread (tr, tc)
// owns rows tr+16i and columns tc+16j (i, j < 4), a layout whose LDS reads are
// conflict-free with the 17-word row stride.  K is staged 16 u64 words at a
// time; split_k blocks along grid.z accumulate with atomics.
constexpr int PK_KW = 16;
__global__ __launch_bounds__(256) void k_popmat_packed(const uint64_t *__restrict__ A,
                                                       const uint64_t *__restrict__ B,
                                                       int32_t *__restrict__ C, int64_t b, int64_t d,
                                                       int64_t pitch, int64_t k_per_split, int atomic) {
    __shared__ uint64_t As[64][PK_KW + 1];
    __shared__ uint64_t Bs[64][PK_KW + 1];
    const int tid = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.y * 64, col0 = (int64_t)blockIdx.x * 64;
    const int64_t kbeg = (int64_t)blockIdx.z * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);
    const int tr = tid >> 4, tc = tid & 15;
    uint32_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    for (int64_t kw = kbeg; kw < kend; kw += PK_KW) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * 256;
            const int r = e / PK_KW, w = e % PK_KW;
            const bool kin = kw + w < kend;
            As[r][w] = (row0 + r < b && kin) ? A[(row0 + r) * pitch + kw + w] : 0ull;
            Bs[r][w] = (col0 + r < d && kin) ? B[(col0 + r) * pitch + kw + w] : 0ull;
        }
        __syncthreads();
#pragma unroll 4
        for (int w = 0; w < PK_KW; w++) {
            uint64_t a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[tr + 16 * i][w];
#pragma unroll
            for (int j = 0; j < 4; j++) bb[j] = Bs[tc + 16 * j][w];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = popc_and64_acc(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t r = row0 + tr + 16 * i;
        if (r >= b) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t c = col0 + tc + 16 * j;
            if (c >= d) continue;
            if (atomic) atomicAdd(&C[r * d + c], (int)acc[i][j]);
            else C[r * d + c] = (int)acc[i][j];
        }
    }
}

void launch_popmat_packed(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                          int64_t pitch_words, int split_k, hipStream_t s) {
    if (split_k < 1) split_k = 1;
    int64_t kps = (pitch_words + split_k - 1) / split_k;
    kps = (kps + PK_KW - 1) / PK_KW * PK_KW;
    const int64_t zs = (pitch_words + kps - 1) / kps;
    dim3 grid((unsigned)((d + 63) / 64), (unsigned)((b + 63) / 64), (unsigned)zs);
    hipLaunchKernelGGL(k_popmat_packed, grid, dim3(256), 0, s, A, B, C, b, d, pitch_words, kps, zs > 1 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea & Flood 2014).  One thread per u64 word: 32 draws,
// two 32-bit halves per draw, bit = draw < thr32.
__device__ __forceinline__ uint64_t sm64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__global__ void k_fill_bernoulli(uint64_t *dst, int64_t rows, int64_t n_bits, int64_t pitch,
                                 uint64_t seed, uint32_t matrix, uint32_t thr32) {
    const uint64_t km = sm64_mix(seed + kGamma * ((uint64_t)matrix + 1));
    const int64_t total = rows * pitch;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / pitch, w = i % pitch;
        uint64_t word = 0;
        const int64_t base = w * 64;
        if (base < n_bits) {
            const uint64_t kr = sm64_mix(km + kGamma * ((uint64_t)row + 1));
#pragma unroll 4
            for (int p = 0; p < 32; p++) {
                const uint64_t pair = (uint64_t)w * 32 + (uint64_t)p;
                const uint64_t x = sm64_mix(kr + kGamma * (pair + 1));
                word |= (uint64_t)((uint32_t)x < thr32) << (2 * p);
                word |= (uint64_t)((uint32_t)(x >> 32) < thr32) << (2 * p + 1);
            }
            const int64_t keep = n_bits - base;
            if (keep < 64) word &= (1ull << keep) - 1;
        }
        dst[i] = word;
    }
}

void launch_fill_bernoulli(uint64_t *dst, int64_t rows, int64_t n_bits, int64_t pitch_words,
                           uint64_t seed, uint32_t matrix, uint32_t thr32, hipStream_t s) {
    const int64_t total = rows * pitch_words;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_bernoulli, dim3((unsigned)blocks), dim3(256), 0, s, dst, rows, n_bits,
                       pitch_words, seed, matrix, thr32);
}

// ---------------------------------------------------------------------------
// Replication (data/make_replica.py on the device): dst bit t of a row is base
// bit (t mod base_bits) for t < replicate*base_bits, zero beyond.
__device__ __forceinline__ uint64_t get_bits64(const uint64_t *row, int64_t pitch, int64_t p) {
    const int64_t w = p >> 6;
    const int sh = (int)(p & 63);
    const uint64_t lo = w < pitch ? row[w] : 0ull;
    if (sh == 0) return lo;
    const uint64_t hi = (w + 1) < pitch ? row[w + 1] : 0ull;
    return (lo >> sh) | (hi << (64 - sh));
}

__global__ void k_replicate(const uint64_t *base, int64_t base_pitch, int64_t base_bits,
                            uint64_t *dst, int64_t dst_pitch, int32_t rows, int32_t replicate) {
    const int64_t total = (int64_t)rows * dst_pitch;
    const int64_t dst_bits = base_bits * replicate;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / dst_pitch, w = i % dst_pitch;
        const uint64_t *src = base + row * base_pitch;
        const int64_t t0 = w * 64;
        uint64_t word = 0;
        if (t0 < dst_bits) {
            if (base_bits >= 64) {
                const int64_t p = t0 % base_bits;
                const int64_t n1 = base_bits - p;   // bits before the wrap
                uint64_t x = get_bits64(src, base_pitch, p);
                if (n1 < 64) {
                    x &= (1ull << n1) - 1;
                    x |= get_bits64(src, base_pitch, 0) << n1;
                }
                word = x;
            } else {
                for (int j = 0; j < 64; j++) {
                    const int64_t p = (t0 + j) % base_bits;
                    word |= ((src[p >> 6] >> (p & 63)) & 1ull) << j;
                }
            }
            const int64_t keep = dst_bits - t0;
            if (keep < 64) word &= (1ull << keep) - 1;
        }
        dst[i] = word;
    }
}

void launch_replicate_rows(const uint64_t *base, int64_t base_pitch, int64_t base_bits,
                           uint64_t *dst, int64_t dst_pitch, int32_t rows, int32_t replicate,
                           hipStream_t s) {
    const int64_t total = (int64_t)rows * dst_pitch;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_replicate, dim3((unsigned)blocks), dim3(256), 0, s, base, base_pitch,
                       base_bits, dst, dst_pitch, rows, replicate);
}

// F1: one block per row (batch_proto.cu:57-65).
__global__ __launch_bounds__(256) void k_row_popcount(const uint4 *bm, int64_t pitch4, int32_t *out) {
    __shared__ int red[4];
    const uint4 *row = bm + (int64_t)blockIdx.x * pitch4;
    uint32_t s = 0;
    for (int64_t i = threadIdx.x; i < pitch4; i += 256) {
        const uint4 x = row[i];
        s += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
    }
    int v = (int)s;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

void launch_row_popcount(const uint64_t *bm, int64_t pitch_words, int32_t rows, int32_t *out,
                         hipStream_t s) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_row_popcount, dim3(rows), dim3(256), 0, s,
                       reinterpret_cast<const uint4 *>(bm), pitch_words / 2, out);
}

}  // namespace tfim
