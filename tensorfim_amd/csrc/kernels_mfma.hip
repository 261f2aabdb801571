// Bits-as-int8 GEMM on CDNA4 matrix cores.
//
// CDNA4 has no 1-bit matrix op (the reference uses b1 wmma with AND/POPC,
// kernel_bmma.cu:636-758, bmma_bench.cu:50-152).  Here each bit becomes a 0/1
// int8, so C[i][j] = sum_t a_it * b_jt is exactly popcount(A_i AND B_j), and
// v_mfma_i32_32x32x32_i8 accumulates it in int32 without rounding.
//
// Tile: 128 rows x 128 columns per 256-thread workgroup; waves in a 2x2 grid
// each own 64x64 = 2x2 MFMA tiles (64 accumulator registers).  K is staged
// 256 bits per step: the packed words are loaded with 128-bit loads and
// unpacked to bytes straight into LDS (32 KB for A, 32 KB for B), with a
// 16-byte-slot XOR swizzle so the ds_read_b128 of a 32-row fragment hits
// distinct banks.
//
// Operand layout.  For v_mfma_i32_32x32x32_i8, lane l supplies 16 bytes of
// row (l % 32) of A and the same 16 k-positions of column (l % 32) of B; the
// half l / 32 selects which 16 of the 32 k.  The instruction sums over all 64
// (half, byte) pairs, so any consistent assignment of k to (half, byte) is
// exact; here byte b of half h is k = 16h + b.  The accumulator layout is the
// 32x32 one of the CDNA3/CDNA4 ISA: register q of lane l holds
// C[8*(q/4) + 4*(l/32) + (q%4)][l % 32].
#include "tfim_internal.h"

namespace tfim {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// 4 bits -> 4 bytes of 0/1: bit j of the nibble lands on bit 8j.
__device__ __forceinline__ uint32_t nib2bytes(uint32_t x) {
    return ((x & 0xFu) * 0x00204081u) & 0x01010101u;
}

// 16 bits -> 16 bytes (one LDS slot).
__device__ __forceinline__ uint4 unpack16(uint32_t x16) {
    return make_uint4(nib2bytes(x16), nib2bytes(x16 >> 4), nib2bytes(x16 >> 8), nib2bytes(x16 >> 12));
}

constexpr int MM_T = 128;     // rows and columns per workgroup tile
constexpr int MM_SLOTS = 16;  // 16-byte slots per row per 256-bit K step

__global__ __launch_bounds__(256) void k_popmat_mfma(const uint64_t *__restrict__ A,
                                                     const uint64_t *__restrict__ B,
                                                     int32_t *__restrict__ C, int64_t b, int64_t d,
                                                     int64_t pitch, int64_t k_per_split, int atomic,
                                                     int tiles_n) {
    __shared__ uint4 As[MM_T * MM_SLOTS];
    __shared__ uint4 Bs[MM_T * MM_SLOTS];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // Consecutive tiles of one row band go to one XCD (blocks are dealt to the
    // 8 XCDs round-robin), so the band's A rows are shared through its L2.
    const int64_t nblk = (int64_t)gridDim.x;
    const int64_t bid = blockIdx.x;
    const int64_t per_xcd = nblk / 8;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * per_xcd + bid / 8 : bid;
    const int64_t tile_m = lin / tiles_n, tile_n = lin % tiles_n;
    const int64_t row0 = tile_m * MM_T, col0 = tile_n * MM_T;
    const int64_t kbeg = (int64_t)blockIdx.y * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);

    // loader role: thread -> (row, half of the 256-bit step)
    const int lr = tid >> 1, lh = tid & 1;
    const bool a_ok = row0 + lr < b, b_ok = col0 + lr < d;
    const uint4 *Arow = reinterpret_cast<const uint4 *>(A + (row0 + lr) * pitch);
    const uint4 *Brow = reinterpret_cast<const uint4 *>(B + (col0 + lr) * pitch);

    const int wr = wave >> 1, wc = wave & 1;
    v16i acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0;

    for (int64_t kw = kbeg; kw < kend; kw += 4) {
        const int64_t u = (kw >> 1) + lh;   // uint4 index of this thread's 128 bits
        const uint4 av = a_ok ? Arow[u] : make_uint4(0, 0, 0, 0);
        const uint4 bv = b_ok ? Brow[u] : make_uint4(0, 0, 0, 0);
        const uint32_t aw[4] = {av.x, av.y, av.z, av.w};
        const uint32_t bw[4] = {bv.x, bv.y, bv.z, bv.w};
        __syncthreads();   // previous step's fragment reads are done
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int slot = lh * 8 + j;
            const uint32_t xa = aw[j >> 1] >> ((j & 1) * 16);
            const uint32_t xb = bw[j >> 1] >> ((j & 1) * 16);
            As[lr * MM_SLOTS + (slot ^ (lr & 15))] = unpack16(xa);
            Bs[lr * MM_SLOTS + (slot ^ (lr & 15))] = unpack16(xb);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const int slot = kk * 2 + (lane >> 5);
            v4i af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wr * 64 + i * 32 + (lane & 31);
                const uint4 x = As[r * MM_SLOTS + (slot ^ (r & 15))];
                af[i] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = wc * 64 + j * 32 + (lane & 31);
                const uint4 x = Bs[c * MM_SLOTS + (slot ^ (c & 15))];
                bf[j] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }

#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int64_t r = row0 + wr * 64 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int64_t c = col0 + wc * 64 + j * 32 + (lane & 31);
                if (r < b && c < d) {
                    if (atomic) atomicAdd(&C[r * d + c], acc[i][j][q]);
                    else C[r * d + c] = acc[i][j][q];
                }
            }
}

// ---------------------------------------------------------------------------
// Large-tile variant: 256x256 per workgroup, each wave 128x128 (4x4 MFMA
// tiles, 256 int32 accumulators), so every 1 KB fragment read from LDS feeds
// 4 MFMAs instead of 2.  Thread t owns row t of the A tile and row t of the B
// tile: it keeps one 1024-bit block of each in registers (eight 128-bit
// loads), prefetches the next block while the current one is consumed, and
// per 128-bit sub-stage unpacks 16 bytes of each row into a double-buffered
// LDS stage (2 x 64 KB, 16-byte slots XOR-swizzled by row & 7).  One barrier
// per sub-stage: a stage is rewritten only after every wave has passed the
// barrier that follows its last read.
constexpr int M2_T = 256;
constexpr int M2_SLOTS = 8;   // 16-byte slots per row per 128-bit sub-stage

__global__ __launch_bounds__(256) void k_popmat_mfma256(const uint64_t *__restrict__ A,
                                                        const uint64_t *__restrict__ B,
                                                        int32_t *__restrict__ C, int64_t b, int64_t d,
                                                        int64_t pitch, int64_t k_per_split, int atomic,
                                                        int tiles_n) {
    extern __shared__ uint4 sm[];
    uint4 *As = sm;                              // [2][256][8]
    uint4 *Bs = sm + 2 * M2_T * M2_SLOTS;        // [2][256][8]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int64_t nblk = (int64_t)gridDim.x, bid = blockIdx.x;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int64_t row0 = (lin / tiles_n) * M2_T, col0 = (lin % tiles_n) * M2_T;
    const int64_t kbeg = (int64_t)blockIdx.y * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);
    const bool a_ok = row0 + tid < b, b_ok = col0 + tid < d;
    const uint4 *Arow = reinterpret_cast<const uint4 *>(A + (row0 + tid) * pitch);
    const uint4 *Brow = reinterpret_cast<const uint4 *>(B + (col0 + tid) * pitch);
    const uint4 z4 = make_uint4(0, 0, 0, 0);

    const int wr = wave >> 1, wc = wave & 1;
    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0;

    uint4 ra[8], rb[8], na[8], nb[8];
    if (kbeg < kend) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            ra[u] = a_ok ? Arow[(kbeg >> 1) + u] : z4;
            rb[u] = b_ok ? Brow[(kbeg >> 1) + u] : z4;
        }
    }
    int buf = 0;
    for (int64_t kw = kbeg; kw < kend; kw += 16) {
        const bool more = kw + 16 < kend;
#pragma unroll
        for (int u = 0; u < 8; u++) {   // prefetch the next 1024-bit block
            na[u] = (more && a_ok) ? Arow[((kw + 16) >> 1) + u] : z4;
            nb[u] = (more && b_ok) ? Brow[((kw + 16) >> 1) + u] : z4;
        }
#pragma unroll
        for (int sub = 0; sub < 8; sub++) {
            uint4 *as = As + buf * (M2_T * M2_SLOTS) + tid * M2_SLOTS;
            uint4 *bs = Bs + buf * (M2_T * M2_SLOTS) + tid * M2_SLOTS;
            const uint32_t aw[4] = {ra[sub].x, ra[sub].y, ra[sub].z, ra[sub].w};
            const uint32_t bw[4] = {rb[sub].x, rb[sub].y, rb[sub].z, rb[sub].w};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                as[j ^ (tid & 7)] = unpack16(aw[j >> 1] >> ((j & 1) * 16));
                bs[j ^ (tid & 7)] = unpack16(bw[j >> 1] >> ((j & 1) * 16));
            }
            __syncthreads();
            const uint4 *ab = As + buf * (M2_T * M2_SLOTS);
            const uint4 *bb = Bs + buf * (M2_T * M2_SLOTS);
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int slot = kk * 2 + (lane >> 5);
                v4i af[4], bf[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = wr * 128 + i * 32 + (lane & 31);
                    const uint4 x = ab[r * M2_SLOTS + (slot ^ (r & 7))];
                    af[i] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int c = wc * 128 + j * 32 + (lane & 31);
                    const uint4 x = bb[c * M2_SLOTS + (slot ^ (c & 7))];
                    bf[j] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
            buf ^= 1;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            ra[u] = na[u];
            rb[u] = nb[u];
        }
    }

#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int64_t r = row0 + wr * 128 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int64_t c = col0 + wc * 128 + j * 32 + (lane & 31);
                if (r < b && c < d) {
                    if (atomic) atomicAdd(&C[r * d + c], acc[i][j][q]);
                    else C[r * d + c] = acc[i][j][q];
                }
            }
}

// Software-pipelined 256x256 variant: while a wave runs the 64 MFMAs of
// sub-stage g from LDS buffer g%2, it unpacks sub-stage g+1 (2 of the 8 slots
// of its A row and of its B row per K=32 step) into the other buffer, so the
// VALU/DS work of the unpack sits in the MFMA shadow of the same wave (one
// wave per SIMD at 256 accumulators).  One barrier per sub-stage orders the
// buffer swap; the raw 1024-bit blocks are prefetched one block ahead.
__device__ __forceinline__ void unpack_pair(uint4 *dst, int row, int j, const uint4 &raw) {
    const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
    dst[j ^ (row & 7)] = unpack16(w[j >> 1] >> ((j & 1) * 16));
}

__global__ __launch_bounds__(256) void k_popmat_mfma256p(const uint64_t *__restrict__ A,
                                                         const uint64_t *__restrict__ B,
                                                         int32_t *__restrict__ C, int64_t b, int64_t d,
                                                         int64_t pitch, int64_t k_per_split, int atomic,
                                                         int tiles_n) {
    extern __shared__ uint4 sm[];
    uint4 *As = sm;                              // [2][256][8]
    uint4 *Bs = sm + 2 * M2_T * M2_SLOTS;        // [2][256][8]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int64_t nblk = (int64_t)gridDim.x, bid = blockIdx.x;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int64_t row0 = (lin / tiles_n) * M2_T, col0 = (lin % tiles_n) * M2_T;
    const int64_t kbeg = (int64_t)blockIdx.y * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);
    const bool a_ok = row0 + tid < b, b_ok = col0 + tid < d;
    const uint4 *Arow = reinterpret_cast<const uint4 *>(A + (row0 + tid) * pitch);
    const uint4 *Brow = reinterpret_cast<const uint4 *>(B + (col0 + tid) * pitch);
    const uint4 z4 = make_uint4(0, 0, 0, 0);
    const int wr = wave >> 1, wc = wave & 1;
    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0;
    const int64_t nsub = (kend - kbeg) / 2;   // 128-bit sub-stages (uint4 per row)
    if (nsub > 0) {
        uint4 ra[8], rb[8], na[8], nb[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            ra[u] = a_ok ? Arow[(kbeg >> 1) + u] : z4;
            rb[u] = b_ok ? Brow[(kbeg >> 1) + u] : z4;
            na[u] = z4;
            nb[u] = z4;
        }
        {   // prologue: sub-stage 0 into buffer 0
            uint4 *as = As + tid * M2_SLOTS, *bs = Bs + tid * M2_SLOTS;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                unpack_pair(as, tid, j, ra[0]);
                unpack_pair(bs, tid, j, rb[0]);
            }
        }
        __syncthreads();
        for (int64_t g0 = 0; g0 < nsub; g0 += 8) {   // one 1024-bit block per iteration
            const bool more = g0 + 8 < nsub;
            const int64_t nxt = (kbeg >> 1) + g0 + 8;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                na[u] = (more && a_ok) ? Arow[nxt + u] : z4;
                nb[u] = (more && b_ok) ? Brow[nxt + u] : z4;
            }
#pragma unroll
            for (int sub = 0; sub < 8; sub++) {
                const int buf = sub & 1;   // g0 is a multiple of 8: parity of g is parity of sub
                const uint4 *ab = As + buf * (M2_T * M2_SLOTS);
                const uint4 *bb = Bs + buf * (M2_T * M2_SLOTS);
                uint4 *an = As + (buf ^ 1) * (M2_T * M2_SLOTS) + tid * M2_SLOTS;
                uint4 *bn = Bs + (buf ^ 1) * (M2_T * M2_SLOTS) + tid * M2_SLOTS;
                const bool has_next = sub < 7 || more;
                const uint4 &rna = sub < 7 ? ra[sub + 1] : na[0];
                const uint4 &rnb = sub < 7 ? rb[sub + 1] : nb[0];
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int slot = kk * 2 + (lane >> 5);
                    v4i af[4], bf[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int r = wr * 128 + i * 32 + (lane & 31);
                        const uint4 x = ab[r * M2_SLOTS + (slot ^ (r & 7))];
                        af[i] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int c = wc * 128 + j * 32 + (lane & 31);
                        const uint4 x = bb[c * M2_SLOTS + (slot ^ (c & 7))];
                        bf[j] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
                    }
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
                    if (has_next) {
                        unpack_pair(an, tid, 2 * kk, rna);
                        unpack_pair(an, tid, 2 * kk + 1, rna);
                        unpack_pair(bn, tid, 2 * kk, rnb);
                        unpack_pair(bn, tid, 2 * kk + 1, rnb);
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                ra[u] = na[u];
                rb[u] = nb[u];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int64_t r = row0 + wr * 128 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int64_t c = col0 + wc * 128 + j * 32 + (lane & 31);
                if (r < b && c < d) {
                    if (atomic) atomicAdd(&C[r * d + c], acc[i][j][q]);
                    else C[r * d + c] = acc[i][j][q];
                }
            }
}

void launch_popmat_mfma256(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                           int64_t pitch_words, int split_k, hipStream_t s, bool pipelined) {
    if (split_k < 1) split_k = 1;
    int64_t kps = (pitch_words + split_k - 1) / split_k;
    kps = (kps + 15) / 16 * 16;
    const int64_t zs = (pitch_words + kps - 1) / kps;
    const int64_t tm = (b + M2_T - 1) / M2_T, tn = (d + M2_T - 1) / M2_T;
    const size_t lds = (size_t)4 * M2_T * M2_SLOTS * sizeof(uint4);   // 128 KB
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute((const void *)k_popmat_mfma256, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipFuncSetAttribute((const void *)k_popmat_mfma256p, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    dim3 grid((unsigned)(tm * tn), (unsigned)zs);
    if (pipelined)
        hipLaunchKernelGGL(k_popmat_mfma256p, grid, dim3(256), lds, s, A, B, C, b, d, pitch_words, kps,
                           zs > 1 ? 1 : 0, (int)tn);
    else
        hipLaunchKernelGGL(k_popmat_mfma256, grid, dim3(256), lds, s, A, B, C, b, d, pitch_words, kps,
                           zs > 1 ? 1 : 0, (int)tn);
}

void launch_popmat_mfma(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                        int64_t pitch_words, int split_k, hipStream_t s) {
    if (split_k < 1) split_k = 1;
    int64_t kps = (pitch_words + split_k - 1) / split_k;
    kps = (kps + 3) / 4 * 4;
    const int64_t zs = (pitch_words + kps - 1) / kps;
    const int64_t tm = (b + MM_T - 1) / MM_T, tn = (d + MM_T - 1) / MM_T;
    dim3 grid((unsigned)(tm * tn), (unsigned)zs);
    hipLaunchKernelGGL(k_popmat_mfma, grid, dim3(256), 0, s, A, B, C, b, d, pitch_words, kps,
                       zs > 1 ? 1 : 0, (int)tn);
}

}  // namespace tfim
