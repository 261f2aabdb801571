// Bits-as-int8 GEMM on CDNA4 matrix cores.
//
// CDNA4 has no 1-bit matrix op (the reference uses b1 wmma with AND/POPC,
// kernel_bmma.cu:636-758, bmma_bench.cu:50-152).  Here each bit becomes a 0/1
// int8, so C[i][j] = sum_t a_it * b_jt is exactly popcount(A_i AND B_j), and
// v_mfma_i32_32x32x32_i8 accumulates it in int32 without rounding.
//
// Two kernels:
//   k_popmat_mfma     128 x 128 outputs per workgroup, bits expanded to bytes
//                     into LDS (every wave reads bytes back);
//   k_popmat_mfma_rp  256 x 128 outputs per workgroup, packed bits in LDS and
//                     each lane expanding its own operands in registers (the
//                     default: LDS traffic per MFMA is 1/8 of the first).
//
// Operand layout.  For v_mfma_i32_32x32x32_i8, lane l supplies 16 bytes of
// row (l % 32) of A and the same 16 k-positions of column (l % 32) of B; the
// half l / 32 selects which 16 of the 32 k.  The instruction sums over all 64
// (half, byte) pairs, so any consistent assignment of k to (half, byte) is
// exact.  The accumulator layout is the 32x32 one of the CDNA3/CDNA4 ISA:
// register q of lane l holds C[8*(q/4) + 4*(l/32) + (q%4)][l % 32].
#include <mutex>

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

// 128 x 128 tile: waves in a 2x2 grid each own 64x64 = 2x2 MFMA tiles (64
// accumulator registers).  K is staged 256 bits per step: the packed words are
// loaded with 128-bit loads and unpacked to bytes straight into LDS (32 KB for
// A, 32 KB for B), byte b of half h being k = 16h + b, with a 16-byte-slot XOR
// swizzle so the ds_read_b128 of a 32-row fragment hits distinct banks.
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
// Register-expanded variant.  LDS holds the PACKED bits (512 per row per
// stage, double-buffered, one barrier per stage); each lane expands its own
// operand bytes in registers, so LDS carries 1/8 of the bytes of the 128x128
// kernel.  A 32-bit word w gives four operand dwords per parity p: dword t
// holds the bits 8i + q (i = 0..3, q = 4p + t), one bit per byte.  Operand A
// keeps each bit in place (weight 2^q, one AND); operand B rotates it to the
// complementary position (weight 2^(6-q), rotate + AND); bit 7 goes to
// position 3 on both sides.  Every matching bit therefore contributes exactly
// 64 and the accumulator is 64 x popcount (recovered with a shift; a split
// holds at most 2^24 bits, so an accumulator stays below 2^30).  All byte
// values are 0..64, inside the signed int8 range.
//
// K map.  Per stage a row has 16 words of 32 bits.  At MFMA step s (0..15)
// lane half h supplies word 8h + s/2 with parity p = s & 1: every bit of the
// stage is used exactly once, by the same (step, half, byte) in A and B.
//
// Tile: waves in a 2x2 grid, each owning TI x TJ fragments of 32x32, so a
// workgroup covers (64 TI) x (64 TJ); TI = 4, TJ = 2 gives 256 x 128 with
// 128 accumulators per lane and two waves per SIMD.  The A side has more
// fragments, so it gets the cheaper expansion.
constexpr int RP_STRIDE = 20;   // u32 per LDS row: 16 data + 4 pad (bank spread)

// A side: bit 8i + q (q < 7) stays in place, byte value 2^q (one AND);
// q = 7 goes to bit 8i + 3 (value 8).
__device__ __forceinline__ v4i rp_expand_a(uint32_t w, int parity) {
    v4i r;
    if (parity == 0) {
        r[0] = (int)(w & 0x01010101u);
        r[1] = (int)(w & 0x02020202u);
        r[2] = (int)(w & 0x04040404u);
        r[3] = (int)(w & 0x08080808u);
    } else {
        r[0] = (int)(w & 0x10101010u);
        r[1] = (int)(w & 0x20202020u);
        r[2] = (int)(w & 0x40404040u);
        r[3] = (int)(__builtin_amdgcn_alignbit(w, w, 4) & 0x08080808u);
    }
    return r;
}

// B side: bit 8i + q (q < 7) goes to bit 8i + 6 - q, byte value 2^(6-q), so
// every matching pair contributes 2^q * 2^(6-q) = 64; q = 7 goes to bit
// 8i + 3 (8 * 8 = 64).  alignbit(w, w, n) rotates right by n.
__device__ __forceinline__ v4i rp_expand_b(uint32_t w, int parity) {
    v4i r;
    if (parity == 0) {
        r[0] = (int)(__builtin_amdgcn_alignbit(w, w, 26) & 0x40404040u);
        r[1] = (int)(__builtin_amdgcn_alignbit(w, w, 28) & 0x20202020u);
        r[2] = (int)(__builtin_amdgcn_alignbit(w, w, 30) & 0x10101010u);
        r[3] = (int)(w & 0x08080808u);
    } else {
        r[0] = (int)(__builtin_amdgcn_alignbit(w, w, 2) & 0x04040404u);
        r[1] = (int)(__builtin_amdgcn_alignbit(w, w, 4) & 0x02020202u);
        r[2] = (int)(__builtin_amdgcn_alignbit(w, w, 6) & 0x01010101u);
        r[3] = (int)(__builtin_amdgcn_alignbit(w, w, 4) & 0x08080808u);
    }
    return r;
}

// ---------------------------------------------------------------------------
// The same kernel on the fp4 matrix cores (FMT = kFmtFp4).  CDNA4's
// v_mfma_f32_32x32x64_f8f6f4 with both operands in e2m1 takes 32 k-values of
// 4 bits per lane and runs at twice the int8 rate.  A 0/1 bit becomes one of
// the exact e2m1 values 0, 0.5, 1.0 or 2.0 (nibble 0b0001 = 0.5, the one
// subnormal; 0b0010 = 1.0; 0b0100 = 2.0; the sign bit is never set): a 32-bit
// word w gives four operand dwords, dword g holding the bits 4n + g (n =
// 0..7) in nibble n.  Side A keeps bits 4n, 4n+1, 4n+2 in place (0.5, 1.0,
// 2.0; one AND each) and moves 4n+3 to 4n+2 (2.0); side B places each at the
// reciprocal value (2.0, 1.0, 0.5, 0.5), so every matching pair contributes
// exactly 1.0 and the fp32 accumulator holds the popcount itself, an integer
// of at most 2^24 (a split holds at most 2^24 bits): nothing rounds.
//
// K map.  One MFMA step consumes a whole 32-bit word per lane: at step s
// (0..7) of a stage lane half h supplies word 8h + s of the row, the same
// word in A and B, so every bit of the stage is used exactly once.
constexpr int kFmtI8 = 0, kFmtFp4 = 1;
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v8i fp4_expand_a(uint32_t w) {
    v8i r = {0, 0, 0, 0, 0, 0, 0, 0};
    r[0] = (int)(w & 0x11111111u);          // bit 4n   -> 0.5
    r[1] = (int)(w & 0x22222222u);          // bit 4n+1 -> 1.0
    r[2] = (int)(w & 0x44444444u);          // bit 4n+2 -> 2.0
    r[3] = (int)((w >> 1) & 0x44444444u);   // bit 4n+3 -> 2.0
    return r;
}

__device__ __forceinline__ v8i fp4_expand_b(uint32_t w) {
    v8i r = {0, 0, 0, 0, 0, 0, 0, 0};
    r[0] = (int)((w << 2) & 0x44444444u);   // bit 4n   -> 2.0
    r[1] = (int)(w & 0x22222222u);          // bit 4n+1 -> 1.0
    r[2] = (int)((w >> 2) & 0x11111111u);   // bit 4n+2 -> 0.5
    r[3] = (int)((w >> 3) & 0x11111111u);   // bit 4n+3 -> 0.5
    return r;
}

template <int FMT> struct RpAcc { typedef v16i type; };
template <> struct RpAcc<kFmtFp4> { typedef v16f type; };

template <int TI, int TJ, int FMT = kFmtI8>
__global__ __launch_bounds__(256, 2) void k_popmat_mfma_rp(const uint64_t *__restrict__ A,
                                                        const uint64_t *__restrict__ B,
                                                        int32_t *__restrict__ C, int64_t b, int64_t d,
                                                        int64_t pitch, int64_t k_per_split, int atomic,
                                                        int tiles_n) {
    constexpr int RA = 64 * TI, RB = 64 * TJ;          // rows of A / B per workgroup
    constexpr int LA = RA * 4 / 256, LB = RB * 4 / 256;  // uint4 loads per thread per stage
    __shared__ uint32_t As[2][RA * RP_STRIDE];
    __shared__ uint32_t Bs[2][RB * RP_STRIDE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t nblk = (int64_t)gridDim.x, bid = blockIdx.x;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int64_t row0 = (lin / tiles_n) * RA, col0 = (lin % tiles_n) * RB;
    const int64_t kbeg = (int64_t)blockIdx.y * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);

    // loader: load j of thread t is uint4 part (idx & 3) of row (idx >> 2), idx = 256 j + t.
    // Rows past the edge load the last row again: an output depends only on
    // its own A row and B row, and the outputs of those rows are never
    // written, so their LDS copies need no zeroing.  A load that is always
    // from the global buffer stays a global load (a choice between the
    // buffer and a zero constant would make it a flat load through a private
    // copy of the constant, and waits on LDS traffic would then also wait for
    // the loads in flight); nothing reads the loaded registers before the
    // stage is written to LDS, so the loads stay in flight during the MFMAs
    // of the current stage.
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(1))) u4v *gptr4;
    u4v ra[LA], rb[LB];
    auto gload = [&](int64_t kw) {
#pragma unroll
        for (int j = 0; j < LA; j++) {
            const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
            ra[j] = *(gptr4)(A + std::min<int64_t>(row0 + r, b - 1) * pitch + kw + 2 * part);
        }
#pragma unroll
        for (int j = 0; j < LB; j++) {
            const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
            rb[j] = *(gptr4)(B + std::min<int64_t>(col0 + r, d - 1) * pitch + kw + 2 * part);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < LA; j++) {
            const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
            *reinterpret_cast<u4v *>(&As[buf][r * RP_STRIDE + 4 * part]) = ra[j];
        }
#pragma unroll
        for (int j = 0; j < LB; j++) {
            const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
            *reinterpret_cast<u4v *>(&Bs[buf][r * RP_STRIDE + 4 * part]) = rb[j];
        }
    };

    typename RpAcc<FMT>::type acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0;

    if (kbeg < kend) {
        gload(kbeg);
        lstore(0);
    }
    __syncthreads();
    int buf = 0;
    for (int64_t kw = kbeg; kw < kend; kw += 8) {
        const bool more = kw + 8 < kend;
        if (more) gload(kw + 8);
#pragma unroll
        for (int hk = 0; hk < 2; hk++) {
            // this lane's 4 words (of its half) for steps 8 hk .. 8 hk + 7
            uint4 pa[TI], pb[TJ];
#pragma unroll
            for (int i = 0; i < TI; i++) {
                const int r = wr * (32 * TI) + i * 32 + l32;
                pa[i] = *reinterpret_cast<const uint4 *>(&As[buf][r * RP_STRIDE + 8 * half + 4 * hk]);
            }
#pragma unroll
            for (int j = 0; j < TJ; j++) {
                const int r = wc * (32 * TJ) + j * 32 + l32;
                pb[j] = *reinterpret_cast<const uint4 *>(&Bs[buf][r * RP_STRIDE + 8 * half + 4 * hk]);
            }
            if constexpr (FMT == kFmtI8) {
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++) {
                    const int wi = s2 >> 1, par = s2 & 1;
                    v4i ea[TI], eb[TJ];
#pragma unroll
                    for (int i = 0; i < TI; i++) {
                        const uint32_t w = wi == 0 ? pa[i].x : wi == 1 ? pa[i].y : wi == 2 ? pa[i].z : pa[i].w;
                        ea[i] = rp_expand_a(w, par);
                    }
#pragma unroll
                    for (int j = 0; j < TJ; j++) {
                        const uint32_t w = wi == 0 ? pb[j].x : wi == 1 ? pb[j].y : wi == 2 ? pb[j].z : pb[j].w;
                        eb[j] = rp_expand_b(w, par);
                    }
#pragma unroll
                    for (int i = 0; i < TI; i++)
#pragma unroll
                        for (int j = 0; j < TJ; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ea[i], eb[j], acc[i][j], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int wi = 0; wi < 4; wi++) {
                    v8i ea[TI], eb[TJ];
#pragma unroll
                    for (int i = 0; i < TI; i++)
                        ea[i] = fp4_expand_a(wi == 0 ? pa[i].x : wi == 1 ? pa[i].y : wi == 2 ? pa[i].z : pa[i].w);
#pragma unroll
                    for (int j = 0; j < TJ; j++)
                        eb[j] = fp4_expand_b(wi == 0 ? pb[j].x : wi == 1 ? pb[j].y : wi == 2 ? pb[j].z : pb[j].w);
                    // cbsz = blgp = 4: both operands e2m1; zero scale operands
                    // select the unscaled instruction
#pragma unroll
                    for (int i = 0; i < TI; i++)
#pragma unroll
                        for (int j = 0; j < TJ; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ea[i], eb[j], acc[i][j], 4,
                                                                                        4, 0, 0, 0, 0);
                }
            }
        }
        // keep the stores of the next stage after this stage's LDS reads, so
        // its global loads stay in flight during the MFMAs instead of being
        // waited on at the top of the loop
        asm volatile("" ::: "memory");
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int64_t r = row0 + wr * (32 * TI) + i * 32 + 8 * (q >> 2) + 4 * half + (q & 3);
                const int64_t c = col0 + wc * (32 * TJ) + j * 32 + l32;
                if (r < b && c < d) {
                    int32_t v;
                    if constexpr (FMT == kFmtI8) v = acc[i][j][q] >> 6;
                    else v = (int32_t)acc[i][j][q];
                    if (atomic) atomicAdd(&C[r * d + c], v);
                    else C[r * d + c] = v;
                }
            }
}

// ---------------------------------------------------------------------------
// fp4 matrix count with 256 x 256 outputs per workgroup (one workgroup per
// CU): NW waves in a 2 x (NW / 2) grid, each owning 4 x TJ fragments of
// 32 x 32.  The form in use is NW = 8, TJ = 2: two waves per SIMD, 128 fp32
// accumulators per lane, 4.25 VALU ops of operand expansion per MFMA that the
// other wave on the SIMD overlaps with its MFMAs.  (NW = 4, TJ = 4 -- one
// wave per SIMD, 256 accumulators, 3 VALU ops per MFMA that its own MFMAs do
// not hide -- measured 1.5 % slower.)  Operand expansion as in
// k_popmat_mfma_rp.
//
// With one workgroup per CU no other workgroup hides memory latency, so
// stages of SW words per row (SW = 8: 512 bits) are fetched straight into LDS by
// global_load_lds_dwordx4 (no registers hold data in flight), double-
// buffered: per stage, wait for this thread's loads, one barrier, read the
// whole stage from LDS into registers, then issue the next stage's loads
// into the other buffer (every wave read it before this barrier) and run
// the stage's MFMAs from registers while those loads travel.  Reading
// before issuing also keeps the compiler's own waits on the LDS loads from
// covering the next stage.
//
// K map: at step s (0..SW/2-1) of a stage, lane half h supplies word
// (SW/2) h + s of the row, the same word in A and B.
//
// LDS layout: a buffer holds 256 rows x 8 SW bytes per side, unpadded (one
// load instruction of a wave fills 1 KB contiguously, lane l at byte 16 l).
// A row has NC = SW / 2 chunks of 16 bytes; chunk c of row r sits in slot
// c ^ f(r), f(r) = (r / (32 / SW)) mod NC, so any 16 consecutive rows of one
// ds_read_b128 pass cover all 64 banks once.  Rows past the edge load the
// last row again (their outputs are never written).
template <int SW, int TJ, int NW = 4>
__global__ __launch_bounds__(64 * NW, (TJ == 4 || NW == 8) ? 1 : 2) void k_popmat_fp4_256(
    const uint64_t *__restrict__ A, const uint64_t *__restrict__ B, int32_t *__restrict__ C, int64_t b, int64_t d,
    int64_t pitch, int64_t k_per_split, int atomic, int tiles_n) {
    // NW waves in a 2 x (NW / 2) grid, each owning 4 x TJ fragments
    constexpr int WC = NW / 2;
    constexpr int TI = 4, TA = 256, TB = 32 * TJ * WC;   // fragments per wave in rows; rows of A and B per tile
    constexpr int RB = SW * 8;               // bytes per row per stage
    constexpr int NC = SW / 2;               // 16-byte chunks per row
    constexpr int RPI = 1024 / RB;           // rows per load instruction (1 KB)
    constexpr int NIA = TA / RPI / NW, NIB = TB / RPI / NW;   // load instructions per wave per side
    constexpr int FS = SW == 16 ? 1 : 2;     // f(r) = (r >> FS) & (NC - 1)
    constexpr int SIDEA = TA * RB, SIDEB = TB * RB;   // bytes per side per buffer
    constexpr int HW = NC / 2;               // uint4 per fragment per lane per stage
    __shared__ __attribute__((aligned(1024))) unsigned char lds[2 * (SIDEA + SIDEB)];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int wr = wave / WC, wc = wave % WC;
    const int64_t nblk = (int64_t)gridDim.x, bid = blockIdx.x;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int64_t row0 = (lin / tiles_n) * TA, col0 = (lin % tiles_n) * TB;
    const int64_t kbeg = (int64_t)blockIdx.y * k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + k_per_split);
    const int nst = (int)((kend - kbeg) / SW);   // kend - kbeg is a multiple of SW

    // byte offsets of this lane's loads from the tile's first row (256 rows
    // x pitch x 8 bytes < 2^32 for pitch < 2^21 words)
    const char *atile = reinterpret_cast<const char *>(A + row0 * pitch);
    const char *btile = reinterpret_cast<const char *>(B + col0 * pitch);
    uint32_t aoff[NIA], boff[NIB];
#pragma unroll
    for (int j = 0; j < NIA; j++) {
        const int r = (NW * j + wave) * RPI + lane / NC;
        const int chunk = (lane % NC) ^ ((r >> FS) & (NC - 1));
        aoff[j] = (uint32_t)((std::min<int64_t>(row0 + r, b - 1) - row0) * pitch * 8 + 16 * chunk);
    }
#pragma unroll
    for (int j = 0; j < NIB; j++) {
        const int r = (NW * j + wave) * RPI + lane / NC;
        const int chunk = (lane % NC) ^ ((r >> FS) & (NC - 1));
        boff[j] = (uint32_t)((std::min<int64_t>(col0 + r, d - 1) - col0) * pitch * 8 + 16 * chunk);
    }
    typedef const __attribute__((address_space(1))) void *gvoid;
    typedef __attribute__((address_space(3))) void *lvoid;
    auto issue = [&](int st) {
        const int64_t kb = (kbeg + SW * (int64_t)st) * 8;
        unsigned char *ba = lds + (st & 1) * (SIDEA + SIDEB), *bb = ba + SIDEA;
#pragma unroll
        for (int j = 0; j < NIA; j++)
            __builtin_amdgcn_global_load_lds((gvoid)(atile + kb + aoff[j]), (lvoid)(ba + (NW * j + wave) * 1024), 16, 0,
                                             0);
#pragma unroll
        for (int j = 0; j < NIB; j++)
            __builtin_amdgcn_global_load_lds((gvoid)(btile + kb + boff[j]), (lvoid)(bb + (NW * j + wave) * 1024), 16, 0,
                                             0);
    };

    v16f acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0.f;

    // this lane's rows r = 128 wr + 32 i + l32 (A), 32 TJ wc + 32 j + l32
    // (B) all have f(r) = f(l32)
    const int swz = (l32 >> FS) & (NC - 1);
    const int arow = (wr * 128 + l32) * RB, brow = (wc * 32 * TJ + l32) * RB;

    if (nst > 0) issue(0);
    for (int st = 0; st < nst; st++) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this thread's loads of stage st have landed
        __syncthreads();
        const unsigned char *ba = lds + (st & 1) * (SIDEA + SIDEB), *bb = ba + SIDEA;
        uint4 pa[HW][TI], pb[HW][TJ];
#pragma unroll
        for (int hk = 0; hk < HW; hk++) {
            const int slot = ((HW * half + hk) ^ swz) * 16;
#pragma unroll
            for (int i = 0; i < TI; i++)
                pa[hk][i] = *reinterpret_cast<const uint4 *>(ba + arow + i * 32 * RB + slot);
#pragma unroll
            for (int j = 0; j < TJ; j++)
                pb[hk][j] = *reinterpret_cast<const uint4 *>(bb + brow + j * 32 * RB + slot);
        }
        if (st + 1 < nst) issue(st + 1);
#pragma unroll
        for (int hk = 0; hk < HW; hk++) {
#pragma unroll
            for (int wi = 0; wi < 4; wi++) {
                v8i ea[TI], eb[TJ];
#pragma unroll
                for (int i = 0; i < TI; i++) {
                    const uint4 &qa = pa[hk][i];
                    ea[i] = fp4_expand_a(wi == 0 ? qa.x : wi == 1 ? qa.y : wi == 2 ? qa.z : qa.w);
                }
#pragma unroll
                for (int j = 0; j < TJ; j++) {
                    const uint4 &qb = pb[hk][j];
                    eb[j] = fp4_expand_b(wi == 0 ? qb.x : wi == 1 ? qb.y : wi == 2 ? qb.z : qb.w);
                }
#pragma unroll
                for (int i = 0; i < TI; i++)
#pragma unroll
                    for (int j = 0; j < TJ; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ea[i], eb[j], acc[i][j], 4, 4, 0,
                                                                                    0, 0, 0);
            }
        }
        // One wave per SIMD issues in order, so the operands of the next
        // step are built between this step's MFMAs (one MFMA, then three
        // VALU ops), not after them while the matrix core idles.
        if constexpr (TJ == 4) {
#pragma unroll
            for (int m = 0; m < HW * 4 * TI * TJ; m++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int64_t r = row0 + wr * (32 * TI) + i * 32 + 8 * (q >> 2) + 4 * half + (q & 3);
                const int64_t c = col0 + wc * (32 * TJ) + j * 32 + l32;
                if (r < b && c < d) {
                    const int32_t v = (int32_t)acc[i][j][q];
                    if (atomic) atomicAdd(&C[r * d + c], v);
                    else C[r * d + c] = v;
                }
            }
}

// ---------------------------------------------------------------------------
// Fused matrix-core wave: one launch per wave does what the reference splits
// into mask build, BMMA and split-k reduce (kernel_bmma.cu:619-768) and what
// k_wave_masks -> k_popmat_mfma_rp -> k_wave_gather do in three launches here,
// without the [rows][pitch] mask buffer and without the count matrix.
//
// Workgroup (row tile tm of RA = 64 TI wave rows, column tile tn of RB = 64 TJ
// survivors, word range blockIdx.y).  Per 512-bit stage:
//   - the words of every distinct prefix item of the wave ("slots") are
//     staged in LDS (Ps), two stages ahead of the MFMAs, double-buffered;
//   - each row's prefix AND is built from Ps into the A operand block (As),
//     one stage ahead, so row masks never leave LDS;
//   - the survivor tile's words are staged as the B operand (Bs);
//   - the MFMA loop of k_popmat_mfma_rp in its fp4 form (register-expanded
//     e2m1 operands on v_mfma_f32_32x32x64_f8f6f4, the fp32 accumulator
//     holding the popcount itself).
// One barrier per stage.  Epilogue: the tile's counts go to LDS, then
//   - explicit rows: the wave's listed outputs that fall in this tile
//     (entries grouped by column tile and row, resolved on the host) are
//     written, with atomics when the word axis is split;
//   - complement rows: the live columns' partial counts are added into the
//     row's full-count accumulator; the last workgroup of the row tile to
//     arrive (arrival counter) thresholds them: support of the prefix and
//     the number of non-excluded survivors reaching min_count.
constexpr int kFusedSlotLoads = 4;   // uint4 slot loads per thread per stage: at most 256 slots
constexpr int kFusedRegSlots = 16;   // prefix slots per row held in registers
// u32 per staged slot: 16 data + 4 pad, so that the 16 rows a wave builds at
// once read 16 different slots from disjoint LDS banks
constexpr int PS_STRIDE = 20;

template <int TI, int TJ, int FMT = kFmtFp4>
__global__ __launch_bounds__(256, 2) void k_wave_mfma(const uint64_t *__restrict__ bm, int64_t pitch,
                                                      WaveMfmaDesc d) {
    constexpr int RA = 64 * TI, RB = 64 * TJ, TS = RB + 1;
    extern __shared__ uint32_t lds_u32[];
    __shared__ int last_arrival, any_compl;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t nblk = (int64_t)gridDim.x, bid = blockIdx.x;
    const int64_t lin = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int tm = (int)(lin / d.tiles_n), tn = (int)(lin % d.tiles_n);
    const int row0 = tm * RA, col0 = tn * RB;
    const int rend = min(d.n_rows, row0 + RA);
    const int64_t kbeg = (int64_t)blockIdx.y * d.k_per_split;
    const int64_t kend = std::min<int64_t>(pitch, kbeg + d.k_per_split);

    uint32_t *As = lds_u32;                              // [2][RA][RP_STRIDE]
    uint32_t *Bs = As + 2 * RA * RP_STRIDE;              // [2][RB][RP_STRIDE]
    uint32_t *Ps = Bs + 2 * RB * RP_STRIDE;              // [2][n_slots][PS_STRIDE]
    int32_t *loff = reinterpret_cast<int32_t *>(Ps + 2 * d.n_slots * PS_STRIDE);   // [RA + 1]
    uint16_t *lsl = reinterpret_cast<uint16_t *>(loff + RA + 1);           // prefix slots of the tile's rows

    // the tile's prefix slot lists -> LDS (rows past the wave's end: empty)
    const int ib0 = d.item_off[row0];
    for (int r = tid; r <= RA; r += 256) loff[r] = d.item_off[min(row0 + r, rend)] - ib0;
    const int nl = d.item_off[rend] - ib0;
    // four loads in flight per thread and round (the last index is read
    // again past the end, not stored)
    for (int i0 = tid; i0 < nl; i0 += 1024) {
        int32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = d.item_slots[ib0 + min(i0 + 256 * q, nl - 1)];
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (i0 + 256 * q < nl) lsl[i0 + 256 * q] = (uint16_t)v[q];
    }

    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(1))) u4v *gptr4;
    // slot loads: load j of thread t is uint4 part (idx & 3) of slot (idx >> 2), idx = 256 j + t.
    // Loads past the last slot read the last slot's last part and are not
    // stored: loads without a branch around them stay in flight together
    // (under a per-lane branch each one waited for the one before).
    const int ns4 = d.n_slots * 4;
    int64_t sbase[kFusedSlotLoads];
#pragma unroll
    for (int j = 0; j < kFusedSlotLoads; j++) {
        const int idx = min(j * 256 + tid, ns4 - 1);
        sbase[j] = (int64_t)d.slot_rows[idx >> 2] * pitch + 2 * (idx & 3);
    }
    // B loads: as k_popmat_mfma_rp (rows past the edge read the last row, zeroed at the store)
    int64_t bbase[TJ];
    uint32_t mb[TJ];
#pragma unroll
    for (int j = 0; j < TJ; j++) {
        const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
        mb[j] = col0 + r < d.n_cols ? ~0u : 0u;
        bbase[j] = (int64_t)min(col0 + r, d.n_cols - 1) * pitch + 2 * part;
    }
    u4v rs[kFusedSlotLoads], rb[TJ];
    auto gload_s = [&](int64_t kw) {
#pragma unroll
        for (int j = 0; j < kFusedSlotLoads; j++) rs[j] = *(gptr4)(bm + sbase[j] + kw);
    };
    auto lstore_s = [&](int buf) {
#pragma unroll
        for (int j = 0; j < kFusedSlotLoads; j++) {
            const int idx = j * 256 + tid;
            if (idx < ns4) *reinterpret_cast<u4v *>(&Ps[(buf * d.n_slots + (idx >> 2)) * PS_STRIDE + 4 * (idx & 3)]) = rs[j];
        }
    };
    auto gload_b = [&](int64_t kw) {
#pragma unroll
        for (int j = 0; j < TJ; j++) rb[j] = *(gptr4)(bm + bbase[j] + kw);
    };
    auto lstore_b = [&](int buf) {
#pragma unroll
        for (int j = 0; j < TJ; j++) {
            const int idx = j * 256 + tid, r = idx >> 2, part = idx & 3;
            *reinterpret_cast<u4v *>(&Bs[(buf * RB + r) * RP_STRIDE + 4 * part]) = rb[j] & mb[j];
        }
    };
    // row masks of the stage in Ps[pbuf] -> As[abuf]; task j of thread t is
    // uint4 part (x & 3) of tile row (x >> 2), x = 256 j + t.  Each task's
    // first kFusedRegSlots prefix slots are held in registers (bytes; a
    // shorter prefix is padded with its first slot, which the AND absorbs),
    // so a stage's loads of one task do not wait on each other; groups of
    // four run up to the largest prefix in the wave.  Longer prefixes read
    // the rest of their list from LDS.
    uint32_t sreg[TI][kFusedRegSlots / 4];
    int kt[TI], kw[TI];
    auto load_lists = [&]() {
#pragma unroll
        for (int j = 0; j < TI; j++) {
            const int r = (j * 256 + tid) >> 2;
            const int b = loff[r], k = loff[r + 1] - b;
            const uint32_t first = k > 0 ? lsl[b] : 0u;
#pragma unroll
            for (int q = 0; q < kFusedRegSlots / 4; q++) {
                uint32_t w = 0;
#pragma unroll
                for (int t = 0; t < 4; t++) w |= (4 * q + t < k ? (uint32_t)lsl[b + 4 * q + t] : first) << (8 * t);
                sreg[j][q] = w;
            }
            kt[j] = k;
            int mx = k;
            for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
            kw[j] = __builtin_amdgcn_readfirstlane(mx);
        }
    };
    auto build = [&](int pbuf, int abuf) {
        const uint32_t *P = Ps + pbuf * d.n_slots * PS_STRIDE;
#pragma unroll
        for (int j = 0; j < TI; j++) {
            const int x = j * 256 + tid, r = x >> 2, part = x & 3;
            const uint32_t *Pp = P + 4 * part;
            auto ld = [&](uint32_t slot) { return *reinterpret_cast<const u4v *>(&Pp[slot * PS_STRIDE]); };
            u4v m = ld(sreg[j][0] & 255u);
#pragma unroll
            for (int q = 0; q < kFusedRegSlots / 4; q++) {
                if (4 * q >= kw[j]) break;   // wave-uniform
                const uint32_t w = sreg[j][q];
                m &= (ld(w & 255u) & ld((w >> 8) & 255u)) & (ld((w >> 16) & 255u) & ld(w >> 24));
            }
            if (kw[j] > kFusedRegSlots) {
                const int b = loff[r];
                for (int i = kFusedRegSlots; i < kt[j]; i++) m &= ld(lsl[b + i]);
            }
            if (kt[j] == 0) m = u4v{0u, 0u, 0u, 0u};   // rows past the wave's end
            *reinterpret_cast<u4v *>(&As[(abuf * RA + r) * RP_STRIDE + 4 * part]) = m;
        }
    };

    typename RpAcc<FMT>::type acc[TI][TJ];   // fp4: the count itself; int8: 64 x the count
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) acc[i][j][q] = 0;

    if (kbeg < kend) {
        gload_s(kbeg);
        lstore_s(0);
        if (kbeg + 8 < kend) {
            gload_s(kbeg + 8);
            lstore_s(1);
        }
        gload_b(kbeg);
        lstore_b(0);
    }
    __syncthreads();   // Ps, Bs of the first stages, loff, lsl
    load_lists();
    if (kbeg < kend) build(0, 0);
    __syncthreads();
    int cur = 0;
    // (the last stages load the last stage again, unused: no branch around
    // the loads, so they are all in flight together)
    const int64_t klast = kend - 8;
    for (int64_t kw = kbeg; kw < kend; kw += 8) {
        const bool n1 = kw + 8 < kend, n2 = kw + 16 < kend;
        gload_s(std::min<int64_t>(kw + 16, klast));
        gload_b(std::min<int64_t>(kw + 8, klast));
        asm volatile("" ::: "memory");   // the loads are issued before this stage's MFMAs
#pragma unroll
        for (int hk = 0; hk < 2; hk++) {
            uint4 pa[TI], pb[TJ];
#pragma unroll
            for (int i = 0; i < TI; i++) {
                const int r = wr * (32 * TI) + i * 32 + l32;
                pa[i] = *reinterpret_cast<const uint4 *>(&As[(cur * RA + r) * RP_STRIDE + 8 * half + 4 * hk]);
            }
#pragma unroll
            for (int j = 0; j < TJ; j++) {
                const int r = wc * (32 * TJ) + j * 32 + l32;
                pb[j] = *reinterpret_cast<const uint4 *>(&Bs[(cur * RB + r) * RP_STRIDE + 8 * half + 4 * hk]);
            }
            if constexpr (FMT == kFmtFp4) {
#pragma unroll
                for (int wi = 0; wi < 4; wi++) {
                    v8i ea[TI], eb[TJ];
#pragma unroll
                    for (int i = 0; i < TI; i++)
                        ea[i] = fp4_expand_a(wi == 0 ? pa[i].x : wi == 1 ? pa[i].y : wi == 2 ? pa[i].z : pa[i].w);
#pragma unroll
                    for (int j = 0; j < TJ; j++)
                        eb[j] = fp4_expand_b(wi == 0 ? pb[j].x : wi == 1 ? pb[j].y : wi == 2 ? pb[j].z : pb[j].w);
#pragma unroll
                    for (int i = 0; i < TI; i++)
#pragma unroll
                        for (int j = 0; j < TJ; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ea[i], eb[j], acc[i][j], 4, 4,
                                                                                        0, 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++) {
                    const int wi = s2 >> 1, par = s2 & 1;
                    v4i ea[TI], eb[TJ];
#pragma unroll
                    for (int i = 0; i < TI; i++)
                        ea[i] = rp_expand_a(wi == 0 ? pa[i].x : wi == 1 ? pa[i].y : wi == 2 ? pa[i].z : pa[i].w, par);
#pragma unroll
                    for (int j = 0; j < TJ; j++)
                        eb[j] = rp_expand_b(wi == 0 ? pb[j].x : wi == 1 ? pb[j].y : wi == 2 ? pb[j].z : pb[j].w, par);
#pragma unroll
                    for (int i = 0; i < TI; i++)
#pragma unroll
                        for (int j = 0; j < TJ; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ea[i], eb[j], acc[i][j], 0, 0, 0);
                }
            }
        }
        // the next stage's masks from its staged slot words (Ps[cur ^ 1],
        // written before the last barrier) into the A block the MFMAs of the
        // previous stage have finished reading
        if (n1) build(cur ^ 1, cur ^ 1);
        asm volatile("" ::: "memory");
        if (n2) lstore_s(cur);        // stage + 2 -> the slot buffer build() read one stage ago
        if (n1) lstore_b(cur ^ 1);    // stage + 1 -> the B buffer the previous stage's MFMAs read
        __syncthreads();
        cur ^= 1;
    }

    // epilogue: the tile's counts in LDS (the staging buffers are free)
    int32_t *T = reinterpret_cast<int32_t *>(lds_u32);   // [RA][TS]
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int r = wr * (32 * TI) + i * 32 + 8 * (q >> 2) + 4 * half + (q & 3);
                const int c = wc * (32 * TJ) + j * 32 + l32;
                if constexpr (FMT == kFmtFp4) T[r * TS + c] = (int32_t)acc[i][j][q];
                else T[r * TS + c] = acc[i][j][q] >> 6;
            }
    __syncthreads();
    const bool split = gridDim.y > 1;
    if (d.seg_off) {
        // listed outputs of this tile: entries of column tile tn, rows row0 .. rend
        const int32_t *seg = d.seg_off + (int64_t)tn * (d.n_rows + 1);
        const int e0 = seg[row0], e1 = seg[rend];
        // four entries per thread and round, their loads in flight together
        for (int b = e0 + tid; b < e1; b += 1024) {
            int2 en[4];
#pragma unroll
            for (int q = 0; q < 4; q++) en[q] = d.entries[min(b + 256 * q, e1 - 1)];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (b + 256 * q >= e1) break;
                const int v = T[((en[q].y >> 8) - row0) * TS + (en[q].y & 255)];
                if (!split) d.out[en[q].x] = v;
                else if (v) atomicAdd(&d.out[en[q].x], v);
            }
        }
    } else if (d.dense_out) {
        for (int i = tid; i < RA * RB; i += 256) {
            const int r = i / RB, c = i % RB;
            if (row0 + r < rend && col0 + c < d.n_cols) {
                const int v = T[r * TS + c];
                int32_t *o = &d.out[(int64_t)(row0 + r) * d.n_cols + col0 + c];
                if (!split) *o = v;
                else if (v) atomicAdd(o, v);
            }
        }
    }
    if (!d.row_compl) return;
    // complement rows: partial counts of the live columns into the full-count
    // accumulator; the tile's complement indices go to LDS first
    int *rcl = T + RA * TS;
    bool mine = false;
    for (int r = tid; r < RA; r += 256) {
        const int j = row0 + r < rend ? d.row_compl[row0 + r] : -1;
        rcl[r] = j;
        mine |= j >= 0;
    }
    if (tid == 0) any_compl = 0;
    __syncthreads();
    if (mine) any_compl = 1;   // every writer stores 1
    __syncthreads();
    const bool any = any_compl != 0;
    if (any)
        for (int i = tid; i < RA * RB; i += 256) {
            const int r = i / RB, c = i % RB, j = rcl[r], e = col0 + c;
            if (j < 0 || e >= d.n_cols) continue;
            const bool live = e == d.compl_x0[j] || !((d.excl[(int64_t)j * d.excl_words + (e >> 6)] >> (e & 63)) & 1ull);
            const int v = T[r * TS + c];
            if (live && v) atomicAdd(&d.acc[(int64_t)j * d.n_cols + e], v);
        }
    if (!any) return;
    __threadfence();
    __syncthreads();
    if (tid == 0) last_arrival = atomicAdd(&d.arrive[tm], 1) == d.tiles_n * (int)gridDim.y - 1;
    __syncthreads();
    if (!last_arrival) return;
    __threadfence();
    // the last workgroup of the row tile: one wave per complement row
    for (int r = row0 + wave; r < rend; r += 4) {
        const int j = d.row_compl[r];
        if (j < 0) continue;
        const int x0 = d.compl_x0[j];
        const int32_t *full = d.acc + (int64_t)j * d.n_cols;
        const uint64_t *ex = d.excl + (int64_t)j * d.excl_words;
        int n = 0;
        for (int w = 0; w < d.excl_words; w++) {
            const int e = w * 64 + lane;
            const bool live = e < d.n_cols && e != x0 && !((ex[w] >> lane) & 1ull);
            if (live && __hip_atomic_load(&full[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= d.min_count) n++;
        }
        for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
        if (lane == 0) {
            d.res[2 * j] = __hip_atomic_load(&full[x0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            d.res[2 * j + 1] = n;
        }
    }
}

size_t wave_mfma_lds_bytes(int ti, int tj, int n_slots, int max_tile_list) {
    const int RA = 64 * ti, RB = 64 * tj;
    const size_t stage = (size_t)4 * (2 * RA * RP_STRIDE + 2 * RB * RP_STRIDE + 2 * n_slots * PS_STRIDE + RA + 1) +
                         (size_t)2 * max_tile_list + 4;
    const size_t tile = (size_t)4 * (RA * (RB + 1) + RA);
    return (std::max(stage, tile) + 15) / 16 * 16;
}

bool plan_wave_mfma(const uint64_t *bm, int64_t pitch_words, const WaveMfmaDesc &d, int ti, int tj, int tiles_m,
                    int splits, size_t lds_bytes, KernelLaunch &L, bool int8) {
    const void *f = nullptr;
    if (int8) {
        if (tj == 1) f = ti == 1 ? (const void *)k_wave_mfma<1, 1, kFmtI8> : ti == 2 ? (const void *)k_wave_mfma<2, 1, kFmtI8>
                                                                                      : (const void *)k_wave_mfma<4, 1, kFmtI8>;
        else f = ti == 1 ? (const void *)k_wave_mfma<1, 2, kFmtI8> : (const void *)k_wave_mfma<2, 2, kFmtI8>;
    } else {
        if (tj == 1) f = ti == 1 ? (const void *)k_wave_mfma<1, 1> : ti == 2 ? (const void *)k_wave_mfma<2, 1>
                                                                              : (const void *)k_wave_mfma<4, 1>;
        else f = ti == 1 ? (const void *)k_wave_mfma<1, 2> : (const void *)k_wave_mfma<2, 2>;
    }
    if ((tj == 2 && ti == 4) || d.n_slots < 1 || d.n_slots > 64 * kFusedSlotLoads || lds_bytes > kFusedMaxLds)
        return false;
    {
        // the kernels may take up to the full 160 KB: set once per device
        static std::mutex mu;
        static uint64_t done = 0;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        std::lock_guard<std::mutex> g(mu);
        if (dev >= 64 || !((done >> dev) & 1)) {
            const void *fn[10] = {(const void *)k_wave_mfma<1, 1>, (const void *)k_wave_mfma<2, 1>,
                                  (const void *)k_wave_mfma<4, 1>, (const void *)k_wave_mfma<1, 2>,
                                  (const void *)k_wave_mfma<2, 2>, (const void *)k_wave_mfma<1, 1, kFmtI8>,
                                  (const void *)k_wave_mfma<2, 1, kFmtI8>, (const void *)k_wave_mfma<4, 1, kFmtI8>,
                                  (const void *)k_wave_mfma<1, 2, kFmtI8>, (const void *)k_wave_mfma<2, 2, kFmtI8>};
            for (const void *fp : fn) {
                // the dynamic limit is the CU's LDS less the kernel's static shared variables
                hipFuncAttributes fa{};
                hipError_t e = hipFuncGetAttributes(&fa, fp);
                if (e == hipSuccess && fa.sharedSizeBytes + kFusedMaxLds > 160 * 1024) e = hipErrorInvalidValue;
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(fp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFusedMaxLds);
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    return false;
                }
            }
            if (dev < 64) done |= 1ull << dev;
        }
    }
    L.func = f;
    L.grid = dim3((unsigned)(tiles_m * d.tiles_n), (unsigned)splits);
    L.block = dim3(256);
    L.shmem = (uint32_t)lds_bytes;
    L.reset();
    L.arg(bm);
    L.arg(pitch_words);
    L.arg(d);
    return true;
}

int rp_rows_per_tile(int64_t b) { return b <= 64 ? 64 : b <= 128 ? 128 : 256; }

int64_t plan_popmat_mfma_rp(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                            int64_t pitch_words, int split_k, KernelLaunch &L, bool fp4, bool wide) {
    wide = wide && fp4;
    if (split_k < 1) split_k = 1;
    int64_t kps = (pitch_words + split_k - 1) / split_k;
    kps = (kps + 15) / 16 * 16;                                  // whole 1,024-bit stages
    kps = std::min<int64_t>(kps, (int64_t)1 << 18);              // <= 2^24 bits per split
    const int64_t zs = (pitch_words + kps - 1) / kps;
    // row tile by the row count: 64, 128 or 256 rows x 128 columns (waves of
    // the engine often have few rows; a 256-row tile would count padding)
    // (a 256 x 128 form of k_popmat_fp4_256 at two workgroups per CU, 512-bit
    // stages, measured 42.5 ms against 41.0 ms on the 65,536 x 512 x 2^22 count)
    // (k_popmat_fp4_256 addresses a tile with 32-bit byte offsets: 256 rows of
    // at most 2^21 words)
    const int RA = rp_rows_per_tile(b),
              RB = wide && RA == 256 && d > 128 && pitch_words < ((int64_t)1 << 21) ? 256 : 128;
    const int64_t tm = (b + RA - 1) / RA, tn = (d + RB - 1) / RB;
    // 256 x 256 tiles: eight waves (two per SIMD) of 4 x 2 fragments with
    // 512-bit stages, 39.6 ms on the 65,536 x 512 x 2^22 count, against 40.2
    // ms for four waves (one per SIMD) of 4 x 4 fragments with 1,024-bit
    // stages (the other wave on a SIMD hides the operand expansion)
    if (RB == 256)
        L.func = (const void *)k_popmat_fp4_256<8, 2, 8>;
    else if (fp4)
        L.func = RA == 64    ? (const void *)k_popmat_mfma_rp<1, 2, kFmtFp4>
                 : RA == 128 ? (const void *)k_popmat_mfma_rp<2, 2, kFmtFp4>
                             : (const void *)k_popmat_mfma_rp<4, 2, kFmtFp4>;
    else
        L.func = RA == 64 ? (const void *)k_popmat_mfma_rp<1, 2>
                          : RA == 128 ? (const void *)k_popmat_mfma_rp<2, 2> : (const void *)k_popmat_mfma_rp<4, 2>;
    L.grid = dim3((unsigned)(tm * tn), (unsigned)zs);
    L.block = dim3(RB == 256 ? 512 : 256);
    L.shmem = 0;
    L.reset();
    L.arg(A);
    L.arg(B);
    L.arg(C);
    L.arg(b);
    L.arg(d);
    L.arg(pitch_words);
    L.arg(kps);
    L.arg(zs > 1 ? 1 : 0);
    L.arg((int)tn);
    return zs;
}

// A[r] = AND of the prefix rows of wave row r, one uint4 per thread.
__global__ __launch_bounds__(256) void k_wave_masks(const uint4 *__restrict__ bm, int64_t pitch4,
                                                    const int32_t *__restrict__ item_ids,
                                                    const int32_t *__restrict__ item_off, uint4 *__restrict__ A) {
    const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (w >= pitch4) return;
    const int ib = item_off[r], ie = item_off[r + 1];
    uint4 m = bm[(int64_t)item_ids[ib] * pitch4 + w];
    // four rows per step, so that their loads are in flight together
    int i = ib + 1;
    for (; i + 3 < ie; i += 4) {
        const uint4 v0 = bm[(int64_t)item_ids[i] * pitch4 + w];
        const uint4 v1 = bm[(int64_t)item_ids[i + 1] * pitch4 + w];
        const uint4 v2 = bm[(int64_t)item_ids[i + 2] * pitch4 + w];
        const uint4 v3 = bm[(int64_t)item_ids[i + 3] * pitch4 + w];
        m.x &= v0.x & v1.x & v2.x & v3.x;
        m.y &= v0.y & v1.y & v2.y & v3.y;
        m.z &= v0.z & v1.z & v2.z & v3.z;
        m.w &= v0.w & v1.w & v2.w & v3.w;
    }
    for (; i < ie; i++) {
        const uint4 v = bm[(int64_t)item_ids[i] * pitch4 + w];
        m.x &= v.x;
        m.y &= v.y;
        m.z &= v.z;
        m.w &= v.w;
    }
    A[(int64_t)r * pitch4 + w] = m;
}

void plan_wave_masks(const uint64_t *bm, int64_t pitch_words, const int32_t *item_ids, const int32_t *item_off,
                     int32_t n_rows, uint64_t *masks, KernelLaunch &L) {
    const int64_t pitch4 = pitch_words / 2;
    L.func = (const void *)k_wave_masks;
    L.grid = dim3((unsigned)((pitch4 + 255) / 256), (unsigned)n_rows);
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(reinterpret_cast<const uint4 *>(bm));
    L.arg(pitch4);
    L.arg(item_ids);
    L.arg(item_off);
    L.arg(reinterpret_cast<uint4 *>(masks));
}

// One 64-lane wave per wave row and slice: an explicit row copies its listed
// columns out of C (slice y of gridDim.y takes every gridDim.y-th group of 64),
// a complement row counts its live columns that reach min_count (slice 0).
__global__ __launch_bounds__(256) void k_wave_gather(WaveDesc d, const int32_t *__restrict__ C, int32_t n_rows,
                                                     int32_t n_cols, const int32_t *__restrict__ row_compl,
                                                     int32_t min_count, int32_t *__restrict__ res) {
    const int lane = threadIdx.x & 63;
    const int r = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (r >= n_rows) return;
    const int32_t *Cr = C + (int64_t)r * n_cols;
    const int j = row_compl[r];
    if (j < 0) {
        for (int q = d.col_off[r] + (int)blockIdx.y * 64 + lane; q < d.col_off[r + 1]; q += 64 * (int)gridDim.y)
            d.out[q] = Cr[d.col_ids[q]];
        return;
    }
    if (blockIdx.y != 0) return;
    const int x0 = d.item_ids[d.item_off[r]];
    const uint64_t *ex = d.excl + (int64_t)j * d.excl_words;
    int n = 0;
    for (int w = 0; w < d.excl_words; w++) {
        const int e = w * 64 + lane;
        const bool live = e < n_cols && e != x0 && !((ex[w] >> lane) & 1ull);
        if (live && Cr[e] >= min_count) n++;
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (lane == 0) {
        res[2 * j] = Cr[x0];
        res[2 * j + 1] = n;
    }
}

void plan_wave_gather(const WaveDesc &d, const int32_t *C, int32_t n_rows, int32_t n_cols, const int32_t *row_compl,
                      int32_t min_count, int32_t *res, int32_t max_row_cols, KernelLaunch &L) {
    L.func = (const void *)k_wave_gather;
    // slices per row: about 256 explicit outputs each, at most 32
    const int slices = std::max(1, std::min(32, (max_row_cols + 255) / 256));
    L.grid = dim3((unsigned)std::max(1, (n_rows + 3) / 4), (unsigned)slices);
    L.block = dim3(256);
    L.shmem = 0;
    L.reset();
    L.arg(d);
    L.arg(C);
    L.arg(n_rows);
    L.arg(n_cols);
    L.arg(row_compl);
    L.arg(min_count);
    L.arg(res);
}

hipError_t launch_popmat_mfma_rp(const uint64_t *A, const uint64_t *B, int32_t *C, int64_t b, int64_t d,
                                 int64_t pitch_words, int split_k, hipStream_t s, bool fp4) {
    KernelLaunch L;
    const int64_t zs = plan_popmat_mfma_rp(A, B, C, b, d, pitch_words, split_k, L, fp4);
    if (zs > 1 && split_k <= 1) launch_zero_i32(C, b * d, s);   // the caller zeroes C only for split_k > 1
    return L.launch(s);
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
