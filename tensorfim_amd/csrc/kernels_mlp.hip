// The search policy's MLP: 14 -> 64 -> 64 -> 2, ReLU, two sigmoid heads,
// trained with binary cross-entropy and Adam on the GPU.
//
// Parameters (fp32, contiguous): W1[64][14] b1[64] W2[64][64] b2[64]
// W3[2][64] b3[2] (5,250 floats).  One 64-lane wave evaluates one sample,
// lane j owning hidden unit j of both layers, so the 64-wide hidden layers map
// one-to-one onto a CDNA wavefront; layer-to-layer products broadcast the
// previous layer through LDS.  Gradients are reduced per parameter by one
// thread looping over samples (deterministic, no atomics).  Inference reads
// a bf16 copy of the weights (round-to-nearest-even) and accumulates in fp32.
#include "tfim_internal.h"

namespace tfim {

constexpr int MLP_IN = 14, MLP_H = 64, MLP_OUT = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_OUT * MLP_H,
              MLP_NPARAM = OFF_B3 + MLP_OUT;
static_assert(MLP_NPARAM == 5250, "parameter count");

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + __expf(-x)); }

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Forward + backward of one sample per wave.  Writes h1, h2, d1 (dL/dz1),
// d2 (dL/dz2) and dy (dL/dy) per sample; loss terms per sample into `lossv`.
// Every wave of a block reaches every barrier; waves past n compute on a
// clamped sample and store nothing.
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float *__restrict__ P, const float *__restrict__ X,
                                                     const float *__restrict__ Y, const float *__restrict__ M,
                                                     int n, float inv_count, float *__restrict__ H1,
                                                     float *__restrict__ H2, float *__restrict__ D1,
                                                     float *__restrict__ D2, float *__restrict__ DY,
                                                     float *__restrict__ lossv) {
    __shared__ float sh[4][MLP_H];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s_raw = blockIdx.x * 4 + w;
    const bool valid = s_raw < n;
    const int s = valid ? s_raw : n - 1;
    const float *x = X + (int64_t)s * MLP_IN;
    float z1 = P[OFF_B1 + lane];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) z1 += P[OFF_W1 + lane * MLP_IN + i] * x[i];
    const float h1 = fmaxf(z1, 0.0f);
    sh[w][lane] = h1;
    __syncthreads();
    float z2 = P[OFF_B2 + lane];
    for (int i = 0; i < MLP_H; i++) z2 += P[OFF_W2 + lane * MLP_H + i] * sh[w][i];
    const float h2 = fmaxf(z2, 0.0f);
    float y[MLP_OUT];
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
        float v = P[OFF_W3 + o * MLP_H + lane] * h2;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        y[o] = v + P[OFF_B3 + o];
    }
    float dy[MLP_OUT], loss = 0.0f;
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
        const float m = M[(int64_t)s * MLP_OUT + o], l = Y[(int64_t)s * MLP_OUT + o];
        const float p = sigmoidf(y[o]);
        dy[o] = m * (p - l) * inv_count;
        // stable BCE with logits: max(y,0) - y*l + log(1 + exp(-|y|))
        loss += m * (fmaxf(y[o], 0.0f) - y[o] * l + __logf(1.0f + __expf(-fabsf(y[o]))));
    }
    float d2 = 0.0f;
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) d2 += P[OFF_W3 + o * MLP_H + lane] * dy[o];
    d2 = z2 > 0.0f ? d2 : 0.0f;
    __syncthreads();
    sh[w][lane] = d2;
    __syncthreads();
    float d1 = 0.0f;
    for (int j = 0; j < MLP_H; j++) d1 += P[OFF_W2 + j * MLP_H + lane] * sh[w][j];
    d1 = z1 > 0.0f ? d1 : 0.0f;
    if (!valid) return;
    H1[(int64_t)s * MLP_H + lane] = h1;
    H2[(int64_t)s * MLP_H + lane] = h2;
    D1[(int64_t)s * MLP_H + lane] = d1;
    D2[(int64_t)s * MLP_H + lane] = d2;
    if (lane < MLP_OUT) DY[(int64_t)s * MLP_OUT + lane] = dy[lane];
    if (lane == 0) lossv[s] = loss * inv_count;
}

// Gradient reduction in two deterministic passes: blockIdx.y = slice z of
// kSlice samples; thread p sums its parameter's terms over the slice in
// sample order into part[z][p]; k_mlp_grad_reduce adds the slices in order.
constexpr int kSlice = 256;
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ H1,
                                                  const float *__restrict__ H2, const float *__restrict__ D1,
                                                  const float *__restrict__ D2, const float *__restrict__ DY,
                                                  int n, float *__restrict__ part) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const int s0 = blockIdx.y * kSlice, s1 = min(n, s0 + kSlice);
    float g = 0.0f;
    if (p < OFF_B1) {
        const int j = (p - OFF_W1) / MLP_IN, i = (p - OFF_W1) % MLP_IN;
        for (int s = s0; s < s1; s++) g += D1[(int64_t)s * MLP_H + j] * X[(int64_t)s * MLP_IN + i];
    } else if (p < OFF_W2) {
        const int j = p - OFF_B1;
        for (int s = s0; s < s1; s++) g += D1[(int64_t)s * MLP_H + j];
    } else if (p < OFF_B2) {
        const int j = (p - OFF_W2) / MLP_H, i = (p - OFF_W2) % MLP_H;
        for (int s = s0; s < s1; s++) g += D2[(int64_t)s * MLP_H + j] * H1[(int64_t)s * MLP_H + i];
    } else if (p < OFF_W3) {
        const int j = p - OFF_B2;
        for (int s = s0; s < s1; s++) g += D2[(int64_t)s * MLP_H + j];
    } else if (p < OFF_B3) {
        const int o = (p - OFF_W3) / MLP_H, j = (p - OFF_W3) % MLP_H;
        for (int s = s0; s < s1; s++) g += DY[(int64_t)s * MLP_OUT + o] * H2[(int64_t)s * MLP_H + j];
    } else {
        const int o = p - OFF_B3;
        for (int s = s0; s < s1; s++) g += DY[(int64_t)s * MLP_OUT + o];
    }
    part[(int64_t)blockIdx.y * MLP_NPARAM + p] = g;
}

__global__ void k_mlp_grad_reduce(const float *__restrict__ part, int slices, float *__restrict__ G) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float g = 0.0f;
    for (int z = 0; z < slices; z++) g += part[(int64_t)z * MLP_NPARAM + p];
    G[p] = g;
}

// Adam (Kingma & Ba 2015) with bias correction; refreshes the bf16 copy.
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ Mo,
                           float *__restrict__ Vo, uint16_t *__restrict__ Pb, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float g = G[p];
    const float m = b1 * Mo[p] + (1.0f - b1) * g;
    const float v = b2 * Vo[p] + (1.0f - b2) * g * g;
    Mo[p] = m;
    Vo[p] = v;
    const float w = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[p] = w;
    const uint32_t u = __float_as_uint(w);
    Pb[p] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__global__ void k_mlp_to_bf16(const float *__restrict__ P, uint16_t *__restrict__ Pb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const uint32_t u = __float_as_uint(P[p]);
    Pb[p] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Inference: one wave per sample; weights from the bf16 copy, or the fp32
// master weights when Pb is null, read straight from memory (21 KB, cached):
// with no weight block in LDS the kernel fits beside the LDS-heavy count
// kernels and runs while they do.  A block handles groups of 4 samples,
// every wave looping the same number of times.
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ P, const uint16_t *__restrict__ Pb,
                                                   const float *__restrict__ X, int n, float *__restrict__ out) {
    __shared__ float sh[4][MLP_H];
    auto W = [&](int p) { return Pb ? bf16_to_f32(Pb[p]) : P[p]; };
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int groups = (n + 3) / 4;
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int s_raw = g * 4 + w;
        const bool valid = s_raw < n;
        const int s = valid ? s_raw : n - 1;
        const float *x = X + (int64_t)s * MLP_IN;
        float z1 = W(OFF_B1 + lane);
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) z1 += W(OFF_W1 + lane * MLP_IN + i) * x[i];
        __syncthreads();
        sh[w][lane] = fmaxf(z1, 0.0f);
        __syncthreads();
        float z2 = W(OFF_B2 + lane);
        for (int i = 0; i < MLP_H; i++) z2 += W(OFF_W2 + lane * MLP_H + i) * sh[w][i];
        const float h2 = fmaxf(z2, 0.0f);
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) {
            float v = W(OFF_W3 + o * MLP_H + lane) * h2;
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0 && valid) out[(int64_t)s * MLP_OUT + o] = v + W(OFF_B3 + o);
        }
    }
}

int mlp_nparam() { return MLP_NPARAM; }

void launch_mlp_fwd_bwd(const float *P, const float *X, const float *Y, const float *M, int n, float inv_count,
                        float *H1, float *H2, float *D1, float *D2, float *DY, float *lossv, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((n + 3) / 4), dim3(256), 0, s, P, X, Y, M, n, inv_count, H1, H2, D1,
                       D2, DY, lossv);
}

int mlp_grad_slices(int n) { return (n + kSlice - 1) / kSlice; }

void launch_mlp_grad(const float *X, const float *H1, const float *H2, const float *D1, const float *D2,
                     const float *DY, int n, float *part, float *G, hipStream_t s) {
    const int slices = mlp_grad_slices(n);
    hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_NPARAM + 255) / 256, slices), dim3(256), 0, s, X, H1, H2, D1, D2, DY,
                       n, part);
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, part, slices, G);
}

void launch_mlp_adam(float *P, const float *G, float *Mo, float *Vo, uint16_t *Pb, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, P, G, Mo, Vo, Pb, lr, b1, b2,
                       eps, c1, c2);
}

void launch_mlp_to_bf16(const float *P, uint16_t *Pb, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, P, Pb);
}

void launch_mlp_infer(const float *P, const uint16_t *Pb, const float *X, int n, float *out, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::min(1024, (n + 3) / 4);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, P, Pb, X, n, out);
}

}  // namespace tfim
