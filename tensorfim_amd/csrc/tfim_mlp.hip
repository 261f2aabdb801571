// Host side of the MLP entry points (include/tfim.h): device buffers, Adam
// state, He-uniform initialisation.
#include <cmath>
#include <cstring>
#include <vector>

#include "tfim_internal.h"

using namespace tfim;

struct tfim_mlp {
    tfim_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;   // own stream: inference never queues behind counting
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *Pb = nullptr;
    int64_t step = 0;
    // batch scratch (grow-only): X, Y, M, H1, H2, D1, D2, DY, loss, out
    float *buf = nullptr;
    int64_t cap = 0;
    // pinned staging (grow-only): inference input/output, training batch;
    // `staged` marks when the last training batch has left its staging
    float *h_x = nullptr, *h_o = nullptr, *h_t = nullptr;
    int64_t h_cap = 0, t_cap = 0;
    hipEvent_t fwd_done = nullptr, staged = nullptr;
    int32_t fwd_n = -1;   // rows of the outstanding forward_submit, -1 = none
};

namespace {

constexpr int IN = 14, H = 64, OUT = 2;

uint64_t sm64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Scratch {
    float *X, *Y, *M, *H1, *H2, *D1, *D2, *DY, *L, *O, *part;
};

int ensure(tfim_mlp *m, int64_t n, Scratch &sc) {
    const int64_t per = IN + OUT + OUT + 4 * H + OUT + 1 + OUT;
    const int64_t part_floats = (int64_t)mlp_grad_slices((int)std::max<int64_t>(n, 4096)) * mlp_nparam();
    if (n > m->cap) {
        hipStreamSynchronize(m->stream);   // queued steps may still use the old block
        hipFree(m->buf);
        m->buf = nullptr;
        m->cap = 0;
        const int64_t c = std::max<int64_t>(n, 4096);
        const int64_t pf = (int64_t)mlp_grad_slices((int)c) * mlp_nparam();
        if (hipMalloc(&m->buf, (size_t)(c * per + pf) * sizeof(float)) != hipSuccess)
            return ctx_fail(m->ctx, TFIM_E_NOMEM, "mlp scratch allocation failed");
        m->cap = c;
    }
    float *p = m->buf;
    const int64_t c = m->cap;
    sc.X = p; p += c * IN;
    sc.Y = p; p += c * OUT;
    sc.M = p; p += c * OUT;
    sc.H1 = p; p += c * H;
    sc.H2 = p; p += c * H;
    sc.D1 = p; p += c * H;
    sc.D2 = p; p += c * H;
    sc.DY = p; p += c * OUT;
    sc.L = p; p += c;
    sc.O = p; p += c * OUT;
    sc.part = p;
    (void)part_floats;
    return TFIM_OK;
}

#define MCHK(m, expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return ctx_fail((m)->ctx, TFIM_E_HIP, hipGetErrorString(e_)); \
    } while (0)

// Stages a training batch in pinned memory and queues its copy to the
// device scratch (so a queued step never reads the caller's buffers; the
// previous batch must have left the staging first).  Returns 1/count of the
// masked outputs in *inv.
int stage_batch(tfim_mlp *m, const float *x, const float *y, const float *mask, int32_t n, Scratch &sc, float *inv) {
    if (int r = ensure(m, n, sc)) return r;
    hipStream_t s = m->stream;
    double cnt = 0;
    for (int64_t i = 0; i < (int64_t)n * OUT; i++) cnt += mask[i];
    *inv = (float)(1.0 / std::max(1.0, cnt));
    const int64_t per = IN + 2 * OUT;
    if (m->t_cap > 0) MCHK(m, hipEventSynchronize(m->staged));
    if (n > m->t_cap) {
        hipHostFree(m->h_t);
        m->h_t = nullptr;
        m->t_cap = 0;
        if (hipHostMalloc(&m->h_t, (size_t)n * per * 4) != hipSuccess)
            return ctx_fail(m->ctx, TFIM_E_NOMEM, "mlp pinned staging allocation failed");
        m->t_cap = n;
    }
    std::memcpy(m->h_t, x, (size_t)n * IN * 4);
    std::memcpy(m->h_t + (size_t)n * IN, y, (size_t)n * OUT * 4);
    std::memcpy(m->h_t + (size_t)n * (IN + OUT), mask, (size_t)n * OUT * 4);
    MCHK(m, hipMemcpyAsync(sc.X, m->h_t, (size_t)n * IN * 4, hipMemcpyHostToDevice, s));
    MCHK(m, hipMemcpyAsync(sc.Y, m->h_t + (size_t)n * IN, (size_t)n * OUT * 4, hipMemcpyHostToDevice, s));
    MCHK(m, hipMemcpyAsync(sc.M, m->h_t + (size_t)n * (IN + OUT), (size_t)n * OUT * 4, hipMemcpyHostToDevice, s));
    MCHK(m, hipEventRecord(m->staged, s));
    return TFIM_OK;
}

// Forward+backward of the staged batch into scratch: leaves G filled; with
// `loss`, waits and returns the batch's summed loss.
int fwd_bwd_staged(tfim_mlp *m, int32_t n, float inv, float *loss, Scratch &sc) {
    hipStream_t s = m->stream;
    launch_mlp_fwd_bwd(m->P, sc.X, sc.Y, sc.M, n, inv, sc.H1, sc.H2, sc.D1, sc.D2, sc.DY, sc.L, s);
    launch_mlp_grad(sc.X, sc.H1, sc.H2, sc.D1, sc.D2, sc.DY, n, sc.part, m->G, s);
    MCHK(m, hipGetLastError());
    if (loss) {
        std::vector<float> lv((size_t)n);
        MCHK(m, hipMemcpyAsync(lv.data(), sc.L, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        MCHK(m, hipStreamSynchronize(s));
        double t = 0;
        for (float v : lv) t += v;
        *loss = (float)t;
    }
    return TFIM_OK;
}

// forward+backward into scratch; returns the loss and leaves G filled
int fwd_bwd(tfim_mlp *m, const float *x, const float *y, const float *mask, int32_t n, float *loss, Scratch &sc) {
    float inv = 0;
    if (int r = stage_batch(m, x, y, mask, n, sc, &inv)) return r;
    return fwd_bwd_staged(m, n, inv, loss, sc);
}

// Room for batches and inference requests of n rows: device scratch and
// pinned staging, allocated once up front.
int reserve(tfim_mlp *m, int64_t n) {
    Scratch sc;
    if (int r = ensure(m, n, sc)) return r;
    if (n > m->t_cap) {
        hipHostFree(m->h_t);
        m->h_t = nullptr;
        m->t_cap = 0;
        if (hipHostMalloc(&m->h_t, (size_t)n * (IN + 2 * OUT) * 4) != hipSuccess)
            return ctx_fail(m->ctx, TFIM_E_NOMEM, "mlp pinned staging allocation failed");
        m->t_cap = n;
    }
    if (n > m->h_cap) {
        hipHostFree(m->h_x);
        hipHostFree(m->h_o);
        m->h_x = m->h_o = nullptr;
        m->h_cap = 0;
        // coherent: the inference kernel reads its rows from and writes its
        // logits to this memory directly (no copies around it)
        if (hipHostMalloc(&m->h_x, (size_t)n * IN * 4, hipHostMallocCoherent) != hipSuccess ||
            hipHostMalloc(&m->h_o, (size_t)n * OUT * 4, hipHostMallocCoherent) != hipSuccess)
            return ctx_fail(m->ctx, TFIM_E_NOMEM, "mlp pinned staging allocation failed");
        m->h_cap = n;
    }
    return TFIM_OK;
}

}  // namespace

extern "C" {

int tfim_mlp_create(tfim_ctx *ctx, uint64_t seed, tfim_mlp **out) {
    if (!ctx || !out) return TFIM_E_INVALID;
    *out = nullptr;
    tfim_mlp *m = new tfim_mlp();
    m->ctx = ctx;
    const int np = mlp_nparam();
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess ||
        hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->fwd_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->staged, hipEventDisableTiming) != hipSuccess || hipMalloc(&m->P, np * 4) != hipSuccess ||
        hipMalloc(&m->G, np * 4) != hipSuccess || hipMalloc(&m->Mo, np * 4) != hipSuccess ||
        hipMalloc(&m->Vo, np * 4) != hipSuccess || hipMalloc(&m->Pb, np * 2) != hipSuccess) {
        tfim_mlp_destroy(m);
        return ctx_fail(ctx, TFIM_E_NOMEM, "mlp allocation failed");
    }
    std::vector<float> p((size_t)np, 0.0f);
    uint64_t s = seed;
    auto init = [&](int off, int rows, int cols) {
        const double a = std::sqrt(6.0 / cols);
        for (int i = 0; i < rows * cols; i++)
            p[off + i] = (float)((2.0 * ((double)(sm64(s) >> 11) / 9007199254740992.0) - 1.0) * a);
    };
    init(0, H, IN);                       // W1
    init(H * IN + H, H, H);               // W2
    init(H * IN + H + H * H + H, OUT, H); // W3
    if (int r = tfim_mlp_set_params(m, p.data())) {
        tfim_mlp_destroy(m);
        return r;
    }
    // the engine's batches and inference requests (up to 8,192 rows) then
    // allocate nothing while it mines
    if (int r = reserve(m, 8192)) {
        tfim_mlp_destroy(m);
        return r;
    }
    *out = m;
    return TFIM_OK;
}

void tfim_mlp_destroy(tfim_mlp *m) {
    if (!m) return;
    if (m->stream) hipStreamSynchronize(m->stream);
    hipFree(m->P); hipFree(m->G); hipFree(m->Mo); hipFree(m->Vo); hipFree(m->Pb); hipFree(m->buf);
    hipHostFree(m->h_x); hipHostFree(m->h_o); hipHostFree(m->h_t);
    if (m->fwd_done) hipEventDestroy(m->fwd_done);
    if (m->staged) hipEventDestroy(m->staged);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
}

int tfim_mlp_get_params(tfim_mlp *m, float *params) {
    if (!m || !params) return TFIM_E_INVALID;
    hipStream_t s = m->stream;
    MCHK(m, hipMemcpyAsync(params, m->P, (size_t)mlp_nparam() * 4, hipMemcpyDeviceToHost, s));
    MCHK(m, hipStreamSynchronize(s));
    return TFIM_OK;
}

int tfim_mlp_set_params(tfim_mlp *m, const float *params) {
    if (!m || !params) return TFIM_E_INVALID;
    hipStream_t s = m->stream;
    const size_t b = (size_t)mlp_nparam() * 4;
    MCHK(m, hipMemcpyAsync(m->P, params, b, hipMemcpyHostToDevice, s));
    MCHK(m, hipMemsetAsync(m->Mo, 0, b, s));
    MCHK(m, hipMemsetAsync(m->Vo, 0, b, s));
    launch_mlp_to_bf16(m->P, m->Pb, s);
    MCHK(m, hipGetLastError());
    MCHK(m, hipStreamSynchronize(s));
    m->step = 0;
    return TFIM_OK;
}

int tfim_mlp_forward_submit(tfim_mlp *m, const float *x, int32_t n, int32_t bf16) {
    if (!m || !x || n < 1) return TFIM_E_INVALID;
    if (m->fwd_n >= 0) return ctx_fail(m->ctx, TFIM_E_STATE, "mlp forward: a request is outstanding; wait first");
    Scratch sc;
    if (int r = ensure(m, n, sc)) return r;
    if (n > m->h_cap) {
        hipHostFree(m->h_x);
        hipHostFree(m->h_o);
        m->h_x = m->h_o = nullptr;
        m->h_cap = 0;
        if (hipHostMalloc(&m->h_x, (size_t)n * IN * 4, hipHostMallocCoherent) != hipSuccess ||
            hipHostMalloc(&m->h_o, (size_t)n * OUT * 4, hipHostMallocCoherent) != hipSuccess)
            return ctx_fail(m->ctx, TFIM_E_NOMEM, "mlp pinned staging allocation failed");
        m->h_cap = n;
    }
    std::memcpy(m->h_x, x, (size_t)n * IN * 4);
    hipStream_t s = m->stream;
    // one launch: the kernel reads the rows from and writes the logits to the
    // coherent pinned staging (a request is tens of rows; two copies around
    // the kernel cost more host time than the kernel)
    (void)sc;
    launch_mlp_infer(m->P, bf16 ? m->Pb : nullptr, m->h_x, n, m->h_o, s);
    MCHK(m, hipGetLastError());
    MCHK(m, hipEventRecord(m->fwd_done, s));
    m->fwd_n = n;
    return TFIM_OK;
}

int tfim_mlp_forward_wait(tfim_mlp *m, float *logits) {
    if (!m || !logits) return TFIM_E_INVALID;
    if (m->fwd_n < 0) return ctx_fail(m->ctx, TFIM_E_STATE, "mlp forward_wait: nothing submitted");
    const int32_t n = m->fwd_n;
    m->fwd_n = -1;
    MCHK(m, hipEventSynchronize(m->fwd_done));
    std::memcpy(logits, m->h_o, (size_t)n * OUT * 4);
    return TFIM_OK;
}

int tfim_mlp_forward(tfim_mlp *m, const float *x, int32_t n, float *logits, int32_t bf16) {
    if (!m || !x || !logits || n < 0) return TFIM_E_INVALID;
    if (n == 0) return TFIM_OK;
    if (int r = tfim_mlp_forward_submit(m, x, n, bf16)) return r;
    return tfim_mlp_forward_wait(m, logits);
}

int tfim_mlp_gradients(tfim_mlp *m, const float *x, const float *y, const float *mask, int32_t n, float *grads,
                       float *loss) {
    if (!m || !x || !y || !mask || !grads || n < 1) return TFIM_E_INVALID;
    Scratch sc;
    if (int r = fwd_bwd(m, x, y, mask, n, loss, sc)) return r;
    hipStream_t s = m->stream;
    MCHK(m, hipMemcpyAsync(grads, m->G, (size_t)mlp_nparam() * 4, hipMemcpyDeviceToHost, s));
    MCHK(m, hipStreamSynchronize(s));
    return TFIM_OK;
}

int tfim_mlp_train(tfim_mlp *m, const float *x, const float *y, const float *mask, int32_t n, float lr,
                   int32_t steps, float *loss) {
    if (!m || !x || !y || !mask || n < 1 || steps < 1) return TFIM_E_INVALID;
    Scratch sc;
    float inv = 0;
    if (int r = stage_batch(m, x, y, mask, n, sc, &inv)) return r;
    hipStream_t s = m->stream;
    for (int32_t k = 0; k < steps; k++) {
        // the last step reports the loss when asked (before its update)
        if (int r = fwd_bwd_staged(m, n, inv, k + 1 == steps ? loss : nullptr, sc)) return r;
        m->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 - std::pow(0.9, (double)m->step));
        const float c2 = (float)(1.0 - std::pow(0.999, (double)m->step));
        launch_mlp_adam(m->P, m->G, m->Mo, m->Vo, m->Pb, lr, b1, b2, 1e-8f, c1, c2, s);
        MCHK(m, hipGetLastError());
    }
    tfim::ctx_model_training(m->ctx, s);
    if (loss) MCHK(m, hipStreamSynchronize(s));   // without a loss request the steps stay queued
    return TFIM_OK;
}

int tfim_mlp_train_step(tfim_mlp *m, const float *x, const float *y, const float *mask, int32_t n, float lr,
                        float *loss) {
    return tfim_mlp_train(m, x, y, mask, n, lr, 1, loss);
}

}  // extern "C"
