// libtfim runtime: the C ABI of include/tfim.h over HIP.
//
// One tfim_ctx per engine holds the device, one default-priority stream, the
// resident item bitmaps and grow-only scratch (device and pinned host), so a
// count call does no allocation once warm.  The reference keeps the same state
// in process globals and thread_locals (kernel_bitmap.cu:334-340,447-449;
// kernel_bmma.cu:345-350,780-797).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "tfim_internal.h"

using namespace tfim;

struct tfim_buf {
    void *ptr = nullptr;
    int64_t bytes = 0;
};

namespace {
// Word-split tuner of a matrix-core wave path: the split is an occupancy
// rule's `base` scaled by 1/2, 1 or 2 (arms 0..2); each arm is tried three
// times, then the cheapest per unit of work runs, another being re-tried
// every 16 waves.
struct SplitTuner {
    double cost[3] = {};
    int64_t samples[3] = {};
    int64_t calls = 0;
    int pick() {
        calls++;
        for (int a = 0; a < 3; a++)
            if (samples[a] < 3) return a;
        int best = 0;
        for (int a = 1; a < 3; a++)
            if (cost[a] < cost[best]) best = a;
        return calls % 16 == 0 ? (best + 1 + (int)((calls / 16) % 2)) % 3 : best;
    }
    static int64_t split_of(int arm, int64_t base, int64_t stages) {
        return arm == 0 ? std::max<int64_t>(1, base / 2) : arm == 2 ? std::min<int64_t>(stages, base * 2) : base;
    }
    void record(int arm, double per) {
        cost[arm] = samples[arm] == 0 ? per : 0.8 * cost[arm] + 0.2 * per;
        samples[arm]++;
    }
};

// Online choice among N arms by measured cost per unit of work: every arm
// once; then up to three samples of the arms within 1.5x of the best cost;
// then the best, another arm being re-tried every 16th call in turn (one
// measured at more than 1.5x the best only every 8th time its turn comes,
// beyond 3x only every 32nd).  Counts the arms issued, since a sample
// arrives only when its wave is waited on.
template <int N>
struct ArmTuner {
    double cost[N] = {};
    int64_t samples[N] = {}, issued[N] = {};
    int64_t calls = 0;
    int pick(const bool *allowed) {
        calls++;
        int a = -1;
        for (int i = 0; i < N && a < 0; i++)
            if (allowed[i] && issued[i] == 0) a = i;
        int best = -1;
        for (int i = 0; i < N; i++)
            if (allowed[i] && samples[i] > 0 && (best < 0 || cost[i] < cost[best])) best = i;
        if (best < 0)
            for (int i = 0; i < N && best < 0; i++)
                if (allowed[i]) best = i;
        for (int i = 0; i < N && a < 0; i++)
            if (allowed[i] && samples[i] > 0 && issued[i] < 3 && cost[i] <= 1.5 * cost[best]) a = i;
        if (a < 0) {
            a = best;
            if (calls % 16 == 0) {
                const int64_t turn = calls / 16;
                const int cand = (best + 1 + (int)(turn % (N - 1))) % N;
                const double ratio = cost[cand] / std::max(1e-30, cost[best]);
                const int64_t period = ratio <= 1.5 ? 1 : ratio <= 3.0 ? 8 : 32;
                if (allowed[cand] && (turn / (N - 1)) % period == 0) a = cand;
            }
        }
        issued[a]++;
        return a;
    }
    void record(int a, double per) {
        cost[a] = samples[a] == 0 ? per : 0.8 * cost[a] + 0.2 * per;
        samples[a]++;
    }
};

// Arms of the mask kernel: (columns per tile, wide staging).
constexpr int kMaskArms = 6;
constexpr int kMaskArmTile[kMaskArms] = {64, 32, 16, 64, 32, 16};
constexpr bool kMaskArmWide[kMaskArms] = {true, true, true, false, false, false};

// Online tuner arms for waves that both batched kernels accept:
// (kernel 1 = L2-resident, 2 = LDS-staged, 5 = fused matrix-core; columns per
// work tile; LDS staging wide = one workgroup per CU with up to 4 x 128 bits
// per lane).
constexpr int kTuneArms = 8;
constexpr int kArmKernel[kTuneArms] = {1, 2, 2, 2, 2, 2, 2, 5};
constexpr int kArmTile[kTuneArms] = {32, 16, 32, 64, 16, 32, 64, 0};
constexpr bool kArmWide[kTuneArms] = {false, false, false, false, true, true, true, false};
// The matrix cores are taken unless the L2 kernel is predicted at less than
// 1/kMfmaBias of their time: the per-unit costs are averages over waves of
// different shapes, and the L2 kernel's time varies more with the shape
// (clickstream, support 0.1 %: 25 ms of count kernels with the plain
// comparison, 16 ms with every wave on the matrix cores).
constexpr double kMfmaBias = 2.0;
// A fused matrix-core wave that does not fit the kernel's limits (the caller
// takes another path); not an error.
constexpr int kFusedNoFit = 1;
}  // namespace

struct tfim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 256;
    char name[256] = {0};
    int64_t hbm = 0;
    int clock_mhz = 0;
    int wall_khz = 0;   // rate of the device wall clock the mask kernel stamps its span with
    std::string err;
    std::mutex err_mu;
    // wave state (slots, tuner, counters): one thread may submit a wave while
    // another waits on the earlier one
    std::mutex wave_mu;

    // resident bitmaps
    uint64_t *d_bm = nullptr;
    bool owns_bm = true;   // false for contexts opened with tfim_open_shared
    int64_t pitch = 0;     // u64 words per row (multiple of 32)
    int64_t n_bits = 0;
    int32_t n_items = 0, n_cols = 0;

    // per-node scratch
    int32_t *d_node_items = nullptr, *d_node_out = nullptr;
    int32_t *h_node_items = nullptr, *h_node_out = nullptr;   // pinned
    int64_t node_items_cap = 0, node_out_cap = 0;

    // wave scratch: kWaveSlots in-flight slots, each a descriptor block and an
    // output block (device + pinned host) and its events
    struct Slot {
        char *d_desc = nullptr, *h_desc = nullptr;
        int64_t desc_cap = 0;
        int32_t *d_out = nullptr, *h_out = nullptr;
        int64_t out_cap = 0;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around the wave's count kernels
        hipEvent_t done = nullptr;                 // after its drain
        // a mask wave's span, stamped by its kernel at output int stamp_at
        // (-1: timed by ev0/ev1)
        int64_t stamp_at = -1;
        bool busy = false;
        int64_t ticket = -1, n_out = 0, n_rows = 0;
        int tuned = -1;
        double work = 0;
        int mf_arm = -1;       // second tuner's arm for this wave (-1: not tuned)
        double mf_work = 0;
        int sk_arm = -1;       // third tuner's arm (matrix-core waves)
        bool sk_fused = false; // ... of the fused path's split tuner
        int mask_arm = -1;     // mask kernel tuner's arm (-1: not tuned)
        double mask_work = 0;
        double sk_work = 0;
        // the model was training when the wave was issued, or a training
        // round was queued before it ended: its time is kept out of the tuners
        bool contended = false;
        int64_t train_seq = 0;
        // one executable graph per kernel sequence: stage -> [fill] -> ev0 ->
        // kernels -> ev1 -> drain
        uint64_t *d_mat = nullptr;   // matrix-core path: prefix masks, then the count matrix
        int64_t mat_cap = 0;
        struct Graph {
            static constexpr int kMaxKernels = 3;
            const void *funcs[kMaxKernels] = {};
            int nk = 0;
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            bool has_fill = false;
            hipGraphNode_t stage = nullptr, fill = nullptr, e0 = nullptr, e1 = nullptr, drain = nullptr;
            hipGraphNode_t kern[kMaxKernels] = {};
        };
        std::vector<Graph> graphs;
    };
    static constexpr int kWaveSlots = 4;
    Slot slots[kWaveSlots];
    // Training rounds of a learned policy's network run on its own stream and
    // share the device with the count kernels; `train_ev` follows the last one
    // queued (guarded by wave_mu).
    hipEvent_t train_ev = nullptr;
    bool train_ev_set = false;
    int64_t train_seq = 0;
    int64_t next_ticket = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    int64_t wave_launches = 0, wave_graph_replays = 0, wave_graphs = 0, wave_rows = 0, wave_lds_launches = 0;
    int64_t wave_contended = 0;   // waves kept out of the tuners (model training alongside)
    double wave_ms = 0;
    int wave_kernel = 0;    // 0 auto, 1 L2 kernel, 2 LDS kernel, 3 LDS kernel with wide staging, 4 matrix cores
    bool mfma_int8 = false; // wave matrix-core paths on the int8 matrix cores (default: fp4)
    // second tuner, for waves outside the LDS kernel: arm 0 = L2 kernel, arm 1
    // = matrix cores in three launches, arm 2 = fused matrix-core launch; EMA
    // of kernel milliseconds per unit of each arm's work
    double mf_cost[3] = {64.0 / 300e12 * 1e3, 64.0 / 3000e12 * 1e3, 64.0 / 1500e12 * 1e3};   // ms per u64 word (priors)
    int64_t mf_samples[3] = {};
    int64_t mf_calls = 0, wave_mfma_launches = 0, wave_fused_launches = 0;
    // third tuner: each matrix-core path's word split, as 1/2, 1 or 2 times
    // the occupancy rule's split; EMA of milliseconds per unit of work
    SplitTuner sk, fk;
    ArmTuner<kMaskArms> mask_tune;   // mask rows: columns per tile and staging
    std::vector<MaskTile> mask_tiles;
    int64_t wave_mask_launches = 0;
    // wave tuner: EMA of kernel milliseconds per (row prefix + column) word,
    // per arm (kernel, columns per tile)
    double tune_cost[kTuneArms] = {};
    int64_t tune_samples[kTuneArms] = {}, tune_issued[kTuneArms] = {};
    int64_t tune_calls = 0;
    int tile_cols = 0;      // 0: chosen by the tuner (32 when it does not run)
    int mfma_tile = 0;      // 0 auto (= 256), 128 or 256
    int mfma_split = 0;     // matrix-core waves: 0 split by the occupancy rule and tuner, 1 no split
    int node_shape = 0;     // per-node kernel: 0 auto, else R * 100 + columns per block
    int node_spread = 0;    // per-node counts: 0 auto, 1 one copy, 8 spread over 8 copies
    // waves replayed as instantiated graphs (off by default: on MI355X with
    // ROCm 7.2 a wave's three to five kernels launched one by one cost about
    // 12 us of host time against 22 us for a graph replay, and run back to
    // back on the device where the graph leaves ~11 us between nodes)
    int use_graphs = 0;
    std::vector<int32_t> slot_of;   // item -> LDS slot of the current wave (-1 = none)
    std::vector<int32_t> fslot_of;  // item -> slot of the fused matrix-core wave being packed
    std::vector<int32_t> tile_bucket;   // counting sort of the LDS kernel's tiles by cost
    const char *fused_nofit = "";   // the limit the last wave refused by the fused kernel exceeded
    std::map<std::string, std::pair<int, int32_t>> popmat_tuned;   // shape -> (tile, split_k)

    int32_t *d_scratch = nullptr;   // this device's `scratch` (tfim_count_node_noalloc)

    // device preprocessing state between tfim_prep_begin and tfim_prep_finish
    int32_t *p_ids = nullptr;
    int64_t *p_off = nullptr;
    int64_t p_n_ids = 0, p_n_rows = -1;
    int32_t p_id_max = -1;
};

namespace {

int fail(tfim_ctx *c, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) {
        std::lock_guard<std::mutex> g(c->err_mu);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, TFIM_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

// grow-only device + pinned host pair
// (host_flags: of the pinned part; zero: the device part starts zeroed,
// cleared on the ctx stream ahead of the work that follows)
template <typename T>
int ensure_pair(tfim_ctx *c, T **d, T **h, int64_t *cap, int64_t need, unsigned host_flags = hipHostMallocDefault,
                bool zero = false) {
    if (need <= *cap) return TFIM_OK;
    int64_t n = std::max<int64_t>(need, std::max<int64_t>(*cap * 2, 1024));
    if (*d) hipFree(*d);
    if (h && *h) hipHostFree(*h);
    *d = nullptr;
    if (h) *h = nullptr;
    *cap = 0;
    if (hipMalloc((void **)d, (size_t)n * sizeof(T)) != hipSuccess)
        return fail(c, TFIM_E_NOMEM, "device scratch allocation of %lld bytes failed", (long long)(n * sizeof(T)));
    if (zero && hipMemsetAsync(*d, 0, (size_t)n * sizeof(T), c->stream) != hipSuccess)
        return fail(c, TFIM_E_HIP, "clearing %lld bytes of device scratch failed", (long long)(n * sizeof(T)));
    if (h && hipHostMalloc((void **)h, (size_t)n * sizeof(T), host_flags) != hipSuccess)
        return fail(c, TFIM_E_NOMEM, "pinned scratch allocation of %lld bytes failed", (long long)(n * sizeof(T)));
    *cap = n;
    return TFIM_OK;
}

int set_device(tfim_ctx *c) {
    HIPCHK(c, hipSetDevice(c->device));
    return TFIM_OK;
}

// A wave slot's descriptor block (read by the wave's stage kernel straight
// from pinned host memory, so that memory is mapped coherent, and whole
// 16-byte units are read) and its output block (the device part is zero
// whenever the slot is idle: cleared when allocated, and by each wave's
// drain kernel after it copies the counts to the coherent pinned part).
template <typename Slot>
int ensure_desc(tfim_ctx *c, Slot &sl, int64_t bytes) {
    return ensure_pair(c, &sl.d_desc, &sl.h_desc, &sl.desc_cap, (bytes + 15) / 16 * 16, hipHostMallocCoherent);
}
template <typename Slot>
int ensure_out(tfim_ctx *c, Slot &sl, int64_t n) {
    return ensure_pair(c, &sl.d_out, &sl.h_out, &sl.out_cap, std::max<int64_t>(1, n), hipHostMallocCoherent, true);
}

}  // namespace

// Scratch of tfim_count_node_noalloc: the 8 spread copies of one node's
// counts (partial results that its kernel's atomics are spread over), for at
// most TFIM_NOALLOC_MAX_COLS survivors: 256 KiB per device, held until the
// process ends.  Host code must not pass `scratch` itself to a kernel (that
// name is a host-side stand-in): it uses the device address
// hipGetSymbolAddress returns, looked up once per device in tfim_open.
static __device__ __align__(256) int32_t scratch[8 * TFIM_NOALLOC_MAX_COLS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the device address of `scratch` on it (once per
// device; later calls return the kept address).
int scratch_setup(int device, int32_t **ptr) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return TFIM_E_RANGE;
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TFIM_E_HIP;
        g_scratch_ptr[device] = static_cast<int32_t *>(p);
    }
    *ptr = g_scratch_ptr[device];
    return TFIM_OK;
}
}  // namespace

namespace tfim {
hipStream_t ctx_stream(tfim_ctx *c) { return c->stream; }
int ctx_device(tfim_ctx *c) { return c->device; }
int ctx_fail(tfim_ctx *c, int code, const char *msg) { return fail(c, code, "%s", msg); }
}  // namespace tfim

extern "C" {

int tfim_open(int device, tfim_ctx **out) {
    if (!out) return TFIM_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TFIM_E_HIP;
    if (device < 0 || device >= n) return TFIM_E_INVALID;
    tfim_ctx *c = new tfim_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return TFIM_E_HIP; }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) { delete c; return TFIM_E_HIP; }
    c->n_cu = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
    if (p.name[0]) snprintf(c->name, sizeof c->name, "%s (%s)", p.name, p.gcnArchName);
    else snprintf(c->name, sizeof c->name, "%s, %d CUs", p.gcnArchName, c->n_cu);   // the runtime gave no marketing name
    c->hbm = (int64_t)p.totalGlobalMem;
    c->clock_mhz = p.clockRate / 1000;
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || c->wall_khz <= 0) {
        delete c;
        return TFIM_E_HIP;
    }
    // device address of the no-allocation entry point's scratch (once per device)
    if (device < kMaxDevices && scratch_setup(device, &c->d_scratch) != TFIM_OK) { delete c; return TFIM_E_HIP; }
    // from here on a failure releases what was created (tfim_close skips null handles)
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { tfim_close(c); return TFIM_E_HIP; }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) { tfim_close(c); return TFIM_E_HIP; }
    for (auto &sl : c->slots)
        if (hipEventCreate(&sl.ev0) != hipSuccess || hipEventCreate(&sl.ev1) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) { tfim_close(c); return TFIM_E_HIP; }
    *out = c;
    return TFIM_OK;
}

extern "C++" {
namespace tfim {
void ctx_model_training(tfim_ctx *c, hipStream_t s) {
    std::lock_guard<std::mutex> g(c->wave_mu);
    if (!c->train_ev && hipEventCreateWithFlags(&c->train_ev, hipEventDisableTiming) != hipSuccess) c->train_ev = nullptr;
    c->train_ev_set = c->train_ev && hipEventRecord(c->train_ev, s) == hipSuccess;
    c->train_seq++;
}
}  // namespace tfim
}

void tfim_close(tfim_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->owns_bm) hipFree(c->d_bm);
    hipFree(c->d_node_items); hipFree(c->d_node_out);
    hipHostFree(c->h_node_items); hipHostFree(c->h_node_out);
    for (auto &sl : c->slots) {
        hipFree(sl.d_desc); hipHostFree(sl.h_desc);
        hipFree(sl.d_out); hipHostFree(sl.h_out);
        hipFree(sl.d_mat);
        if (sl.ev0) hipEventDestroy(sl.ev0);
        if (sl.ev1) hipEventDestroy(sl.ev1);
        if (sl.done) hipEventDestroy(sl.done);
        for (auto &g : sl.graphs) {
            if (g.x) hipGraphExecDestroy(g.x);
            if (g.g) hipGraphDestroy(g.g);
        }
    }
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->train_ev) hipEventDestroy(c->train_ev);
    if (c->ev1) hipEventDestroy(c->ev1);
    hipFree(c->p_ids); hipFree(c->p_off);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int tfim_open_shared(tfim_ctx *parent, tfim_ctx **out) {
    if (!parent || !out) return TFIM_E_INVALID;
    if (!parent->d_bm) return fail(parent, TFIM_E_STATE, "open_shared: parent has no bitmaps loaded");
    tfim_ctx *c = nullptr;
    if (int r = tfim_open(parent->device, &c)) return r;
    c->d_bm = parent->d_bm;
    c->owns_bm = false;
    c->pitch = parent->pitch;
    c->n_bits = parent->n_bits;
    c->n_items = parent->n_items;
    c->n_cols = parent->n_cols;
    *out = c;
    return TFIM_OK;
}

const char *tfim_last_error(const tfim_ctx *c) { return c ? c->err.c_str() : "null context"; }

int tfim_device_info(tfim_ctx *c, char *name, int name_len, int32_t *cus, int64_t *hbm, int32_t *clk) {
    if (!c) return TFIM_E_INVALID;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->name);
    if (cus) *cus = c->n_cu;
    if (hbm) *hbm = c->hbm;
    if (clk) *clk = c->clock_mhz;
    return TFIM_OK;
}

int64_t tfim_pitch_words(int64_t n_bits) { return pitch_for_bits(n_bits); }

static int load_common(tfim_ctx *c, int32_t n_items, int64_t n_bits, int32_t n_cols) {
    if (n_items < 1 || n_bits < 1 || n_cols < 0 || n_cols > n_items)
        return fail(c, TFIM_E_INVALID, "load_bitmaps: n_items=%d n_bits=%lld n_cols=%d", n_items,
                    (long long)n_bits, n_cols);
    if (n_bits > (int64_t)INT32_MAX)
        return fail(c, TFIM_E_RANGE, "load_bitmaps: %lld transactions exceed int32 supports", (long long)n_bits);
    if (!c->owns_bm) return fail(c, TFIM_E_STATE, "load_bitmaps: context shares its parent's bitmaps");
    if (int r = set_device(c)) return r;
    hipFree(c->d_bm);
    c->d_bm = nullptr;
    c->pitch = pitch_for_bits(n_bits);
    const size_t bytes = (size_t)n_items * (size_t)c->pitch * 8;
    if (hipMalloc(&c->d_bm, bytes) != hipSuccess)
        return fail(c, TFIM_E_NOMEM, "bitmap allocation of %zu bytes failed", bytes);
    c->n_bits = n_bits;
    c->n_items = n_items;
    c->n_cols = n_cols;
    return TFIM_OK;
}

int tfim_load_bitmaps(tfim_ctx *c, const uint64_t *rows, int32_t n_items, int64_t n_bits,
                      int64_t words_per_row, int32_t n_cols) {
    if (!c || !rows) return fail(c, TFIM_E_INVALID, "load_bitmaps: null argument");
    if (words_per_row < (n_bits + 63) / 64)
        return fail(c, TFIM_E_INVALID, "load_bitmaps: words_per_row %lld < ceil(n_bits/64)", (long long)words_per_row);
    if (int r = load_common(c, n_items, n_bits, n_cols)) return r;
    const int64_t used = (n_bits + 63) / 64;
    HIPCHK(c, hipMemsetAsync(c->d_bm, 0, (size_t)n_items * c->pitch * 8, c->stream));
    HIPCHK(c, hipMemcpy2DAsync(c->d_bm, (size_t)c->pitch * 8, rows, (size_t)words_per_row * 8,
                               (size_t)used * 8, (size_t)n_items, hipMemcpyHostToDevice, c->stream));
    // bits beyond n_bits in the last used word must be zero for exact counts
    if (n_bits % 64) {
        std::vector<uint64_t> tail((size_t)n_items);
        const uint64_t keep = (1ull << (n_bits % 64)) - 1;
        for (int32_t i = 0; i < n_items; i++) tail[i] = rows[(size_t)i * words_per_row + used - 1] & keep;
        HIPCHK(c, hipMemcpy2DAsync(c->d_bm + used - 1, (size_t)c->pitch * 8, tail.data(), 8, 8,
                                   (size_t)n_items, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TFIM_OK;
}

int tfim_load_bitmaps_replicated(tfim_ctx *c, const uint64_t *rows, int32_t n_items, int64_t n_bits,
                                 int64_t words_per_row, int32_t n_cols, int32_t replicate) {
    if (!c || !rows || replicate < 1) return fail(c, TFIM_E_INVALID, "load_bitmaps_replicated: bad argument");
    if (words_per_row < (n_bits + 63) / 64)
        return fail(c, TFIM_E_INVALID, "load_bitmaps_replicated: words_per_row too small");
    const int64_t total_bits = n_bits * (int64_t)replicate;
    if (int r = load_common(c, n_items, total_bits, n_cols)) return r;
    // stage the base rows (zero beyond n_bits) on the device, then expand
    const int64_t bp = pitch_for_bits(n_bits);
    std::vector<uint64_t> base((size_t)n_items * bp, 0ull);
    const int64_t used = (n_bits + 63) / 64;
    for (int32_t i = 0; i < n_items; i++) {
        memcpy(&base[(size_t)i * bp], rows + (size_t)i * words_per_row, (size_t)used * 8);
        if (n_bits % 64) base[(size_t)i * bp + used - 1] &= (1ull << (n_bits % 64)) - 1;
    }
    uint64_t *d_base = nullptr;
    if (hipMalloc(&d_base, base.size() * 8) != hipSuccess) return fail(c, TFIM_E_NOMEM, "base staging allocation failed");
    hipError_t e = hipMemcpyAsync(d_base, base.data(), base.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_replicate_rows(d_base, bp, n_bits, c->d_bm, c->pitch, n_items, replicate, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d_base);
    if (e != hipSuccess) return fail(c, TFIM_E_HIP, "replicate: %s", hipGetErrorString(e));
    return TFIM_OK;
}

int tfim_item_supports(tfim_ctx *c, int32_t *out) {
    if (!c || !out) return fail(c, TFIM_E_INVALID, "item_supports: null argument");
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "item_supports: no bitmaps loaded");
    if (int r = set_device(c)) return r;
    if (int r = ensure_pair(c, &c->d_node_out, &c->h_node_out, &c->node_out_cap, c->n_items)) return r;
    launch_row_popcount(c->d_bm, c->pitch, c->n_items, c->d_node_out, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_node_out, c->d_node_out, (size_t)c->n_items * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    memcpy(out, c->h_node_out, (size_t)c->n_items * 4);
    return TFIM_OK;
}

namespace {
// Checks a node's items and stages them on the device for k_count_node.
// Spread stride of the per-node counts (0 = one copy): 8 copies 4 KB apart
// for nodes of at most 512 survivors (the copy back and the host sum stay
// small), unless the "node_spread" knob says otherwise.
int64_t node_spread_stride(const tfim_ctx *c) {
    const bool on = c->node_spread == 8 || (c->node_spread == 0 && c->n_cols <= 512);
    return on ? ((int64_t)c->n_cols + 1023) / 1024 * 1024 : 0;
}

int node_stage(tfim_ctx *c, const char *what, const int32_t *items, int32_t k) {
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "%s: no bitmaps loaded", what);
    for (int32_t i = 0; i < k; i++)
        if (items[i] < 0 || items[i] >= c->n_items)
            return fail(c, TFIM_E_INVALID, "%s: item %d out of range [0,%d)", what, items[i], c->n_items);
    if (int r = set_device(c)) return r;
    if (int r = ensure_pair(c, &c->d_node_items, &c->h_node_items, &c->node_items_cap, k)) return r;
    if (int r = ensure_pair(c, &c->d_node_out, &c->h_node_out, &c->node_out_cap, 8 * node_spread_stride(c) +
                                                                                    std::max(1, c->n_cols)))
        return r;
    memcpy(c->h_node_items, items, (size_t)k * 4);
    HIPCHK(c, hipMemcpyAsync(c->d_node_items, c->h_node_items, (size_t)k * 4, hipMemcpyHostToDevice, c->stream));
    return TFIM_OK;
}
}  // namespace

int tfim_count_node(tfim_ctx *c, const int32_t *items, int32_t k, int32_t *out) {
    if (!c || !items || !out || k < 1) return fail(c, TFIM_E_INVALID, "count_node: null argument or k < 1");
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "count_node: no bitmaps loaded");
    if (c->n_cols == 0) return TFIM_OK;
    if (int r = node_stage(c, "count_node", items, k)) return r;
    const int64_t ss = node_spread_stride(c);
    launch_zero_i32(c->d_node_out, ss ? 8 * ss : c->n_cols, c->stream);
    launch_count_node(c->d_bm, c->pitch, c->d_node_items, k, c->n_cols, c->d_node_out, ss, c->n_cu, c->node_shape,
                      c->stream);
    HIPCHK(c, hipGetLastError());
    if (ss) {   // the 8 copies, packed, then summed here
        HIPCHK(c, hipMemcpy2DAsync(c->h_node_out, (size_t)c->n_cols * 4, c->d_node_out, (size_t)ss * 4,
                                   (size_t)c->n_cols * 4, 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(c->h_node_out, c->d_node_out, (size_t)c->n_cols * 4, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ss) {
        for (int32_t e = 0; e < c->n_cols; e++) {
            int32_t t = 0;
            for (int j = 0; j < 8; j++) t += c->h_node_out[(size_t)j * c->n_cols + e];
            out[e] = t;
        }
    } else {
        memcpy(out, c->h_node_out, (size_t)c->n_cols * 4);
    }
    return TFIM_OK;
}

int tfim_count_node_noalloc(tfim_ctx *c, const int32_t *items, int32_t k, int32_t *out) {
    if (!c || !items || !out) return fail(c, TFIM_E_INVALID, "count_node_noalloc: null argument");
    if (k < 1 || k > TFIM_NOALLOC_MAX_ITEMS)
        return fail(c, TFIM_E_RANGE, "count_node_noalloc: %d prefix items (1..%d accepted)", k, TFIM_NOALLOC_MAX_ITEMS);
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "count_node_noalloc: no bitmaps loaded");
    if (c->n_cols > TFIM_NOALLOC_MAX_COLS)
        return fail(c, TFIM_E_RANGE, "count_node_noalloc: %d survivors exceed the scratch limit of %d", c->n_cols,
                    TFIM_NOALLOC_MAX_COLS);
    if (!c->d_scratch) return fail(c, TFIM_E_STATE, "count_node_noalloc: no scratch on device %d", c->device);
    NodeItems it{};
    it.n = k;
    for (int32_t i = 0; i < k; i++) {
        if (items[i] < 0 || items[i] >= c->n_items)
            return fail(c, TFIM_E_INVALID, "count_node_noalloc: item %d out of range [0,%d)", items[i], c->n_items);
        it.id[i] = items[i];
    }
    if (c->n_cols == 0) return TFIM_OK;
    if (int r = set_device(c)) return r;
    // the 8 copies, 1024-int aligned; cleared by a fill kernel on every call
    // (the array keeps the previous call's partial counts)
    const int64_t stride = ((int64_t)c->n_cols + 1023) / 1024 * 1024;
    launch_zero_i32(c->d_scratch, 8 * stride, c->stream);
    launch_count_node_args(c->d_bm, c->pitch, it, c->n_cols, c->d_scratch, stride, c->n_cu, c->node_shape, c->stream);
    launch_sum_copies8(c->d_scratch, stride, c->n_cols, out, c->stream);
    HIPCHK(c, hipGetLastError());
    return TFIM_OK;
}

int tfim_time_count_node(tfim_ctx *c, const int32_t *items, int32_t k, int32_t warmup, int32_t reps,
                         double *avg_ms) {
    if (!c || !items || k < 1 || reps < 1 || warmup < 0 || !avg_ms)
        return fail(c, TFIM_E_INVALID, "time_count_node: null argument, k < 1 or reps < 1");
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "time_count_node: no bitmaps loaded");
    if (c->n_cols == 0) return fail(c, TFIM_E_STATE, "time_count_node: no survivors to count");
    if (int r = node_stage(c, "time_count_node", items, k)) return r;
    // the launches accumulate into the library's node buffer, whose contents
    // are not returned
    for (int i = 0; i < warmup; i++)
        launch_count_node(c->d_bm, c->pitch, c->d_node_items, k, c->n_cols, c->d_node_out, node_spread_stride(c), c->n_cu,
                          c->node_shape, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; i++)
        launch_count_node(c->d_bm, c->pitch, c->d_node_items, k, c->n_cols, c->d_node_out, node_spread_stride(c), c->n_cu,
                          c->node_shape, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = (double)ms / reps;
    return TFIM_OK;
}

namespace {

hipKernelNodeParams node_params(KernelLaunch &L) {
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(L.func);
    p.gridDim = L.grid;
    p.blockDim = L.block;
    p.sharedMemBytes = L.shmem;
    p.kernelParams = L.params();
    p.extra = nullptr;
    return p;
}

// Issues one wave on a slot, every step a kernel on the ctx stream (no copy
// calls, no second stream, no cross-stream waits, which cost more host time
// per wave than the kernels' own launches):
//   stage: the descriptor block from pinned host memory to the device;
//   fill (optional): zero a buffer the wave accumulates into other than the
//         output block (the matrix-core path's count matrix);
//   the wave's kernels (1 to 3: count kernel, or count kernel and complement
//         reduction, or masks, matrix count and gather), between the slot's
//         timing events (so the tuners see the count kernels alone);
//   drain: the first n_out counts to pinned host memory, and the output
//         block's [0, max(n_out, n_zero)) back to zero;
// then the slot's `done` event.  With graphs on, stage to drain form one
// graph per (slot, fill or not, kernel sequence), instantiated on first use
// and replayed with this wave's node parameters.  The slot is idle (its last
// wave waited on, so its buffers are free and its device output block is
// zero) whenever it is issued again.
int issue_wave(tfim_ctx *c, tfim_ctx::Slot &sl, int64_t desc_bytes, int32_t *fill, int64_t n_fill, int64_t n_out,
               int64_t n_zero, KernelLaunch *const *ks, int nk, bool events = true) {
    for (int i = 0; i < nk; i++)
        if (ks[i]->overflow || !ks[i]->func) return fail(c, TFIM_E_INVALID, "wave kernel %d: bad launch description", i);
    // (the callers hold wave_mu)
    sl.train_seq = c->train_seq;
    sl.contended = c->train_ev_set && hipEventQuery(c->train_ev) == hipErrorNotReady;
    KernelLaunch S, Z, D;
    plan_stage16(sl.h_desc, sl.d_desc, desc_bytes, S);
    const bool has_fill = fill && n_fill > 0;
    if (has_fill) plan_zero_i32(fill, n_fill, Z);
    plan_drain_i32(sl.d_out, sl.h_out, n_out, n_zero, D);
    if (!c->use_graphs) {
        HIPCHK(c, S.launch(c->stream));
        if (has_fill) HIPCHK(c, Z.launch(c->stream));
        // (a kernel that stamps its own span needs no events around it:
        // each costs about 10 us of the gap between two waves' kernels)
        if (events) HIPCHK(c, hipEventRecord(sl.ev0, c->stream));
        for (int i = 0; i < nk; i++) HIPCHK(c, ks[i]->launch(c->stream));
        if (events) HIPCHK(c, hipEventRecord(sl.ev1, c->stream));
        HIPCHK(c, D.launch(c->stream));
    } else {
        hipKernelNodeParams sp = node_params(S), zp{}, dp = node_params(D);
        if (has_fill) zp = node_params(Z);
        hipKernelNodeParams kp[tfim_ctx::Slot::Graph::kMaxKernels];
        for (int i = 0; i < nk; i++) kp[i] = node_params(*ks[i]);
        tfim_ctx::Slot::Graph *G = nullptr;
        for (auto &g : sl.graphs) {
            bool same = g.nk == nk && g.has_fill == has_fill;
            for (int i = 0; i < nk && same; i++) same = g.funcs[i] == ks[i]->func;
            if (same) G = &g;
        }
        if (!G) {
            // build and instantiate; only a complete graph is kept
            tfim_ctx::Slot::Graph g;
            g.nk = nk;
            g.has_fill = has_fill;
            for (int i = 0; i < nk; i++) g.funcs[i] = ks[i]->func;
            hipError_t e = hipGraphCreate(&g.g, 0);
            if (e == hipSuccess) e = hipGraphAddKernelNode(&g.stage, g.g, nullptr, 0, &sp);
            hipGraphNode_t prev = g.stage;
            if (e == hipSuccess && has_fill) {
                e = hipGraphAddKernelNode(&g.fill, g.g, &prev, 1, &zp);
                prev = g.fill;
            }
            if (e == hipSuccess) e = hipGraphAddEventRecordNode(&g.e0, g.g, &prev, 1, sl.ev0);
            prev = g.e0;
            for (int i = 0; i < nk && e == hipSuccess; i++) {
                e = hipGraphAddKernelNode(&g.kern[i], g.g, &prev, 1, &kp[i]);
                prev = g.kern[i];
            }
            if (e == hipSuccess) e = hipGraphAddEventRecordNode(&g.e1, g.g, &prev, 1, sl.ev1);
            if (e == hipSuccess) e = hipGraphAddKernelNode(&g.drain, g.g, &g.e1, 1, &dp);
            if (e == hipSuccess) e = hipGraphInstantiate(&g.x, g.g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                if (g.g) hipGraphDestroy(g.g);
                return fail(c, TFIM_E_HIP, "wave graph construction failed: %s", hipGetErrorString(e));
            }
            sl.graphs.push_back(g);
            G = &sl.graphs.back();
            c->wave_graphs++;
        } else {
            HIPCHK(c, hipGraphExecKernelNodeSetParams(G->x, G->stage, &sp));
            if (has_fill) HIPCHK(c, hipGraphExecKernelNodeSetParams(G->x, G->fill, &zp));
            for (int i = 0; i < nk; i++) HIPCHK(c, hipGraphExecKernelNodeSetParams(G->x, G->kern[i], &kp[i]));
            HIPCHK(c, hipGraphExecKernelNodeSetParams(G->x, G->drain, &dp));
        }
        HIPCHK(c, hipGraphLaunch(G->x, c->stream));
        c->wave_graph_replays++;
    }
    HIPCHK(c, hipEventRecord(sl.done, c->stream));
    return TFIM_OK;
}

}  // namespace

namespace {

// Matrix-core path of a wave (arguments already checked by wave_submit):
// descriptor = item ids | item_off | col_off | col ids | exclusion bitmaps |
// row -> complement index; d_mat = prefix masks [n_rows][pitch] then the
// count matrix [n_rows][n_cols].  Marks the slot busy with its ticket.
int wave_submit_mfma(tfim_ctx *c, tfim_ctx::Slot &sl, const int32_t *item_ids, const int32_t *item_off,
                     int32_t n_rows, const int32_t *col_ids, const int32_t *col_off, const int32_t *compl_rows,
                     int32_t n_compl, const uint64_t *excl, int32_t min_count, int64_t n_exp, int64_t n_out) {
    const int64_t n_item_ids = item_off[n_rows];
    const int64_t excl_words = ((int64_t)c->n_cols + 63) / 64;
    auto align16 = [](int64_t x) { return (x + 15) / 16 * 16; };
    const int64_t o_items = 0;
    const int64_t o_ioff = align16(o_items + n_item_ids * 4);
    const int64_t o_coff = align16(o_ioff + (int64_t)(n_rows + 1) * 4);
    const int64_t o_cids = align16(o_coff + (int64_t)(n_rows + 1) * 4);
    const int64_t o_excl = align16(o_cids + n_exp * 4);
    const int64_t o_rc = align16(o_excl + (int64_t)n_compl * excl_words * 8);
    const int64_t total = o_rc + (int64_t)n_rows * 4;
    if (int r = ensure_desc(c, sl, total)) return r;
    if (int r = ensure_out(c, sl, n_out)) return r;
    const int64_t mask_words = (int64_t)n_rows * c->pitch;
    const int64_t c_ints = (int64_t)n_rows * c->n_cols;
    const int64_t c_off = (mask_words + 31) / 32 * 32;   // 256-byte aligned
    if (int r = ensure_pair<uint64_t>(c, &sl.d_mat, nullptr, &sl.mat_cap, c_off + (c_ints + 1) / 2)) return r;
    char *h = sl.h_desc;
    memcpy(h + o_items, item_ids, (size_t)n_item_ids * 4);
    memcpy(h + o_ioff, item_off, (size_t)(n_rows + 1) * 4);
    int32_t *coff = reinterpret_cast<int32_t *>(h + o_coff);
    int32_t *cids = reinterpret_cast<int32_t *>(h + o_cids);
    if (col_ids) {
        memcpy(coff, col_off, (size_t)(n_rows + 1) * 4);
        if (n_exp) memcpy(cids, col_ids, (size_t)n_exp * 4);
    } else {   // every row against all survivors: out = the count matrix itself
        for (int32_t r = 0; r <= n_rows; r++) coff[r] = r * c->n_cols;
        for (int32_t r = 0; r < n_rows; r++)
            for (int32_t j = 0; j < c->n_cols; j++) cids[(int64_t)r * c->n_cols + j] = j;
    }
    if (n_compl) memcpy(h + o_excl, excl, (size_t)(n_compl * excl_words * 8));
    int32_t *rc = reinterpret_cast<int32_t *>(h + o_rc);
    for (int32_t r = 0; r < n_rows; r++) rc[r] = -1;
    for (int32_t j = 0; j < n_compl; j++) rc[compl_rows[j]] = j;

    uint64_t *masks = sl.d_mat;
    int32_t *C = reinterpret_cast<int32_t *>(sl.d_mat + c_off);
    const int32_t *d_items = reinterpret_cast<const int32_t *>(sl.d_desc + o_items);
    const int32_t *d_ioff = reinterpret_cast<const int32_t *>(sl.d_desc + o_ioff);
    KernelLaunch M, P, G;
    plan_wave_masks(c->d_bm, c->pitch, d_items, d_ioff, n_rows, masks, M);
    // split the words so that about three workgroups per CU run, keeping at
    // least 16 stages of 512 bits per split (short rows are not split); the
    // third tuner scales that split by 1/2, 1 or 2 (each tried three times,
    // then the cheapest per unit of work, another re-tried every 16 waves)
    const int ra = rp_rows_per_tile(n_rows);
    const int64_t tiles = ((n_rows + ra - 1) / ra) * ((c->n_cols + 127) / 128);
    const int64_t stages = c->pitch / 8;
    const int64_t base = c->mfma_split == 1
                             ? 1
                             : std::max<int64_t>(1, std::min<int64_t>((3 * c->n_cu + tiles - 1) / tiles, stages / 16));
    int sk_arm = -1;
    int64_t split = base;
    if (base > 1) {
        sk_arm = c->sk.pick();
        split = SplitTuner::split_of(sk_arm, base, stages);
    }
    // fp4 matrix cores (twice the int8 rate, same counts; int8 if configured)
    // with the 128-column tiles the occupancy rule above assumes
    const int64_t zs =
        plan_popmat_mfma_rp(masks, c->d_bm, C, n_rows, c->n_cols, c->pitch, (int)split, P, !c->mfma_int8, false);
    WaveDesc d{};
    d.item_ids = d_items;
    d.item_off = d_ioff;
    d.col_off = reinterpret_cast<const int32_t *>(sl.d_desc + o_coff);
    d.col_ids = reinterpret_cast<const int32_t *>(sl.d_desc + o_cids);
    d.out = sl.d_out;
    d.excl = reinterpret_cast<const uint64_t *>(sl.d_desc + o_excl);
    d.excl_words = (int32_t)excl_words;
    int32_t max_row_cols = 0;
    for (int32_t r = 0; r < n_rows; r++) max_row_cols = std::max(max_row_cols, coff[r + 1] - coff[r]);
    plan_wave_gather(d, C, n_rows, c->n_cols, reinterpret_cast<const int32_t *>(sl.d_desc + o_rc), min_count,
                     sl.d_out + n_exp, max_row_cols, G);
    KernelLaunch *ks[3] = {&M, &P, &G};
    if (int r = issue_wave(c, sl, total, zs > 1 ? C : nullptr, zs > 1 ? c_ints : 0, n_out, n_out, ks, 3)) return r;
    sl.busy = true;
    sl.stamp_at = -1;
    sl.ticket = c->next_ticket++;
    sl.n_out = n_out;
    sl.n_rows = n_rows;
    sl.sk_arm = sk_arm;
    sl.mask_arm = -1;
    sl.sk_fused = false;
    sl.sk_work = (double)((n_rows + ra - 1) / ra * ra) * (double)((c->n_cols + 127) / 128 * 128) * (double)c->pitch;
    c->wave_mfma_launches++;
    return TFIM_OK;
}

// Fused matrix-core path of a wave (arguments already checked by
// wave_submit): one k_wave_mfma launch.  Descriptor = slot rows | item_off |
// prefix slots | gather segments | gather entries | row -> complement index |
// complement x0 | exclusion bitmaps; output block = explicit outputs | two
// results per complement row | (device only, 64-aligned) full counts of the
// complement rows | arrival counters of the row tiles, all zeroed by the
// wave's fill kernel.  Returns kFusedNoFit, with nothing issued, when the
// wave has more than kFusedMaxSlots distinct prefix items or needs more LDS
// than a CU has.  Marks the slot busy with its ticket.
int wave_submit_fused(tfim_ctx *c, tfim_ctx::Slot &sl, const int32_t *item_ids, const int32_t *item_off,
                      int32_t n_rows, const int32_t *col_ids, const int32_t *col_off, const int32_t *compl_rows,
                      int32_t n_compl, const uint64_t *excl, int32_t min_count, int64_t n_exp, int64_t n_out) {
    c->fused_nofit = "a wave of 1 to 2^23 - 1 rows over at least one survivor";
    if (c->n_cols < 1 || n_rows < 1 || (int64_t)n_rows >= (1 << 23)) return kFusedNoFit;
    const int64_t n_item_ids = item_off[n_rows];
    // distinct prefix items -> slots (fslot_of is cleared again below)
    if ((int32_t)c->fslot_of.size() < c->n_items) c->fslot_of.assign(c->n_items, -1);
    std::vector<int32_t> slot_rows;
    std::vector<int32_t> islots((size_t)n_item_ids);
    bool fits = true;
    for (int64_t i = 0; i < n_item_ids; i++) {
        int32_t &so = c->fslot_of[item_ids[i]];
        if (so < 0) {
            if ((int)slot_rows.size() >= kFusedMaxSlots) {
                fits = false;
                break;
            }
            so = (int32_t)slot_rows.size();
            slot_rows.push_back(item_ids[i]);
        }
        islots[(size_t)i] = so;
    }
    for (int32_t it : slot_rows) c->fslot_of[it] = -1;
    c->fused_nofit = "at most 256 distinct prefix items";
    if (!fits) return kFusedNoFit;
    // tile shape: 64- or 128-column survivor tiles, 64 to 256 rows
    const int tj = c->n_cols <= 64 ? 1 : 2;
    const int ti = n_rows <= 64 ? 1 : (n_rows <= 128 || tj == 2) ? 2 : 4;
    const int RA = 64 * ti, RB = 64 * tj;
    const int tiles_m = (n_rows + RA - 1) / RA, tiles_n = (c->n_cols + RB - 1) / RB;
    int max_list = 0;
    for (int tm = 0; tm < tiles_m; tm++)
        max_list = std::max(max_list, item_off[std::min(n_rows, (tm + 1) * RA)] - item_off[tm * RA]);
    const size_t lds = wave_mfma_lds_bytes(ti, tj, (int)slot_rows.size(), max_list);
    c->fused_nofit = "prefix lists of a row tile that fit the LDS";
    if (lds > kFusedMaxLds) return kFusedNoFit;

    const bool csr = col_ids || n_compl > 0;
    const int64_t excl_words = ((int64_t)c->n_cols + 63) / 64;
    auto align16 = [](int64_t x) { return (x + 15) / 16 * 16; };
    const int64_t n_seg = csr ? (int64_t)tiles_n * (n_rows + 1) : 0;
    const int64_t o_slots = 0;
    const int64_t o_ioff = align16(o_slots + (int64_t)slot_rows.size() * 4);
    const int64_t o_isl = align16(o_ioff + (int64_t)(n_rows + 1) * 4);
    const int64_t o_seg = align16(o_isl + n_item_ids * 4);
    const int64_t o_ent = align16(o_seg + n_seg * 4);
    const int64_t o_rc = align16(o_ent + (csr ? n_exp : 0) * 8);
    const int64_t o_x0 = align16(o_rc + (n_compl ? (int64_t)n_rows * 4 : 0));
    const int64_t o_excl = align16(o_x0 + (int64_t)n_compl * 4);
    const int64_t total = o_excl + (int64_t)n_compl * excl_words * 8;
    // outputs: explicit | results | complement full counts | arrival counters
    const int64_t o_acc = (n_out + 63) / 64 * 64;
    const int64_t o_arr = o_acc + (int64_t)n_compl * c->n_cols;
    const int64_t n_dev = n_compl ? o_arr + tiles_m : n_out;
    c->fused_nofit = "outputs within int32 indices";
    if (n_dev > INT32_MAX) return kFusedNoFit;
    if (int r = ensure_desc(c, sl, total)) return r;
    if (int r = ensure_out(c, sl, n_dev)) return r;
    char *h = sl.h_desc;
    memcpy(h + o_slots, slot_rows.data(), slot_rows.size() * 4);
    memcpy(h + o_ioff, item_off, (size_t)(n_rows + 1) * 4);
    memcpy(h + o_isl, islots.data(), (size_t)n_item_ids * 4);
    if (csr) {
        // listed outputs grouped by (column tile, row): counting sort
        int32_t *seg = reinterpret_cast<int32_t *>(h + o_seg);
        int2 *ent = reinterpret_cast<int2 *>(h + o_ent);
        std::fill(seg, seg + n_seg, 0);
        for (int32_t r = 0; r < n_rows; r++)
            for (int32_t q = col_off[r]; q < col_off[r + 1]; q++) seg[(int64_t)(col_ids[q] / RB) * (n_rows + 1) + r]++;
        int32_t run = 0;
        for (int64_t i = 0; i < n_seg; i++) {
            const int32_t n = seg[i];
            seg[i] = run;
            run += n;
        }
        // seg[t][r] is the start of (t, r); entries are written in order
        std::vector<int32_t> fillp(seg, seg + n_seg);
        for (int32_t r = 0; r < n_rows; r++)
            for (int32_t q = col_off[r]; q < col_off[r + 1]; q++) {
                const int32_t col = col_ids[q], t = col / RB;
                ent[fillp[(size_t)t * (n_rows + 1) + r]++] = make_int2(q, (r << 8) | (col - t * RB));
            }
        // (seg[t][n_rows], an empty bucket, is the end of tile t's entries)
    }
    if (n_compl) {
        int32_t *rc = reinterpret_cast<int32_t *>(h + o_rc);
        int32_t *x0 = reinterpret_cast<int32_t *>(h + o_x0);
        for (int32_t r = 0; r < n_rows; r++) rc[r] = -1;
        for (int32_t j = 0; j < n_compl; j++) {
            rc[compl_rows[j]] = j;
            x0[j] = item_ids[item_off[compl_rows[j]]];
        }
        memcpy(h + o_excl, excl, (size_t)(n_compl * excl_words * 8));
    }
    // word split: the workgroups that fit at once (two per CU), at least 32
    // stages per split, scaled by the fused path's own split tuner (each
    // workgroup's set-up and output pass are worth a few stages)
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    const int64_t stages = c->pitch / 8;
    const int64_t base = c->mfma_split == 1
                             ? 1
                             : std::max<int64_t>(1, std::min<int64_t>((2 * c->n_cu + tiles - 1) / tiles, stages / 32));
    int sk_arm = -1;
    int64_t split = base;
    if (base > 1) {
        sk_arm = c->fk.pick();
        split = SplitTuner::split_of(sk_arm, base, stages);
    }
    int64_t kps = (c->pitch + split - 1) / split;
    kps = (kps + 7) / 8 * 8;
    kps = std::min<int64_t>(kps, (int64_t)1 << 18);   // <= 2^24 bits per split: the fp32 accumulators stay exact
    const int64_t zs = (c->pitch + kps - 1) / kps;
    WaveMfmaDesc d{};
    const char *dd = sl.d_desc;
    d.slot_rows = reinterpret_cast<const int32_t *>(dd + o_slots);
    d.item_off = reinterpret_cast<const int32_t *>(dd + o_ioff);
    d.item_slots = reinterpret_cast<const int32_t *>(dd + o_isl);
    d.seg_off = csr ? reinterpret_cast<const int32_t *>(dd + o_seg) : nullptr;
    d.entries = reinterpret_cast<const int2 *>(dd + o_ent);
    d.row_compl = n_compl ? reinterpret_cast<const int32_t *>(dd + o_rc) : nullptr;
    d.compl_x0 = reinterpret_cast<const int32_t *>(dd + o_x0);
    d.excl = reinterpret_cast<const uint64_t *>(dd + o_excl);
    d.out = sl.d_out;
    d.res = sl.d_out + n_exp;
    d.acc = sl.d_out + o_acc;
    d.arrive = sl.d_out + o_arr;
    d.k_per_split = kps;
    d.n_rows = n_rows;
    d.n_cols = c->n_cols;
    d.n_slots = (int32_t)slot_rows.size();
    d.excl_words = (int32_t)excl_words;
    d.min_count = min_count;
    d.tiles_n = tiles_n;
    d.dense_out = csr ? 0 : 1;
    KernelLaunch K;
    c->fused_nofit = "a launch the kernel accepts (tile, slots, LDS attribute)";
    if (!plan_wave_mfma(c->d_bm, c->pitch, d, ti, tj, tiles_m, (int)zs, lds, K, c->mfma_int8)) return kFusedNoFit;
    KernelLaunch *ks[1] = {&K};
    if (int r = issue_wave(c, sl, total, nullptr, 0, n_out, n_dev, ks, 1)) return r;
    sl.busy = true;
    sl.stamp_at = -1;
    sl.ticket = c->next_ticket++;
    sl.n_out = n_out;
    sl.n_rows = n_rows;
    sl.sk_arm = sk_arm;
    sl.mask_arm = -1;
    sl.sk_fused = true;
    sl.sk_work = (double)tiles_m * RA * (double)tiles_n * RB * (double)c->pitch;
    c->wave_fused_launches++;
    return TFIM_OK;
}

// Common body of tfim_count_nodes_submit and tfim_count_nodes_ex_submit.
// Complement rows (n_compl > 0) need explicit column offsets (col_off), and
// their waves run on the L2 kernel followed by the complement reduction, or
// on the matrix-core path.  path TFIM_PATH_PACKED keeps a wave on the packed
// kernels, TFIM_PATH_MFMA_I8 sends it to the matrix cores.
int wave_submit(tfim_ctx *c, const int32_t *item_ids, const int32_t *item_off, int32_t n_rows,
                const int32_t *col_ids, const int32_t *col_off, const int32_t *compl_rows, int32_t n_compl,
                const uint64_t *excl, int32_t min_count, int32_t path, int64_t *ticket) {
    if (!c || !item_ids || !item_off || !ticket || n_rows < 0 || n_compl < 0)
        return fail(c, TFIM_E_INVALID, "count_nodes: null argument");
    std::lock_guard<std::mutex> wave_lock(c->wave_mu);
    *ticket = -1;
    if (col_ids && !col_off) return fail(c, TFIM_E_INVALID, "count_nodes: col_ids without col_off");
    if (n_compl > 0) {
        if (!col_off || !compl_rows || !excl)
            return fail(c, TFIM_E_INVALID, "count_nodes_ex: complement rows need col_off, compl_rows and excl");
        if (!col_ids && col_off[n_rows] != 0)
            return fail(c, TFIM_E_INVALID, "count_nodes_ex: explicit columns without col_ids");
    }
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "count_nodes: no bitmaps loaded");
    if (path != TFIM_PATH_AUTO && path != TFIM_PATH_PACKED && path != TFIM_PATH_MFMA_I8)
        return fail(c, TFIM_E_INVALID, "count_nodes: unknown path %d", path);
    // the path argument overrides the configured wave kernel for this wave:
    // TFIM_PATH_MFMA_I8 takes the fused matrix-core kernel, or the
    // three-launch form when the wave is outside the fused kernel's limits
    const bool mfma_any = path == TFIM_PATH_MFMA_I8;
    const int wave_kernel = mfma_any ? 5
                            : path == TFIM_PATH_PACKED && c->wave_kernel >= 4 ? 0 : c->wave_kernel;
    const bool allow_mfma = path != TFIM_PATH_PACKED;
    if (n_rows == 0) return TFIM_OK;
    tfim_ctx::Slot &sl = c->slots[c->next_ticket % tfim_ctx::kWaveSlots];
    if (sl.busy)
        return fail(c, TFIM_E_STATE, "count_nodes_submit: %d waves in flight; wait on a ticket first",
                    tfim_ctx::kWaveSlots);
    if (item_off[0] != 0) return fail(c, TFIM_E_INVALID, "count_nodes: item_off[0] != 0");
    for (int32_t r = 0; r < n_rows; r++)
        if (item_off[r + 1] <= item_off[r]) return fail(c, TFIM_E_INVALID, "count_nodes: row %d has no prefix item", r);
    const int64_t n_item_ids = item_off[n_rows];
    for (int64_t i = 0; i < n_item_ids; i++)
        if (item_ids[i] < 0 || item_ids[i] >= c->n_items)
            return fail(c, TFIM_E_INVALID, "count_nodes: item %d out of range", item_ids[i]);
    int64_t n_out;
    const bool csr = col_ids || n_compl > 0;
    if (csr) {
        if (col_off[0] != 0) return fail(c, TFIM_E_INVALID, "count_nodes: col_off[0] != 0");
        for (int32_t r = 0; r < n_rows; r++)
            if (col_off[r + 1] < col_off[r]) return fail(c, TFIM_E_INVALID, "count_nodes: col_off decreasing at %d", r);
        n_out = col_off[n_rows];
        for (int64_t i = 0; i < n_out; i++)
            if (col_ids[i] < 0 || col_ids[i] >= c->n_cols)
                return fail(c, TFIM_E_INVALID, "count_nodes: column %d out of range [0,%d)", col_ids[i], c->n_cols);
    } else {
        n_out = (int64_t)n_rows * c->n_cols;
    }
    for (int32_t j = 0; j < n_compl; j++) {
        const int32_t r = compl_rows[j];
        if (r < 0 || r >= n_rows || (j > 0 && r <= compl_rows[j - 1]))
            return fail(c, TFIM_E_INVALID, "count_nodes_ex: complement rows must be ascending row indices");
        if (col_off[r + 1] != col_off[r])
            return fail(c, TFIM_E_INVALID, "count_nodes_ex: complement row %d has explicit columns", r);
        if (item_ids[item_off[r]] >= c->n_cols)
            return fail(c, TFIM_E_INVALID, "count_nodes_ex: complement row %d starts with a non-survivor item", r);
    }
    // output: explicit counts | two results per complement row | (device
    // only) every survivor's count of each complement row
    const int64_t n_exp = n_out;
    const int64_t excl_words = ((int64_t)c->n_cols + 63) / 64;
    const int64_t full_base = (n_exp + 2 * (int64_t)n_compl + 63) / 64 * 64;
    const int64_t n_dev = n_compl ? full_base + (int64_t)n_compl * c->n_cols : n_exp;
    n_out = n_exp + 2 * (int64_t)n_compl;
    if (n_dev > INT32_MAX) return fail(c, TFIM_E_RANGE, "count_nodes: output too large");
    if (n_out == 0) return TFIM_OK;
    if (int r = set_device(c)) return r;

    // work of this wave in row-words: sum over rows of (prefix + columns) x pitch
    const double work = ((double)n_item_ids + (double)n_out) * (double)c->pitch;
    // LDS slots: the distinct rows this wave touches; slot_of is cleared
    // again on every way out of this call
    std::vector<int32_t> slot_rows;
    struct SlotReset {
        tfim_ctx *c;
        const std::vector<int32_t> &rows;
        ~SlotReset() {
            for (int32_t it : rows) c->slot_of[it] = -1;
        }
    } slot_reset{c, slot_rows};
    bool lds = (wave_kernel == 0 || wave_kernel == 2 || wave_kernel == 3) && n_compl == 0 &&
               n_out <= kLdsMaxOut;
    if (lds) {
        if ((int32_t)c->slot_of.size() < c->n_items) c->slot_of.assign(c->n_items, -1);
        auto slot = [&](int32_t item) {
            int32_t &sl = c->slot_of[item];
            if (sl < 0) {
                sl = (int32_t)slot_rows.size();
                slot_rows.push_back(item);
            }
            return sl;
        };
        for (int64_t i = 0; i < n_item_ids && (int)slot_rows.size() <= kLdsMaxSlots; i++) slot(item_ids[i]);
        if (col_ids)
            for (int64_t i = 0; i < n_out && (int)slot_rows.size() <= kLdsMaxSlots; i++) slot(col_ids[i]);
        else
            for (int32_t i = 0; i < c->n_cols && (int)slot_rows.size() <= kLdsMaxSlots; i++) slot(i);
        lds = (int)slot_rows.size() <= kLdsMaxSlots;
        if (!lds)
            for (int32_t it : slot_rows) c->slot_of[it] = -1;
    }
    if ((wave_kernel == 2 || wave_kernel == 3) && !lds && n_compl == 0)
        return fail(c, TFIM_E_RANGE, "count_nodes: wave does not fit the LDS kernel (%zu items, %lld outputs)",
                    slot_rows.size(), (long long)n_out);
    // autotune over (kernel, columns per tile) for waves both kernels accept:
    // each arm is sampled three times, then the cheapest per row-word is
    // taken, one other arm being re-sampled every 16 calls (round robin)
    int tuned = -1;
    if (lds && wave_kernel == 0) {
        c->tune_calls++;
        // every arm once; then up to three samples of the arms within 1.5x
        // of the best one's cost
        // of the best one's cost (counting the waves issued, whose timings
        // arrive when they are waited on, and waiting for the first ones)
        for (int a = 0; a < kTuneArms && tuned < 0; a++)
            if (c->tune_issued[a] == 0) tuned = a;
        int best = -1;
        for (int a = 0; a < kTuneArms; a++)
            if (c->tune_samples[a] > 0 && (best < 0 || c->tune_cost[a] < c->tune_cost[best])) best = a;
        if (best < 0) best = 2;   // nothing measured yet: an LDS arm
        for (int a = 0; a < kTuneArms && tuned < 0; a++)
            if (c->tune_samples[a] > 0 && c->tune_issued[a] < 3 && c->tune_cost[a] <= 1.5 * c->tune_cost[best])
                tuned = a;
        if (tuned < 0) {
            tuned = best;
            if (c->tune_calls % 16 == 0) {
                // round robin over the other arms; one measured at more than
                // 1.5x the best one's cost is re-tried only every 8th time
                // its turn comes, beyond 3x only every 32nd
                const int64_t turn = c->tune_calls / 16;
                const int cand = (best + 1 + (int)(turn % (kTuneArms - 1))) % kTuneArms;
                const double ratio = c->tune_cost[cand] / std::max(1e-30, c->tune_cost[best]);
                const int64_t period = ratio <= 1.5 ? 1 : ratio <= 3.0 ? 8 : 32;
                if ((turn / (kTuneArms - 1)) % period == 0) tuned = cand;
            }
        }
        c->tune_issued[tuned]++;
        if (kArmKernel[tuned] == 1) {
            lds = false;
            for (int32_t it : slot_rows) c->slot_of[it] = -1;
            slot_rows.clear();
        } else if (kArmKernel[tuned] == 5 && allow_mfma) {
            // the fused matrix-core kernel on a wave the LDS kernel also takes
            const int r = wave_submit_fused(c, sl, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl,
                                            excl, min_count, n_out, n_out);
            if (r == TFIM_OK) {
                sl.tuned = tuned;
                sl.work = work;
                sl.mf_arm = -1;
                *ticket = sl.ticket;
                return TFIM_OK;
            }
            if (r != kFusedNoFit) return r;
            tuned = 2;   // outside the fused kernel's limits: an LDS arm counts it
        } else if (kArmKernel[tuned] == 5) {
            tuned = 2;
        }
    }
    // Waves outside the LDS kernel go to the L2 kernel (counts only the
    // listed pairs) or to the matrix cores (masks + fp4 matrix count of every
    // (row, survivor) pair + gather).  wave_kernel 4 forces the second; in
    // auto mode a second tuner predicts each arm's time as its work times its
    // measured cost per unit (priors from the two paths' rates on MI355X
    // until measured; the other arm re-tried every 16 waves) and takes the
    // cheaper.
    int mf_arm = -1;
    double mf_work = 0;
    if (!lds && allow_mfma && wave_kernel == 5) {
        // the fused matrix-core kernel, forced (or asked for by the path)
        const int r = wave_submit_fused(c, sl, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl,
                                        excl, min_count, n_exp, n_out);
        if (r == TFIM_OK) {
            sl.tuned = -1;
            sl.mf_arm = -1;
            *ticket = sl.ticket;
            return TFIM_OK;
        }
        if (r != kFusedNoFit) return r;
        if (!mfma_any)
            return fail(c, TFIM_E_RANGE, "count_nodes: wave outside the fused matrix-core kernel's limits (it needs %s)",
                        c->fused_nofit);
    }
    if (!lds && allow_mfma && (wave_kernel == 0 || wave_kernel == 4 || mfma_any)) {
        const int64_t mask_bytes = (int64_t)n_rows * c->pitch * 8;
        // (the mask kernel's grid has one row per y index, at most 65535)
        const bool fits = c->n_cols > 0 && n_rows <= 65535 && mask_bytes <= kMfmaMaskBytes &&
                          (int64_t)n_rows * c->n_cols <= INT32_MAX;
        if ((wave_kernel == 4 || mfma_any) && !fits)
            return fail(c, TFIM_E_RANGE, "count_nodes: wave too large for the matrix-core path (%d rows)", n_rows);
        if (fits || wave_kernel == 0) {
            double pairs = (double)n_exp;
            for (int32_t j = 0; j < n_compl; j++) {
                const uint64_t *ex = excl + (int64_t)j * excl_words;
                int64_t dead = 0;
                for (int64_t w = 0; w < excl_words; w++) dead += __builtin_popcountll(ex[w]);
                pairs += (double)std::max<int64_t>(1, c->n_cols - dead);
            }
            const double w_l2 = ((double)n_item_ids + pairs) * (double)c->pitch;
            // padded matrix pairs, plus the mask build's traffic weighted by
            // the ratio of the matrix cores' word-pair rate to HBM word rate (~39)
            const int ra = rp_rows_per_tile(n_rows);
            const double w_mf = ((double)((n_rows + ra - 1) / ra * ra) * (double)((c->n_cols + 127) / 128 * 128) +
                                 39.0 * (double)(n_item_ids + n_rows)) *
                                (double)c->pitch;
            // fused: padded (row, survivor) pairs of its tiles, plus the
            // prefix ANDs built from LDS (an eighth of a pair per prefix item
            // and tile column)
            const int tjf = c->n_cols <= 64 ? 1 : 2;
            const int raf = 64 * (n_rows <= 64 ? 1 : (n_rows <= 128 || tjf == 2) ? 2 : 4), rbf = 64 * tjf;
            const double tnf = (double)((c->n_cols + rbf - 1) / rbf);
            const double w_fu = ((double)((n_rows + raf - 1) / raf * raf) * tnf * rbf +
                                 8.0 * (double)n_item_ids * tnf) * (double)c->pitch;
            const double w[3] = {w_l2, w_mf, w_fu};
            int arm = 1;
            if (wave_kernel == 0) {
                c->mf_calls++;
                // the costs start from priors (the L2 kernel's and the fp4
                // matrix cores' rates on MI355X, about 300 and 3,000
                // Tbit-op/s over their units of work) and become measured
                // EMAs; the other arm is tried on every 16th wave (one
                // predicted at more than twice the chosen one's time only
                // every eighth turn, unless it has fewer than three samples:
                // one wave of an unusual shape, or timed beside other device
                // work, would otherwise shut it out for the rest of a run).
                // The fused matrix-core kernel is not among the automatic
                // arms: its prefix build from LDS made it the slowest arm on
                // every workload measured (clickstream: 39 fused waves took
                // 49 ms of count kernels where the three launches took 16;
                // gene 5.9 ms against 1.8), so it runs only when asked for.
                arm = fits && w[1] * c->mf_cost[1] < kMfmaBias * w[0] * c->mf_cost[0] ? 1 : 0;
                if (c->mf_calls % 16 == 0 && fits) {
                    const int64_t turn = c->mf_calls / 16;
                    const int cand = 1 - arm;
                    if (c->mf_samples[cand] < 3 || w[cand] * c->mf_cost[cand] <= 2.0 * w[arm] * c->mf_cost[arm] ||
                        turn % 8 == 0)
                        arm = cand;
                }
            }
            if (arm == 2) {
                const int r = wave_submit_fused(c, sl, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl,
                                                excl, min_count, n_exp, n_out);
                if (r == TFIM_OK) {
                    sl.tuned = tuned;
                    sl.work = tuned >= 0 ? work : 0;
                    sl.mf_arm = 2;
                    sl.mf_work = w_fu;
                    *ticket = sl.ticket;
                    return TFIM_OK;
                }
                if (r != kFusedNoFit) return r;
                arm = fits && w_mf * c->mf_cost[1] < w_l2 * c->mf_cost[0] ? 1 : 0;
            }
            if (wave_kernel == 0) {
                mf_arm = arm;
                mf_work = w[arm];
            }
            if (arm == 1) {
                int r = wave_submit_mfma(c, sl, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl,
                                         excl, min_count, n_exp, n_out);
                if (r) return r;
                // a wave the first tuner sent off the LDS kernel is timed
                // for that arm as well, at what leaving the LDS kernel
                // actually cost (here the matrix cores)
                sl.tuned = tuned;
                sl.work = tuned >= 0 ? work : 0;
                sl.mf_arm = mf_arm;
                sl.mf_work = mf_work;
                *ticket = sl.ticket;
                return TFIM_OK;
            }
        }
    }
    const int tile_cols = c->tile_cols ? c->tile_cols : (tuned >= 0 ? kArmTile[tuned] : 32);

    // tiles of <= tile_cols columns per row
    std::vector<WaveTile> tiles;
    tiles.reserve((size_t)n_rows * 2);
    for (int32_t r = 0; r < n_rows; r++) {
        const int32_t nc = csr ? col_off[r + 1] - col_off[r] : c->n_cols;
        for (int32_t b = 0; b < nc; b += tile_cols) tiles.push_back({r, b, std::min(nc, b + tile_cols), 0});
    }
    // complement rows: one tile per 64 survivor ids holding a column to count
    for (int32_t j = 0; j < n_compl; j++) {
        const int32_t r = compl_rows[j], x0 = item_ids[item_off[r]];
        const uint64_t *ex = excl + (int64_t)j * excl_words;
        for (int64_t w = 0; w < excl_words; w++) {
            uint64_t live = ~ex[w];
            if (x0 >> 6 == w) live |= 1ull << (x0 & 63);
            const int32_t b = (int32_t)(w * 64), e = std::min<int32_t>(c->n_cols, b + 64);
            if (e - b < 64) live &= (1ull << (e - b)) - 1;
            if (live) tiles.push_back({r, b, e, j + 1});
        }
    }
    if (tiles.empty()) {   // only empty tails: n_out == 0, handled above
        for (int32_t it : slot_rows) c->slot_of[it] = -1;
        return TFIM_OK;
    }

    // pack the descriptor block: [slot_rows] | item ids or slots | item_off | col_off | col ids or slots | tiles
    auto align16 = [](int64_t x) { return (x + 15) / 16 * 16; };
    const int64_t o_slots = 0;
    const int64_t o_items = align16(o_slots + (int64_t)slot_rows.size() * 4);
    const int64_t o_ioff = align16(o_items + n_item_ids * 4);
    const int64_t o_coff = align16(o_ioff + (int64_t)(n_rows + 1) * 4);
    const int64_t o_cids = align16(o_coff + (int64_t)(n_rows + 1) * 4);
    const bool need_cols = col_ids != nullptr || lds;
    const int64_t o_tiles = align16(o_cids + (need_cols ? n_exp * 4 : 0));
    const int64_t o_excl = align16(o_tiles + (int64_t)tiles.size() * (int64_t)sizeof(WaveTile));
    const int64_t o_crows = align16(o_excl + (int64_t)n_compl * excl_words * 8);
    const int64_t total = o_crows + (int64_t)n_compl * 4;
    if (int r = ensure_desc(c, sl, total)) return r;
    if (int r = ensure_out(c, sl, n_dev)) return r;
    char *h = sl.h_desc;
    int32_t *hi = reinterpret_cast<int32_t *>(h + o_items);
    int32_t *hc = reinterpret_cast<int32_t *>(h + o_cids);
    if (lds) {
        memcpy(h + o_slots, slot_rows.data(), slot_rows.size() * 4);
        for (int64_t i = 0; i < n_item_ids; i++) hi[i] = c->slot_of[item_ids[i]];
        if (col_ids)
            for (int64_t i = 0; i < n_out; i++) hc[i] = c->slot_of[col_ids[i]];
        else
            for (int32_t r = 0; r < n_rows; r++)
                for (int32_t j = 0; j < c->n_cols; j++) hc[(int64_t)r * c->n_cols + j] = c->slot_of[j];
        for (int32_t it : slot_rows) c->slot_of[it] = -1;
    } else {
        memcpy(hi, item_ids, (size_t)n_item_ids * 4);
        if (col_ids) memcpy(hc, col_ids, (size_t)n_exp * 4);
    }
    memcpy(h + o_ioff, item_off, (size_t)(n_rows + 1) * 4);
    int32_t *coff = reinterpret_cast<int32_t *>(h + o_coff);
    if (csr) {
        memcpy(coff, col_off, (size_t)(n_rows + 1) * 4);
    } else {
        for (int32_t r = 0; r <= n_rows; r++) coff[r] = r * c->n_cols;
    }
    // a complement row's counts go to its block of the device-only region
    for (int32_t j = 0; j < n_compl; j++) coff[compl_rows[j]] = (int32_t)(full_base + (int64_t)j * c->n_cols);
    if (n_compl) {
        memcpy(h + o_excl, excl, (size_t)(n_compl * excl_words * 8));
        memcpy(h + o_crows, compl_rows, (size_t)n_compl * 4);
    }
    if (lds) {
        // the LDS kernel takes tiles resolved to prefix and output ranges
        static_assert(sizeof(LdsTile) == sizeof(WaveTile), "tile records share the descriptor layout");
        // by falling cost 2 * columns + prefix items (a column reads and
        // counts its words, a prefix item only reads and ANDs them), so that
        // the waves of a workgroup, taking tiles in this order, finish a
        // chunk together: a stable counting sort on the cost, so equal costs
        // keep the ascending output order they are generated in
        LdsTile *lt = reinterpret_cast<LdsTile *>(h + o_tiles);
        std::vector<int32_t> &start = c->tile_bucket;
        int32_t max_cost = 0;
        for (const WaveTile &w : tiles)
            max_cost = std::max(max_cost, 2 * (w.col_end - w.col_begin) + item_off[w.row + 1] - item_off[w.row]);
        start.assign((size_t)max_cost + 2, 0);
        for (const WaveTile &w : tiles)
            start[(size_t)(max_cost - (2 * (w.col_end - w.col_begin) + item_off[w.row + 1] - item_off[w.row])) + 1]++;
        for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
        for (const WaveTile &w : tiles) {
            const int32_t k = item_off[w.row + 1] - item_off[w.row], nc = w.col_end - w.col_begin;
            LdsTile &t = lt[start[(size_t)(max_cost - (2 * nc + k))]++];
            t.ib = item_off[w.row];
            t.k = k;
            t.cb = coff[w.row] + w.col_begin;
            t.nc = nc;
        }
    } else {
        memcpy(h + o_tiles, tiles.data(), tiles.size() * sizeof(WaveTile));
    }
    KernelLaunch K, R;
    if (lds) {
        WaveLdsDesc d;
        d.slot_rows = reinterpret_cast<const int32_t *>(sl.d_desc + o_slots);
        d.item_slots = reinterpret_cast<const int32_t *>(sl.d_desc + o_items);
        d.item_off = reinterpret_cast<const int32_t *>(sl.d_desc + o_ioff);
        d.col_slots = reinterpret_cast<const int32_t *>(sl.d_desc + o_cids);
        d.col_off = reinterpret_cast<const int32_t *>(sl.d_desc + o_coff);
        d.tiles = reinterpret_cast<const LdsTile *>(sl.d_desc + o_tiles);
        d.out = sl.d_out;
        d.n_slots = (int32_t)slot_rows.size();
        d.n_tiles = (int32_t)tiles.size();
        d.n_out = (int32_t)n_out;
        const bool wide = wave_kernel == 3 || (tuned >= 0 && kArmWide[tuned]);
        if (!plan_count_wave_lds(c->d_bm, c->pitch, d, c->n_cu, wide, K))
            return fail(c, TFIM_E_RANGE, "count_nodes: LDS kernel rejected %d slots", d.n_slots);
        c->wave_lds_launches++;
    } else {
        WaveDesc d;
        d.item_ids = reinterpret_cast<const int32_t *>(sl.d_desc + o_items);
        d.item_off = reinterpret_cast<const int32_t *>(sl.d_desc + o_ioff);
        d.col_off = reinterpret_cast<const int32_t *>(sl.d_desc + o_coff);
        d.col_ids = col_ids ? reinterpret_cast<const int32_t *>(sl.d_desc + o_cids) : nullptr;
        d.tiles = reinterpret_cast<const WaveTile *>(sl.d_desc + o_tiles);
        d.out = sl.d_out;
        d.excl = reinterpret_cast<const uint64_t *>(sl.d_desc + o_excl);
        d.excl_words = (int32_t)excl_words;
        d.n_tiles = (int32_t)tiles.size();
        const int rw = c->pitch >= 2048 ? 4 : (c->pitch >= 512 ? 2 : 1);
        const int64_t cw = wave_chunk_words(rw);
        d.n_chunks = (int32_t)((c->pitch + cw - 1) / cw);
        const int64_t units = (int64_t)d.n_chunks * d.n_tiles;
        int64_t blocks = (units + 3) / 4;
        blocks = std::min<int64_t>(blocks, (int64_t)c->n_cu * 8);
        blocks = std::max<int64_t>(8, (blocks + 7) / 8 * 8);
        plan_count_wave(c->d_bm, c->pitch, d, (int)blocks, rw, K);
        if (n_compl)
            plan_compl_reduce(d, reinterpret_cast<const int32_t *>(sl.d_desc + o_crows), n_compl, c->n_cols,
                              min_count, sl.d_out + n_exp, R);
    }
    KernelLaunch *ks[2] = {&K, &R};
    if (int r = issue_wave(c, sl, total, nullptr, 0, n_out, n_dev, ks, n_compl ? 2 : 1)) return r;
    sl.busy = true;
    sl.stamp_at = -1;
    sl.ticket = c->next_ticket++;
    sl.n_out = n_out;
    sl.n_rows = n_rows;
    sl.tuned = tuned;
    sl.work = work;
    sl.mf_arm = mf_arm;
    sl.mf_work = mf_work;
    sl.sk_arm = -1;
    sl.mask_arm = -1;
    *ticket = sl.ticket;
    return TFIM_OK;
}

// Mask rows (tfim_count_masks_submit).  Waves the mask kernel does not take
// (more than kLdsMaxOut outputs, the matrix-core path asked for, or a wave
// kernel forced other than the LDS one) are turned into id lists (ascending,
// so the outputs keep the mask order) and counted by wave_submit.
int wave_submit_masks(tfim_ctx *c, const uint64_t *pmask, const uint64_t *cmask, int32_t n_rows, int32_t path,
                      int64_t *ticket) {
    if (!c || !pmask || !cmask || !ticket || n_rows < 0) return fail(c, TFIM_E_INVALID, "count_masks: null argument");
    *ticket = -1;
    if (!c->d_bm) return fail(c, TFIM_E_STATE, "count_masks: no bitmaps loaded");
    if (c->n_items > kMaskMaxItems)
        return fail(c, TFIM_E_RANGE, "count_masks: %d dense ids loaded, masks cover at most %d", c->n_items,
                    kMaskMaxItems);
    if (path != TFIM_PATH_AUTO && path != TFIM_PATH_PACKED && path != TFIM_PATH_MFMA_I8)
        return fail(c, TFIM_E_INVALID, "count_masks: unknown path %d", path);
    const uint64_t ids_ok = c->n_items >= 64 ? ~0ull : (1ull << c->n_items) - 1;
    const uint64_t cols_ok = c->n_cols >= 64 ? ~0ull : (1ull << c->n_cols) - 1;
    int64_t n_out = 0, n_pref = 0;
    for (int32_t r = 0; r < n_rows; r++) {
        if (!pmask[r] || (pmask[r] & ~ids_ok))
            return fail(c, TFIM_E_INVALID, "count_masks: row %d prefix mask is empty or names an id >= %d", r,
                        c->n_items);
        if (cmask[r] & ~cols_ok)
            return fail(c, TFIM_E_INVALID, "count_masks: row %d column mask names an id >= %d", r, c->n_cols);
        n_out += __builtin_popcountll(cmask[r]);
        n_pref += __builtin_popcountll(pmask[r]);
    }
    if (n_out > INT32_MAX) return fail(c, TFIM_E_RANGE, "count_masks: output too large");
    if (n_rows == 0 || n_out == 0) return TFIM_OK;
    const bool mask_kernel = n_out <= kLdsMaxOut && path != TFIM_PATH_MFMA_I8 &&
                             (c->wave_kernel == 0 || c->wave_kernel == 2 || c->wave_kernel == 3);
    if (!mask_kernel) {
        std::vector<int32_t> ids, ioff(1, 0), cols, coff(1, 0);
        ids.reserve((size_t)n_pref);
        cols.reserve((size_t)n_out);
        for (int32_t r = 0; r < n_rows; r++) {
            for (uint64_t b = pmask[r]; b; b &= b - 1) ids.push_back(__builtin_ctzll(b));
            for (uint64_t b = cmask[r]; b; b &= b - 1) cols.push_back(__builtin_ctzll(b));
            ioff.push_back((int32_t)ids.size());
            coff.push_back((int32_t)cols.size());
        }
        return wave_submit(c, ids.data(), ioff.data(), n_rows, cols.data(), coff.data(), nullptr, 0, nullptr, 0, path,
                           ticket);
    }
    std::lock_guard<std::mutex> wave_lock(c->wave_mu);
    tfim_ctx::Slot &sl = c->slots[c->next_ticket % tfim_ctx::kWaveSlots];
    if (sl.busy)
        return fail(c, TFIM_E_STATE, "count_masks_submit: %d waves in flight; wait on a ticket first",
                    tfim_ctx::kWaveSlots);
    if (int r = set_device(c)) return r;
    // arm: columns per tile and staging (two workgroups per CU only where a
    // chunk of every id and the accumulators fit half the LDS)
    bool allowed[kMaskArms];
    const bool narrow_fits = (size_t)c->n_items * 1024 + (size_t)(n_out + 3) / 4 * 16 + 16 <= 80 * 1024;
    for (int a = 0; a < kMaskArms; a++) allowed[a] = kMaskArmWide[a] || narrow_fits;
    int arm;
    if (c->wave_kernel == 2 || c->wave_kernel == 3) {
        arm = c->wave_kernel == 3 || !narrow_fits ? 0 : 3;
        if (c->tile_cols) arm += c->tile_cols >= 64 ? 0 : c->tile_cols >= 32 ? 1 : 2;
    } else {
        arm = c->mask_tune.pick(allowed);
    }
    const int tile_cols = c->tile_cols ? c->tile_cols : kMaskArmTile[arm];
    // tiles of <= tile_cols columns (the lowest ids first), outputs CSR by row
    std::vector<MaskTile> &tiles = c->mask_tiles;
    tiles.clear();
    int32_t ob = 0, max_cost = 0;
    for (int32_t r = 0; r < n_rows; r++) {
        const uint64_t pm = pmask[r];
        const int32_t k = __builtin_popcountll(pm);
        uint64_t cm = cmask[r];
        while (cm) {
            uint64_t part = cm;
            int32_t n = __builtin_popcountll(cm);
            if (n > tile_cols) {
                part = 0;
                for (int i = 0; i < tile_cols; i++) {
                    const uint64_t low = cm & (~cm + 1);
                    part |= low;
                    cm ^= low;
                }
                n = tile_cols;
            } else {
                cm = 0;
            }
            tiles.push_back({pm, part, ob, k, n, 0});
            ob += n;
            max_cost = std::max(max_cost, 2 * n + k);
        }
    }
    const int64_t n_tiles = (int64_t)tiles.size();
    const int64_t desc_bytes = n_tiles * (int64_t)sizeof(MaskTile);
    if (int r = ensure_desc(c, sl, desc_bytes)) return r;
    // outputs, then (16-byte aligned) the kernel's two wall-clock stamps
    const int64_t stamp_at = (n_out + 3) / 4 * 4, n_block = stamp_at + 4;
    if (int r = ensure_out(c, sl, n_block)) return r;
    // by falling cost 2 * columns + prefix ids, stable (see the list path)
    MaskTile *ht = reinterpret_cast<MaskTile *>(sl.h_desc);
    std::vector<int32_t> &start = c->tile_bucket;
    start.assign((size_t)max_cost + 2, 0);
    for (const MaskTile &t : tiles) start[(size_t)(max_cost - (2 * t.nc + t.k)) + 1]++;
    for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
    for (const MaskTile &t : tiles) ht[start[(size_t)(max_cost - (2 * t.nc + t.k))]++] = t;
    WaveMaskDesc d;
    d.tiles = reinterpret_cast<const MaskTile *>(sl.d_desc);
    d.out = sl.d_out;
    d.stamp = reinterpret_cast<unsigned long long *>(sl.d_out + stamp_at);
    d.n_items = c->n_items;
    d.n_tiles = (int32_t)n_tiles;
    d.n_out = (int32_t)n_out;
    KernelLaunch K;
    if (!plan_count_wave_mask(c->d_bm, c->pitch, d, c->n_cu, kMaskArmWide[arm], K))
        return fail(c, TFIM_E_RANGE, "count_masks: the mask kernel rejected a wave of %lld outputs over %d ids",
                    (long long)n_out, c->n_items);
    KernelLaunch *ks[1] = {&K};
    if (int r = issue_wave(c, sl, desc_bytes, nullptr, 0, n_block, n_block, ks, 1, false)) return r;
    sl.busy = true;
    sl.stamp_at = stamp_at;
    sl.ticket = c->next_ticket++;
    sl.n_out = n_out;
    sl.n_rows = n_rows;
    sl.tuned = -1;
    sl.mf_arm = -1;
    sl.sk_arm = -1;
    sl.mask_arm = (c->wave_kernel == 2 || c->wave_kernel == 3) ? -1 : arm;
    sl.mask_work = (double)(n_pref + n_out) * (double)c->pitch;
    c->wave_lds_launches++;
    c->wave_mask_launches++;
    *ticket = sl.ticket;
    return TFIM_OK;
}

}  // namespace

int tfim_count_nodes_submit(tfim_ctx *c, const int32_t *item_ids, const int32_t *item_off, int32_t n_rows,
                            const int32_t *col_ids, const int32_t *col_off, int32_t path, int64_t *ticket) {
    return wave_submit(c, item_ids, item_off, n_rows, col_ids, col_off, nullptr, 0, nullptr, 0, path, ticket);
}

int tfim_count_nodes_ex_submit(tfim_ctx *c, const int32_t *item_ids, const int32_t *item_off, int32_t n_rows,
                               const int32_t *col_ids, const int32_t *col_off, const int32_t *compl_rows,
                               int32_t n_compl, const uint64_t *excl, int32_t min_count, int32_t path,
                               int64_t *ticket) {
    if (!col_off || n_rows < 0) return fail(c, TFIM_E_INVALID, "count_nodes_ex: col_off is required");
    static const int32_t no_cols = 0;
    if (!col_ids) {
        if (col_off[n_rows] != 0) return fail(c, TFIM_E_INVALID, "count_nodes_ex: explicit columns without col_ids");
        col_ids = &no_cols;   // rows are CSR either way
    }
    return wave_submit(c, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl, excl, min_count, path,
                       ticket);
}

int tfim_count_nodes_ex(tfim_ctx *c, const int32_t *item_ids, const int32_t *item_off, int32_t n_rows,
                        const int32_t *col_ids, const int32_t *col_off, const int32_t *compl_rows, int32_t n_compl,
                        const uint64_t *excl, int32_t min_count, int32_t *out, int32_t path) {
    if (!out) return fail(c, TFIM_E_INVALID, "count_nodes_ex: null output");
    int64_t ticket = -1;
    if (int r = tfim_count_nodes_ex_submit(c, item_ids, item_off, n_rows, col_ids, col_off, compl_rows, n_compl,
                                           excl, min_count, path, &ticket))
        return r;
    return tfim_count_nodes_wait(c, ticket, out);
}

int tfim_count_masks_submit(tfim_ctx *c, const uint64_t *prefix_mask, const uint64_t *col_mask, int32_t n_rows,
                            int32_t path, int64_t *ticket) {
    return wave_submit_masks(c, prefix_mask, col_mask, n_rows, path, ticket);
}

int tfim_count_masks(tfim_ctx *c, const uint64_t *prefix_mask, const uint64_t *col_mask, int32_t n_rows,
                     int32_t *out, int32_t path) {
    if (!out) return fail(c, TFIM_E_INVALID, "count_masks: null output");
    int64_t ticket = -1;
    if (int r = tfim_count_masks_submit(c, prefix_mask, col_mask, n_rows, path, &ticket)) return r;
    return tfim_count_nodes_wait(c, ticket, out);
}

int tfim_count_nodes_wait(tfim_ctx *c, int64_t ticket, int32_t *out) {
    if (!c) return TFIM_E_INVALID;
    if (ticket < 0) return TFIM_OK;   // nothing was submitted (no rows or no columns)
    tfim_ctx::Slot &sl = c->slots[ticket % tfim_ctx::kWaveSlots];
    {
        std::lock_guard<std::mutex> g(c->wave_mu);
        if (!sl.busy || sl.ticket != ticket)
            return fail(c, TFIM_E_STATE, "count_nodes_wait: unknown ticket %lld", (long long)ticket);
    }
    if (!out) return fail(c, TFIM_E_INVALID, "count_nodes_wait: null output");
    if (int r = set_device(c)) return r;
    // the slot stays busy (a submit cannot take it) until its counts are copied
    struct Release {
        tfim_ctx *c;
        tfim_ctx::Slot &sl;
        ~Release() {
            std::lock_guard<std::mutex> g(c->wave_mu);
            sl.busy = false;
        }
    } release{c, sl};
    // poll rather than block in the runtime, so that a submit from another
    // thread is not held up behind this wait
    for (hipError_t q; (q = hipEventQuery(sl.done)) != hipSuccess;) {
        if (q != hipErrorNotReady) HIPCHK(c, q);
        std::this_thread::yield();
    }
    memcpy(out, sl.h_out, (size_t)sl.n_out * 4);
    std::lock_guard<std::mutex> g(c->wave_mu);
    float ms = 0;
    if (sl.stamp_at >= 0) {
        // first workgroup start (kept as its complement under a max) to
        // last workgroup end, in wall-clock ticks
        uint64_t st[2];
        memcpy(st, sl.h_out + sl.stamp_at, sizeof st);
        const uint64_t t0 = ~st[0], t1 = st[1];
        ms = st[0] && t1 > t0 ? (float)((double)(t1 - t0) / c->wall_khz) : 0.0f;
    } else {
        hipEventElapsedTime(&ms, sl.ev0, sl.ev1);
    }
    c->wave_ms += ms;
    c->wave_launches++;
    c->wave_rows += sl.n_rows;
    if (sl.contended || sl.train_seq != c->train_seq) {
        c->wave_contended++;   // timed while the model trained: not a sample of its kernel's cost
        return TFIM_OK;
    }
    if (sl.sk_arm >= 0 && sl.sk_work > 0) (sl.sk_fused ? c->fk : c->sk).record(sl.sk_arm, (double)ms / sl.sk_work);
    if (sl.mask_arm >= 0 && sl.mask_work > 0) c->mask_tune.record(sl.mask_arm, (double)ms / sl.mask_work);
    if (sl.mf_arm >= 0 && sl.mf_work > 0) {
        const double per = (double)ms / sl.mf_work;
        double &cst = c->mf_cost[sl.mf_arm];
        cst = c->mf_samples[sl.mf_arm] == 0 ? per : 0.8 * cst + 0.2 * per;
        c->mf_samples[sl.mf_arm]++;
    }
    if (sl.tuned >= 0 && sl.work > 0) {
        const double per = (double)ms / sl.work;
        double &cst = c->tune_cost[sl.tuned];
        cst = c->tune_samples[sl.tuned] == 0 ? per : 0.8 * cst + 0.2 * per;
        c->tune_samples[sl.tuned]++;
    }
    return TFIM_OK;
}

int tfim_count_nodes(tfim_ctx *c, const int32_t *item_ids, const int32_t *item_off, int32_t n_rows,
                     const int32_t *col_ids, const int32_t *col_off, int32_t *out, int32_t path) {
    if (!out) return fail(c, TFIM_E_INVALID, "count_nodes: null output");
    int64_t ticket = -1;
    if (int r = tfim_count_nodes_submit(c, item_ids, item_off, n_rows, col_ids, col_off, path, &ticket)) return r;
    return tfim_count_nodes_wait(c, ticket, out);
}

int tfim_config(tfim_ctx *c, const char *key, int64_t value) {
    if (!c || !key) return TFIM_E_INVALID;
    const std::string k(key);
    if (k == "wave_kernel" && value >= 0 && value <= 5) {
        c->wave_kernel = (int)value;
        return TFIM_OK;
    }
    if (k == "node_shape" && (value == 0 || ((value / 100 == 1 || value / 100 == 2 || value / 100 == 4) &&
                                             (value % 100 == 8 || value % 100 == 16 || value % 100 == 32 ||
                                              value % 100 == 64)))) {
        c->node_shape = (int)value;
        return TFIM_OK;
    }
    if (k == "node_spread" && (value == 0 || value == 1 || value == 8)) {
        c->node_spread = (int)value;
        return TFIM_OK;
    }
    if (k == "mfma_tile" && (value == 0 || value == 128 || value == 256)) {
        c->mfma_tile = (int)value;
        return TFIM_OK;
    }
    if (k == "mfma_format" && (value == 0 || value == 1)) {
        c->mfma_int8 = value == 1;
        return TFIM_OK;
    }
    if (k == "mfma_split" && (value == 0 || value == 1)) {
        c->mfma_split = (int)value;
        return TFIM_OK;
    }
    if (k == "graphs" && (value == 0 || value == 1)) {
        c->use_graphs = (int)value;
        return TFIM_OK;
    }
    if (k == "tile_cols" && (value == 0 || (value >= 8 && value <= 64))) {
        c->tile_cols = (int)value;
        return TFIM_OK;
    }
    return fail(c, TFIM_E_INVALID, "config: unknown key or value %s=%lld", key, (long long)value);
}

int tfim_wave_stats(tfim_ctx *c, int64_t *launches, int64_t *lds_launches, int64_t *rows, double *ms) {
    if (!c) return TFIM_E_INVALID;
    std::lock_guard<std::mutex> g(c->wave_mu);
    if (lds_launches) *lds_launches = c->wave_lds_launches;
    if (launches) *launches = c->wave_launches;
    if (rows) *rows = c->wave_rows;
    if (ms) *ms = c->wave_ms;
    return TFIM_OK;
}

int tfim_wave_path_stats(tfim_ctx *c, int64_t *l2, int64_t *lds, int64_t *mfma, int64_t *fused) {
    if (!c) return TFIM_E_INVALID;
    std::lock_guard<std::mutex> g(c->wave_mu);
    if (l2) *l2 = c->wave_launches - c->wave_lds_launches - c->wave_mfma_launches - c->wave_fused_launches;
    if (lds) *lds = c->wave_lds_launches;
    if (mfma) *mfma = c->wave_mfma_launches;
    if (fused) *fused = c->wave_fused_launches;
    return TFIM_OK;
}

int tfim_wave_graph_stats(tfim_ctx *c, int64_t *graphs, int64_t *replays) {
    if (!c) return TFIM_E_INVALID;
    std::lock_guard<std::mutex> g(c->wave_mu);
    if (graphs) *graphs = c->wave_graphs;
    if (replays) *replays = c->wave_graph_replays;
    return TFIM_OK;
}

int tfim_buf_alloc(tfim_ctx *c, int64_t bytes, tfim_buf **out) {
    if (!c || !out || bytes <= 0) return fail(c, TFIM_E_INVALID, "buf_alloc: bad argument");
    if (int r = set_device(c)) return r;
    tfim_buf *b = new tfim_buf();
    if (hipMalloc(&b->ptr, (size_t)bytes) != hipSuccess) {
        delete b;
        return fail(c, TFIM_E_NOMEM, "buf_alloc: %lld bytes failed", (long long)bytes);
    }
    b->bytes = bytes;
    *out = b;
    return TFIM_OK;
}

void tfim_buf_free(tfim_ctx *c, tfim_buf *b) {
    if (!b) return;
    if (c) {
        hipSetDevice(c->device);
        hipStreamSynchronize(c->stream);
    }
    hipFree(b->ptr);
    delete b;
}

int tfim_buf_write(tfim_ctx *c, tfim_buf *b, int64_t off, const void *src, int64_t bytes) {
    if (!c || !b || !src || off < 0 || bytes < 0 || off + bytes > b->bytes)
        return fail(c, TFIM_E_INVALID, "buf_write: range");
    if (int r = set_device(c)) return r;
    HIPCHK(c, hipMemcpyAsync((char *)b->ptr + off, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TFIM_OK;
}

int tfim_buf_read(tfim_ctx *c, const tfim_buf *b, int64_t off, void *dst, int64_t bytes) {
    if (!c || !b || !dst || off < 0 || bytes < 0 || off + bytes > b->bytes)
        return fail(c, TFIM_E_INVALID, "buf_read: range");
    if (int r = set_device(c)) return r;
    HIPCHK(c, hipMemcpyAsync(dst, (const char *)b->ptr + off, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TFIM_OK;
}

int tfim_fill_bernoulli(tfim_ctx *c, tfim_buf *dst, int64_t rows, int64_t n_bits, int64_t pitch,
                        uint64_t seed, uint32_t matrix, uint32_t thr32) {
    if (!c || !dst || rows < 0 || n_bits < 0 || pitch < (n_bits + 63) / 64)
        return fail(c, TFIM_E_INVALID, "fill_bernoulli: bad argument");
    if (rows * pitch * 8 > dst->bytes) return fail(c, TFIM_E_INVALID, "fill_bernoulli: buffer too small");
    if (int r = set_device(c)) return r;
    launch_fill_bernoulli((uint64_t *)dst->ptr, rows, n_bits, pitch, seed, matrix, thr32, c->stream);
    HIPCHK(c, hipGetLastError());
    return TFIM_OK;
}

static int popmat_check(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b,
                        int64_t d, int64_t pitch, int32_t path) {
    if (!c || !A || !B || !C || b < 1 || d < 1 || pitch < 1) return fail(c, TFIM_E_INVALID, "popcount_matrix: bad argument");
    if (pitch % kPitchWords) return fail(c, TFIM_E_INVALID, "popcount_matrix: pitch %lld not a multiple of 32 words", (long long)pitch);
    if (b * pitch * 8 > A->bytes || d * pitch * 8 > B->bytes || b * d * 4 > C->bytes)
        return fail(c, TFIM_E_INVALID, "popcount_matrix: buffer too small");
    if (pitch * 64 > (int64_t)INT32_MAX) return fail(c, TFIM_E_RANGE, "popcount_matrix: rows too long for int32 counts");
    if (path != TFIM_PATH_AUTO && path != TFIM_PATH_PACKED && path != TFIM_PATH_MFMA_I8 && path != TFIM_PATH_MFMA_FP4)
        return fail(c, TFIM_E_INVALID, "popcount_matrix: unknown path %d", path);
    return TFIM_OK;
}

static hipError_t popmat_launch(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b,
                                int64_t d, int64_t pitch, int32_t path, int32_t split_k, int tile) {
    if (split_k <= 0) split_k = 1;
    if (split_k > 1) launch_zero_i32((int32_t *)C->ptr, b * d, c->stream);
    const uint64_t *a = (const uint64_t *)A->ptr, *bb = (const uint64_t *)B->ptr;
    int32_t *cc = (int32_t *)C->ptr;
    if (path == TFIM_PATH_PACKED) {
        launch_popmat_packed(a, bb, cc, b, d, pitch, split_k, c->stream);
    } else if (path == TFIM_PATH_MFMA_FP4) {
        if (hipError_t e = launch_popmat_mfma_rp(a, bb, cc, b, d, pitch, split_k, c->stream, true)) return e;
    } else {
        if (tile == 256) {
            if (hipError_t e = launch_popmat_mfma_rp(a, bb, cc, b, d, pitch, split_k, c->stream)) return e;
        } else {
            launch_popmat_mfma(a, bb, cc, b, d, pitch, split_k, c->stream);
        }
    }
    return hipGetLastError();
}

// Tile and split-k of a matrix count.  split_k >= 1 and a configured
// mfma_tile are taken as given; split_k == 0 asks the tuner, which times every
// candidate once on the caller's buffers the first time a (path, b, d, pitch)
// shape is seen (blocking) and keeps the fastest for that shape.
static int popmat_plan(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b, int64_t d,
                       int64_t pitch, int32_t path, int32_t split_k, int *tile, int32_t *split) {
    *tile = c->mfma_tile ? c->mfma_tile : 256;
    *split = split_k;
    if (split_k >= 1) return TFIM_OK;
    const std::string key = std::to_string(path) + ":" + std::to_string(b) + ":" + std::to_string(d) + ":" +
                            std::to_string(pitch) + ":" + std::to_string(c->mfma_tile);
    auto it = c->popmat_tuned.find(key);
    if (it != c->popmat_tuned.end()) {
        *tile = it->second.first;
        *split = it->second.second;
        return TFIM_OK;
    }
    std::vector<int> tiles;
    if (path == TFIM_PATH_PACKED || path == TFIM_PATH_MFMA_FP4) tiles = {0};
    else if (c->mfma_tile) tiles = {c->mfma_tile};
    else tiles = {256, 128};
    // splits 1, 2, 4, ... while a split keeps at least 64 words: shapes of few
    // tiles (128 rows x 256 items) need far more than 16 to fill the CUs
    std::vector<int32_t> splits;
    for (int32_t sk = 1; sk <= 512 && (sk == 1 || pitch / sk >= 64); sk *= 2) splits.push_back(sk);
    double best = 1e300;
    for (int t : tiles)
        for (int32_t sk : splits) {
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            HIPCHK(c, popmat_launch(c, A, B, C, b, d, pitch, path, sk, t));
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            if (ms < best) {
                best = ms;
                *tile = t;
                *split = sk;
            }
        }
    c->popmat_tuned[key] = {*tile, *split};
    return TFIM_OK;
}

int tfim_popcount_matrix(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b,
                         int64_t d, int64_t pitch, int32_t path, int32_t split_k) {
    if (int r = popmat_check(c, A, B, C, b, d, pitch, path)) return r;
    if (int r = set_device(c)) return r;
    int tile;
    int32_t split;
    if (int r = popmat_plan(c, A, B, C, b, d, pitch, path, split_k, &tile, &split)) return r;
    HIPCHK(c, popmat_launch(c, A, B, C, b, d, pitch, path, split, tile));
    return TFIM_OK;
}

int tfim_popcount_matrix_plan(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b,
                              int64_t d, int64_t pitch, int32_t path, int32_t *tile, int32_t *split_k) {
    if (!tile || !split_k) return fail(c, TFIM_E_INVALID, "popcount_matrix_plan: null output");
    if (int r = popmat_check(c, A, B, C, b, d, pitch, path)) return r;
    if (int r = set_device(c)) return r;
    int t;
    int32_t sk;
    if (int r = popmat_plan(c, A, B, C, b, d, pitch, path, 0, &t, &sk)) return r;
    *tile = path == TFIM_PATH_PACKED ? 64 : path == TFIM_PATH_MFMA_FP4 ? 256 : t;
    *split_k = sk;
    return TFIM_OK;
}

int tfim_time_popcount_matrix(tfim_ctx *c, const tfim_buf *A, const tfim_buf *B, tfim_buf *C, int64_t b,
                              int64_t d, int64_t pitch, int32_t path, int32_t split_k, int32_t warmup,
                              int32_t reps, double *avg_ms) {
    if (int r = popmat_check(c, A, B, C, b, d, pitch, path)) return r;
    if (reps < 1 || !avg_ms) return fail(c, TFIM_E_INVALID, "time_popcount_matrix: reps < 1");
    if (int r = set_device(c)) return r;
    int tile;
    int32_t split;
    if (int r = popmat_plan(c, A, B, C, b, d, pitch, path, split_k, &tile, &split)) return r;
    for (int i = 0; i < warmup; i++) HIPCHK(c, popmat_launch(c, A, B, C, b, d, pitch, path, split, tile));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < reps; i++) HIPCHK(c, popmat_launch(c, A, B, C, b, d, pitch, path, split, tile));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = (double)ms / reps;
    return TFIM_OK;
}

namespace {
void prep_release(tfim_ctx *c) {
    hipFree(c->p_ids);
    hipFree(c->p_off);
    c->p_ids = nullptr;
    c->p_off = nullptr;
    c->p_n_ids = 0;
    c->p_n_rows = -1;
    c->p_id_max = -1;
}
}  // namespace

namespace {
// Host -> device copy of a large pageable block on the ctx stream.  HIP's
// pageable copy stages the data through pinned memory from one thread; here
// up to 8 threads each take a contiguous share and move it through two
// pinned buffers of their own, so host copies run in parallel and overlap
// the DMA of the previous piece.  Returns after the copy has completed.
// Blocks below 64 MB take hipMemcpyAsync directly.
int upload_staged(tfim_ctx *c, void *dst, const void *src, size_t bytes) {
    constexpr size_t kPiece = (size_t)8 << 20;
    if (bytes < ((size_t)64 << 20)) {
        HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TFIM_OK;
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::min<unsigned>(8u, hw);
    std::vector<hipError_t> err((size_t)T, hipSuccess);
    auto worker = [&](int t) {
        const size_t share = (bytes / T + kPiece - 1) / kPiece * kPiece;
        const size_t lo = std::min(bytes, (size_t)t * share), hi = std::min(bytes, lo + share);
        char *stage[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        hipError_t e = hipSetDevice(c->device);
        for (int b = 0; b < 2 && e == hipSuccess; b++) {
            e = hipHostMalloc((void **)&stage[b], kPiece, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
        }
        size_t k = 0;
        for (size_t off = lo; off < hi && e == hipSuccess; off += kPiece, k++) {
            const int b = (int)(k & 1);
            const size_t n = std::min(kPiece, hi - off);
            if (k >= 2) e = hipEventSynchronize(ev[b]);   // the piece that last used this buffer has landed
            if (e != hipSuccess) break;
            memcpy(stage[b], static_cast<const char *>(src) + off, n);
            e = hipMemcpyAsync(static_cast<char *>(dst) + off, stage[b], n, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[b], c->stream);
        }
        for (int b = 0; b < 2; b++) {
            if (ev[b]) {
                const hipError_t w = hipEventSynchronize(ev[b]);
                if (e == hipSuccess && k > (size_t)b) e = w;
                hipEventDestroy(ev[b]);
            }
            if (stage[b]) hipHostFree(stage[b]);
        }
        err[(size_t)t] = e;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; t++) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    for (hipError_t e : err)
        if (e != hipSuccess) return fail(c, TFIM_E_HIP, "staged upload failed: %s", hipGetErrorString(e));
    return TFIM_OK;
}
}  // namespace

int tfim_prep_begin(tfim_ctx *c, const int32_t *ids, int64_t n_ids, const int64_t *row_off, int64_t n_rows,
                    int32_t id_max, int32_t *freq) {
    if (!c || !row_off || !freq || n_ids < 0 || n_rows < 0 || id_max < 0 || (n_ids > 0 && !ids))
        return fail(c, TFIM_E_INVALID, "prep_begin: bad argument");
    if (row_off[0] != 0 || row_off[n_rows] != n_ids)
        return fail(c, TFIM_E_INVALID, "prep_begin: row offsets must run from 0 to n_ids");
    if (n_rows > (int64_t)INT32_MAX) return fail(c, TFIM_E_RANGE, "prep_begin: %lld rows exceed int32", (long long)n_rows);
    // the histogram holds id_max + 2 int32 entries (counts and a flag)
    if (id_max > INT32_MAX - 2) return fail(c, TFIM_E_RANGE, "prep_begin: id_max %d above %d", id_max, INT32_MAX - 2);
    if (!c->owns_bm) return fail(c, TFIM_E_STATE, "prep_begin: context shares its parent's bitmaps");
    if (int r = set_device(c)) return r;
    prep_release(c);
    int32_t *d_freq = nullptr, *d_bad = nullptr;
    if (hipMalloc(&c->p_ids, (size_t)std::max<int64_t>(1, n_ids) * 4) != hipSuccess ||
        hipMalloc(&c->p_off, (size_t)(n_rows + 1) * 8) != hipSuccess ||
        hipMalloc(&d_freq, (size_t)(id_max + 1) * 4 + 4) != hipSuccess) {
        hipFree(d_freq);
        prep_release(c);
        return fail(c, TFIM_E_NOMEM, "prep_begin: device allocation for %lld ids failed", (long long)n_ids);
    }
    d_bad = d_freq + (id_max + 1);
    int rc = TFIM_OK;
    int32_t bad = 0;
    do {
        if (n_ids > 0 && upload_staged(c, c->p_ids, ids, (size_t)n_ids * 4) != TFIM_OK) { rc = TFIM_E_HIP; break; }
        if (upload_staged(c, c->p_off, row_off, (size_t)(n_rows + 1) * 8) != TFIM_OK) { rc = TFIM_E_HIP; break; }
        if (hipMemsetAsync(d_freq, 0, (size_t)(id_max + 2) * 4, c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
        launch_prep_hist(c->p_ids, n_ids, id_max + 1, d_freq, d_bad, c->n_cu, c->stream);
        if (hipGetLastError() != hipSuccess) { rc = TFIM_E_HIP; break; }
        if (hipMemcpyAsync(freq, d_freq, (size_t)(id_max + 1) * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
    } while (0);
    hipFree(d_freq);
    if (rc) {
        prep_release(c);
        return fail(c, rc, "prep_begin: HIP error");
    }
    if (bad) {
        prep_release(c);
        return fail(c, TFIM_E_INVALID, "prep_begin: an id lies outside [0, %d]", id_max);
    }
    c->p_n_ids = n_ids;
    c->p_n_rows = n_rows;
    c->p_id_max = id_max;
    return TFIM_OK;
}

int tfim_prep_finish(tfim_ctx *c, const int32_t *dense_of, int32_t n_items, int32_t n_cols, int64_t *rows_reduced,
                     int64_t *len_min, int64_t *len_max, int64_t *len_sum, int32_t *duplicates) {
    if (!c || !dense_of || !rows_reduced || n_items < 0 || n_cols < 0 || n_cols > n_items)
        return fail(c, TFIM_E_INVALID, "prep_finish: bad argument");
    if (c->p_n_rows < 0) return fail(c, TFIM_E_STATE, "prep_finish: no tfim_prep_begin before");
    for (int32_t i = 0; i <= c->p_id_max; i++)
        if (dense_of[i] < -1 || dense_of[i] >= n_items) {
            prep_release(c);
            return fail(c, TFIM_E_INVALID, "prep_finish: dense_of[%d] = %d outside [-1, %d)", i, dense_of[i], n_items);
        }
    if (int r = set_device(c)) return r;
    const int64_t n_rows = c->p_n_rows, n_blk = prep_blocks(n_rows);
    int32_t *d_map = nullptr, *d_cnt = nullptr, *d_blk = nullptr, *d_kept = nullptr, *d_dup = nullptr;
    unsigned long long *d_stats = nullptr;
    int64_t *d_total = nullptr;
    int rc = TFIM_OK;
    int64_t n_red = 0;
    unsigned long long st[3] = {0, 0, 0};
    int32_t dup = 0;
    const unsigned long long init[3] = {0ull, ~0ull, 0ull};
    do {
        if (hipMalloc(&d_map, (size_t)(c->p_id_max + 1) * 4) != hipSuccess ||
            hipMalloc(&d_cnt, (size_t)std::max<int64_t>(1, n_rows) * 4) != hipSuccess ||
            hipMalloc(&d_blk, (size_t)std::max<int64_t>(1, n_blk) * 4) != hipSuccess ||
            hipMalloc(&d_kept, (size_t)std::max<int64_t>(1, n_rows) * 4) != hipSuccess ||
            hipMalloc(&d_stats, sizeof init + 8 + 4) != hipSuccess) { rc = TFIM_E_NOMEM; break; }
        d_total = reinterpret_cast<int64_t *>(d_stats + 3);
        d_dup = reinterpret_cast<int32_t *>(d_total + 1);
        if (hipMemcpyAsync(d_map, dense_of, (size_t)(c->p_id_max + 1) * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(d_stats, init, sizeof init, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(d_total, 0, 12, c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
        if (n_rows > 0) {
            launch_prep_rows(c->p_ids, c->p_off, n_rows, d_map, d_cnt, d_blk, d_stats, c->stream);
            launch_prep_scan(d_blk, n_blk, d_total, c->stream);
            if (hipGetLastError() != hipSuccess) { rc = TFIM_E_HIP; break; }
        }
        if (hipMemcpyAsync(&n_red, d_total, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
        if (n_red > 0 && n_items > 0) {
            if (int r = load_common(c, n_items, n_red, n_cols)) { rc = r; break; }
            if (hipMemsetAsync(c->d_bm, 0, (size_t)n_items * c->pitch * 8, c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
            launch_prep_index(d_cnt, n_rows, d_blk, d_kept, c->stream);
            launch_prep_build(c->p_ids, c->p_off, d_kept, n_red, d_map, c->d_bm, c->pitch, d_dup, c->stream);
            if (hipGetLastError() != hipSuccess) { rc = TFIM_E_HIP; break; }
        }
        if (hipMemcpyAsync(st, d_stats, sizeof st, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(&dup, d_dup, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = TFIM_E_HIP; break; }
    } while (0);
    hipFree(d_map); hipFree(d_cnt); hipFree(d_blk); hipFree(d_kept); hipFree(d_stats);
    prep_release(c);
    if (rc == TFIM_E_NOMEM) return fail(c, rc, "prep_finish: device allocation failed");
    if (rc == TFIM_E_HIP) return fail(c, rc, "prep_finish: HIP error");
    if (rc) return rc;   // message set by load_common
    *rows_reduced = n_red;
    if (len_sum) *len_sum = (int64_t)st[0];
    if (len_min) *len_min = n_red > 0 ? (int64_t)st[1] : 0;
    if (len_max) *len_max = (int64_t)st[2];
    if (duplicates) *duplicates = dup;
    return TFIM_OK;
}

int tfim_sync(tfim_ctx *c) {
    if (!c) return TFIM_E_INVALID;
    if (int r = set_device(c)) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TFIM_OK;
}

}  // extern "C"
