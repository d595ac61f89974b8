// capi.cpp -- the C-ABI of libksvd_amd.so (declared in include/ksvd_amd.h):
// handle, configuration checks, scratch ownership and the host-side driver
// loop.  Replaces the reference's KSVD / OMP / DictionaryUpdate classes over
// the DeviceMemory arena (S/KSVD.cuh, S/OMP.cuh, S/DictionaryUpdate.cuh,
// S/DeviceMemory.cu) with caller-owned buffers and a per-handle stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ksvd_amd.h"
#include "common.hpp"

namespace ksvd {
hipError_t launch_fill_f32(float* p, long n, float v, hipStream_t st);
}

struct ksvd_handle {
    ksvd_config cfg;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<void*> allocs;
    ksvd::DuWork du{};
    bool du_ready = false;
    float* dtx = nullptr;      // K x N correlation scratch
    float* gram = nullptr;     // K x K
    float* errv = nullptr;     // N
    int32_t* counter = nullptr;
    int32_t* cidx = nullptr;   // compact scratch (dense entry points)
    float* cval = nullptr;
    int32_t* ccnt = nullptr;
    double* rmse_all = nullptr;
    int rmse_cap = 0;
    uint32_t* perm_all = nullptr;
    int perm_cap = 0;
    double* xnorm2 = nullptr;  // N: ||x_s||^2 in FP64 (error mode from signals)
    float* adj_z = nullptr;    // T x N (adjoint: z per support slot)
    unsigned long long* omp_stats = nullptr;  // [0] candidates, [1] fallbacks
    int last_variant = 0;
    bool last_noalloc = false;      // the last Batch-OMP call kept no statistics
    ksvd_tuning tune{};             // launch shapes (all zero: the built-in rules)
    std::string tune_log;           // the last autotune's candidates and their times
    float* tune_D = nullptr;        // autotuner scratch: normalised D, a working copy, coefficients
    float* tune_D2 = nullptr;
    float* tune_val = nullptr;
    hipStream_t gstream = nullptr;  // private non-blocking stream for graph capture/replay
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

// Device buffer of the allocation-free entry point
// (ksvd_batch_omp_from_signals_noalloc): DtX = D^T X of one chunk of signals.
// Every GPU the library runs on holds its own copy for the life of the
// process.  In host code `scratch` is only a symbol: its device address on the
// current GPU comes from hipGetSymbolAddress (scratch_ptr below).
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, resolved once per device
// (at ksvd_create, or at the first allocation-free call on a device).
hipError_t scratch_ptr(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

// ||x_s||^2 per signal in FP64 (one wave per signal).
__global__ __launch_bounds__(256) void sumsq_cols_kernel(const float* __restrict__ X, int n, int N,
                                                         double* __restrict__ out) {
    const long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= N) return;
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int t = lane; t < n; t += 64) {
        const double v = X[s * n + t];
        acc += v * v;
    }
    acc = ksvd::wave_sum(acc);
    if (lane == 0) out[s] = acc;
}

namespace {

int fail(ksvd_handle* h, int code, const char* fmt, ...) {
    if (h) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        h->err = buf;
    }
    return code;
}

int hip_check(ksvd_handle* h, hipError_t e, const char* what) {
    if (e == hipSuccess) return KSVD_OK;
    return fail(h, KSVD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
int alloc(ksvd_handle* h, T** p, size_t count, const char* what) {
    if (*p) return KSVD_OK;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 256);
    if (e != hipSuccess) return fail(h, KSVD_ERR_ALLOC, "hipMalloc(%s, %zu bytes): %s", what, count * sizeof(T),
                                     hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return KSVD_OK;
}

#define TRY(x)                        \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != KSVD_OK) return rc_; \
    } while (0)
#define HIPTRY(h, x, what) TRY(hip_check((h), (x), (what)))

int check_cfg(ksvd_handle* h, const ksvd_config* c) {
    if (c->n < 1 || c->K < 1 || c->N < 1 || c->T < 1)
        return fail(h, KSVD_ERR_INVALID, "n, K, N, T must be positive (n=%d K=%d N=%d T=%d)", c->n, c->K, c->N, c->T);
    if (c->T > 64) return fail(h, KSVD_ERR_LIMIT, "T=%d exceeds 64 (one support slot per lane)", c->T);
    if (c->T > c->K) return fail(h, KSVD_ERR_INVALID, "T=%d exceeds K=%d", c->T, c->K);
    if (c->K > 2048) return fail(h, KSVD_ERR_LIMIT, "K=%d exceeds 2048 (alpha held in VGPRs)", c->K);
    if (c->n > 4096) return fail(h, KSVD_ERR_LIMIT, "n=%d exceeds 4096", c->n);
    if ((long)c->N * c->T > 0x7fffffffL) return fail(h, KSVD_ERR_LIMIT, "N*T exceeds 2^31");
    if (c->gemm_form < 0 || c->gemm_form > 3) return fail(h, KSVD_ERR_INVALID, "gemm_form=%d", c->gemm_form);
    if (c->omp_variant != KSVD_OMP_AUTO && c->omp_variant != KSVD_OMP_WAVE && c->omp_variant != KSVD_OMP_QFORM)
        return fail(h, KSVD_ERR_INVALID, "omp_variant=%d (KSVD_OMP_AUTO, _WAVE or _QFORM)", c->omp_variant);
    if (c->omp_variant == KSVD_OMP_QFORM && ksvd::qform_shape(c->K, c->T).kpl == 0)
        return fail(h, KSVD_ERR_LIMIT, "register-basis Batch-OMP needs T <= 32 when K > 1024 (K=%d T=%d)", c->K, c->T);
    if (c->mu_thresh < 0) return fail(h, KSVD_ERR_INVALID, "mu_thresh must be >= 0");
    if (c->max_signals < 1) return fail(h, KSVD_ERR_INVALID, "max_signals must be >= 1");
    return KSVD_OK;
}

int ensure_du(ksvd_handle* h) {
    if (h->du_ready) return KSVD_OK;
    ksvd::DuWork& w = h->du;
    const ksvd_config& c = h->cfg;
    w.n = c.n; w.K = c.K; w.N = c.N; w.T = c.T;
    w.max_signals = c.max_signals;
    w.use_thresh = c.use_thresh;
    w.mu_thresh = c.mu_thresh;
    w.chunk = 256;
    w.nchunks = (c.N + w.chunk - 1) / w.chunk;
    w.rmse_parts = 2048;
    TRY(alloc(h, &w.count, c.K, "count"));
    TRY(alloc(h, &w.ptr, c.K + 1, "ptr"));
    TRY(alloc(h, &w.chunk_hist, (size_t)c.K * w.nchunks, "chunk_hist"));
    TRY(alloc(h, &w.list_pos, (size_t)c.N * c.T, "list_pos"));
    TRY(alloc(h, &w.list_rec, (size_t)c.N * c.T, "list_rec"));
    TRY(alloc(h, &w.R, (size_t)c.n * c.N, "R"));
    w.blocks = h->tune.du_blocks > 0 ? h->tune.du_blocks : ksvd::kDuBlocks;
    TRY(alloc(h, &w.apart, (size_t)2 * ksvd::kDuBlocksMax * ((c.n + 3) & ~3), "apart"));
    TRY(alloc(h, &w.bpart, (size_t)8 * ksvd::kDuBlocksMax, "bpart"));
    TRY(alloc(h, &w.dnew, (size_t)2 * c.n + 32, "dnew"));
    TRY(alloc(h, &w.dacc, (size_t)3 * ksvd::du_set_words((c.n + 3) & ~3), "dacc"));
    TRY(alloc(h, &w.dstate, 2, "dstate"));
    TRY(alloc(h, &w.unused, c.N, "unused"));
    TRY(alloc(h, &w.us_counter, 1, "us_counter"));
    TRY(alloc(h, &w.bitmap, c.N, "bitmap"));
    TRY(alloc(h, &w.replaced, c.K, "replaced"));
    TRY(alloc(h, &w.perm_dev, c.K, "perm"));
    TRY(alloc(h, &w.usecount, c.K, "usecount"));
    TRY(alloc(h, &w.Gtmp, (size_t)c.K * c.K, "Gtmp"));
    TRY(alloc(h, &w.rmse_part, w.rmse_parts, "rmse_part"));
    TRY(alloc(h, &w.rmse_dev, 1, "rmse"));
    TRY(alloc(h, &w.clear_flag, c.K, "clear_flag"));
    TRY(alloc(h, &w.sweep, (size_t)2 * c.K, "sweep"));
    h->du_ready = true;
    HIPTRY(h, ksvd::launch_du_init(w, h->stream), "du_init");
    return KSVD_OK;
}

// SplitMix64 (Steele, Lea & Flood 2014) driving a Fisher-Yates shuffle
// (Knuth TAOCP 3.4.2 Algorithm P) of a running buffer, one shuffle per
// iteration -- the running-buffer semantics of
// Algorithms::RandomPermutation::generateRandPerm (S/RandPermutation.cu:35-42)
// with a seeded, documented generator instead of std::random_shuffle.
uint64_t splitmix64(uint64_t* st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void shuffle(uint32_t* buf, int K, uint64_t* st) {
    for (int i = K - 1; i >= 1; --i) {
        const int j = (int)(splitmix64(st) % (uint64_t)(i + 1));
        const uint32_t t = buf[i];
        buf[i] = buf[j];
        buf[j] = t;
    }
}

ksvd::OmpArgs omp_args(ksvd_handle* h, const float* DtX, const float* G) {
    ksvd::OmpArgs a{};
    a.DtX = DtX;
    a.G = G;
    a.K = h->cfg.K;
    a.N = h->cfg.N;
    a.T = h->cfg.T;
    a.pivot_fp64 = (float)h->cfg.pivot_fp64;
    a.q_kpl = h->tune.omp_kpl;
    a.q_waves = h->tune.omp_waves;
    return a;
}

bool use_qform(const ksvd_handle* h, bool eps = false) {
    const int v = h->cfg.omp_variant;
    // error mode, KSVD_OMP_AUTO: the tuning may pick the wave form's persistent launch
    if (eps && v == KSVD_OMP_AUTO && h->tune.omp_eps_launch == 1) return false;
    return (v == KSVD_OMP_QFORM || v == KSVD_OMP_AUTO) && ksvd::qform_shape(h->cfg.K, h->cfg.T).kpl != 0;
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Chooses the Batch-OMP kernel (register basis, else the G-streaming wave
// form) and arms its statistics counters.
int prepare_wave(ksvd_handle* h, const float* G, ksvd::OmpArgs& a, bool eps = false) {
    (void)G;
    h->last_noalloc = false;
    TRY(alloc(h, &h->omp_stats, 2, "omp stats"));
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, h->stream, h->omp_stats, 2);
    a.stats = h->omp_stats;
    h->last_variant = use_qform(h, eps) ? KSVD_OMP_QFORM : KSVD_OMP_WAVE;
    return KSVD_OK;
}

// ||x_s||^2 of every signal in FP64 into h->xnorm2.
int compute_xnorm2(ksvd_handle* h, const float* X) {
    const ksvd_config& c = h->cfg;
    TRY(alloc(h, &h->xnorm2, c.N, "xnorm2"));
    hipLaunchKernelGGL(sumsq_cols_kernel, dim3((unsigned)((c.N + 3) / 4)), dim3(256), 0, h->stream, X, c.n, c.N,
                       h->xnorm2);
    return hip_check(h, hipGetLastError(), "xnorm2");
}

// Error mode: the per-signal stop data, and the work-stealing counter the
// G-streaming wave form's persistent launch draws signals from (the
// register-basis form runs one workgroup per signal).
void set_eps(ksvd_handle* h, ksvd::OmpArgs& a, const double* xnorm2, const float* eps2) {
    a.xnorm2 = xnorm2;
    a.eps2 = eps2;
    a.work_counter = h->counter;
}

// Launches the wave-family kernel that prepare_wave() chose.
hipError_t launch_wave_family(const ksvd_handle* h, const ksvd::OmpArgs& a) {
    if (h->last_variant == KSVD_OMP_QFORM) return ksvd::launch_batch_omp_qform(a, h->stream);
    return ksvd::launch_batch_omp(a, h->stream);
}

// OMP::BatchOMP from D and X with the configured variant (compact output);
// error mode when eps2 != NULL, with ||x_s||^2 already in h->xnorm2.
int omp_from_signals(ksvd_handle* h, const float* D, const float* X, const float* G, const float* eps2,
                     int32_t* idx, float* val, int32_t* cnt) {
    const ksvd_config& c = h->cfg;
    TRY(alloc(h, &h->dtx, (size_t)c.K * c.N, "DtX"));
    HIPTRY(h, ksvd::launch_at_times_b(D, c.n, c.K, X, c.N, h->dtx, c.gemm_form, h->stream, h->tune.gemm_wpe), "correlate");
    ksvd::OmpArgs a = omp_args(h, h->dtx, G);
    a.idx = idx;
    a.val = val;
    a.cnt = cnt;
    TRY(prepare_wave(h, G, a, eps2 != nullptr));
    if (eps2) set_eps(h, a, h->xnorm2, eps2);
    return hip_check(h, launch_wave_family(h, a), "batch_omp (wave)");
}

// ---- autotuner ---------------------------------------------------------------
struct TuneKey {
    int n, K, N, T, form, eps;
    std::string arch;
    bool operator<(const TuneKey& o) const {
        return std::tie(n, K, N, T, form, eps, arch) < std::tie(o.n, o.K, o.N, o.T, o.form, o.eps, o.arch);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, ksvd_tuning> g_tune_cache;
bool g_tune_file_read = false;

// The cache file named by KSVD_TUNE_CACHE: one line per tuning,
//   ksvd_tune1 n K N T form eps arch kpl waves eps_launch wpe du_blocks omp gemm du domp dgemm ddu
void tune_file_read_locked() {
    if (g_tune_file_read) return;
    g_tune_file_read = true;
    const char* path = std::getenv("KSVD_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char arch[128];
    TuneKey k;
    ksvd_tuning t{};
    while (std::fscanf(f, " ksvd_tune1 %d %d %d %d %d %d %127s %d %d %d %d %d %lf %lf %lf %lf %lf %lf", &k.n, &k.K,
                       &k.N, &k.T, &k.form, &k.eps, arch, &t.omp_kpl, &t.omp_waves, &t.omp_eps_launch, &t.gemm_wpe,
                       &t.du_blocks, &t.omp_ms, &t.gemm_ms, &t.du_ms, &t.default_omp_ms, &t.default_gemm_ms,
                       &t.default_du_ms) == 18) {
        k.arch = arch;
        t.source = 2;
        g_tune_cache[k] = t;
    }
    std::fclose(f);
}

void tune_file_append_locked(const TuneKey& k, const ksvd_tuning& t) {
    const char* path = std::getenv("KSVD_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "ksvd_tune1 %d %d %d %d %d %d %s %d %d %d %d %d %.6f %.6f %.6f %.6f %.6f %.6f\n", k.n, k.K, k.N, k.T,
                 k.form, k.eps, k.arch.c_str(), t.omp_kpl, t.omp_waves, t.omp_eps_launch, t.gemm_wpe, t.du_blocks,
                 t.omp_ms, t.gemm_ms, t.du_ms, t.default_omp_ms, t.default_gemm_ms, t.default_du_ms);
    std::fclose(f);
}

// Shortest time of one call of `fn` (enqueued on h->stream) over `reps`
// calls after one warm-up call, in ms.
template <class F>
int time_ms(ksvd_handle* h, int reps, double* ms, F&& fn) {
    hipEvent_t e0, e1;
    HIPTRY(h, hipEventCreate(&e0), "event");
    HIPTRY(h, hipEventCreate(&e1), "event");
    int rc = fn();
    double best = 1e30;
    for (int r = 0; r < reps && rc == KSVD_OK; ++r) {
        rc = hip_check(h, hipEventRecord(e0, h->stream), "record");
        if (rc == KSVD_OK) rc = fn();
        if (rc == KSVD_OK) rc = hip_check(h, hipEventRecord(e1, h->stream), "record");
        if (rc == KSVD_OK) rc = hip_check(h, hipEventSynchronize(e1), "sync");
        float t = 0.f;
        if (rc == KSVD_OK) rc = hip_check(h, hipEventElapsedTime(&t, e0, e1), "elapsed");
        if (rc == KSVD_OK && t < best) best = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return rc;
}

int autotune(ksvd_handle* h, const float* D, const float* X, const float* eps2, bool force) {
    const ksvd_config& c = h->cfg;
    TuneKey key{c.n, c.K, c.N, c.T, c.gemm_form, eps2 ? 1 : 0, ""};
    {
        int dev = 0;
        hipDeviceProp_t prop;
        HIPTRY(h, hipGetDevice(&dev), "device");
        HIPTRY(h, hipGetDeviceProperties(&prop, dev), "device properties");
        key.arch = prop.gcnArchName;
        for (char& ch : key.arch)
            if (ch == ' ') ch = '_';
    }
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_file_read_locked();
        const auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            h->tune = it->second;
            h->tune.source = 3;
            h->tune_log = "cache hit";
            if (h->du_ready) h->du.blocks = h->tune.du_blocks > 0 ? h->tune.du_blocks : ksvd::kDuBlocks;
            return KSVD_OK;
        }
    }
    const long N = c.N;
    TRY(alloc(h, &h->tune_D, (size_t)c.n * c.K, "tune D"));
    TRY(alloc(h, &h->tune_D2, (size_t)c.n * c.K, "tune D2"));
    TRY(alloc(h, &h->tune_val, (size_t)N * c.T, "tune val"));
    TRY(alloc(h, &h->gram, (size_t)c.K * c.K, "G"));
    TRY(alloc(h, &h->dtx, (size_t)c.K * N, "DtX"));
    TRY(alloc(h, &h->cidx, (size_t)N * c.T, "cidx"));
    TRY(alloc(h, &h->cval, (size_t)N * c.T, "cval"));
    TRY(alloc(h, &h->ccnt, N, "ccnt"));
    TRY(ensure_du(h));
    if (eps2) TRY(compute_xnorm2(h, X));
    HIPTRY(h, ksvd::launch_normcols(D, h->tune_D, c.n, c.K, h->stream), "normcols");
    // about 100 ms of GEMM work first, so that the GPU's clocks have ramped up
    // before anything is timed
    {
        const long Nw = N < 131072 ? N : 131072;
        for (int r = 0; r < 60; ++r)
            HIPTRY(h, ksvd::launch_at_times_b(h->tune_D, c.n, c.K, X, Nw, h->dtx, c.gemm_form, h->stream), "warm-up");
        HIPTRY(h, hipStreamSynchronize(h->stream), "warm-up");
    }
    ksvd_tuning best{};
    std::string log;
    auto note = [&](const char* fmt, ...) {
        char buf[256];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        log += buf;
    };
    // 1. GEMM builds, on the correlation of up to 131072 signals (times scaled to N)
    const long Ng = N < 131072 ? N : 131072;
    double bestg = 0.0;
    for (int wpe : {3, 2}) {
        double ms = 0.0;
        TRY(time_ms(h, 5, &ms, [&] {
            return hip_check(h, ksvd::launch_at_times_b(h->tune_D, c.n, c.K, X, Ng, h->dtx, c.gemm_form, h->stream, wpe),
                             "correlate");
        }));
        ms *= (double)N / (double)Ng;
        note("gemm wpe=%d %.4f ms; ", wpe, ms);
        if (wpe == 3) best.default_gemm_ms = ms;
        if (wpe == 3 || ms < bestg) { bestg = ms; best.gemm_wpe = wpe; }
    }
    best.gemm_ms = bestg;
    HIPTRY(h, ksvd::launch_at_times_b(h->tune_D, c.n, c.K, h->tune_D, c.K, h->gram, c.gemm_form, h->stream,
                                      best.gemm_wpe), "gram");
    // 2. Batch-OMP shapes (and error-mode launches), on up to 65536 signals
    const int No = N < 65536 ? (int)N : 65536;
    HIPTRY(h, ksvd::launch_at_times_b(h->tune_D, c.n, c.K, X, No, h->dtx, c.gemm_form, h->stream, best.gemm_wpe),
           "correlate");
    auto omp_run = [&](int kpl, int waves, int eps_launch, int nsig) {
        ksvd::OmpArgs a = omp_args(h, h->dtx, h->gram);
        a.N = nsig;
        a.idx = h->cidx;
        a.val = h->cval;
        a.cnt = h->ccnt;
        a.q_kpl = kpl;
        a.q_waves = waves;
        const bool qf = use_qform(h) && eps_launch == 0;
        if (eps2) {
            a.xnorm2 = h->xnorm2;
            a.eps2 = eps2;
            a.work_counter = h->counter;
        }
        return hip_check(h, qf ? ksvd::launch_batch_omp_qform(a, h->stream) : ksvd::launch_batch_omp(a, h->stream),
                         "batch_omp (autotune)");
    };
    {
        ksvd::QformShape cand[8];
        const int nc = use_qform(h) ? ksvd::qform_candidates(c.K, c.T, cand, 8) : 0;
        const ksvd::QformShape def = ksvd::qform_shape(c.K, c.T);
        double besto = 0.0;
        bool first = true;
        // the register-basis shapes, then (error mode, automatic variant, or
        // no register basis) the G-streaming wave form with its persistent launch
        for (int i = 0; i <= nc; ++i) {
            const bool wave = i == nc;
            if (wave && nc > 0 && !(eps2 && c.omp_variant == KSVD_OMP_AUTO)) break;
            double ms = 0.0;
            TRY(time_ms(h, 3, &ms, [&] {
                return wave ? omp_run(0, 0, 1, No) : omp_run(cand[i].kpl, cand[i].waves, 0, No);
            }));
            ms *= (double)N / (double)No;
            if (wave)
                note("batch_omp wave form%s %.4f ms; ", eps2 ? " (persistent)" : "", ms);
            else
                note("batch_omp kpl=%d waves=%d %.4f ms; ", cand[i].kpl, cand[i].waves, ms);
            const bool is_def = wave ? nc == 0 : (cand[i].kpl == def.kpl && cand[i].waves == def.waves);
            if (is_def) best.default_omp_ms = ms;
            if (first || ms < besto) {
                besto = ms;
                best.omp_kpl = wave ? 0 : cand[i].kpl;
                best.omp_waves = wave ? 0 : cand[i].waves;
                best.omp_eps_launch = (wave && nc > 0) ? 1 : 0;
                first = false;
            }
        }
        best.omp_ms = besto;
    }
    // 3. dictionary-update grids, on a full sweep over all N signals coded
    //    with the chosen shapes (identity atom order)
    HIPTRY(h, ksvd::launch_at_times_b(h->tune_D, c.n, c.K, X, N, h->dtx, c.gemm_form, h->stream, best.gemm_wpe),
           "correlate");
    TRY(omp_run(best.omp_kpl, best.omp_waves, best.omp_eps_launch, (int)N));
    {
        std::vector<uint32_t> ident(c.K);
        for (int k = 0; k < c.K; ++k) ident[k] = (uint32_t)k;
        HIPTRY(h, hipMemcpyAsync(h->du.perm_dev, ident.data(), sizeof(uint32_t) * c.K, hipMemcpyHostToDevice, h->stream),
               "perm upload");
        HIPTRY(h, hipStreamSynchronize(h->stream), "perm upload sync");
    }
    double bestd = 0.0;
    // the built-in grid first (it is the reference time), then the others
    for (int blocks : {ksvd::kDuBlocks, 256, 512, 768, 1024}) {
        if (blocks == ksvd::kDuBlocks && bestd > 0.0) continue;
        h->du.blocks = blocks;
        double ms = 1e30;
        for (int rep = 0; rep < 2; ++rep) {  // the first sweep warms up
            HIPTRY(h, hipMemcpyAsync(h->tune_D2, h->tune_D, sizeof(float) * c.n * c.K, hipMemcpyDeviceToDevice,
                                     h->stream), "copy D");
            HIPTRY(h, hipMemcpyAsync(h->tune_val, h->cval, sizeof(float) * N * c.T, hipMemcpyDeviceToDevice, h->stream),
                   "copy val");
            HIPTRY(h, ksvd::launch_du_init(h->du, h->stream), "du_init");
            double one = 0.0;
            hipEvent_t e0, e1;
            HIPTRY(h, hipEventCreate(&e0), "event");
            HIPTRY(h, hipEventCreate(&e1), "event");
            HIPTRY(h, hipEventRecord(e0, h->stream), "record");
            HIPTRY(h, ksvd::launch_du_sweep(h->du, X, h->tune_D2, h->cidx, h->tune_val, h->ccnt, h->stream),
                   "dictionary update (autotune)");
            HIPTRY(h, hipEventRecord(e1, h->stream), "record");
            HIPTRY(h, hipEventSynchronize(e1), "sync");
            float t = 0.f;
            HIPTRY(h, hipEventElapsedTime(&t, e0, e1), "elapsed");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            one = t;
            if (rep > 0 && one < ms) ms = one;
        }
        note("du blocks=%d %.4f ms; ", blocks, ms);
        if (blocks == ksvd::kDuBlocks) best.default_du_ms = ms;
        if (bestd == 0.0 || ms < bestd) { bestd = ms; best.du_blocks = blocks; }
    }
    best.du_ms = bestd;
    h->du.blocks = best.du_blocks;
    HIPTRY(h, ksvd::launch_du_init(h->du, h->stream), "du_init");
    HIPTRY(h, hipStreamSynchronize(h->stream), "sync");
    best.source = 2;
    h->tune = best;
    h->tune_log = log;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_file_read_locked();
    g_tune_cache[key] = best;
    tune_file_append_locked(key, best);
    return KSVD_OK;
}

}  // namespace

extern "C" {

void ksvd_config_default(ksvd_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->T = 6;
    c->iterations = 30;
    c->use_thresh = 4;
    c->max_signals = 5000;
    c->gemm_form = KSVD_FORM_FP32;
    c->omp_variant = KSVD_OMP_AUTO;
    c->mu_thresh = 0.8;
    c->pivot_fp64 = 1e-3;
    c->perm_seed = 0;
    c->use_graph = 1;
}

const char* ksvd_version(void) { return "ksvd_amd 0.1 (gfx950)"; }

int ksvd_create(ksvd_handle** out, const ksvd_config* cfg) {
    if (!out || !cfg) return KSVD_ERR_NULL;
    *out = nullptr;
    ksvd_handle* h = new ksvd_handle();
    h->cfg = *cfg;
    int rc = check_cfg(h, cfg);
    if (rc != KSVD_OK) {
        // keep the handle so the caller can read the message, as documented
        *out = h;
        return rc;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        *out = h;
        return fail(h, KSVD_ERR_HIP, "no HIP device available");
    }
    rc = alloc(h, &h->counter, 4, "counter");
    *out = h;
    if (rc != KSVD_OK) return rc;
    float* sp = nullptr;
    return hip_check(h, scratch_ptr(&sp), "hipGetSymbolAddress(scratch)");
}

int ksvd_destroy(ksvd_handle* h) {
    if (!h) return KSVD_ERR_NULL;
    if (!h->allocs.empty()) (void)hipStreamSynchronize(h->stream);
    if (h->gstream) {
        (void)hipStreamSynchronize(h->gstream);
        (void)hipStreamDestroy(h->gstream);
    }
    if (h->ev_in) (void)hipEventDestroy(h->ev_in);
    if (h->ev_out) (void)hipEventDestroy(h->ev_out);
    for (void* p : h->allocs) (void)hipFree(p);
    delete h;
    return KSVD_OK;
}

int ksvd_set_stream(ksvd_handle* h, void* stream) {
    if (!h) return KSVD_ERR_NULL;
    h->stream = static_cast<hipStream_t>(stream);
    return KSVD_OK;
}

const char* ksvd_last_error(const ksvd_handle* h) { return h ? h->err.c_str() : "null handle"; }

int ksvd_sync(ksvd_handle* h) {
    if (!h) return KSVD_ERR_NULL;
    return hip_check(h, hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

int ksvd_normcols(ksvd_handle* h, const float* D_in, float* D_out) {
    if (!h) return KSVD_ERR_NULL;
    if (!D_in || !D_out) return fail(h, KSVD_ERR_NULL, "normcols: null pointer");
    return hip_check(h, ksvd::launch_normcols(D_in, D_out, h->cfg.n, h->cfg.K, h->stream), "normcols");
}

int ksvd_at_times_b(ksvd_handle* h, const float* A, int rows, int colsA, const float* B, int colsB, float* C) {
    if (!h) return KSVD_ERR_NULL;
    if (!A || !B || !C) return fail(h, KSVD_ERR_NULL, "at_times_b: null pointer");
    if (rows < 1 || colsA < 1 || colsB < 1)
        return fail(h, KSVD_ERR_INVALID, "at_times_b: bad shape rows=%d colsA=%d colsB=%d", rows, colsA, colsB);
    return hip_check(h, ksvd::launch_at_times_b(A, rows, colsA, B, colsB, C, h->cfg.gemm_form, h->stream, h->tune.gemm_wpe),
                     "at_times_b");
}

int ksvd_gram(ksvd_handle* h, const float* D, float* G) {
    if (!h) return KSVD_ERR_NULL;
    return ksvd_at_times_b(h, D, h->cfg.n, h->cfg.K, D, h->cfg.K, G);
}

int ksvd_correlate(ksvd_handle* h, const float* D, const float* X, float* DtX) {
    if (!h) return KSVD_ERR_NULL;
    return ksvd_at_times_b(h, D, h->cfg.n, h->cfg.K, X, h->cfg.N, DtX);
}

int ksvd_batch_omp(ksvd_handle* h, const float* DtX, const float* G, float* selected_atoms, float* c) {
    if (!h) return KSVD_ERR_NULL;
    if (!DtX || !G || !selected_atoms || !c) return fail(h, KSVD_ERR_NULL, "batch_omp: null pointer");
    ksvd::OmpArgs a = omp_args(h, DtX, G);
    a.sel = selected_atoms;
    a.c = c;
    TRY(prepare_wave(h, G, a));
    return hip_check(h, launch_wave_family(h, a), "batch_omp");
}

int ksvd_batch_omp_compact(ksvd_handle* h, const float* DtX, const float* G, int32_t* idx, float* val,
                           int32_t* cnt) {
    if (!h) return KSVD_ERR_NULL;
    if (!DtX || !G || !idx || !val || !cnt) return fail(h, KSVD_ERR_NULL, "batch_omp_compact: null pointer");
    ksvd::OmpArgs a = omp_args(h, DtX, G);
    a.idx = idx;
    a.val = val;
    a.cnt = cnt;
    TRY(prepare_wave(h, G, a));
    return hip_check(h, launch_wave_family(h, a), "batch_omp_compact");
}

int ksvd_batch_omp_compact_factors(ksvd_handle* h, const float* DtX, const float* G, int32_t* idx, float* val,
                                   int32_t* cnt, float* Lfac) {
    if (!h) return KSVD_ERR_NULL;
    if (!DtX || !G || !idx || !val || !cnt || !Lfac)
        return fail(h, KSVD_ERR_NULL, "batch_omp_compact_factors: null pointer");
    if (!use_qform(h))
        return fail(h, KSVD_ERR_INVALID, "batch_omp_compact_factors: the factors come from the register-basis form "
                    "(variant %d, K=%d, T=%d does not use it)", h->cfg.omp_variant, h->cfg.K, h->cfg.T);
    ksvd::OmpArgs a = omp_args(h, DtX, G);
    a.idx = idx;
    a.val = val;
    a.cnt = cnt;
    a.Lfac = Lfac;
    TRY(prepare_wave(h, G, a));
    return hip_check(h, launch_wave_family(h, a), "batch_omp_compact_factors");
}

int ksvd_batch_omp_from_signals(ksvd_handle* h, const float* D, const float* X, const float* G, int32_t* idx,
                                float* val, int32_t* cnt) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !G || !idx || !val || !cnt) return fail(h, KSVD_ERR_NULL, "batch_omp_from_signals: null pointer");
    return omp_from_signals(h, D, X, G, nullptr, idx, val, cnt);
}

int ksvd_batch_omp_from_signals_dense(ksvd_handle* h, const float* D, const float* X, const float* G,
                                      float* selected_atoms, float* c) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !G || !selected_atoms || !c)
        return fail(h, KSVD_ERR_NULL, "batch_omp_from_signals_dense: null pointer");
    const ksvd_config& cf = h->cfg;
    TRY(alloc(h, &h->cidx, (size_t)cf.N * cf.T, "cidx"));
    TRY(alloc(h, &h->cval, (size_t)cf.N * cf.T, "cval"));
    TRY(alloc(h, &h->ccnt, cf.N, "ccnt"));
    TRY(ksvd_batch_omp_from_signals(h, D, X, G, h->cidx, h->cval, h->ccnt));
    return hip_check(h, ksvd::launch_compact_to_slotmap(h->cidx, h->cval, h->ccnt, cf.K, cf.N, cf.T, selected_atoms, c,
                                                        h->stream),
                     "compact_to_slotmap");
}

int ksvd_batch_omp_eps_from_signals(ksvd_handle* h, const float* D, const float* X, const float* G,
                                    const float* eps2, int32_t* idx, float* val, int32_t* cnt) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !G || !eps2 || !idx || !val || !cnt)
        return fail(h, KSVD_ERR_NULL, "batch_omp_eps_from_signals: null pointer");
    TRY(compute_xnorm2(h, X));
    return omp_from_signals(h, D, X, G, eps2, idx, val, cnt);
}

int ksvd_batch_omp_from_signals_noalloc(ksvd_handle* h, const float* D, const float* X, const float* G, int N,
                                        int32_t* idx, float* val, int32_t* cnt) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !G || !idx || !val || !cnt)
        return fail(h, KSVD_ERR_NULL, "batch_omp_from_signals_noalloc: null pointer");
    const ksvd_config& c = h->cfg;
    if (N < 0 || N > c.N)
        return fail(h, KSVD_ERR_INVALID, "batch_omp_from_signals_noalloc: N=%d outside 0..%d (the handle's N)", N, c.N);
    const size_t cap = kScratchBytes / sizeof(float);
    if ((size_t)c.K > cap)
        return fail(h, KSVD_ERR_LIMIT, "batch_omp_from_signals_noalloc: one signal's correlations (K=%d floats) exceed "
                    "the %zu-byte scratch buffer", c.K, kScratchBytes);
    if (!use_qform(h) && c.omp_variant != KSVD_OMP_WAVE && c.omp_variant != KSVD_OMP_AUTO)
        return fail(h, KSVD_ERR_INVALID, "batch_omp_from_signals_noalloc: Batch-OMP variant %d needs handle scratch",
                    c.omp_variant);
    float* dtx = nullptr;
    HIPTRY(h, scratch_ptr(&dtx), "hipGetSymbolAddress(scratch)");
    // chunks of `chunk` signals: DtX of the chunk into the scratch buffer, then
    // Batch-OMP on it (stream order keeps the next chunk's GEMM behind this
    // chunk's Batch-OMP)
    const long chunk = (long)(cap / (size_t)c.K);
    const bool qf = use_qform(h);
    h->last_variant = qf ? KSVD_OMP_QFORM : KSVD_OMP_WAVE;
    h->last_noalloc = true;
    for (long s0 = 0; s0 < N; s0 += chunk) {
        const long nc = (N - s0 < chunk) ? N - s0 : chunk;
        HIPTRY(h, ksvd::launch_at_times_b(D, c.n, c.K, X + s0 * c.n, nc, dtx, c.gemm_form, h->stream, h->tune.gemm_wpe),
               "correlate (noalloc)");
        ksvd::OmpArgs a = omp_args(h, dtx, G);
        a.N = (int)nc;
        a.idx = idx + s0 * c.T;
        a.val = val + s0 * c.T;
        a.cnt = cnt + s0;
        HIPTRY(h, qf ? ksvd::launch_batch_omp_qform(a, h->stream) : ksvd::launch_batch_omp(a, h->stream),
               "batch_omp (noalloc)");
    }
    return KSVD_OK;
}

int ksvd_omp_stats(ksvd_handle* h, int64_t* candidates, int64_t* fallbacks) {
    if (!h) return KSVD_ERR_NULL;
    unsigned long long st[2] = {0ull, 0ull};
    if (!h->last_noalloc && h->omp_stats)  // the allocation-free entry point counts nothing
        HIPTRY(h, hipMemcpyAsync(st, h->omp_stats, sizeof(st), hipMemcpyDeviceToHost, h->stream), "download");
    HIPTRY(h, hipStreamSynchronize(h->stream), "sync");
    if (candidates) *candidates = (int64_t)st[0];
    if (fallbacks) *fallbacks = (int64_t)st[1];
    return KSVD_OK;
}

int ksvd_batch_omp_eps(ksvd_handle* h, const float* DtX, const float* G, const double* xnorm2, const float* eps2,
                       int32_t* idx, float* val, int32_t* cnt) {
    if (!h) return KSVD_ERR_NULL;
    if (!DtX || !G || !xnorm2 || !eps2 || !idx || !val || !cnt)
        return fail(h, KSVD_ERR_NULL, "batch_omp_eps: null pointer");
    ksvd::OmpArgs a = omp_args(h, DtX, G);
    a.idx = idx;
    a.val = val;
    a.cnt = cnt;
    TRY(prepare_wave(h, G, a, true));
    set_eps(h, a, xnorm2, eps2);
    return hip_check(h, launch_wave_family(h, a), "batch_omp_eps");
}

int ksvd_du_init(ksvd_handle* h) {
    if (!h) return KSVD_ERR_NULL;
    if (!h->du_ready) return ensure_du(h);  // allocation also resets the state
    return hip_check(h, ksvd::launch_du_init(h->du, h->stream), "du_init");
}

int ksvd_dictionary_update(ksvd_handle* h, const float* X, float* D, const int32_t* idx, float* val,
                           const int32_t* cnt, const uint32_t* perm_host) {
    if (!h) return KSVD_ERR_NULL;
    if (!X || !D || !idx || !val || !cnt || !perm_host)
        return fail(h, KSVD_ERR_NULL, "dictionary_update: null pointer");
    std::vector<char> seen(h->cfg.K, 0);
    for (int k = 0; k < h->cfg.K; ++k) {
        if (perm_host[k] >= (uint32_t)h->cfg.K || seen[perm_host[k]])
            return fail(h, KSVD_ERR_INVALID, "dictionary_update: perm is not a permutation of 0..K-1");
        seen[perm_host[k]] = 1;
    }
    TRY(ensure_du(h));
    HIPTRY(h, hipMemcpyAsync(h->du.perm_dev, perm_host, sizeof(uint32_t) * h->cfg.K, hipMemcpyHostToDevice, h->stream),
           "perm upload");
    // the host copy of perm may be reused by the caller right after return
    HIPTRY(h, hipStreamSynchronize(h->stream), "perm upload sync");
    return hip_check(h, ksvd::launch_du_sweep(h->du, X, D, idx, val, cnt, h->stream), "dictionary_update");
}

int ksvd_dictionary_update_dense(ksvd_handle* h, const float* X, float* D, float* selected_atoms, const float* c,
                                 const uint32_t* perm_host) {
    if (!h) return KSVD_ERR_NULL;
    if (!selected_atoms || !c) return fail(h, KSVD_ERR_NULL, "dictionary_update_dense: null pointer");
    const ksvd_config& cf = h->cfg;
    TRY(alloc(h, &h->cidx, (size_t)cf.N * cf.T, "cidx"));
    TRY(alloc(h, &h->cval, (size_t)cf.N * cf.T, "cval"));
    TRY(alloc(h, &h->ccnt, cf.N, "ccnt"));
    HIPTRY(h, ksvd::launch_dense_to_compact(selected_atoms, c, cf.K, cf.N, cf.T, h->cidx, h->cval, h->ccnt, h->stream),
           "dense_to_compact");
    TRY(ksvd_dictionary_update(h, X, D, h->cidx, h->cval, h->ccnt, perm_host));
    return hip_check(h, ksvd::launch_compact_to_dense(h->cidx, h->cval, h->ccnt, cf.K, cf.N, cf.T, selected_atoms,
                                                      h->stream),
                     "compact_to_dense");
}

int ksvd_error(ksvd_handle* h, const float* X, const float* D, const int32_t* idx, const float* val,
               const int32_t* cnt, float* err, double* rmse_host) {
    if (!h) return KSVD_ERR_NULL;
    if (!X || !D || !idx || !val || !cnt || !err) return fail(h, KSVD_ERR_NULL, "error: null pointer");
    TRY(ensure_du(h));
    HIPTRY(h, ksvd::launch_error(h->du, X, D, idx, val, cnt, err, h->stream), "error");
    if (rmse_host) {
        HIPTRY(h, hipMemcpyAsync(rmse_host, h->du.rmse_dev, sizeof(double), hipMemcpyDeviceToHost, h->stream),
               "rmse download");
        HIPTRY(h, hipStreamSynchronize(h->stream), "rmse sync");
    }
    return KSVD_OK;
}

int ksvd_clear_dict(ksvd_handle* h, const float* X, float* D, const int32_t* idx, const float* val,
                    const int32_t* cnt, float* err) {
    if (!h) return KSVD_ERR_NULL;
    if (!X || !D || !idx || !val || !cnt || !err) return fail(h, KSVD_ERR_NULL, "clear_dict: null pointer");
    TRY(ensure_du(h));
    return hip_check(h, ksvd::launch_clear_dict(h->du, X, D, idx, val, cnt, err, h->stream), "clear_dict");
}

// The iteration loop of ksvd_execute_iterations[_eps] (eps2 NULL: sparsity mode).
static int execute_iterations(ksvd_handle* h, const float* X, const float* D_in, float* D_out, const float* eps2,
                              int32_t* idx, float* val, int32_t* cnt, int iterations, double* rmse_host) {
    if (!X || !D_in || !D_out || !idx || !val || !cnt)
        return fail(h, KSVD_ERR_NULL, "execute_iterations: null pointer");
    if (iterations < 0) return fail(h, KSVD_ERR_INVALID, "iterations=%d", iterations);
    const ksvd_config& c = h->cfg;
    TRY(ensure_du(h));
    if (eps2) TRY(compute_xnorm2(h, X));  // X is fixed: once, outside the captured iteration
    TRY(alloc(h, &h->gram, (size_t)c.K * c.K, "G"));
    TRY(alloc(h, &h->errv, c.N, "err"));
    if (iterations > h->rmse_cap || iterations > h->perm_cap) {
        // grow: the previous buffers stay owned by the handle until destroy
        h->rmse_all = nullptr;
        h->perm_all = nullptr;
        TRY(alloc(h, &h->rmse_all, iterations > 0 ? iterations : 1, "rmse_all"));
        TRY(alloc(h, &h->perm_all, (size_t)(iterations > 0 ? iterations : 1) * c.K, "perm_all"));
        h->rmse_cap = h->perm_cap = iterations;
    }
    // all permutations up front: the loop below never waits on the host
    std::vector<uint32_t> perms((size_t)(iterations > 0 ? iterations : 1) * c.K);
    TRY(ksvd_perm_sequence(c.perm_seed, c.K, iterations, perms.data()));
    if (iterations > 0)
        HIPTRY(h, hipMemcpyAsync(h->perm_all, perms.data(), sizeof(uint32_t) * perms.size(), hipMemcpyHostToDevice,
                                 h->stream),
               "perm upload");
    HIPTRY(h, hipStreamSynchronize(h->stream), "perm upload sync");  // `perms` is pageable and local
    HIPTRY(h, ksvd::launch_normcols(D_in, D_out, c.n, c.K, h->stream), "normcols");

    // One iteration with fixed buffers (the iteration's permutation is copied
    // into du.perm_dev and its RMSE read from du.rmse_dev around it), so the
    // same launch sequence can be captured once and replayed.
    auto enqueue = [&]() -> int {
        HIPTRY(h, ksvd::launch_at_times_b(D_out, c.n, c.K, D_out, c.K, h->gram, c.gemm_form, h->stream, h->tune.gemm_wpe), "gram");
        TRY(omp_from_signals(h, D_out, X, h->gram, eps2, idx, val, cnt));
        HIPTRY(h, ksvd::launch_du_init(h->du, h->stream), "du_init");
        HIPTRY(h, ksvd::launch_du_sweep(h->du, X, D_out, idx, val, cnt, h->stream), "dictionary_update");
        HIPTRY(h, ksvd::launch_error_from_residual(h->du, h->errv, h->stream), "error");
        HIPTRY(h, ksvd::launch_clear_dict(h->du, X, D_out, idx, val, cnt, h->errv, h->stream), "clear_dict");
        return KSVD_OK;
    };
    auto stage_in = [&](int it) {
        return hip_check(h, hipMemcpyAsync(h->du.perm_dev, h->perm_all + (size_t)it * c.K, sizeof(uint32_t) * c.K,
                                           hipMemcpyDeviceToDevice, h->stream),
                         "perm stage");
    };
    auto stage_out = [&](int it) {
        return hip_check(h, hipMemcpyAsync(h->rmse_all + it, h->du.rmse_dev, sizeof(double), hipMemcpyDeviceToDevice,
                                           h->stream),
                         "rmse stage");
    };

    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int rc = KSVD_OK;
    for (int it = 0; it < iterations && rc == KSVD_OK; ++it) {
        if ((rc = stage_in(it)) != KSVD_OK) break;
        if (it == 0 || !c.use_graph) {
            rc = enqueue();  // first iteration eagerly: every buffer gets allocated here
        } else {
            if (!exec) {
                if (!h->gstream) {
                    HIPTRY(h, hipStreamCreateWithFlags(&h->gstream, hipStreamNonBlocking), "stream");
                    HIPTRY(h, hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming), "event");
                    HIPTRY(h, hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming), "event");
                }
                hipStream_t user = h->stream;
                HIPTRY(h, hipStreamBeginCapture(h->gstream, hipStreamCaptureModeThreadLocal), "begin capture");
                h->stream = h->gstream;
                const int crc = enqueue();
                h->stream = user;
                const hipError_t ec = hipStreamEndCapture(h->gstream, &graph);
                if (crc != KSVD_OK) { rc = crc; break; }
                if (ec != hipSuccess) { rc = hip_check(h, ec, "end capture"); break; }
                if ((rc = hip_check(h, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate")) !=
                    KSVD_OK)
                    break;
            }
            // replay on the private stream, ordered after the caller's stream
            if ((rc = hip_check(h, hipEventRecord(h->ev_in, h->stream), "record")) != KSVD_OK) break;
            if ((rc = hip_check(h, hipStreamWaitEvent(h->gstream, h->ev_in, 0), "wait")) != KSVD_OK) break;
            if ((rc = hip_check(h, hipGraphLaunch(exec, h->gstream), "graph launch")) != KSVD_OK) break;
            if ((rc = hip_check(h, hipEventRecord(h->ev_out, h->gstream), "record")) != KSVD_OK) break;
            if ((rc = hip_check(h, hipStreamWaitEvent(h->stream, h->ev_out, 0), "wait")) != KSVD_OK) break;
        }
        if (rc == KSVD_OK) rc = stage_out(it);
    }
    if (exec) {
        (void)hipStreamSynchronize(h->gstream);
        (void)hipGraphExecDestroy(exec);
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != KSVD_OK) return rc;
    if (rmse_host && iterations > 0) {
        HIPTRY(h, hipMemcpyAsync(rmse_host, h->rmse_all, sizeof(double) * iterations, hipMemcpyDeviceToHost, h->stream),
               "rmse download");
        HIPTRY(h, hipStreamSynchronize(h->stream), "rmse sync");
    }
    return KSVD_OK;
}

int ksvd_execute_iterations(ksvd_handle* h, const float* X, const float* D_in, float* D_out, int32_t* idx, float* val,
                            int32_t* cnt, int iterations, double* rmse_host) {
    if (!h) return KSVD_ERR_NULL;
    return execute_iterations(h, X, D_in, D_out, nullptr, idx, val, cnt, iterations, rmse_host);
}

int ksvd_execute_iterations_eps(ksvd_handle* h, const float* X, const float* D_in, float* D_out, const float* eps2,
                                int32_t* idx, float* val, int32_t* cnt, int iterations, double* rmse_host) {
    if (!h) return KSVD_ERR_NULL;
    if (!eps2) return fail(h, KSVD_ERR_NULL, "execute_iterations_eps: null eps2");
    return execute_iterations(h, X, D_in, D_out, eps2, idx, val, cnt, iterations, rmse_host);
}

int ksvd_get_tuning(const ksvd_handle* h, ksvd_tuning* out) {
    if (!h || !out) return KSVD_ERR_NULL;
    *out = h->tune;
    return KSVD_OK;
}

const char* ksvd_tuning_log(const ksvd_handle* h) { return h ? h->tune_log.c_str() : ""; }

int ksvd_set_tuning(ksvd_handle* h, const ksvd_tuning* t) {
    if (!h || !t) return KSVD_ERR_NULL;
    const ksvd_config& c = h->cfg;
    if ((t->omp_kpl != 0) != (t->omp_waves != 0))
        return fail(h, KSVD_ERR_INVALID, "set_tuning: omp_kpl and omp_waves are set together");
    if (t->omp_kpl && !ksvd::qform_shape_with(c.K, c.T, t->omp_kpl, t->omp_waves).kpl)
        return fail(h, KSVD_ERR_INVALID, "set_tuning: register-basis shape kpl=%d waves=%d is not compiled for "
                    "K=%d T=%d", t->omp_kpl, t->omp_waves, c.K, c.T);
    if (t->omp_eps_launch != 0 && t->omp_eps_launch != 1)
        return fail(h, KSVD_ERR_INVALID, "set_tuning: omp_eps_launch=%d", t->omp_eps_launch);
    if (t->gemm_wpe != 0 && t->gemm_wpe != 2 && t->gemm_wpe != 3)
        return fail(h, KSVD_ERR_INVALID, "set_tuning: gemm_wpe=%d (2 or 3)", t->gemm_wpe);
    if (t->du_blocks != 0 && (t->du_blocks < 64 || t->du_blocks > ksvd::kDuBlocksMax))
        return fail(h, KSVD_ERR_INVALID, "set_tuning: du_blocks=%d (64..%d)", t->du_blocks, ksvd::kDuBlocksMax);
    h->tune = *t;
    h->tune.source = 1;
    if (h->du_ready) h->du.blocks = t->du_blocks > 0 ? t->du_blocks : ksvd::kDuBlocks;
    return KSVD_OK;
}

int ksvd_autotune(ksvd_handle* h, const float* D, const float* X, const float* eps2, int force) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X) return fail(h, KSVD_ERR_NULL, "autotune: null pointer");
    return autotune(h, D, X, eps2, force != 0);
}

int ksvd_perm_sequence(uint64_t seed, int K, int iterations, uint32_t* out) {
    if (!out) return KSVD_ERR_NULL;
    if (K < 1 || iterations < 0) return KSVD_ERR_INVALID;
    std::vector<uint32_t> buf(K);
    for (int k = 0; k < K; ++k) buf[k] = (uint32_t)k;
    uint64_t st = seed;
    for (int it = 0; it < iterations; ++it) {
        shuffle(buf.data(), K, &st);
        std::memcpy(out + (size_t)it * K, buf.data(), sizeof(uint32_t) * K);
    }
    return KSVD_OK;
}

static int omp_adjoint(ksvd_handle* h, const float* D, const float* X, const int32_t* idx, const float* val,
                       const int32_t* cnt, const float* Lfac, const float* gval, float* gD, float* gX) {
    const ksvd_config& c = h->cfg;
    TRY(ensure_du(h));
    TRY(alloc(h, &h->adj_z, (size_t)c.N * c.T, "adjoint z"));
    // per-atom lists (ascending signal order), as the dictionary update builds them
    HIPTRY(h, ksvd::launch_du_build_lists(h->du, idx, cnt, h->stream), "adjoint lists");
    return hip_check(h,
                     ksvd::launch_omp_adjoint(D, X, idx, val, cnt, Lfac, gval, gD, gX, h->du.ptr, h->du.list_pos,
                                              h->du.R, h->adj_z, c.n, c.K, c.N, c.T, h->stream),
                     "omp_adjoint");
}

int ksvd_omp_adjoint(ksvd_handle* h, const float* D, const float* X, const int32_t* idx, const float* val,
                     const int32_t* cnt, const float* gval, float* gD, float* gX) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !idx || !val || !cnt || !gval || !gD || !gX)
        return fail(h, KSVD_ERR_NULL, "omp_adjoint: null pointer");
    return omp_adjoint(h, D, X, idx, val, cnt, nullptr, gval, gD, gX);
}

int ksvd_omp_adjoint_factored(ksvd_handle* h, const float* D, const float* X, const int32_t* idx, const float* val,
                              const int32_t* cnt, const float* Lfac, const float* gval, float* gD, float* gX) {
    if (!h) return KSVD_ERR_NULL;
    if (!D || !X || !idx || !val || !cnt || !Lfac || !gval || !gD || !gX)
        return fail(h, KSVD_ERR_NULL, "omp_adjoint_factored: null pointer");
    return omp_adjoint(h, D, X, idx, val, cnt, Lfac, gval, gD, gX);
}

int ksvd_normcols_adjoint(ksvd_handle* h, const float* D_in, const float* gD_out, float* gD_in) {
    if (!h) return KSVD_ERR_NULL;
    if (!D_in || !gD_out || !gD_in) return fail(h, KSVD_ERR_NULL, "normcols_adjoint: null pointer");
    return hip_check(h, ksvd::launch_normcols_adjoint(D_in, gD_out, gD_in, h->cfg.n, h->cfg.K, h->stream),
                     "normcols_adjoint");
}

int ksvd_compact_to_dense(ksvd_handle* h, const int32_t* idx, const float* val, const int32_t* cnt,
                          float* gamma_dense) {
    if (!h) return KSVD_ERR_NULL;
    if (!idx || !val || !cnt || !gamma_dense) return fail(h, KSVD_ERR_NULL, "compact_to_dense: null pointer");
    const ksvd_config& c = h->cfg;
    return hip_check(h, ksvd::launch_compact_to_dense(idx, val, cnt, c.K, c.N, c.T, gamma_dense, h->stream),
                     "compact_to_dense");
}

int ksvd_compact_gather(ksvd_handle* h, const float* dense, const int32_t* idx, const int32_t* cnt, float* out) {
    if (!h) return KSVD_ERR_NULL;
    if (!dense || !idx || !cnt || !out) return fail(h, KSVD_ERR_NULL, "compact_gather: null pointer");
    const ksvd_config& c = h->cfg;
    return hip_check(h, ksvd::launch_gather_compact(dense, idx, cnt, c.K, c.N, c.T, out, h->stream), "compact_gather");
}

}  // extern "C"
