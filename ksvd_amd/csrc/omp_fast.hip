// omp_fast.hip -- Batch-OMP with the correlation update on MFMA.
//
// The reference updates alpha = alpha0 - G(:,I) gamma per signal (addMul,
// S/OMP.cu:234-247): K*i multiply-adds per step on gathered columns of G that
// no two signals share, so on a GPU the loop is bound by L2 gathers of G
// (~K*T^2/2*4 bytes per signal).  The same vector is alpha = D^T r with the
// residual r = x - D_I gamma.  This kernel computes it for a tile of 64
// signals at once as an MFMA product of the shared dictionary with the
// residuals, in fp16 with FP32 accumulation, and then makes the selection
// exact:
//   * |alpha~_k - alpha_k| <= eps_s, a rigorous bound from fp16 rounding
//     (2u + u^2 with u = 2^-11 per product, FP32 accumulation, flushed
//     denormals) in the tile's scaled units;
//   * every atom within 2 eps_s of the largest |alpha~| is re-evaluated in
//     FP32 as d_k^T r and the arg-max taken over those (lowest index on ties),
//     which is the arg-max of the FP32 correlations;
//   * when the candidate lists kept during the MFMA sweep may be truncated,
//     the signal falls back to FP32 correlations with every atom.
// Everything else follows parOMP (S/OMP.cu:254-405): stop on a re-selected
// atom or alpha^2 < 1e-14, Cholesky row by w = L^-1 G(I,pos) with pivot
// 1 - ||w||^2 (FP64 below pivot_fp64), gamma = L^-T L^-1 alpha0_I.  alpha0_I
// is d_pos^T x evaluated when pos is selected, so DtX is never materialised.
//
// Tile layout (one workgroup = 4 waves, 64 signal slots):
//   LDS   r (64 x NP FP32 residuals), per-slot scale / eps / state
//   VGPR  fp16 residual fragments (64 signals x NP) for the whole MFMA sweep
//   L2    Dh: fp16 copy of 256*D (Kp x NP), streamed once per step per tile
//   HBM   per-slot Cholesky factor and support state (workgroup scratch)
// Slots are refilled from a global queue as signals finish (work stealing),
// so uneven support sizes (error-constrained mode) keep the tile full.
#include "common.hpp"

namespace ksvd {

typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;

namespace {

constexpr int S = 64;        // signal slots per workgroup
constexpr int NT = 256;      // threads
constexpr float DSCALE = 256.0f;  // Dh = fp16(DSCALE * D)
constexpr float RMAX_EXP = 7;     // scaled residual max in [2^6, 2^7)

struct FastArgs {
    const float* X;      // n x N
    const float* D;      // n x K (FP32)
    const _Float16* Dh;  // Kp x NP fp16, DSCALE * D, zero padded
    const float* G;      // K x K
    const float* dmax;   // [0] = max_k ||d_k||, [1] = max_k ||d_k||_1
    int n, NP, K, Kp, N, T;
    int32_t* idx;        // T x N
    float* val;          // T x N
    int32_t* cnt;        // N
    const float* eps2;   // error mode: stop when ||r||^2 <= eps2[s] (NULL: sparsity mode)
    int32_t* queue;      // [0] next signal
    float* Lws;          // per workgroup: S x T x (T+1)
    float* Sws;          // per workgroup: S x 4 x T  (a0I, y, gamma, diag)
    int32_t* Iws;        // per workgroup: S x T      (support)
    int32_t* fallbacks;  // [0] count of full-correlation fallbacks (diagnostic)
    float pivot_fp64;
};

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// -------------------------------------------------------------------------
// D -> Dh (fp16, scaled, padded) and the norms used by the error bound.
// -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void prep_dict_kernel(const float* __restrict__ D, int n, int NP, int K, int Kp,
                                                        _Float16* __restrict__ Dh, float* __restrict__ dnorm) {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= Kp) return;
    float s2 = 0.f, s1 = 0.f;
    for (int t = lane; t < NP; t += 64) {
        const float v = (k < K && t < n) ? D[(long)k * n + t] : 0.f;
        Dh[(long)k * NP + t] = (_Float16)(v * DSCALE);
        s2 += v * v;
        s1 += fabsf(v);
    }
    s2 = wave_sum(s2);
    s1 = wave_sum(s1);
    if (lane == 0 && k < K) {
        dnorm[2 * k] = sqrtf(s2);
        dnorm[2 * k + 1] = s1;
    }
}

__global__ __launch_bounds__(1024) void dict_max_kernel(const float* __restrict__ dnorm, int K, float* __restrict__ dmax,
                                                        int32_t* __restrict__ queue, int32_t* __restrict__ fallbacks) {
    __shared__ float r2[16], r1[16];
    float m2 = 0.f, m1 = 0.f;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        m2 = fmaxf(m2, dnorm[2 * k]);
        m1 = fmaxf(m1, dnorm[2 * k + 1]);
    }
    m2 = wave_max(m2);
    m1 = wave_max(m1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { r2[w] = m2; r1[w] = m1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { m2 = fmaxf(r2[0], r2[i]); r2[0] = m2; m1 = fmaxf(r1[0], r1[i]); r1[0] = m1; }
        dmax[0] = r2[0] * 1.0000005f;  // bound of ||d_k|| (rounding of the norm itself)
        dmax[1] = r1[0] * 1.0000005f;
        queue[0] = 0;
        fallbacks[0] = 0;
    }
}

// -------------------------------------------------------------------------
// The tile kernel.
// -------------------------------------------------------------------------
template <int NPB>  // NP / 16 (k-steps of 16 along the signal length), NP <= 256
__global__ __launch_bounds__(NT, 2) void omp_fast_kernel(FastArgs a) {
    constexpr int NP = NPB * 16;
    __shared__ float rs[S * NP];           // residuals (FP32)
    __shared__ float cv[S * 16];           // candidate values (scaled |alpha~|)
    __shared__ int ci[S * 16];             // candidate atoms
    __shared__ float rscale[S];            // per-slot residual scale (power of two)
    __shared__ float reps[S];              // per-slot error bound eps (scaled units)
    __shared__ int slot_sig[S];            // signal in slot (-1: empty)
    __shared__ int slot_i[S];              // atoms accepted so far
    __shared__ int active_count;

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int n = a.n, K = a.K, T = a.T, LD = T + 1;
    float* Lw = a.Lws + (long)blockIdx.x * S * T * LD;
    float* Sw = a.Sws + (long)blockIdx.x * S * 4 * T;
    int32_t* Iw = a.Iws + (long)blockIdx.x * S * T;
    const float dm = a.dmax[0], dm1 = a.dmax[1];
    const bool eps_mode = a.eps2 != nullptr;

    // ---- slot management ------------------------------------------------------
    // Loads signal `sig` into slot `sl` (executed by one wave): r = x, state reset.
    auto load_slot = [&](int sl, int sig) {
        float mx = 0.f, ss = 0.f;
        for (int t = lane; t < NP; t += 64) {
            const float v = (t < n) ? a.X[(long)sig * n + t] : 0.f;
            rs[sl * NP + t] = v;
            mx = fmaxf(mx, fabsf(v));
            ss += v * v;
        }
        mx = wave_max(mx);
        ss = wave_sum(ss);
        int e = 0;
        frexpf(mx, &e);
        const float sc = (mx > 0.f) ? ldexpf(1.0f, (int)RMAX_EXP - e) : 1.0f;
        if (lane == 0) {
            slot_sig[sl] = sig;
            slot_i[sl] = 0;
            rscale[sl] = sc;
            // |alpha~ - alpha| in scaled units (see header)
            reps[sl] = sc * DSCALE * (1.1e-3f * dm * sqrtf(ss)) + 6.2e-5f * (n * 128.0f + DSCALE * dm1);
        }
        if (lane < T) Iw[sl * T + lane] = -1;
    };

    auto claim = [&]() -> int {
        int v = 0;
        if (lane == 0) v = atomicAdd(a.queue, 1);
        v = readlane(v, 0);
        return v < a.N ? v : -1;
    };

    // initial fill: wave w fills slots w, w+4, ...
    for (int sl = w; sl < S; sl += 4) {
        const int sig = claim();
        if (sig >= 0) {
            load_slot(sl, sig);
        } else {
            for (int t = lane; t < NP; t += 64) rs[sl * NP + t] = 0.f;
            if (lane == 0) { slot_sig[sl] = -1; rscale[sl] = 0.f; reps[sl] = 0.f; slot_i[sl] = 0; }
        }
    }
    __syncthreads();

    for (;;) {
        // ---- phase 1: alpha~ = Dh^T r~ for all slots, per-lane top-2 ---------------
        {
            // residual fragments: lane supplies signal column (lane & 31) and
            // eight consecutive samples starting at 16*kk + 8*(lane >> 5)
            f16x8 rf[2][NPB];
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) {
                const int sl = sb * 32 + (lane & 31);
                const float sc = rscale[sl];
#pragma unroll
                for (int kk = 0; kk < NPB; ++kk) {
                    const float* p = rs + sl * NP + kk * 16 + (lane >> 5) * 8;
#pragma unroll
                    for (int e = 0; e < 8; ++e) rf[sb][kk][e] = (_Float16)(p[e] * sc);
                }
            }
            float v1[2], v2[2];
            int i1[2], i2[2];
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) { v1[sb] = -1.f; v2[sb] = -1.f; i1[sb] = 0x7fffffff; i2[sb] = 0x7fffffff; }
            const int kw = a.Kp / 4;
            const int k0 = w * kw;
            for (int ab = 0; ab < kw; ab += 32) {
                const int atom = k0 + ab + (lane & 31);
                const _Float16* dp = a.Dh + (long)atom * NP + (lane >> 5) * 8;
                f32x16 accs[2];
#pragma unroll
                for (int sb = 0; sb < 2; ++sb)
#pragma unroll
                    for (int r = 0; r < 16; ++r) accs[sb][r] = 0.f;
#pragma unroll
                for (int kk = 0; kk < NPB; ++kk) {
                    const f16x8 df = *reinterpret_cast<const f16x8*>(dp + kk * 16);
                    accs[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(df, rf[0][kk], accs[0], 0, 0, 0);
                    accs[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(df, rf[1][kk], accs[1], 0, 0, 0);
                }
#pragma unroll
                for (int sb = 0; sb < 2; ++sb) {
                    const f32x16 acc = accs[sb];
                    // register r: atom row 8*(r/4) + 4*(lane/32) + r%4 of this block
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int k = k0 + ab + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        const float v = fabsf(acc[r]);
                        if (k < K) {
                            if (v > v2[sb]) {
                                if (v > v1[sb]) { v2[sb] = v1[sb]; i2[sb] = i1[sb]; v1[sb] = v; i1[sb] = k; }
                                else { v2[sb] = v; i2[sb] = k; }
                            }
                        }
                    }
                }
            }
            // candidate slot of this lane: 4 waves x 2 lane halves = 8 lists per signal
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) {
                const int sl = sb * 32 + (lane & 31);
                const int q = (w * 2 + (lane >> 5)) * 2;
                cv[sl * 16 + q] = v1[sb];
                ci[sl * 16 + q] = i1[sb];
                cv[sl * 16 + q + 1] = v2[sb];
                ci[sl * 16 + q + 1] = i2[sb];
            }
        }
        if (threadIdx.x == 0) active_count = 0;
        __syncthreads();

        // ---- phase 2: per-slot selection, Cholesky, solves, residual -----------------
        for (int sl = w; sl < S; sl += 4) {
            int sig = slot_sig[sl];
            if (sig < 0) continue;
            const int i = slot_i[sl];
            int cfinal = i;
            float* L = Lw + (long)sl * T * LD;
            float* st = Sw + (long)sl * 4 * T;
            int32_t* Is = Iw + sl * T;
            const float* x = a.X + (long)sig * n;
            const float* r = rs + sl * NP;
            bool finish = false;

            // per-slot state, slot j in lane j
            const int ind_l = (lane < T) ? Is[lane] : -1;
            float a0I_l = (lane < T) ? st[lane] : 0.f;
            float y_l = (lane < T) ? st[T + lane] : 0.f;
            float gam_l = (lane < T) ? st[2 * T + lane] : 0.f;
            float diag_l = (lane < T) ? st[3 * T + lane] : 1.f;
            int ind_new = ind_l;

            // residual norm^2 (FP64) for the error-mode stop
            if (eps_mode) {
                double ss = 0.0;
                for (int t = lane; t < n; t += 64) ss += (double)r[t] * (double)r[t];
                ss = uniform(wave_sum(ss));
                if (ss <= (double)a.eps2[sig]) finish = true;
            }
            int pos = -1;
            float apos = 0.f;
            if (!finish && i < T) {
                // candidates within 2 eps of the largest |alpha~|
                const float cvl = (lane < 16) ? cv[sl * 16 + lane] : -1.f;
                const int cil = (lane < 16) ? ci[sl * 16 + lane] : 0x7fffffff;
                const float M = uniform(wave_max(cvl));
                const float W = M - 2.f * reps[sl];
                const bool inwin = lane < 16 && cil < K && cvl >= W;
                // a list's second entry inside the window may hide a third
                const bool trunc = inwin && (lane & 1);
                const unsigned long long tb = __ballot(trunc);
                float best = -1.f;
                int bestk = 0x7fffffff;
                if (tb == 0ull) {
                    unsigned long long m = __ballot(inwin);
                    while (m) {
                        const int l = __builtin_ctzll(m);
                        m &= m - 1;
                        const int k = readlane(cil, l);
                        float d = 0.f;
                        for (int t = lane; t < n; t += 64) d += a.D[(long)k * n + t] * r[t];
                        d = wave_sum(d);
                        const float v = d * d;
                        if (v > best || (v == best && k < bestk)) { best = v; bestk = k; apos = d; }
                    }
                } else {
                    if (lane == 0) atomicAdd(a.fallbacks, 1);
                    for (int k = 0; k < K; ++k) {
                        float d = 0.f;
                        for (int t = lane; t < n; t += 64) d += a.D[(long)k * n + t] * r[t];
                        d = wave_sum(d);
                        const float v = d * d;
                        if (v > best) { best = v; bestk = k; apos = d; }
                    }
                }
                pos = uniform(bestk);
                apos = uniform(apos);
                best = uniform(best);
                if (pos >= K) finish = true;  // no candidate (cannot happen for K >= 1)
                // stop: atom selected twice, or inner product too small (S/OMP.cu:292,304)
                const bool again = __ballot(lane < i && ind_l == pos) != 0ull;
                if (again || (double)best < 1e-14) finish = true;
            } else {
                finish = true;
            }

            if (!finish && pos < K) {
                // Cholesky row: w = L^-1 G(I, pos), pivot 1 - ||w||^2
                bool ok = true;
                if (i == 0) {
                    if (lane == 0) { L[0] = 1.f; }
                    if (lane == 0) diag_l = 1.f;
                } else {
                    const float bsrc = (lane < i) ? a.G[(long)pos * K + ind_l] : 0.f;
                    float b = bsrc, wv = 0.f;
                    for (int j = 0; j < i; ++j) {
                        const float wj = readlane(b, j) / readlane(diag_l, j);
                        if (lane == j) wv = wj;
                        if (lane > j && lane < i) b = b - L[j * LD + lane] * wj;
                    }
                    const float ssq = uniform(wave_sum((lane < i) ? wv * wv : 0.f));
                    float d = 1.f - ssq;
                    double dd = (double)d;
                    if (d < a.pivot_fp64) {
                        double b64 = (double)bsrc, w64 = 0.0;
                        for (int j = 0; j < i; ++j) {
                            const double wj = readlane(b64, j) / (double)readlane(diag_l, j);
                            if (lane == j) w64 = wj;
                            if (lane > j && lane < i) b64 = b64 - (double)L[j * LD + lane] * wj;
                        }
                        dd = uniform(1.0 - wave_sum((lane < i) ? w64 * w64 : 0.0));
                        wv = (float)w64;
                    }
                    if (dd <= 1e-14) {
                        ok = false;  // S/OMP.cu:331-333
                    } else {
                        const float lii = (float)sqrt(dd);
                        if (lane < i) L[lane * LD + i] = wv;
                        if (lane == i) { L[i * LD + i] = lii; diag_l = lii; }
                    }
                }
                if (!ok) {
                    finish = true;
                } else {
                    // alpha0_pos = d_pos^T x
                    float a0p = 0.f;
                    for (int t = lane; t < n; t += 64) a0p += a.D[(long)pos * n + t] * x[t];
                    a0p = wave_sum(a0p);
                    if (lane == i) { a0I_l = a0p; ind_new = pos; }
                    __threadfence_block();
                    wave_lds_fence();
                    const int i1 = i + 1;
                    // y_i = (alpha0_I[i] - L(i, 0:i-1) y) / L(i,i)
                    {
                        const float part = (lane < i) ? L[lane * LD + i] * y_l : 0.f;
                        const float t = wave_sum(part);
                        const float yn = (readlane(a0I_l, i) - t) / readlane(diag_l, i);
                        if (lane == i) y_l = yn;
                    }
                    // gamma = L^-T y
                    {
                        float rhs = y_l;
                        for (int l = i1 - 1; l >= 0; --l) {
                            const float gl = readlane(rhs, l) / readlane(diag_l, l);
                            if (lane == l) gam_l = gl;
                            if (lane < l) rhs = rhs - L[lane * LD + l] * gl;
                        }
                    }
                    // residual r = x - D_I gamma, its scale and error bound
                    float* rw = rs + sl * NP;
                    float mx = 0.f, ss = 0.f;
                    for (int t = lane; t < n; t += 64) {
                        float acc = 0.f;
                        for (int j = 0; j < i1; ++j) {
                            const int c = readlane(ind_new, j);
                            acc += a.D[(long)c * n + t] * readlane(gam_l, j);
                        }
                        const float v = x[t] - acc;
                        rw[t] = v;
                        mx = fmaxf(mx, fabsf(v));
                        ss += v * v;
                    }
                    mx = wave_max(mx);
                    ss = wave_sum(ss);
                    int e = 0;
                    frexpf(mx, &e);
                    const float sc = (mx > 0.f) ? ldexpf(1.0f, (int)RMAX_EXP - e) : 1.0f;
                    if (lane < T) {
                        Is[lane] = ind_new;
                        st[lane] = a0I_l;
                        st[T + lane] = y_l;
                        st[2 * T + lane] = gam_l;
                        st[3 * T + lane] = diag_l;
                    }
                    cfinal = i1;
                    if (lane == 0) {
                        slot_i[sl] = i1;
                        rscale[sl] = sc;
                        reps[sl] = sc * DSCALE * (1.1e-3f * dm * sqrtf(ss)) + 6.2e-5f * (n * 128.0f + DSCALE * dm1);
                    }
                    if (i1 >= T) finish = true;
                    if (mx == 0.f) finish = true;  // exact fit: every correlation is 0
                }
            }

            if (finish) {
                // write the compact result and refill the slot
                const int c = cfinal;
                if (lane < T) {
                    a.idx[(long)sig * T + lane] = (lane < c) ? ind_new : -1;
                    a.val[(long)sig * T + lane] = (lane < c) ? gam_l : 0.f;
                }
                if (lane == 0) a.cnt[sig] = c;
                wave_lds_fence();
                const int nsig = claim();
                if (nsig >= 0) load_slot(sl, nsig);
                else if (lane == 0) slot_sig[sl] = -1;
            }
            wave_lds_fence();
            if (lane == 0 && slot_sig[sl] >= 0) atomicAdd(&active_count, 1);
        }
        __syncthreads();
        if (active_count == 0) break;
    }
}

}  // namespace

static int fast_npb(int n) {
    const int npb = (n + 15) / 16;
    return npb <= 4 ? 4 : npb <= 8 ? 8 : npb <= 12 ? 12 : 16;
}

size_t fast_scratch_bytes(int n, int K, int T, int grid) {
    const int NP = fast_npb(n) * 16;
    const int Kp = ((K + 127) / 128) * 128;
    size_t b = 0;
    b += (size_t)Kp * NP * sizeof(_Float16);         // Dh
    b += (size_t)2 * K * sizeof(float) + 64;         // dnorm
    b += 64;                                         // dmax
    b += (size_t)grid * S * T * (T + 1) * sizeof(float);  // L
    b += (size_t)grid * S * 4 * T * sizeof(float);        // state
    b += (size_t)grid * S * T * sizeof(int32_t);          // support
    return b + 16 * 256;
}

int fast_grid() {
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus * 2;
}

hipError_t launch_batch_omp_fast(const float* X, const float* D, const float* G, int n, int K, int N, int T,
                                 const float* eps2, int32_t* idx, float* val, int32_t* cnt, float pivot_fp64,
                                 void* scratch, int grid, int32_t* counters, hipStream_t st) {
    if (n > 256) return hipErrorInvalidValue;
    const int npb = fast_npb(n);
    const int NP = npb * 16;
    const int Kp = ((K + 127) / 128) * 128;
    char* p = static_cast<char*>(scratch);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) & ~size_t(255);
        return q;
    };
    _Float16* Dh = reinterpret_cast<_Float16*>(take((size_t)Kp * NP * sizeof(_Float16)));
    float* dnorm = reinterpret_cast<float*>(take((size_t)2 * K * sizeof(float) + 64));
    float* dmax = reinterpret_cast<float*>(take(64));
    float* Lws = reinterpret_cast<float*>(take((size_t)grid * S * T * (T + 1) * sizeof(float)));
    float* Sws = reinterpret_cast<float*>(take((size_t)grid * S * 4 * T * sizeof(float)));
    int32_t* Iws = reinterpret_cast<int32_t*>(take((size_t)grid * S * T * sizeof(int32_t)));
    hipLaunchKernelGGL(prep_dict_kernel, dim3((Kp + 3) / 4), dim3(256), 0, st, D, n, NP, K, Kp, Dh, dnorm);
    hipLaunchKernelGGL(dict_max_kernel, dim3(1), dim3(1024), 0, st, dnorm, K, dmax, counters, counters + 1);
    FastArgs a{};
    a.X = X; a.D = D; a.Dh = Dh; a.G = G; a.dmax = dmax;
    a.n = n; a.NP = NP; a.K = K; a.Kp = Kp; a.N = N; a.T = T;
    a.idx = idx; a.val = val; a.cnt = cnt; a.eps2 = eps2;
    a.queue = counters; a.fallbacks = counters + 1;
    a.Lws = Lws; a.Sws = Sws; a.Iws = Iws;
    a.pivot_fp64 = pivot_fp64;
    dim3 g(grid), b(NT);
    if (npb <= 4) hipLaunchKernelGGL(omp_fast_kernel<4>, g, b, 0, st, a);
    else if (npb <= 8) hipLaunchKernelGGL(omp_fast_kernel<8>, g, b, 0, st, a);
    else if (npb <= 12) hipLaunchKernelGGL(omp_fast_kernel<12>, g, b, 0, st, a);
    else hipLaunchKernelGGL(omp_fast_kernel<16>, g, b, 0, st, a);
    return hipGetLastError();
}

}  // namespace ksvd
