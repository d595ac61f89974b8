// omp.hip -- Batch-OMP on wave64 (CDNA4).
//
// Reference: parOMP (S/OMP.cu:254-405), one CUDA thread per signal with
// per-thread arrays alpha[32], Lchol[T*T], ind[T] (K <= 32).  Here one wave64
// owns one signal:
//   * alpha0 and alpha are spread over the lanes in VGPRs: lane l holds atoms
//     k = m*64 + l, m < KPL, so every column of G is read coalesced (256 B per
//     load instruction) in the alpha update alpha = alpha0 - G(:,I) gamma;
//   * slot j of the support lives in lane j (T <= 64): atom index, alpha0(I_j),
//     y_j = (L^-1 alpha0_I)_j, gamma_j and the diagonal L_jj are registers;
//   * the Cholesky factor L (T x T) lives in LDS with an odd column pitch T+1,
//     so the row-wise reads of the back substitution are bank-conflict free;
//   * triangular solves are lane-parallel right-looking sweeps: one broadcast
//     (v_readlane) and one FMA per step;
//   * the slot map (SELECTED_ATOMS column) is a byte array in LDS, which also
//     answers the reference's "atom selected twice" test.
// The pivot d = 1 - ||w||^2 is recomputed in FP64 when it falls below
// pivot_fp64 (ill-conditioned: the new atom is nearly in the span of I).
//
// Error-constrained mode (EPS) stops when ||x||^2 - alpha0_I^T gamma_I (FP64)
// <= eps^2 and runs persistently: each wave takes the next signal from a
// global counter (work stealing), which balances the uneven support sizes.
#include "omp_common.hpp"

namespace ksvd {

namespace {

constexpr int WPB = 4;  // waves per block

// ---------------------------------------------------------------------------
// Plain FP32 form: alpha = alpha0 - G(:,I) gamma recomputed every step from
// FP32 columns of G (addMul, S/OMP.cu:234-247); lane l holds atoms m*64 + l.
// ---------------------------------------------------------------------------
template <int KPL, bool EPS>
__global__ __launch_bounds__(256) void omp_wave_kernel(OmpArgs a) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int K = a.K, T = a.T, LD = T + 1;
    float* L = smem + w * (T * LD);
    unsigned char* slot = reinterpret_cast<unsigned char*>(smem + WPB * T * LD) + w * (KPL * 64);
    const float* __restrict__ G = a.G;

    for (long it = 0;; ++it) {
        long s;
        if constexpr (EPS) {
            int v = 0;
            if (lane == 0) v = atomicAdd(a.work_counter, 1);
            s = readlane(v, 0);
        } else {
            if (it > 0) break;
            s = (long)blockIdx.x * WPB + w;
        }
        if (s >= a.N) break;

        const float* __restrict__ a0p = a.DtX + s * (long)K;
        float a0[KPL], al[KPL];
#pragma unroll
        for (int m = 0; m < KPL; ++m) {
            const int k = m * 64 + lane;
            a0[m] = (k < K) ? a0p[k] : 0.f;
            al[m] = 0.f;
            slot[k] = 0;
        }
        Support sp;
        double xn2 = 0.0, e2 = 0.0;
        if constexpr (EPS) {
            xn2 = (double)a.xnorm2[s];
            e2 = (double)a.eps2[s];
        }
        wave_lds_fence();

        int i = 0;
        while (i < T) {
            if constexpr (EPS) {
                const double part = (lane < i) ? (double)sp.a0I * (double)sp.gam : 0.0;
                const double r2 = uniform(xn2 - wave_sum(part));
                if (r2 <= e2) break;
            }
            // pos = argmax_k cur_k^2, lowest index on ties (maxabs, S/OMP.cu:201-213)
            float bv = -1.f;
            int bi = 0x7fffffff;
#pragma unroll
            for (int m = 0; m < KPL; ++m) {
                const int k = m * 64 + lane;
                const float c = (i == 0) ? a0[m] : al[m];
                const float v = c * c;
                if (k < K && v > bv) { bv = v; bi = k; }
            }
            wave_argmax(bv, bi);
            const int pos = uniform(bi);
            bv = uniform(bv);
            // stop: selected twice, or inner product too small (S/OMP.cu:292,304)
            if (uniform((int)slot[pos]) != 0 || (double)bv < 1e-14) break;
            if (lane == i) sp.ind = pos;
            if (lane == 0) slot[pos] = (unsigned char)(i + 1);  // marked before the pivot test (S/OMP.cu:313)
            if (!chol_append(L, LD, i, pos, G, K, sp, a.pivot_fp64, lane)) break;
            if (lane == i) sp.a0I = a0p[pos];
            wave_lds_fence();
            i++;
            solve_gamma(L, LD, i, sp, lane);
            // alpha = alpha0 - G(:, I) gamma  (addMul, S/OMP.cu:234-247)
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = 0.f;
            int j = 0;
            for (; j + 1 < i; j += 2) {
                const int c0 = readlane(sp.ind, j), c1 = readlane(sp.ind, j + 1);
                const float g0 = readlane(sp.gam, j), g1 = readlane(sp.gam, j + 1);
                const float* __restrict__ p0 = G + (long)c0 * K + lane;
                const float* __restrict__ p1 = G + (long)c1 * K + lane;
                float v0[KPL], v1[KPL];
#pragma unroll
                for (int m = 0; m < KPL; ++m) {
                    const bool ok = (m * 64 + lane) < K;
                    v0[m] = ok ? p0[m * 64] : 0.f;
                    v1[m] = ok ? p1[m * 64] : 0.f;
                }
#pragma unroll
                for (int m = 0; m < KPL; ++m) {
                    al[m] = al[m] + v0[m] * g0;
                    al[m] = al[m] + v1[m] * g1;
                }
            }
            if (j < i) {
                const int c0 = readlane(sp.ind, j);
                const float g0 = readlane(sp.gam, j);
                const float* __restrict__ p0 = G + (long)c0 * K + lane;
#pragma unroll
                for (int m = 0; m < KPL; ++m)
                    if ((m * 64 + lane) < K) al[m] = al[m] + p0[m * 64] * g0;
            }
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = a0[m] - al[m];
        }
        wave_lds_fence();
        write_result(a, s, i, sp, slot, KPL * 64, lane);
        wave_lds_fence();
    }
}

// ---------------------------------------------------------------------------
// Compressed-G form.  The alpha update streams an fp16 copy of G (half the
// bytes of the FP32 form, 16 consecutive atoms per lane read with two 16-byte
// loads per column) and is followed by an exact selection:
//   |alpha~_k - alpha_k| <= eps_k = c1 (g_k ||gamma||_1 + d_k max|gamma|)
//                                   + 3e-8 ||gamma||_1 + 6e-8 |alpha~_k|,
//   c1 = 2^-11 (1 + 2^-11) + (i+2) 2^-24,
// where g_k = max_{j != k} |G_jk| and d_k = |G_kk| (fp16 rounding of each
// entry, FP32 accumulation of i terms, flushed fp16 denormals below 2^-24).
// Every atom whose interval [|alpha~|-eps, |alpha~|+eps] reaches the largest
// lower end is re-evaluated in FP32 as alpha0_k - G(k, I) gamma and the
// arg-max taken over those (lowest index on ties).  More than CMAX candidates
// (or an fp16-unsafe G) falls back to the FP32 update for that step.
// ---------------------------------------------------------------------------
constexpr int CMAX = 24;

// al[m] += g * Gh_col[kb + m] with the column kept packed (two halves per VGPR)
// until each multiply-add.
template <int KPL>
__device__ __forceinline__ void axpy_h16(const _Float16* __restrict__ p, float g, float* al) {
    if constexpr (KPL % 8 == 0) {
        uint4 u[KPL / 8];
#pragma unroll
        for (int q = 0; q < KPL / 8; ++q) u[q] = *reinterpret_cast<const uint4*>(p + 8 * q);
#pragma unroll
        for (int q = 0; q < KPL / 8; ++q) {
            const _Float16* h = reinterpret_cast<const _Float16*>(&u[q]);
#pragma unroll
            for (int e = 0; e < 8; ++e) al[8 * q + e] = fmaf((float)h[e], g, al[8 * q + e]);
        }
    } else {
        uint2 u[KPL / 4];
#pragma unroll
        for (int q = 0; q < KPL / 4; ++q) u[q] = *reinterpret_cast<const uint2*>(p + 4 * q);
#pragma unroll
        for (int q = 0; q < KPL / 4; ++q) {
            const _Float16* h = reinterpret_cast<const _Float16*>(&u[q]);
#pragma unroll
            for (int e = 0; e < 4; ++e) al[4 * q + e] = fmaf((float)h[e], g, al[4 * q + e]);
        }
    }
}

template <int KPL>
__device__ __forceinline__ void load_f32(const float* __restrict__ p, float* out) {
#pragma unroll
    for (int q = 0; q < KPL / 4; ++q) {
        const float4 u = *reinterpret_cast<const float4*>(p + 4 * q);
        out[4 * q] = u.x; out[4 * q + 1] = u.y; out[4 * q + 2] = u.z; out[4 * q + 3] = u.w;
    }
}

template <int KPL, bool EPS>
__global__ __launch_bounds__(256) void omp_wave16_kernel(OmpArgs a) {
    extern __shared__ float smem[];
    constexpr int KP = KPL * 64;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int K = a.K, T = a.T, LD = T + 1;
    float* L = smem + w * (T * LD);
    float* a0s = smem + WPB * T * LD + w * KP;
    float* gbs = smem + WPB * T * LD + WPB * KP;  // g_k, then d_k (shared by the block)
    float* gds = gbs + KP;
    unsigned char* slot = reinterpret_cast<unsigned char*>(gds + KP) + w * KP;
    const float* __restrict__ G = a.G;
    const _Float16* __restrict__ Gh = a.Gh;
    const int Kp = a.Kpad;
    const int kb = lane * KPL;  // this lane's atoms: kb .. kb + KPL - 1

    for (int k = threadIdx.x; k < KP; k += blockDim.x) {
        gbs[k] = (k < K) ? a.gbound[2 * k] : 0.f;
        gds[k] = (k < K) ? a.gbound[2 * k + 1] : 0.f;
    }
    __syncthreads();
    long n_cand = 0, n_fall = 0;

    for (long it = 0;; ++it) {
        long s;
        if constexpr (EPS) {
            int v = 0;
            if (lane == 0) v = atomicAdd(a.work_counter, 1);
            s = readlane(v, 0);
        } else {
            if (it > 0) break;
            s = (long)blockIdx.x * WPB + w;
        }
        if (s >= a.N) break;

        const float* __restrict__ a0p = a.DtX + s * (long)K;
        float al[KPL];
        if (kb + KPL <= K) {
            load_f32<KPL>(a0p + kb, al);
        } else {
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = (kb + m < K) ? a0p[kb + m] : 0.f;
        }
#pragma unroll
        for (int m = 0; m < KPL; ++m) {
            a0s[kb + m] = al[m];
            slot[kb + m] = 0;
        }
        Support sp;
        double xn2 = 0.0, e2 = 0.0;
        if constexpr (EPS) {
            xn2 = (double)a.xnorm2[s];
            e2 = (double)a.eps2[s];
        }
        wave_lds_fence();

        int i = 0;
        while (i < T) {
            if constexpr (EPS) {
                const double part = (lane < i) ? (double)sp.a0I * (double)sp.gam : 0.0;
                const double r2 = uniform(xn2 - wave_sum(part));
                if (r2 <= e2) break;
            }
            int pos;
            float bv;
            bool exact = (i == 0);
            if (!exact) {
                const float ag = fabsf(sp.gam);
                const float gsum = uniform(wave_sum((lane < i) ? ag : 0.f));
                const float gmax = uniform(wave_max((lane < i) ? ag : 0.f));
                const float c1 = 4.8852e-4f + (float)(i + 2) * 6.0e-8f;
                const float c0 = 3.0e-8f * gsum;
                auto eps = [&](int m) {
                    return c1 * (gbs[kb + m] * gsum + gds[kb + m] * gmax) + c0 + 6.0e-8f * fabsf(al[m]);
                };
                float lo = -1.f;
#pragma unroll
                for (int m = 0; m < KPL; ++m)
                    if (kb + m < K) lo = fmaxf(lo, fabsf(al[m]) - eps(m));
                const float Mlow = uniform(wave_max(lo));
                int ncand = 0;
#pragma unroll
                for (int m = 0; m < KPL; ++m)
                    ncand += __popcll(__ballot(kb + m < K && fabsf(al[m]) + eps(m) >= Mlow));
                ncand = uniform(ncand);
                if (ncand > CMAX || !(Mlow == Mlow)) {
                    exact = true;  // recompute alpha in FP32 for this step
                    n_fall++;
#pragma unroll
                    for (int m = 0; m < KPL; ++m) al[m] = 0.f;
                    for (int j = 0; j < i; ++j) {
                        const int c = readlane(sp.ind, j);
                        const float g = readlane(sp.gam, j);
                        float gv[KPL];
                        if (kb + KPL <= K) {
                            load_f32<KPL>(G + (long)c * K + kb, gv);
                        } else {
#pragma unroll
                            for (int m = 0; m < KPL; ++m) gv[m] = (kb + m < K) ? G[(long)c * K + kb + m] : 0.f;
                        }
#pragma unroll
                        for (int m = 0; m < KPL; ++m) al[m] = al[m] + gv[m] * g;
                    }
#pragma unroll
                    for (int m = 0; m < KPL; ++m) al[m] = a0s[kb + m] - al[m];
                } else {
                    n_cand += ncand;
                    float best = -1.f;
                    int bestk = 0x7fffffff;
#pragma unroll
                    for (int m = 0; m < KPL; ++m) {
                        unsigned long long mask = __ballot(kb + m < K && fabsf(al[m]) + eps(m) >= Mlow);
                        while (mask) {
                            const int l = __builtin_ctzll(mask);
                            mask &= mask - 1;
                            const int k = l * KPL + m;
                            // exact FP32 alpha_k = alpha0_k - sum_j G(k, I_j) gamma_j
                            const float part = (lane < i) ? G[(long)sp.ind * K + k] * sp.gam : 0.f;
                            const float ak = a0s[k] - wave_sum(part);
                            const float v2 = uniform(ak * ak);
                            if (v2 > best || (v2 == best && k < bestk)) { best = v2; bestk = k; }
                        }
                    }
                    pos = bestk;
                    bv = best;
                }
            }
            if (exact) {
                // arg-max over exact FP32 values (step 0: alpha0; fallback: FP32 update)
                bv = -1.f;
                int bi = 0x7fffffff;
#pragma unroll
                for (int m = 0; m < KPL; ++m) {
                    const float v = al[m] * al[m];
                    if (kb + m < K && v > bv) { bv = v; bi = kb + m; }
                }
                wave_argmax(bv, bi);
                pos = uniform(bi);
                bv = uniform(bv);
            }
            // stop: selected twice, or inner product too small (S/OMP.cu:292,304)
            if (uniform((int)slot[pos]) != 0 || (double)bv < 1e-14) break;
            if (lane == i) sp.ind = pos;
            if (lane == 0) slot[pos] = (unsigned char)(i + 1);
            if (!chol_append(L, LD, i, pos, G, K, sp, a.pivot_fp64, lane)) break;
            if (lane == i) sp.a0I = a0s[pos];
            wave_lds_fence();
            i++;
            solve_gamma(L, LD, i, sp, lane);
            // alpha~ = alpha0 - Gh(:, I) gamma, fp16 columns of G
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = 0.f;
            for (int j = 0; j < i; ++j) {
                const int c0 = readlane(sp.ind, j);
                const float g0 = readlane(sp.gam, j);
                axpy_h16<KPL>(Gh + (long)c0 * Kp + kb, g0, al);
            }
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = a0s[kb + m] - al[m];
        }
        wave_lds_fence();
        write_result(a, s, i, sp, slot, KP, lane);
        wave_lds_fence();
    }
    if (a.stats && lane == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(a.stats), (unsigned long long)n_cand);
        atomicAdd(reinterpret_cast<unsigned long long*>(a.stats) + 1, (unsigned long long)n_fall);
    }
}

// G (K x K FP32) -> Gh (K columns of Kpad fp16, zero padded) and per-atom
// bounds g_k = max_{j != k} |G_jk|, d_k = |G_kk| (row k across columns j).
// Entries beyond the fp16 range make the bound infinite (always fall back).
__global__ __launch_bounds__(256) void g16_prep_kernel(const float* __restrict__ G, int K, int Kpad,
                                                       _Float16* __restrict__ Gh, float* __restrict__ gbound) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)K * Kpad;
    if (idx < total) {
        const int c = (int)(idx / Kpad), r = (int)(idx % Kpad);
        Gh[idx] = (r < K) ? (_Float16)G[(long)c * K + r] : (_Float16)0.f;
    }
    if (idx < K) {
        const int k = (int)idx;
        float g = 0.f;
        bool big = false;
        for (int j = 0; j < K; ++j) {
            const float v = fabsf(G[(long)j * K + k]);
            big |= !(v < 60000.f);
            if (j != k) g = fmaxf(g, v);
        }
        const float d = fabsf(G[(long)k * K + k]);
        gbound[2 * k] = big ? __builtin_huge_valf() : g;
        gbound[2 * k + 1] = big ? __builtin_huge_valf() : d;
    }
}

__global__ void zero_counter_kernel(int32_t* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

template <int KPL, bool G16>
hipError_t launch_kpl(const OmpArgs& a, hipStream_t st) {
    const size_t lds = G16 ? (size_t)WPB * a.T * (a.T + 1) * sizeof(float) + (size_t)(WPB + 2) * KPL * 64 * sizeof(float) +
                                 (size_t)WPB * KPL * 64
                           : (size_t)WPB * a.T * (a.T + 1) * sizeof(float) + (size_t)WPB * KPL * 64;
    void (*kern_eps)(OmpArgs);
    void (*kern)(OmpArgs);
    if constexpr (G16) {
        kern_eps = omp_wave16_kernel<KPL, true>;
        kern = omp_wave16_kernel<KPL, false>;
    } else {
        kern_eps = omp_wave_kernel<KPL, true>;
        kern = omp_wave_kernel<KPL, false>;
    }
    if (a.eps2) {
        int dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int per_cu = 0;
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern_eps, 256, lds);
        if (per_cu < 1) per_cu = 1;
        long want = (a.N + WPB - 1) / WPB;
        long grid = (long)cus * per_cu;
        if (grid > want) grid = want;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, st, a.work_counter);
        hipLaunchKernelGGL(kern_eps, dim3((unsigned)grid), dim3(256), lds, st, a);
    } else {
        const long grid = (a.N + WPB - 1) / WPB;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, st, a);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_omp(const OmpArgs& a_in, hipStream_t st) {
    OmpArgs a = a_in;
    if (a.N <= 0) return hipSuccess;
    const int kpl = (a.K + 63) / 64;
    // the compressed-G kernel reads 16-byte-aligned runs of KPL atoms per lane
    if (a.Gh && (kpl < 4 || kpl > 16 || (a.K & 3) != 0)) a.Gh = nullptr;
    if (a.Gh) {
        if (kpl <= 4) return launch_kpl<4, true>(a, st);
        if (kpl <= 8) return launch_kpl<8, true>(a, st);
        return launch_kpl<16, true>(a, st);
    }
    if (kpl <= 1) return launch_kpl<1, false>(a, st);
    if (kpl <= 2) return launch_kpl<2, false>(a, st);
    if (kpl <= 4) return launch_kpl<4, false>(a, st);
    if (kpl <= 8) return launch_kpl<8, false>(a, st);
    if (kpl <= 16) return launch_kpl<16, false>(a, st);
    if (kpl <= 32) return launch_kpl<32, false>(a, st);
    return hipErrorInvalidValue;
}

int g16_kpad(int K) {
    const int kpl = (K + 63) / 64;
    const int k = kpl <= 4 ? 4 : kpl <= 8 ? 8 : 16;
    return 64 * k;
}

bool g16_supported(int K) {
    const int kpl = (K + 63) / 64;
    return kpl >= 4 && kpl <= 16 && (K & 3) == 0;
}

hipError_t launch_g16_prep(const float* G, int K, _Float16* Gh, float* gbound, hipStream_t st) {
    const int Kpad = g16_kpad(K);
    const long total = (long)K * Kpad;
    const long n = total > K ? total : K;
    hipLaunchKernelGGL(g16_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, G, K, Kpad, Gh, gbound);
    return hipGetLastError();
}

}  // namespace ksvd
