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
            xn2 = a.xnorm2[s];
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

__global__ void zero_counter_kernel(int32_t* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

template <int KPL>
hipError_t launch_kpl(const OmpArgs& a, hipStream_t st) {
    const size_t lds = (size_t)WPB * a.T * (a.T + 1) * sizeof(float) + (size_t)WPB * KPL * 64;
    if (a.eps2) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int per_cu = 0;
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, omp_wave_kernel<KPL, true>, 256, lds);
        if (e != hipSuccess) return e;
        if (per_cu < 1) per_cu = 1;
        long want = (a.N + WPB - 1) / WPB;
        long grid = (long)cus * per_cu;
        if (grid > want) grid = want;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, st, a.work_counter);
        hipLaunchKernelGGL((omp_wave_kernel<KPL, true>), dim3((unsigned)grid), dim3(256), lds, st, a);
    } else {
        const long grid = (a.N + WPB - 1) / WPB;
        hipLaunchKernelGGL((omp_wave_kernel<KPL, false>), dim3((unsigned)grid), dim3(256), lds, st, a);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_zero_counter(int32_t* p, hipStream_t st) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_batch_omp(const OmpArgs& a, hipStream_t st) {
    if (a.N <= 0) return hipSuccess;
    const int kpl = (a.K + 63) / 64;
    if (kpl <= 1) return launch_kpl<1>(a, st);
    if (kpl <= 2) return launch_kpl<2>(a, st);
    if (kpl <= 4) return launch_kpl<4>(a, st);
    if (kpl <= 8) return launch_kpl<8>(a, st);
    if (kpl <= 16) return launch_kpl<16>(a, st);
    if (kpl <= 32) return launch_kpl<32>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace ksvd
