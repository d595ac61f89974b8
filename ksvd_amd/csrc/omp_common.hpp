// omp_common.hpp -- per-signal Batch-OMP steps shared by the wave kernels.
//
// Layout of one signal's state in a wave64 (support slot j in lane j, T <= 64):
//   ind   atom of slot j                 a0I   alpha0[ind_j]
//   y     (L^-1 alpha0_I)_j              gam   gamma_j
//   diag  L_jj
// and the Cholesky factor L (T x T, lower) in LDS, element (row r, col c) at
// L[c*LD + r] with an odd pitch LD = T + 1: column reads (forward sweep) are
// contiguous and row reads (backward sweep) are stride-LD, both conflict free.
#pragma once

#include "common.hpp"

namespace ksvd {

struct Support {
    int ind = -1;
    float a0I = 0.f, y = 0.f, gam = 0.f, diag = 1.f;
};

// Makes this wave's earlier LDS writes visible to all of its lanes and keeps
// the compiler from moving LDS accesses across the call.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends atom `pos` as slot i (i >= 0): for i > 0 the new row of L is
// w = L^-1 G(I, pos) (backsubt_indexed_append_L, S/OMP.cu:169-182) with pivot
// d = 1 - ||w||^2 (sumSquareValues, S/OMP.cu:221-228), recomputed in FP64
// when d < pivot_fp64.  Returns false, leaving L untouched, when d <= 1e-14
// (S/OMP.cu:331-333).  Slot i's atom index is recorded by the caller.
__device__ __forceinline__ bool chol_append(float* L, int LD, int i, int pos, const float* __restrict__ G, int K,
                                            Support& sp, float pivot_fp64, int lane) {
    if (i == 0) {
        if (lane == 0) { L[0] = 1.f; sp.diag = 1.f; }  // S/OMP.cu:317
        return true;
    }
    const float bsrc = (lane < i) ? G[(long)pos * K + sp.ind] : 0.f;
    float b = bsrc, wv = 0.f;
    for (int j = 0; j < i; ++j) {
        const float wj = readlane(b, j) / readlane(sp.diag, j);
        if (lane == j) wv = wj;
        if (lane > j && lane < i) b = b - L[j * LD + lane] * wj;
    }
    const float ss = uniform(wave_sum((lane < i) ? wv * wv : 0.f));
    const float d = 1.f - ss;
    double dd = (double)d;
    if (d < pivot_fp64) {
        double b64 = (double)bsrc, w64 = 0.0;
        for (int j = 0; j < i; ++j) {
            const double wj = readlane(b64, j) / (double)readlane(sp.diag, j);
            if (lane == j) w64 = wj;
            if (lane > j && lane < i) b64 = b64 - (double)L[j * LD + lane] * wj;
        }
        dd = uniform(1.0 - wave_sum((lane < i) ? w64 * w64 : 0.0));
        wv = (float)w64;
    }
    if (dd <= 1e-14) return false;
    const float lii = (float)sqrt(dd);
    if (lane < i) L[lane * LD + i] = wv;
    if (lane == i) { L[i * LD + i] = lii; sp.diag = lii; }
    return true;
}

// With slots 0..i1-1 accepted (a0I of slot i1-1 set): y_{i1-1} by one forward
// row (backsubt_indexed_L, S/OMP.cu:147-159: earlier rows are unchanged) and
// gamma = L^-T y by a lane-parallel backward sweep (backsubst_Lt,
// S/OMP.cu:126-138).
__device__ __forceinline__ void solve_gamma(const float* L, int LD, int i1, Support& sp, int lane) {
    const int r = i1 - 1;
    const float part = (lane < r) ? L[lane * LD + r] * sp.y : 0.f;
    const float t = wave_sum(part);
    const float yn = (readlane(sp.a0I, r) - t) / readlane(sp.diag, r);
    if (lane == r) sp.y = yn;
    float rhs = sp.y;
    for (int l = i1 - 1; l >= 0; --l) {
        const float gl = readlane(rhs, l) / readlane(sp.diag, l);
        if (lane == l) sp.gam = gl;
        if (lane < l) rhs = rhs - L[lane * LD + l] * gl;
    }
}

// Writes one signal's result: compact (idx/val/cnt) and/or the reference
// layout (SELECTED_ATOMS slot map from the LDS byte map, c).
__device__ __forceinline__ void write_result(const OmpArgs& a, long s, int i, const Support& sp,
                                             const unsigned char* slot, int kpl64, int lane) {
    const int T = a.T, K = a.K;
    if (a.idx && lane < T) {
        a.idx[s * T + lane] = (lane < i) ? sp.ind : -1;
        a.val[s * T + lane] = (lane < i) ? sp.gam : 0.f;
    }
    if (a.cnt && lane == 0) a.cnt[s] = i;
    if (a.sel) {
        float* selp = a.sel + s * (long)K;
        for (int k = lane; k < kpl64; k += 64)
            if (k < K) selp[k] = (float)slot[k];
    }
    if (a.c && lane < i) a.c[s * T + lane] = sp.gam;
}

}  // namespace ksvd
