// adjoint.hip -- reverse-mode derivative of the fixed-support least squares
//   gamma_I = (D_I^T D_I)^{-1} D_I^T x
// (no reference counterpart).  With g = dL/dgamma_I, z = (D_I^T D_I)^{-1} g and
// r = x - D_I gamma_I:
//   dL/dx   = D_I z
//   dL/dD_I = r z^T - (D_I z) gamma_I^T      (zero outside the support)
// One wave per signal: slot j of the support in lane j, the Gram block
// G_II and its Cholesky factor in LDS (pitch T+1), lane-parallel triangular
// solves, and gD accumulated with float atomics on the support columns.
#include "common.hpp"

namespace ksvd {

namespace {

constexpr int WPB = 4;

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(256) void adjoint_kernel(const float* __restrict__ D, const float* __restrict__ X,
                                                      const int32_t* __restrict__ idx,
                                                      const float* __restrict__ val,
                                                      const int32_t* __restrict__ cnt,
                                                      const float* __restrict__ gval, float* __restrict__ gD,
                                                      float* __restrict__ gX, int n, int N, int T) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int LD = T + 1;
    float* A = smem + w * (T * LD);  // Gram block, overwritten by its Cholesky factor (lower)
    const long s = (long)blockIdx.x * WPB + w;
    if (s >= N) return;
    const int m = cnt[s];
    const int il = (lane < m) ? idx[s * T + lane] : 0;
    const float gam = (lane < m) ? val[s * T + lane] : 0.f;
    const float gl = (lane < m) ? gval[s * T + lane] : 0.f;

    // G_II(i, j) = d_i^T d_j -> A[j*LD + i]   (column j, row i)
    for (int i = 0; i < m; ++i) {
        const int ci = readlane(il, i);
        for (int j = 0; j <= i; ++j) {
            const int cj = readlane(il, j);
            float sacc = 0.f;
            for (int t = lane; t < n; t += 64) sacc += D[(long)ci * n + t] * D[(long)cj * n + t];
            sacc = wave_sum(sacc);
            if (lane == 0) { A[j * LD + i] = sacc; A[i * LD + j] = sacc; }
        }
    }
    wave_lds_fence();
    // right-looking Cholesky: lane i owns row i
    float diag = 1.f;
    for (int k = 0; k < m; ++k) {
        const float akk = A[k * LD + k];
        const float lkk = sqrtf(akk);
        if (lane == k) diag = lkk;
        float lik = 0.f;
        if (lane > k && lane < m) {
            lik = A[k * LD + lane] / lkk;
            A[k * LD + lane] = lik;
        }
        wave_lds_fence();
        // trailing update A(i, j) -= L(i,k) L(j,k) for k < j <= i
        if (lane > k && lane < m)
            for (int j = k + 1; j <= lane; ++j) A[j * LD + lane] -= lik * A[k * LD + j];
        wave_lds_fence();
    }
    // z = L^-T L^-1 g
    float y = 0.f;
    {
        float b = gl;
        for (int j = 0; j < m; ++j) {
            const float yj = readlane(b, j) / readlane(diag, j);
            if (lane == j) y = yj;
            if (lane > j && lane < m) b -= A[j * LD + lane] * yj;
        }
    }
    float z = 0.f;
    {
        float b = y;
        for (int l = m - 1; l >= 0; --l) {
            const float zl = readlane(b, l) / readlane(diag, l);
            if (lane == l) z = zl;
            if (lane < l) b -= A[lane * LD + l] * zl;
        }
    }
    // r = x - D_I gamma, dz = D_I z; gX = dz; gD(:, I_j) += r z_j - dz gamma_j
    for (int t0 = 0; t0 < n; t0 += 64) {
        const int t = t0 + lane;
        float r = (t < n) ? X[s * n + t] : 0.f;
        float dz = 0.f;
        for (int j = 0; j < m; ++j) {
            const int cj = readlane(il, j);
            const float d = (t < n) ? D[(long)cj * n + t] : 0.f;
            r -= d * readlane(gam, j);
            dz += d * readlane(z, j);
        }
        if (t < n) {
            gX[s * n + t] = dz;
            for (int j = 0; j < m; ++j) {
                const int cj = readlane(il, j);
                atomicAdd(&gD[(long)cj * n + t], r * readlane(z, j) - dz * readlane(gam, j));
            }
        }
    }
}

__global__ void zero_kernel(float* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace

hipError_t launch_omp_adjoint(const float* D, const float* X, const int32_t* idx, const float* val,
                              const int32_t* cnt, const float* gval, float* gD, float* gX, int n, int K, int N, int T,
                              hipStream_t st) {
    const long nd = (long)n * K;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, gD, nd);
    const size_t lds = (size_t)WPB * T * (T + 1) * sizeof(float);
    hipLaunchKernelGGL(adjoint_kernel, dim3((unsigned)((N + WPB - 1) / WPB)), dim3(256), lds, st, D, X, idx, val, cnt,
                       gval, gD, gX, n, N, T);
    return hipGetLastError();
}

}  // namespace ksvd
