// adjoint.hip -- reverse-mode derivative of the fixed-support least squares
//   gamma_I = (D_I^T D_I)^{-1} D_I^T x
// (no reference counterpart).  With g = dL/dgamma_I, z = (D_I^T D_I)^{-1} g and
// r = x - D_I gamma_I:
//   dL/dx   = D_I z
//   dL/dD_I = r z^T - (D_I z) gamma_I^T      (zero outside the support)
//
// Two launches, no atomics, so the gradients are the same bit for bit on
// every run:
//   1. one wave per signal (slot j of the support in lane j): the Cholesky
//      factor L of G_II in LDS (pitch T+1) -- either the factor Batch-OMP
//      built and stored (S/OMP.cu:316-336 builds the same factor and drops it)
//      or, in the low-memory mode, G_II recomputed from D and factored here --
//      then z = L^-T L^-1 g by lane-parallel triangular solves, gX_s = D_I z,
//      and r_s and z into scratch;
//   2. one workgroup per atom k: gD(:, k) = sum over the signals using k, in
//      ascending signal order, of r_s z_{s,k} - gX_s gamma_{s,k}, each wave a
//      contiguous part of the atom's list and the waves' partial sums added
//      in wave order.
#include "common.hpp"

namespace ksvd {

namespace {

constexpr int ATOM_WAVES = 16;  // waves per atom in the per-atom pass

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one wave per signal, blockDim.x / 64 signals per workgroup
template <bool FACT>
__global__ __launch_bounds__(256) void adjoint_signal_kernel(const float* __restrict__ D, const float* __restrict__ X,
                                                             const int32_t* __restrict__ idx,
                                                             const float* __restrict__ val,
                                                             const int32_t* __restrict__ cnt,
                                                             const float* __restrict__ Lfac,
                                                             const float* __restrict__ gval, float* __restrict__ gX,
                                                             float* __restrict__ R, float* __restrict__ zbuf, int n,
                                                             int N, int T) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int LD = T + 1;
    float* A = smem + w * (T * LD);  // L(row i, col j) at A[j*LD + i]
    const long s = (long)blockIdx.x * (blockDim.x >> 6) + w;
    if (s >= N) return;
    const int m = cnt[s];
    const int il = (lane < m) ? idx[s * T + lane] : 0;
    const float gam = (lane < m) ? val[s * T + lane] : 0.f;
    const float gl = (lane < m) ? gval[s * T + lane] : 0.f;
    float diag = 1.f;  // L(lane, lane)
    if constexpr (FACT) {
        // packed column by column with pitch T: L(r, c) at c*T - c(c-1)/2 + (r - c)
        const float* Lp = Lfac + s * (long)(T * (T + 1) / 2);
        for (int c = 0; c < m; ++c) {
            const int r = c + lane;
            if (r < m) A[c * LD + r] = Lp[c * T - ((c * (c - 1)) >> 1) + lane];
        }
        wave_lds_fence();
        if (lane < m) diag = A[lane * LD + lane];
    } else {
        // G_II(i, j) = d_i^T d_j -> A[j*LD + i]
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
        for (int k = 0; k < m; ++k) {
            const float lkk = sqrtf(A[k * LD + k]);
            if (lane == k) diag = lkk;
            float lik = 0.f;
            if (lane > k && lane < m) {
                lik = A[k * LD + lane] / lkk;
                A[k * LD + lane] = lik;
            }
            wave_lds_fence();
            if (lane > k && lane < m)
                for (int j = k + 1; j <= lane; ++j) A[j * LD + lane] -= lik * A[k * LD + j];
            wave_lds_fence();
        }
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
    if (lane < T) zbuf[s * T + lane] = (lane < m) ? z : 0.f;
    // r = x - D_I gamma and gX = D_I z
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
            R[s * n + t] = r;
        }
    }
}

// gD(:, k) = sum_{(s, j) on atom k's list} r_s z_{s,j} - gX_s gamma_{s,j}.
// Wave w of the workgroup takes list entries [w L / W, (w+1) L / W); each lane
// keeps 4 elements of the column per 256 (t = t0 + lane + 64 q).
__global__ __launch_bounds__(64 * ATOM_WAVES) void adjoint_atom_kernel(const int32_t* __restrict__ list_ptr,
                                                                      const int32_t* __restrict__ list_pos,
                                                                      const float* __restrict__ val,
                                                                      const float* __restrict__ zbuf,
                                                                      const float* __restrict__ R,
                                                                      const float* __restrict__ gX,
                                                                      float* __restrict__ gD, int n, int T) {
    __shared__ float part[ATOM_WAVES][256];
    const int k = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int beg = list_ptr[k], len = list_ptr[k + 1] - beg;
    const int e0 = beg + (int)((long)len * w / ATOM_WAVES), e1 = beg + (int)((long)len * (w + 1) / ATOM_WAVES);
    for (int t0 = 0; t0 < n; t0 += 256) {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (int b = e0; b < e1; b += 64) {
            // 64 entries lane-parallel: position, gamma and z
            const int e = b + lane;
            const int pos = (e < e1) ? list_pos[e] : 0;
            const float g = (e < e1) ? val[pos] : 0.f;
            const float zz = (e < e1) ? zbuf[pos] : 0.f;
            const int sl = pos / T;
            const int nb = min(64, e1 - b);
            constexpr int U = 4;  // rows in flight
            for (int u0 = 0; u0 < nb; u0 += U) {
                float rv[U][4], xv[U][4], gu[U], zu[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool ok = u0 + u < nb;
                    const long s = readlane(sl, ok ? u0 + u : 0);
                    gu[u] = ok ? readlane(g, u0 + u) : 0.f;
                    zu[u] = ok ? readlane(zz, u0 + u) : 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int t = t0 + lane + 64 * q;
                        rv[u][q] = (ok && t < n) ? R[s * n + t] : 0.f;
                        xv[u][q] = (ok && t < n) ? gX[s * n + t] : 0.f;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] += rv[u][q] * zu[u] - xv[u][q] * gu[u];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) part[w][lane + 64 * q] = acc[q];
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = t0 + lane + 64 * q;
                float sum = 0.f;
                for (int v = 0; v < ATOM_WAVES; ++v) sum += part[v][lane + 64 * q];
                if (t < n) gD[(long)k * n + t] = sum;
            }
        }
        __syncthreads();
    }
}

// d = D_in / ||D_in|| per column; gin = (gout - d (d^T gout)) / ||D_in||.
__global__ __launch_bounds__(256) void normcols_adjoint_kernel(const float* __restrict__ Din,
                                                               const float* __restrict__ gout,
                                                               float* __restrict__ gin, int n, int K) {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= K) return;
    const int lane = threadIdx.x & 63;
    const float* d = Din + (long)k * n;
    const float* g = gout + (long)k * n;
    float ss = 0.f;
    double dg = 0.0;  // the projection is removed from g: its sum in FP64
    for (int t = lane; t < n; t += 64) {
        ss += d[t] * d[t];
        dg += (double)d[t] * (double)g[t];
    }
    ss = wave_sum(ss);  // the forward normcols' sum of squares
    dg = wave_sum(dg);
    const float inv = 1.0f / sqrtf(ss);
    const double c = dg * (double)inv * (double)inv;  // d_n^T g / ||d|| with d_n = d * inv
    for (int t = lane; t < n; t += 64)
        gin[(long)k * n + t] = (float)(((double)g[t] - (double)d[t] * c) * (double)inv);
}

// out[s, j] = dense[s, idx[s, j]] for j < cnt[s], 0 beyond.
__global__ __launch_bounds__(256) void gather_compact_kernel(const float* __restrict__ dense,
                                                             const int32_t* __restrict__ idx,
                                                             const int32_t* __restrict__ cnt, int K, int N, int T,
                                                             float* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)N * T) return;
    const long s = e / T;
    const int j = (int)(e - s * T);
    out[e] = (j < cnt[s]) ? dense[s * K + idx[e]] : 0.f;
}

}  // namespace

hipError_t launch_omp_adjoint(const float* D, const float* X, const int32_t* idx, const float* val,
                              const int32_t* cnt, const float* Lfac, const float* gval, float* gD, float* gX,
                              const int32_t* list_ptr, const int32_t* list_pos, float* R, float* zbuf, int n,
                              int K, int N, int T, hipStream_t st) {
    // four signals per workgroup while their factors fit 64 kB of LDS (T <= 63), else two
    const int wpb = (size_t)4 * T * (T + 1) * sizeof(float) <= 65536 ? 4 : 2;
    const size_t lds = (size_t)wpb * T * (T + 1) * sizeof(float);
    const dim3 grid((unsigned)((N + wpb - 1) / wpb));
    if (Lfac)
        hipLaunchKernelGGL(adjoint_signal_kernel<true>, grid, dim3(64 * wpb), lds, st, D, X, idx, val, cnt, Lfac, gval, gX,
                           R, zbuf, n, N, T);
    else
        hipLaunchKernelGGL(adjoint_signal_kernel<false>, grid, dim3(64 * wpb), lds, st, D, X, idx, val, cnt, Lfac, gval,
                           gX, R, zbuf, n, N, T);
    hipLaunchKernelGGL(adjoint_atom_kernel, dim3((unsigned)K), dim3(64 * ATOM_WAVES), 0, st, list_ptr, list_pos, val,
                       zbuf, R, gX, gD, n, T);
    return hipGetLastError();
}

hipError_t launch_normcols_adjoint(const float* Din, const float* gout, float* gin, int n, int K, hipStream_t st) {
    hipLaunchKernelGGL(normcols_adjoint_kernel, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, st, Din, gout, gin, n, K);
    return hipGetLastError();
}

hipError_t launch_gather_compact(const float* dense, const int32_t* idx, const int32_t* cnt, int K, int N, int T,
                                 float* out, hipStream_t st) {
    const long ne = (long)N * T;
    hipLaunchKernelGGL(gather_compact_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, dense, idx, cnt, K,
                       N, T, out);
    return hipGetLastError();
}

}  // namespace ksvd
