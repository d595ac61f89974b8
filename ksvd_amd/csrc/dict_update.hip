// dict_update.hip -- approximate K-SVD dictionary update, error and
// dictionary clearing on CDNA4.
//
// Reference: DictionaryUpdate::updateDictionary (S/DictionaryUpdate.cu:45-203)
// launches ~12 kernels per atom including a full n x N x K cuBLAS GEMM for the
// unused-atom test, and builds its per-atom lists with atomicAdd slots
// (nondeterministic order).  This engine works on the compact Gamma
// (idx/val/cnt) and the restricted residuals R = X - D Gamma, formed once per
// sweep (du_residual4_kernel).  The per-atom lists are built once per sweep in
// ascending signal order by a two-level counting sort (no order-dependent
// atomics).  For n <= 1024 the sweep is pipelined by one atom, one launch
// per atom:
//   du_ba_kernel(jj)       every workgroup first forms a = E gamma^T of the
//                          atom at sweep position jj from the integer
//                          accumulator set the previous launch filled, and
//                          d_new = a / ||a||; then phase B of that atom
//                          (gamma_new = d_new^T (r_s + d_p gamma_s), residual
//                          update), fused with phase A of the atom at jj + 1
//                          (its block partials of a, added as fixed-point
//                          int64 words: integer sums, so the result does not
//                          depend on the order of the additions), and the
//                          column commit of the atom at jj - 1, one element
//                          per workgroup.  An unused atom takes the special
//                          case (sparse errors from R, no GEMM): block maxima
//                          here, the pick and the replacement in the next
//                          launch.
// Above n = 1024 the unpipelined du_phase_a_kernel, du_a_reduce_kernel (a
// fixed-order sum of FP32 block partials) and du_phase_b_kernel do the same
// work, three launches per atom.  No kernel synchronises across workgroups
// (a device-scope fence per workgroup costs about 0.1 ms per launch here):
// every cross-workgroup result goes through a kernel boundary.  All
// data-dependent branching is device-side (early exits on N_p), so the host
// loop never synchronises and the sequence is graph-capturable.
#include <algorithm>

#include "common.hpp"

namespace ksvd {

namespace {

constexpr int RES_BLOCKS = 2048;   // grid of the residual initialisation

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- state reset: S/KSVD.cu:442-457 ---------------------------------------
__global__ void du_init_kernel(int32_t* unused, int32_t* usc, int32_t* bitmap, int32_t* replaced, int N, int K) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) { unused[i] = (int32_t)i; bitmap[i] = 1; }
    if (i < K) replaced[i] = 0;
    if (i == 0) *usc = 0;
}

__global__ void fill_i32_kernel(int32_t* p, long n, int32_t v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_f32_kernel(float* p, long n, float v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- per-atom lists (replaces sprow_kernel, S/DictionaryUpdate.cu:378-395) --
// chunk_hist[c*K + k] = number of entries of atom k among signals of chunk c.
__global__ __launch_bounds__(256) void du_hist_kernel(const int32_t* __restrict__ idx, const int32_t* __restrict__ cnt,
                                                      int N, int T, int K, int CH, int32_t* __restrict__ chunk_hist) {
    extern __shared__ int32_t h[];
    for (int k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const long s0 = (long)blockIdx.x * CH;
    const long s1 = min((long)N, s0 + CH);
    for (long s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const int c = cnt[s];
        for (int j = 0; j < c; ++j) atomicAdd(&h[idx[s * T + j]], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) chunk_hist[(long)blockIdx.x * K + k] = h[k];
}

// Per atom: exclusive scan of chunk_hist over chunks (in place) and the total.
__global__ __launch_bounds__(256) void du_scan_chunks_kernel(int32_t* __restrict__ chunk_hist, int nch, int K,
                                                             int32_t* __restrict__ count) {
    __shared__ int32_t buf[256];
    const int k = blockIdx.x;
    int32_t carry = 0;
    for (int c0 = 0; c0 < nch; c0 += 256) {
        const int c = c0 + threadIdx.x;
        const int32_t v = (c < nch) ? chunk_hist[(long)c * K + k] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {  // Hillis-Steele inclusive scan
            const int32_t t = (threadIdx.x >= o) ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (c < nch) chunk_hist[(long)c * K + k] = carry + buf[threadIdx.x] - v;
        carry += buf[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) count[k] = carry;
}

// ptr = exclusive scan of count (single block).
__global__ __launch_bounds__(1024) void du_scan_ptr_kernel(const int32_t* __restrict__ count, int K,
                                                           int32_t* __restrict__ ptr) {
    __shared__ int32_t buf[1024];
    int32_t carry = 0;
    for (int k0 = 0; k0 < K; k0 += 1024) {
        const int k = k0 + threadIdx.x;
        const int32_t v = (k < K) ? count[k] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int32_t t = (threadIdx.x >= o) ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (k < K) ptr[k] = carry + buf[threadIdx.x] - v;
        carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) ptr[K] = carry;
}

// One wave per chunk: signals in ascending order, lane j = slot j; the atoms of
// one signal are distinct, so the LDS cursors never collide within a step.
// With `list_rec` (the pipelined sweep) each entry (signal s, atom p) is one
// 16-byte record (one store, one load): its position s*T + slot, whether s
// uses the atom updated just before p in the sweep order `perm` (bit 31: its
// phase-B pass already covers s), its coefficient, and the coefficient of s
// on the atom updated just after p (0 when s does not use it).  Otherwise
// only the position goes to list_pos.
constexpr int32_t kPrevFlag = (int32_t)0x80000000;
__device__ __forceinline__ int32_t entry_pos(int32_t e) { return e & 0x7FFFFFFF; }

// Places the entries of signal s (count cc, lane j < cc holds slot j: atom kk,
// coefficient vl) at the atoms' LDS cursors.
__device__ __forceinline__ void du_fill_place(long s, int lane, int cc, int kk, float vl, int T, int K,
                                              int32_t* __restrict__ cur, int32_t* __restrict__ rk,
                                              int32_t* __restrict__ at, int32_t* __restrict__ list_pos,
                                              int4* __restrict__ list_rec) {
    if (list_rec) {
        const int r = (lane < cc) ? rk[kk] : 0;
        if (lane < cc) at[r] = lane + 1;
        wave_lds_fence();
        const int pv = (lane < cc && r > 0) ? at[r - 1] : 0;
        const int nx = (lane < cc && r + 1 < K) ? at[r + 1] : 0;
        const float gx = __shfl(vl, nx > 0 ? nx - 1 : 0);
        wave_lds_fence();
        if (lane < cc) {
            at[r] = 0;
            const int p = cur[kk];
            cur[kk] = p + 1;
            list_rec[p] = make_int4((int32_t)(s * T + lane) | (pv ? kPrevFlag : 0), __float_as_int(vl),
                                    __float_as_int(nx ? gx : 0.f), 0);
        }
    } else if (lane < cc) {
        const int p = cur[kk];
        cur[kk] = p + 1;
        list_pos[p] = (int32_t)(s * T + lane);
    }
    wave_lds_fence();
}

__global__ __launch_bounds__(64) void du_fill_kernel(const int32_t* __restrict__ idx, const int32_t* __restrict__ cnt,
                                                     int N, int T, int K, int CH,
                                                     const int32_t* __restrict__ chunk_off,
                                                     const int32_t* __restrict__ ptr,
                                                     int32_t* __restrict__ list_pos,
                                                     const float* __restrict__ val,
                                                     const uint32_t* __restrict__ perm, int4* __restrict__ list_rec) {
    extern __shared__ int32_t cur[];
    int32_t* rk = cur + K;  // sweep rank of each atom
    int32_t* at = rk + K;   // per signal: slot + 1 of its atom at each rank, 0 elsewhere
    const int lane = threadIdx.x;
    const long c = blockIdx.x;
    for (int k = lane; k < K; k += 64) cur[k] = ptr[k] + chunk_off[c * K + k];
    if (list_rec) {
        for (int k = lane; k < K; k += 64) {
            rk[perm[k]] = k;
            at[k] = 0;
        }
    }
    wave_lds_fence();
    const long s0 = c * CH;
    const int ns = (int)max(0L, min((long)N, s0 + CH) - s0);
    // FB signals per batch: their counts, slots and values are fetched with
    // independent loads (all T slots, masked by the count afterwards), so a
    // batch costs one memory latency instead of one per signal
    constexpr int FB = 16;
    for (int b = 0; b < ns; b += FB) {
        const int cl = (lane < FB && b + lane < ns) ? cnt[s0 + b + lane] : 0;
        int kq[FB];
        float vq[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const bool ok = b + u < ns && lane < T;
            kq[u] = ok ? idx[(s0 + b + u) * T + lane] : 0;
            vq[u] = (ok && list_rec) ? val[(s0 + b + u) * T + lane] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            if (b + u >= ns) continue;  // (uniform) past the chunk's last signal
            const long s = s0 + b + u;
            const int cc = readlane(cl, u);
            const int kk = (lane < cc) ? kq[u] : 0;
            du_fill_place(s, lane, cc, kk, vq[u], T, K, cur, rk, at, list_pos, list_rec);
        }
    }
}

// ---- restricted residuals -------------------------------------------------
// Sparse reconstruction error of one signal by one wave: ||x - D gamma_s||^2.
__device__ __forceinline__ float wave_signal_error(const float* __restrict__ X, const float* __restrict__ D,
                                                   const int32_t* __restrict__ idx,
                                                   const float* __restrict__ val, int c, long s, int n, int T,
                                                   int lane) {
    const int il = (lane < c) ? idx[s * T + lane] : 0;
    const float vl = (lane < c) ? val[s * T + lane] : 0.f;
    float e = 0.f;
    for (int t0 = 0; t0 < n; t0 += 64) {
        const int t = t0 + lane;
        float xh = 0.f;
        for (int j = 0; j < c; ++j) {
            const int col = readlane(il, j);
            const float g = readlane(vl, j);
            if (t < n) xh += D[(long)col * n + t] * g;
        }
        if (t < n) {
            const float d = X[s * n + t] - xh;
            e += d * d;
        }
    }
    return wave_sum(e);
}

// R = X - D Gamma, one wave per signal (grid-stride).
__global__ __launch_bounds__(256) void du_residual_kernel(const float* __restrict__ X, const float* __restrict__ D,
                                                          const int32_t* __restrict__ idx,
                                                          const float* __restrict__ val,
                                                          const int32_t* __restrict__ cnt, int n, int N, int T,
                                                          float* __restrict__ R) {
    const int lane = threadIdx.x & 63;
    for (long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6); s < N; s += (long)gridDim.x * 4) {
        const int c = cnt[s];
        const int il = (lane < c) ? idx[s * T + lane] : 0;
        const float vl = (lane < c) ? val[s * T + lane] : 0.f;
        for (int t0 = 0; t0 < n; t0 += 64) {
            const int t = t0 + lane;
            float xh = 0.f;
            for (int j = 0; j < c; ++j) {
                const int col = readlane(il, j);
                const float g = readlane(vl, j);
                if (t < n) xh += D[(long)col * n + t] * g;
            }
            if (t < n) R[s * n + t] = X[s * n + t] - xh;
        }
    }
}


// (defined with the pipelined sweep below)
__device__ void du_special_partial(const float* __restrict__ R, const int32_t* __restrict__ unused,
                                   const int32_t* __restrict__ usc, int n, int N, int M, float* __restrict__ bpart,
                                   float* rv, int* ri);
__device__ __forceinline__ float du_anorm(int jj, const float* __restrict__ anorm);
__device__ __forceinline__ void du_commit_column(int jj, const uint32_t* __restrict__ perm,
                                                 const int32_t* __restrict__ count, float* __restrict__ D,
                                                 const float* __restrict__ abuf, const float* __restrict__ anorm,
                                                 int n);

// Per-atom phase A (the sweep for n > 1024: per atom phase A, the reduction
// du_a_reduce_kernel, phase B).  Block 0 first commits the new column of the
// atom at position jj - 1.
//   N_p > 0:  block partials of a = sum_{s in I_p} gamma_s (r_s + d_p gamma_s)
//             ( = E_p gamma^T, S/Multiplication.cu:177-193 computes the same
//             vector as Y_p - D v + d_p ||gamma||^2 ), rows of apart padded to
//             n rounded up to 4; the reduction forms a and ||a||.
//   N_p = 0:  the unused-atom special case (S/DictionaryUpdate.cu:238-302,
//             SCupdate :406-427): block maxima of the errors, finished by the
//             reduction's workgroup 0.
__global__ __launch_bounds__(256) void du_phase_a_kernel(int jj, const uint32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ count,
                                                         const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ list_pos,
                                                         const float* __restrict__ val, const float* __restrict__ X,
                                                         float* __restrict__ D, const float* __restrict__ R,
                                                         int32_t* __restrict__ unused, int32_t* __restrict__ usc,
                                                         int32_t* __restrict__ bitmap, int32_t* __restrict__ replaced,
                                                         int n, int N, int T, int M, float* __restrict__ apart,
                                                         float* __restrict__ bpart, const float* __restrict__ abuf,
                                                         const float* __restrict__ anorm) {
    extern __shared__ float sh[];  // 4 x n (per-wave partial a)
    __shared__ float rv[4];
    __shared__ int ri[4];
    const int p = (int)perm[jj];
    const int Np = count[p];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int G = gridDim.x;
    if (jj > 0 && blockIdx.x == 0) du_commit_column(jj - 1, perm, count, D, abuf, anorm, n);
    if (Np == 0) {
        du_special_partial(R, unused, usc, n, N, M, bpart, rv, ri);
        return;
    }
    const int beg = ptr[p];
    const float* dp = D + (long)p * n;
    for (int t0 = 0; t0 < n; t0 += 256) {
        // lanes cover t0 + lane + 64 q, q < 4
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        float dv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = t0 + lane + 64 * q;
            dv[q] = (t < n) ? dp[t] : 0.f;
        }
        for (int m = blockIdx.x * 4 + w; m < Np; m += G * 4) {
            const int pos = list_pos[beg + m];
            const long s = pos / T;
            const float g = val[pos];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = t0 + lane + 64 * q;
                if (t < n) acc[q] += g * (R[s * n + t] + dv[q] * g);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = t0 + lane + 64 * q;
            if (t < n) sh[w * n + t] = acc[q];
        }
    }
    __syncthreads();
    const int np4 = (n + 3) & ~3;
    for (int t = threadIdx.x; t < n; t += blockDim.x)
        apart[(long)blockIdx.x * np4 + t] = ((sh[t] + sh[n + t]) + sh[2 * n + t]) + sh[3 * n + t];
}

// Per-atom phase B: gamma_new = d_new^T (r_s + d_p gamma_s) and the residual
// update r_s <- r_s + d_p gamma_s - d_new gamma_new for every signal on the
// list (S/Multiplication.cu:198-225 evaluates the same gamma_new as
// (a^T x - u^T Gamma_s + u_p gamma_s) / ||a||).  d_p = d_new
// (S/Multiplication.cu:221) is written by the next phase A, or
// du_tail_kernel after the last atom: every block of this launch reads d_p.
__global__ __launch_bounds__(256) void du_phase_b_kernel(int jj, const uint32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ count,
                                                         const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ list_pos, float* __restrict__ val,
                                                         float* __restrict__ D, float* __restrict__ R,
                                                         const float* __restrict__ abuf,
                                                         const float* __restrict__ anorm, int n, int T) {
    extern __shared__ float sdn[];  // d_new = a / ||a|| (n floats)
    const int p = (int)perm[jj];
    const int Np = count[p];
    if (Np == 0) return;
    {
        const float nrm = du_anorm(jj, anorm);
        for (int t = threadIdx.x; t < n; t += blockDim.x) sdn[t] = abuf[(long)(jj & 1) * n + t] / nrm;
        __syncthreads();
    }
    const float* dnew = sdn;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int beg = ptr[p];
    const float* dp = D + (long)p * n;
    if (n <= 256) {
        float dv[4], dn[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = lane + 64 * q;
            dv[q] = (t < n) ? dp[t] : 0.f;
            dn[q] = (t < n) ? dnew[t] : 0.f;
        }
        for (int m = blockIdx.x * 4 + w; m < Np; m += G * 4) {
            const int pos = list_pos[beg + m];
            const long s = pos / T;
            const float g = val[pos];
            float e[4];
            float dot = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = lane + 64 * q;
                e[q] = (t < n) ? R[s * n + t] + dv[q] * g : 0.f;
                dot += dn[q] * e[q];
            }
            const float gn = wave_sum(dot);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = lane + 64 * q;
                if (t < n) R[s * n + t] = e[q] - dn[q] * gn;
            }
            if (lane == 0) val[pos] = gn;
        }
    } else {
        for (int m = blockIdx.x * 4 + w; m < Np; m += G * 4) {
            const int pos = list_pos[beg + m];
            const long s = pos / T;
            const float g = val[pos];
            float dot = 0.f;
            for (int t = lane; t < n; t += 64) dot += dnew[t] * (R[s * n + t] + dp[t] * g);
            const float gn = wave_sum(dot);
            for (int t = lane; t < n; t += 64) R[s * n + t] = R[s * n + t] + dp[t] * g - dnew[t] * gn;
            if (lane == 0) val[pos] = gn;
        }
    }
}

// ---- per-atom phases, n <= 1024: batched, latency-hiding, pipelined ---------
// Same arithmetic as du_phase_a/b_kernel above, organised for memory-level
// parallelism: each wave loads up to 64 of its list entries (position and
// gamma) lane-parallel, then streams batches of residual rows (one float4 per
// lane per 256 columns); 8 waves per block.  Phase A's block partials are
// reduced by du_a_reduce_kernel in a fixed order (deterministic).  Phase B
// does not write d_p: the next launch commits the new column (thread 0 of
// workgroups 0 .. n-1, one element each), and du_tail_kernel does it for the
// last atom of the sweep; nothing reads d_p in between.
constexpr int DU_WAVES = 8;
constexpr int FQ = 2;   // rows per batch in du_ba_kernel, divided by the float4 chunks per row (59 VGPRs at n <= 256)
constexpr int FQA = 8;  // residual rows in flight per wave in du_a_kernel (phase A of the first atom)
constexpr int DU_RB = 16;  // workgroups of the phase A reduction

__device__ __forceinline__ float4 ld4(const float* __restrict__ p, int t, int n, bool vec) {
    if (vec) return (t < n) ? *reinterpret_cast<const float4*>(p + t) : make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4(t < n ? p[t] : 0.f, t + 1 < n ? p[t + 1] : 0.f, t + 2 < n ? p[t + 2] : 0.f,
                       t + 3 < n ? p[t + 3] : 0.f);
}

__device__ __forceinline__ void st4(float* __restrict__ p, int t, int n, bool vec, float4 v) {
    if (vec) {
        if (t < n) *reinterpret_cast<float4*>(p + t) = v;
        return;
    }
    if (t < n) p[t] = v.x;
    if (t + 1 < n) p[t + 1] = v.y;
    if (t + 2 < n) p[t + 2] = v.z;
    if (t + 3 < n) p[t + 3] = v.w;
}

// ld4 with a streaming (non-temporal) load: phase B reads a row for the last
// time this stretch of the sweep
__device__ __forceinline__ float4 ld4_nt(const float* __restrict__ p, int t, int n, bool vec) {
    if (vec) {
        if (t >= n) return make_float4(0.f, 0.f, 0.f, 0.f);
        const f32x4 u = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p + t));
        return make_float4(u[0], u[1], u[2], u[3]);
    }
    return ld4(p, t, n, vec);
}

// st4 with a streaming (non-temporal) store: a residual row written by phase
// B is not read again for about T / K of a sweep, so it should not displace
// the rows phase A just read, which the next launch's phase B reads again.
__device__ __forceinline__ void st4_nt(float* __restrict__ p, int t, int n, bool vec, float4 v) {
    if (vec) {
        if (t < n) __builtin_nontemporal_store(f32x4{v.x, v.y, v.z, v.w}, reinterpret_cast<f32x4*>(p + t));
        return;
    }
    st4(p, t, n, vec, v);
}

__device__ __forceinline__ float4 f4fma(float a, float4 x, float4 y) {
    return make_float4(fmaf(a, x.x, y.x), fmaf(a, x.y, y.y), fmaf(a, x.z, y.z), fmaf(a, x.w, y.w));
}

// ||a|| of the atom at sweep position jj from the DU_RB per-block partial sums
// of squares (fixed order: every consumer computes the same value).
__device__ __forceinline__ float du_anorm(int jj, const float* __restrict__ anorm) {
    const float* q = anorm + (jj & 1) * DU_RB;
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < DU_RB; ++r) s += q[r];
    return sqrtf(s);
}

// Copies the new column d_new = a / ||a|| of the atom updated at sweep
// position jj (if it was a regular update) into D.
__device__ __forceinline__ void du_commit_column(int jj, const uint32_t* __restrict__ perm,
                                                 const int32_t* __restrict__ count, float* __restrict__ D,
                                                 const float* __restrict__ abuf, const float* __restrict__ anorm,
                                                 int n) {
    const int pp = (int)perm[jj];
    if (count[pp] > 0) {
        const float nrm = du_anorm(jj, anorm);
        for (int t = threadIdx.x; t < n; t += blockDim.x) D[(long)pp * n + t] = abuf[(long)(jj & 1) * n + t] / nrm;
    }
}

// The unused-atom special case (S/DictionaryUpdate.cu:238-302, SCupdate
// :406-427) for atom p, in two parts.  du_special_partial, run by every block
// of the launch that meets the unused atom: errors ||r_b||^2 of the first
// min(N - used, M) live unused-list entries and each block's arg-max (lowest
// index on ties) in bpart (rv, ri: one entry per wave).  du_special_finish,
// run by one workgroup of the next launch (the kernel boundary publishes
// bpart; a last-block ticket would need a device-scope fence in every block,
// ~0.1 ms per atom here): the largest block maximum, then d_p = x / ||x||,
// swap-remove, replaced, bitmap.
__device__ void du_special_partial(const float* __restrict__ R, const int32_t* __restrict__ unused,
                                   const int32_t* __restrict__ usc, int n, int N, int M, float* __restrict__ bpart,
                                   float* rv, int* ri) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int G = gridDim.x;
    const int counter = *usc;
    const int lim = min(M, N - counter);
    // every signal already replaced an atom this sweep: no candidate, the atom
    // stays as it is (the reference would read before its unused list)
    if (lim <= 0) return;
    float bv = 0.f;  // the reference seeds its reduction with value 0 / index 0
    int bi = 0;
    for (int b = blockIdx.x * nw + w; b < lim; b += G * nw) {
        const long sig = unused[b];
        float e = 0.f;
        for (int t = lane; t < n; t += 64) {
            const float r = R[sig * n + t];
            e += r * r;
        }
        e = wave_sum_dpp(e);
        if (e > bv) { bv = e; bi = b; }  // b ascending per wave: keeps the lowest on ties
    }
    if (lane == 0) { rv[w] = bv; ri[w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = rv[0];
        int i = ri[0];
        for (int q = 1; q < nw; ++q)
            if (rv[q] > v || (rv[q] == v && ri[q] < i)) { v = rv[q]; i = ri[q]; }
        bpart[2 * blockIdx.x] = v;
        bpart[2 * blockIdx.x + 1] = __int_as_float(i);
    }
}

// G: the grid of the launch that ran du_special_partial.  Uniform control
// flow (one whole workgroup, any size up to 1024); ri and wred hold 16
// entries.
__device__ void du_special_finish(int p, int G, const float* __restrict__ X, float* __restrict__ D,
                                  int32_t* __restrict__ unused, int32_t* __restrict__ usc,
                                  int32_t* __restrict__ bitmap, int32_t* __restrict__ replaced, int n, int N, int M,
                                  const float* __restrict__ bpart, int* ri, float* wred) {
    const int counter = *usc;
    const int lim = min(M, N - counter);
    if (lim <= 0) return;  // the partial step wrote nothing either
    // the largest block maximum, lowest entry among equal ones (the order of
    // the comparisons does not matter: every block maximum is >= 0 and the
    // entries of different blocks differ), in parallel over the workgroup
    float v = -1.f;
    int i = 0x7fffffff;
    for (int q = threadIdx.x; q < G; q += blockDim.x) {
        const float vq = bpart[2 * q];
        const int iq = __float_as_int(bpart[2 * q + 1]);
        if (vq > v || (vq == v && iq < i)) { v = vq; i = iq; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float vo = __shfl_xor(v, off);
        const int io = __shfl_xor(i, off);
        if (vo > v || (vo == v && io < i)) { v = vo; i = io; }
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) { wred[w] = v; ri[w] = i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float vb = wred[0];
        int ib = ri[0];
        for (int q = 1; q < nw; ++q)
            if (wred[q] > vb || (wred[q] == vb && ri[q] < ib)) { vb = wred[q]; ib = ri[q]; }
        ri[0] = ib;
    }
    __syncthreads();
    const int best = ri[0];
    const long sig = unused[best];
    float sq = 0.f;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const float x = X[sig * n + t];
        sq += x * x;
    }
    sq = block_sum(sq, wred);
    // the pick is a zero signal (every candidate error 0): x / ||x|| is
    // undefined, the atom stays as it is
    if (sq == 0.f) return;
    const float nrm = sqrtf(sq);
    for (int t = threadIdx.x; t < n; t += blockDim.x) D[(long)p * n + t] = X[sig * n + t] / nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unused[best] = unused[N - counter - 1];
        *usc = counter + 1;
        replaced[p] = 1;
        bitmap[sig] = 0;
    }
}

// ---- the pipelined sweep's cross-workgroup sums (n <= 1024) ---------------
// Each workgroup adds its block partial of a, component by component, into
// bucket blockIdx mod kDuBuckets of an accumulator set as two int64 words,
// h = rint(2^8 f) and l = rint(2^60 (f - 2^-8 h)): exact for
// 2^-37 <= |f| < 2^44 (smaller parts below 2^-60 are rounded away).  Integer
// sums do not depend on the order of the additions, so every run gives the
// same a, and a is the sum of all block partials rounded once.  A partial
// outside that range (or NaN) sets the set's flag word instead, and the
// consumers then sum the FP32 block partials (also written, to ap) in block
// order.  Sets go by sweep position mod 3: launch jj consumes set jj % 3,
// fills set (jj + 1) % 3 and clears set (jj + 2) % 3; the FP32 partials and
// the special case's block records by position parity.  Every workgroup of the
// consuming launch sums the kDuBuckets buckets itself (no separate reduction
// launch), in the same order, so all of them hold the same a and ||a||.
constexpr float kDuFixMax = 0x1p44f;

__device__ __forceinline__ void du_publish(float f, int t, int np4, unsigned long long* __restrict__ set,
                                           float* __restrict__ ap) {
    ap[t] = f;
    if (f == 0.f) return;  // adds nothing
    if (!(fabsf(f) < kDuFixMax)) {
        atomicOr(set + 2L * kDuBuckets * np4, 1ull);
        return;
    }
    const double fd = (double)f;
    const double hd = rint(fd * 256.0);
    const long long h = (long long)hd;
    const long long l = (long long)rint((fd - hd * 0.00390625) * 0x1p60);  // both steps exact in FP64
    const int b = blockIdx.x & (kDuBuckets - 1);
    atomicAdd(set + (2L * b) * np4 + t, (unsigned long long)h);
    atomicAdd(set + (2L * b + 1) * np4 + t, (unsigned long long)l);
}

// a into sa (LDS, n floats) and ||a||, the same values in every workgroup;
// ap: the G FP32 block partials of the position (read after an overflow
// only).  Uniform control flow; red holds 16 floats.
__device__ float du_form_a(const unsigned long long* __restrict__ set, const float* __restrict__ ap, int G, int n,
                           int np4, float* sa, float* red) {
    const bool ovf = set[2L * kDuBuckets * np4] != 0ull;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        float a;
        if (!ovf) {
            long long hs = 0, ls = 0;
#pragma unroll
            for (int b = 0; b < kDuBuckets; ++b) {
                hs += (long long)set[(2L * b) * np4 + t];
                ls += (long long)set[(2L * b + 1) * np4 + t];
            }
            a = (float)((double)hs * 0.00390625 + (double)ls * 0x1p-60);
        } else {
            a = 0.f;
            for (int b = 0; b < G; ++b) a += ap[(long)b * np4 + t];
        }
        sa[t] = a;
    }
    float sq = 0.f;
    for (int t = threadIdx.x; t < n; t += blockDim.x) sq += sa[t] * sa[t];
    return sqrtf(block_sum(sq, red));
}

// The special case in the pipelined sweep: du_special_partial_at runs in the
// launch that meets the unused atom, with the counter C of the unused list
// given and entry sb read as sv (the swap-remove of the previous position's
// pick, which workgroup 0 of the same launch writes); each block's record
// {largest error, its entry, ||x||^2 of that entry's signal, 0} goes to bp.
// The next launch's workgroups pick from the records alone
// (du_special_pick), so no workgroup but the one applying the pick
// (du_special_apply, workgroup 0) reads the list entry it rewrites.
__device__ void du_special_partial_at(const float* __restrict__ R, const float* __restrict__ X,
                                      const int32_t* __restrict__ unused, int C, int sb, int sv, int n, int N, int M,
                                      float* __restrict__ bp, float* rv, int* ri) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int G = gridDim.x;
    const int lim = min(M, N - C);
    if (lim <= 0) return;  // no candidate: the atom stays as it is
    float bv = 0.f;        // the reference seeds its reduction with value 0 / index 0
    int bi = 0;
    for (int b = blockIdx.x * nw + w; b < lim; b += G * nw) {
        const long sig = (b == sb) ? sv : unused[b];
        float e = 0.f;
        for (int t = lane; t < n; t += 64) {
            const float r = R[sig * n + t];
            e += r * r;
        }
        e = wave_sum_dpp(e);
        if (e > bv) { bv = e; bi = b; }  // b ascending per wave: keeps the lowest on ties
    }
    if (lane == 0) { rv[w] = bv; ri[w] = bi; }
    __syncthreads();
    if (w == 0) {
        float v = rv[0];
        int i = ri[0];
        for (int q = 1; q < nw; ++q)
            if (rv[q] > v || (rv[q] == v && ri[q] < i)) { v = rv[q]; i = ri[q]; }
        const long sig = (i == sb) ? sv : unused[i];
        float xs = 0.f;
        for (int t = lane; t < n; t += 64) {
            const float x = X[sig * n + t];
            xs += x * x;
        }
        xs = wave_sum_dpp(xs);
        if (lane == 0) *reinterpret_cast<float4*>(bp + 4 * blockIdx.x) = make_float4(v, __int_as_float(i), xs, 0.f);
    }
}

// The entry that replaces the atom (the largest block maximum, lowest entry
// among equal ones), or -1: no candidate left, or the pick is a zero signal
// (x / ||x|| undefined).  G: the grid of the launch that wrote the records.
// Uniform control flow; ri holds 16 ints, wred 32 floats.
__device__ int du_special_pick(int G, const float* __restrict__ bp, int C, int N, int M, int* ri, float* wred) {
    if (min(M, N - C) <= 0) return -1;
    float v = -1.f, xs = 0.f;
    int i = 0x7fffffff;
    for (int q = threadIdx.x; q < G; q += blockDim.x) {
        const float4 r = *reinterpret_cast<const float4*>(bp + 4 * q);
        const int iq = __float_as_int(r.y);
        if (r.x > v || (r.x == v && iq < i)) { v = r.x; i = iq; xs = r.z; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float vo = __shfl_xor(v, off), xo = __shfl_xor(xs, off);
        const int io = __shfl_xor(i, off);
        if (vo > v || (vo == v && io < i)) { v = vo; i = io; xs = xo; }
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { wred[w] = v; ri[w] = i; wred[16 + w] = xs; }
    __syncthreads();
    float vb = wred[0], xb = wred[16];
    int ib = ri[0];
    for (int q = 1; q < nw; ++q)
        if (wred[q] > vb || (wred[q] == vb && ri[q] < ib)) { vb = wred[q]; ib = ri[q]; xb = wred[16 + q]; }
    __syncthreads();
    return xb != 0.f ? ib : -1;
}

// One workgroup: d_p = x / ||x|| for the picked entry's signal, then the
// swap-remove of the entry (C: the list's counter before the pick),
// replaced[p] and bitmap (S/DictionaryUpdate.cu:238-302).
__device__ void du_special_apply(int p, int best, int C, const float* __restrict__ X, float* __restrict__ D,
                                 int32_t* __restrict__ unused, int32_t* __restrict__ bitmap,
                                 int32_t* __restrict__ replaced, int n, int N, float* wred) {
    const long sig = unused[best];
    float sq = 0.f;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const float x = X[sig * n + t];
        sq += x * x;
    }
    const float nrm = sqrtf(block_sum(sq, wred));
    for (int t = threadIdx.x; t < n; t += blockDim.x) D[(long)p * n + t] = X[sig * n + t] / nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unused[best] = unused[N - C - 1];
        replaced[p] = 1;
        bitmap[sig] = 0;
    }
}

// Phase A of the atom at sweep position 0 (its set 0 and FP32 block
// partials, or the special case's block records), and the unused list's
// counter at the start of the sweep into dstate[0].
template <int NCH>
__global__ __launch_bounds__(512) void du_a_kernel(const uint32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ count, const int32_t* __restrict__ ptr,
                                                   const int4* __restrict__ list_rec, const float* __restrict__ X,
                                                   const float* __restrict__ D, const float* __restrict__ R,
                                                   const int32_t* __restrict__ unused, const int32_t* __restrict__ usc,
                                                   int n, int N, int T, int M, float* __restrict__ apart,
                                                   float* __restrict__ bpart, unsigned long long* __restrict__ dacc,
                                                   int32_t* __restrict__ dstate) {
    __shared__ float4 part[DU_WAVES][64 * NCH];  // per-wave partial a
    __shared__ float rv[DU_WAVES];
    __shared__ int ri[DU_WAVES];
    const int p = (int)perm[0];
    const int Np = count[p];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int G = gridDim.x;
    const bool vec = (n & 3) == 0;
    const int C = *usc;
    if (blockIdx.x == 0 && threadIdx.x == 0) dstate[0] = C;
    if (Np == 0) {
        du_special_partial_at(R, X, unused, C, -1, 0, n, N, M, bpart, rv, ri);  // picked by du_ba_kernel(1)
        return;
    }
    // a = sum_{s in I_p} gamma_s (r_s + d_p gamma_s)
    const int beg = ptr[p];
    const int gw = blockIdx.x * DU_WAVES + w, GW = G * DU_WAVES;
    const float* dp = D + (long)p * n;
    float4 dv[NCH], acc[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        dv[c] = ld4(dp, 256 * c + 4 * lane, n, vec);
        acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    for (int base = gw; base < Np; base += 64 * GW) {
        const int mk = base + lane * GW;
        const int4 rec = (mk < Np) ? list_rec[beg + mk] : make_int4(0, 0, 0, 0);
        const int pos = entry_pos(rec.x);
        const float gl = __int_as_float(rec.y);
        const int sl = pos / T;
        const int nk = min(64, (Np - base + GW - 1) / GW);
        for (int k = 0; k < nk; k += FQA) {
            float4 r[FQA][NCH];
            float g[FQA];
#pragma unroll
            for (int u = 0; u < FQA; ++u) {
                const bool ok = k + u < nk;
                const long s = readlane(sl, ok ? k + u : 0);
                g[u] = ok ? readlane(gl, k + u) : 0.f;
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    r[u][c] = ok ? ld4(R + s * n, 256 * c + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < FQA; ++u)
#pragma unroll
                for (int c = 0; c < NCH; ++c) acc[c] = f4fma(g[u], f4fma(g[u], dv[c], r[u][c]), acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) part[w][64 * c + lane] = acc[c];
    __syncthreads();
    // block partial, rows padded to np4 = n rounded up to 4 floats
    const int np4 = (n + 3) & ~3;
    const float* pf = reinterpret_cast<const float*>(&part[0][0]);
    for (int t = threadIdx.x; t < np4; t += blockDim.x) {
        float sacc = 0.f;
#pragma unroll
        for (int q = 0; q < DU_WAVES; ++q) sacc += pf[q * 256 * NCH + t];
        du_publish(sacc, t, np4, dacc, apart + (long)blockIdx.x * np4);
    }
}

// Sum of phase A's G block partials, a = sum_b apart[b], in a fixed order,
// spread over DU_RB workgroups by column: workgroup r owns float4 column
// groups [r*ng4/DU_RB, (r+1)*ng4/DU_RB); thread = (column group, slice), slice
// sl sums partials sl, sl+S, sl+2S, ...; slices are combined by a pairwise
// tree in LDS.  Writes a into abuf[jj & 1] and workgroup r's sum of a^2 into
// anorm[(jj & 1)*DU_RB + r]; consumers form ||a|| and d_new = a/||a||
// (S/Multiplication.cu:221).  A separate launch instead of a last-block
// ticket keeps device-scope fences out of phase A.  It runs for an unused
// atom too (nothing reads a or ||a|| of an unused atom), so it starts with
// no dependent load.
__global__ __launch_bounds__(512) void du_a_reduce_kernel(int jj, const float* __restrict__ apart,
                                                          float* __restrict__ abuf, float* __restrict__ anorm, int n,
                                                          int G, int fpos, const int4* __restrict__ sweep,
                                                          const float* __restrict__ X, float* __restrict__ D,
                                                          int32_t* __restrict__ unused, int32_t* __restrict__ usc,
                                                          int32_t* __restrict__ bitmap,
                                                          int32_t* __restrict__ replaced, int N, int M,
                                                          const float* __restrict__ bpart) {
    __shared__ float4 red[512];
    __shared__ int fri[16];
    __shared__ float fwred[16];
    // workgroup 0 first finishes the special case of sweep position fpos
    // when its atom is unused (begun by the previous launch)
    if (fpos >= 0 && blockIdx.x == 0) {
        const int4 d = sweep[2 * fpos];
        if (d.y == 0) du_special_finish(d.x, G, X, D, unused, usc, bitmap, replaced, n, N, M, bpart, fri, fwred);
    }
    const int np4 = (n + 3) & ~3;
    const int ng4 = np4 / 4;
    const int r = blockIdx.x;
    const int c0 = r * ng4 / DU_RB, c1 = (r + 1) * ng4 / DU_RB;
    int ncs = 1;
    while (ncs < (ng4 + DU_RB - 1) / DU_RB) ncs *= 2;  // column-group slots (<= 16 for n <= 1024)
    const int S = 512 / ncs;                            // slices per column group
    const int cl = threadIdx.x % ncs, sl = threadIdx.x / ncs;
    const int cg = c0 + cl;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cg < c1) {
        // RU partials in flight per thread (the partials come from other
        // XCDs' L2s, so each load is a long round trip), added in order b
        constexpr int RU = 8;
        for (int b0 = sl; b0 < G; b0 += RU * S) {
            float4 v4[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int b = b0 + u * S;
                v4[u] = (b < G) ? *reinterpret_cast<const float4*>(apart + (long)b * np4 + 4 * cg)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < RU; ++u)
                if (b0 + u * S < G) acc = make_float4(acc.x + v4[u].x, acc.y + v4[u].y, acc.z + v4[u].z, acc.w + v4[u].w);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int half = 256; half >= 64; half >>= 1) {  // pairwise tree over slices, across waves
        if ((int)threadIdx.x < half) {
            const float4 u = red[threadIdx.x], w4 = red[threadIdx.x + half];
            red[threadIdx.x] = make_float4(u.x + w4.x, u.y + w4.y, u.z + w4.z, u.w + w4.w);
        }
        __syncthreads();
    }
    if (threadIdx.x >= 64) return;  // the same tree's levels below 64 (ncs <= 16) inside wave 0
    float4 v = red[threadIdx.x];
    for (int half = 32; half >= ncs; half >>= 1) {
        const float4 w4 = make_float4(__shfl_down(v.x, half), __shfl_down(v.y, half), __shfl_down(v.z, half),
                                      __shfl_down(v.w, half));
        if ((int)threadIdx.x < half) v = make_float4(v.x + w4.x, v.y + w4.y, v.z + w4.z, v.w + w4.w);
    }
    float sq = 0.f;
    if ((int)threadIdx.x < ncs && cg < c1) {
        const float4 a4 = v;
        const float av[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int t = 4 * cg + e;
            if (t < n) {
                abuf[(long)(jj & 1) * n + t] = av[e];
                sq += av[e] * av[e];
            }
        }
    }
    sq = wave_sum(sq);  // only wave 0 holds squares
    if (threadIdx.x == 0) anorm[(jj & 1) * DU_RB + r] = sq;
}

// Phase B of the atom at sweep position jj fused with phase A of the atom at
// jj + 1 (the pipelined sweep).  Phase B (S/Multiplication.cu:198-225): for
// s in I_p, e = r_s + d_p gamma_s, gamma_new = d_new^T e, r_s <- e - d_new
// gamma_new, with d_new = a / ||a|| formed by every workgroup from set
// jj % 3.  A signal the next atom q also uses contributes gamma_q (r_s + d_q
// gamma_q) to q's a from its updated residual, right here (the entry record
// holds gamma_q); every other signal of I_q is unaffected by this update and
// contributes from R.  Block partials of q's a go to set (jj + 1) % 3.
// First, the atom at position jj - 1: its new column (written by workgroup 0
// of the previous launch to dbuf) is committed, one element per workgroup, or
// its special case is picked (by every workgroup that needs the unused
// list's state after it) and applied by workgroup 0.  An unused atom p
// (N_p = 0, jj > 0; position 0 is handled by du_a_kernel) writes its special
// case's block records last.
template <int NCH>
__global__ __launch_bounds__(512) void du_ba_kernel(int jj, const int4* __restrict__ list_rec,
                                                    const float* __restrict__ X, float* __restrict__ val,
                                                    float* __restrict__ D, float* __restrict__ R,
                                                    int32_t* __restrict__ unused, int32_t* __restrict__ bitmap,
                                                    int32_t* __restrict__ replaced, int n, int N, int T, int M,
                                                    float* __restrict__ apart, float* __restrict__ bpart,
                                                    unsigned long long* __restrict__ dacc, float* __restrict__ dbuf,
                                                    int32_t* __restrict__ dstate, const int4* __restrict__ sweep) {
    __shared__ float4 part[DU_WAVES][64 * NCH];  // per-wave partial a of the next atom
    __shared__ float sa[256 * NCH];              // a of this atom
    __shared__ float rv[DU_WAVES], fwred[32];
    __shared__ int ri[DU_WAVES], fri[16];
    // this position's atom p and the next one q, their list lengths and offsets
    const int4 d0 = sweep[2 * jj], d1 = sweep[2 * jj + 1];
    const int p = d0.x, Np = d0.y, q = d0.w, Nq = d1.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int G = gridDim.x;
    const bool vec = (n & 3) == 0;
    const int gw = blockIdx.x * DU_WAVES + w, GW = G * DU_WAVES;
    const int np4 = (n + 3) & ~3;
    const long sw = du_set_words(np4);
    const long apw = (long)kDuBlocksMax * np4;  // FP32 block partials per position parity
    // the atom at position jj - 1; C: the unused list's counter after it, and
    // the entry sb that its pick rewrites to sv
    int C, sb = -1, sv = 0;
    if (jj > 0) {
        const int4 dp = sweep[2 * (jj - 1)];
        C = dstate[(jj - 1) & 1];
        if (dp.y > 0) {
            if (threadIdx.x == 0 && (int)blockIdx.x < n)
                for (int t = blockIdx.x; t < n; t += G) D[(long)dp.x * n + t] = dbuf[(long)((jj - 1) & 1) * n + t];
        } else if (blockIdx.x == 0 || Np == 0) {
            const int best = du_special_pick(G, bpart + (long)((jj - 1) & 1) * 4 * kDuBlocksMax, C, N, M, fri, fwred);
            if (best >= 0) {
                sb = best;
                sv = unused[N - C - 1];  // workgroup 0 writes only entry best
                if (blockIdx.x == 0) du_special_apply(dp.x, best, C, X, D, unused, bitmap, replaced, n, N, fwred);
                C += 1;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) dstate[jj & 1] = C;
    } else {
        C = dstate[0];
    }
    // d_new = a / ||a|| (every workgroup; workgroup 0 also keeps it for the commit)
    float nrm = 1.f;
    if (Np > 0) {
        nrm = du_form_a(dacc + (jj % 3) * sw, apart + (jj & 1) * apw, G, n, np4, sa, fwred);
        if (blockIdx.x == 0)
            for (int t = threadIdx.x; t < n; t += blockDim.x) dbuf[(long)(jj & 1) * n + t] = sa[t] / nrm;
    }
    float4 dq[NCH], acc[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        dq[c] = (Nq > 0) ? ld4(D + (long)q * n, 256 * c + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
        acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // per wave, in step: a batch of up to FB phase-B entries of p and up to FB
    // phase-A entries of q, all rows loaded before either is used
    constexpr int FB = FQ / NCH > 0 ? FQ / NCH : 1;
    float4 dv[NCH], dn[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        dv[c] = (Np > 0) ? ld4(D + (long)p * n, 256 * c + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 a4 = (Np > 0) ? ld4(sa, 256 * c + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
        dn[c] = make_float4(a4.x / nrm, a4.y / nrm, a4.z / nrm, a4.w / nrm);  // d_new = a / ||a||
    }
    const int bp = d0.z, bq = d1.y;
    const int Nmax = max(Np, Nq);
    for (int base = gw; base < Nmax; base += 64 * GW) {
        const int mk = base + lane * GW;
        const int4 rb = (mk < Np) ? list_rec[bp + mk] : make_int4(0, 0, 0, 0);
        const int posb = entry_pos(rb.x);
        const float gb = __int_as_float(rb.y);
        const float xb = __int_as_float(rb.z);
        const int slb = posb / T;
        const int nkb = (base < Np) ? min(64, (Np - base + GW - 1) / GW) : 0;
        const int4 ra = (mk < Nq) ? list_rec[bq + mk] : make_int4(kPrevFlag, 0, 0, 0);
        const int32_t ent = ra.x;
        const int sla = entry_pos(ent) / T;
        const int sha = ent < 0;  // shared with p: added by the phase-B part
        const float ga = __int_as_float(ra.y);
        const int nka = (base < Nq) ? min(64, (Nq - base + GW - 1) / GW) : 0;
        for (int k = 0; k < max(nkb, nka); k += FB) {
            float4 e[FB][NCH], r[FB][NCH];
            float g2[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                const bool okb = k + u < nkb;
                const long s = readlane(slb, okb ? k + u : 0);
                const float g = okb ? readlane(gb, k + u) : 0.f;
#pragma unroll
                for (int c = 0; c < NCH; ++c)  // e = r_s + d_p gamma_s
                    e[u][c] = okb ? f4fma(g, dv[c], ld4_nt(R + s * n, 256 * c + 4 * lane, n, vec))
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
                const bool oka = k + u < nka && !readlane(sha, k + u < nka ? k + u : 0);
                const long sa = readlane(sla, oka ? k + u : 0);
                g2[u] = oka ? readlane(ga, k + u) : 0.f;
#pragma unroll
                for (int c = 0; c < NCH; ++c)
                    r[u][c] = oka ? ld4(R + sa * n, 256 * c + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            // phase B: gamma_new = d_new^T e, r_s <- e - d_new gamma_new, and
            // gamma_q (r_s + d_q gamma_q) for the signals q also uses
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                if (k + u < nkb) {
                    float dot = 0.f;
#pragma unroll
                    for (int c = 0; c < NCH; ++c) {
                        dot = fmaf(dn[c].x, e[u][c].x, dot);
                        dot = fmaf(dn[c].y, e[u][c].y, dot);
                        dot = fmaf(dn[c].z, e[u][c].z, dot);
                        dot = fmaf(dn[c].w, e[u][c].w, dot);
                    }
                    dot = wave_sum_dpp(dot);
                    const long s = readlane(slb, k + u);
                    const float gq = readlane(xb, k + u);
#pragma unroll
                    for (int c = 0; c < NCH; ++c) {
                        const float4 rn = f4fma(-dot, dn[c], e[u][c]);
                        st4_nt(R + s * n, 256 * c + 4 * lane, n, vec, rn);
                        if (gq != 0.f) acc[c] = f4fma(gq, f4fma(gq, dq[c], rn), acc[c]);
                    }
                    if (lane == 0) val[readlane(posb, k + u)] = dot;
                }
            }
            // phase A of q over its signals not shared with p
#pragma unroll
            for (int u = 0; u < FB; ++u)
#pragma unroll
                for (int c = 0; c < NCH; ++c) acc[c] = f4fma(g2[u], f4fma(g2[u], dq[c], r[u][c]), acc[c]);
        }
    }
    if (Nq > 0) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) part[w][64 * c + lane] = acc[c];
        __syncthreads();
        // block partial, rows padded to np4 = n rounded up to 4 floats
        const float* pf = reinterpret_cast<const float*>(&part[0][0]);
        unsigned long long* setq = dacc + ((jj + 1) % 3) * sw;
        float* apq = apart + ((jj + 1) & 1) * apw + (long)blockIdx.x * np4;
        for (int t = threadIdx.x; t < np4; t += blockDim.x) {
            float sacc = 0.f;
#pragma unroll
            for (int qq = 0; qq < DU_WAVES; ++qq) sacc += pf[qq * 256 * NCH + t];
            du_publish(sacc, t, np4, setq, apq);
        }
    }
    // clear the set of position jj + 2 (last read by the previous launch)
    {
        unsigned long long* setz = dacc + ((jj + 2) % 3) * sw;
        for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < sw; i += (long)G * blockDim.x) setz[i] = 0ull;
    }
    // an unused atom: picked by the next launch (du_ba_kernel(jj + 1), or du_tail_pipe_kernel)
    if (Np == 0 && jj > 0)
        du_special_partial_at(R, X, unused, C, sb, sv, n, N, M, bpart + (long)(jj & 1) * 4 * kDuBlocksMax, rv, ri);
}

// The last sweep position jj: its new column, or its special case (picked
// and applied here); then the unused list's counter after the sweep.
__global__ void du_tail_pipe_kernel(int jj, const int4* __restrict__ sweep, float* __restrict__ D,
                                    const float* __restrict__ dbuf, int n, int G, const float* __restrict__ X,
                                    int32_t* __restrict__ unused, int32_t* __restrict__ usc,
                                    int32_t* __restrict__ bitmap, int32_t* __restrict__ replaced, int N, int M,
                                    const float* __restrict__ bpart, const int32_t* __restrict__ dstate) {
    __shared__ int fri[16];
    __shared__ float fwred[32];
    const int4 d = sweep[2 * jj];
    int C = dstate[jj & 1];
    if (d.y > 0) {
        for (int t = threadIdx.x; t < n; t += blockDim.x) D[(long)d.x * n + t] = dbuf[(long)(jj & 1) * n + t];
    } else {
        const int best = du_special_pick(G, bpart + (long)(jj & 1) * 4 * kDuBlocksMax, C, N, M, fri, fwred);
        if (best >= 0) {
            du_special_apply(d.x, best, C, X, D, unused, bitmap, replaced, n, N, fwred);
            C += 1;
        }
    }
    if (threadIdx.x == 0) *usc = C;
}

// The last sweep position: its new column, or (an unused atom, jj > 0, G > 0:
// the grid of du_ba_kernel) the special case begun by du_ba_kernel(jj).
__global__ void du_tail_kernel(int jj, const uint32_t* __restrict__ perm, const int32_t* __restrict__ count,
                               float* __restrict__ D, const float* __restrict__ abuf,
                               const float* __restrict__ anorm, int n, int G, const float* __restrict__ X,
                               int32_t* __restrict__ unused, int32_t* __restrict__ usc,
                               int32_t* __restrict__ bitmap, int32_t* __restrict__ replaced, int N, int M,
                               const float* __restrict__ bpart) {
    __shared__ int fri[16];
    __shared__ float fwred[16];
    du_commit_column(jj, perm, count, D, abuf, anorm, n);
    const int p = (int)perm[jj];
    if (G > 0 && jj > 0 && count[p] == 0)
        du_special_finish(p, G, X, D, unused, usc, bitmap, replaced, n, N, M, bpart, fri, fwred);
}

// R = X - D Gamma for n <= 1024: one wave per signal, each lane owns 4
// consecutive rows per 256 (float4), RES_COLS dictionary columns in flight
// (8 measured slower than 4 at n = 256); the signal's own row of X is fetched
// first so its latency overlaps the gathers.  X is read and R written with
// streaming accesses (each touched once here), so that L2 keeps the
// dictionary columns the gathers share.
template <int NCH>
__global__ __launch_bounds__(256) void du_residual4_kernel(const float* __restrict__ X, const float* __restrict__ D,
                                                           const int32_t* __restrict__ idx,
                                                           const float* __restrict__ val,
                                                           const int32_t* __restrict__ cnt, int n, int N, int T,
                                                           float* __restrict__ R) {
    constexpr int RES_COLS = 4;
    const int lane = threadIdx.x & 63;
    const bool vec = (n & 3) == 0;
    for (long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6); s < N; s += (long)gridDim.x * 4) {
        // all T slots are fetched together with the count (only lanes below
        // the count are read), so the loads do not wait on each other
        const int c = cnt[s];
        const int il = (lane < T) ? idx[s * T + lane] : 0;
        const float vl = (lane < T) ? val[s * T + lane] : 0.f;
        float4 x4[NCH], xh[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            x4[ch] = ld4_nt(X + s * n, 256 * ch + 4 * lane, n, vec);
            xh[ch] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (int j = 0; j < c; j += RES_COLS) {
            float4 d4[RES_COLS][NCH];
            float g[RES_COLS];
#pragma unroll
            for (int u = 0; u < RES_COLS; ++u) {
                const bool ok = j + u < c;
                const long col = readlane(il, ok ? j + u : j);
                g[u] = ok ? readlane(vl, j + u) : 0.f;
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch)
                    d4[u][ch] = ok ? ld4(D + col * n, 256 * ch + 4 * lane, n, vec) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < RES_COLS; ++u)
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) xh[ch] = f4fma(g[u], d4[u][ch], xh[ch]);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
            st4_nt(R + s * n, 256 * ch + 4 * lane, n, vec,
                   make_float4(x4[ch].x - xh[ch].x, x4[ch].y - xh[ch].y, x4[ch].z - xh[ch].z, x4[ch].w - xh[ch].w));
    }
}

// Per sweep position jj: {p, N_p, ptr_p, q, N_q, ptr_q, 0, 0} with p = perm[jj]
// and q = perm[jj + 1] (-1, 0, 0 past the end; ptr 0 for an empty list).
__global__ void du_sweep_desc_kernel(const uint32_t* __restrict__ perm, const int32_t* __restrict__ count,
                                     const int32_t* __restrict__ ptr, int K, int4* __restrict__ sweep) {
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= K) return;
    const int p = (int)perm[jj];
    const int q = (jj + 1 < K) ? (int)perm[jj + 1] : -1;
    const int Np = count[p], Nq = (q >= 0) ? count[q] : 0;
    sweep[2 * jj] = make_int4(p, Np, Np > 0 ? ptr[p] : 0, q);
    sweep[2 * jj + 1] = make_int4(Nq, Nq > 0 ? ptr[q] : 0, 0, 0);
}

template <int NCH>
void launch_du_atoms(DuWork& w, const float* X, float* D, float* val, int M, hipStream_t st) {
    // sweep position 0: phase A (or the special case's block records); then per
    // position jj: phase B of jj fused with phase A of jj + 1 (one launch per
    // atom); the tail commits the last column.  The accumulator sets start
    // cleared (a fill kernel: graph-capturable).
    const long sw = du_set_words((w.n + 3) & ~3);
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((6 * sw + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<int32_t*>(w.dacc), 6 * sw, 0);
    hipLaunchKernelGGL(du_sweep_desc_kernel, dim3((w.K + 255) / 256), dim3(256), 0, st, w.perm_dev, w.count, w.ptr,
                       w.K, w.sweep);
    const int nb = w.blocks;  // grid of the per-atom kernels (grid-stride over the atom's list)
    float* dbuf = w.dnew;     // 2 x n: d_new by sweep-position parity
    hipLaunchKernelGGL((du_a_kernel<NCH>), dim3(nb), dim3(512), 0, st, w.perm_dev, w.count, w.ptr, w.list_rec, X,
                       D, w.R, w.unused, w.us_counter, w.n, w.N, w.T, M, w.apart, w.bpart, w.dacc, w.dstate);
    for (int jj = 0; jj < w.K; ++jj)
        hipLaunchKernelGGL((du_ba_kernel<NCH>), dim3(nb), dim3(512), 0, st, jj, w.list_rec, X, val, D, w.R,
                           w.unused, w.bitmap, w.replaced, w.n, w.N, w.T, M, w.apart, w.bpart, w.dacc, dbuf,
                           w.dstate, w.sweep);
    hipLaunchKernelGGL(du_tail_pipe_kernel, dim3(1), dim3(256), 0, st, w.K - 1, w.sweep, D, dbuf, w.n, nb, X,
                       w.unused, w.us_counter, w.bitmap, w.replaced, w.N, M, w.bpart, w.dstate);
}

// ---- error: S/KSVD.cu:235-262, S/ReductionKernels.cu:257-322 ---------------
constexpr int ERR_BLOCKS = 2048;

__global__ __launch_bounds__(256) void error_kernel(const float* __restrict__ X, const float* __restrict__ D,
                                                    const int32_t* __restrict__ idx, const float* __restrict__ val,
                                                    const int32_t* __restrict__ cnt,
                                                    const int32_t* __restrict__ bitmap, int n, int N, int T,
                                                    float* __restrict__ err, double* __restrict__ part) {
    __shared__ double wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc = 0.0;
    for (long s = (long)blockIdx.x * 4 + w; s < N; s += (long)gridDim.x * 4) {
        const float e = wave_signal_error(X, D, idx, val, cnt[s], s, n, T, lane);
        if (lane == 0) err[s] = e * (float)bitmap[s];
        acc += (double)e;
    }
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// err_s = ||r_s||^2 from the maintained residuals R = X - D Gamma (after a
// dictionary-update sweep R holds exactly the reconstruction error the
// reference forms with a dense D * Gamma GEMM, S/KSVD.cu:235-262).
__global__ __launch_bounds__(256) void error_from_residual_kernel(const float* __restrict__ R,
                                                                  const int32_t* __restrict__ bitmap, int n, int N,
                                                                  float* __restrict__ err, double* __restrict__ part) {
    __shared__ double wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc = 0.0;
    for (long s = (long)blockIdx.x * 4 + w; s < N; s += (long)gridDim.x * 4) {
        // four loads of the row in flight, summed in the order t = lane, lane + 64, ...
        const float* __restrict__ r_s = R + s * n;
        float e = 0.f;
        int t = lane;
        for (; t + 192 < n; t += 256) {
            const float r0 = __builtin_nontemporal_load(r_s + t), r1 = __builtin_nontemporal_load(r_s + t + 64),
                        r2 = __builtin_nontemporal_load(r_s + t + 128), r3 = __builtin_nontemporal_load(r_s + t + 192);
            e += r0 * r0;
            e += r1 * r1;
            e += r2 * r2;
            e += r3 * r3;
        }
        for (; t < n; t += 64) {
            const float r = r_s[t];
            e += r * r;
        }
        e = wave_sum(e);
        if (lane == 0) err[s] = e * (float)bitmap[s];
        acc += (double)e;
    }
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

__global__ __launch_bounds__(1024) void error_finish_kernel(const double* __restrict__ part, int nparts, double denom,
                                                            double* __restrict__ rmse) {
    __shared__ double wred[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    s = block_sum(s, wred);
    if (threadIdx.x == 0) *rmse = sqrt(s / denom);
}

// ---- clearDict: S/KSVD.cu:138-226 ------------------------------------------
// usecount_k = #{s : gamma_ks^2 > 1e-7}  (use_count_kernel, S/KSVD.cu:403-415):
// per-block LDS histogram over a grid-stride range of compact entries (thread
// = entry s*T + j, so the idx/val loads are coalesced), then one global add
// per atom and block (integer counts: order independent).  The stride is the
// largest multiple of T the grid covers, so each thread keeps one slot j and
// steps its signal by stride / T (no division per entry; N*T < 2^31).
__global__ __launch_bounds__(256) void usecount_kernel(const int32_t* __restrict__ idx, const float* __restrict__ val,
                                                       const int32_t* __restrict__ cnt, int N, int T, int K,
                                                       int32_t* __restrict__ usecount) {
    extern __shared__ int32_t hist[];
    for (int k = threadIdx.x; k < K; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int sstep = (int)(gridDim.x * blockDim.x) / T;  // signals per sweep of the grid
    if (gt < sstep * T) {
        const int j = gt % T;
        // four signals per round: their count, coefficient and atom loads all
        // in flight together (slots past the count are in bounds, unused)
        constexpr int SU = 4;
        for (int s = gt / T; s < N; s += SU * sstep) {
            int c[SU], k[SU];
            float v[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int su = s + u * sstep;
                const bool in = su < N;
                c[u] = in ? __builtin_nontemporal_load(cnt + su) : 0;
                v[u] = in ? __builtin_nontemporal_load(val + su * T + j) : 0.f;
                k[u] = in ? __builtin_nontemporal_load(idx + su * T + j) : 0;
            }
#pragma unroll
            for (int u = 0; u < SU; ++u)
                if (j < c[u] && v[u] * v[u] > 1e-7f) atomicAdd(&hist[k[u]], 1);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        if (hist[k]) atomicAdd(&usecount[k], hist[k]);
}

// clearDict's test for every atom j on the Gram at entry, one wave per atom:
// flag0_j = (max_{i != j} G_j[i]^2 > mu^2 or usecount_j < use_thresh) and not
// replaced_j (S/KSVD.cu:148-160).  The sequential sweep uses it for as long
// as no atom has been replaced (until then G is still the Gram at entry).
__global__ __launch_bounds__(256) void clear_flags_kernel(const float* __restrict__ G,
                                                          const int32_t* __restrict__ usecount,
                                                          const int32_t* __restrict__ replaced, int K, double mu2,
                                                          int use_thresh, int32_t* __restrict__ flag0) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= K) return;
    float mx = 0.f;
    for (int i = lane; i < K; i += 64) {
        const float g = G[(long)j * K + i];
        if (i != j) mx = fmaxf(mx, g * g);
    }
    mx = wave_max(mx);
    if (lane == 0) flag0[j] = (((double)mx > mu2) || (usecount[j] < use_thresh)) && (replaced[j] == 0);
}

// Sequential sweep over atoms in one workgroup.  G holds D^T D for the D at
// entry; when atom j is replaced, row and column j of G are recomputed so that
// later steps see the current dictionary (the reference recomputes
// G_j = D^T d_j with a GEMV per atom, S/KSVD.cu:164-173).  Until the first
// replacement the per-atom test is flag0 (held in LDS), so atoms that keep
// their column cost no global read and no barrier.
__global__ __launch_bounds__(1024) void clear_sweep_kernel(const float* __restrict__ X, float* __restrict__ D,
                                                           float* __restrict__ G,
                                                           const int32_t* __restrict__ usecount,
                                                           const int32_t* __restrict__ replaced,
                                                           const int32_t* __restrict__ flag0,
                                                           float* __restrict__ err, int n, int K, int N,
                                                           double mu2, int use_thresh) {
    extern __shared__ int32_t fl[];  // flag0, K entries
    __shared__ float rv[16];
    __shared__ int ri[16];
    __shared__ float wred[16];
    __shared__ int flag_sh;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int k = threadIdx.x; k < K; k += blockDim.x) fl[k] = flag0[k];
    __syncthreads();
    int nrep = 0;  // atoms replaced so far in this sweep (uniform)
    for (int j = 0; j < K; ++j) {
        if (nrep == 0 && !fl[j]) continue;
        // max_{i != j} G_j[i]^2   (device_max, S/ReductionKernels.cu:393-408)
        float mx = 0.f;
        for (int i = threadIdx.x; i < K; i += blockDim.x) {
            const float g = G[(long)j * K + i];
            if (i != j) mx = fmaxf(mx, g * g);
        }
        mx = wave_max(mx);
        if (lane == 0) rv[w] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = rv[0];
            for (int q = 1; q < nw; ++q) m = fmaxf(m, rv[q]);
            flag_sh = (((double)m > mu2) || (usecount[j] < use_thresh)) && (replaced[j] == 0);
        }
        __syncthreads();
        if (!flag_sh) continue;
        // arg-max of err (S/ReductionKernels.cu:329-386), lowest index on ties
        float bv = -1.f;
        int bi = 0;
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            const float e = err[i];
            if (e > bv) { bv = e; bi = i; }
        }
        wave_argmax(bv, bi);
        if (lane == 0) { rv[w] = bv; ri[w] = bi; }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v = rv[0];
            int i = ri[0];
            for (int q = 1; q < nw; ++q)
                if (rv[q] > v || (rv[q] == v && ri[q] < i)) { v = rv[q]; i = ri[q]; }
            ri[0] = i;
            err[i] = 0.f;
        }
        __syncthreads();
        const long best = ri[0];
        float sq = 0.f;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const float x = X[best * n + t];
            sq += x * x;
        }
        sq = block_sum(sq, wred);
        // the pick is a zero signal (every remaining error 0): x / ||x|| is
        // undefined, the atom stays as it is
        if (sq == 0.f) continue;
        ++nrep;
        const float nrm = sqrtf(sq);  // EUnorm S/ReductionKernels.cu:413-430
        for (int t = threadIdx.x; t < n; t += blockDim.x) D[(long)j * n + t] = X[best * n + t] / nrm;
        __threadfence_block();
        __syncthreads();
        // refresh row/column j of G for the new atom
        for (int i = w; i < K; i += nw) {
            float s = 0.f;
            for (int t = lane; t < n; t += 64) s += D[(long)i * n + t] * D[(long)j * n + t];
            s = wave_sum(s);
            if (lane == 0) { G[(long)j * K + i] = s; G[(long)i * K + j] = s; }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// ---- layout conversions ------------------------------------------------------
// Slot map + c -> compact (every nonzero slot, as sprow_kernel reads them).
__global__ __launch_bounds__(256) void dense_to_compact_kernel(const float* __restrict__ sel,
                                                               const float* __restrict__ c, int K, int N, int T,
                                                               int32_t* __restrict__ idx, float* __restrict__ val,
                                                               int32_t* __restrict__ cnt) {
    const long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= N) return;
    const int lane = threadIdx.x & 63;
    int mxs = 0;
    for (int k = lane; k < K; k += 64) {
        const int slot = (int)sel[s * K + k];
        if (slot > 0 && slot <= T) {
            idx[s * T + slot - 1] = k;
            val[s * T + slot - 1] = c[s * T + slot - 1];
            mxs = max(mxs, slot);
        }
    }
    for (int o = 32; o > 0; o >>= 1) mxs = max(mxs, __shfl_xor(mxs, o, 64));
    if (lane >= mxs && lane < T) { idx[s * T + lane] = -1; val[s * T + lane] = 0.f; }
    if (lane == 0) cnt[s] = mxs;
}

// compact -> dense Gamma values (K x N); each entry written once.
__global__ __launch_bounds__(256) void compact_to_dense_kernel(const int32_t* __restrict__ idx,
                                                               const float* __restrict__ val,
                                                               const int32_t* __restrict__ cnt, int K, int N, int T,
                                                               float* __restrict__ out) {
    const long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= N) return;
    const int lane = threadIdx.x & 63;
    const int c = cnt[s];
    const int il = (lane < c) ? idx[s * T + lane] : -1;
    const float vl = (lane < c) ? val[s * T + lane] : 0.f;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const int k = k0 + lane;
        float v = 0.f;
        for (int j = 0; j < c; ++j)
            if (readlane(il, j) == k) v = readlane(vl, j);
        if (k < K) out[s * K + k] = v;
    }
}

// compact -> reference slot map (K x N, slot+1 or 0) and c (T x N); slots of c
// beyond the count are left untouched, as parOMP leaves them.
__global__ __launch_bounds__(256) void compact_to_slotmap_kernel(const int32_t* __restrict__ idx,
                                                                 const float* __restrict__ val,
                                                                 const int32_t* __restrict__ cnt, int K, int N, int T,
                                                                 float* __restrict__ sel, float* __restrict__ c) {
    const long s = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= N) return;
    const int lane = threadIdx.x & 63;
    const int cc = cnt[s];
    const int il = (lane < cc) ? idx[s * T + lane] : -1;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const int k = k0 + lane;
        float v = 0.f;
        for (int j = 0; j < cc; ++j)
            if (readlane(il, j) == k) v = (float)(j + 1);
        if (k < K) sel[s * K + k] = v;
    }
    if (lane < cc) c[s * T + lane] = val[s * T + lane];
}

}  // namespace

// ---- launchers ---------------------------------------------------------------
static inline unsigned blocks_for(long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_du_init(DuWork& w, hipStream_t st) {
    const long m = w.N > w.K ? w.N : w.K;
    hipLaunchKernelGGL(du_init_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, st, w.unused, w.us_counter, w.bitmap,
                       w.replaced, w.N, w.K);
    return hipGetLastError();
}

hipError_t launch_du_build_lists(DuWork& w, const int32_t* idx, const int32_t* cnt, hipStream_t st,
                                 const float* val) {
    hipLaunchKernelGGL(du_hist_kernel, dim3(w.nchunks), dim3(256), (size_t)w.K * 4, st, idx, cnt, w.N, w.T, w.K,
                       w.chunk, w.chunk_hist);
    hipLaunchKernelGGL(du_scan_chunks_kernel, dim3(w.K), dim3(256), 0, st, w.chunk_hist, w.nchunks, w.K, w.count);
    hipLaunchKernelGGL(du_scan_ptr_kernel, dim3(1), dim3(1024), 0, st, w.count, w.K, w.ptr);
    // the pipelined sweep (n <= 1024) also records each entry's neighbours in the sweep order
    int4* rec = (val && w.n <= 1024) ? w.list_rec : nullptr;
    hipLaunchKernelGGL(du_fill_kernel, dim3(w.nchunks), dim3(64), (size_t)w.K * 4 * (rec ? 3 : 1), st, idx, cnt, w.N,
                       w.T, w.K, w.chunk, w.chunk_hist, w.ptr, w.list_pos, val, w.perm_dev, rec);
    return hipGetLastError();
}

hipError_t launch_du_sweep(DuWork& w, const float* X, float* D, const int32_t* idx, float* val, const int32_t* cnt,
                           hipStream_t st) {
    hipError_t e = launch_du_build_lists(w, idx, cnt, st, val);
    if (e != hipSuccess) return e;
    const int M = w.N < w.max_signals ? w.N : w.max_signals;
    if (w.n <= 1024) {
        const int nch = (w.n + 255) / 256;
        if (nch == 1) {
            hipLaunchKernelGGL((du_residual4_kernel<1>), dim3(RES_BLOCKS), dim3(256), 0, st, X, D, idx, val, cnt, w.n,
                               w.N, w.T, w.R);
            launch_du_atoms<1>(w, X, D, val, M, st);
        } else if (nch == 2) {
            hipLaunchKernelGGL((du_residual4_kernel<2>), dim3(RES_BLOCKS), dim3(256), 0, st, X, D, idx, val, cnt, w.n,
                               w.N, w.T, w.R);
            launch_du_atoms<2>(w, X, D, val, M, st);
        } else {
            hipLaunchKernelGGL((du_residual4_kernel<4>), dim3(RES_BLOCKS), dim3(256), 0, st, X, D, idx, val, cnt, w.n,
                               w.N, w.T, w.R);
            launch_du_atoms<4>(w, X, D, val, M, st);
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(du_residual_kernel, dim3(RES_BLOCKS), dim3(256), 0, st, X, D, idx, val, cnt, w.n, w.N, w.T, w.R);
    // per atom: phase A (block partials, or the special case's block maxima),
    // their reduction (a and ||a||; workgroup 0 finishes the special case),
    // phase B; the new column is committed by the next phase A or the tail
    float* abuf = w.dnew;             // 2 x n: a by sweep-position parity
    float* anorm = w.dnew + 2 * w.n;  // 2 x DU_RB: partial sums of a^2
    hipLaunchKernelGGL(du_sweep_desc_kernel, dim3((w.K + 255) / 256), dim3(256), 0, st, w.perm_dev, w.count, w.ptr,
                       w.K, w.sweep);
    const size_t lds_a = (size_t)4 * w.n * sizeof(float);
    const size_t lds_b = (size_t)w.n * sizeof(float);
    for (int jj = 0; jj < w.K; ++jj) {
        hipLaunchKernelGGL(du_phase_a_kernel, dim3(w.blocks), dim3(256), lds_a, st, jj, w.perm_dev, w.count, w.ptr,
                           w.list_pos, val, X, D, w.R, w.unused, w.us_counter, w.bitmap, w.replaced, w.n, w.N, w.T, M,
                           w.apart, w.bpart, abuf, anorm);
        hipLaunchKernelGGL(du_a_reduce_kernel, dim3(DU_RB), dim3(512), 0, st, jj, w.apart, abuf, anorm, w.n, w.blocks,
                           jj, w.sweep, X, D, w.unused, w.us_counter, w.bitmap, w.replaced, w.N, M, w.bpart);
        hipLaunchKernelGGL(du_phase_b_kernel, dim3(w.blocks), dim3(256), lds_b, st, jj, w.perm_dev, w.count, w.ptr,
                           w.list_pos, val, D, w.R, abuf, anorm, w.n, w.T);
    }
    // the last column (the special cases were all finished by the reductions)
    hipLaunchKernelGGL(du_tail_kernel, dim3(1), dim3(256), 0, st, w.K - 1, w.perm_dev, w.count, D, abuf, anorm, w.n,
                       0, X, w.unused, w.us_counter, w.bitmap, w.replaced, w.N, M, w.bpart);
    return hipGetLastError();
}

hipError_t launch_error(DuWork& w, const float* X, const float* D, const int32_t* idx, const float* val,
                        const int32_t* cnt, float* err, hipStream_t st) {
    hipLaunchKernelGGL(error_kernel, dim3(ERR_BLOCKS), dim3(256), 0, st, X, D, idx, val, cnt, w.bitmap, w.n, w.N, w.T,
                       err, w.rmse_part);
    hipLaunchKernelGGL(error_finish_kernel, dim3(1), dim3(1024), 0, st, w.rmse_part, ERR_BLOCKS,
                       (double)w.n * (double)w.N, w.rmse_dev);
    return hipGetLastError();
}

hipError_t launch_error_from_residual(DuWork& w, float* err, hipStream_t st) {
    hipLaunchKernelGGL(error_from_residual_kernel, dim3(ERR_BLOCKS), dim3(256), 0, st, w.R, w.bitmap, w.n, w.N, err,
                       w.rmse_part);
    hipLaunchKernelGGL(error_finish_kernel, dim3(1), dim3(1024), 0, st, w.rmse_part, ERR_BLOCKS,
                       (double)w.n * (double)w.N, w.rmse_dev);
    return hipGetLastError();
}

hipError_t launch_clear_dict(DuWork& w, const float* X, float* D, const int32_t* idx, const float* val,
                             const int32_t* cnt, float* err, hipStream_t st) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(w.K, 256)), dim3(256), 0, st, w.usecount, (long)w.K, 0);
    const unsigned ub = (unsigned)std::min<long>(blocks_for((long)w.N * w.T, 256), 512);
    hipLaunchKernelGGL(usecount_kernel, dim3(ub), dim3(256), (size_t)w.K * 4, st, idx, val, cnt, w.N, w.T, w.K,
                       w.usecount);
    hipError_t e = launch_at_times_b(D, w.n, w.K, D, w.K, w.Gtmp, 0, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(clear_flags_kernel, dim3(blocks_for(w.K, 4)), dim3(256), 0, st, w.Gtmp, w.usecount,
                       w.replaced, w.K, w.mu_thresh * w.mu_thresh, w.use_thresh, w.clear_flag);
    hipLaunchKernelGGL(clear_sweep_kernel, dim3(1), dim3(1024), (size_t)w.K * 4, st, X, D, w.Gtmp, w.usecount,
                       w.replaced, w.clear_flag, err, w.n, w.K, w.N, w.mu_thresh * w.mu_thresh, w.use_thresh);
    return hipGetLastError();
}

hipError_t launch_dense_to_compact(const float* sel, const float* c, int K, int N, int T, int32_t* idx, float* val,
                                   int32_t* cnt, hipStream_t st) {
    hipLaunchKernelGGL(dense_to_compact_kernel, dim3(blocks_for(N, 4)), dim3(256), 0, st, sel, c, K, N, T, idx, val,
                       cnt);
    return hipGetLastError();
}

hipError_t launch_compact_to_dense(const int32_t* idx, const float* val, const int32_t* cnt, int K, int N, int T,
                                   float* out, hipStream_t st) {
    hipLaunchKernelGGL(compact_to_dense_kernel, dim3(blocks_for(N, 4)), dim3(256), 0, st, idx, val, cnt, K, N, T,
                       out);
    return hipGetLastError();
}

hipError_t launch_compact_to_slotmap(const int32_t* idx, const float* val, const int32_t* cnt, int K, int N, int T,
                                     float* sel, float* c, hipStream_t st) {
    hipLaunchKernelGGL(compact_to_slotmap_kernel, dim3(blocks_for(N, 4)), dim3(256), 0, st, idx, val, cnt, K, N, T,
                       sel, c);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, long n, float v, hipStream_t st) {
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace ksvd
