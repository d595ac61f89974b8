// common.hpp -- shared device helpers and internal launcher declarations for
// the MI355X K-SVD engine.  gfx950 (CDNA4) only: wave64, MFMA, 160 KB LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ksvd {

constexpr int WAVE = 64;

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// ---------------------------------------------------------------------------
// Wave64 cross-lane reductions.  The reference's warp32 __shfl_down trees
// (S/ReductionBase.cuh:40-136) become butterfly reductions over 64 lanes
// (ds_swizzle / DPP via __shfl_xor); every lane ends with the result.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Arg-max with the reference's tie rule: larger value wins, equal values keep
// the lower index (S/OMP.cu:201-213 scans with strict '>';
// S/ReductionBase.cuh:40-51 keeps its own lane on ties).
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(v, o, WAVE);
        int oi = __shfl_xor(i, o, WAVE);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// DPP row permutation of a float (all 64 lanes active; every source lane valid
// for the controls used here, so with all rows and banks enabled no lane keeps
// the old value and bound_ctrl lets the compiler leave it unset).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// DPP row permutation of an int (same controls as dpp_f).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

// BCAST = 15: rows 1 and 3 add the last lane of rows 0 and 2 (row_bcast:15,
// row mask 0xA); BCAST = 31: rows 2 and 3 add lane 31 (row_bcast:31, row
// mask 0xC); the other rows keep v.  One v_add_f32_dpp: the rows it leaves
// out keep the destination, which is v (the compiler does not fold a -0.0
// identity into the add and materialises it instead).  The s_nop gives the
// two wait states a DPP read needs after the VALU write of v.
template <int BCAST>
__device__ __forceinline__ float add_row_bcast(float v) {
    static_assert(BCAST == 15 || BCAST == 31, "row_bcast:15 or row_bcast:31");
    if constexpr (BCAST == 15)
        asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(v));
    else
        asm("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(v));
    return v;
}

// Signed-int wave max (DPP within rows, v_readlane across the four rows);
// on the bit patterns of floats that are >= 0 or negative sentinels this is
// the float max.  Call from uniform control flow.
__device__ __forceinline__ int wave_max_i32_dpp(int v) {
    v = max(v, dpp_i<0xB1>(v));
    v = max(v, dpp_i<0x4E>(v));
    v = max(v, dpp_i<0x141>(v));
    v = max(v, dpp_i<0x140>(v));
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return max(max(r0, r1), max(r2, r3));
}

// The same wave max with the row maxima combined by DPP row_bcast:15 (rows
// 1 and 3 take the last lane of rows 0 and 2) and row_bcast:31 (row 3 takes
// lane 31): lane 63 ends with the wave max, read with one v_readlane.
__device__ __forceinline__ int wave_max_i32_bcast(int v) {
    v = max(v, dpp_i<0xB1>(v));
    v = max(v, dpp_i<0x4E>(v));
    v = max(v, dpp_i<0x141>(v));
    v = max(v, dpp_i<0x140>(v));
    // INT_MIN (the identity of max) in the rows left out lets the compiler
    // fold each step into one v_max_i32_dpp
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// lanes_sum_dpp with the row sums combined by row_bcast:15 / row_bcast:31
// (the same additions as lanes_sum_dpp, one v_readlane; the rows left out
// keep their value)
template <int NR>
__device__ __forceinline__ float lanes_sum_bcast(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    if constexpr (NR == 1) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    v = add_row_bcast<15>(v);
    if constexpr (NR == 2) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    v = add_row_bcast<31>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Wave sum with DPP row operations (quad swaps, half-row and row mirrors:
// every lane of a row ends with the same row sum, the adds being commutative
// pairs), then the four row sums in fixed order.  Deterministic; uniform
// control flow only.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
}

// Sum over the first 16*NR lanes (NR = 1, 2 or 4 rows of 16 lanes): DPP
// within each row as in wave_sum_dpp, then the row sums in fixed order.
// Deterministic; uniform control flow only.
template <int NR>
__device__ __forceinline__ float lanes_sum_dpp(float v) {
    static_assert(NR == 1 || NR == 2 || NR == 4, "rows of 16 lanes");
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    if constexpr (NR == 1) return r0;
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    if constexpr (NR == 2) return r0 + r1;
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
}

// Wave max without LDS traffic: quad swaps (quad_perm [1,0,3,2], [2,3,0,1]),
// then row_half_mirror and row_mirror make every 16-lane row uniform; the four
// row maxima are combined from v_readlane.  Call from uniform control flow.
__device__ __forceinline__ float wave_max_dpp(float v) {
    v = fmaxf(v, dpp_f<0xB1>(v));
    v = fmaxf(v, dpp_f<0x4E>(v));
    v = fmaxf(v, dpp_f<0x141>(v));
    v = fmaxf(v, dpp_f<0x140>(v));
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}



__device__ __forceinline__ float readlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double readlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Wave-uniform copies (scalar registers) of values every lane already agrees
// on, so that control flow on them compiles to scalar branches.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) { return readlane(v, 0); }

// Block-wide sum for blockDim.x <= 1024 (any multiple of 64); result valid in
// every thread.  `red` must hold 16 entries.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    __syncthreads();
    if (lane_id() == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    for (int i = 0; i < nw; ++i) s += red[i];  // fixed order: deterministic
    return s;
}

}  // namespace ksvd

// ---------------------------------------------------------------------------
// Internal launchers (implemented in the .hip files, called by capi.cpp).
// All enqueue on `st`; return a hipError_t.
// ---------------------------------------------------------------------------
namespace ksvd {

// C (colsA x colsB) = A^T B, column-major, lda = ldb = rows.  form: 0 fp32,
// 1 bf16, 2 bf16x3, 3 bf16x6.
// wpe: workgroups per CU the kernel is compiled for, 2 or 3 (0: 3).
hipError_t launch_at_times_b(const float* A, int rows, int colsA, const float* B, long colsB, float* C,
                             int form, hipStream_t st, int wpe = 0);
hipError_t launch_normcols(const float* Din, float* Dout, int n, int K, hipStream_t st);

struct OmpArgs {
    const float* DtX;   // K x N
    const float* G;     // K x K
    int K, N, T;
    float* sel;         // optional K x N slot map
    float* c;           // optional T x N coefficients
    int32_t* idx;       // optional T x N compact
    float* val;         // optional T x N compact
    int32_t* cnt;       // optional N
    const double* xnorm2;  // error mode: ||x_s||^2 in FP64 (else NULL)
    const float* eps2;
    int32_t* work_counter;  // persistent work-stealing counter (error mode)
    float pivot_fp64;
    unsigned long long* stats;  // [1]: pivots recomputed in FP64 (register-basis form); [0] unused
    int q_kpl, q_waves;         // register-basis shape (0: the default rule)
    float* Lfac;                // optional N x T(T+1)/2: each signal's Cholesky factor L (L L^T = G_II),
                                // lower triangle packed column by column (register-basis form)
};
hipError_t launch_batch_omp(const OmpArgs& a, hipStream_t st);
hipError_t launch_zero_counter(int32_t* p, hipStream_t st);  // *p = 0 (a kernel: graph-capturable)

// Batch-OMP with the Gram-domain basis Q = G(:,I) L^-T in registers
// (omp_qform.hip): one workgroup of `waves` waves per signal, `kpl` atoms per
// lane, Q sized for T <= tt.  kpl == 0: (K, T) outside the supported range.
struct QformShape {
    int kpl, tt, waves;
};
QformShape qform_shape(int K, int T);
// (kpl, waves) as given when compiled and covering K, else kpl == 0
QformShape qform_shape_with(int K, int T, int kpl, int waves);
// the compiled shapes for (K, T): the fewest waves for each kpl; returns the count
int qform_candidates(int K, int T, QformShape* out, int cap);
hipError_t launch_batch_omp_qform(const OmpArgs& a, hipStream_t st);

hipError_t launch_compact_to_slotmap(const int32_t* idx, const float* val, const int32_t* cnt, int K, int N, int T,
                                     float* sel, float* c, hipStream_t st);

// Dictionary update / error / clear (dict_update.hip)
constexpr int kDuBlocks = 768;      // default workgroups of the per-atom phase kernels (one block partial each)
constexpr int kDuBlocksMax = 1024;  // the most the scratch (apart, bpart) is sized for
constexpr int kDuBuckets = 8;       // buckets of an accumulator set of the pipelined sweep (n <= 1024)
// int64 words of one accumulator set: kDuBuckets x 2 words x np4, a flag word, padding
__host__ __device__ constexpr long du_set_words(int np4) { return 2L * kDuBuckets * np4 + 2; }
struct DuWork {
    int n, K, N, T, max_signals, use_thresh;
    double mu_thresh;
    // scratch (device)
    int32_t* count;      // K
    int32_t* ptr;        // K + 1
    int32_t* chunk_hist; // K x nchunks
    int32_t* list_pos;   // N*T: entry positions s*T + slot (the sweep without pipelining, n > 1024)
    int4* list_rec;      // N*T records of the pipelined sweep: {position | previous-atom flag (bit 31),
                         // gamma bits, bits of the signal's gamma on the next atom in the sweep order (0 if
                         // unused), 0}
    float* R;            // n x N restricted residuals X - D Gamma
    int blocks;          // workgroups of the per-atom kernels (<= kDuBlocksMax)
    float* apart;        // 2 x kDuBlocksMax x (n rounded up to 4) FP32 block partials of a, by sweep-position
                         // parity (the pipelined sweep; above n = 1024 the first half, one position)
    float* bpart;        // 2 x kDuBlocksMax x 4 special-case block records, by sweep-position parity
    float* dnew;         // 2 x n + 2 x 16: a and its partial sums of squares (above n = 1024), or d_new
                         // (the pipelined sweep), by sweep-position parity
    unsigned long long* dacc;  // 3 accumulator sets of du_set_words(np4) (the pipelined sweep)
    int32_t* dstate;     // 2: the unused list's counter after a sweep position, by position parity
    int32_t* unused;     // N
    int32_t* us_counter; // 1
    int32_t* bitmap;     // N
    int32_t* replaced;   // K
    uint32_t* perm_dev;  // K
    int32_t* usecount;   // K
    float* Gtmp;         // K x K (clearDict Gram)
    double* rmse_part;   // nparts
    double* rmse_dev;    // 1
    int32_t* clear_flag; // K: clearDict's per-atom test on the Gram at entry
    int4* sweep;         // 2K: per sweep position {p, N_p, ptr_p, q}, {N_q, ptr_q, 0, 0} (pipelined sweep)
    int nchunks, chunk;
    int rmse_parts;
};
hipError_t launch_du_init(DuWork& w, hipStream_t st);
hipError_t launch_du_build_lists(DuWork& w, const int32_t* idx, const int32_t* cnt, hipStream_t st,
                                 const float* val = nullptr);
hipError_t launch_du_sweep(DuWork& w, const float* X, float* D, const int32_t* idx, float* val,
                           const int32_t* cnt, hipStream_t st);
hipError_t launch_error(DuWork& w, const float* X, const float* D, const int32_t* idx, const float* val,
                        const int32_t* cnt, float* err, hipStream_t st);
hipError_t launch_error_from_residual(DuWork& w, float* err, hipStream_t st);
hipError_t launch_clear_dict(DuWork& w, const float* X, float* D, const int32_t* idx, const float* val,
                             const int32_t* cnt, float* err, hipStream_t st);
hipError_t launch_dense_to_compact(const float* sel, const float* c, int K, int N, int T, int32_t* idx,
                                   float* val, int32_t* cnt, hipStream_t st);
hipError_t launch_compact_to_dense(const int32_t* idx, const float* val, const int32_t* cnt, int K, int N,
                                   int T, float* gamma_dense, hipStream_t st);

// Adjoint (adjoint.hip).  Lfac NULL: G_II and its factor recomputed per
// signal; else the packed factors Batch-OMP stored.  list_ptr / list_pos: the
// per-atom lists of (signal, slot) positions in ascending signal order
// (launch_du_build_lists); R, zbuf: n x N and T x N scratch.
hipError_t launch_omp_adjoint(const float* D, const float* X, const int32_t* idx, const float* val,
                              const int32_t* cnt, const float* Lfac, const float* gval, float* gD, float* gX,
                              const int32_t* list_ptr, const int32_t* list_pos, float* R, float* zbuf, int n,
                              int K, int N, int T, hipStream_t st);
hipError_t launch_normcols_adjoint(const float* Din, const float* gout, float* gin, int n, int K, hipStream_t st);
hipError_t launch_gather_compact(const float* dense, const int32_t* idx, const int32_t* cnt, int K, int N, int T,
                                 float* out, hipStream_t st);

}  // namespace ksvd
