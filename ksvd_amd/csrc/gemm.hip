// gemm.hip -- C = A^T B on CDNA4 MFMA (Gram G = D^T D and correlation
// DtX = D^T X).  Replaces the cuBLAS ?gemm(T,N) of
// Algorithms::Multiplication::AT_times_B (S/Multiplication.cu:40-63).
//
// Layout: A is rows x colsA, B is rows x colsB, both column-major, so both
// operands are contiguous along the reduction dimension.  C is colsA x colsB
// column-major: C[j*colsA + m] = <A(:,m), B(:,j)>.
//
// Block tile 128 (columns of B) x 128 (columns of A), BK = 32, 256 threads =
// 4 waves in a 2 x 2 arrangement, each wave 64 x 64 = 2 x 2 MFMA 32x32 tiles.
// The MFMA row operand is B (signals) and the column operand is A (atoms), so
// each accumulator register holds 32 consecutive atoms of one signal: the
// column-major store of C is one 128-byte segment per half-wave.
//
// Forms (MI355X dense peaks):
//   FP32   v_mfma_f32_32x32x2_f32   (157 TF)  FP32 products, FP32 accumulate
//   BF16   v_mfma_f32_32x32x16_bf16 (2.5 PF)  RNE-rounded BF16 operands
//   BF16X3 three BF16 MFMAs: hi*hi + hi*lo + lo*hi with x = hi + lo, hi =
//          bf16(x), lo = bf16(x - hi)  (Henry, Tang & Heinecke, ARITH 2019):
//          x is carried to 16 significant bits, |x - hi - lo| <= 2^-16 |x|
//   BF16X6 the three-part split of the same paper, x = hi + mid + lo (24
//          bits, FP32's precision), six BF16 MFMAs: every product term down
//          to 2^-16 (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi); the
//          terms left out (mid*lo, lo*mid, lo*lo) are below 2^-23 relative
// Operands are fetched as FP32; the BF16 forms convert them once, on their way
// into LDS (bf16 hi, and lo for the split form), so one kernel serves all
// three forms.
#include "common.hpp"

namespace ksvd {

namespace {

constexpr int BM = 128;    // tile extent along colsB (MFMA rows)
constexpr int BN = 128;    // tile extent along colsA (MFMA cols)
constexpr int BK = 32;     // reduction chunk
constexpr int PITCH = 36;  // LDS row pitch in floats (pads 128 B rows: no bank conflicts)
constexpr int PH = 32;     // LDS row pitch in bf16 elements for the BF16 forms (64 B rows, swizzled)

// BF16 tiles: a 64-byte row is four 16-byte units, and unit u of row r is
// stored at unit u ^ bf16_swz(r).  A bf16x8 MFMA operand read has lanes
// 0..31 on 32 consecutive rows at one unit; the swizzle puts any 8 (or 16)
// consecutive rows at distinct 16-byte bank groups of a 128 (or 256) byte
// bank row, so the reads are conflict-free without padding the pitch, and
// the split-BF16x6 form's six tiles take 48 kB (three workgroups per CU)
// instead of 60 kB.
__device__ __forceinline__ int bf16_swz(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }

// One 128 x 32 tile (rows of the tile are columns of the matrix): each
// thread fetches 4 float4s into registers (fetch_tile), later stored to LDS
// (put_tile), so the next chunk's global loads are in flight while the MFMAs
// of the current one run.  Mt is the tile's first column (uniform), thread
// rows are tid/8 + 32 r at 32-bit offsets from it, and a whole tile (in
// range, rows a multiple of 4 and of BK) takes no bounds test: FULL = 1 at
// compile time, FULL = 2 when `full` (uniform) says so at run time.
template <int FULL>
__device__ __forceinline__ void fetch_tile(float4* v, const float* __restrict__ Mt, int rows, int nvalid, int t0,
                                           bool vec, bool full) {
    const int tid = threadIdx.x;
    const int row0 = tid >> 3;
    const int t = t0 + (tid & 7) * 4;
    if (FULL == 1 || (FULL == 2 && full)) {
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = *reinterpret_cast<const float4*>(Mt + ((row0 + 32 * r) * rows + t));
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = row0 + 32 * r;
        v[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < nvalid) {
            const float* src = Mt + (row * rows + t);
            if (vec && t + 3 < rows) {
                v[r] = *reinterpret_cast<const float4*>(src);
            } else {
                if (t + 0 < rows) v[r].x = src[0];
                if (t + 1 < rows) v[r].y = src[1];
                if (t + 2 < rows) v[r].z = src[2];
                if (t + 3 < rows) v[r].w = src[3];
            }
        }
    }
}

// BF16 forms: the tile converted once, at the store to LDS (hi = bf16(x); for
// the split forms lo = bf16(x - hi), or mid = bf16(x - hi) and
// lo = bf16(x - hi - mid); the differences are exact in FP32).
// Pairs are rounded together (one packed RNE conversion per pair) and the
// two BF16 values are widened back by a shift and a mask of the packed word,
// so the word that goes to LDS is the one the residual was formed from.
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
typedef __attribute__((ext_vector_type(2))) float f32x2_t;

__device__ __forceinline__ bf16x2_t bf16_pair(float a, float b, float& fa, float& fb) {
    const f32x2_t x = {a, b};
    const bf16x2_t p = __builtin_convertvector(x, bf16x2_t);
    const unsigned w = __builtin_bit_cast(unsigned, p);
    fa = __uint_as_float(w << 16);
    fb = __uint_as_float(w & 0xFFFF0000u);
    return p;
}

template <int FORM>
__device__ __forceinline__ void put_tile_bf16(__bf16* __restrict__ sh, __bf16* __restrict__ sl,
                                              __bf16* __restrict__ sm, const float4* v) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int q = tid + 256 * r;
        const float x[4] = {v[r].x, v[r].y, v[r].z, v[r].w};
        bf16x4 h, l, m;
#pragma unroll
        for (int e = 0; e < 4; e += 2) {
            float h0, h1;
            const bf16x2_t hp = bf16_pair(x[e], x[e + 1], h0, h1);
            h[e] = hp.x;
            h[e + 1] = hp.y;
            const float r0 = x[e] - h0, r1 = x[e + 1] - h1;
            if constexpr (FORM == 2) {
                const f32x2_t rr = {r0, r1};
                const bf16x2_t lp = __builtin_convertvector(rr, bf16x2_t);
                l[e] = lp.x;
                l[e + 1] = lp.y;
            }
            if constexpr (FORM == 3) {
                float m0, m1;
                const bf16x2_t mp = bf16_pair(r0, r1, m0, m1);
                m[e] = mp.x;
                m[e + 1] = mp.y;
                const f32x2_t rr = {r0 - m0, r1 - m1};
                const bf16x2_t lp = __builtin_convertvector(rr, bf16x2_t);
                l[e] = lp.x;
                l[e + 1] = lp.y;
            }
        }
        // row q / 8, unit (q % 8) / 2, half q % 2 of the unit
        const int row = q >> 3;
        const int off = row * PH + ((((q & 7) >> 1) ^ bf16_swz(row)) << 3) + (q & 1) * 4;
        *reinterpret_cast<bf16x4*>(sh + off) = h;
        if constexpr (FORM >= 2) *reinterpret_cast<bf16x4*>(sl + off) = l;
        if constexpr (FORM == 3) *reinterpret_cast<bf16x4*>(sm + off) = m;
    }
}

__device__ __forceinline__ void put_tile(float* __restrict__ s, const float4* v) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int q = tid + 256 * r;
        *reinterpret_cast<float4*>(s + (q >> 3) * PITCH + (q & 7) * 4) = v[r];
    }
}

// The chunk loop of at_times_b_kernel.  The BF16 forms instantiate it for
// whole tiles (F = 1) and for edge tiles (F = 0) apart: with one loop for
// both, the compiler merged the two fetch paths and waited for the B tile's
// loads to land before it issued the A tile's.  The FP32 form keeps one loop
// with the test at run time (F = 2): its whole-tile instance at three
// workgroups per CU kept the fetch registers in scratch memory.
template <int FORM, int F>
__device__ __forceinline__ void chunk_loop(f32x16 (&acc)[2][2], const float* __restrict__ Bt,
                                           const float* __restrict__ At, int rows, int nB, int nA, bool vec, bool full,
                                           float* __restrict__ sB, float* __restrict__ sA, __bf16* __restrict__ sBh,
                                           __bf16* __restrict__ sAh, __bf16* __restrict__ sBl, __bf16* __restrict__ sAl,
                                           __bf16* __restrict__ sBm, __bf16* __restrict__ sAm, int wi, int wj,
                                           int li, int lk) {
    float4 pb[4], pa[4];
    fetch_tile<F>(pb, Bt, rows, nB, 0, vec, full);
    fetch_tile<F>(pa, At, rows, nA, 0, vec, full);
    for (int t0 = 0; t0 < rows; t0 += BK) {
        if constexpr (FORM == 0) {
            put_tile(sB, pb);
            put_tile(sA, pa);
        } else {
            put_tile_bf16<FORM>(sBh, sBl, sBm, pb);
            put_tile_bf16<FORM>(sAh, sAl, sAm, pa);
        }
        __syncthreads();
        if (t0 + BK < rows) {
            fetch_tile<F>(pb, Bt, rows, nB, t0 + BK, vec, full);
            fetch_tile<F>(pa, At, rows, nA, t0 + BK, vec, full);
        }
        if constexpr (FORM == 0) {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 2) {
                float ra[2], rb[2];
#pragma unroll
                for (int a = 0; a < 2; ++a) ra[a] = sB[(wi * 64 + a * 32 + li) * PITCH + ks + lk];
#pragma unroll
                for (int b = 0; b < 2; ++b) rb[b] = sA[(wj * 64 + b * 32 + li) * PITCH + ks + lk];
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ra[a], rb[b], acc[a][b], 0, 0, 0);
            }
        } else {
            const int sw = bf16_swz(li);  // rows wi*64 + a*32 + li: the swizzle depends on li only
#pragma unroll
            for (int ks = 0; ks < BK; ks += 16) {
                bf16x8 ah[2], al[2], am[2], bh[2], bl[2], bm[2];
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int off = (wi * 64 + a * 32 + li) * PH + ((((ks >> 3) + lk) ^ sw) << 3);
                    ah[a] = *reinterpret_cast<const bf16x8*>(sBh + off);
                    if constexpr (FORM >= 2) al[a] = *reinterpret_cast<const bf16x8*>(sBl + off);
                    if constexpr (FORM == 3) am[a] = *reinterpret_cast<const bf16x8*>(sBm + off);
                }
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int off = (wj * 64 + b * 32 + li) * PH + ((((ks >> 3) + lk) ^ sw) << 3);
                    bh[b] = *reinterpret_cast<const bf16x8*>(sAh + off);
                    if constexpr (FORM >= 2) bl[b] = *reinterpret_cast<const bf16x8*>(sAl + off);
                    if constexpr (FORM == 3) bm[b] = *reinterpret_cast<const bf16x8*>(sAm + off);
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        if constexpr (FORM == 3) {  // smallest terms first
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[a], bm[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[a], bh[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bm[b], acc[a][b], 0, 0, 0);
                        }
                        if constexpr (FORM == 2) {
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                        }
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
}

// Compiled for WPE workgroups per CU (3 by default: the split form needs 192
// registers per lane at the compiler's default and would run two -- 3.03 ms
// against 2.70 ms for the n=256, K=1024, N=10^6 correlation, with a few
// spilled registers; the FP32 and BF16 forms fit three either way, and so
// does the three-part split's 48 kB of LDS).  WPE changes the register
// budget only, not the arithmetic: the autotuner may pick 2.
template <int FORM, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void at_times_b_kernel(const float* __restrict__ A, int rows, int colsA,
                                                         const float* __restrict__ B, long colsB,
                                                         float* __restrict__ C, int tilesA) {
    // FP32: tiles in LDS as FP32; BF16 forms: as bf16 hi (and lo)
    constexpr int NF = FORM == 0 ? BM * PITCH : 1;
    constexpr int NH = FORM == 0 ? 1 : BM * PH;
    constexpr int NL = FORM >= 2 ? BM * PH : 1;
    constexpr int NM = FORM == 3 ? BM * PH : 1;
    __shared__ float sB[NF], sA[NF];
    __shared__ __align__(16) __bf16 sBh[NH], sAh[NH], sBl[NL], sAl[NL], sBm[NM], sAm[NM];
    // XCD-aware tile order: workgroups are dispatched round-robin over the 8
    // XCDs (block b on XCD b % 8), each with its own L2.  Block b works on
    // XCD x = b % 8 through its r = b / 8-th tile: signal tile
    // tb = 8 * (r / tilesA) + x and atom tile r % tilesA, so the tilesA
    // blocks that share a signal tile run on one XCD and read it from HBM
    // once (a plain b / tilesA order spreads them over all 8 L2s).  The grid
    // holds tilesB rounded up to a multiple of 8 signal tiles.
    const long blk = blockIdx.x;
    const long r8 = blk >> 3;
    const int ta = (int)(r8 % tilesA);
    const long tb = 8 * (r8 / tilesA) + (blk & 7);
    if (tb * BM >= colsB) return;
    const long j0 = tb * BM;
    const int m0 = ta * BN;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int wi = w & 1, wj = w >> 1;
    const bool vec = (rows & 3) == 0;
    const int nB = (int)min((long)BM, colsB - j0), nA = min(BN, colsA - m0);
    const bool full = vec && rows % BK == 0 && nB == BM && nA == BN;
    const float* Bt = B + j0 * (long)rows;
    const float* At = A + (long)m0 * rows;

    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

    const int li = lane & 31;
    const int lk = lane >> 5;
    if (FORM == 0)
        chunk_loop<FORM, 2>(acc, Bt, At, rows, nB, nA, vec, full, sB, sA, sBh, sAh, sBl, sAl, sBm, sAm, wi, wj, li, lk);
    else if (full)
        chunk_loop<FORM, 1>(acc, Bt, At, rows, nB, nA, vec, full, sB, sA, sBh, sAh, sBl, sAl, sBm, sAm, wi, wj, li, lk);
    else
        chunk_loop<FORM, 0>(acc, Bt, At, rows, nB, nA, vec, full, sB, sA, sBh, sAh, sBl, sAl, sBm, sAm, wi, wj, li, lk);

    // Accumulator register r of lane l holds MFMA row 8*(r/4) + 4*(l/32) + r%4
    // (a signal) and column l%32 (an atom); offsets from the tile's corner.
    float* Ct = C + j0 * (long)colsA + m0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int ml = wj * 64 + b * 32 + li;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int jl = wi * 64 + a * 32 + 8 * (r >> 2) + 4 * lk + (r & 3);
                if (full || (ml < nA && jl < nB)) __builtin_nontemporal_store(acc[a][b][r], Ct + (jl * colsA + ml));
            }
        }
}

__global__ __launch_bounds__(256) void normcols_kernel(const float* __restrict__ Din, float* __restrict__ Dout,
                                                       int n, int K) {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= K) return;
    const int lane = threadIdx.x & 63;
    const float* d = Din + (long)k * n;
    float s = 0.f;
    for (int t = lane; t < n; t += 64) s += d[t] * d[t];
    s = wave_sum(s);
    // invSQRT: 1/sqrt(sum)  (S/ReductionBase.cuh:158-161), then diagMul
    const float coeff = 1.0f / sqrtf(s);
    for (int t = lane; t < n; t += 64) Dout[(long)k * n + t] = d[t] * coeff;
}

template <int WPE>
hipError_t launch_atb(const float* A, int rows, int colsA, const float* B, long colsB, float* C, int form,
                      dim3 grid, int tilesA, hipStream_t st) {
    const dim3 block(256);
    switch (form) {
        case 0: hipLaunchKernelGGL((at_times_b_kernel<0, WPE>), grid, block, 0, st, A, rows, colsA, B, colsB, C, tilesA); break;
        case 1: hipLaunchKernelGGL((at_times_b_kernel<1, WPE>), grid, block, 0, st, A, rows, colsA, B, colsB, C, tilesA); break;
        case 2: hipLaunchKernelGGL((at_times_b_kernel<2, WPE>), grid, block, 0, st, A, rows, colsA, B, colsB, C, tilesA); break;
        case 3: hipLaunchKernelGGL((at_times_b_kernel<3, WPE>), grid, block, 0, st, A, rows, colsA, B, colsB, C, tilesA); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_at_times_b(const float* A, int rows, int colsA, const float* B, long colsB, float* C, int form,
                             hipStream_t st, int wpe) {
    if (rows <= 0 || colsA <= 0 || colsB <= 0) return hipSuccess;
    // tile-relative offsets are 32-bit: 128 columns of A, B and C
    if ((long)rows * BM > 0x7fffffffL || (long)colsA * BM > 0x7fffffffL) return hipErrorInvalidValue;
    const int tilesA = (colsA + BN - 1) / BN;
    const long tilesB = ((colsB + BM - 1) / BM + 7) / 8 * 8;  // whole groups of 8 (one per XCD)
    const long blocks = tilesB * tilesA;
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks);
    if (wpe == 2) return launch_atb<2>(A, rows, colsA, B, colsB, C, form, grid, tilesA, st);
    return launch_atb<3>(A, rows, colsA, B, colsB, C, form, grid, tilesA, st);
}

hipError_t launch_normcols(const float* Din, float* Dout, int n, int K, hipStream_t st) {
    dim3 grid((K + 3) / 4), block(256);
    hipLaunchKernelGGL(normcols_kernel, grid, block, 0, st, Din, Dout, n, K);
    return hipGetLastError();
}

}  // namespace ksvd
