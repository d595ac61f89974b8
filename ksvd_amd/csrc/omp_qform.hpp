// omp_qform.hpp -- Batch-OMP with the Gram-domain basis held on chip: the
// kernel template and its launchers, instantiated per basis size TT by
// omp_qform_t{8,16,32,48,64}.hip (separate translation units compile in
// parallel) and dispatched by omp_qform.hip.
//
// Same selections and results as parOMP (S/OMP.cu:254-405), computed through
// a different but algebraically identical recurrence.  With the Cholesky
// factor L L^T = G_II of the support and y = L^-1 alpha0_I, the reference
// forms gamma = L^-T y and alpha = alpha0 - G(:,I) gamma every step (a K x i
// product read from G).  Writing Q = G(:,I) L^-T (K x i), column j of Q is
// fixed once atom j is in the support, and
//   alpha^(i)   = alpha^(i-1) - q_i y_i
//   w           = L^-1 G(I, k_i) = (q_0(k_i), ..., q_{i-1}(k_i))
//   l_ii        = sqrt(1 - ||w||^2)
//   q_i         = (G(:, k_i) - sum_j q_j w_j) / l_ii
//   y_i         = alpha^(i-1)(k_i) / l_ii
// so a step needs one column of G, the i columns of Q it already holds, and
// no triangular solve; gamma = L^-T y is solved once at the end.
//
// Layout: one workgroup of W waves per signal.  Wave v, lane l holds atoms
// k = (v*64 + l)*KPL + m, m < KPL, contiguous per lane (vector loads of the
// DtX row and of G columns), with alpha[m] and Q[j][m] (j < TT-1) in VGPRs:
// TT*KPL <= 128 registers, the per-signal working set T*K*4 bytes spread over
// the workgroup's register files (and, for the largest basis, LDS) instead
// of being re-streamed from L2.  All steps are unrolled, so every index into
// Q is a compile-time constant.
// Per step, one barrier: each wave's arg-max of |alpha| (max3 on magnitudes,
// DPP max, ballot) goes to LDS as one 12-byte record (|alpha| bits, atom,
// alpha) with its winning lane's
// row of Q (double-buffered by step parity) -> barrier -> lanes w < W of
// every wave read the W records and reduce them by DPP (same data, same
// order: identical decisions, so every wave leaves the loop at the same
// step), load their slice of G(:, k_i), read w from the winner's row and
// form q_i.
//
// The pivot d = 1 - ||w||^2 is recomputed in FP64 from G(I, k_i) and L when
// it falls below pivot_fp64, as in the wave kernels.
#pragma once
#include "omp_common.hpp"

namespace ksvd {

namespace qform_detail {

// out[m] = p[kb + m] for kb + m < K, else 0.  `vec`: 16-byte (or KPL*4-byte)
// aligned vector loads are legal for this row (K a multiple of min(KPL, 4)).
template <int KPL>
__device__ __forceinline__ void load_slice(const float* __restrict__ p, int kb, int K, bool vec, bool full,
                                           float* out) {
    if (full) {
        if constexpr (KPL == 1) {
            out[0] = p[kb];
        } else if constexpr (KPL == 2) {
            const float2 u = *reinterpret_cast<const float2*>(p + kb);
            out[0] = u.x; out[1] = u.y;
        } else {
#pragma unroll
            for (int q = 0; q < KPL / 4; ++q) {
                const float4 u = *reinterpret_cast<const float4*>(p + kb + 4 * q);
                out[4 * q] = u.x; out[4 * q + 1] = u.y; out[4 * q + 2] = u.z; out[4 * q + 3] = u.w;
            }
        }
        return;
    }
    if (vec && kb + KPL <= K) {
        if constexpr (KPL == 1) {
            out[0] = p[kb];
        } else if constexpr (KPL == 2) {
            const float2 u = *reinterpret_cast<const float2*>(p + kb);
            out[0] = u.x; out[1] = u.y;
        } else {
#pragma unroll
            for (int q = 0; q < KPL / 4; ++q) {
                const float4 u = *reinterpret_cast<const float4*>(p + kb + 4 * q);
                out[4 * q] = u.x; out[4 * q + 1] = u.y; out[4 * q + 2] = u.z; out[4 * q + 3] = u.w;
            }
        }
    } else {
#pragma unroll
        for (int m = 0; m < KPL; ++m) out[m] = (kb + m < K) ? p[kb + m] : 0.f;
    }
}

// Basis columns j < qform_reg_cols stay in VGPRs.  At KPL = 4, TT = 32 in
// sparsity mode the last 9 columns live in LDS instead, lane-contiguous
// float4s per column and wave, so that a wave fits 128 VGPRs (88 for the
// register columns, ~40 others): four waves per SIMD (four signals per CU
// at four waves per signal) instead of three, for 9 KB of LDS per wave
// (with the packed L and the exchange rows a workgroup stays under 40 KB, a
// quarter of the CU's 160 KB).  The other large bases would need more LDS
// than a CU has at four waves per SIMD (KPL = 8: 2 KB per column and wave;
// TT = 64: 17+ columns), and error mode keeps FP64 stopping state in
// registers as well: those keep the whole basis in registers, sized for
// 512 / (TT*KPL + 3*KPL + 26) waves per SIMD (Q, three KPL-wide arrays and
// ~26 others): three waves, or four for the 48-slot basis (TT*KPL = 96, a
// few spilled registers, measured faster than three waves without spills).
constexpr bool qform_split(int kpl, int tt, bool eps) { return kpl == 4 && tt == 32 && !eps; }
// |alpha| >= this float (bits) <=> (double)fl32(alpha^2) >= 1e-14, the
// reference's stop test (S/OMP.cu:292), fl32 squaring being monotone:
// 0x33D6BF96 is the smallest float whose rounded square reaches 1e-14.
constexpr int kAbsAlphaMin = 0x33D6BF96;
constexpr int qform_reg_cols(int kpl, int tt, bool eps) { return qform_split(kpl, tt, eps) ? 22 : tt - 1; }
constexpr int qform_waves_per_eu(int kpl, int tt, bool eps) {
    return tt * kpl < 96 ? 1 : qform_split(kpl, tt, eps) ? 4 : 512 / (tt * kpl + 3 * kpl + 26);
}

// KPL lane-contiguous floats from / to LDS
template <int KPL>
__device__ __forceinline__ void lds_get(const float* p, float* out) {
    if constexpr (KPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < KPL / 4; ++q) {
            const float4 u = reinterpret_cast<const float4*>(p)[q];
            out[4 * q] = u.x; out[4 * q + 1] = u.y; out[4 * q + 2] = u.z; out[4 * q + 3] = u.w;
        }
    } else if constexpr (KPL == 2) {
        const float2 u = *reinterpret_cast<const float2*>(p);
        out[0] = u.x; out[1] = u.y;
    } else {
        out[0] = p[0];
    }
}
template <int KPL>
__device__ __forceinline__ void lds_put(float* p, const float* in) {
    if constexpr (KPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < KPL / 4; ++q)
            reinterpret_cast<float4*>(p)[q] = make_float4(in[4 * q], in[4 * q + 1], in[4 * q + 2], in[4 * q + 3]);
    } else if constexpr (KPL == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(in[0], in[1]);
    } else {
        p[0] = in[0];
    }
}

// max_m |a[m]| with the magnitudes taken by the max instructions' abs source
// modifiers: one v_max3_f32 per two further values.  (fmaxf(fabsf(.), .)
// compiles to a canonicalising max(x, x) per operand first.)  Same value as
// the fmaxf chain: maxNum returns one of its operands, a NaN only when all
// are NaN, and no sNaN reaches it (every operand is a computed value).
template <int KPL>
__device__ __forceinline__ float abs_max(const float* a) {
    if constexpr (KPL == 1) {
        return fabsf(a[0]);
    } else if constexpr (KPL == 2) {
        float t;
        asm("v_max_f32_e64 %0, |%1|, |%2|" : "=v"(t) : "v"(a[0]), "v"(a[1]));
        return t;
    } else {
        float t;
        asm("v_max3_f32 %0, |%1|, |%2|, |%3|" : "=v"(t) : "v"(a[0]), "v"(a[1]), "v"(a[2]));
#pragma unroll
        for (int m = 3; m < KPL; m += 2) {
            if (m + 1 < KPL)
                asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(t) : "v"(t), "v"(a[m]), "v"(a[m + 1]));
            else
                asm("v_max_f32_e64 %0, %1, |%2|" : "=v"(t) : "v"(t), "v"(a[m]));
        }
        return t;
    }
}

template <int KPL, int TT, int W, bool EPS>
__global__ __launch_bounds__(64 * W, qform_waves_per_eu(KPL, TT, EPS)) void omp_qform_kernel(OmpArgs a) {
    // L packed column by column (lower triangle): L(r, c), r >= c, at Lp(r, c)
    auto Lp = [](int r, int c) { return c * TT - ((c * (c - 1)) >> 1) + (r - c); };
    constexpr int UG = 4;        // rows of Q per group (one float4 of w)
    constexpr int NG = TT / UG;  // groups of UG support slots
    __shared__ float Ls[TT * (TT + 1) / 2];
    __shared__ float ys[TT];       // y = L^-1 alpha0_I
    // per step parity: each wave's candidate (|alpha|, atom, signed alpha) and
    // the candidate's row of Q (zero beyond the step), which is w if it wins
    __shared__ int3 cand[2][W];
    __shared__ __align__(16) float wrow[2][W][TT];  // rows < QR used
    __shared__ __align__(16) float wfix[W][TT];  // per wave: w recomputed in FP64 (zero beyond the step)
    // columns QR .. QT-1 of Q: column j of wave v at Qs[v][j - QR][lane*KPL + m]
    constexpr int QT = TT - 1;
    constexpr int QR = qform_reg_cols(KPL, TT, EPS);
    constexpr int QL = QT - QR;
    __shared__ __align__(16) float Qs[W][QL > 0 ? QL : 1][64 * KPL];
    extern __shared__ unsigned char slot[];      // K bytes, dense output only
    const int lane = threadIdx.x & 63;
    const int v = threadIdx.x >> 6;
    const int K = a.K, T = a.T;
    const long s = blockIdx.x;
    const int kb = (v * 64 + lane) * KPL;
    const float* __restrict__ G = a.G;
    const bool vec = (K % (KPL < 4 ? KPL : 4)) == 0;
    const bool full = vec && 64 * W * KPL <= K;

    // alpha of this lane's atoms (0 for padding atoms k >= K: they sit above
    // every real atom, so with the lowest-index tie rule they never win
    // against one)
    float al[KPL];
    {
        // the DtX row is read once: non-temporal, so it does not evict G
        const float* __restrict__ p = a.DtX + s * (long)K;
        if (vec && kb + KPL <= K) {
            if constexpr (KPL == 1) {
                al[0] = __builtin_nontemporal_load(p + kb);
            } else if constexpr (KPL == 2) {
                al[0] = __builtin_nontemporal_load(p + kb);
                al[1] = __builtin_nontemporal_load(p + kb + 1);
            } else {
#pragma unroll
                for (int q = 0; q < KPL / 4; ++q) {
                    const f32x4 u = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p + kb + 4 * q));
                    al[4 * q] = u[0]; al[4 * q + 1] = u[1]; al[4 * q + 2] = u[2]; al[4 * q + 3] = u[3];
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < KPL; ++m) al[m] = (kb + m < K) ? p[kb + m] : 0.f;
        }
    }
    // columns 0 .. QR-1 of Q (the column of step TT-1, necessarily the last
    // step since T <= TT, is never read)
    float Q[QR][KPL];  // column j written at step j, read only by later steps
    for (int j = threadIdx.x; j < 2 * W * TT; j += 64 * W) (&wrow[0][0][0])[j] = 0.f;
    for (int j = threadIdx.x; j < W * TT; j += 64 * W) (&wfix[0][0])[j] = 0.f;
    Support sp;
    double xn2 = 0.0, e2 = 0.0, ysum = 0.0;
    if constexpr (EPS) {
        xn2 = a.xnorm2[s];
        e2 = (double)a.eps2[s];
    }
    unsigned n_redo = 0;
    int brk = -1;  // atom marked in SELECTED_ATOMS by a pivot break (S/OMP.cu:313,331-333)
    __syncthreads();

    // Steps fully unrolled: i is a compile-time constant, so every index into
    // Q and every loop over the support is static (no compare chains).
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < TT; ++i) {
        {
            // groups of UG rows of Q holding slots < i
            const int ng = (i + UG - 1) / UG;
            if (i >= T) break;
            // error mode: ||x - D_I gamma_I||^2 = ||x||^2 - ||y||^2
            if constexpr (EPS) {
                if (xn2 - ysum <= e2) break;
            }
            const int b = i & 1;  // step parity
            // this wave's arg-max of |alpha|: largest value, lowest atom on ties
            // (maxabs, S/OMP.cu:201-213; lanes own increasing atom ranges).
            // |alpha| >= 0, so its bit pattern orders like a signed int; the
            // magnitudes come from the max instructions' abs modifiers
            const float mf = abs_max<KPL>(al);
            const int bv = __float_as_int(mf);
            const int M = wave_max_i32_bcast(bv);
            const int wl = __builtin_ctzll(__builtin_amdgcn_ballot_w64(bv == M));
            if (lane == wl) {
                // lowest slot holding the lane's maximum
                int bm = KPL - 1;
#pragma unroll
                for (int m = KPL - 2; m >= 0; --m)
                    if (fabsf(al[m]) == mf) bm = m;
                float av = al[0];
#pragma unroll
                for (int m = 1; m < KPL; ++m)
                    if (m == bm) av = al[m];
                cand[b][v] = make_int3(M, kb + bm, __float_as_int(av));
                // rows j < min(i, QR) of Q for atom kb + bm (groups of UG; rows >= i
                // are zero); rows QR.. are read from Qs by the other waves
#pragma unroll
                for (int m = 0; m < KPL; ++m) {
                    if (m == bm) {
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
                            if (g < ng && UG * g < QR) {
#pragma unroll
                                for (int e = 0; e < UG; ++e) {
                                    const int j = UG * g + e;
                                    wrow[b][v][j] = (j < i && j < QR) ? Q[j < QR ? j : 0][m] : 0.f;
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();

            // the block's arg-max over the W records, lanes w < W holding
            // record w: max by DPP, then the lowest wave holding it (the lowest
            // atom among equal values)
            // every lane reads record lane mod W (broadcast reads): each group
            // of W lanes holds all W records, so the lowest lane holding the
            // maximum is below W
            const int3 rec = cand[b][lane & (W - 1)];
            const int cv = rec.x, ck = rec.y, ca = rec.z;
            int mx = cv;
            if constexpr (W >= 2) mx = max(mx, dpp_i<0xB1>(mx));
            if constexpr (W >= 4) mx = max(mx, dpp_i<0x4E>(mx));
            if constexpr (W >= 8) mx = max(mx, dpp_i<0x141>(mx));
            const int best = __builtin_amdgcn_readfirstlane(mx);
            const int vo = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(cv == best));
            const int pos = __builtin_amdgcn_readlane(ck, vo);
            const float apos = __int_as_float(__builtin_amdgcn_readlane(ca, vo));
            // stop: inner product too small, or selected twice (S/OMP.cu:292,304);
            // pos >= K: only padding atoms left (every real alpha NaN or 0)
            if ((unsigned)pos >= (unsigned)K ||
                (unsigned)(best - kAbsAlphaMin) > (unsigned)(0x7F800000 - kAbsAlphaMin))
                break;
            if ((__builtin_amdgcn_ballot_w64(sp.ind == pos) & ((1ull << i) - 1ull)) != 0ull) break;
            float g[KPL];
            load_slice<KPL>(G + (long)pos * K, kb, K, vec, full, g);
            // w_j = q_j(pos), one per lane j < i; d = 1 - ||w||^2 from a
            // lane-parallel sum
            const float* wsrc = wrow[b][vo];
            float wv = 0.f;  // row i of L (written by wave 0)
            if (lane < i) {
                const float* wp = wsrc + (lane < TT ? lane : 0);
                // rows QR <= j < i <= QT: the winner's own LDS column (its lane
                // and slot are pos mod 64*KPL)
                if constexpr (QL > 0)
                    if (lane >= QR) wp = &Qs[vo][lane - QR][pos & (64 * KPL - 1)];
                wv = *wp;
            }
            const float d = 1.f - lanes_sum_bcast<(TT <= 16 ? 1 : TT <= 32 ? 2 : 4)>(wv * wv);
            double dd = (double)d;
            const bool redo = i > 0 && d < a.pivot_fp64;  // both FP32: exact
            if (redo) {
                // ill-conditioned: w = L^-1 G(I, pos) and d again in FP64
                // (backsubt_indexed_append_L / sumSquareValues, S/OMP.cu:169-182, 221-228)
                n_redo++;
                double b64 = (lane < i) ? (double)G[(long)pos * K + sp.ind] : 0.0, w64 = 0.0;
#pragma unroll 1
                for (int j = 0; j < i; ++j) {
                    const double wj = readlane(b64, j) / (double)Ls[Lp(j, j)];
                    if (lane == j) w64 = wj;
                    if (lane > j && lane < i) b64 = b64 - (double)Ls[Lp(lane, j)] * wj;
                }
                dd = uniform(1.0 - wave_sum((lane < i) ? w64 * w64 : 0.0));
                wv = (float)w64;
                if (lane < i) wfix[v][lane] = wv;
                wave_lds_fence();
                wsrc = wfix[v];
            }
            if (dd <= 1e-14) {  // S/OMP.cu:331-333
                brk = pos;
                break;
            }
            // l_ii, 1/l_ii and y_i, and row i of L, before the Q w pass (w_lane
            // and the pivot are dead during it).  1/l_ii from v_rsq_f32 (1 ulp)
            // on the FP32 pivot, from FP64 after a recompute
            float rl, l;
            if (redo) {
                const double sq = sqrt(dd);
                l = (float)sq;
                rl = (float)(1.0 / sq);
            } else {
                rl = __builtin_amdgcn_rsqf(d);
                l = d * rl;
            }
            const float y = apos * rl;
            // w_j for the LDS-held rows, now (scalar registers), so that w_lane is
            // dead during the Q w pass
            float wq[QL > 0 ? QL : 1];
#pragma unroll
            for (int j = QR; j < QT; ++j)
                if (j < i) wq[j - QR] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), j));
            if (v == 0) {
                if (lane < i) Ls[Lp(i, lane)] = wv;
                if (lane == i) { Ls[Lp(i, i)] = l; ys[i] = y; }
            }
            if constexpr (EPS) ysum += (double)y * (double)y;
            // acc = Q w (rows j < i)
            float acc[KPL];
#pragma unroll
            for (int m = 0; m < KPL; ++m) acc[m] = 0.f;
#pragma unroll
            for (int gg = 0; gg < NG; ++gg) {
                if (gg < ng && UG * gg < QR) {
                    const float4 w4 = *reinterpret_cast<const float4*>(wsrc + UG * gg);
                    const int j = UG * gg;
                    const float wj[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (j + e < i && j + e < QR) {
#pragma unroll
                            for (int m = 0; m < KPL; ++m) acc[m] = fmaf(Q[j + e][m], wj[e], acc[m]);
                        }
                }
            }
            // rows QR <= j < i from LDS
#pragma unroll
            for (int j = QR; j < QT; ++j) {
                if (j < i) {
                    const float wj = wq[j - QR];
                    float qv[KPL];
                    lds_get<KPL>(&Qs[v][j - QR][lane * KPL], qv);
#pragma unroll
                    for (int m = 0; m < KPL; ++m) acc[m] = fmaf(qv[m], wj, acc[m]);
                }
            }
            // q_i = (G(:, pos) - Q w) / l_ii,  alpha -= q_i y_i
#pragma unroll
            for (int m = 0; m < KPL; ++m) {
                const float q = (g[m] - acc[m]) * rl;
                al[m] = al[m] - q * y;
                acc[m] = q;
            }
            // Q(:, i) = q
            if (i < QR) {
#pragma unroll
                for (int m = 0; m < KPL; ++m) Q[i < QR ? i : 0][m] = acc[m];
            } else if (i < QT) {
                lds_put<KPL>(&Qs[v][i > QR ? i - QR : 0][lane * KPL], acc);
            }
            // support slot i <- pos (lane i), one v_writelane
            asm("v_writelane_b32 %0, %1, %2" : "+v"(sp.ind) : "s"(pos), "n"(i < 64 ? i : 0));
            cnt = i + 1;
        }
    }
    const int i = cnt;

    if (v == 0) {
        wave_lds_fence();
        // gamma = L^-T y, lane-parallel backward sweep (backsubst_Lt, S/OMP.cu:126-138).
        // Wave 0 runs it alone while the workgroup's LDS waits for the next
        // signal, so its chain is kept short: the reciprocals 1/L(r, r) are
        // formed in parallel first (lane r), each step multiplies instead of
        // dividing, and the steps are unrolled (guarded by the count), so
        // the rows of L are read ahead of the chain.
        const int ld = lane < TT ? lane : 0;
        float rhs = (lane < i) ? ys[ld] : 0.f;
        const float rdiag = (lane < i) ? 1.0f / Ls[Lp(ld, ld)] : 0.f;
#pragma unroll
        for (int r = TT - 1; r >= 0; --r) {
            if (r < i) {
                const float lrow = (lane < r) ? Ls[Lp(r, ld)] : 0.f;  // L(r, lane)
                const float gr = readlane(rhs, r) * readlane(rdiag, r);
                if (lane == r) sp.gam = gr;
                if (lane < r) rhs = rhs - lrow * gr;
            }
        }
        // streaming stores: the results are not re-read by this kernel and
        // should not displace G in L2
        if (a.idx && lane < T) {
            __builtin_nontemporal_store((lane < i) ? sp.ind : -1, a.idx + s * T + lane);
            __builtin_nontemporal_store((lane < i) ? sp.gam : 0.f, a.val + s * T + lane);
        }
        if (a.cnt && lane == 0) __builtin_nontemporal_store(i, a.cnt + s);
        if (a.c && lane < i) a.c[s * T + lane] = sp.gam;
        if (a.stats && lane == 0 && n_redo) atomicAdd(a.stats + 1, (unsigned long long)n_redo);
        if (a.Lfac) {
            // L of the accepted atoms, packed column by column with pitch T:
            // L(r, c) at c*T - c(c-1)/2 + (r - c); zero beyond the count
            float* Lo = a.Lfac + s * (long)(T * (T + 1) / 2);
            for (int c = 0; c < T; ++c) {
                const int r = c + lane;
                if (r < T) Lo[c * T - ((c * (c - 1)) >> 1) + lane] = (r < i && c < i) ? Ls[Lp(r, c)] : 0.f;
            }
        }
    }
    if (a.sel) {
        // SELECTED_ATOMS column: slot+1 of each support atom, and the atom a
        // pivot break marked before stopping
        for (int k = threadIdx.x; k < K; k += 64 * W) slot[k] = 0;
        __syncthreads();
        if (v == 0) {
            if (lane < i) slot[sp.ind] = (unsigned char)(lane + 1);
            if (lane == 0 && brk >= 0) slot[brk] = (unsigned char)(i + 1);
        }
        __syncthreads();
        float* selp = a.sel + s * (long)K;
        for (int k = threadIdx.x; k < K; k += 64 * W) selp[k] = (float)slot[k];
    }
}

template <int KPL, int TT, int W>
hipError_t launch_q(const OmpArgs& a, hipStream_t st) {
    const size_t lds = a.sel ? (size_t)((a.K + 15) & ~15) : 0;
    if (a.eps2)
        hipLaunchKernelGGL((omp_qform_kernel<KPL, TT, W, true>), dim3((unsigned)a.N), dim3(64 * W), lds, st, a);
    else
        hipLaunchKernelGGL((omp_qform_kernel<KPL, TT, W, false>), dim3((unsigned)a.N), dim3(64 * W), lds, st, a);
    return hipGetLastError();
}

template <int KPL, int TT>
hipError_t launch_qw(const OmpArgs& a, int W, hipStream_t st) {
    switch (W) {
        case 1: return launch_q<KPL, TT, 1>(a, st);
        case 2: return launch_q<KPL, TT, 2>(a, st);
        case 4: return launch_q<KPL, TT, 4>(a, st);
        case 8: return launch_q<KPL, TT, 8>(a, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace qform_detail

}  // namespace ksvd
