/*
 * ksvd_oracle.c -- CPU restatement of the reference's Batch-OMP / K-SVD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke check
 * and the CPU-baseline leg of bench.py load it.  The product (libksvd_amd.so)
 * never links, loads or calls it, and no product path falls back to it.
 *
 * Built twice from this one source (see oracle/Makefile):
 *   liboracle_f64.so  REAL=double  -- the FP64 oracle every GPU result is checked against
 *   liboracle_f32.so  REAL=float   -- the FP32 "port" timed as the CPU baseline
 *
 * Matrix layouts follow the reference (all column-major, MATLAB order):
 *   X n x N, D n x K, G K x K, DtX K x N, SELECTED_ATOMS K x N (slot+1 or 0),
 *   c T x N.  The compact Gamma used by the engine is idx/val (T x N) + cnt (N).
 *
 * Citations are to the reference tree; S/ = K-SVD/K-SVD/.
 *
 * Parity pins (what checks each function -- tests/test_oracle.py):
 *   or_normcols        unit column norms; closed form on scaled identity
 *   or_at_times_b      brute force long-double triple loop; D = I gives G = I
 *   or_batch_omp       orthonormal-D closed form (top-T of |DtX|), exact recovery
 *                      (mu(D) < 1/(2T-1)), textbook OMP by explicit least squares,
 *                      invariants D_I^T r = 0 and L L^T = G_II, zero signal,
 *                      duplicated-atom pivot break, reference fixture structure
 *   or_batch_omp_eps   residual <= eps on stop, monotone residual, T_max cap,
 *                      equality with sparsity mode when eps = 0
 *   or_dict_update     unit atoms, per-atom error non-increasing, power iteration
 *                      converges to the rank-1 SVD of E_k (numpy.linalg.svd),
 *                      support unchanged, unused-atom replacement rule
 *   or_error           closed form on constructed Gamma
 *   or_clear_dict      coherent pair replaced, usecount < 4 replaced,
 *                      replaced atoms skipped
 *   or_ls_adjoint      FP64 central finite differences of or_ls_fixed
 *   or_perm            golden permutation vectors (SplitMix64 Fisher-Yates)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif
#ifndef SUFFIX
#define SUFFIX _f64
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

#define SQR(x) ((x) * (x))

/* ------------------------------------------------------------------------- */
/* threading                                                                  */
/* ------------------------------------------------------------------------- */
int FN(or_set_threads)(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
    return omp_get_max_threads();
#else
    (void)t;
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* normcols: D_out = D_in * diag(1/sqrt(sum(D_in.^2)))                        */
/* S/KSVD.cu:267-286, S/ReductionKernels.cu:179-190 (invSQRT at              */
/* S/ReductionBase.cuh:158-161), S/KSVD.cu:388-395 (diagMul)                  */
/* ------------------------------------------------------------------------- */
void FN(or_normcols)(int n, int K, const REAL* Din, REAL* Dout) {
    for (int k = 0; k < K; ++k) {
        REAL sum = 0;
        for (int t = 0; t < n; ++t) sum += SQR(Din[(size_t)k * n + t]);
        REAL coeff = (REAL)1.0 / sqrt(sum);
        for (int t = 0; t < n; ++t) Dout[(size_t)k * n + t] = Din[(size_t)k * n + t] * coeff;
    }
}

/* ------------------------------------------------------------------------- */
/* C (colsA x colsB) = A^T B, column-major, lda = ldb = rows                  */
/* S/Multiplication.cu:40-63 (AT_times_B: cublas?gemm(T,N))                  */
/* ------------------------------------------------------------------------- */
void FN(or_at_times_b)(int rows, int colsA, int colsB, const REAL* A, const REAL* B, REAL* C) {
#pragma omp parallel for schedule(static)
    for (long j = 0; j < colsB; ++j) {
        const REAL* b = B + (size_t)j * rows;
        for (int m = 0; m < colsA; ++m) {
            const REAL* a = A + (size_t)m * rows;
            REAL sum = 0;
            for (int t = 0; t < rows; ++t) sum += a[t] * b[t];
            C[(size_t)j * colsA + m] = sum;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* The same product, blocked for the CPU baseline: OpenMP over blocks of 64  */
/* columns of B, 2 x 4 dot products per register block so that each loaded  */
/* element feeds several products (the compiler vectorises the shared-      */
/* dimension loop; the FP32 port is built with reassociation allowed).  It  */
/* rounds differently from or_at_times_b; the FP32 port uses it             */
/* (KSVD_PORT), the FP64 checker keeps the plain loop.                      */
/* ------------------------------------------------------------------------- */
static void FN(dot_2x4)(int rows, const REAL* restrict a0, const REAL* restrict a1, const REAL* restrict b0,
                        const REAL* restrict b1, const REAL* restrict b2, const REAL* restrict b3, REAL* out) {
    REAL s00 = 0, s01 = 0, s02 = 0, s03 = 0, s10 = 0, s11 = 0, s12 = 0, s13 = 0;
    for (int t = 0; t < rows; ++t) {
        const REAL x0 = a0[t], x1 = a1[t];
        s00 += x0 * b0[t]; s01 += x0 * b1[t]; s02 += x0 * b2[t]; s03 += x0 * b3[t];
        s10 += x1 * b0[t]; s11 += x1 * b1[t]; s12 += x1 * b2[t]; s13 += x1 * b3[t];
    }
    out[0] = s00; out[1] = s01; out[2] = s02; out[3] = s03;
    out[4] = s10; out[5] = s11; out[6] = s12; out[7] = s13;
}

void FN(or_at_times_b_blocked)(int rows, int colsA, int colsB, const REAL* A, const REAL* B, REAL* C) {
#pragma omp parallel for schedule(dynamic, 1)
    for (long jb = 0; jb < colsB; jb += 64) {
        const long je = jb + 64 < colsB ? jb + 64 : colsB;
        for (int m0 = 0; m0 < colsA; m0 += 2) {
            const int mc = m0 + 1 < colsA ? 2 : 1;
            const REAL* a0 = A + (size_t)m0 * rows;
            const REAL* a1 = A + (size_t)(m0 + mc - 1) * rows;
            for (long j0 = jb; j0 < je; j0 += 4) {
                const int jc = j0 + 4 <= je ? 4 : (int)(je - j0);
                const REAL* bp[4];
                for (int b = 0; b < 4; ++b) bp[b] = B + (size_t)(j0 + (b < jc ? b : 0)) * rows;
                REAL out[8];
                FN(dot_2x4)(rows, a0, a1, bp[0], bp[1], bp[2], bp[3], out);
                for (int a = 0; a < mc; ++a)
                    for (int b = 0; b < jc; ++b) C[(size_t)(j0 + b) * colsA + m0 + a] = out[4 * a + b];
            }
        }
    }
}

#ifdef KSVD_PORT
#define AT_TIMES_B FN(or_at_times_b_blocked)
#else
#define AT_TIMES_B FN(or_at_times_b)
#endif

/* ------------------------------------------------------------------------- */
/* Batch-OMP, one signal.  S/OMP.cu:254-405 (parOMP) with helpers            */
/*   maxabs                    S/OMP.cu:201-213  (strict '>': lowest index)  */
/*   backsubt_indexed_append_L S/OMP.cu:169-182                              */
/*   sumSquareValues           S/OMP.cu:221-228  (returns 1 - sum)           */
/*   backsubt_indexed_L        S/OMP.cu:147-159                              */
/*   backsubst_Lt              S/OMP.cu:126-138                              */
/*   addMul                    S/OMP.cu:234-247                              */
/* L is T x T column-major, L[col*T + row], as in the reference.            */
/*                                                                           */
/* Error mode (no reference counterpart beyond the comment at S/KSVD.cu:79): */
/* when eps2 >= 0 the loop also stops once ||r||^2 <= eps2, with            */
/* ||r||^2 = ||x||^2 - alpha0_I^T gamma_I evaluated in double, checked      */
/* before each selection (OMP-Box "error" coding mode), and T acts as T_max. */
/*                                                                           */
/* gap[j] (optional) records, at step j, the absolute gap between the        */
/* largest and second-largest |alpha_k| -- the oracle's tie diagnostic.     */
/* Returns the accepted atom count; *pending gets the atom marked in the    */
/* slot map when the loop broke on the pivot test (else -1).                */
/* ------------------------------------------------------------------------- */
static int FN(omp_one)(int K, int T, const REAL* a0, const REAL* G, REAL* alpha, REAL* L, REAL* chol,
                       int* ind, REAL* sel /* K, slot+1 or 0 */, REAL* cs /* T */,
                       double xnorm2, double eps2, double* gap, int* pending) {
    int i = 0;
    *pending = -1;
    for (int k = 0; k < K; ++k) sel[k] = 0; /* zeroOut S/OMP.cu:276 */
    if (gap)
        for (int j = 0; j < T; ++j) gap[j] = -1.0;
    while (i < T) {
        if (eps2 >= 0.0) {
            double r2 = xnorm2;
            for (int j = 0; j < i; ++j) r2 -= (double)a0[ind[j]] * (double)cs[j];
            if (r2 <= eps2) break;
        }
        const REAL* cur = (i == 0) ? a0 : alpha;
        /* pos = maxabs(cur, K) */
        REAL mx = SQR(cur[0]);
        int pos = 0;
        for (int k = 1; k < K; ++k) {
            REAL v = SQR(cur[k]);
            if (v > mx) { mx = v; pos = k; }
        }
        if (gap) {
            double b1 = -1.0, b2 = -1.0;
            for (int k = 0; k < K; ++k) {
                double v = fabs((double)cur[k]);
                if (v > b1) { b2 = b1; b1 = v; }
                else if (v > b2) b2 = v;
            }
            gap[i] = b1 - b2;
        }
        /* stop: same atom twice, or inner product too small  S/OMP.cu:292,304 */
        if (sel[pos] != 0 || SQR(cur[pos]) < 1e-14) break;
        ind[i] = pos;
        sel[pos] = (REAL)(i + 1); /* S/OMP.cu:313 -- marked before the pivot test */
        if (i == 0) {
            L[0] = 1; /* S/OMP.cu:317 */
        } else {
            /* backsubt_indexed_append_L(L, G + pos*K, ind, T, i, chol) */
            const REAL* b = G + (size_t)pos * K;
            for (int r = 0; r < i; ++r) {
                REAL rhs = b[ind[r]];
                for (int j = 0; j < r; ++j) rhs -= L[j * T + r] * chol[j];
                chol[r] = rhs / L[r * T + r];
                L[r * T + i] = chol[r];
            }
            /* sumSquareValues(L, i, T) = 1 - sum_k L[k*T + i]^2 */
            REAL sum = 0;
            for (int k = 0; k < i; ++k) sum += SQR(L[k * T + i]);
            sum = 1 - sum;
            if (sum <= 1e-14) { *pending = pos; break; } /* S/OMP.cu:331-333 */
            L[i * T + i] = sqrt(sum);
        }
        i++;
        /* backsubt_indexed_L(L, a0, chol, ind, T, i) */
        for (int r = 0; r < i; ++r) {
            REAL rhs = a0[ind[r]];
            for (int j = 0; j < r; ++j) rhs -= L[j * T + r] * chol[j];
            chol[r] = rhs / L[r * T + r];
        }
        /* backsubst_Lt(L, chol, cs, T, i) */
        for (int r = i; r >= 1; --r) {
            REAL rhs = chol[r - 1];
            for (int j = r; j < i; ++j) rhs -= L[(r - 1) * T + j] * cs[j];
            cs[r - 1] = rhs / L[(r - 1) * T + r - 1];
        }
        /* addMul(G, K, i, cs, ind, alpha, a0):  alpha = a0 - G(:,ind)*cs.     */
        /* Loops interchanged (column-axpy form) so the k loop vectorises; each */
        /* alpha[k] still accumulates its sum over j in ascending order.        */
        for (int k = 0; k < K; ++k) alpha[k] = 0;
        for (int j = 0; j < i; ++j) {
            const REAL* gcol = G + (size_t)ind[j] * K;
            const REAL cj = cs[j];
            for (int k = 0; k < K; ++k) alpha[k] += gcol[k] * cj;
        }
        for (int k = 0; k < K; ++k) alpha[k] = a0[k] - alpha[k];
    }
    return i;
}

/*
 * Batch-OMP over N signals.  S/OMP.cu:42-59 (BatchOMP), S/OMP.cu:100-113.
 *   DtX, G      inputs (K x N, K x K)
 *   selected    optional K x N slot map output (reference SELECTED_ATOMS)
 *   c           optional T x N coefficients in selection order; slots beyond
 *               the accepted count are left untouched, as in the reference
 *   idx,val,cnt optional compact outputs (T x N, T x N, N); idx slots beyond
 *               cnt are -1 and val slots 0
 *   xnorm2,eps2 error mode when eps2 != NULL (per-signal ||x||^2, eps^2)
 *   gap         optional T x N tie diagnostic
 *   pending     optional N: atom marked by a pivot break, else -1
 */
void FN(or_batch_omp)(int K, int N, int T, const REAL* DtX, const REAL* G, REAL* selected, REAL* c,
                      int32_t* idx, REAL* val, int32_t* cnt, const REAL* xnorm2, const REAL* eps2,
                      double* gap, int32_t* pending) {
#pragma omp parallel
    {
        REAL* alpha = (REAL*)malloc(sizeof(REAL) * (size_t)K);
        REAL* L = (REAL*)calloc((size_t)T * T, sizeof(REAL));
        REAL* chol = (REAL*)malloc(sizeof(REAL) * (size_t)T);
        REAL* csl = (REAL*)malloc(sizeof(REAL) * (size_t)T);
        REAL* sel = (REAL*)malloc(sizeof(REAL) * (size_t)K);
        int* ind = (int*)malloc(sizeof(int) * (size_t)T);
#pragma omp for schedule(dynamic, 16)
        for (long s = 0; s < N; ++s) {
            REAL* cs = c ? c + (size_t)s * T : csl;
            if (!c)
                for (int j = 0; j < T; ++j) cs[j] = 0;
            REAL* selp = selected ? selected + (size_t)s * K : sel;
            int pend;
            double xn = xnorm2 ? (double)xnorm2[s] : 0.0;
            double e2 = eps2 ? (double)eps2[s] : -1.0;
            int i = FN(omp_one)(K, T, DtX + (size_t)s * K, G, alpha, L, chol, ind, selp, cs, xn, e2,
                                gap ? gap + (size_t)s * T : NULL, &pend);
            if (idx)
                for (int j = 0; j < T; ++j) idx[(size_t)s * T + j] = j < i ? ind[j] : -1;
            if (val)
                for (int j = 0; j < T; ++j) val[(size_t)s * T + j] = j < i ? cs[j] : 0;
            if (cnt) cnt[s] = i;
            if (pending) pending[s] = pend;
        }
        free(alpha); free(L); free(chol); free(csl); free(sel); free(ind);
    }
}

/* ------------------------------------------------------------------------- */
/* Permutation of 0..K-1 for the atom visiting order.                        */
/* The reference shuffles a running buffer in place each iteration with     */
/* unseeded std::random_shuffle (S/RandPermutation.cu:35-42, buffer set to   */
/* the identity once at :48-56).  The engine keeps the running-buffer        */
/* semantics but draws from SplitMix64 (Steele, Lea & Flood, OOPSLA 2014)   */
/* with a Fisher-Yates shuffle (Knuth TAOCP 3.4.2 Algorithm P), j = z mod    */
/* (i+1).  The product's C++ host code implements the same published        */
/* generator; the golden vectors in tests/golden/ pin both.                 */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void FN(or_perm_shuffle)(uint32_t* buf, int K, uint64_t* state) {
    for (int i = K - 1; i >= 1; --i) {
        uint64_t z = splitmix64_next(state);
        int j = (int)(z % (uint64_t)(i + 1));
        uint32_t t = buf[i]; buf[i] = buf[j]; buf[j] = t;
    }
}

/* ------------------------------------------------------------------------- */
/* Dictionary-update state reset: S/KSVD.cu:442-457 (init)                  */
/* ------------------------------------------------------------------------- */
void FN(or_du_init)(int N, int K, int32_t* unused, int32_t* us_counter, int32_t* bitmap,
                    int32_t* replaced) {
    for (int s = 0; s < N; ++s) { unused[s] = s; bitmap[s] = 1; }
    for (int k = 0; k < K; ++k) replaced[k] = 0;
    *us_counter = 0;
}

/* ---- helpers over the compact Gamma -------------------------------------- */
/* Per-atom lists in ascending signal order.  The reference builds them with */
/* atomicAdd slots (S/DictionaryUpdate.cu:378-395), i.e. in a nondeterministic */
/* order; the engine and this oracle both use ascending signal order.        */
typedef struct { int32_t* ptr; int32_t* pos; } FN(lists_t);

static FN(lists_t) FN(build_lists)(int K, int N, int T, const int32_t* idx, const int32_t* cnt) {
    FN(lists_t) l;
    l.ptr = (int32_t*)calloc((size_t)K + 1, sizeof(int32_t));
    for (long s = 0; s < N; ++s)
        for (int j = 0; j < cnt[s]; ++j) l.ptr[idx[s * T + j] + 1]++;
    for (int k = 0; k < K; ++k) l.ptr[k + 1] += l.ptr[k];
    l.pos = (int32_t*)malloc(sizeof(int32_t) * ((size_t)l.ptr[K] + 1));
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
    for (int k = 0; k < K; ++k) cur[k] = l.ptr[k];
    for (long s = 0; s < N; ++s)
        for (int j = 0; j < cnt[s]; ++j) l.pos[cur[idx[s * T + j]]++] = (int32_t)(s * T + j);
    free(cur);
    return l;
}

/* x_hat = D * gamma_s (sparse over the signal's support) */
static void FN(reconstruct)(int n, int T, const REAL* D, const int32_t* idx, const REAL* val, int c,
                            long s, REAL* out) {
    for (int t = 0; t < n; ++t) out[t] = 0;
    for (int j = 0; j < c; ++j) {
        const REAL* d = D + (size_t)idx[s * T + j] * n;
        REAL g = val[s * T + j];
        for (int t = 0; t < n; ++t) out[t] += d[t] * g;
    }
}

/* ------------------------------------------------------------------------- */
/* Dictionary update (approximate K-SVD, one sweep).                         */
/* S/DictionaryUpdate.cu:45-203 (updateDictionary)                          */
/*   precompute ||gamma_k||^2 and Y_k = X(:,I_k) gamma_k                     */
/*      S/ReductionKernels.cu:31-98                                          */
/*   per atom pJ = perm[j]:                                                   */
/*     specialCase  S/DictionaryUpdate.cu:238-302, S/ReductionKernels.cu:436- */
/*                  539, S/DictionaryUpdate.cu:406-427 (SCupdate)            */
/*     v = Gamma(:,I) gamma^T, v[pJ] = ||gamma||^2                            */
/*                  S/ReductionKernels.cu:105-168                            */
/*     a = Y - D v + d_pJ ||gamma||^2   S/Multiplication.cu:177-193          */
/*     w = a^T X_I, u = a^T D, ||a||    S/ReductionKernels.cu:196-250        */
/*     gamma_new = (w - u^T Gamma_I + u_pJ gamma)/||a||, d_pJ = a/||a||       */
/*                  S/Multiplication.cu:198-225                              */
/* The compact Gamma (idx/cnt fixed, val updated in place) replaces the      */
/* reference's dense SELECTED_ATOMS; sums over Gamma's zero entries vanish.  */
/* ------------------------------------------------------------------------- */
void FN(or_dict_update)(int n, int K, int N, int T, const REAL* X, REAL* D, const int32_t* idx,
                        REAL* val, const int32_t* cnt, const uint32_t* perm, int32_t* unused,
                        int32_t* us_counter, int32_t* bitmap, int32_t* replaced, int max_signals) {
    FN(lists_t) l = FN(build_lists)(K, N, T, idx, cnt);
    REAL* gn2 = (REAL*)calloc((size_t)K, sizeof(REAL));
    REAL* Y = (REAL*)calloc((size_t)K * n, sizeof(REAL));
    /* the parallel loops below each give every output to one thread, which
       sums in the serial order: the results do not depend on the thread count */
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < K; ++k) {
        REAL sum = 0;
        for (int m = l.ptr[k]; m < l.ptr[k + 1]; ++m) sum += SQR(val[l.pos[m]]);
        gn2[k] = sum;
        for (int t = 0; t < n; ++t) {
            REAL acc = 0;
            for (int m = l.ptr[k]; m < l.ptr[k + 1]; ++m) {
                long s = l.pos[m] / T;
                acc += X[(size_t)s * n + t] * val[l.pos[m]];
            }
            Y[(size_t)k * n + t] = acc;
        }
    }
    REAL* v = (REAL*)malloc(sizeof(REAL) * (size_t)K);
    REAL* u = (REAL*)malloc(sizeof(REAL) * (size_t)K);
    REAL* a = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* xh = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* gnew = (REAL*)malloc(sizeof(REAL) * ((size_t)N + 1));
    for (int jj = 0; jj < K; ++jj) {
        int p = (int)perm[jj];
        int Np = l.ptr[p + 1] - l.ptr[p];
        if (Np == 0) {
            /* special case: candidates are the first min(N, max_signals) live   */
            /* entries of the unused list (identity myGenerator,                 */
            /* S/ReductionBase.cuh:176-180); error vs the current D*Gamma.       */
            int live = N - *us_counter;
            int M = N < max_signals ? N : max_signals;
            /* every signal already replaced an atom in this sweep: no candidate */
            /* left, the atom stays as it is (the reference's reduction over an  */
            /* empty range returns index 0 and it then reads unused[-1]: no      */
            /* defined result to follow)                                         */
            if (live <= 0) continue;
            int best = 0;
            REAL bestv = 0;
            for (int b = 0; b < M && b < live; ++b) {
                long sig = unused[b];
                FN(reconstruct)(n, T, D, idx, val, cnt[sig], sig, xh);
                REAL e = 0;
                for (int t = 0; t < n; ++t) e += SQR(X[(size_t)sig * n + t] - xh[t]);
                if (e > bestv) { bestv = e; best = b; } /* keeps lowest index on ties */
            }
            long sig = unused[best];
            REAL nrm = 0;
            for (int t = 0; t < n; ++t) nrm += SQR(X[(size_t)sig * n + t]);
            nrm = sqrt(nrm);
            /* a zero signal (every candidate error 0): x / ||x|| is undefined,  */
            /* the atom stays as it is                                           */
            if (nrm == 0) continue;
            for (int t = 0; t < n; ++t) D[(size_t)p * n + t] = X[(size_t)sig * n + t] / nrm;
            int counter = *us_counter;
            unused[best] = unused[N - counter - 1];
            *us_counter = counter + 1;
            replaced[p] = 1;
            bitmap[sig] = 0;
            continue;
        }
        /* v = Gamma(:, I_p) gamma_p^T over rows r != p; v[p] = ||gamma_p||^2 */
        for (int r = 0; r < K; ++r) v[r] = 0;
        for (int m = l.ptr[p]; m < l.ptr[p + 1]; ++m) {
            long s = l.pos[m] / T;
            REAL gp = val[l.pos[m]];
            for (int j = 0; j < cnt[s]; ++j) {
                int r = idx[s * T + j];
                if (r != p) v[r] += val[s * T + j] * gp;
            }
        }
        v[p] = gn2[p];
        /* a = Y_p - D v + d_p ||gamma_p||^2 */
#pragma omp parallel for schedule(static) if ((long)n * K > 65536)
        for (int t = 0; t < n; ++t) {
            REAL sum = 0;
            for (int r = 0; r < K; ++r) sum += D[(size_t)r * n + t] * v[r];
            a[t] = Y[(size_t)p * n + t] - sum + D[(size_t)p * n + t] * gn2[p];
        }
        /* u = a^T D (old d_p), ||a|| */
        REAL nrm = 0;
        for (int t = 0; t < n; ++t) nrm += SQR(a[t]);
        nrm = sqrt(nrm);
#pragma omp parallel for schedule(static) if ((long)n * K > 65536)
        for (int r = 0; r < K; ++r) {
            REAL sum = 0;
            for (int t = 0; t < n; ++t) sum += D[(size_t)r * n + t] * a[t];
            u[r] = sum;
        }
        /* gamma_new for every signal on the list */
#pragma omp parallel for schedule(static) if ((long)Np * n > 65536)
        for (int m = l.ptr[p]; m < l.ptr[p + 1]; ++m) {
            long s = l.pos[m] / T;
            REAL w = 0;
            for (int t = 0; t < n; ++t) w += X[(size_t)s * n + t] * a[t];
            REAL sum = 0;
            for (int j = 0; j < cnt[s]; ++j) sum += val[s * T + j] * u[idx[s * T + j]];
            gnew[m - l.ptr[p]] = (w - sum + u[p] * val[l.pos[m]]) / nrm;
        }
        for (int m = l.ptr[p]; m < l.ptr[p + 1]; ++m) val[l.pos[m]] = gnew[m - l.ptr[p]];
        for (int t = 0; t < n; ++t) D[(size_t)p * n + t] = a[t] / nrm;
    }
    free(v); free(u); free(a); free(xh); free(gnew); free(gn2); free(Y);
    free(l.ptr); free(l.pos);
}

/* ------------------------------------------------------------------------- */
/* errorComputation: err_s = ||x_s - D gamma_s||^2, RMSE over all signals,  */
/* then err_s *= bitmap_s.  S/KSVD.cu:235-262, S/ReductionKernels.cu:257-322 */
/* ------------------------------------------------------------------------- */
double FN(or_error)(int n, int K, int N, int T, const REAL* X, const REAL* D, const int32_t* idx,
                    const REAL* val, const int32_t* cnt, const int32_t* bitmap, REAL* err) {
    (void)K;
    double total = 0;
#pragma omp parallel
    {
        REAL* xh = (REAL*)malloc(sizeof(REAL) * (size_t)n);
#pragma omp for reduction(+ : total) schedule(static)
        for (long s = 0; s < N; ++s) {
            FN(reconstruct)(n, T, D, idx, val, cnt[s], s, xh);
            REAL e = 0;
            for (int t = 0; t < n; ++t) e += SQR(X[(size_t)s * n + t] - xh[t]);
            total += (double)e;
            err[s] = e * (REAL)(bitmap ? bitmap[s] : 1);
        }
        free(xh);
    }
    return sqrt(total / ((double)n * (double)N));
}

/* ------------------------------------------------------------------------- */
/* clearDict.  S/KSVD.cu:138-226; use_count S/KSVD.cu:403-415 (squared       */
/* gamma > 1e-7); device_max S/ReductionKernels.cu:393-408; Maximum          */
/* S/ReductionKernels.cu:329-386 (lowest index on ties); EUnorm :413-430;     */
/* replaceDj S/KSVD.cu:422-434.                                              */
/* ------------------------------------------------------------------------- */
void FN(or_clear_dict)(int n, int K, int N, int T, const REAL* X, REAL* D, const int32_t* idx,
                       const REAL* val, const int32_t* cnt, const int32_t* replaced, REAL* err,
                       double mu_thresh, int use_thresh, int32_t* flags_out) {
    int32_t* usecount = (int32_t*)calloc((size_t)K, sizeof(int32_t));
    for (long s = 0; s < N; ++s)
        for (int j = 0; j < cnt[s]; ++j)
            if (SQR(val[s * T + j]) > 1e-7) usecount[idx[s * T + j]]++;
    REAL* Gj = (REAL*)malloc(sizeof(REAL) * (size_t)K);
    double sqr_mu = mu_thresh * mu_thresh;
    for (int j = 0; j < K; ++j) {
        const REAL* dj = D + (size_t)j * n;
        for (int r = 0; r < K; ++r) {
            REAL sum = 0;
            for (int t = 0; t < n; ++t) sum += D[(size_t)r * n + t] * dj[t];
            Gj[r] = sum;
        }
        REAL mx = 0;
        for (int r = 0; r < K; ++r)
            if (r != j && SQR(Gj[r]) > mx) mx = SQR(Gj[r]);
        int flag = ((double)mx > sqr_mu || usecount[j] < use_thresh) && replaced[j] == 0;
        if (flags_out) flags_out[j] = flag;
        if (!flag) continue;
        long best = 0;
        REAL bestv = err[0];
        for (long s = 1; s < N; ++s)
            if (err[s] > bestv) { bestv = err[s]; best = s; }
        err[best] = 0;
        REAL nrm = 0;
        for (int t = 0; t < n; ++t) nrm += SQR(X[(size_t)best * n + t]);
        nrm = sqrt(nrm);
        if (nrm == 0) continue; /* a zero signal: the atom stays as it is */
        for (int t = 0; t < n; ++t) D[(size_t)j * n + t] = X[(size_t)best * n + t] / nrm;
    }
    free(Gj); free(usecount);
}

/* ------------------------------------------------------------------------- */
/* Full loop: KSVD::ExecuteIterations  S/KSVD.cu:55-131                      */
/* D_out receives the working dictionary; idx/val/cnt the last Gamma;       */
/* rmse[it] the per-iteration RMSE printed by the reference.                */
/* ------------------------------------------------------------------------- */
/* ExecuteIterations (S/KSVD.cu:55-131).  eps2 != NULL: the error-constrained */
/* coding the comment at S/KSVD.cu:79 names (sparsecode with a threshold):   */
/* or_batch_omp in error mode with ||x_s||^2 summed in double, T as T_max.   */
int FN(or_ksvd)(int n, int K, int N, int T, const REAL* X, const REAL* Din, REAL* Dout, int iters,
                uint64_t seed, double mu_thresh, int use_thresh, int max_signals, int32_t* idx,
                REAL* val, int32_t* cnt, double* rmse, const REAL* eps2) {
    REAL* xn2 = NULL;
    if (eps2) {
        xn2 = (REAL*)malloc(sizeof(REAL) * (size_t)N);
        for (long s = 0; s < N; ++s) {
            double acc = 0;
            for (int t = 0; t < n; ++t) acc += (double)X[(size_t)s * n + t] * (double)X[(size_t)s * n + t];
            xn2[s] = (REAL)acc;
        }
    }
    FN(or_normcols)(n, K, Din, Dout);
    REAL* G = (REAL*)malloc(sizeof(REAL) * (size_t)K * K);
    REAL* DtX = (REAL*)malloc(sizeof(REAL) * (size_t)K * N);
    REAL* err = (REAL*)malloc(sizeof(REAL) * (size_t)N);
    int32_t* unused = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    int32_t* bitmap = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    int32_t* replaced = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K);
    int32_t usc = 0;
    uint64_t state = seed;
    for (int k = 0; k < K; ++k) perm[k] = (uint32_t)k;
    for (int it = 0; it < iters; ++it) {
        AT_TIMES_B(n, K, K, Dout, Dout, G);
        AT_TIMES_B(n, K, N, Dout, X, DtX);
        FN(or_batch_omp)(K, N, T, DtX, G, NULL, NULL, idx, val, cnt, xn2, eps2, NULL, NULL);
        FN(or_du_init)(N, K, unused, &usc, bitmap, replaced);
        FN(or_perm_shuffle)(perm, K, &state);
        FN(or_dict_update)(n, K, N, T, X, Dout, idx, val, cnt, perm, unused, &usc, bitmap, replaced,
                           max_signals);
        double r = FN(or_error)(n, K, N, T, X, Dout, idx, val, cnt, bitmap, err);
        if (rmse) rmse[it] = r;
        FN(or_clear_dict)(n, K, N, T, X, Dout, idx, val, cnt, replaced, err, mu_thresh, use_thresh,
                          NULL);
    }
    free(G); free(DtX); free(err); free(unused); free(bitmap); free(replaced); free(perm); free(xn2);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Fixed-support least squares and its adjoint (no reference counterpart).   */
/*   gamma_I = (D_I^T D_I)^{-1} D_I^T x   via Cholesky of G_II = D_I^T D_I   */
/*   with g = dL/dgamma_I:  z = G_II^{-1} g, r = x - D_I gamma_I              */
/*     dL/dx   = D_I z                                                       */
/*     dL/dD_I = r z^T - (D_I z) gamma_I^T                                   */
/* ------------------------------------------------------------------------- */
static void FN(chol_factor)(int m, const REAL* A, REAL* L) { /* A, L: m x m row-major */
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            REAL s = A[i * m + j];
            for (int k = 0; k < j; ++k) s -= L[i * m + k] * L[j * m + k];
            L[i * m + j] = (i == j) ? sqrt(s) : s / L[j * m + j];
        }
}

static void FN(chol_solve)(int m, const REAL* L, const REAL* b, REAL* x, REAL* y) {
    for (int i = 0; i < m; ++i) {
        REAL s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * m + k] * y[k];
        y[i] = s / L[i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {
        REAL s = y[i];
        for (int k = i + 1; k < m; ++k) s -= L[k * m + i] * x[k];
        x[i] = s / L[i * m + i];
    }
}

void FN(or_ls_fixed)(int n, int K, int N, int T, const REAL* D, const REAL* X, const int32_t* idx,
                     const int32_t* cnt, REAL* val) {
    (void)K;
#pragma omp parallel
    {
        REAL* A = (REAL*)malloc(sizeof(REAL) * (size_t)T * T);
        REAL* L = (REAL*)calloc((size_t)T * T, sizeof(REAL));
        REAL* b = (REAL*)malloc(sizeof(REAL) * (size_t)T);
        REAL* y = (REAL*)malloc(sizeof(REAL) * (size_t)T);
        REAL* xs = (REAL*)malloc(sizeof(REAL) * (size_t)T);
#pragma omp for schedule(static)
        for (long s = 0; s < N; ++s) {
            int m = cnt[s];
            const REAL* x = X + (size_t)s * n;
            for (int i = 0; i < m; ++i) {
                const REAL* di = D + (size_t)idx[s * T + i] * n;
                REAL bi = 0;
                for (int t = 0; t < n; ++t) bi += di[t] * x[t];
                b[i] = bi;
                for (int j = 0; j < m; ++j) {
                    const REAL* dj = D + (size_t)idx[s * T + j] * n;
                    REAL g = 0;
                    for (int t = 0; t < n; ++t) g += di[t] * dj[t];
                    A[i * m + j] = g;
                }
            }
            FN(chol_factor)(m, A, L);
            FN(chol_solve)(m, L, b, xs, y);
            for (int j = 0; j < T; ++j) val[s * T + j] = j < m ? xs[j] : 0;
        }
        free(A); free(L); free(b); free(y); free(xs);
    }
}

/* gD (n x K) and gX (n x N) are overwritten. */
void FN(or_ls_adjoint)(int n, int K, int N, int T, const REAL* D, const REAL* X, const int32_t* idx,
                       const int32_t* cnt, const REAL* val, const REAL* gval, REAL* gD, REAL* gX) {
    memset(gD, 0, sizeof(REAL) * (size_t)n * K);
    REAL* A = (REAL*)malloc(sizeof(REAL) * (size_t)T * T);
    REAL* L = (REAL*)calloc((size_t)T * T, sizeof(REAL));
    REAL* y = (REAL*)malloc(sizeof(REAL) * (size_t)T);
    REAL* z = (REAL*)malloc(sizeof(REAL) * (size_t)T);
    REAL* r = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* dz = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    for (long s = 0; s < N; ++s) {
        int m = cnt[s];
        const REAL* x = X + (size_t)s * n;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const REAL* di = D + (size_t)idx[s * T + i] * n;
                const REAL* dj = D + (size_t)idx[s * T + j] * n;
                REAL g = 0;
                for (int t = 0; t < n; ++t) g += di[t] * dj[t];
                A[i * m + j] = g;
            }
        for (int t = 0; t < n; ++t) { r[t] = x[t]; dz[t] = 0; }
        if (m > 0) {
            FN(chol_factor)(m, A, L);
            FN(chol_solve)(m, L, gval + (size_t)s * T, z, y);
            for (int j = 0; j < m; ++j) {
                const REAL* dj = D + (size_t)idx[s * T + j] * n;
                for (int t = 0; t < n; ++t) {
                    r[t] -= dj[t] * val[s * T + j];
                    dz[t] += dj[t] * z[j];
                }
            }
        }
        for (int t = 0; t < n; ++t) gX[(size_t)s * n + t] = dz[t];
        for (int j = 0; j < m; ++j) {
            REAL* g = gD + (size_t)idx[s * T + j] * n;
            for (int t = 0; t < n; ++t) g[t] += r[t] * z[j] - dz[t] * val[s * T + j];
        }
    }
    free(A); free(L); free(y); free(z); free(r); free(dz);
}
