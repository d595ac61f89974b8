/*
 * ksvd_amd.h -- C-ABI of the MI355X K-SVD engine (libksvd_amd.so).
 *
 * Drop-in boundary for the reference's Batch-OMP / K-SVD path.  Each entry
 * point names the reference interface it replaces (S/ = K-SVD/K-SVD/ in the
 * reference tree).  The reference exposes C++ classes over a global
 * DeviceMemory arena; here every buffer is caller-owned device memory passed
 * as a plain pointer, with the reference's layouts:
 *
 *   X  n x N  column-major float   (signal s = X + s*n)        X_ARRAY
 *   D  n x K  column-major float   (atom k   = D + k*n)        TEMP_ROWSD_BY_COLSD
 *   G  K x K  column-major float                               G
 *   DtX K x N column-major float   (alpha0 of signal s)        DtX
 *   SELECTED_ATOMS K x N float: slot+1 (1..T) or 0; after the
 *     dictionary update, the dense Gamma values                SELECTED_ATOMS
 *   c  T x N float: coefficients in selection order             c
 *   (S/DeviceMemory.cuh:36-64, S/DeviceMemory.cu:55-137)
 *
 * The engine's own fast layout is the compact Gamma: idx (T x N, int32, -1
 * padded), val (T x N, float, 0 padded), cnt (N, int32).
 *
 * Conventions: every call returns KSVD_OK (0) or a negative status;
 * ksvd_last_error() gives the text.  Calls are asynchronous on the handle's
 * stream unless stated; the handle owns its scratch memory.  No globals:
 * one handle per host thread.
 */
#ifndef KSVD_AMD_H
#define KSVD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ksvd_handle ksvd_handle;

enum {
    KSVD_OK = 0,
    KSVD_ERR_INVALID = -1, /* bad argument or shape */
    KSVD_ERR_LIMIT = -2,   /* size beyond what the kernels support */
    KSVD_ERR_HIP = -3,     /* HIP runtime error */
    KSVD_ERR_ALLOC = -4,   /* device allocation failed */
    KSVD_ERR_NULL = -5     /* required pointer is NULL */
};

/* Precision form of the two GEMMs (G = D^T D, DtX = D^T X). */
enum {
    KSVD_FORM_FP32 = 0,   /* FP32 operands, FP32 MFMA, FP32 accumulate      */
    KSVD_FORM_BF16 = 1,   /* BF16 operands on MFMA, FP32 accumulate         */
    KSVD_FORM_BF16X3 = 2, /* split-BF16 x3: hi*hi + hi*lo + lo*hi, FP32 acc  */
                          /* (operands carried to 16 significant bits)      */
    KSVD_FORM_BF16X6 = 3  /* three-part split x = hi + mid + lo (24 bits),  */
                          /* six BF16 products down to 2^-16, FP32 acc      */
};

/* Batch-OMP kernel variant (values 2 and 3 are not used). */
enum {
    KSVD_OMP_AUTO = 0,   /* KSVD_OMP_QFORM, except T > 32 with K > 1024 (its basis */
                         /* would not fit 8 waves' VGPRs): KSVD_OMP_WAVE          */
    KSVD_OMP_WAVE = 1,   /* one wave64 per signal: alpha = alpha0 - G(:,I) gamma   */
                         /* from FP32 columns of G in VGPRs, L in LDS (the plain  */
                         /* FP32 form, and the fallback for the largest shapes)  */
    KSVD_OMP_QFORM = 4   /* one workgroup per signal holding Q = G(:,I) L^-T in    */
                         /* VGPRs: alpha -= q_i y_i, one G column per step, no     */
                         /* per-step triangular solves; FP32 throughout            */
};

/*
 * Runtime configuration.  Replaces the reference's compile-time parameters
 * (S/ProgramParameters.cuh:17-23, S/GlobalDeclarations.cuh:27-63) and the
 * dimensions read from MAT files (S/Globals.cpp:10-37).
 */
typedef struct ksvd_config {
    int32_t n;            /* signal length  = rowsX = rowsD                     */
    int32_t K;            /* atoms          = colsD                             */
    int32_t N;            /* signals        = colsX                             */
    int32_t T;            /* sparsity Tdata (T_max in error mode), <= 64        */
    int32_t iterations;   /* NUMBERofITERATIONS                                 */
    int32_t use_thresh;   /* USE_THRESH (4)                                     */
    int32_t max_signals;  /* MAX_SIGNALS (5000)                                 */
    int32_t gemm_form;    /* KSVD_FORM_*                                        */
    int32_t omp_variant;  /* KSVD_OMP_*                                         */
    int32_t use_graph;    /* ksvd_execute_iterations: replay iterations 2.. as */
                          /* one captured hipGraph (1, default) or launch them  */
                          /* kernel by kernel (0)                              */
    double mu_thresh;     /* muTHRESH (0.8)                                     */
    double pivot_fp64;    /* Cholesky pivot 1-||w||^2 below this is recomputed  */
                          /* in FP64 (0 disables the fallback)                  */
    uint64_t perm_seed;   /* seed of the atom-order permutation stream          */
} ksvd_config;

/*
 * Launch shapes.  The reference fixes its launches (S/OMP.cu:109-111: one
 * thread per signal in blocks of 512; S/FileManager.cu:53 TPB_rowsX); here
 * they are per-handle settings, chosen by measurement with ksvd_autotune or
 * set by hand.  All zero selects the built-in rules.  The Batch-OMP shape and
 * the GEMM build change no result (each value is computed with the same
 * operations in the same order); du_blocks changes the order in which the
 * dictionary update sums its block partials (results are repeatable for a
 * given setting).
 */
typedef struct ksvd_tuning {
    int32_t omp_kpl;        /* register-basis Batch-OMP: atoms per lane (1, 2, 4, 8; 0: rule) */
    int32_t omp_waves;      /* register-basis Batch-OMP: waves per signal (1, 2, 4, 8; 0: rule)   */
    int32_t omp_eps_launch; /* error mode with KSVD_OMP_AUTO: 0 the register-basis form (one     */
                            /* workgroup per signal, the hardware dispatcher starting the next */
                            /* as each finishes), 1 the G-streaming wave form's persistent     */
                            /* launch with work stealing over signals                          */
    int32_t gemm_wpe;       /* GEMM kernel built for 3 (0 or 3) or 2 workgroups per CU           */
    int32_t du_blocks;      /* workgroups of the per-atom dictionary-update kernels, 64..1024   */
                            /* (0: 512)                                                       */
    int32_t source;         /* 0 defaults, 1 ksvd_set_tuning, 2 measured by ksvd_autotune,      */
                            /* 3 ksvd_autotune cache hit                                       */
    double omp_ms, gemm_ms, du_ms;                 /* autotune: the chosen candidates' times */
    double default_omp_ms, default_gemm_ms, default_du_ms; /* and the built-in rules' times   */
} ksvd_tuning;

/* Fills the reference defaults (T=6, 30 iterations, mu=0.8, use_thresh=4,
 * max_signals=5000), FP32 form, automatic Batch-OMP variant, pivot_fp64 =
 * 1e-3, perm_seed = 0, use_graph = 1. */
void ksvd_config_default(ksvd_config* cfg);

/* Handle lifetime.  Replaces the KSVD/CudaAlgorithm constructors
 * (S/KSVD.cu:42-44, S/CudaAlgorithm.cu:23-33) and DeviceMemory set-up. */
int ksvd_create(ksvd_handle** h, const ksvd_config* cfg);
int ksvd_destroy(ksvd_handle* h);
/* stream is a hipStream_t (NULL = default stream). */
int ksvd_set_stream(ksvd_handle* h, void* stream);
const char* ksvd_last_error(const ksvd_handle* h);
const char* ksvd_version(void);
/* Blocks until the handle's stream is idle. */
int ksvd_sync(ksvd_handle* h);

/* Current launch shapes of the handle. */
int ksvd_get_tuning(const ksvd_handle* h, ksvd_tuning* out);

/* Sets the launch shapes (time fields ignored); KSVD_ERR_INVALID for a shape
 * that is not compiled or does not cover K. */
int ksvd_set_tuning(ksvd_handle* h, const ksvd_tuning* t);

/* Project autotuner: times the candidate launch shapes on the caller's data
 * -- the GEMM builds on the correlation of up to 131072 signals, the
 * register-basis Batch-OMP shapes (and, with eps2, both error-mode launches)
 * on up to 65536 signals, the dictionary-update grids on a full sweep over
 * all N signals -- and keeps the fastest of each in the handle.  D is
 * normalised first (into handle scratch); eps2 may be NULL (sparsity mode).
 * Results are cached for the process by (n, K, N, T, GEMM form, mode, device)
 * and, when the environment variable KSVD_TUNE_CACHE names a file, in that
 * file too; force != 0 measures again.  Synchronises.  Replaces the
 * reference's fixed launch shapes (S/OMP.cu:109-111, S/FileManager.cu:53). */
int ksvd_autotune(ksvd_handle* h, const float* D, const float* X, const float* eps2, int force);

/* Text of the last ksvd_autotune's measurements: every candidate and its time
 * ("cache hit" when the choice came from the cache). */
const char* ksvd_tuning_log(const ksvd_handle* h);

/* ---- primitives ---------------------------------------------------------- */

/* D_out = D_in * diag(1/||d_k||).  KSVD::normcols  S/KSVD.cu:267-286. */
int ksvd_normcols(ksvd_handle* h, const float* D_in, float* D_out);

/* C (colsA x colsB) = A^T B, column-major, lda = ldb = rows, in the handle's
 * GEMM form.  Algorithms::Multiplication::AT_times_B  S/Multiplication.cu:40-63. */
int ksvd_at_times_b(ksvd_handle* h, const float* A, int rows, int colsA, const float* B, int colsB,
                    float* C);

/* G = D^T D (K x K).  AT_times_B call at S/KSVD.cu:64-73. */
int ksvd_gram(ksvd_handle* h, const float* D, float* G);

/* DtX = D^T X (K x N).  AT_times_B call at S/OMP.cu:44-53. */
int ksvd_correlate(ksvd_handle* h, const float* D, const float* X, float* DtX);

/* ---- Batch-OMP ----------------------------------------------------------- */

/* Reference-layout Batch-OMP: reads DtX and G, writes SELECTED_ATOMS (slot map)
 * and c.  Slots of c beyond a signal's accepted count are left untouched, as in
 * the reference.  parallelOMP / parOMP  S/OMP.cu:100-113, S/OMP.cu:254-405. */
int ksvd_batch_omp(ksvd_handle* h, const float* DtX, const float* G, float* selected_atoms,
                   float* c);

/* Same algorithm, compact output. */
int ksvd_batch_omp_compact(ksvd_handle* h, const float* DtX, const float* G, int32_t* idx,
                           float* val, int32_t* cnt);

/* OMP::BatchOMP  S/OMP.cu:42-59: DtX = D^T X in the handle's GEMM form (into
 * handle scratch) followed by Batch-OMP; compact output. */
int ksvd_batch_omp_from_signals(ksvd_handle* h, const float* D, const float* X, const float* G,
                                int32_t* idx, float* val, int32_t* cnt);

/* OMP::BatchOMP with the reference output layout: SELECTED_ATOMS slot map and
 * c (slots beyond a signal's count untouched).  Uses the configured variant. */
int ksvd_batch_omp_from_signals_dense(ksvd_handle* h, const float* D, const float* X,
                                      const float* G, float* selected_atoms, float* c);

/* Error-constrained Batch-OMP (no reference counterpart beyond the comment at
 * S/KSVD.cu:79): per signal, stop when ||r||^2 = ||x||^2 - alpha0_I^T gamma_I
 * (evaluated in FP64) <= eps2[s], or at T atoms (T_max).  xnorm2[s] =
 * ||x_s||^2 in FP64.  The register-basis form runs one workgroup per signal:
 * the uneven support sizes are balanced by the hardware dispatcher, which
 * starts the next queued workgroup on a CU as soon as one finishes.  The
 * G-streaming wave form runs one persistent launch (as many workgroups as the
 * GPU holds) whose waves take the next signal from a device counter until all
 * are done.  With KSVD_OMP_AUTO the tuning (omp_eps_launch) picks between
 * them; ksvd_autotune measures both. */
int ksvd_batch_omp_eps(ksvd_handle* h, const float* DtX, const float* G, const double* xnorm2,
                       const float* eps2, int32_t* idx, float* val, int32_t* cnt);

/* Error-constrained Batch-OMP from D and X with the configured variant:
 * stop when ||r||^2 <= eps2[s] (FP64 sum, ||x_s||^2 formed in FP64 from X) or
 * at T atoms; launches as ksvd_batch_omp_eps. */
int ksvd_batch_omp_eps_from_signals(ksvd_handle* h, const float* D, const float* X,
                                    const float* G, const float* eps2, int32_t* idx, float* val,
                                    int32_t* cnt);

/* Allocation-free Batch-OMP from signals.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * results, bit for bit, as ksvd_batch_omp_from_signals on the first N signals
 * of X (0 <= N <= the handle's N; compact output for those N signals).  Instead
 * of handle scratch it forms DtX = D^T X chunk by chunk in one library-owned
 * static device buffer of 32 MiB per GPU (held until the process ends):
 * floor(8M / K) signals per chunk, KSVD_ERR_LIMIT if one signal's K
 * correlations do not fit.  The call allocates nothing and keeps no
 * statistics (ksvd_omp_stats then reports 0).  Every call on a GPU uses that
 * GPU's one buffer: calls that overlap on one GPU (other streams, other
 * handles, other threads) race and must be serialised by the caller.
 * Variants KSVD_OMP_AUTO, KSVD_OMP_QFORM and KSVD_OMP_WAVE. */
int ksvd_batch_omp_from_signals_noalloc(ksvd_handle* h, const float* D, const float* X, const float* G,
                                        int N, int32_t* idx, float* val, int32_t* cnt);

/* Diagnostics of the last Batch-OMP call (synchronises): *fallbacks = the
 * Cholesky pivots recomputed in FP64 (register-basis form); *candidates is
 * always 0 (kept for the layout of earlier versions).  NULL outputs allowed. */
int ksvd_omp_stats(ksvd_handle* h, int64_t* candidates, int64_t* fallbacks);

/* ---- dictionary update / error / clear ------------------------------------ */

/* Resets the unused-signal list, bitmap, replaced flags and counter.
 * initialize()/init  S/KSVD.cu:368-375, S/KSVD.cu:442-457. */
int ksvd_du_init(ksvd_handle* h);

/* One approximate K-SVD sweep over atoms in the order perm[0..K-1] (a HOST
 * pointer, as in the reference).  D and val are updated in place; idx/cnt are
 * read only.  DictionaryUpdate::updateDictionary  S/DictionaryUpdate.cu:45-203. */
int ksvd_dictionary_update(ksvd_handle* h, const float* X, float* D, const int32_t* idx,
                           float* val, const int32_t* cnt, const uint32_t* perm_host);

/* Reference-layout variant: selected_atoms holds the slot map on entry and the
 * dense Gamma values on exit (sprow_kernel S/DictionaryUpdate.cu:378-395); c is
 * read.  The pivot-break slot (marked but never written by parOMP) reads the
 * stale c entry, as the reference does. */
int ksvd_dictionary_update_dense(ksvd_handle* h, const float* X, float* D, float* selected_atoms,
                                 const float* c, const uint32_t* perm_host);

/* err[s] = ||x_s - D gamma_s||^2 masked by the unused-signal bitmap; *rmse_host
 * (if not NULL) receives sqrt(sum of unmasked err / (n N)) -- this call then
 * synchronises.  KSVD::errorComputation  S/KSVD.cu:235-262. */
int ksvd_error(ksvd_handle* h, const float* X, const float* D, const int32_t* idx,
               const float* val, const int32_t* cnt, float* err, double* rmse_host);

/* Replaces atoms that are too coherent with another atom (|d_i^T d_j| > mu) or
 * used by fewer than use_thresh signals, with the worst-represented signal.
 * KSVD::clearDict  S/KSVD.cu:138-226. */
int ksvd_clear_dict(ksvd_handle* h, const float* X, float* D, const int32_t* idx,
                    const float* val, const int32_t* cnt, float* err);

/* Full loop: normcols(D_in) -> D_out, then per iteration G, Batch-OMP, init,
 * permutation, dictionary update, error, clearDict.  rmse_host[it] receives
 * the per-iteration RMSE (host array of `iterations` doubles, may be NULL).
 * KSVD::ExecuteIterations  S/KSVD.cu:55-131. */
int ksvd_execute_iterations(ksvd_handle* h, const float* X, const float* D_in, float* D_out,
                            int32_t* idx, float* val, int32_t* cnt, int iterations,
                            double* rmse_host);

/* The same loop with error-constrained Batch-OMP (the comment at S/KSVD.cu:79:
 * sparsecode with a threshold instead of a sparsity): each signal stops when
 * ||x_s - D gamma_s||^2 <= eps2[s] (FP64 test) or at T atoms (T_max).
 * Iterations after the first are replayed from one captured hipGraph. */
int ksvd_execute_iterations_eps(ksvd_handle* h, const float* X, const float* D_in, float* D_out,
                                const float* eps2, int32_t* idx, float* val, int32_t* cnt,
                                int iterations, double* rmse_host);

/* Host-only (no device work): the atom-order permutations ksvd_execute_iterations
 * uses -- iterations x K values, row it = the running buffer after it+1
 * shuffles.  Replaces Algorithms::RandomPermutation::generateRandPerm
 * (S/RandPermutation.cu:35-42: std::random_shuffle of a running buffer) with a
 * seeded SplitMix64 Fisher-Yates shuffle of the same running buffer. */
int ksvd_perm_sequence(uint64_t seed, int K, int iterations, uint32_t* out);

/* ---- MAT-file I/O (host only, no device work) --------------------------- */

/* Reads the real 2-D numeric variable `name` of a level-5 MAT-file (any numeric
 * storage type, miCOMPRESSED or not) as column-major doubles.  With out == NULL
 * only *rows / *cols are returned; otherwise capacity must be >= rows*cols.
 * Replaces readMat (S/FileManager.cu:93-122: matOpen, matGetVariable, mxGetPr,
 * mxGetM/N of the MATLAB MAT API). */
int ksvd_mat_read(const char* path, const char* name, double* out, int64_t capacity, int64_t* rows,
                  int64_t* cols);

/* Writes a rows x cols column-major double matrix as variable `name`
 * (uncompressed level 5), creating the file (append == 0) or adding to it
 * (append != 0).  global != 0 sets the global flag as matPutVariableAsGlobal
 * does.  Replaces writeMat (S/FileManager.cu:126-152). */
int ksvd_mat_write(const char* path, const char* name, const double* data, int64_t rows, int64_t cols,
                   int append, int global);

/* Text of the last MAT-file error on the calling thread. */
const char* ksvd_mat_last_error(void);

/* ---- differentiation ----------------------------------------------------- */

/* Batch-OMP (compact output) that also stores each signal's Cholesky factor:
 * Lfac (N x T(T+1)/2 floats), L L^T = G_II of the accepted atoms in slot
 * order, lower triangle packed column by column with pitch T -- L(r, c) at
 * c*T - c(c-1)/2 + (r - c), zero for rows or columns at or beyond the
 * signal's count.  The reference builds this factor (Lchol,
 * S/OMP.cu:316-336) and drops it; here it feeds ksvd_omp_adjoint_factored.
 * Register-basis variant only (KSVD_ERR_INVALID otherwise). */
int ksvd_batch_omp_compact_factors(ksvd_handle* h, const float* DtX, const float* G, int32_t* idx,
                                   float* val, int32_t* cnt, float* Lfac);

/* Adjoint of the fixed-support least squares gamma_I = (D_I^T D_I)^{-1} D_I^T x
 * (no reference counterpart).  gval (T x N) = dL/dgamma in slot order.
 * Writes gX (n x N) = D_I z and gD (n x K) = sum_s (r z^T - D_I z gamma^T)
 * on the support columns (zero elsewhere), z = (D_I^T D_I)^{-1} g,
 * r = x - D_I gamma.  Low-memory mode: G_II and its factor are recomputed
 * per signal.  No atomics: the same inputs give the same bits. */
int ksvd_omp_adjoint(ksvd_handle* h, const float* D, const float* X, const int32_t* idx,
                     const float* val, const int32_t* cnt, const float* gval, float* gD,
                     float* gX);

/* The same adjoint reusing the factors ksvd_batch_omp_compact_factors stored
 * (Lfac as described there) instead of recomputing G_II. */
int ksvd_omp_adjoint_factored(ksvd_handle* h, const float* D, const float* X, const int32_t* idx,
                              const float* val, const int32_t* cnt, const float* Lfac,
                              const float* gval, float* gD, float* gX);

/* Adjoint of ksvd_normcols: with d_k = D_in_k / ||D_in_k||,
 * gD_in_k = (gD_out_k - d_k (d_k^T gD_out_k)) / ||D_in_k||.  The forward is
 * KSVD::normcols (S/KSVD.cu:267-286); the derivative has no counterpart. */
int ksvd_normcols_adjoint(ksvd_handle* h, const float* D_in, const float* gD_out, float* gD_in);

/* Compact Gamma -> dense Gamma values (K x N column-major, zeros off the
 * support): the layout sprow_kernel leaves in SELECTED_ATOMS
 * (S/DictionaryUpdate.cu:378-395). */
int ksvd_compact_to_dense(ksvd_handle* h, const int32_t* idx, const float* val, const int32_t* cnt,
                          float* gamma_dense);

/* The reverse gather: out (T x N) = dense(idx[s, j], s) for j < cnt[s], 0
 * beyond (a dense K x N gradient brought to slot order). */
int ksvd_compact_gather(ksvd_handle* h, const float* dense, const int32_t* idx, const int32_t* cnt,
                        float* out);

#ifdef __cplusplus
}
#endif
#endif /* KSVD_AMD_H */
