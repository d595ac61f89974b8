/* ksvd_amd.hpp -- header-only C++ facade over the C-ABI of ksvd_amd.h with the
 * reference's class names and bool returns, for a maintainer who keeps the
 * reference's main-style host program:
 *
 *   class CudaAlgorithm                          S/CudaAlgorithm.cuh (common base)
 *   class OMP : virtual CudaAlgorithm            S/OMP.cuh:12-25       BatchOMP()
 *   class DictionaryUpdate : virtual ...         S/DictionaryUpdate.cuh:11-25
 *                                                                      updateDictionary(p)
 *   class KSVD : OMP, DictionaryUpdate           S/KSVD.cuh:11-29      ExecuteIterations(n),
 *                                                                      isValid()
 *   Algorithms::Multiplication::AT_times_B       S/Algorithms.cuh:99
 *
 * The reference's classes work on a DeviceMemory arena they own
 * (S/DeviceMemory.cuh:36-64); here the caller owns the device buffers, in the
 * same column-major layouts, and names them in a DeviceMemory record.  The
 * phases the reference keeps private (normcols, the Gram, initialize,
 * errorComputation, clearDict) are public, so a caller can also drive the
 * loop phase by phase.  Failures return false and print the library's
 * message on stderr, as the reference reports on std::cerr; lastError()
 * returns it.  No HIP header is needed: streams are passed as void*.
 */
#ifndef KSVD_AMD_HPP
#define KSVD_AMD_HPP

#include <cstdint>
#include <cstdio>
#include <vector>

#include "ksvd_amd.h"

namespace ksvd_amd {

/* The device buffers of the path (caller-owned; NULL where a phase is not
 * used).  Names follow the reference's arena entries. */
struct DeviceMemory {
    const float* X = nullptr;         /* n x N signals                      X_ARRAY             */
    const float* D_in = nullptr;      /* n x K dictionary as read           D_ARRAY             */
    float* D = nullptr;               /* n x K working dictionary           TEMP_ROWSD_BY_COLSD */
    float* G = nullptr;               /* K x K Gram D^T D                   G                   */
    float* DtX = nullptr;             /* K x N correlations D^T X           DtX                 */
    float* selected_atoms = nullptr;  /* K x N slot map (1-based into c)    SELECTED_ATOMS      */
    float* c = nullptr;               /* T x N coefficients                 c                   */
    int32_t* idx = nullptr;           /* T x N compact Gamma: atoms (-1 past the count)         */
    float* val = nullptr;             /* T x N compact Gamma: coefficients                      */
    int32_t* cnt = nullptr;           /* N     compact Gamma: support sizes                     */
    float* err = nullptr;             /* N     per-signal squared errors                        */
};

class CudaAlgorithm {
  public:
    virtual ~CudaAlgorithm() {
        if (h_) ksvd_destroy(h_);
    }
    CudaAlgorithm(const CudaAlgorithm&) = delete;
    CudaAlgorithm& operator=(const CudaAlgorithm&) = delete;

    bool isValid() const { return valid_; }
    ksvd_handle* handle() const { return h_; }
    const ksvd_config& config() const { return cfg_; }
    const DeviceMemory& memory() const { return mem_; }
    const char* lastError() const { return h_ ? ksvd_last_error(h_) : "ksvd_create failed"; }

  protected:
    CudaAlgorithm() = default;
    /* Called once by the most derived class (the base is virtual). */
    bool init(const ksvd_config& cfg, const DeviceMemory& mem, void* stream) {
        cfg_ = cfg;
        mem_ = mem;
        if (ksvd_create(&h_, &cfg) != KSVD_OK) {
            report();
            return valid_ = false;
        }
        valid_ = true;
        return check(ksvd_set_stream(h_, stream));
    }
    bool check(int rc) {
        if (rc == KSVD_OK) return true;
        valid_ = false;
        report();
        return false;
    }
    bool need(bool present, const char* what) {
        if (present) return true;
        std::fprintf(stderr, "ksvd_amd: %s not set in DeviceMemory\n", what);
        return valid_ = false;
    }
    void report() const { std::fprintf(stderr, "%s\n", lastError()); }

    ksvd_handle* h_ = nullptr;
    ksvd_config cfg_{};
    DeviceMemory mem_{};
    bool valid_ = false;
};

class OMP : public virtual CudaAlgorithm {
  public:
    /* OMP::BatchOMP (S/OMP.cu:42-59): DtX = D^T X, then Batch-OMP with G.
     * With selected_atoms and c set: the reference's slot map and c (DtX set
     * too); otherwise the compact idx / val / cnt. */
    bool BatchOMP() {
        const DeviceMemory& m = mem_;
        if (!need(m.X && m.D && m.G, "X, D and G")) return false;
        if (m.selected_atoms && m.c) {
            if (!need(m.DtX != nullptr, "DtX")) return false;
            return check(ksvd_correlate(h_, m.D, m.X, m.DtX)) &&
                   check(ksvd_batch_omp(h_, m.DtX, m.G, m.selected_atoms, m.c));
        }
        if (!need(m.idx && m.val && m.cnt, "idx, val and cnt")) return false;
        return check(ksvd_batch_omp_from_signals(h_, m.D, m.X, m.G, m.idx, m.val, m.cnt));
    }

  protected:
    OMP() = default;
};

class DictionaryUpdate : public virtual CudaAlgorithm {
  public:
    /* DictionaryUpdate::updateDictionary(p) (S/DictionaryUpdate.cu:45-203):
     * one sweep in the atom order p (host array of K indices), after
     * initialize().  The reference layout (slot map on entry, dense Gamma
     * values on exit) when selected_atoms and c are set, else the compact
     * Gamma (val updated in place). */
    bool updateDictionary(const unsigned int* p) {
        const DeviceMemory& m = mem_;
        if (!need(m.X && m.D && p, "X, D and the permutation")) return false;
        const uint32_t* perm = reinterpret_cast<const uint32_t*>(p);
        if (m.selected_atoms && m.c) return check(ksvd_dictionary_update_dense(h_, m.X, m.D, m.selected_atoms, m.c, perm));
        if (!need(m.idx && m.val && m.cnt, "idx, val and cnt")) return false;
        return check(ksvd_dictionary_update(h_, m.X, m.D, m.idx, m.val, m.cnt, perm));
    }
    /* initialize / init (S/KSVD.cu:368-375, 442-457): the unused-signal list,
     * bitmap and replaced flags, before each sweep. */
    bool initialize() { return check(ksvd_du_init(h_)); }

  protected:
    DictionaryUpdate() = default;
};

class KSVD : public OMP, public DictionaryUpdate {
  public:
    /* KSVD(DeviceMemory*, HostMemory*) (S/KSVD.cuh:18): cfg carries the sizes
     * and the parameters the reference fixes at compile time. */
    KSVD(const ksvd_config& cfg, const DeviceMemory& mem, void* stream = nullptr) { init(cfg, mem, stream); }

    /* KSVD::ExecuteIterations (S/KSVD.cu:55-131): D = normcols(D_in), then
     * iternum iterations of Gram, Batch-OMP, initialize, dictionary update,
     * error and clearDict; the final compact Gamma in idx / val / cnt and the
     * per-iteration RMSE in rmse(). */
    bool ExecuteIterations(unsigned int iternum) {
        const DeviceMemory& m = mem_;
        if (!need(m.X && m.D_in && m.D && m.idx && m.val && m.cnt, "X, D_in, D, idx, val and cnt")) return false;
        rmse_.assign(iternum, 0.0);
        return check(ksvd_execute_iterations(h_, m.X, m.D_in, m.D, m.idx, m.val, m.cnt, (int)iternum,
                                             rmse_.empty() ? nullptr : rmse_.data()));
    }
    const std::vector<double>& rmse() const { return rmse_; }

    /* The phases (private in the reference). */
    bool normcols() { return need(mem_.D_in && mem_.D, "D_in and D") && check(ksvd_normcols(h_, mem_.D_in, mem_.D)); }
    bool gram() { return need(mem_.D && mem_.G, "D and G") && check(ksvd_gram(h_, mem_.D, mem_.G)); }
    /* KSVD::errorComputation (S/KSVD.cu:235-262) on the compact Gamma. */
    bool errorComputation(double* rmse) {
        const DeviceMemory& m = mem_;
        return need(m.X && m.D && m.idx && m.val && m.cnt && m.err, "X, D, idx, val, cnt and err") &&
               check(ksvd_error(h_, m.X, m.D, m.idx, m.val, m.cnt, m.err, rmse));
    }
    /* KSVD::clearDict (S/KSVD.cu:138-226), after errorComputation. */
    bool clearDict() {
        const DeviceMemory& m = mem_;
        return need(m.X && m.D && m.idx && m.val && m.cnt && m.err, "X, D, idx, val, cnt and err") &&
               check(ksvd_clear_dict(h_, m.X, m.D, m.idx, m.val, m.cnt, m.err));
    }

  private:
    std::vector<double> rmse_;
};

namespace Algorithms {
namespace Multiplication {
/* AT_times_B(handle, A, rowsA, colsA, B, rowsB, colsB, C) (S/Algorithms.cuh:99,
 * S/Multiplication.cu:40-63): C (colsA x colsB) = A^T B, column-major, in the
 * handle's GEMM form on the handle's stream. */
inline bool AT_times_B(ksvd_handle* h, const float* A, int rowsA, int colsA, const float* B, int rowsB, int colsB,
                       float* C) {
    if (rowsA != rowsB) {
        std::fprintf(stderr, "AT_times_B: rowsA = %d and rowsB = %d differ\n", rowsA, rowsB);
        return false;
    }
    const int rc = ksvd_at_times_b(h, A, rowsA, colsA, B, colsB, C);
    if (rc != KSVD_OK) std::fprintf(stderr, "%s\n", ksvd_last_error(h));
    return rc == KSVD_OK;
}
}  // namespace Multiplication
}  // namespace Algorithms

}  // namespace ksvd_amd

#endif /* KSVD_AMD_HPP */
