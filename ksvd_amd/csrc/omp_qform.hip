// omp_qform.hip -- shape choice and dispatch of the register-basis Batch-OMP
// kernel (omp_qform.hpp); one translation unit per basis size TT
// (omp_qform_t{8,16,32,64}.hip) holds its instantiations.
#include "omp_common.hpp"

namespace ksvd {

hipError_t launch_qform_t8(const OmpArgs& a, const QformShape& sh, hipStream_t st);
hipError_t launch_qform_t16(const OmpArgs& a, const QformShape& sh, hipStream_t st);
hipError_t launch_qform_t32(const OmpArgs& a, const QformShape& sh, hipStream_t st);
hipError_t launch_qform_t48(const OmpArgs& a, const QformShape& sh, hipStream_t st);
hipError_t launch_qform_t64(const OmpArgs& a, const QformShape& sh, hipStream_t st);

QformShape qform_shape(int K, int T) {
    QformShape sh{0, 0, 0};
    if (K < 1 || T < 1 || T > 64) return sh;
    const int TT = T <= 8 ? 8 : T <= 16 ? 16 : T <= 32 ? 32 : T <= 48 ? 48 : 64;
    const int kmax = (128 / TT) < 8 ? 128 / TT : 8;
    const int need = (K + 63) / 64;
    int kpl = 1;
    while (kpl < need && kpl < kmax) kpl *= 2;
    const int waves = (K + 64 * kpl - 1) / (64 * kpl);
    int W = 1;
    while (W < waves) W *= 2;
    if (W > 8) return sh;  // 16 waves would cap a wave at 128 VGPRs: Q alone needs them
    sh.kpl = kpl;
    sh.tt = TT;
    sh.waves = W;
    return sh;
}

hipError_t launch_batch_omp_qform(const OmpArgs& a, hipStream_t st) {
    if (a.N <= 0) return hipSuccess;
    const QformShape sh = qform_shape(a.K, a.T);
    if (!sh.kpl) return hipErrorInvalidValue;
    switch (sh.tt) {
        case 8: return launch_qform_t8(a, sh, st);
        case 16: return launch_qform_t16(a, sh, st);
        case 32: return launch_qform_t32(a, sh, st);
        case 48: return launch_qform_t48(a, sh, st);
        case 64: return launch_qform_t64(a, sh, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace ksvd
