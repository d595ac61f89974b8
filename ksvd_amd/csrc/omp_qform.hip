// omp_qform.hip -- shape choice and dispatch of the register-basis Batch-OMP
// kernel (omp_qform.hpp); one translation unit per basis size TT
// (omp_qform_t{8,16,32,48,64}.hip) holds its instantiations.  The default
// shape is a fixed rule (qform_shape); the autotuner picks among
// qform_candidates by measurement.  Every shape computes each atom's values
// with the same operations in the same order, so the results do not depend
// on the shape.
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

// The (atoms per lane, waves per signal) pairs compiled for each basis size
// (omp_qform_t*.hip); a shape also needs 64 * kpl * waves >= K.
static bool qform_compiled(int tt, int kpl, int waves) {
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return false;
    switch (tt) {
        case 8:
        case 16: return kpl == 1 || kpl == 2 || kpl == 4 || kpl == 8;
        case 32: return kpl == 1 || kpl == 2 || kpl == 4;
        case 48: return kpl == 1 || kpl == 2;
        case 64: return kpl == 2 || (kpl == 1 && waves == 1);
    }
    return false;
}

QformShape qform_shape_with(int K, int T, int kpl, int waves) {
    QformShape sh = qform_shape(K, T);
    if (!sh.kpl || !qform_compiled(sh.tt, kpl, waves) || 64L * kpl * waves < K) return QformShape{0, 0, 0};
    sh.kpl = kpl;
    sh.waves = waves;
    return sh;
}

int qform_candidates(int K, int T, QformShape* out, int cap) {
    const QformShape d = qform_shape(K, T);
    if (!d.kpl) return 0;
    int n = 0;
    for (int kpl = 1; kpl <= 8; kpl *= 2) {
        // the fewest waves that cover K with this many atoms per lane
        int W = 1;
        while (64L * kpl * W < K && W < 8) W *= 2;
        const QformShape sh = qform_shape_with(K, T, kpl, W);
        if (sh.kpl && n < cap) out[n++] = sh;
    }
    return n;
}

hipError_t launch_batch_omp_qform(const OmpArgs& a, hipStream_t st) {
    if (a.N <= 0) return hipSuccess;
    const QformShape sh = a.q_kpl ? qform_shape_with(a.K, a.T, a.q_kpl, a.q_waves) : qform_shape(a.K, a.T);
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
