// omp_qform_t64.hip -- register-basis Batch-OMP instantiations for bases of
// up to 64 atoms (omp_qform.hpp): 1 atom per lane at 1 wave, 2 atoms per lane at 1, 2, 4 or 8 waves.
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t64(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1 && sh.waves == 1) return launch_q<1, 64, 1>(a, st);
    if (sh.kpl == 2) return launch_qw<2, 64>(a, sh.waves, st);
    return hipErrorInvalidValue;
}

}  // namespace ksvd
