// omp_qform_t64.hip -- register-basis Batch-OMP instantiations for bases of
// up to 64 atoms (omp_qform.hpp).
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t64(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1) return launch_q<1, 64, 1>(a, st);
    return launch_qw<2, 64>(a, sh.waves, st);
}

}  // namespace ksvd
