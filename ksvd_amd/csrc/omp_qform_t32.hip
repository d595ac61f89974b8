// omp_qform_t32.hip -- register-basis Batch-OMP instantiations for bases of
// up to 32 atoms (omp_qform.hpp).
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t32(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1) return launch_q<1, 32, 1>(a, st);
    if (sh.kpl == 2) return launch_q<2, 32, 1>(a, st);
    return launch_qw<4, 32>(a, sh.waves, st);
}

}  // namespace ksvd
