// omp_qform_t16.hip -- register-basis Batch-OMP instantiations for bases of
// up to 16 atoms (omp_qform.hpp).
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t16(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1) return launch_q<1, 16, 1>(a, st);
    if (sh.kpl == 2) return launch_q<2, 16, 1>(a, st);
    if (sh.kpl == 4) return launch_q<4, 16, 1>(a, st);
    return launch_qw<8, 16>(a, sh.waves, st);
}

}  // namespace ksvd
