// omp_qform_t16.hip -- register-basis Batch-OMP instantiations for bases of
// up to 16 atoms (omp_qform.hpp): 1, 2, 4 or 8 atoms per lane, 1, 2, 4 or 8 waves per signal.
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t16(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1) return launch_qw<1, 16>(a, sh.waves, st);
    if (sh.kpl == 2) return launch_qw<2, 16>(a, sh.waves, st);
    if (sh.kpl == 4) return launch_qw<4, 16>(a, sh.waves, st);
    if (sh.kpl == 8) return launch_qw<8, 16>(a, sh.waves, st);
    return hipErrorInvalidValue;
}

}  // namespace ksvd
