// omp_qform_t48.hip -- register-basis Batch-OMP instantiations for bases of
// up to 48 atoms (omp_qform.hpp): 1 or 2 atoms per lane, 1, 2, 4 or 8 waves per signal.
#include "omp_qform.hpp"

namespace ksvd {

hipError_t launch_qform_t48(const OmpArgs& a, const QformShape& sh, hipStream_t st) {
    using namespace qform_detail;
    if (sh.kpl == 1) return launch_qw<1, 48>(a, sh.waves, st);
    if (sh.kpl == 2) return launch_qw<2, 48>(a, sh.waves, st);
    return hipErrorInvalidValue;
}

}  // namespace ksvd
