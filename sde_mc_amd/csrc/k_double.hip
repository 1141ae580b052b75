// Plain-path kernels of the double model family (see sdemc_kernels.h).
#include "sdemc_kernels.h"

namespace sdemc {

int launch_rng_double(const sdemc_job& job, const LaunchArgs& a) {
  return launch_rng_fam<FAM_DOUBLE>(job, a);
}

int launch_fixed_double(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  return launch_fixed_fam<FAM_DOUBLE>(job, in, a);
}

}  // namespace sdemc
