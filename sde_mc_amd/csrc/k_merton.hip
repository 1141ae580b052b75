// Plain-path kernels of the merton model family (see sdemc_kernels.h).
#include "sdemc_kernels.h"

namespace sdemc {

int launch_rng_merton(const sdemc_job& job, const LaunchArgs& a) {
  return launch_rng_fam<FAM_MERTON>(job, a);
}

int launch_fixed_merton(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  return launch_fixed_fam<FAM_MERTON>(job, in, a);
}

}  // namespace sdemc
