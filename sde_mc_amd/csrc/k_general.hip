// Plain-path kernels of the 'general' noise family (see sdemc_kernels.h): a
// (dim, brown_dim) diffusion matrix applied to the correlated Brownian
// increments, uniform grid.
#include "sdemc_kernels.h"

namespace sdemc {

int launch_rng_general(const sdemc_job& job, const LaunchArgs& a) {
  return launch_rng_fam<FAM_GENERAL>(job, a);
}

int launch_fixed_general(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  return launch_fixed_fam<FAM_GENERAL>(job, in, a);
}

}  // namespace sdemc
