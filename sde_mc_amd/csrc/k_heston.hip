// Plain-path kernels of the Heston family (see sdemc_kernels.h): state
// (S, v), dimension 2 only.
#include "sdemc_kernels.h"

namespace sdemc {

int launch_rng_heston(const sdemc_job& job, const LaunchArgs& a) {
  if (job.model.dim != 2) return SDEMC_ERR_UNSUPPORTED;
  return launch_rng_d<2, FAM_HESTON>(job, a);
}

int launch_fixed_heston(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  if (job.model.dim != 2) return SDEMC_ERR_UNSUPPORTED;
  return launch_fixed_d<2, FAM_HESTON>(job, in, a);
}

}  // namespace sdemc
