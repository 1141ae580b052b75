// Batched persistent kernels of the 2-D 'indep' uniform single-level classes (per-step counts, arrival clock).
#include "k_batch_kernels.h"

namespace sdemc {

template SDEMC_LAUNCH_GROUP_SIG(4, 1)
template SDEMC_LAUNCH_GROUP_SIG(4, 2)
template SDEMC_LAUNCH_GROUP_SIG(SDEMC_CLASS_INDEP2_CLOCK, 1)
template SDEMC_LAUNCH_GROUP_SIG(SDEMC_CLASS_INDEP2_CLOCK, 2)

}  // namespace sdemc
