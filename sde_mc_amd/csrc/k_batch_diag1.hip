// Batched persistent kernels of the 1-D 'diag' classes 0-3.
#include "k_batch_kernels.h"

namespace sdemc {

template SDEMC_LAUNCH_GROUP_SIG(0, 1)
template SDEMC_LAUNCH_GROUP_SIG(0, 2)
template SDEMC_LAUNCH_GROUP_SIG(1, 1)
template SDEMC_LAUNCH_GROUP_SIG(2, 1)
template SDEMC_LAUNCH_GROUP_SIG(3, 1)

}  // namespace sdemc
