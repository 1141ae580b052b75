// Batched persistent kernels of the other 2-D 'indep' classes 5-7.
#include "k_batch_kernels.h"

namespace sdemc {

template SDEMC_LAUNCH_GROUP_SIG(5, 1)
template SDEMC_LAUNCH_GROUP_SIG(6, 1)
template SDEMC_LAUNCH_GROUP_SIG(7, 1)

}  // namespace sdemc
