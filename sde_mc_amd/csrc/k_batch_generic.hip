// Batched persistent kernels of the generic class (every supported job).
#include "k_batch_kernels.h"

namespace sdemc {

template SDEMC_LAUNCH_GROUP_SIG(SDEMC_CLASS_GENERIC, 1)

}  // namespace sdemc
