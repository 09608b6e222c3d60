// Batched prescan/postscan kernels with 8 keys per thread.
#include "gms_batch_kernels.hpp"

namespace gms {
GMS_DEFINE_BATCH_LAUNCHERS(8)
}  // namespace gms
