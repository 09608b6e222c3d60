// Batched prescan/postscan kernels with 32 keys per thread.
#include "gms_batch_kernels.hpp"

namespace gms {
GMS_DEFINE_BATCH_LAUNCHERS(32)
}  // namespace gms
