// Kernel instantiations for the DeltaF bucket identifier (all bucket counts,
// key-only and key-value).
#include "gms_launch.hpp"

namespace gms {
GMS_DEFINE_LAUNCHERS(DeltaF)
}  // namespace gms
