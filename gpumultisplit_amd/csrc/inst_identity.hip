// Kernel instantiations for the IdentF bucket identifier (all bucket counts,
// key-only and key-value).
#include "gms_launch.hpp"

namespace gms {
GMS_DEFINE_LAUNCHERS(IdentF)
}  // namespace gms
