// Kernel instantiations for the RangeF bucket identifier (all bucket counts,
// key-only and key-value).
#include "gms_launch.hpp"

namespace gms {
GMS_DEFINE_LAUNCHERS(RangeF)
}  // namespace gms
