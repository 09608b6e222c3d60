// Kernel instantiations for the DigitF bucket identifier (all bucket counts,
// key-only and key-value).
#include "gms_launch.hpp"

namespace gms {
GMS_DEFINE_LAUNCHERS(DigitF)
}  // namespace gms
