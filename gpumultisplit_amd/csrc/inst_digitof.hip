// Kernel instantiations for DigitOfF: digits of a wide bucket id (the LSD
// passes of multisplits into more than 256 buckets).
#include "gms_launch.hpp"

namespace gms {
GMS_DEFINE_LAUNCHERS(DigitOfF)
}  // namespace gms
