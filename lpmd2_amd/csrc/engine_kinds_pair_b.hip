// Force-kernel launchers of the potential kinds kFastLJ kMorse (engine_launch.hpp),
// compiled on their own so that the kinds build in parallel.
#include "engine_launch.hpp"

LPMDX_INSTANTIATE_KIND(kFastLJ)
LPMDX_INSTANTIATE_KIND(kMorse)
