// Force-kernel launchers of the potential kinds kFinnisSinclair kGupta (engine_launch.hpp),
// compiled on their own so that the kinds build in parallel.
#include "engine_launch.hpp"

LPMDX_INSTANTIATE_KIND(kFinnisSinclair)
LPMDX_INSTANTIATE_KIND(kGupta)
