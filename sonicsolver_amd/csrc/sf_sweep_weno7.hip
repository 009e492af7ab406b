// sf_sweep_weno7.hip -- the grid-path sweep kernels of the WENO7 reconstruction
// (scheme id 3) for all six face fluxes.
#include "sf_sweep.h"

namespace sfamd {
template int launch_sweep<3>(const SweepCall&, int, int, int);
}  // namespace sfamd
