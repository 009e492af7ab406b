// sf_sweep_weno5.hip -- the grid-path sweep kernels of the WENO5-JS reconstruction
// (scheme id 0) for all six face fluxes.
#include "sf_sweep.h"

namespace sfamd {
template int launch_sweep<0>(const SweepCall&, int, int, int);
}  // namespace sfamd
