// sf_sweep_weno3.hip -- the grid-path sweep kernels of the WENO3 reconstruction
// (scheme id 2) for all six face fluxes.
#include "sf_sweep.h"

namespace sfamd {
template int launch_sweep<2>(const SweepCall&, int, int, int);
}  // namespace sfamd
