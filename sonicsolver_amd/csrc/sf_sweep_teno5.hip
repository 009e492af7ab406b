// sf_sweep_teno5.hip -- the grid-path sweep kernels of the TENO5 reconstruction
// (scheme id 1) for all six face fluxes.
#include "sf_sweep.h"

namespace sfamd {
template int launch_sweep<1>(const SweepCall&, int, int, int);
}  // namespace sfamd
