// 2-D hydrodynamics stage kernels (ndir = 2).
#include "fv_launch.h"
namespace FV_NS {
hipError_t launch_sweep_2d_hd(const KernelKey& k, const SweepArgs& a, const FVParams& fp,
                                       int grid, hipStream_t s) {
  if (k.ndir == 2) return launch_prec<Phys<2, 2, false, DIVB_NONE>>(k, a, fp, grid, s);
  return hipErrorInvalidValue;
}
}  // namespace FV_NS
