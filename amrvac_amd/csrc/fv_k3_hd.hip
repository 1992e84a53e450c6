// 3-D hydrodynamics stage kernels (ndir = 3).
#include "fv_launch.h"
namespace FV_NS {
hipError_t launch_sweep_3d_hd(const KernelKey& k, const SweepArgs& a, const FVParams& fp,
                                       int grid, hipStream_t s) {
  if (k.ndir == 3) return launch_prec<Phys<3, 3, false, DIVB_NONE>>(k, a, fp, grid, s);
  return hipErrorInvalidValue;
}
}  // namespace FV_NS
