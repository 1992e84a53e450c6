// 3-D ideal MHD stage kernels (ndir = 3), GLM div B control or none in the
// stage (Linde's source is a kernel of its own).
#include "fv_launch.h"
namespace FV_NS {
hipError_t launch_sweep_3d_mhd(const KernelKey& k, const SweepArgs& a, const FVParams& fp,
                                        int grid, hipStream_t s) {
  if (k.ndir != 3) return hipErrorInvalidValue;
  if (k.divb == DIVB_GLM) return launch_prec<Phys<3, 3, true, DIVB_GLM>>(k, a, fp, grid, s);
  if (k.divb == DIVB_NONE) return launch_prec<Phys<3, 3, true, DIVB_NONE>>(k, a, fp, grid, s);
  return hipErrorInvalidValue;
}
}  // namespace FV_NS
