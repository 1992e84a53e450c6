// 2-D ideal MHD stage kernels with two vector components (ndir = 2, as
// setup.pl -d=2 builds the reference), GLM or no div B control in the
// stage (Linde's source is a kernel of its own).
#include "fv_launch.h"
namespace FV_NS {
hipError_t launch_sweep_2d_mhd_ndir2(const KernelKey& k, const SweepArgs& a, const FVParams& fp, int grid,
                                     hipStream_t s) {
  if (k.ndir != 2) return hipErrorInvalidValue;
  if (k.divb == DIVB_GLM) return launch_prec<Phys<2, 2, true, DIVB_GLM>>(k, a, fp, grid, s);
  if (k.divb == DIVB_NONE) return launch_prec<Phys<2, 2, true, DIVB_NONE>>(k, a, fp, grid, s);
  return hipErrorInvalidValue;
}
}  // namespace FV_NS
