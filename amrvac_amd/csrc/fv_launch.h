// fv_launch.h -- kernel selection between the C-ABI layer and the
// per-physics translation units (each compiled twice: with FMA contraction,
// and with -ffp-contract=off for the bit-reproducible FP64 mode).
#pragma once

#include <hip/hip_runtime.h>

#include "fv_sweep.h"

namespace fvc {
struct KernelKey {
  int ndim, physics, ndir, divb;
  int mixed;   // 0: FP64 reconstruction/Riemann, 1: FP32
  int riem;    // RIEM_*
  int weno;    // 1: WENO5 (NG = 3), 0: TVD limiters (NG = 2)
  int limiter; // runtime limiter code (TVD path)
};

// Adjoint stage operands (device arrays in the interior-only block layout
// [nb][nw][bnz][16][16]):
//   lam_in = coef * lam_out + ca * add_a + cb * add_b + (d stage / d src)^T lam_out
// add_a / add_b may be null; gz is the 3-D z-face buffer, nw*nb*cells
// values of the face precision (double or float).
struct AdjArgs {
  const double* lam_out;
  double coef;
  const double* add_a;
  double ca;
  const double* add_b;
  double cb;
  double* lam_in;
  void* gz;
};

// Returns hipSuccess on launch, hipErrorInvalidValue when no kernel matches.
using SweepLauncher = hipError_t (*)(const KernelKey&, const SweepArgs&, const FVParams&, int grid,
                                     hipStream_t);
}  // namespace fvc

// launchers, defined once per namespace (fv: FMA contraction; fvx:
// -ffp-contract=off, FP64 only)
#define FV_DECLARE_LAUNCHERS(NS)                                                                     \
  namespace NS {                                                                                     \
  hipError_t launch_sweep_2d_hd(const fvc::KernelKey&, const fvc::SweepArgs&, const fvc::FVParams&, \
                                int, hipStream_t);                                                   \
  hipError_t launch_sweep_2d_mhd(const fvc::KernelKey&, const fvc::SweepArgs&, const fvc::FVParams&,\
                                 int, hipStream_t);                                                  \
  hipError_t launch_sweep_2d_mhd_ndir2(const fvc::KernelKey&, const fvc::SweepArgs&,                  \
                                       const fvc::FVParams&, int, hipStream_t);                      \
  hipError_t launch_sweep_3d_hd(const fvc::KernelKey&, const fvc::SweepArgs&, const fvc::FVParams&, \
                                int, hipStream_t);                                                   \
  hipError_t launch_sweep_3d_mhd(const fvc::KernelKey&, const fvc::SweepArgs&, const fvc::FVParams&,\
                                 int, hipStream_t);                                                  \
  hipError_t launch_ct_stage(int mixed, int riem, int lim, int ct, const fvc::SweepArgs& a,             \
                             const double* ws_src,                                                    \
                             const double* ws_base, const double* ws_dst, const double* ws_block,      \
                             int bx_mode, double* faces, double* ws_new, const fvc::FVParams& fp,      \
                             hipStream_t s);                                                         \
  size_t ct_face_buffer_doubles(int nblocks);                                                        \
  hipError_t launch_ct3_stage(int mixed, int riem, int lim, int ct, const fvc::SweepArgs& a,            \
                              const double* ws_src, const double* ws_base, const double* ws_dst,       \
                              const double* ws_block, int bx_mode, double* scratch, double* ws_new,    \
                              const fvc::FVParams& fp, hipStream_t s);                                 \
  size_t ct3_scratch_doubles(int nblocks);                                                           \
  }
FV_DECLARE_LAUNCHERS(fv)
FV_DECLARE_LAUNCHERS(fvx)

namespace fv {
// adjoint stage (hydrodynamics, or ideal MHD without divergence control or
// with GLM, 2-D / 3-D, FP64 or FP32 faces)
hipError_t launch_adjoint_stage(int ndim, int physics, int divb, int mixed, int riem, int lim,
                                const fvc::SweepArgs& a, const fvc::FVParams& fp, const fvc::AdjArgs& g,
                                hipStream_t s);
}

namespace FV_NS {
using fvc::KernelKey;


template <class P, class R, int RIEM, int LIM, int NG>
hipError_t launch_one(const SweepArgs& a, const FVParams& fp, int grid, hipStream_t s) {
  if (a.blevel)
    hipLaunchKernelGGL((sweep_kernel<P, R, RIEM, LIM, NG, true, 16>), dim3(grid), dim3(kLanes<16>), 0, s, a, fp);
  else
    hipLaunchKernelGGL((sweep_kernel<P, R, RIEM, LIM, NG, false, 16>), dim3(grid), dim3(kLanes<16>), 0, s, a,
                       fp);
  return hipGetLastError();
}

// 8^3 and 10^3 blocks (the block sizes of the reference's 3-D regression
// setups): FP64, the runtime-switched TVD limiters or WENO5, one kernel for
// uniform and AMR levels
template <class P, class R, int RIEM, int NG, int BXc>
hipError_t launch_small(const SweepArgs& a, const FVParams& fp, int grid, hipStream_t s) {
  hipLaunchKernelGGL((sweep_kernel<P, R, RIEM, NG == 3 ? LIM_WENO5 : 0, NG, true, BXc>), dim3(grid),
                     dim3(kLanes<BXc>), 0, s, a, fp);
  return hipGetLastError();
}

// TVD (runtime limiter) or WENO5 for one physics / precision / solver
template <class P, class R, int RIEM>
hipError_t launch_lim(const KernelKey& k, const SweepArgs& a, const FVParams& fp, int grid, hipStream_t s) {
  if (a.bx != 16) {
    if constexpr (!std::is_same<R, double>::value || P::NDIM != 3) {
      return hipErrorInvalidValue;
    } else if constexpr (RIEM == RIEM_HANCOCK) {
      if (k.weno) return hipErrorInvalidValue;
      if (a.bx == 8) return launch_small<P, R, RIEM, 2, 8>(a, fp, grid, s);
      if (a.bx == 10) return launch_small<P, R, RIEM, 2, 10>(a, fp, grid, s);
      return hipErrorInvalidValue;
    } else {
      if (a.bx == 8) return k.weno ? launch_small<P, R, RIEM, 3, 8>(a, fp, grid, s)
                                   : launch_small<P, R, RIEM, 2, 8>(a, fp, grid, s);
      if (a.bx == 10) return k.weno ? launch_small<P, R, RIEM, 3, 10>(a, fp, grid, s)
                                    : launch_small<P, R, RIEM, 2, 10>(a, fp, grid, s);
      return hipErrorInvalidValue;
    }
  }
  if (k.weno) {
    if constexpr (RIEM == RIEM_HANCOCK) return hipErrorInvalidValue;
    else return launch_one<P, R, RIEM, LIM_WENO5, 3>(a, fp, grid, s);
  }
  // compile-time limiter for the schemes of the benchmark configurations,
  // a runtime switch for the others and for runs with fix_small_values
  // (the compile-time-limiter kernels carry no small-value checks)
  if (fp.fix_small) return launch_one<P, R, RIEM, 0, 2>(a, fp, grid, s);
  switch (k.limiter) {
    case LIM_MINMOD: return launch_one<P, R, RIEM, LIM_MINMOD, 2>(a, fp, grid, s);
    case LIM_VANLEER: return launch_one<P, R, RIEM, LIM_VANLEER, 2>(a, fp, grid, s);
    case LIM_CADA3: return launch_one<P, R, RIEM, LIM_CADA3, 2>(a, fp, grid, s);
    default: return launch_one<P, R, RIEM, 0, 2>(a, fp, grid, s);
  }
}

template <class P, class R>
hipError_t launch_riem(const KernelKey& k, const SweepArgs& a, const FVParams& fp, int grid, hipStream_t s) {
  switch (k.riem) {
    case RIEM_HLL: return launch_lim<P, R, RIEM_HLL>(k, a, fp, grid, s);
    case RIEM_TVDLF: return launch_lim<P, R, RIEM_TVDLF>(k, a, fp, grid, s);
    case RIEM_HANCOCK: return launch_lim<P, R, RIEM_HANCOCK>(k, a, fp, grid, s);
    case RIEM_HLLC: return launch_lim<P, R, RIEM_HLLC>(k, a, fp, grid, s);
    case RIEM_HLLD:
      if constexpr (P::MHD && P::NDIR == 3) return launch_lim<P, R, RIEM_HLLD>(k, a, fp, grid, s);
      else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <class P>
hipError_t launch_prec(const KernelKey& k, const SweepArgs& a, const FVParams& fp, int grid, hipStream_t s) {
#ifdef FV_EXACT
  if (k.mixed) return hipErrorInvalidValue;
#else
  if (k.mixed) return launch_riem<P, float>(k, a, fp, grid, s);
#endif
  return launch_riem<P, double>(k, a, fp, grid, s);
}

}  // namespace FV_NS
