// fv_ct_common.h -- pieces shared by the 2-D and 3-D constrained-transport
// stage kernels (fv_ct.hip, fv_ct3.hip): the edge-field types, the argument
// block and the limited reconstruction of face quantities.
#pragma once

#include "fv_launch.h"

namespace FV_NS {
namespace {

constexpr int CT_AVERAGE = 1, CT_CONTACT = 2, CT_HLL = 3;

struct CtArgs {
  const double* ws_src;    // face field of the stage input
  const double* ws_base;   // of the RK base slot (comb_mode 1, 2)
  const double* ws_dst;    // of psb before the stage (comb_mode 0)
  const double* ws_block;  // of ps (HLLD normal field) when bx_mode < 0
  int bx_mode;             // < 0: ws_block as is; else the comb_mode whose result ps holds
  double* faces;           // per-block face buffer
  double* ws_new;          // [nb][ndim][bnz][16][16] new upper faces
  int ct;                  // CT_*
  int tgc;                 // transverse_ghost_cells
  int tvdlf;               // 1: no cmin (cbarmin = cbarmax)
  // 3-D scratch (per block, cells -2..17 in each direction)
  void* prim;              // primitive state, face precision
  double* ecc;             // b x v at cell centres (uct_contact)
  double* edges;           // edge electric fields times qdt and edge length
};

// reconstruct src/mod_constrained_transport.t:177-237 (TVD limiters) of four
// values q(-1..2) around the edge between q(0) and q(1)
template <int LIM>
__device__ __forceinline__ void ct_recon(const FVParams& fp, double rdel, double qm, double q0, double q1, double q2,
                                         double& qL, double& qR) {
  const double dm = q0 - qm, d0 = q1 - q0, dp = q2 - q1;
  double ldq, rdq, dummy;
  dwlimiter<double, LIM>(fp.limiter, dm, d0, rdel, ldq, dummy);
  dwlimiter<double, LIM>(fp.limiter, d0, dp, rdel, dummy, rdq);
  qL = q0 + 0.5 * ldq;
  qR = q1 - 0.5 * rdq;
}

}  // namespace
}  // namespace FV_NS
