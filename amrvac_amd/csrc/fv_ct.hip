// Constrained transport on a uniform periodic 2-D level (ndir = 3, ideal MHD
// with the staggered face field, typedivbfix = 'ct').
//
// One stage is two launches:
//   ct_face_kernel    every face flux of the block plus transverse_ghost_cells
//                     rows (finite_volume src/mod_finite_volume.t:183-190 with
//                     mhd_modify_wLR :6338-6345), with the face data the edge
//                     fields need (mhd_get_ct_velocity_* src/mhd/
//                     mod_mhd_phys.t:3261-3326), into a per-block face buffer;
//   ct_update_kernel  edge electric fields (mhd_update_faces_average /
//                     _contact / _hll :6951-7479), the face-field update, the
//                     flux-divergence update of the cells, and the cell field
//                     from the faces (mhd_face_to_center :7586-7648);
//   the new face field then replaces psb's (the shared face of two blocks
//                     is computed by both; each stores its upper faces).
//
// Face field layout: ws[nb][2][16][16], component d of cell (i, j) is the
// field on the upper d face of that cell; a face buffer holds per block and
// direction a 20 x 20 grid (faces -2..17 in both indices) of CT_FK values:
// the NW unscaled fluxes, then vnorm, cbarmin, cbarmax, vbarC(1), vbarC(2).
#include "fv_ct_common.h"

namespace FV_NS {
namespace {

using fvc::BX;
constexpr int CTG = 20;            // face grid edge
// per block, in doubles: the fluxes [2][NW][20][20] in the face precision
// (an FP32 flux is stored as the FP32 value it is), then the face data of
// the edge fields [2][5][20][20] in FP64
template <class P>
struct CtLayout {
  static constexpr int FK = P::NW + 5;
  static constexpr int VN = 0, CBMIN = 1, CBMAX = 2, VB1 = 3, VB2 = 4;
  static constexpr size_t K = (size_t)CTG * CTG;
  static constexpr size_t per_block = 2 * (size_t)FK * K;
};

__device__ __forceinline__ double ws_at(const SweepArgs& a, const double* ws, int b, int d, int x, int y) {
  int z = 0;
  double sign;
  const int sb = locate<2>(a, b, x, y, z, sign);
  return ws[((size_t)sb * 2 + d) * BX * BX + (size_t)y * BX + x];
}

// the face field a slot combination gives (the RK initial state of psb)
__device__ __forceinline__ double ws_comb(const SweepArgs& a, const CtArgs& t, int mode, int b, int d, int x,
                                          int y) {
  if (mode == 0) return ws_at(a, t.ws_dst, b, d, x, y);
  if (mode == 1) return ws_at(a, t.ws_base, b, d, x, y);
  if (mode == 2) return a.comb_a * ws_at(a, t.ws_base, b, d, x, y) + a.comb_b * ws_at(a, t.ws_src, b, d, x, y);
  return ws_at(a, t.ws_src, b, d, x, y);
}

template <class P>
__device__ __forceinline__ void prim_at(const SweepArgs& a, int b, int x, int y, const FVParams& fp, double* w) {
  double c[P::NW];
  load_cons_at<P, 2>(a, b, x, y, 0, c);
  to_prim<P>(c, w, fp);
}

constexpr int CTS = BX + 4;  // staged cells -2..17 per direction

// the stage input's face field of faces -2..17 of a block, staged in LDS
__device__ __forceinline__ void stage_ws(const SweepArgs& a, const CtArgs& t, int b, double* wsl) {
  for (int q = threadIdx.x; q < 2 * CTS * CTS; q += fvc::NTHREADS) {
    const int d = q / (CTS * CTS), r = q % (CTS * CTS);
    wsl[q] = ws_at(a, t.ws_src, b, d, r % CTS - 2, r / CTS - 2);
  }
}
__device__ __forceinline__ double wsl_at(const double* wsl, int d, int x, int y) {
  return wsl[d * CTS * CTS + (y + 2) * CTS + (x + 2)];
}

template <class P, class R, int RIEM, int LIM, int ID>
__device__ void ct_one_face(const SweepArgs& a, const CtArgs& t, const FVParams& fp, const R* lds,
                            const double* wsl, int b, int i, int j, int& crash) {
  using L = CtLayout<P>;
  constexpr int NW = P::NW;
  R st[4][NW];
  // cells i-1 .. i+2 along ID through the upper ID face of cell (i, j),
  // primitive, from the block's staged state
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int x = ID == 0 ? i - 1 + s : i, y = ID == 1 ? j - 1 + s : j;
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) st[s][iw] = lds[iw * CTS * CTS + (y + 2) * CTS + (x + 2)];
  }
  CtIO io;
  io.bn = wsl_at(wsl, ID, i, j);
  if constexpr (RIEM == RIEM_HLLD)
    io.bx = t.bx_mode < 0 ? ws_at(a, t.ws_block, b, ID, i, j) : ws_comb(a, t, t.bx_mode, b, ID, i, j);
  R fC[NW];
  face_flux<P, R, RIEM, ID, LIM, 2, true>(st, fp, R(0), fC, crash, &io);
  constexpr size_t K = L::K;
  const size_t c = (size_t)(j + 2) * CTG + (i + 2);
  R* fr = reinterpret_cast<R*>(t.faces + (size_t)b * L::per_block) + (size_t)ID * NW * K + c;
#pragma unroll
  // the normal field's flux is read by nothing: the edge fields take the
  // tangential components and the cell field is the face average
  for (int iw = 0; iw < NW; ++iw)
    if (iw != P::MAG + ID) fr[iw * K] = fC[iw];
  double* f = t.faces + (size_t)b * L::per_block + 2 * NW * K + (size_t)ID * 5 * K + c;
  if (t.ct == CT_CONTACT) f[L::VN * K] = 0.5 * (io.vL[ID] + io.vR[ID]);
  if (t.ct == CT_HLL) {
    const double cbmax = io.cmax > 0.0 ? io.cmax : 0.0;
    const double cbmin = t.tvdlf ? cbmax : (-io.cmin > 0.0 ? -io.cmin : 0.0);
    constexpr int iN = (ID + 1) % P::NDIR, iE = (ID + 2) % P::NDIR;
    f[L::CBMIN * K] = cbmin;
    f[L::CBMAX * K] = cbmax;
    f[L::VB1 * K] = (cbmax * io.vL[iN] + cbmin * io.vR[iN]) / (cbmax + cbmin);
    // the second average takes vbarRC(:,idim,1), as the reference does
    f[L::VB2 * K] = (cbmax * io.vL[iE] + cbmin * io.vR[iN]) / (cbmax + cbmin);
  }
}

// one workgroup per block: the primitive state of cells -2..17 (halo from
// the neighbours) staged in LDS once (in the face precision R: FP64, or FP32
// for the mixed form), then both directions' faces: -1..15 along the
// direction, -tgc..15+tgc across it
template <class P, class R, int RIEM, int LIM>
__global__ void __launch_bounds__(fvc::NTHREADS) ct_face_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  __shared__ R lds[P::NW * CTS * CTS];
  __shared__ double wsl[2 * CTS * CTS];
  const int b = wg_block(a.per_xcd);
  if (b >= a.nblocks) return;
  stage_ws(a, t, b, wsl);
  for (int q = threadIdx.x; q < CTS * CTS; q += fvc::NTHREADS) {
    double w[P::NW];
    prim_at<P>(a, b, q % CTS - 2, q / CTS - 2, fp, w);
#pragma unroll
    for (int iw = 0; iw < P::NW; ++iw) lds[iw * CTS * CTS + q] = R(w[iw]);
  }
  __syncthreads();
  const int nacross = BX + 2 * t.tgc, nalong = BX + 1, nf = nacross * nalong;
  int crash = 0;
  for (int q = threadIdx.x; q < 2 * nf; q += fvc::NTHREADS) {
    const int r = q < nf ? q : q - nf;
    const int along = r % nalong - 1, across = r / nalong - t.tgc;
    if (q < nf) ct_one_face<P, R, RIEM, LIM, 0>(a, t, fp, lds, wsl, b, along, across, crash);
    else ct_one_face<P, R, RIEM, LIM, 1>(a, t, fp, lds, wsl, b, across, along, crash);
  }
  if (crash) *a.crash = 1;
}

// flux iw through the dim faces (stored in the face precision R)
template <class P, class R>
__device__ __forceinline__ double fget(const CtArgs& t, int b, int dim, int iw, int i, int j) {
  using L = CtLayout<P>;
  const R* f = reinterpret_cast<const R*>(t.faces + (size_t)b * L::per_block);
  return double(f[((size_t)dim * P::NW + iw) * L::K + (size_t)(j + 2) * CTG + (i + 2)]);
}

// face datum k (VN .. VB2) of the dim faces (FP64)
template <class P>
__device__ __forceinline__ double dget(const CtArgs& t, int b, int dim, int k, int i, int j) {
  using L = CtLayout<P>;
  return t.faces[(size_t)b * L::per_block + 2 * P::NW * L::K + ((size_t)dim * 5 + k) * L::K + (size_t)(j + 2) * CTG +
                 (i + 2)];
}

template <class P, class R, int LIM>
__global__ void __launch_bounds__(fvc::NTHREADS) ct_update_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  using L = CtLayout<P>;
  constexpr int NW = P::NW, B1 = P::MAG, B2 = P::MAG + 1;
  constexpr int NE = BX + 1;  // edges -1..15
  __shared__ double fE[NE * NE];
  __shared__ double wsx[NE * BX];  // x faces i = -1..15, rows 0..15
  __shared__ double wsy[BX * NE];  // y faces rows j = -1..15, columns 0..15
  __shared__ double ecc[(BX + 2) * (BX + 2)];  // cells -1..16
  __shared__ double wsl[2 * CTS * CTS];         // stage input's faces -2..17 (uct_hll)
  const int b = wg_block(a.per_xcd);
  if (b >= a.nblocks) return;
  if (t.ct == CT_HLL) {
    stage_ws(a, t, b, wsl);
    __syncthreads();
  }
  if (t.ct == CT_CONTACT) {
    for (int q = threadIdx.x; q < (BX + 2) * (BX + 2); q += fvc::NTHREADS) {
      const int x = q % (BX + 2) - 1, y = q / (BX + 2) - 1;
      double w[NW];
      prim_at<P>(a, b, x, y, fp, w);
      ecc[q] = w[P::MAG] * w[P::MOM + 1] - w[P::MAG + 1] * w[P::MOM];
    }
    __syncthreads();
  }
  // ---- edge electric fields (along z, at the upper x/y corner of cell (i, j))
  for (int q = threadIdx.x; q < NE * NE; q += fvc::NTHREADS) {
    const int i = q % NE - 1, j = q / NE - 1;
    double e;
    if (t.ct != CT_HLL) {
      // fC(b1) through y faces, fC(b2) through x faces
      const double fy0 = fget<P, R>(t, b, 1, B1, i, j), fy1 = fget<P, R>(t, b, 1, B1, i + 1, j);
      const double fx0 = fget<P, R>(t, b, 0, B2, i, j), fx1 = fget<P, R>(t, b, 0, B2, i, j + 1);
      e = 0.25 * (fy0 + fy1 - fx0 - fx1);
      if (t.ct == CT_CONTACT) {
        auto E = [&](int x, int y) { return ecc[(y + 1) * (BX + 2) + (x + 1)]; };
        const double vx0 = dget<P>(t, b, 0, L::VN, i, j), vx1 = dget<P>(t, b, 0, L::VN, i, j + 1);
        const double vy0 = dget<P>(t, b, 1, L::VN, i, j), vy1 = dget<P>(t, b, 1, L::VN, i + 1, j);
        // first pair: y-face fluxes of b1 at (i, j), (i+1, j)
        {
          const double ELc = fy0 - E(i, j), ELp = fy1 - E(i + 1, j);
          const double ERc = fy0 - E(i, j + 1), ERp = fy1 - E(i + 1, j + 1);
          double ELC, ERC;
          if (vx0 > 0.0) ELC = ELc;
          else if (vx0 < 0.0) ELC = ELp;
          else ELC = 0.5 * (ELc + ELp);
          if (vx1 > 0.0) ERC = ERc;
          else if (vx1 < 0.0) ERC = ERp;
          else ERC = 0.5 * (ERc + ERp);
          e = e + 0.25 * (ELC + ERC);
        }
        // second pair: x-face fluxes of b2 at (i, j), (i, j+1)
        {
          const double ELc = -fx0 - E(i, j), ELp = -fx1 - E(i, j + 1);
          const double ERc = -fx0 - E(i + 1, j), ERp = -fx1 - E(i + 1, j + 1);
          double ELC, ERC;
          if (vy0 > 0.0) ELC = ELc;
          else if (vy0 < 0.0) ELC = ELp;
          else ELC = 0.5 * (ELc + ELp);
          if (vy1 > 0.0) ERC = ERc;
          else if (vy1 < 0.0) ERC = ERp;
          else ERC = 0.5 * (ERc + ERp);
          e = e + 0.25 * (ELC + ERC);
        }
      }
      e = e * a.qdt;  // times the edge length 1 of the ignorable direction
    } else {
      // idir = z: idim1 = x, idim2 = y
      const double rx = fp.rdelinv[0], ry = fp.rdelinv[1];
      double vtL1, vtR1, vtL2, vtR2, btLx, btRx, btLy, btRy;
      // vbarC(x faces, 1) along y -> vtil(:, 2)
      ct_recon<LIM>(fp, ry, dget<P>(t, b, 0, L::VB1, i, j - 1), dget<P>(t, b, 0, L::VB1, i, j),
                    dget<P>(t, b, 0, L::VB1, i, j + 1), dget<P>(t, b, 0, L::VB1, i, j + 2), vtL2, vtR2);
      // vbarC(y faces, 2) along x -> vtil(:, 1)
      ct_recon<LIM>(fp, rx, dget<P>(t, b, 1, L::VB2, i - 1, j), dget<P>(t, b, 1, L::VB2, i, j),
                    dget<P>(t, b, 1, L::VB2, i + 1, j), dget<P>(t, b, 1, L::VB2, i + 2, j), vtL1, vtR1);
      // face field of the stage input: x faces along y, y faces along x
      ct_recon<LIM>(fp, ry, wsl_at(wsl, 0, i, j - 1), wsl_at(wsl, 0, i, j), wsl_at(wsl, 0, i, j + 1),
                    wsl_at(wsl, 0, i, j + 2), btLx, btRx);
      ct_recon<LIM>(fp, rx, wsl_at(wsl, 1, i - 1, j), wsl_at(wsl, 1, i, j), wsl_at(wsl, 1, i + 1, j),
                    wsl_at(wsl, 1, i + 2, j), btLy, btRy);
      auto mx = [](double u, double v) { return u > v ? u : v; };
      const double cm1 = mx(dget<P>(t, b, 0, L::CBMIN, i, j + 1), dget<P>(t, b, 0, L::CBMIN, i, j));
      const double cp1 = mx(dget<P>(t, b, 0, L::CBMAX, i, j + 1), dget<P>(t, b, 0, L::CBMAX, i, j));
      const double cm2 = mx(dget<P>(t, b, 1, L::CBMIN, i + 1, j), dget<P>(t, b, 1, L::CBMIN, i, j));
      const double cp2 = mx(dget<P>(t, b, 1, L::CBMAX, i + 1, j), dget<P>(t, b, 1, L::CBMAX, i, j));
      e = -(cp1 * vtL1 * btLy + cm1 * vtR1 * btRy - cp1 * cm1 * (btRy - btLy)) / (cp1 + cm1) +
          (cp2 * vtL2 * btLx + cm2 * vtR2 * btRx - cp2 * cm2 * (btRx - btLx)) / (cp2 + cm2);
      e = a.qdt * e;
    }
    fE[q] = e;
  }
  __syncthreads();
  auto FE = [&](int i, int j) { return fE[(j + 1) * NE + (i + 1)]; };
  // ---- face field: B_face -= circulation / face area
  const double sCx = fp.dx[1], sCy = fp.dx[0];
  for (int q = threadIdx.x; q < NE * BX; q += fvc::NTHREADS) {
    {  // x face (i, j), i = -1..15
      const int i = q % NE - 1, j = q / NE;
      const double circ = 0.0 + (FE(i, j) - FE(i, j - 1));
      const double v = ws_comb(a, t, a.comb_mode, b, 0, i, j) - circ / sCx;
      wsx[j * NE + (i + 1)] = v;
      if (i >= 0) t.ws_new[((size_t)b * 2 + 0) * BX * BX + (size_t)j * BX + i] = v;
    }
    {  // y face (i, j), j = -1..15
      const int i = q % BX, j = q / BX - 1;
      const double circ = 0.0 - (FE(i, j) - FE(i - 1, j));
      const double v = ws_comb(a, t, a.comb_mode, b, 1, i, j) - circ / sCy;
      wsy[(j + 1) * BX + i] = v;
      if (j >= 0) t.ws_new[((size_t)b * 2 + 1) * BX * BX + (size_t)j * BX + i] = v;
    }
  }
  __syncthreads();
  // ---- cells: RK initial state + flux divergence, then the field from the faces
  const int i = threadIdx.x % BX, j = threadIdx.x / BX;
  const size_t cells = (size_t)BX * BX;
  const size_t o = (size_t)b * NW * cells + (size_t)j * BX + i;
  double acc[NW];
  if (a.comb_mode == 0 || a.comb_mode == 1) {
    const double* s0 = a.comb_mode == 0 ? a.dst : a.base;
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) acc[iw] = s0[o + iw * cells];
  } else if (a.comb_mode == 2) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) acc[iw] = a.comb_a * a.base[o + iw * cells] + a.comb_b * a.src[o + iw * cells];
  } else {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) acc[iw] = a.src[o + iw * cells];
  }
  // B1 and B2 come from the faces below, B3 from its fluxes
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    if (iw == B1 || iw == B2) continue;
    const double fx1 = a.dq[0] * fget<P, R>(t, b, 0, iw, i, j), fx0 = a.dq[0] * fget<P, R>(t, b, 0, iw, i - 1, j);
    acc[iw] = acc[iw] + fx1 - fx0;
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    if (iw == B1 || iw == B2) continue;
    const double fy1 = a.dq[1] * fget<P, R>(t, b, 1, iw, i, j), fy0 = a.dq[1] * fget<P, R>(t, b, 1, iw, i, j - 1);
    acc[iw] = acc[iw] + fy1 - fy0;
  }
  acc[B1] = 0.5 / sCx * (wsx[j * NE + (i + 1)] * sCx + wsx[j * NE + i] * sCx);
  acc[B2] = 0.5 / sCy * (wsy[(j + 1) * BX + i] * sCy + wsy[j * BX + i] * sCy);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) a.dst[o + iw * cells] = acc[iw];
}

template <class P, int RIEM, int LIM>
hipError_t launch_ct_lim(int mixed, const SweepArgs& a, const CtArgs& t, const FVParams& fp, hipStream_t s) {
  if (mixed) {
    if constexpr (kExact) return hipErrorInvalidValue;  // the bit-reproducible build is FP64 only
    else {
      hipLaunchKernelGGL((ct_face_kernel<P, float, RIEM, LIM>), dim3(8 * a.per_xcd), dim3(fvc::NTHREADS), 0, s, a,
                         t, fp);
      hipLaunchKernelGGL((ct_update_kernel<P, float, LIM>), dim3(8 * a.per_xcd), dim3(fvc::NTHREADS), 0, s, a, t,
                         fp);
    }
  } else {
    hipLaunchKernelGGL((ct_face_kernel<P, double, RIEM, LIM>), dim3(8 * a.per_xcd), dim3(fvc::NTHREADS), 0, s, a,
                       t, fp);
    hipLaunchKernelGGL((ct_update_kernel<P, double, LIM>), dim3(8 * a.per_xcd), dim3(fvc::NTHREADS), 0, s, a, t, fp);
  }
  return hipGetLastError();
}

template <class P, int RIEM>
hipError_t launch_ct_riem(int mixed, int lim, const SweepArgs& a, const CtArgs& t, const FVParams& fp,
                          hipStream_t s) {
  switch (lim) {
    case LIM_MINMOD: return launch_ct_lim<P, RIEM, LIM_MINMOD>(mixed, a, t, fp, s);
    case LIM_VANLEER: return launch_ct_lim<P, RIEM, LIM_VANLEER>(mixed, a, t, fp, s);
    case LIM_CADA3: return launch_ct_lim<P, RIEM, LIM_CADA3>(mixed, a, t, fp, s);
    default: return launch_ct_lim<P, RIEM, 0>(mixed, a, t, fp, s);
  }
}

}  // namespace

// one constrained-transport stage (2-D ideal MHD, ndir = 3, TVD limiters;
// mixed: FP32 reconstruction and Riemann solver, FP64 state and edge fields)
hipError_t launch_ct_stage(int mixed, int riem, int lim, int ct, const fvc::SweepArgs& a, const double* ws_src,
                           const double* ws_base, const double* ws_dst, const double* ws_block, int bx_mode,
                           double* faces, double* ws_new, const fvc::FVParams& fp, hipStream_t s) {
  using P = Phys<2, 3, true, DIVB_NONE>;
  CtArgs t{};
  t.ws_src = ws_src;
  t.ws_base = ws_base;
  t.ws_dst = ws_dst;
  t.ws_block = ws_block;
  t.bx_mode = bx_mode;
  t.faces = faces;
  t.ws_new = ws_new;
  t.ct = ct;
  t.tgc = ct == CT_HLL ? 2 : 1;
  t.tvdlf = riem == RIEM_TVDLF;
  switch (riem) {
    case RIEM_HLL: return launch_ct_riem<P, RIEM_HLL>(mixed, lim, a, t, fp, s);
    case RIEM_TVDLF: return launch_ct_riem<P, RIEM_TVDLF>(mixed, lim, a, t, fp, s);
    case RIEM_HLLC: return launch_ct_riem<P, RIEM_HLLC>(mixed, lim, a, t, fp, s);
    case RIEM_HLLD: return launch_ct_riem<P, RIEM_HLLD>(mixed, lim, a, t, fp, s);
    default: return hipErrorInvalidValue;
  }
}

size_t ct_face_buffer_doubles(int nblocks) {
  return (size_t)nblocks * CtLayout<Phys<2, 3, true, DIVB_NONE>>::per_block;
}

}  // namespace FV_NS
