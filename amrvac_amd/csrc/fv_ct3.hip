// Constrained transport on a uniform periodic 3-D level (ideal MHD, ndir = 3,
// staggered face field, typedivbfix = 'ct').
//
// A 16^3 block with its 2-cell halo is 20^3 cells, too many to stage in LDS
// as the 2-D kernels do, so a stage is four launches over per-block scratch
// in HBM, each indexed by the owner cell -2..17 in every direction:
//   ct3_prim_kernel    the primitive state of the stage input (face
//                      precision), and b x v at cell centres for uct_contact
//                      (FP64);
//   ct3_face_kernel    one thread per face: the flux and the face data of the
//                      edge fields (finite_volume src/mod_finite_volume.t:
//                      183-190 with mhd_modify_wLR :6338-6345 and
//                      mhd_get_ct_velocity_* src/mhd/mod_mhd_phys.t:
//                      3261-3326), faces -1..15 along the direction and
//                      -tgc..15+tgc across it;
//   ct3_edge_kernel    one thread per edge: the edge electric field times qdt
//                      and the edge length (mhd_update_faces_average /
//                      _contact / _hll :6951-7479), edges along idir at the
//                      upper corner of cells -1..15 in the two other
//                      directions and 0..15 along idir;
//   ct3_update_kernel  one thread per cell: its three upper faces and three
//                      lower faces from the circulation of the edge fields
//                      (the lower ones in the same operations as their
//                      owner, which stores them), the flux-divergence update,
//                      and the cell field from the faces
//                      (mhd_face_to_center :7586-7648).
// The new upper faces go to ws_new, which then replaces psb's face field.
//
// Face field layout: ws[nb][3][16][16][16], component d of cell (i, j, k) is
// the field on the upper d face of that cell.
#include "fv_ct_common.h"

namespace FV_NS {
namespace {

using fvc::BX;
constexpr int G = BX + 4;            // cells -2..17 per direction
constexpr size_t G3 = (size_t)G * G * G;

template <class P>
struct Ct3Layout {
  static constexpr int FK = P::NW + 5;
  // face data of the edge fields (FP64), after the fluxes
  static constexpr int VN = 0, CBMIN = 1, CBMAX = 2, VB1 = 3, VB2 = 4;
  // scratch per block, in doubles: faces [3][FK][G3] (the fluxes [3][NW][G3]
  // in the face precision, whose values they are, then the face data
  // [3][5][G3] in FP64), primitives [NW][G3] (FP64 or FP32), b x v [3][G3],
  // edge fields [3][G3]
  static constexpr size_t faces = 3 * FK * G3;
  static constexpr size_t prim = P::NW * G3;
  static constexpr size_t ecc = 3 * G3;
  static constexpr size_t edges = 3 * G3;
};

__device__ __forceinline__ size_t gidx(int x, int y, int z) {
  return ((size_t)(z + 2) * G + (y + 2)) * G + (x + 2);
}

__device__ __forceinline__ double ws_at3(const SweepArgs& a, const double* ws, int b, int d, int x, int y, int z) {
  double sign;
  const int sb = locate<3>(a, b, x, y, z, sign);
  return ws[(((size_t)sb * 3 + d) * BX + z) * BX * BX + (size_t)y * BX + x];
}

// the face field a slot combination gives (the RK initial state of psb)
__device__ __forceinline__ double ws_comb3(const SweepArgs& a, const CtArgs& t, int mode, int b, int d, int x, int y,
                                           int z) {
  if (mode == 0) return ws_at3(a, t.ws_dst, b, d, x, y, z);
  if (mode == 1) return ws_at3(a, t.ws_base, b, d, x, y, z);
  if (mode == 2)
    return a.comb_a * ws_at3(a, t.ws_base, b, d, x, y, z) + a.comb_b * ws_at3(a, t.ws_src, b, d, x, y, z);
  return ws_at3(a, t.ws_src, b, d, x, y, z);
}

template <class P, class R>
__global__ void __launch_bounds__(fvc::NTHREADS) ct3_prim_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  using L = Ct3Layout<P>;
  const int b = blockIdx.y;
  const size_t q = (size_t)blockIdx.x * fvc::NTHREADS + threadIdx.x;
  if (q >= G3) return;
  const int x = (int)(q % G) - 2, y = (int)(q / G % G) - 2, z = (int)(q / (G * G)) - 2;
  double c[P::NW], w[P::NW];
  load_cons_at<P, 3>(a, b, x, y, z, c);
  to_prim<P>(c, w, fp);
  R* pr = reinterpret_cast<R*>(t.prim) + (size_t)b * P::NW * G3;
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) pr[iw * G3 + q] = R(w[iw]);
  if (t.ct == CT_CONTACT) {
    // ECC = b x v (components x, y, z)
    double* e = t.ecc + (size_t)b * L::ecc;
    const int b1 = P::MAG, b2 = P::MAG + 1, b3 = P::MAG + 2, m1 = P::MOM, m2 = P::MOM + 1, m3 = P::MOM + 2;
    e[0 * G3 + q] = w[b2] * w[m3] - w[b3] * w[m2];
    e[1 * G3 + q] = w[b3] * w[m1] - w[b1] * w[m3];
    e[2 * G3 + q] = w[b1] * w[m2] - w[b2] * w[m1];
  }
}

// flux iw through the d faces, in the face precision R (an FP32 flux is
// stored as the FP32 value it is)
template <class P, class R>
__device__ __forceinline__ R* flux_ptr(const CtArgs& t, int b, int d, int iw) {
  using L = Ct3Layout<P>;
  return reinterpret_cast<R*>(t.faces + (size_t)b * L::faces) + ((size_t)d * P::NW + iw) * G3;
}

// face datum k (VN .. VB2) of the d faces, FP64
template <class P>
__device__ __forceinline__ double* ctd_ptr(const CtArgs& t, int b, int d, int k) {
  using L = Ct3Layout<P>;
  return t.faces + (size_t)b * L::faces + (size_t)3 * P::NW * G3 + ((size_t)d * 5 + k) * G3;
}

template <class P, class R, int RIEM, int LIM, int ID>
__device__ void ct3_one_face(const SweepArgs& a, const CtArgs& t, const FVParams& fp, int b, int i, int j, int k,
                             int& crash) {
  using L = Ct3Layout<P>;
  constexpr int NW = P::NW;
  const R* pr = reinterpret_cast<const R*>(t.prim) + (size_t)b * NW * G3;
  R st[4][NW];
  // cells i-1 .. i+2 along ID through the upper ID face of cell (i, j, k)
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const size_t c = gidx(ID == 0 ? i - 1 + s : i, ID == 1 ? j - 1 + s : j, ID == 2 ? k - 1 + s : k);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) st[s][iw] = pr[iw * G3 + c];
  }
  CtIO io;
  io.bn = ws_at3(a, t.ws_src, b, ID, i, j, k);
  if constexpr (RIEM == RIEM_HLLD)
    io.bx = t.bx_mode < 0 ? ws_at3(a, t.ws_block, b, ID, i, j, k) : ws_comb3(a, t, t.bx_mode, b, ID, i, j, k);
  R fC[NW];
  face_flux<P, R, RIEM, ID, LIM, 2, true>(st, fp, R(0), fC, crash, &io);
  const size_t c = gidx(i, j, k);
  R* fr = flux_ptr<P, R>(t, b, ID, 0) + c;
  // the normal field's flux is read by nothing: the edge fields take the
  // tangential components and the cell field is the face average
#pragma unroll
  for (int iw = 0; iw < NW; ++iw)
    if (iw != P::MAG + ID) fr[iw * G3] = fC[iw];
  double* f = ctd_ptr<P>(t, b, ID, 0) + c;
  if (t.ct == CT_CONTACT) f[L::VN * G3] = 0.5 * (io.vL[ID] + io.vR[ID]);
  if (t.ct == CT_HLL) {
    const double cbmax = io.cmax > 0.0 ? io.cmax : 0.0;
    const double cbmin = t.tvdlf ? cbmax : (-io.cmin > 0.0 ? -io.cmin : 0.0);
    constexpr int iN = (ID + 1) % 3, iE = (ID + 2) % 3;
    f[L::CBMIN * G3] = cbmin;
    f[L::CBMAX * G3] = cbmax;
    f[L::VB1 * G3] = (cbmax * io.vL[iN] + cbmin * io.vR[iN]) / (cbmax + cbmin);
    // the second average takes vbarRC(:,idim,1), as the reference does
    f[L::VB2 * G3] = (cbmax * io.vL[iE] + cbmin * io.vR[iN]) / (cbmax + cbmin);
  }
}

// thread q of a block's faces (or edges) -> cell (x, y, z), x fastest so
// that neighbouring lanes touch neighbouring addresses for every direction;
// lo / n: first index and count per direction
__device__ __forceinline__ bool box_cell(int q, const int* lo, const int* n, int& x, int& y, int& z) {
  if (q >= n[0] * n[1] * n[2]) return false;
  x = q % n[0] + lo[0];
  y = q / n[0] % n[1] + lo[1];
  z = q / (n[0] * n[1]) + lo[2];
  return true;
}

// blockIdx.y: direction, blockIdx.z: block; faces -1..15 along the
// direction, -tgc..15+tgc across it
template <class P, class R, int RIEM, int LIM>
__global__ void __launch_bounds__(fvc::NTHREADS) ct3_face_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  const int d = blockIdx.y, b = blockIdx.z;
  int lo[3], n[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    lo[e] = e == d ? -1 : -t.tgc;
    n[e] = e == d ? BX + 1 : BX + 2 * t.tgc;
  }
  int x, y, z;
  if (!box_cell(blockIdx.x * fvc::NTHREADS + threadIdx.x, lo, n, x, y, z)) return;
  int crash = 0;
  if (d == 0) ct3_one_face<P, R, RIEM, LIM, 0>(a, t, fp, b, x, y, z, crash);
  else if (d == 1) ct3_one_face<P, R, RIEM, LIM, 1>(a, t, fp, b, x, y, z, crash);
  else ct3_one_face<P, R, RIEM, LIM, 2>(a, t, fp, b, x, y, z, crash);
  if (crash) *a.crash = 1;
}

// blockIdx.y: edge direction idir, blockIdx.z: block; edges at the upper
// (idim1, idim2) corner of cells -1..15 across and 0..15 along idir, with
// (idim1, idim2, idir) cyclic (lvc = 1)
template <class P, class R, int LIM>
__global__ void __launch_bounds__(fvc::NTHREADS) ct3_edge_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  using L = Ct3Layout<P>;
  const int idir = blockIdx.y, b = blockIdx.z;
  int lo[3], n[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    lo[e] = e == idir ? 0 : -1;
    n[e] = e == idir ? BX : BX + 1;
  }
  int c[3];
  if (!box_cell(blockIdx.x * fvc::NTHREADS + threadIdx.x, lo, n, c[0], c[1], c[2])) return;
  const int idim1 = (idir + 1) % 3, idim2 = (idir + 2) % 3;
  int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
  e1[idim1] = 1;
  e2[idim2] = 1;
  auto at = [&](int s1, int s2) { return gidx(c[0] + s1 * e1[0] + s2 * e2[0], c[1] + s1 * e1[1] + s2 * e2[1],
                                              c[2] + s1 * e1[2] + s2 * e2[2]); };
  // FM(cell, component, face direction): flux of B_component through the
  // upper face of the cell
  auto FM = [&](size_t cell, int comp, int fd) { return double(flux_ptr<P, R>(t, b, fd, P::MAG + comp)[cell]); };
  auto FQ = [&](size_t cell, int k, int fd) { return ctd_ptr<P>(t, b, fd, k)[cell]; };
  double e;
  if (t.ct != CT_HLL) {
    const double fy0 = FM(at(0, 0), idim1, idim2), fy1 = FM(at(1, 0), idim1, idim2);
    const double fx0 = FM(at(0, 0), idim2, idim1), fx1 = FM(at(0, 1), idim2, idim1);
    e = 0.25 * (fy0 + fy1 - fx0 - fx1);
    if (t.ct == CT_CONTACT) {
      const double* ecc = t.ecc + (size_t)b * L::ecc + (size_t)idir * G3;
      auto E = [&](int s1, int s2) { return ecc[at(s1, s2)]; };
      const double v10 = FQ(at(0, 0), L::VN, idim1), v11 = FQ(at(0, 1), L::VN, idim1);
      const double v20 = FQ(at(0, 0), L::VN, idim2), v21 = FQ(at(1, 0), L::VN, idim2);
      // first pair: idim2-face fluxes of B_idim1 at c, c + e1
      {
        const double ELc = fy0 - E(0, 0), ELp = fy1 - E(1, 0);
        const double ERc = fy0 - E(0, 1), ERp = fy1 - E(1, 1);
        double ELC, ERC;
        if (v10 > 0.0) ELC = ELc;
        else if (v10 < 0.0) ELC = ELp;
        else ELC = 0.5 * (ELc + ELp);
        if (v11 > 0.0) ERC = ERc;
        else if (v11 < 0.0) ERC = ERp;
        else ERC = 0.5 * (ERc + ERp);
        e = e + 0.25 * (ELC + ERC);
      }
      // second pair: idim1-face fluxes of B_idim2 at c, c + e2
      {
        const double ELc = -fx0 - E(0, 0), ELp = -fx1 - E(0, 1);
        const double ERc = -fx0 - E(1, 0), ERp = -fx1 - E(1, 1);
        double ELC, ERC;
        if (v20 > 0.0) ELC = ELc;
        else if (v20 < 0.0) ELC = ELp;
        else ELC = 0.5 * (ELc + ELp);
        if (v21 > 0.0) ERC = ERc;
        else if (v21 < 0.0) ERC = ERp;
        else ERC = 0.5 * (ERc + ERp);
        e = e + 0.25 * (ELC + ERC);
      }
    }
    e = e * a.qdt * fp.dx[idir];
  } else {
    const double r1 = fp.rdelinv[idim1], r2 = fp.rdelinv[idim2];
    double vtL1, vtR1, vtL2, vtR2, bL1, bR1, bL2, bR2;
    // vbarC(idim1 faces, 1) along idim2 -> vtil(:, 2)
    ct_recon<LIM>(fp, r2, FQ(at(0, -1), L::VB1, idim1), FQ(at(0, 0), L::VB1, idim1), FQ(at(0, 1), L::VB1, idim1),
                  FQ(at(0, 2), L::VB1, idim1), vtL2, vtR2);
    // vbarC(idim2 faces, 2) along idim1 -> vtil(:, 1)
    ct_recon<LIM>(fp, r1, FQ(at(-1, 0), L::VB2, idim2), FQ(at(0, 0), L::VB2, idim2), FQ(at(1, 0), L::VB2, idim2),
                  FQ(at(2, 0), L::VB2, idim2), vtL1, vtR1);
    // face field of the stage input: B_idim1 along idim2, B_idim2 along idim1
    auto W = [&](int comp, int s1, int s2) {
      return ws_at3(a, t.ws_src, b, comp, c[0] + s1 * e1[0] + s2 * e2[0], c[1] + s1 * e1[1] + s2 * e2[1],
                    c[2] + s1 * e1[2] + s2 * e2[2]);
    };
    ct_recon<LIM>(fp, r2, W(idim1, 0, -1), W(idim1, 0, 0), W(idim1, 0, 1), W(idim1, 0, 2), bL1, bR1);
    ct_recon<LIM>(fp, r1, W(idim2, -1, 0), W(idim2, 0, 0), W(idim2, 1, 0), W(idim2, 2, 0), bL2, bR2);
    auto mx = [](double u, double v) { return u > v ? u : v; };
    const double cm1 = mx(FQ(at(0, 1), L::CBMIN, idim1), FQ(at(0, 0), L::CBMIN, idim1));
    const double cp1 = mx(FQ(at(0, 1), L::CBMAX, idim1), FQ(at(0, 0), L::CBMAX, idim1));
    const double cm2 = mx(FQ(at(1, 0), L::CBMIN, idim2), FQ(at(0, 0), L::CBMIN, idim2));
    const double cp2 = mx(FQ(at(1, 0), L::CBMAX, idim2), FQ(at(0, 0), L::CBMAX, idim2));
    e = -(cp1 * vtL1 * bL2 + cm1 * vtR1 * bR2 - cp1 * cm1 * (bR2 - bL2)) / (cp1 + cm1) +
        (cp2 * vtL2 * bL1 + cm2 * vtR2 * bR1 - cp2 * cm2 * (bR1 - bL1)) / (cp2 + cm2);
    e = a.qdt * fp.dx[idir] * e;
  }
  t.edges[(size_t)b * L::edges + (size_t)idir * G3 + at(0, 0)] = e;
}

// one thread per interior cell, blockIdx.y: block
template <class P, class R>
__global__ void __launch_bounds__(fvc::NTHREADS) ct3_update_kernel(SweepArgs a, CtArgs t, FVParams fp) {
  using L = Ct3Layout<P>;
  constexpr int NW = P::NW;
  const int b = blockIdx.y;
  const int q = blockIdx.x * fvc::NTHREADS + threadIdx.x;
  const int i = q % BX, j = q / BX % BX, k = q / (BX * BX);
  const double* E = t.edges + (size_t)b * L::edges;
  auto EF = [&](int idir, int x, int y, int z) { return E[(size_t)idir * G3 + gidx(x, y, z)]; };
  const double sC[3] = {fp.dx[1] * fp.dx[2], fp.dx[0] * fp.dx[2], fp.dx[0] * fp.dx[1]};
  // circulation around the upper d face of cell (x, y, z): the terms in the
  // order of idim2, idir ascending, with the sign lvc(d, idim2, idir)
  auto circ = [&](int d, int x, int y, int z) {
    double s = 0.0;
    if (d == 0) {
      s = s + (EF(2, x, y, z) - EF(2, x, y - 1, z));
      s = s - (EF(1, x, y, z) - EF(1, x, y, z - 1));
    } else if (d == 1) {
      s = s - (EF(2, x, y, z) - EF(2, x - 1, y, z));
      s = s + (EF(0, x, y, z) - EF(0, x, y, z - 1));
    } else {
      s = s + (EF(1, x, y, z) - EF(1, x - 1, y, z));
      s = s - (EF(0, x, y, z) - EF(0, x, y - 1, z));
    }
    return s;
  };
  double up[3], lo[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int x = i - (d == 0), y = j - (d == 1), z = k - (d == 2);
    up[d] = ws_comb3(a, t, a.comb_mode, b, d, i, j, k) - circ(d, i, j, k) / sC[d];
    lo[d] = ws_comb3(a, t, a.comb_mode, b, d, x, y, z) - circ(d, x, y, z) / sC[d];
    t.ws_new[(((size_t)b * 3 + d) * BX + k) * BX * BX + (size_t)j * BX + i] = up[d];
  }
  const size_t cells = (size_t)BX * BX * BX;
  const size_t o = (size_t)b * NW * cells + ((size_t)k * BX + j) * BX + i;
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
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R* f = flux_ptr<P, R>(t, b, d, 0);
    const size_t c1 = gidx(i, j, k), c0 = gidx(i - (d == 0), j - (d == 1), k - (d == 2));
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      if (iw >= P::MAG && iw < P::MAG + 3) continue;  // the field comes from the faces below
      const double f1 = a.dq[d] * double(f[iw * G3 + c1]), f0 = a.dq[d] * double(f[iw * G3 + c0]);
      acc[iw] = acc[iw] + f1 - f0;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) acc[P::MAG + d] = 0.5 / sC[d] * (up[d] * sC[d] + lo[d] * sC[d]);
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) a.dst[o + iw * cells] = acc[iw];
}

template <class P, int RIEM, int LIM>
hipError_t launch_ct3_lim(int mixed, const SweepArgs& a, const CtArgs& t, const FVParams& fp, hipStream_t s) {
  const int nb = a.nblocks;
  const int nac = BX + 2 * t.tgc;
  const dim3 gp((unsigned)((G3 + fvc::NTHREADS - 1) / fvc::NTHREADS), nb);
  const dim3 gf((unsigned)(((BX + 1) * nac * nac + fvc::NTHREADS - 1) / fvc::NTHREADS), 3, nb);
  const dim3 ge((unsigned)((BX * (BX + 1) * (BX + 1) + fvc::NTHREADS - 1) / fvc::NTHREADS), 3, nb);
  const dim3 gu((unsigned)(BX * BX * BX / fvc::NTHREADS), nb);
  if (mixed) {
    if constexpr (kExact) return hipErrorInvalidValue;  // the bit-reproducible build is FP64 only
    else {
      hipLaunchKernelGGL((ct3_prim_kernel<P, float>), gp, dim3(fvc::NTHREADS), 0, s, a, t, fp);
      hipLaunchKernelGGL((ct3_face_kernel<P, float, RIEM, LIM>), gf, dim3(fvc::NTHREADS), 0, s, a, t, fp);
      hipLaunchKernelGGL((ct3_edge_kernel<P, float, LIM>), ge, dim3(fvc::NTHREADS), 0, s, a, t, fp);
      hipLaunchKernelGGL((ct3_update_kernel<P, float>), gu, dim3(fvc::NTHREADS), 0, s, a, t, fp);
    }
  } else {
    hipLaunchKernelGGL((ct3_prim_kernel<P, double>), gp, dim3(fvc::NTHREADS), 0, s, a, t, fp);
    hipLaunchKernelGGL((ct3_face_kernel<P, double, RIEM, LIM>), gf, dim3(fvc::NTHREADS), 0, s, a, t, fp);
    hipLaunchKernelGGL((ct3_edge_kernel<P, double, LIM>), ge, dim3(fvc::NTHREADS), 0, s, a, t, fp);
    hipLaunchKernelGGL((ct3_update_kernel<P, double>), gu, dim3(fvc::NTHREADS), 0, s, a, t, fp);
  }
  return hipGetLastError();
}

template <class P, int RIEM>
hipError_t launch_ct3_riem(int mixed, int lim, const SweepArgs& a, const CtArgs& t, const FVParams& fp,
                           hipStream_t s) {
  switch (lim) {
    case LIM_MINMOD: return launch_ct3_lim<P, RIEM, LIM_MINMOD>(mixed, a, t, fp, s);
    case LIM_VANLEER: return launch_ct3_lim<P, RIEM, LIM_VANLEER>(mixed, a, t, fp, s);
    case LIM_CADA3: return launch_ct3_lim<P, RIEM, LIM_CADA3>(mixed, a, t, fp, s);
    default: return launch_ct3_lim<P, RIEM, 0>(mixed, a, t, fp, s);
  }
}

using P3 = Phys<3, 3, true, DIVB_NONE>;

}  // namespace

// one constrained-transport stage in 3-D (ideal MHD, ndir = 3, 16^3 blocks,
// TVD limiters; mixed: FP32 reconstruction and Riemann solver, FP64 state,
// face field and edge fields).  scratch: ct3_scratch_doubles(nblocks)
// doubles.
hipError_t launch_ct3_stage(int mixed, int riem, int lim, int ct, const fvc::SweepArgs& a, const double* ws_src,
                            const double* ws_base, const double* ws_dst, const double* ws_block, int bx_mode,
                            double* scratch, double* ws_new, const fvc::FVParams& fp, hipStream_t s) {
  using L = Ct3Layout<P3>;
  if (a.bx != BX || a.bnz != BX) return hipErrorInvalidValue;
  CtArgs t{};
  t.ws_src = ws_src;
  t.ws_base = ws_base;
  t.ws_dst = ws_dst;
  t.ws_block = ws_block;
  t.bx_mode = bx_mode;
  const size_t nb = (size_t)a.nblocks;
  t.faces = scratch;
  t.prim = scratch + nb * L::faces;
  t.ecc = scratch + nb * (L::faces + L::prim);
  t.edges = scratch + nb * (L::faces + L::prim + L::ecc);
  t.ws_new = ws_new;
  t.ct = ct;
  t.tgc = ct == CT_HLL ? 2 : 1;
  t.tvdlf = riem == RIEM_TVDLF;
  switch (riem) {
    case RIEM_HLL: return launch_ct3_riem<P3, RIEM_HLL>(mixed, lim, a, t, fp, s);
    case RIEM_TVDLF: return launch_ct3_riem<P3, RIEM_TVDLF>(mixed, lim, a, t, fp, s);
    case RIEM_HLLC: return launch_ct3_riem<P3, RIEM_HLLC>(mixed, lim, a, t, fp, s);
    case RIEM_HLLD: return launch_ct3_riem<P3, RIEM_HLLD>(mixed, lim, a, t, fp, s);
    default: return hipErrorInvalidValue;
  }
}

size_t ct3_scratch_doubles(int nblocks) {
  using L = Ct3Layout<P3>;
  return (size_t)nblocks * (L::faces + L::prim + L::ecc + L::edges);
}

}  // namespace FV_NS
