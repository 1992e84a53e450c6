// fv_sweep.h -- the fused finite-volume stage kernel.
//
// One workgroup (256 lanes = 16x16) updates one 16x16(x16) block of the
// level.  The block's stencil data never round-trips through HBM ghost
// cells: halo cells are read straight from the face neighbours' interiors
// through the level's neighbour table (or from a physical-boundary rule),
// converted to primitive variables once and staged in LDS.  In 3-D the
// workgroup marches along z with a per-lane register window of 2*NG
// primitive planes; x/y faces of the current plane come from LDS, the z face
// from the window, and each z-face flux is computed once and carried.
//
// Per cell and stage the kernel does what finite_volume
// (src/mod_finite_volume.t:116-339) or hancock (:19-113) plus addsource2
// (src/mod_source.t:130-160) do for one block, fused with the RK
// combination of src/mod_advance.t:251-267:
//   wnew = init + sum_d [ (-qdt/dx_d) fC_d(i+1/2) - (-qdt/dx_d) fC_d(i-1/2) ]
//          + sources(wCT)
// with init = psb, psb = base, or a*base + b*psa.
#pragma once

#include <type_traits>

#include "fv_device.h"

namespace fvc {
constexpr int BX = 16;  // block_nx(1) = block_nx(2) = 16
constexpr int NTHREADS = 256;

struct SweepArgs {
  const double* src;   // psa (stage input, wCT)
  const double* base;  // RK base slot (comb_mode 1, 2)
  double* dst;         // psb
  const int* nbr;      // [nb][3^ndim]
  int nblocks;
  int bnz;             // block_nx(3) (1 in 2-D)
  int per_xcd;         // ceil(nblocks/8)
  int comb_mode;       // 0: init = dst, 1: init = base, 2: init = a*base + b*src
  double comb_a, comb_b;
  double dq[3];        // -qdt/dx_d
  double qdt;
  double ch;           // cmax_global (GLM)
  double glm_decay;    // exp(-qdt*ch*alpha/min dx)
  int bc[6];           // physical boundary type per side
  int* crash;
};

}  // namespace fvc

namespace FV_NS {
using fvc::BX;
using fvc::NTHREADS;
using fvc::SweepArgs;

// Block id of a workgroup: consecutive blocks of the Morton-ordered level go
// to the same XCD (workgroups are dealt round-robin over the 8 XCDs), so that
// halo reads of neighbouring blocks hit the XCD's own L2.
__device__ __forceinline__ int wg_block(int per_xcd) {
  const int g = blockIdx.x;
  return (g & 7) * per_xcd + (g >> 3);
}

// Resolve the source of cell (x,y,z) (block-local, may lie outside the
// block by up to NG) for block b: returns the source block and rewrites the
// coordinates into it; *sign = -1 for an antisymmetric boundary.
template <int NDIM>
__device__ __forceinline__ int locate(const SweepArgs& a, int b, int& x, int& y, int& z, double& sign) {
  int c[3] = {x, y, z};
  const int n[3] = {BX, BX, a.bnz};
  int dir[3] = {0, 0, 0};
  sign = 1.0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) dir[d] = c[d] < 0 ? -1 : (c[d] >= n[d] ? 1 : 0);
  int nn = 1, idx = 0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) { idx += (dir[d] + 1) * nn; nn *= 3; }
  int nb = a.nbr[b * nn + idx];
  if (nb >= 0) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) c[d] -= dir[d] * n[d];
    x = c[0]; y = c[1]; z = c[2];
    return nb;
  }
  // physical boundary: apply cont / symm / asymm in every direction whose
  // face neighbour is a boundary, then take the remaining offset from the
  // neighbour in the other directions
  // (bc_phys src/mod_boundary_conditions.t:12-120)
  int rem[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    if (dir[d] == 0) continue;
    int fidx = 0, m = 1;
#pragma unroll
    for (int e = 0; e < NDIM; ++e) { fidx += ((e == d ? dir[d] : 0) + 1) * m; m *= 3; }
    if (a.nbr[b * nn + fidx] >= 0) { rem[d] = dir[d]; continue; }
    const int t = a.bc[2 * d + (dir[d] > 0 ? 1 : 0)];
    if (t == 2) {  // cont
      c[d] = dir[d] < 0 ? 0 : n[d] - 1;
    } else {       // symm / asymm
      c[d] = dir[d] < 0 ? -1 - c[d] : 2 * n[d] - 1 - c[d];
      if (t == 4) sign = -sign;
    }
  }
  int ridx = 0, m = 1;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) { ridx += (rem[d] + 1) * m; m *= 3; }
  int src = b;
  if (ridx != (nn - 1) / 2) {
    src = a.nbr[b * nn + ridx];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) c[d] -= rem[d] * n[d];
  }
  x = c[0]; y = c[1]; z = c[2];
  return src;
}

template <class P>
__device__ __forceinline__ void load_cons(const double* __restrict__ w, int b, int x, int y, int z,
                                          int bnz, double* c) {
  const size_t cells = (size_t)BX * BX * bnz;
  const size_t off = (size_t)b * P::NW * cells + ((size_t)z * BX + y) * BX + x;
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) c[iw] = w[off + iw * cells];
}

// primitive state of an arbitrary (possibly halo) cell
template <class P, int NDIM>
__device__ __forceinline__ void load_prim(const SweepArgs& a, const FVParams& fp, int b, int x, int y,
                                          int z, double* w) {
  double sign;
  const int sb = locate<NDIM>(a, b, x, y, z, sign);
  double c[P::NW];
  load_cons<P>(a.src, sb, x, y, z, a.bnz, c);
  if (sign < 0.0) {
#pragma unroll
    for (int iw = 0; iw < P::NW; ++iw) c[iw] = -c[iw];
  }
  to_prim<P>(c, w, fp);
}

template <class P, class R, int RIEM, int LIM, int NG>
__global__ void __launch_bounds__(NTHREADS)
sweep_kernel(SweepArgs a, FVParams fp) {
  constexpr int NDIM = P::NDIM;
  constexpr int NW = P::NW;
  constexpr int PX = BX + 2 * NG;
  constexpr bool CORNERS = P::LINDE;  // Linde needs div B one cell out
  constexpr bool HANCOCK = RIEM == RIEM_HANCOCK;
  static_assert(!(P::LINDE && NDIM == 3), "Linde cleaning is 2-D only on the device path");

  __shared__ R sp[NW][PX][PX];          // primitive plane with halo
  __shared__ R fxs[HANCOCK ? 1 : NW][BX][BX + 1];  // x faces: fxs[iw][y][f]
  __shared__ R fys[HANCOCK ? 1 : NW][BX + 1][BX];  // y faces: fys[iw][f][x]

  const int b = wg_block(a.per_xcd);
  if (b >= a.nblocks) return;
  const int tid = threadIdx.x;
  const int tx = tid & (BX - 1);
  const int ty = tid >> 4;
  const R ch = R(a.ch);
  int crash = 0;
  const size_t cells = (size_t)BX * BX * a.bnz;
  const size_t boff = (size_t)b * NW * cells;

  // z window (3-D): W[m] = primitive plane zb+m of this lane's column
  constexpr int NWIN = NDIM == 3 ? 2 * NG : 1;
  R W[NWIN][NW];
  R fz_carry[NW];
  if constexpr (NDIM == 3) {
    // prologue: planes -NG .. NG-1, carried face -1/2
#pragma unroll
    for (int m = 0; m < 2 * NG; ++m) {
      double w[NW];
      load_prim<P, NDIM>(a, fp, b, tx, ty, m - NG, w);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) W[m][iw] = R(w[iw]);
    }
    if constexpr (!HANCOCK) face_flux<P, R, RIEM, 2, LIM, NG>(W, fp, ch, fz_carry, crash);
  }

  const int nz = NDIM == 3 ? a.bnz : 1;
  for (int k = 0; k < nz; ++k) {
    // ---- stage plane k (primitive, with x/y halo) in LDS -----------------
    double cons_k[NW];  // wCT of this lane's cell (for init when comb uses it)
    if constexpr (NDIM == 3) {
      // advance the window: drop plane k-NG, load plane k+NG, so that
      // W[m] = plane k-NG+1+m and plane k = W[NG-1]
#pragma unroll
      for (int m = 0; m < 2 * NG - 1; ++m)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) W[m][iw] = W[m + 1][iw];
      {
        double w[NW];
        load_prim<P, NDIM>(a, fp, b, tx, ty, k + NG, w);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) W[2 * NG - 1][iw] = R(w[iw]);
      }
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) sp[iw][ty + NG][tx + NG] = W[NG - 1][iw];
      load_cons<P>(a.src, b, tx, ty, k, a.bnz, cons_k);
    } else {
      load_cons<P>(a.src, b, tx, ty, 0, 1, cons_k);
      double w[NW];
      to_prim<P>(cons_k, w, fp);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) sp[iw][ty + NG][tx + NG] = R(w[iw]);
    }
    // halo strips: 4 strips of NG x 16 cells (+ 4 NG x NG corners for Linde)
    {
      constexpr int NSTRIP = 4 * NG * BX;
      constexpr int NCORN = CORNERS ? 4 * NG * NG : 0;
      for (int t = tid; t < NSTRIP + NCORN; t += NTHREADS) {
        int x, y;
        if (t < NSTRIP) {
          const int s = t / (NG * BX), r = t % (NG * BX);
          const int along = r % BX, depth = r / BX;
          if (s == 0) { x = -NG + depth; y = along; }
          else if (s == 1) { x = BX + depth; y = along; }
          else if (s == 2) { x = along; y = -NG + depth; }
          else { x = along; y = BX + depth; }
        } else {
          const int u = t - NSTRIP;
          const int q = u / (NG * NG), r = u % (NG * NG);
          const int ox = r % NG, oy = r / NG;
          x = (q & 1) ? BX + ox : -NG + ox;
          y = (q & 2) ? BX + oy : -NG + oy;
        }
        double w[NW];
        load_prim<P, NDIM>(a, fp, b, x, y, k, w);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) sp[iw][y + NG][x + NG] = R(w[iw]);
      }
    }
    __syncthreads();

    double acc[NW];
    // ---- init: the RK combination (src/mod_advance.t:251-267) ------------
    {
      const size_t o = boff + ((size_t)k * BX + ty) * BX + tx;
      if (a.comb_mode == 0) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = a.dst[o + iw * cells];
      } else if (a.comb_mode == 1) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = a.base[o + iw * cells];
      } else {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = a.comb_a * a.base[o + iw * cells] + a.comb_b * cons_k[iw];
      }
    }

    if constexpr (!HANCOCK) {
      // ---- face fluxes of this plane --------------------------------------
      constexpr int NXF = BX * (BX + 1);      // 272 x faces
      constexpr int XSLOTS = 320;             // padded to a wave multiple
      constexpr int NSLOT = XSLOTS + NXF;     // + 272 y faces
      for (int t = tid; t < NSLOT; t += NTHREADS) {
        if (t < NXF) {
          const int y = t / (BX + 1), f = t % (BX + 1);  // face between f-1 and f
          R st[2 * NG][NW];
#pragma unroll
          for (int s = 0; s < 2 * NG; ++s)
#pragma unroll
            for (int iw = 0; iw < NW; ++iw) st[s][iw] = sp[iw][y + NG][f + s];
          R fc[NW];
          face_flux<P, R, RIEM, 0, LIM, NG>(st, fp, ch, fc, crash);
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) fxs[iw][y][f] = fc[iw];
        } else if (t >= XSLOTS) {
          const int u = t - XSLOTS;
          const int f = u / BX, x = u % BX;  // face between f-1 and f
          R st[2 * NG][NW];
#pragma unroll
          for (int s = 0; s < 2 * NG; ++s)
#pragma unroll
            for (int iw = 0; iw < NW; ++iw) st[s][iw] = sp[iw][f + s][x + NG];
          R fc[NW];
          face_flux<P, R, RIEM, 1, LIM, NG>(st, fp, ch, fc, crash);
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) fys[iw][f][x] = fc[iw];
        }
      }
      R fz_new[NW];
      if constexpr (NDIM == 3) face_flux<P, R, RIEM, 2, LIM, NG>(W, fp, ch, fz_new, crash);
      __syncthreads();
      // ---- flux divergence, dimension by dimension (:269-284) ------------
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        acc[iw] = acc[iw] + a.dq[0] * double(fxs[iw][ty][tx + 1]) - a.dq[0] * double(fxs[iw][ty][tx]);
        acc[iw] = acc[iw] + a.dq[1] * double(fys[iw][ty + 1][tx]) - a.dq[1] * double(fys[iw][ty][tx]);
        if constexpr (NDIM == 3) {
          acc[iw] = acc[iw] + a.dq[2] * double(fz_new[iw]) - a.dq[2] * double(fz_carry[iw]);
          fz_carry[iw] = fz_new[iw];
        }
      }
    } else {
      // ---- Hancock predictor (src/mod_finite_volume.t:19-113) -------------
      const R inv_g1 = R(fp.inv_gamma_1);
      auto hancock_dim = [&](auto IDc, const R* wm, const R* w0, const R* wp) {
        constexpr int ID = decltype(IDc)::value;
        R wL[NW], wR[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) {
          R ldw, rdw;
          dwlimiter<R, LIM>(fp.limiter, w0[iw] - wm[iw], wp[iw] - w0[iw], R(fp.rdelinv[ID]), ldw, rdw);
          wL[iw] = w0[iw] + R(0.5) * ldw;
          wR[iw] = w0[iw] - R(0.5) * rdw;
        }
        R wLC[NW], wRC[NW], fL[NW], fR[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) { wLC[iw] = wL[iw]; wRC[iw] = wR[iw]; }
        to_cons<P, R>(wLC, inv_g1);
        to_cons<P, R>(wRC, inv_g1);
        phys_flux<P, R, ID>(wRC, wR, ch, fR);
        phys_flux<P, R, ID>(wLC, wL, ch, fL);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = acc[iw] + a.dq[ID] * double(fL[iw] - fR[iw]);
      };
      R wm[NW], w0[NW], wp[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        wm[iw] = sp[iw][ty + NG][tx + NG - 1];
        w0[iw] = sp[iw][ty + NG][tx + NG];
        wp[iw] = sp[iw][ty + NG][tx + NG + 1];
      }
      hancock_dim(std::integral_constant<int, 0>(), wm, w0, wp);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        wm[iw] = sp[iw][ty + NG - 1][tx + NG];
        wp[iw] = sp[iw][ty + NG + 1][tx + NG];
      }
      hancock_dim(std::integral_constant<int, 1>(), wm, w0, wp);
      if constexpr (NDIM == 3)
        hancock_dim(std::integral_constant<int, 2>(), W[NG - 2], W[NG - 1], W[NG]);
    }

    // ---- sources: addsource2 -> mhd_add_source (src/mod_source.t:130-160,
    //      src/mhd/mod_mhd_phys.t:4565-4695) -----------------------------
    if constexpr (P::GLM) {
      // add_source_glm src/mhd/mod_mhd_phys.t:5472-5531
      constexpr int S = P::PSI;
      acc[S] = a.glm_decay * acc[S];
      if (fp.glm_extended) {
        const int cx = tx + NG, cy = ty + NG;
        double gp[3], bn[3];
        gp[0] = ((double(sp[S][cy][cx + 1]) - double(sp[S][cy][cx - 1])) * 0.5) / fp.dx[0];
        gp[1] = ((double(sp[S][cy + 1][cx]) - double(sp[S][cy - 1][cx])) * 0.5) / fp.dx[1];
        double divb = 0.0;
        divb = divb + 0.5 * (double(sp[P::MAG][cy][cx + 1]) - double(sp[P::MAG][cy][cx - 1])) * fp.dxinv[0];
        divb = divb + 0.5 * (double(sp[P::MAG + 1][cy + 1][cx]) - double(sp[P::MAG + 1][cy - 1][cx])) * fp.dxinv[1];
        if constexpr (NDIM == 3) {
          const int c = NG - 1;
          gp[2] = ((double(W[c + 1][S]) - double(W[c - 1][S])) * 0.5) / fp.dx[2];
          divb = divb + 0.5 * (double(W[c + 1][P::MAG + 2]) - double(W[c - 1][P::MAG + 2])) * fp.dxinv[2];
        }
#pragma unroll
        for (int d = 0; d < P::NDIR; ++d) bn[d] = double(sp[P::MAG + d][cy][cx]);
#pragma unroll
        for (int d = 0; d < NDIM; ++d) acc[P::E] = acc[P::E] - a.qdt * bn[d] * gp[d];
#pragma unroll
        for (int d = 0; d < P::NDIR; ++d) acc[P::MOM + d] = acc[P::MOM + d] - a.qdt * bn[d] * divb;
      }
    }
    if constexpr (P::LINDE) {
      // add_source_linde src/mhd/mod_mhd_phys.t:5599-5664 (2-D, one level)
      const int cx = tx + NG, cy = ty + NG;
      auto divb_at = [&](int x, int y) {
        double d = 0.0;
        d = d + 0.5 * (double(sp[P::MAG][y][x + 1]) - double(sp[P::MAG][y][x - 1])) * fp.dxinv[0];
        d = d + 0.5 * (double(sp[P::MAG + 1][y + 1][x]) - double(sp[P::MAG + 1][y - 1][x])) * fp.dxinv[1];
        return d;
      };
      const double dxp = divb_at(cx + 1, cy), dxm = divb_at(cx - 1, cy);
      const double dyp = divb_at(cx, cy + 1), dym = divb_at(cx, cy - 1);
      double g = ((dxp - dxm) * 0.5) / fp.dx[0];
      g = g * fp.divbdiff / fp.invds2;
      acc[P::MAG] = acc[P::MAG] + g;
      acc[P::E] = acc[P::E] + double(sp[P::MAG][cy][cx]) * g;
      g = ((dyp - dym) * 0.5) / fp.dx[1];
      g = g * fp.divbdiff / fp.invds2;
      acc[P::MAG + 1] = acc[P::MAG + 1] + g;
      acc[P::E] = acc[P::E] + double(sp[P::MAG + 1][cy][cx]) * g;
    }

    {
      const size_t o = boff + ((size_t)k * BX + ty) * BX + tx;
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) a.dst[o + iw * cells] = acc[iw];
    }
    __syncthreads();  // LDS plane is rewritten next iteration
  }
  if (crash) *a.crash = 1;
}

}  // namespace FV_NS
