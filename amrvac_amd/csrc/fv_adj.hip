// fv_adj.hip -- adjoint of one finite-volume stage on a uniform periodic
// level, hydrodynamics or ideal MHD without divergence cleaning (the reverse
// of the stage kernel's flux divergence):
//   lam_in = coef * lam_out + ca * add_a + cb * add_b
//            + sum_faces (d fC_f / d u_in)^T mu_f,
//   mu_f = (-qdt/dx_d) (lam_out(left of f) - lam_out(right of f)).
//
// adj_plane_kernel: one workgroup per (block, z plane).  The plane's
// primitive states with a 3-cell x/y halo and lam_out with a 2-cell halo
// are staged in LDS (each cell converted once); every x and y face whose
// 4-cell stencil reaches the tile (BX+3 faces per row) is recomputed and
// reversed once, and its four primitive adjoints are added to per-cell LDS
// sums in four passes (one per stencil position; within a pass the faces
// write distinct cells), so the summation order is fixed and the result
// deterministic without atomics.  The summed primitive adjoint is carried to
// the conserved state once per cell.  In 3-D the z faces are reversed
// beforehand by adj_zface_kernel (each lane one column of a ZC-cell chunk),
// which leaves each cell's summed z-face contribution in a global buffer in
// the face precision; the per-cell LDS sums start from it.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fv_adjoint.h"
#include "fv_launch.h"

namespace fv {

using fvc::AdjArgs;

constexpr int AH = 3;             // primitive halo of the tile
constexpr int AW = BX + 2 * AH;   // staged row length
constexpr int LH = 2;             // lam_out halo of the tile
constexpr int LW = BX + 2 * LH;
constexpr int NF = BX + 3;        // faces per row touching the tile: f in [-1, BX+1]
constexpr int NFACE = NF * BX;    // per dimension and plane
constexpr int ZC = 16;            // cells per lane and chunk in adj_zface_kernel

// lam_out at an arbitrary (possibly neighbouring) cell of plane z
template <int NW, int NDIM>
__device__ __forceinline__ void load_lam(const SweepArgs& a, const double* __restrict__ lam, int b, int x, int y,
                                         int z, double* l) {
  double sign;
  const int sb = locate<NDIM>(a, b, x, y, z, sign);
  const size_t cells = (size_t)BX * BX * a.bnz;
  const size_t off = (size_t)sb * NW * cells + ((size_t)z * BX + y) * BX + x;
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) l[iw] = lam[off + iw * cells];
}

// z faces of one column per lane, for a chunk of ZC cells: every face whose
// 4-cell stencil reaches a cell of the chunk (faces k0-1 .. k0+ZC+1, the
// face k lying between cells k-1 and k) is reversed with a sliding window of
// four primitive states and two lam_out values, and its four contributions
// are added to a window of four running per-cell sums.  A cell's sum is
// complete after its fourth face (faces in increasing z, so the order is
// fixed) and is written once: gz[iw * ncells + cell], in the face precision.
// (the FP32-face kernels are held to 2 waves per SIMD, the launch's one
// round; the FP64 ones may take the whole register file)
template <class P, class R, int RIEM, int LIM>
__global__ void __launch_bounds__(NTHREADS, (std::is_same<R, float>::value ? 2 : 1))
    adj_zface_kernel(SweepArgs a, FVParams fp, AdjArgs g) {
  constexpr int NW = P::NW, NDIM = P::NDIM;
  static_assert(NDIM == 3, "z faces exist in 3-D");
  const int nch = a.bnz / ZC;
  const int b = blockIdx.x / nch, k0 = (blockIdx.x % nch) * ZC;
  if (b >= a.nblocks) return;
  const int tx = threadIdx.x & (BX - 1), ty = threadIdx.x >> 4;
  const size_t cells = (size_t)BX * BX * a.bnz;
  const size_t ntot = (size_t)a.nblocks * cells;
  R* gz = static_cast<R*>(g.gz);
  // st[s] = cell k-2+s of the face k; ll/lr = lam_out(k-1), (k), in the
  // face precision as the plane kernel's tile holds them; sum[m] = running
  // sum of cell k-2+m, in the face precision it is stored in
  R st[4][NW];
  R ll[NW], lr[NW];
  R sum[4][NW];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) sum[m][iw] = R(0);
#pragma unroll
  for (int s = 0; s < 3; ++s) load_prim<P, R, NDIM>(a, fp, b, tx, ty, k0 - 3 + s, st[s + 1]);
  auto load_lam_r = [&](int z, R* l) {
    double t[NW];
    load_lam<NW, NDIM>(a, g.lam_out, b, tx, ty, z, t);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) l[iw] = R(t[iw]);
  };
  load_lam_r(k0 - 2, lr);
  for (int k = k0 - 1; k <= k0 + ZC + 1; ++k) {
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      st[0][iw] = st[1][iw];
      st[1][iw] = st[2][iw];
      st[2][iw] = st[3][iw];
      ll[iw] = lr[iw];
    }
    load_prim<P, R, NDIM>(a, fp, b, tx, ty, k + 1, st[3]);
    load_lam_r(k, lr);
    R mu[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) mu[iw] = R(a.dq[2] * (double(ll[iw]) - double(lr[iw])));
    face_adj_prim<P, R, RIEM, 2, LIM, true, R>(st, fp, R(a.ch), mu, sum);
    // cell k-2 has all four of its faces
    if (k - 2 >= k0) {
      const size_t o = (size_t)b * cells + ((size_t)(k - 2) * BX + ty) * BX + tx;
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) gz[(size_t)iw * ntot + o] = sum[0][iw];
    }
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      sum[0][iw] = sum[1][iw];
      sum[1][iw] = sum[2][iw];
      sum[2][iw] = sum[3][iw];
      sum[3][iw] = R(0);
    }
  }
}

// Reverse of the extended GLM sources of a stage (add_source_glm with
// mhd_glm_extended, src/mhd/mod_mhd_phys.t:5472-5531, as the stage kernel
// adds them to the updated cell from the stage input's primitive state):
//   e   -= qdt sum_d B_d (psi(+e_d) - psi(-e_d)) / (2 dx_d),
//   m_k -= qdt B_k divB,  divB = sum_d (B_d(+e_d) - B_d(-e_d)) / (2 dx_d).
// They are linear in B and psi (conserved = primitive), so the own cell's
// B and the neighbours' B and psi collect, for cell c and its neighbours
// p = c + e_d, m = c - e_d:
//   lam[B_d](c)  += -qdt g_d(c) lam_e(c) - qdt divB(c) lam_m_d(c)
//                   + qdt/(2 dx_d) sum_k (B_k lam_m_k)(p) - (B_k lam_m_k)(m),
//   lam[psi](c)  += qdt/(2 dx_d) sum_d (B_d lam_e)(p) - (B_d lam_e)(m).
// x / y neighbours come from the workgroup's tiles, z neighbours from the
// planes above and below.  The psi decay of the stage is a diagonal factor
// on the psi row that the caller applies to lam_out beforehand.
template <class P, class R>
__device__ __forceinline__ void glm_source_rev(const SweepArgs& a, const FVParams& fp, const AdjArgs& g,
                                               const R (*Wl)[AW * AW], const R (*Ll)[LW * LW], const double* lo,
                                               int b, int k, int tx, int ty, double* acc) {
  constexpr int NDIM = P::NDIM, ND = P::NDIR, E = P::E, MAG = P::MAG, MOM = P::MOM, S = P::PSI;
  const double qdt = a.qdt;
  // B_0..ND-1, psi and lam_e, lam_m_0..ND-1 at the own cell's neighbours
  // n = 0: -e_d, 1: +e_d per dimension d
  double Bn[3][2][3], psin[3][2], le[3][2], lm[3][2][3];
#pragma unroll
  for (int d = 0; d < 2; ++d)
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const int sx = d == 0 ? 2 * n - 1 : 0, sy = d == 1 ? 2 * n - 1 : 0;
      const int wi = (ty + AH + sy) * AW + tx + AH + sx, li = (ty + LH + sy) * LW + tx + LH + sx;
#pragma unroll
      for (int q = 0; q < ND; ++q) {
        Bn[d][n][q] = double(Wl[MAG + q][wi]);
        lm[d][n][q] = double(Ll[MOM + q][li]);
      }
      psin[d][n] = double(Wl[S][wi]);
      le[d][n] = double(Ll[E][li]);
    }
  if constexpr (NDIM == 3) {
    // the planes above and below: only the field and psi (their primitive
    // and conserved values are the same, rounded to R like the tiles') and
    // the energy and momentum adjoints are read
    const size_t cells = (size_t)BX * BX * a.bnz;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      int x = tx, y = ty, z = k + 2 * n - 1;
      double sign;
      const int sb = locate<NDIM>(a, b, x, y, z, sign);
      const size_t off = (size_t)sb * P::NW * cells + ((size_t)z * BX + y) * BX + x;
#pragma unroll
      for (int q = 0; q < ND; ++q) {
        Bn[2][n][q] = double(R(sign * a.src[off + (MAG + q) * cells]));
        lm[2][n][q] = g.lam_out[off + (MOM + q) * cells];
      }
      psin[2][n] = double(R(sign * a.src[off + S * cells]));
      le[2][n] = g.lam_out[off + E * cells];
    }
  }
  double gp[3], divb = 0.0;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    gp[d] = ((psin[d][1] - psin[d][0]) * 0.5) / fp.dx[d];
    divb = divb + 0.5 * (Bn[d][1][d] - Bn[d][0][d]) * fp.dxinv[d];
  }
#pragma unroll
  for (int d = 0; d < NDIM; ++d) acc[MAG + d] = acc[MAG + d] - qdt * gp[d] * lo[E];
#pragma unroll
  for (int q = 0; q < ND; ++q) acc[MAG + q] = acc[MAG + q] - qdt * divb * lo[MOM + q];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) {
    const double h = 0.5 * qdt / fp.dx[d];
    acc[S] = acc[S] + h * (Bn[d][1][d] * le[d][1] - Bn[d][0][d] * le[d][0]);
    double sp = 0.0, sm = 0.0;
#pragma unroll
    for (int q = 0; q < ND; ++q) {
      sp = sp + Bn[d][1][q] * lm[d][1][q];
      sm = sm + Bn[d][0][q] * lm[d][0][q];
    }
    acc[MAG + d] = acc[MAG + d] + 0.5 * qdt * fp.dxinv[d] * (sp - sm);
  }
}

// occupancy: the FP32-face kernels keep to 3 workgroups per CU (GLM's
// reverse then spills 48 B per lane, still 1.1x faster than at 2), except
// those of cada3 (held to 3, its reverse with GLM spills 312 B per lane)
template <class P, class R, int LIM>
constexpr int kAdjWaves = std::is_same<R, float>::value ? (LIM == LIM_CADA3 ? 2 : 3)
                                                        : (P::NW > 5 ? 1 : 2);

template <class P, class R, int RIEM, int LIM>
__global__ void __launch_bounds__(NTHREADS, (kAdjWaves<P, R, LIM>))
    adj_plane_kernel(SweepArgs a, FVParams fp, AdjArgs g) {
  constexpr int NW = P::NW, NDIM = P::NDIM;
  // lam_out is kept in the face precision: mu is rounded to R anyway
  __shared__ R Wl[NW][AW * AW];
  __shared__ R Ll[NW][LW * LW];
  // per-cell sums of the faces' primitive adjoints, filled in four passes
  // per face round (stencil position s = 0..3): within a pass the faces of
  // a round write distinct cells, so the order of the sums is fixed
  __shared__ double Acc[NW][BX * BX];
  const int b = blockIdx.x / a.bnz, k = blockIdx.x % a.bnz;
  if (b >= a.nblocks) return;
  const int t = threadIdx.x;
  const int tx = t & (BX - 1), ty = t >> 4;
  const size_t cells = (size_t)BX * BX * a.bnz;
  const size_t o = (size_t)b * NW * cells + ((size_t)k * BX + ty) * BX + tx;

  // lam_out and the z-face sum of the own cell are loaded first, so that
  // their latency overlaps the staging; the per-cell sums start from the
  // z-face contributions (3-D)
  double lo[NW];
  R zs[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    lo[iw] = g.lam_out[o + iw * cells];
    if constexpr (NDIM == 3) {
      const size_t ntot = (size_t)a.nblocks * cells;
      zs[iw] = static_cast<const R*>(g.gz)[(size_t)iw * ntot + (size_t)b * cells + (size_t)k * BX * BX + t];
    } else {
      zs[iw] = R(0);
    }
  }
  for (int i = t; i < AW * AW; i += NTHREADS) {
    const int x = i % AW - AH, y = i / AW - AH;
    if ((x < 0 || x >= BX) && (y < 0 || y >= BX)) continue;
    R w[NW];
    load_prim<P, R, NDIM>(a, fp, b, x, y, k, w);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) Wl[iw][i] = w[iw];
  }
  for (int i = t; i < LW * LW; i += NTHREADS) {
    const int x = i % LW - LH, y = i / LW - LH;
    if ((x < 0 || x >= BX) && (y < 0 || y >= BX)) continue;
    double l[NW];
    load_lam<NW, NDIM>(a, g.lam_out, b, x, y, k, l);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) Ll[iw][i] = R(l[iw]);
  }
  double acc[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    double v = g.coef * lo[iw];
    if (g.add_a) v = v + g.ca * g.add_a[o + iw * cells];
    if (g.add_b) v = v + g.cb * g.add_b[o + iw * cells];
    acc[iw] = v;
    Acc[iw][t] = double(zs[iw]);
  }
  __syncthreads();

  // Faces of dimension D: x face j = (f, row) with f = j % NF - 1, row =
  // j / NF; y face j = (col, f) with col = j % BX, f = j / BX - 1.  A face's
  // primitive adjoints go to the cells at stencil positions s = 0..3.
  auto reverse_face = [&](auto Dc, int j, double (*wb)[NW]) {
    constexpr int D = decltype(Dc)::value;
    R st[4][NW];
    R mu[NW];
    if constexpr (D == 0) {
      const int f = j % NF - 1, row = j / NF;
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) st[s][iw] = Wl[iw][(row + AH) * AW + f - 2 + s + AH];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw)
        mu[iw] = R(a.dq[0] * (double(Ll[iw][(row + LH) * LW + f - 1 + LH]) - double(Ll[iw][(row + LH) * LW + f + LH])));
    } else {
      const int col = j % BX, f = j / BX - 1;
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) st[s][iw] = Wl[iw][(f - 2 + s + AH) * AW + col + AH];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw)
        mu[iw] = R(a.dq[1] * (double(Ll[iw][(f - 1 + LH) * LW + col + LH]) - double(Ll[iw][(f + LH) * LW + col + LH])));
    }
    face_adj_prim<P, R, RIEM, D, LIM>(st, fp, R(a.ch), mu, wb);
  };
  // the cell of stencil position s of face j (-1: outside the tile)
  auto cell_of = [&](auto Dc, int j, int s) {
    constexpr int D = decltype(Dc)::value;
    if constexpr (D == 0) {
      const int x = j % NF - 3 + s, row = j / NF;
      return (x >= 0 && x < BX) ? row * BX + x : -1;
    } else {
      const int y = j / BX - 3 + s, col = j % BX;
      return (y >= 0 && y < BX) ? y * BX + col : -1;
    }
  };
  // four passes over the stencil positions: within a pass the faces of one
  // dimension and round write distinct cells, so the order of every cell's
  // sum is fixed
  auto scatter = [&](auto Dc, int j, bool on, double (*wb)[NW]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      __syncthreads();
      const int c = on ? cell_of(Dc, j, s) : -1;
      if (c >= 0)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) Acc[iw][c] += wb[s][iw];
    }
  };
  constexpr auto DX = std::integral_constant<int, 0>();
  constexpr auto DY = std::integral_constant<int, 1>();
  static_assert(NFACE > NTHREADS && NFACE - NTHREADS <= 64, "one full round and a partial wave per dimension");
  // Rounds: the first NTHREADS x faces, the first NTHREADS y faces, then
  // the remaining NFACE - NTHREADS faces of both dimensions together, the
  // x ones in wave rx and the y ones in wave ry (a different wave, varying
  // over the planes), so that no workgroup spends a round on one partial
  // wave per dimension and no wave takes both code paths.
  {
    double wb[4][NW];
    reverse_face(DX, t, wb);
    scatter(DX, t, true, wb);
    reverse_face(DY, t, wb);
    scatter(DY, t, true, wb);
  }
  {
    const int rx = k & 3, ry = (k + 2) & 3;
    const int wave = t >> 6, lane = t & 63;
    const int j = NTHREADS + lane;
    const bool onx = wave == rx && j < NFACE, ony = wave == ry && j < NFACE;
    double wb[4][NW];
    if (onx) reverse_face(DX, j, wb);
    if (ony) reverse_face(DY, j, wb);
    scatter(DX, j, onx, wb);
    scatter(DY, j, ony, wb);
  }
  __syncthreads();
  double wsum[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) wsum[iw] = Acc[iw][t];

  // ---- to the conserved state; lam_in -------------------------------------
  double cons[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) cons[iw] = a.src[o + iw * cells];
  prim_rev<P>(cons, wsum, fp, acc);
  if constexpr (P::GLM) {
    if (fp.glm_extended) glm_source_rev<P, R>(a, fp, g, Wl, Ll, lo, b, k, tx, ty, acc);
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) g.lam_in[o + iw * cells] = acc[iw];
}

template <class P, class R, int RIEM, int LIM>
hipError_t launch_adj(const SweepArgs& a, const FVParams& fp, const AdjArgs& g, hipStream_t s) {
  const dim3 grid(a.nblocks * a.bnz);
  if constexpr (P::NDIM == 3) {
    if (a.bnz % ZC != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((adj_zface_kernel<P, R, RIEM, LIM>), dim3(a.nblocks * (a.bnz / ZC)), dim3(NTHREADS), 0, s, a,
                       fp, g);
  }
  hipLaunchKernelGGL((adj_plane_kernel<P, R, RIEM, LIM>), grid, dim3(NTHREADS), 0, s, a, fp, g);
  return hipGetLastError();
}

template <class P, class R, int RIEM>
hipError_t adj_lim(int lim, const SweepArgs& a, const FVParams& fp, const AdjArgs& g, hipStream_t s) {
  switch (lim) {
    case LIM_VANLEER: return launch_adj<P, R, RIEM, LIM_VANLEER>(a, fp, g, s);
    case LIM_MINMOD: return launch_adj<P, R, RIEM, LIM_MINMOD>(a, fp, g, s);
    case LIM_CADA3: return launch_adj<P, R, RIEM, LIM_CADA3>(a, fp, g, s);
    default: return hipErrorInvalidValue;
  }
}

template <class P, class R>
hipError_t adj_riem(int riem, int lim, const SweepArgs& a, const FVParams& fp, const AdjArgs& g, hipStream_t s) {
  switch (riem) {
    case RIEM_HLL: return adj_lim<P, R, RIEM_HLL>(lim, a, fp, g, s);
    case RIEM_HLLC:
      if constexpr (P::MHD) return hipErrorInvalidValue;
      else return adj_lim<P, R, RIEM_HLLC>(lim, a, fp, g, s);
    case RIEM_TVDLF: return adj_lim<P, R, RIEM_TVDLF>(lim, a, fp, g, s);
    default: return hipErrorInvalidValue;
  }
}

template <class P>
hipError_t adj_prec(int mixed, int riem, int lim, const SweepArgs& a, const FVParams& fp, const AdjArgs& g,
                    hipStream_t s) {
  if (P::NDIM == 3 && !g.gz) return hipErrorInvalidValue;
  return mixed ? adj_riem<P, float>(riem, lim, a, fp, g, s) : adj_riem<P, double>(riem, lim, a, fp, g, s);
}

// hydrodynamics (physics 1) and ideal MHD (physics 2, ndir 3) without
// divergence control or with GLM (divb 1), 2-D / 3-D
hipError_t launch_adjoint_stage(int ndim, int physics, int divb, int mixed, int riem, int lim, const SweepArgs& a,
                                const FVParams& fp, const AdjArgs& g, hipStream_t s) {
  if (physics == 1) {
    if (ndim == 2) return adj_prec<Phys<2, 2, false, DIVB_NONE>>(mixed, riem, lim, a, fp, g, s);
    if (ndim == 3) return adj_prec<Phys<3, 3, false, DIVB_NONE>>(mixed, riem, lim, a, fp, g, s);
  } else if (physics == 2 && divb == DIVB_GLM) {
    if (ndim == 2) return adj_prec<Phys<2, 3, true, DIVB_GLM>>(mixed, riem, lim, a, fp, g, s);
    if (ndim == 3) return adj_prec<Phys<3, 3, true, DIVB_GLM>>(mixed, riem, lim, a, fp, g, s);
  } else if (physics == 2 && divb == DIVB_NONE) {
    if (ndim == 2) return adj_prec<Phys<2, 3, true, DIVB_NONE>>(mixed, riem, lim, a, fp, g, s);
    if (ndim == 3) return adj_prec<Phys<3, 3, true, DIVB_NONE>>(mixed, riem, lim, a, fp, g, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fv
