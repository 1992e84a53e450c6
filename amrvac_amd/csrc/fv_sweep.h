// fv_sweep.h -- the fused finite-volume stage kernel.
//
// One workgroup updates one block of the level; a lane owns one (x, y)
// column of the block (BX x BX lanes: 256 for 16^2 blocks, rounded up to
// whole waves for 8^2 and 10^2).  The block's stencil data never round-trips through HBM ghost
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
constexpr int BX = 16;  // block_nx(1) = block_nx(2) of the fastest kernels
constexpr int NTHREADS = 256;

struct SweepArgs {
  const double* src;   // psa (stage input, wCT)
  const double* base;  // RK base slot (comb_mode 1, 2)
  double* dst;         // psb
  const int* nbr;      // [nb][3^ndim]
  int nblocks;
  int bx;              // block_nx(1) = block_nx(2)
  int bnz;             // block_nx(3) (1 in 2-D)
  int per_xcd;         // ceil(nblocks/8)
  int comb_mode;       // 0: init = dst, 1: init = base, 2: init = a*base + b*src,
                       // 3: init = src (base is the stage input itself)
  double comb_a, comb_b;
  double dq[3];        // -qdt/dx_d
  double qdt;
  double ch;           // cmax_global (GLM)
  double glm_decay;    // exp(-qdt*ch*alpha/min dx)
  int bc[6];           // physical boundary type per side
  int* crash;
  // block-structured AMR (all nullptr / unused on a uniform level)
  const int* blevel;        // level-1 of each block: cell size dx/2^blevel
  double glm_decay_lev[8];  // psi decay factor per level
  double* freg;             // scaled block-boundary face fluxes (refluxing)
  const int* fregoff;       // [nb][2*ndim] register offset of a face, -1: none
  int nface;                // transverse cells of a face (register stride)
  int linde;                // Linde's source follows the stage (linde_kernel)
  // CFL partials of the output (uniform level): per block the max over
  // cells of sum_d cmax_d/dx_d and of cmax, as getdt_courant needs them;
  // nullptr: not wanted
  double* dt_part;
};

constexpr int MAXLEV = 8;

}  // namespace fvc

namespace FV_NS {
using fvc::BX;
using fvc::NTHREADS;
using fvc::SweepArgs;

// lanes of the stage kernel for BX x BX blocks (whole waves)
template <int BXc>
constexpr int kLanes = (BXc * BXc + 63) / 64 * 64;

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
// (BXc = 0: the block size a.bx of the arguments; else the compile-time one)
template <int NDIM, int BXc = 0>
__device__ __forceinline__ int locate(const SweepArgs& a, int b, int& x, int& y, int& z, double& sign) {
  int c[3] = {x, y, z};
  const int bx = BXc ? BXc : a.bx;
  const int n[3] = {bx, bx, a.bnz};
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
__device__ __forceinline__ void load_cons(const double* __restrict__ w, int b, int x, int y, int z, int bx,
                                          int bnz, double* c) {
  const size_t cells = (size_t)bx * bx * bnz;
  const size_t off = (size_t)b * P::NW * cells + ((size_t)z * bx + y) * bx + x;
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) c[iw] = w[off + iw * cells];
}

// conserved state of an arbitrary (possibly halo) cell, with the sign of an
// antisymmetric boundary applied
template <class P, int NDIM, int BXc = 0>
__device__ __forceinline__ void load_cons_at(const SweepArgs& a, int b, int x, int y, int z, double* c) {
  double sign;
  const int sb = locate<NDIM, BXc>(a, b, x, y, z, sign);
  load_cons<P>(a.src, sb, x, y, z, BXc ? BXc : a.bx, a.bnz, c);
  if (sign < 0.0) {
#pragma unroll
    for (int iw = 0; iw < P::NW; ++iw) c[iw] = -c[iw];
  }
}

// primitive state (FP64 conversion, rounded to R) of conserved state c
template <class P, class R>
__device__ __forceinline__ void prim_of(const double* c, R* out, const FVParams& fp) {
  double w[P::NW];
  to_prim<P>(c, w, fp);
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) out[iw] = R(w[iw]);
}

// primitive state of an arbitrary (possibly halo) cell
template <class P, class R, int NDIM, int BXc = 0>
__device__ __forceinline__ void load_prim(const SweepArgs& a, const FVParams& fp, int b, int x, int y,
                                          int z, R* w) {
  double c[P::NW];
  load_cons_at<P, NDIM, BXc>(a, b, x, y, z, c);
  prim_of<P, R>(c, w, fp);
}

// The RK base and the output are touched once per launch: they stream
// through L2 (non-temporal) so that they do not displace the stage input,
// whose planes are re-read as halos by neighbouring workgroups.
__device__ __forceinline__ double ld_stream(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(double v, double* p) { __builtin_nontemporal_store(v, p); }

// Cell t of the four NG-deep halo strips of a plane (x and y block-local,
// outside the block on one side), t < 4*NG*BX; then the 4 NG x NG corners.
template <int NG, int BX>
__device__ __forceinline__ void strip_cell(int t, int& x, int& y) {
  constexpr int NSTRIP = 4 * NG * BX;
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
}

// Occupancy: FP64 kernels run 2 waves per SIMD (LDS and registers), FP32
// kernels are held to 3 (at most 168 VGPRs).
template <class R>
constexpr int kMinWaves = std::is_same<R, float>::value ? 3 : 2;

// Scheme constants of a block's level: every length scales by 2^-lev
// (dx(level) = dx(level-1)/2, src/mod_initialize.t:101), so the products
// below are exact and match a level-wise evaluation bit for bit.
__device__ __forceinline__ void level_params(const SweepArgs& a, int b, FVParams& fp, double* dq,
                                             double& glm_decay) {
  dq[0] = a.dq[0]; dq[1] = a.dq[1]; dq[2] = a.dq[2];
  glm_decay = a.glm_decay;
  if (!a.blevel) return;
  const int lev = a.blevel[b];
  if (lev == 0) return;
  const double sc = double(1 << lev), sc2 = sc * sc;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    fp.dx[d] = fp.dx[d] / sc;
    fp.dxinv[d] = fp.dxinv[d] * sc;
    fp.rdelinv[d] = fp.rdelinv[d] * sc2;
    dq[d] = dq[d] * sc;
  }
  fp.invds2 = fp.invds2 * sc2;
  fp.mindx = fp.mindx / sc;
  glm_decay = a.glm_decay_lev[lev];
}

// AMR: level-dependent scheme constants and the refluxing registers; the
// uniform-level kernels are compiled without them (their registers would
// spill in the 3-D MHD kernels).
template <class P, class R, int RIEM, int LIM, int NG, bool AMR, int BXc>
__global__ void __launch_bounds__(kLanes<BXc>, kMinWaves<R>)
sweep_kernel(SweepArgs a, FVParams fp0) {
  constexpr int BX = BXc;                 // block_nx(1) = block_nx(2)
  constexpr int NTHREADS = kLanes<BXc>;   // lanes of the workgroup
  constexpr bool FULL = BX * BX == NTHREADS;
  constexpr int NDIM = P::NDIM;
  constexpr int NW = P::NW;
  constexpr int PX = BX + 2 * NG;
  constexpr bool HANCOCK = RIEM == RIEM_HANCOCK;
  // Any extra runtime branch in the z loop of the specialised kernels (a
  // compile-time TVD limiter on a uniform level of 16^2 cross-sections)
  // costs the FP64 form about 12% (measured): they carry no small-value
  // checks (runs with fix_small_values take the runtime-limiter kernel),
  // and only the FP32-reconstruction form, where the fused CFL partials cost
  // 2% and save the separate pass over ps, computes them.
  constexpr bool SPECIAL = LIM != 0 && LIM != LIM_WENO5 && !AMR && BXc == 16;
  constexpr bool SMALLCHK = !SPECIAL;
  constexpr bool CFL = !AMR && BXc == 16 && std::is_same<R, float>::value;
  constexpr int NSTRIP = 4 * NG * BX;
  constexpr int NHALO = NSTRIP;
  // 3-D, FP32 reconstruction: the halo strips of the NG planes ahead are
  // loaded together with the window's lookahead plane and kept in an LDS ring
  // (the face neighbours read the same planes at the same time, so the loads
  // hit L2).  The FP64 kernel keeps to 2 workgroups per CU within 160 KB of
  // LDS by loading the halo of the current plane instead.
  constexpr bool HRING = NDIM == 3 && std::is_same<R, float>::value;

  __shared__ R sp[NW][PX][PX];          // primitive plane with halo
  __shared__ R fxs[HANCOCK ? 1 : NW][BX][BX + 1];  // x faces: fxs[iw][y][f]
  __shared__ R fys[HANCOCK ? 1 : NW][BX + 1][BX];  // y faces: fys[iw][f][x]
  __shared__ R hring[HRING ? NG : 1][HRING ? NW : 1][HRING ? NSTRIP : 1];

  const int b = wg_block(a.per_xcd);
  if (b >= a.nblocks) return;
  FVParams fp = fp0;
  double dq[3], glm_decay;
  // refluxing registers of the block's faces (x-, x+, y-, y+, z-, z+)
  int freg[2 * NDIM];
  if constexpr (AMR) {
    level_params(a, b, fp, dq, glm_decay);
#pragma unroll
    for (int f = 0; f < 2 * NDIM; ++f) freg[f] = a.fregoff ? a.fregoff[b * 2 * NDIM + f] : -1;
  } else {
    dq[0] = a.dq[0]; dq[1] = a.dq[1]; dq[2] = a.dq[2];
    glm_decay = a.glm_decay;
  }
  const int tid = threadIdx.x;
  // lane -> column; the lanes past BX*BX (8^2 and 10^2 blocks) repeat
  // column 0's arithmetic and store nothing of it
  const bool own = FULL || tid < BX * BX;
  const int cid = own ? tid : 0;
  const int tx = cid % BX;
  const int ty = cid / BX;
  const R ch = R(a.ch);
  int crash = 0;
  double dt_maxtot = -1.0e99, dt_cmax = 0.0;  // CFL partials of this lane's column
  const size_t cells = (size_t)BX * BX * a.bnz;
  const size_t boff = (size_t)b * NW * cells;

  // z window (3-D): Wp[m] = primitive plane k-NG+1+m (m < NG, the current
  // plane last), Wc[m] = conserved plane k+1+m as stored (FP64).  The current
  // plane's conserved state is the RK input of comb modes 2 and 3, so the
  // block's own planes are read from HBM once.
  constexpr int NWP = NDIM == 3 ? NG : 1;
  R Wp[NWP][NW];
  double Wc[NWP][NW];
  R fz_carry[NW];
  // primitive planes k-NG+1 .. k+NG of this lane's column
  auto zstencil = [&](R (*Z)[NW]) {
#pragma unroll
    for (int m = 0; m < NG; ++m)
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) Z[m][iw] = Wp[m][iw];
#pragma unroll
    for (int m = 0; m < NG; ++m) prim_of<P, R>(Wc[m], Z[NG + m], fp);
  };
  if constexpr (NDIM == 3) {
    // prologue: planes -NG .. NG-1, halo ring of planes 0 .. NG-1, carried
    // face -1/2
#pragma unroll
    for (int m = 0; m < NG; ++m) load_prim<P, R, NDIM, BX>(a, fp, b, tx, ty, m - NG, Wp[m]);
#pragma unroll
    for (int m = 0; m < NG; ++m) load_cons_at<P, NDIM, BX>(a, b, tx, ty, m, Wc[m]);
    if constexpr (HRING) {
      for (int p = 0; p < NG && p < a.bnz; ++p)
        for (int t = tid; t < NSTRIP; t += NTHREADS) {
          int x, y;
          strip_cell<NG, BX>(t, x, y);
          R w[NW];
          load_prim<P, R, NDIM, BX>(a, fp, b, x, y, p, w);
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) hring[p][iw][t] = w[iw];
        }
    }
    if constexpr (!HANCOCK) {
      R Z[2 * NG][NW];
      zstencil(Z);
      face_flux<P, R, RIEM, 2, LIM, NG>(Z, fp, ch, fz_carry, crash);
    }
    if constexpr (HRING) __syncthreads();
  }

  const int nz = NDIM == 3 ? a.bnz : 1;
  for (int k = 0; k < nz; ++k) {
    const size_t o = boff + ((size_t)k * BX + ty) * BX + tx;
    double cons_k[NW];  // wCT of this lane's cell
    double acc[NW];
    // RK base (src/mod_advance.t:251-267), streamed; the combination with
    // the stage input is formed once the base has arrived
    auto rk_init = [&]() {
      if (a.comb_mode == 0) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = ld_stream(a.dst + o + iw * cells);
      } else if (a.comb_mode == 1) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = ld_stream(a.base + o + iw * cells);
      } else if (a.comb_mode == 2) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = a.comb_a * ld_stream(a.base + o + iw * cells) + a.comb_b * cons_k[iw];
      } else {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) acc[iw] = cons_k[iw];
      }
    };
    // ---- stage plane k (primitive, with x/y halo) in LDS -----------------
    if constexpr (NDIM == 3) {
      // advance the window and issue the lookahead plane k+NG
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) cons_k[iw] = Wc[0][iw];
      // fix_small_values: phys_to_primitive checks every cell it converts
      // (src/hd/mod_hd_phys.t:713-715, src/mhd/mod_mhd_phys.t:1750-1753);
      // over the level, the blocks' own cells are every converted cell
      if (SMALLCHK && fp.fix_small && small_cons<P>(cons_k, fp)) crash = 1;
#pragma unroll
      for (int m = 0; m < NG - 1; ++m)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) Wp[m][iw] = Wp[m + 1][iw];
      prim_of<P, R>(cons_k, Wp[NG - 1], fp);
#pragma unroll
      for (int m = 0; m < NG - 1; ++m)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) Wc[m][iw] = Wc[m + 1][iw];
      load_cons_at<P, NDIM, BX>(a, b, tx, ty, k + NG, Wc[NG - 1]);
      if (own)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) sp[iw][ty + NG][tx + NG] = Wp[NG - 1][iw];
    } else {
      load_cons<P>(a.src, b, tx, ty, 0, BX, 1, cons_k);
      if (SMALLCHK && fp.fix_small && small_cons<P>(cons_k, fp)) crash = 1;
      R w[NW];
      prim_of<P, R>(cons_k, w, fp);
      if (own)
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) sp[iw][ty + NG][tx + NG] = w[iw];
    }
    if constexpr (HRING) {
      const int slot = k % NG;
      for (int t = tid; t < NSTRIP; t += NTHREADS) {
        int x, y;
        strip_cell<NG, BX>(t, x, y);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) sp[iw][y + NG][x + NG] = hring[slot][iw][t];
      }
    } else {
      // halo strips of plane k (FP64: loaded by alternate wave pairs from
      // plane to plane)
      const int th = std::is_same<R, double>::value && NTHREADS == 256 ? (tid + 128 * (k & 1)) % NTHREADS : tid;
      for (int t = th; t < NHALO; t += NTHREADS) {
        int x, y;
        strip_cell<NG, BX>(t, x, y);
        R w[NW];
        load_prim<P, R, NDIM, BX>(a, fp, b, x, y, k, w);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) sp[iw][y + NG][x + NG] = w[iw];
      }
    }
    __syncthreads();
    if constexpr (HRING) {
      // halo of plane k+NG into the ring slot just copied out
      if (k + NG < nz) {
        const int slot = k % NG;
        // by the upper waves: the lower ones take the partial face pass
        const int th = NTHREADS == 256 ? (tid + 128) % NTHREADS : tid;
        for (int t = th; t < NSTRIP; t += NTHREADS) {
          int x, y;
          strip_cell<NG, BX>(t, x, y);
          R w[NW];
          load_prim<P, R, NDIM, BX>(a, fp, b, x, y, k + NG, w);
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) hring[slot][iw][t] = w[iw];
        }
      }
    }

    // z neighbours of this lane's cell for the sources and the Hancock
    // predictor: primitive planes k-1 and k+1
    R zm[NW], zp[NW];
    if constexpr (NDIM == 3) {
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) zm[iw] = Wp[NG - 2][iw];
    }

    if constexpr (!HANCOCK) {
      // ---- face fluxes of this plane --------------------------------------
      constexpr int NXF = BX * (BX + 1);             // x faces (272 for 16^2)
      constexpr int XSLOTS = (NXF + 63) / 64 * 64;   // padded to a wave multiple
      constexpr int NSLOT = XSLOTS + NXF;            // + the y faces
      // FP64: the waves that take the third, partial pass change from plane
      // to plane, so that the co-resident workgroups' heavy waves spread
      // over the SIMDs (measured 1.6% faster; the FP32 form is not helped)
      const int t0 = std::is_same<R, double>::value ? (tid + 64 * (k & 3)) % NTHREADS : tid;
      for (int t = t0; t < NSLOT; t += NTHREADS) {
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
      rk_init();
      // z face k+1/2 last, so that the lookahead load has landed
      R fz_new[NW];
      if constexpr (NDIM == 3) {
        R Z[2 * NG][NW];
        zstencil(Z);
        face_flux<P, R, RIEM, 2, LIM, NG>(Z, fp, ch, fz_new, crash);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) zp[iw] = Z[NG][iw];
      }
      __syncthreads();
      // ---- store_flux (src/mod_fix_conserve.t:597-648): the scaled fluxes
      //      of the block-boundary faces, for refluxing --------------------
      if (AMR && a.freg && own) {
        const size_t nf = (size_t)a.nface;
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) {
          if (tx == 0 && freg[0] >= 0)
            a.freg[freg[0] + iw * nf + ty + BX * k] = dq[0] * double(fxs[iw][ty][0]);
          if (tx == BX - 1 && freg[1] >= 0)
            a.freg[freg[1] + iw * nf + ty + BX * k] = dq[0] * double(fxs[iw][ty][BX]);
          if (ty == 0 && freg[2] >= 0)
            a.freg[freg[2] + iw * nf + tx + BX * k] = dq[1] * double(fys[iw][0][tx]);
          if (ty == BX - 1 && freg[3] >= 0)
            a.freg[freg[3] + iw * nf + tx + BX * k] = dq[1] * double(fys[iw][BX][tx]);
          if constexpr (NDIM == 3) {
            if (k == 0 && freg[4] >= 0) a.freg[freg[4] + iw * nf + tx + BX * ty] = dq[2] * double(fz_carry[iw]);
            if (k == nz - 1 && freg[5] >= 0) a.freg[freg[5] + iw * nf + tx + BX * ty] = dq[2] * double(fz_new[iw]);
          }
        }
      }
      // ---- flux divergence, dimension by dimension (:269-284) ------------
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        acc[iw] = acc[iw] + dq[0] * double(fxs[iw][ty][tx + 1]) - dq[0] * double(fxs[iw][ty][tx]);
        acc[iw] = acc[iw] + dq[1] * double(fys[iw][ty + 1][tx]) - dq[1] * double(fys[iw][ty][tx]);
        if constexpr (NDIM == 3) {
          acc[iw] = acc[iw] + dq[2] * double(fz_new[iw]) - dq[2] * double(fz_carry[iw]);
          fz_carry[iw] = fz_new[iw];
        }
      }
    } else {
      // ---- Hancock predictor (src/mod_finite_volume.t:19-113) -------------
      rk_init();
      if constexpr (NDIM == 3) prim_of<P, R>(Wc[0], zp, fp);
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
        for (int iw = 0; iw < NW; ++iw) acc[iw] = acc[iw] + dq[ID] * double(fL[iw] - fR[iw]);
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
      if constexpr (NDIM == 3) hancock_dim(std::integral_constant<int, 2>(), zm, w0, zp);
    }

    // ---- unphysical values after the flux update (finite_volume
    //      :325-333, hancock :109-111: small_values_method 'error' sets
    //      crash, after which the run stops).  The reconstructed face
    //      states of the TVD limiters lie between their cells' values, so
    //      the face checks of reconstruct_LR (:1213-1216) flag nothing more
    // (hancock checks only after its sources, GLM here or Linde's kernel)
    if constexpr (!(HANCOCK && P::GLM))
      if (!(HANCOCK && a.linde) && SMALLCHK && fp.fix_small && small_cons<P>(acc, fp)) crash = 1;
    // ---- sources: addsource2 -> mhd_add_source (src/mod_source.t:130-160,
    //      src/mhd/mod_mhd_phys.t:4565-4695) -----------------------------
    if constexpr (P::GLM) {
      // add_source_glm src/mhd/mod_mhd_phys.t:5472-5531
      constexpr int S = P::PSI;
      acc[S] = glm_decay * acc[S];
      if (fp.glm_extended) {
        const int cx = tx + NG, cy = ty + NG;
        double gp[3], bn[3];
        gp[0] = ((double(sp[S][cy][cx + 1]) - double(sp[S][cy][cx - 1])) * 0.5) / fp.dx[0];
        gp[1] = ((double(sp[S][cy + 1][cx]) - double(sp[S][cy - 1][cx])) * 0.5) / fp.dx[1];
        double divb = 0.0;
        divb = divb + 0.5 * (double(sp[P::MAG][cy][cx + 1]) - double(sp[P::MAG][cy][cx - 1])) * fp.dxinv[0];
        divb = divb + 0.5 * (double(sp[P::MAG + 1][cy + 1][cx]) - double(sp[P::MAG + 1][cy - 1][cx])) * fp.dxinv[1];
        if constexpr (NDIM == 3) {
          gp[2] = ((double(zp[S]) - double(zm[S])) * 0.5) / fp.dx[2];
          divb = divb + 0.5 * (double(zp[P::MAG + 2]) - double(zm[P::MAG + 2])) * fp.dxinv[2];
        }
#pragma unroll
        for (int d = 0; d < P::NDIR; ++d) bn[d] = double(sp[P::MAG + d][cy][cx]);
#pragma unroll
        for (int d = 0; d < NDIM; ++d) acc[P::E] = acc[P::E] - a.qdt * bn[d] * gp[d];
#pragma unroll
        for (int d = 0; d < P::NDIR; ++d) acc[P::MOM + d] = acc[P::MOM + d] - a.qdt * bn[d] * divb;
      }
    }
    // Linde's source (add_source_linde) follows in linde_kernel: div B at
    // the neighbours needs the field beyond this kernel's planes and rim
    // the GLM source checks again (src/mhd/mod_mhd_phys.t:5529)
    if constexpr (P::GLM)
      if (SMALLCHK && fp.fix_small && small_cons<P>(acc, fp)) crash = 1;
    if (own)
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) st_stream(acc[iw], a.dst + o + iw * cells);
    if constexpr (CFL)
      if (a.dt_part) cfl_cell<P>(acc, fp, dt_maxtot, dt_cmax);
    __syncthreads();  // LDS plane is rewritten next iteration
  }
  if (crash) *a.crash = 1;
  if constexpr (CFL) {
    if (a.dt_part) {
      // block max: wave64 shuffles, then the 4 waves through LDS
      for (int off = 32; off > 0; off >>= 1) {
        dt_maxtot = fmax(dt_maxtot, __shfl_xor(dt_maxtot, off, 64));
        dt_cmax = fmax(dt_cmax, __shfl_xor(dt_cmax, off, 64));
      }
      __shared__ double dtr[2][NTHREADS / 64];
      if ((tid & 63) == 0) { dtr[0][tid >> 6] = dt_maxtot; dtr[1][tid >> 6] = dt_cmax; }
      __syncthreads();
      if (tid == 0) {
        double m = dtr[0][0], c = dtr[1][0];
        for (int i = 1; i < NTHREADS / 64; ++i) { m = fmax(m, dtr[0][i]); c = fmax(c, dtr[1][i]); }
        a.dt_part[2 * b] = m;
        a.dt_part[2 * b + 1] = c;
      }
    }
  }
}

}  // namespace FV_NS
