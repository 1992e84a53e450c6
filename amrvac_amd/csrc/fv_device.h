// fv_device.h -- device-side physics, limiters and Riemann solvers of the
// finite-volume block update, templated on the physics (HD / ideal MHD with
// GLM or Linde div B control), the direction of the face (compile-time, so
// every per-variable array stays in registers) and the working precision R
// of reconstruction / fluxes / wave speeds.
//
// Each routine restates the reference routine cited beside it; operand
// order follows the reference's Fortran so that FP64 results agree with the
// reference to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef FV_EXACT
#define FV_NS fvx
#else
#define FV_NS fv
#endif

namespace fvc {
// scalar parameters of the scheme (kernel argument, by value)
struct FVParams {
  double gamma, gamma_1, inv_gamma_1;
  double glm_alpha;
  int glm_extended;
  double divbdiff;
  int limiter;
  int boundspeed;
  double tvdlfeps;
  double small_pressure, small_density;
  int check_small;
  int fix_small;       // fix_small_values (small_values_method 'error')
  double dx[3];
  double dxinv[3];     // 1/dx
  double rdelinv[3];   // 1/(cada3_radius*dx)^2, src/limiter/mod_limiter.t:254
  double invds2;       // sum_d 1/dx_d^2 (Linde)
  double mindx;        // minval(dxlevel) (GLM decay)
};

}  // namespace fvc

namespace FV_NS {
using fvc::FVParams;

// Arithmetic helpers.  The contraction-free build (FV_EXACT) keeps IEEE
// division / square root and the reference's MIN/MAX semantics, so that every
// operation rounds like the CPU restatement; the production builds use the
// hardware min/max and, in FP32, the single-instruction reciprocal and square
// root (about 1 ulp).  In FP64 they refine the hardware reciprocal and
// reciprocal square root (about 23 correct bits) by Newton / Goldschmidt
// steps to within an ulp or two, in half the instructions of the IEEE
// sequences; arguments are finite and positive (densities, sound speeds,
// wave-speed spans), or zero for the square root.
#ifdef FV_EXACT
constexpr bool kExact = true;
#else
constexpr bool kExact = false;
#endif

template <class R>
__device__ __forceinline__ R mn2(R a, R b) {
  if constexpr (kExact) return a < b ? a : b;
  else return fmin(a, b);
}
template <class R>
__device__ __forceinline__ R mx2(R a, R b) {
  if constexpr (kExact) return a > b ? a : b;
  else return fmax(a, b);
}
__device__ __forceinline__ double frcp(double a) {
  if constexpr (kExact) {
    return 1.0 / a;
  } else {
    double r = __builtin_amdgcn_rcp(a);
    double e = fma(-a, r, 1.0);  // relative error of r, about 2^-23
    r = fma(r, e, r);            // about 2^-46
    e = fma(-a, r, 1.0);
    return fma(r, e, r);
  }
}
__device__ __forceinline__ float frcp(float a) {
  if constexpr (kExact) return 1.0f / a;
  else return __builtin_amdgcn_rcpf(a);
}
__device__ __forceinline__ double fdiv(double a, double b) {
  if constexpr (kExact) return a / b;
  else return a * frcp(b);
}
__device__ __forceinline__ float fdiv(float a, float b) {
  if constexpr (kExact) return a / b;
  else return a * __builtin_amdgcn_rcpf(b);
}
__device__ __forceinline__ double fsqrt(double a) {
  if constexpr (kExact) {
    return sqrt(a);
  } else {
    // Goldschmidt: g -> sqrt(a), h -> 1/(2 sqrt(a)) from y = rsq(a)
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, a);
    g = fma(d, h, g);
    return a == 0.0 ? 0.0 : g;
  }
}
__device__ __forceinline__ float fsqrt(float a) {
  if constexpr (kExact) return sqrtf(a);
  else return __builtin_amdgcn_sqrtf(a);
}

// ---------------------------------------------------------------------------
// physics traits: variable indices as in hd_phys_init / mhd_phys_init
// (rho_, mom(:), e_, mag(:), psi_; src/hd/mod_hd_phys.t:217-228,
// src/mhd/mod_mhd_phys.t:478-503)
// ---------------------------------------------------------------------------
enum { DIVB_NONE = 0, DIVB_GLM = 1, DIVB_LINDE = 4 };
enum { RIEM_HLL = 1, RIEM_HLLC = 2, RIEM_HLLD = 3, RIEM_TVDLF = 5, RIEM_HANCOCK = 8 };
enum {
  LIM_MINMOD = 1, LIM_WOODWARD = 2, LIM_MCBETA = 3, LIM_SUPERBEE = 4,
  LIM_VANLEER = 5, LIM_ALBADA = 6, LIM_KOREN = 7, LIM_CADA = 8, LIM_CADA3 = 9, LIM_WENO5 = 16
};

template <int NDIM_, int NDIR_, bool MHD_, int DIVB_>
struct Phys {
  static constexpr int NDIM = NDIM_;
  static constexpr int NDIR = NDIR_;
  static constexpr bool MHD = MHD_;
  static constexpr bool GLM = MHD_ && DIVB_ == DIVB_GLM;
  static constexpr bool LINDE = MHD_ && DIVB_ == DIVB_LINDE;
  static constexpr int RHO = 0;
  static constexpr int MOM = 1;
  static constexpr int E = NDIR_ + 1;
  static constexpr int MAG = NDIR_ + 2;
  static constexpr int PSI = 2 * NDIR_ + 2;
  static constexpr int NW = MHD_ ? (2 * NDIR_ + 2 + (GLM ? 1 : 0)) : (NDIR_ + 2);
};

// ---------------------------------------------------------------------------
// conversions
// ---------------------------------------------------------------------------

// hd_to_primitive src/hd/mod_hd_phys.t:703-734,
// mhd_to_primitive_origin src/mhd/mod_mhd_phys.t:1741-1765 (always FP64:
// the pressure recovery from total energy cancels at low beta)
template <class P>
__device__ __forceinline__ void to_prim(const double* c, double* w, const FVParams& fp) {
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) w[iw] = c[iw];
  const double inv_rho = 1.0 / w[P::RHO];
#pragma unroll
  for (int d = 0; d < P::NDIR; ++d) w[P::MOM + d] = w[P::MOM + d] * inv_rho;
  double vsq = w[P::MOM] * w[P::MOM];
#pragma unroll
  for (int d = 1; d < P::NDIR; ++d) vsq = vsq + w[P::MOM + d] * w[P::MOM + d];
  if constexpr (!P::MHD) {
    w[P::E] = (fp.gamma - 1.0) * (w[P::E] - 0.5 * w[P::RHO] * vsq);
  } else {
    double bsq = w[P::MAG] * w[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bsq = bsq + w[P::MAG + d] * w[P::MAG + d];
    w[P::E] = fp.gamma_1 * (w[P::E] - 0.5 * (w[P::RHO] * vsq + bsq));
  }
}

// phys_check_w on a conserved FP64 cell (hd_check_w src/hd/mod_hd_phys.t:
// 648-674, mhd_check_w_origin src/mhd/mod_mhd_phys.t:1452-1473 with small_e
// = small_pressure/(gamma-1)).  With fix_small_values and the default
// small_values_method 'error' a flagged cell sets crash
// (src/mod_small_values.t:25-87).
template <class P>
__device__ __forceinline__ bool small_cons(const double* c, const FVParams& fp) {
  if (c[P::RHO] < fp.small_density) return true;
  double msq = c[P::MOM] * c[P::MOM];
#pragma unroll
  for (int d = 1; d < P::NDIR; ++d) msq = msq + c[P::MOM + d] * c[P::MOM + d];
  if constexpr (!P::MHD) {
    return (fp.gamma - 1.0) * (c[P::E] - 0.5 * msq / c[P::RHO]) < fp.small_pressure;
  } else {
    double bsq = c[P::MAG] * c[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bsq = bsq + c[P::MAG + d] * c[P::MAG + d];
    return c[P::E] - 0.5 * (msq / c[P::RHO] + bsq) < fp.small_pressure * fp.inv_gamma_1;
  }
}

// hd_to_conserved src/hd/mod_hd_phys.t:677-700,
// mhd_to_conserved_origin src/mhd/mod_mhd_phys.t:1562-1579
template <class P, class R>
__device__ __forceinline__ void to_cons(R* w, R inv_gamma_1) {
  R vsq = w[P::MOM] * w[P::MOM];
#pragma unroll
  for (int d = 1; d < P::NDIR; ++d) vsq = vsq + w[P::MOM + d] * w[P::MOM + d];
  if constexpr (!P::MHD) {
    w[P::E] = w[P::E] * inv_gamma_1 + R(0.5) * vsq * w[P::RHO];
  } else {
    R bsq = w[P::MAG] * w[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bsq = bsq + w[P::MAG + d] * w[P::MAG + d];
    w[P::E] = w[P::E] * inv_gamma_1 + R(0.5) * (vsq * w[P::RHO] + bsq);
  }
#pragma unroll
  for (int d = 0; d < P::NDIR; ++d) w[P::MOM + d] = w[P::RHO] * w[P::MOM + d];
}

// hd_get_flux src/hd/mod_hd_phys.t:1097-1146,
// mhd_get_flux src/mhd/mod_mhd_phys.t:3810-3889 (GLM :3868-3874)
template <class P, class R, int ID>
__device__ __forceinline__ void phys_flux(const R* wC, const R* w, R ch, R* f) {
  const R vn = w[P::MOM + ID];
  if constexpr (!P::MHD) {
    f[P::RHO] = vn * w[P::RHO];
#pragma unroll
    for (int d = 0; d < P::NDIR; ++d) f[P::MOM + d] = vn * wC[P::MOM + d];
    f[P::MOM + ID] = f[P::MOM + ID] + w[P::E];
    f[P::E] = vn * (wC[P::E] + w[P::E]);
  } else {
    const R bn = w[P::MAG + ID];
    f[P::RHO] = vn * w[P::RHO];
#pragma unroll
    for (int d = 0; d < P::NDIR; ++d) f[P::MOM + d] = wC[P::MOM + ID] * w[P::MOM + d] - bn * w[P::MAG + d];
    R bsq = w[P::MAG] * w[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bsq = bsq + w[P::MAG + d] * w[P::MAG + d];
    const R ptotal = w[P::E] + R(0.5) * bsq;
    f[P::MOM + ID] = f[P::MOM + ID] + ptotal;
    R bv = w[P::MAG] * w[P::MOM];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bv = bv + w[P::MAG + d] * w[P::MOM + d];
    f[P::E] = vn * (wC[P::E] + ptotal) - bn * bv;
#pragma unroll
    for (int d = 0; d < P::NDIR; ++d) f[P::MAG + d] = vn * w[P::MAG + d] - bn * w[P::MOM + d];
    if constexpr (P::GLM) {
      f[P::MAG + ID] = w[P::PSI];
      f[P::PSI] = ch * ch * w[P::MAG + ID];
    }
  }
}

// mhd_get_csound_prim src/mhd/mod_mhd_phys.t:3329-3379 (no B0, no Hall)
template <class P, class R, int ID>
__device__ __forceinline__ R fast_speed(const R* w, R gamma) {
  const R inv_rho = frcp(w[P::RHO]);
  const R cs = gamma * w[P::E] * inv_rho;
  R b2 = w[P::MAG] * w[P::MAG];
#pragma unroll
  for (int d = 1; d < P::NDIR; ++d) b2 = b2 + w[P::MAG + d] * w[P::MAG + d];
  const R cfast2 = b2 * inv_rho + cs;
  const R bn = w[P::MAG + ID];
  R av = cfast2 * cfast2 - R(4) * cs * (bn * bn) * inv_rho;
  if (av < R(0)) av = R(0);
  return fsqrt(R(0.5) * (cfast2 + fsqrt(av)));
}

// hd_get_cbounds src/hd/mod_hd_phys.t:891-1001,
// mhd_get_cbounds src/mhd/mod_mhd_phys.t:3037-3127
template <class P, class R, int ID, bool WANT_CMIN>
__device__ __forceinline__ void cbounds(const R* wLC, const R* wRC, const R* wLp, const R* wRp,
                                        const FVParams& fp, R& cmax, R& cmin, int& crash) {
  const R g = R(fp.gamma);
  constexpr int MN = P::MOM + ID;
  if constexpr (!P::MHD) {
    if (fp.boundspeed == 1) {
      const R tmp1 = fsqrt(wLp[P::RHO]);
      const R tmp2 = fsqrt(wRp[P::RHO]);
      const R tmp3 = frcp(tmp1 + tmp2);
      const R umean = (wLp[MN] * tmp1 + wRp[MN] * tmp2) * tmp3;
      const R csL = fdiv(g * wLp[P::E], wLp[P::RHO]);
      const R csR = fdiv(g * wRp[P::E], wRp[P::RHO]);
      const R dv = wRp[MN] - wLp[MN];
      R dmean = (tmp1 * csL + tmp2 * csR) * tmp3 + R(0.5) * tmp1 * tmp2 * (tmp3 * tmp3) * (dv * dv);
      dmean = fsqrt(dmean);
      if constexpr (WANT_CMIN) {
        cmin = umean - dmean;
        cmax = umean + dmean;
      } else {
        cmax = fabs(umean) + dmean;
      }
    } else if (fp.boundspeed == 2) {
      // conserved mean, hd_get_csound2 -> hd_get_pthermal (:1021-1064)
      const R rho = R(0.5) * (wLC[P::RHO] + wRC[P::RHO]);
      const R mn = R(0.5) * (wLC[MN] + wRC[MN]);
      R msq = R(0);
#pragma unroll
      for (int d = 0; d < P::NDIR; ++d) {
        const R m = R(0.5) * (wLC[P::MOM + d] + wRC[P::MOM + d]);
        msq = msq + m * m;
      }
      const R e = R(0.5) * (wLC[P::E] + wRC[P::E]);
      const R tmp1 = fdiv(mn, rho);
      R pth = (g - R(1)) * (e - fdiv(R(0.5) * msq, rho));
      if (fp.fix_small) {
        if (pth < R(fp.small_pressure)) pth = R(fp.small_pressure);
      } else if (fp.check_small && pth < R(fp.small_pressure)) {
        crash = 1;
      }
      const R cs = fsqrt(fdiv(g * pth, rho));
      if constexpr (WANT_CMIN) {
        cmax = mx2(tmp1 + cs, R(0));
        cmin = mn2(tmp1 - cs, R(0));
      } else {
        cmax = fabs(tmp1) + cs;
      }
    } else {
      const R csL = fdiv(g * wLp[P::E], wLp[P::RHO]);
      const R csR = fdiv(g * wRp[P::E], wRp[P::RHO]);
      const R cs = mx2(fsqrt(csL), fsqrt(csR));
      if constexpr (WANT_CMIN) cmin = mn2(wLp[MN], wRp[MN]) - cs;
      cmax = mx2(wLp[MN], wRp[MN]) + cs;
    }
  } else {
    if (fp.boundspeed == 1) {
      const R csL = fast_speed<P, R, ID>(wLp, g);
      const R csR = fast_speed<P, R, ID>(wRp, g);
      const R tmp1 = fsqrt(wLp[P::RHO]);
      const R tmp2 = fsqrt(wRp[P::RHO]);
      const R tmp3 = frcp(tmp1 + tmp2);
      const R umean = (wLp[MN] * tmp1 + wRp[MN] * tmp2) * tmp3;
      const R dv = wRp[MN] - wLp[MN];
      const R dmean = fsqrt((tmp1 * (csL * csL) + tmp2 * (csR * csR)) * tmp3 +
                           R(0.5) * tmp1 * tmp2 * (tmp3 * tmp3) * (dv * dv));
      if constexpr (WANT_CMIN) {
        cmin = umean - dmean;
        cmax = umean + dmean;
      } else {
        cmax = fabs(umean) + dmean;
      }
    } else if (fp.boundspeed == 2) {
      R wm[P::NW];
#pragma unroll
      for (int iw = 0; iw < P::NW; ++iw) wm[iw] = R(0.5) * (wLp[iw] + wRp[iw]);
      const R cs = fast_speed<P, R, ID>(wm, g);
      if constexpr (WANT_CMIN) {
        cmax = mx2(wm[MN] + cs, R(0));
        cmin = mn2(wm[MN] - cs, R(0));
      } else {
        cmax = fabs(wm[MN]) + cs;
      }
    } else {
      const R csL = fast_speed<P, R, ID>(wLp, g);
      const R csR = fast_speed<P, R, ID>(wRp, g);
      const R cs = mx2(csL, csR);
      if constexpr (WANT_CMIN) cmin = mn2(wLp[MN], wRp[MN]) - cs;
      cmax = mx2(wLp[MN], wRp[MN]) + cs;
    }
  }
}

// ---------------------------------------------------------------------------
// slope limiters, dwlimiter2 src/limiter/mod_limiter.t:128-334 for one cell:
// a = dwC(i-1), b = dwC(i); gives ldw(i) and rdw(i)
// ---------------------------------------------------------------------------
// the limited third-order slope factor of cada (src/limiter/mod_limiter.t:
// 231-252) for theta = th
template <class R>
__device__ __forceinline__ R cada_limited(R th) {
  const R cadalfa = R(0.5), cadbeta = R(2), cadgamma = R(1.6);
  const R A = (R(2) + th) * R(1.0 / 3.0);
  return mx2(R(0), mn2(A, mx2(-cadalfa * th, mn2(mn2(cadbeta * th, A), cadgamma))));
}

template <class R>
__device__ __forceinline__ void cada3_side(R th, R eta, R& out) {
  const R A = (R(2) + th) * R(1.0 / 3.0);
  const R lim = cada_limited(th);
  if (eta <= R(1) - R(1.0e-14)) {
    out = A;
  } else if (eta >= R(1) + R(1.0e-14)) {
    out = lim;
  } else {
    const R t2 = (eta - R(1)) * R(1.0e14);
    out = R(0.5) * ((R(1) - t2) * A + (R(1) + t2) * lim);
  }
}

template <class R, int LIM>
__device__ __forceinline__ void dwlimiter(int lim_rt, R a, R b, R rdelinv, R& ldw, R& rdw) {
  const int lim = LIM > 0 ? LIM : lim_rt;
  if (lim == LIM_MINMOD) {
    R t = copysign(R(1), b);
    t = t * mx2(R(0), mn2(fabs(b), t * a));
    ldw = t; rdw = t;
  } else if (lim == LIM_VANLEER) {
    const R t = fdiv(R(2) * mx2(a * b, R(0)), b + a + R(1.0e-12));
    ldw = t; rdw = t;
  } else if (lim == LIM_CADA3) {
    const R eps = R(1.4901161193847656e-08);  // sqrt(epsilon(1d0))
    const R eta = (b * b + a * a) * rdelinv;
    R l, r;
    cada3_side(fdiv(a, b + copysign(eps, b)), eta, l);
    cada3_side(fdiv(b, a + copysign(eps, a)), eta, r);
    ldw = l * b;
    rdw = r * a;
  } else if (lim == LIM_CADA) {
    const R eps = R(1.4901161193847656e-08);  // sqrt(epsilon(1d0))
    ldw = cada_limited(fdiv(a, b + copysign(eps, b))) * b;
    rdw = cada_limited(fdiv(b, a + copysign(eps, a))) * a;
  } else if (lim == LIM_KOREN) {
    const R t = copysign(R(1), b);
    const R t2 = mn2(R(2) * fabs(b), R(2) * t * a);
    ldw = t * mx2(R(0), mn2(t2, (a * t + R(2) * fabs(b)) * R(1.0 / 3.0)));
    rdw = t * mx2(R(0), mn2(t2, (R(2) * a * t + fabs(b)) * R(1.0 / 3.0)));
  } else if (lim == LIM_WOODWARD) {
    R t = copysign(R(1), b);
    t = R(2) * t * mx2(R(0), mn2(mn2(fabs(b), t * a), t * R(0.25) * (a + b)));
    ldw = t; rdw = t;
  } else if (lim == LIM_MCBETA) {
    R t = copysign(R(1), b);
    t = t * mx2(R(0), mn2(mn2(R(1.4) * fabs(b), R(1.4) * t * a), t * R(0.5) * (a + b)));
    ldw = t; rdw = t;
  } else if (lim == LIM_SUPERBEE) {
    R t = copysign(R(1), b);
    t = t * mx2(mx2(R(0), mn2(R(2) * fabs(b), t * a)), mn2(fabs(b), R(2) * t * a));
    ldw = t; rdw = t;
  } else {  // albada
    const R t = fdiv(a * (b * b + R(1.0e-12)) + b * (a * a + R(1.0e-12)), b * b + a * a + R(2.0e-12));
    ldw = t; rdw = t;
  }
}

// Limited slopes of one face: ldw(i) from (dwC(i-1), dwC(i)) and rdw(i+1)
// from (dwC(i), dwC(i+1)).  For cada3 both quotients share the denominator
// dwC(i) + sign(eps, dwC(i)); outside the bit-reproducible build it is
// inverted once.
template <class R, int LIM>
__device__ __forceinline__ void face_limit(int lim_rt, R dm, R d0, R dp, R rdelinv, R& ldw, R& rdw) {
  const int lim = LIM > 0 ? LIM : lim_rt;
  if (!kExact && lim == LIM_CADA3) {
    const R eps = R(1.4901161193847656e-08);
    const R inv = frcp(d0 + copysign(eps, d0));
    R l, r;
    cada3_side(dm * inv, (d0 * d0 + dm * dm) * rdelinv, l);
    cada3_side(dp * inv, (dp * dp + d0 * d0) * rdelinv, r);
    ldw = l * d0;
    rdw = r * d0;
    return;
  }
  R r2, l2;
  dwlimiter<R, LIM>(lim_rt, dm, d0, rdelinv, ldw, r2);
  dwlimiter<R, LIM>(lim_rt, d0, dp, rdelinv, l2, rdw);
}

// WENO5limiter var=1 src/limiter/mod_weno.t:254-377, one variable, face
// between cells 0 and 1 of the stencil s[-2..3] (passed as s0..s5 =
// w(i-2)..w(i+3))
template <class R>
__device__ __forceinline__ void weno5(R wmm, R wm, R w0, R wp, R wpp, R wppp, R& wl, R& wr) {
  const R bc1 = R(1.0833333333333333), bc2 = R(0.25);
  const R d1 = R(1.0 / 10.0), d2 = R(3.0 / 5.0), d3 = R(3.0 / 10.0);
  const R u11 = R(1.0 / 3.0), u12 = R(-7.0 / 6.0), u13 = R(11.0 / 6.0);
  const R u21 = R(-1.0 / 6.0), u22 = R(5.0 / 6.0), u23 = R(1.0 / 3.0);
  const R u31 = R(1.0 / 3.0), u32 = R(5.0 / 6.0), u33 = R(-1.0 / 6.0);
  const R weps = R(1.0e-18);
  {
    const R f1 = u11 * wmm + u12 * wm + u13 * w0;
    const R f2 = u21 * wm + u22 * w0 + u23 * wp;
    const R f3 = u31 * w0 + u32 * wp + u33 * wpp;
    R t = wmm + w0 - R(2) * wm, q = wmm - R(4) * wm + R(3) * w0;
    const R b1 = bc1 * (t * t) + bc2 * (q * q);
    t = wm + wp - R(2) * w0; q = wm - wp;
    const R b2 = bc1 * (t * t) + bc2 * (q * q);
    t = w0 + wpp - R(2) * wp; q = R(3) * w0 - R(4) * wp + wpp;
    const R b3 = bc1 * (t * t) + bc2 * (q * q);
    t = b1 + weps; const R a1 = fdiv(d1, t * t);
    t = b2 + weps; const R a2 = fdiv(d2, t * t);
    t = b3 + weps; const R a3 = fdiv(d3, t * t);
    const R as = a1 + a2 + a3;
    wl = fdiv(f1 * a1, as) + fdiv(f2 * a2, as) + fdiv(f3 * a3, as);
  }
  {
    const R f1 = u11 * wppp + u12 * wpp + u13 * wp;
    const R f2 = u21 * wpp + u22 * wp + u23 * w0;
    const R f3 = u31 * wp + u32 * w0 + u33 * wm;
    R t = wppp + wp - R(2) * wpp, q = wppp - R(4) * wpp + R(3) * wp;
    const R b1 = bc1 * (t * t) + bc2 * (q * q);
    t = wpp + w0 - R(2) * wp; q = wpp - w0;
    const R b2 = bc1 * (t * t) + bc2 * (q * q);
    t = wp + wm - R(2) * w0; q = R(3) * wp - R(4) * w0 + wm;
    const R b3 = bc1 * (t * t) + bc2 * (q * q);
    t = b1 + weps; const R a1 = fdiv(d1, t * t);
    t = b2 + weps; const R a2 = fdiv(d2, t * t);
    t = b3 + weps; const R a3 = fdiv(d3, t * t);
    const R as = a1 + a2 + a3;
    wr = fdiv(f1 * a1, as) + fdiv(f2 * a2, as) + fdiv(f3 * a3, as);
  }
}

// get_Riemann_flux_hlld src/mod_finite_volume.t:505-718 (Miyoshi & Kusano
// 2005 with Guo 2016's star pressure; no B0, no stagger grid, ndir = 3):
// the single-star states w1 across the fast waves, the double-star states w2
// across the Alfven waves, then the region of the fan at the face.  wLp/wRp
// are the primitive states after modify_wLR, wLC/wRC the conserved ones.
template <class P, class R, int ID, bool CT = false>
__device__ __forceinline__ void hlld_flux(const R* wLC, const R* wRC, const R* wLp, const R* wRp, const R* fL,
                                          const R* fR, R sR, R sL, R* fC, R bx_ct = R(0)) {
  static_assert(P::MHD && P::NDIR == 3, "HLLD is provided for MHD with three vector components");
  constexpr int NW = P::NW, RHO = P::RHO, E = P::E, MG = P::MAG, MO = P::MOM;
  constexpr int I2 = (ID + 1) % 3, I3 = (ID + 2) % 3;
  constexpr int M1 = MO + ID, M2 = MO + I2, M3 = MO + I3;
  constexpr int B1 = MG + ID, B2 = MG + I2, B3 = MG + I3;
  // HLL estimate of the normal field (Li 2005 eq. 33); with constrained
  // transport the face field (src/mod_finite_volume.t:536-537)
  R Bx;
  if constexpr (CT) Bx = bx_ct;
  else Bx = fdiv(sR * wRC[B1] - sL * wLC[B1], sR - sL);
  const R ptR = wRp[E] + R(0.5) * (wRC[MG] * wRC[MG] + wRC[MG + 1] * wRC[MG + 1] + wRC[MG + 2] * wRC[MG + 2]);
  const R ptL = wLp[E] + R(0.5) * (wLC[MG] * wLC[MG] + wLC[MG + 1] * wLC[MG + 1] + wLC[MG + 2] * wLC[MG + 2]);
  const R vR = wRp[M1], vL = wLp[M1];
  const R suR = (sR - vR) * wRC[RHO];
  const R suL = (sL - vL) * wLC[RHO];
  const R sm = fdiv(suR * vR - suL * vL - ptR + ptL, suR - suL);
  // single-star states (velocities in the momentum slots until the end)
  R w1R[NW], w1L[NW], w2R[NW], w2L[NW];
  w1R[M1] = sm; w1L[M1] = sm; w2R[M1] = sm; w2L[M1] = sm;
  w1R[B1] = Bx; w1L[B1] = Bx;
  w1R[RHO] = fdiv(suR, sR - sm);
  w1L[RHO] = fdiv(suL, sL - sm);
  R r1R = suR * (sR - sm) - Bx * Bx;
  if (r1R != R(0)) r1R = frcp(r1R);
  R r1L = suL * (sL - sm) - Bx * Bx;
  if (r1L != R(0)) r1L = frcp(r1L);
  w1R[M2] = wRp[M2] - Bx * wRC[B2] * (sm - vR) * r1R;
  w1L[M2] = wLp[M2] - Bx * wLC[B2] * (sm - vL) * r1L;
  const R qR = (suR * (sR - vR) - Bx * Bx) * r1R;
  const R qL = (suL * (sL - vL) - Bx * Bx) * r1L;
  w1R[M3] = wRp[M3] - Bx * wRC[B3] * (sm - vR) * r1R;
  w1L[M3] = wLp[M3] - Bx * wLC[B3] * (sm - vL) * r1L;
  w1R[B3] = wRC[B3] * qR;
  w1L[B3] = wLC[B3] * qL;
  w1R[B2] = wRC[B2] * qR;
  w1L[B2] = wLC[B2] * qL;
  // total pressure of the star region (Guo eq. 25), then the energies
  const R pts = suR * (sm - vR) + ptR;
  {
    const R vbR = wRp[MO] * wRC[MG] + wRp[MO + 1] * wRC[MG + 1] + wRp[MO + 2] * wRC[MG + 2];
    const R vb1R = w1R[MO] * w1R[MG] + w1R[MO + 1] * w1R[MG + 1] + w1R[MO + 2] * w1R[MG + 2];
    w1R[E] = fdiv((sR - vR) * wRC[E] - ptR * vR + pts * sm + Bx * (vbR - vb1R), sR - sm);
    const R vbL = wLp[MO] * wLC[MG] + wLp[MO + 1] * wLC[MG + 1] + wLp[MO + 2] * wLC[MG + 2];
    const R vb1L = w1L[MO] * w1L[MG] + w1L[MO + 1] * w1L[MG + 1] + w1L[MO + 2] * w1L[MG + 2];
    w1L[E] = fdiv((sL - vL) * wLC[E] - ptL * vL + pts * sm + Bx * (vbL - vb1L), sL - sm);
  }
  // double-star states
  w2R[RHO] = w1R[RHO]; w2L[RHO] = w1L[RHO];
  w2R[B1] = Bx; w2L[B1] = Bx;
  const R sR1 = fsqrt(w1R[RHO]);
  const R sL1 = fsqrt(w1L[RHO]);
  const R tmp = frcp(sR1 + sL1);
  const R sgn = copysign(R(1), Bx);
  const R s1R = sm + fdiv(fabs(Bx), sR1);
  const R s1L = sm - fdiv(fabs(Bx), sL1);
  w2R[M2] = (sL1 * w1L[M2] + sR1 * w1R[M2] + (w1R[B2] - w1L[B2]) * sgn) * tmp;
  w2L[M2] = w2R[M2];
  w2R[B2] = (sL1 * w1R[B2] + sR1 * w1L[B2] + sL1 * sR1 * (w1R[M2] - w1L[M2]) * sgn) * tmp;
  w2L[B2] = w2R[B2];
  w2R[M3] = (sL1 * w1L[M3] + sR1 * w1R[M3] + (w1R[B3] - w1L[B3]) * sgn) * tmp;
  w2L[M3] = w2R[M3];
  w2R[B3] = (sL1 * w1R[B3] + sR1 * w1L[B3] + sL1 * sR1 * (w1R[M3] - w1L[M3]) * sgn) * tmp;
  w2L[B3] = w2R[B3];
  {
    const R d1R = w1R[MO] * w1R[MG] + w1R[MO + 1] * w1R[MG + 1] + w1R[MO + 2] * w1R[MG + 2];
    const R d2R = w2R[MO] * w2R[MG] + w2R[MO + 1] * w2R[MG + 1] + w2R[MO + 2] * w2R[MG + 2];
    w2R[E] = w1R[E] + sR1 * (d1R - d2R) * sgn;
    const R d1L = w1L[MO] * w1L[MG] + w1L[MO + 1] * w1L[MG + 1] + w1L[MO + 2] * w1L[MG + 2];
    const R d2L = w2L[MO] * w2L[MG] + w2L[MO + 1] * w2L[MG + 1] + w2L[MO + 2] * w2L[MG + 2];
    w2L[E] = w1L[E] - sL1 * (d1L - d2L) * sgn;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    w1R[MO + c] = w1R[MO + c] * w1R[RHO];
    w1L[MO + c] = w1L[MO + c] * w1L[RHO];
    w2R[MO + c] = w2R[MO + c] * w2R[RHO];
    w2L[MO + c] = w2L[MO + c] * w2L[RHO];
  }
  // flux_type: GLM's psi and normal B take the known flux
  // (src/mhd/mod_mhd_phys.t:580-586)
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    if (P::GLM && (iw == P::PSI || iw == B1)) {
      fC[iw] = fL[iw];
    } else if (sL > R(0)) {
      fC[iw] = fL[iw];
    } else if (s1L >= R(0)) {
      fC[iw] = fL[iw] + sL * (w1L[iw] - wLC[iw]);
    } else if (sm >= R(0)) {
      fC[iw] = fL[iw] + sL * (w1L[iw] - wLC[iw]) + s1L * (w2L[iw] - w1L[iw]);
    } else if (s1R >= R(0)) {
      fC[iw] = fR[iw] + sR * (w1R[iw] - wRC[iw]) + s1R * (w2R[iw] - w1R[iw]);
    } else if (sR >= R(0)) {
      fC[iw] = fR[iw] + sR * (w1R[iw] - wRC[iw]);
    } else {
      fC[iw] = fR[iw];
    }
  }
}

// ---------------------------------------------------------------------------
// one face: reconstruct_LR (src/mod_finite_volume.t:1110-1228), modify_wLR
// (GLM, src/mhd/mod_mhd_phys.t:6326-6377), fluxes, bounds and the Riemann
// solver (src/mod_finite_volume.t:348-504, src/hd/mod_hd_hllc.t:56-257).
// st[k][iw] are primitive values of cells i-NG+1+k (face between i and i+1,
// NG = 2: st = w(i-1), w(i), w(i+1), w(i+2); NG = 3 adds w(i-2), w(i+3)).
// ---------------------------------------------------------------------------
// Constrained transport: the face's normal field in (the stage input's face
// field, and that of ps for the HLLD star states), the reconstructed
// velocities and signal speeds out for the edge electric fields.
struct CtIO {
  double bn, bx;
  double vL[3], vR[3];
  double cmax, cmin;
};

template <class P, class R, int RIEM, int ID, int LIM, int NG, bool CT = false>
__device__ __forceinline__ void face_flux(const R (*st)[P::NW], const FVParams& fp, R ch, R* fC,
                                          int& crash, CtIO* ctio = nullptr) {
  constexpr int NW = P::NW;
  constexpr int C0 = NG - 1;  // index of cell i in st
  R wLp[NW], wRp[NW];
  if constexpr (LIM == LIM_WENO5) {
    R tl[NW], tr[NW];
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      wLp[iw] = st[C0][iw];
      wRp[iw] = st[C0 + 1][iw];
      weno5(st[C0 - 2][iw], st[C0 - 1][iw], st[C0][iw], st[C0 + 1][iw], st[C0 + 2][iw], st[C0 + 3][iw],
            tl[iw], tr[iw]);
    }
    // fix_limiter src/limiter/mod_weno.t:100-133 with hd/mhd check_w
    const bool bad = tl[P::RHO] < R(fp.small_density) || tl[P::E] < R(fp.small_pressure) ||
                     tr[P::RHO] < R(fp.small_density) || tr[P::E] < R(fp.small_pressure);
    if (!bad) {
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) { wLp[iw] = tl[iw]; wRp[iw] = tr[iw]; }
    }
  } else {
    const R rdel = R(fp.rdelinv[ID]);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      const R dm = st[C0][iw] - st[C0 - 1][iw];
      const R d0 = st[C0 + 1][iw] - st[C0][iw];
      const R dp = st[C0 + 2][iw] - st[C0 + 1][iw];
      R ldw, rdw;
      face_limit<R, LIM>(fp.limiter, dm, d0, dp, rdel, ldw, rdw);
      wLp[iw] = st[C0][iw] + R(0.5) * ldw;
      wRp[iw] = st[C0 + 1][iw] - R(0.5) * rdw;
    }
  }
  R wLC[NW], wRC[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) { wLC[iw] = wLp[iw]; wRC[iw] = wRp[iw]; }
  const R inv_g1 = R(fp.inv_gamma_1);
  to_cons<P, R>(wLC, inv_g1);
  to_cons<P, R>(wRC, inv_g1);
  if constexpr (P::GLM) {
    constexpr int B = P::MAG + ID, S = P::PSI, E = P::E;
    const R dB = wRp[B] - wLp[B];
    const R dPsi = wRp[S] - wLp[S];
    wLp[B] = R(0.5) * (wRp[B] + wLp[B] - fdiv(dPsi, ch));
    wLp[S] = R(0.5) * (wRp[S] + wLp[S] - dB * ch);
    wRp[B] = wLp[B];
    wRp[S] = wLp[S];
    wRC[E] = wRC[E] - R(0.5) * (wRC[B] * wRC[B]);
    wLC[E] = wLC[E] - R(0.5) * (wLC[B] * wLC[B]);
    wRC[B] = wLp[B]; wRC[S] = wLp[S];
    wLC[B] = wLp[B]; wLC[S] = wLp[S];
    wRC[E] = wRC[E] + R(0.5) * (wRC[B] * wRC[B]);
    wLC[E] = wLC[E] + R(0.5) * (wLC[B] * wLC[B]);
  }
  if constexpr (CT) {
    // mhd_modify_wLR with stagger_grid (src/mhd/mod_mhd_phys.t:6338-6345)
    constexpr int B = P::MAG + ID;
    wLC[B] = R(ctio->bn);
    wRC[B] = wLC[B];
    wLp[B] = wLC[B];
    wRp[B] = wLC[B];
#pragma unroll
    for (int d = 0; d < P::NDIR; ++d) {
      ctio->vL[d] = wLp[P::MOM + d];
      ctio->vR[d] = wRp[P::MOM + d];
    }
  }
  R fL[NW], fR[NW];
  phys_flux<P, R, ID>(wLC, wLp, ch, fL);
  phys_flux<P, R, ID>(wRC, wRp, ch, fR);
  R cmax, cmin = R(0);
  if constexpr (RIEM == RIEM_TVDLF) {
    cbounds<P, R, ID, false>(wLC, wRC, wLp, wRp, fp, cmax, cmin, crash);
    if constexpr (CT) { ctio->cmax = cmax; ctio->cmin = cmin; }
    const R fac = R(-0.5) * R(fp.tvdlfeps) * cmax;
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      fC[iw] = R(0.5) * (fL[iw] + fR[iw]);
      fC[iw] = fC[iw] + fac * (wRC[iw] - wLC[iw]);
    }
  } else {
    cbounds<P, R, ID, true>(wLC, wRC, wLp, wRp, fp, cmax, cmin, crash);
    if constexpr (CT) { ctio->cmax = cmax; ctio->cmin = cmin; }
    // flux_type: normal B gets Lax-Friedrichs dissipation only without GLM
    // (src/mhd/mod_mhd_phys.t:580-588)
    constexpr int BTV = (P::MHD && !P::GLM && P::NDIM > 1) ? (P::MAG + ID) : -1;
    if constexpr (RIEM == RIEM_HLLD) {
      if constexpr (CT) hlld_flux<P, R, ID, true>(wLC, wRC, wLp, wRp, fL, fR, cmax, cmin, fC, R(ctio->bx));
      else hlld_flux<P, R, ID>(wLC, wRC, wLp, wRp, fL, fR, cmax, cmin, fC);
    } else if constexpr (RIEM == RIEM_HLL) {
      const R inv = frcp(cmax - cmin);
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) {
        if (iw == BTV) {
          fC[iw] = -R(fp.tvdlfeps) * R(0.5) * mx2(cmax, R(fabs(cmin))) * (wRC[iw] - wLC[iw]);
        } else if (cmin >= R(0)) {
          fC[iw] = fL[iw];
        } else if (cmax <= R(0)) {
          fC[iw] = fR[iw];
        } else {
          const R num = cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * (wRC[iw] - wLC[iw]);
          if constexpr (kExact) fC[iw] = num / (cmax - cmin);
          else fC[iw] = num * inv;
        }
      }
    } else {  // HLLC: hd_get_lCD / hd_get_wCD or mhd_get_lCD / mhd_get_wCD
      constexpr int RHO = P::RHO, MN = P::MOM + ID, E = P::E;
      int patchf = 1;
      if (cmin >= R(0)) patchf = -2;
      else if (cmax <= R(0)) patchf = 2;
      R lambdaCD = R(0);
      R Fhll[NW];
      // the HLL state of the magnetic field and momentum (MHD contact state)
      R whll[P::MHD ? NW : 1];
      if (patchf == 1) {
        R whll_m = R(0), whll_r = R(0);
        const R inv = frcp(cmax - cmin);
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) {
          const R nf = cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * (wRC[iw] - wLC[iw]);
          const R nw = cmax * wRC[iw] - cmin * wLC[iw] + fL[iw] - fR[iw];
          R wh;
          if constexpr (kExact) {
            Fhll[iw] = nf / (cmax - cmin);
            wh = nw / (cmax - cmin);
          } else {
            Fhll[iw] = nf * inv;
            wh = nw * inv;
          }
          if (iw == MN) whll_m = wh;
          if (iw == RHO) whll_r = wh;
          if constexpr (P::MHD) whll[iw] = wh;
        }
        lambdaCD = fdiv(whll_m, whll_r);
        bool cond = true;
        if (cmin < R(0) && lambdaCD > R(0) && lambdaCD < cmax) {
          patchf = -1;
        } else if (cmax > R(0) && lambdaCD < R(0) && lambdaCD > cmin) {
          patchf = 1;
        } else if (lambdaCD >= cmax || lambdaCD <= cmin) {
          lambdaCD = R(0);
          patchf = 3;
          cond = false;
        }
        if (cond && lambdaCD == R(0)) patchf = (-cmin >= cmax) ? 1 : -1;
      } else {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) Fhll[iw] = R(0);
      }
      if (patchf == 1 || patchf == -1) {
        const bool right = patchf == 1;
        const R cspeed = right ? cmax : cmin;
        R wSub[NW], fSub[NW], wCD[NW];
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) {
          wSub[iw] = right ? wRC[iw] : wLC[iw];
          fSub[iw] = right ? fR[iw] : fL[iw];
        }
        const R vSub = fdiv(wSub[MN], wSub[RHO]);
        if constexpr (P::MHD) {
          // mhd_get_wCD src/mhd/mod_mhd_hllc.t:147-240 (Li 2005)
          constexpr int BN = P::MAG + ID;
          const R den = cspeed - lambdaCD;
          wCD[RHO] = fdiv(wSub[RHO] * (cspeed - vSub), den);
#pragma unroll
          for (int d = 0; d < P::NDIR; ++d) wCD[P::MAG + d] = whll[P::MAG + d];
#pragma unroll
          for (int d = 0; d < P::NDIR; ++d) {
            if (d != ID)
              wCD[P::MOM + d] = fdiv(cspeed * wSub[P::MOM + d] - fSub[P::MOM + d] - wCD[BN] * wCD[P::MAG + d], den);
            else
              wCD[P::MOM + d] = wCD[RHO] * lambdaCD;
          }
          R vb = whll[P::MOM] * whll[P::MAG];
#pragma unroll
          for (int d = 1; d < P::NDIR; ++d) vb = vb + whll[P::MOM + d] * whll[P::MAG + d];
          const R vdotb = fdiv(vb, whll[RHO]);
          const R pCD = wSub[RHO] * (cspeed - vSub) * (lambdaCD - vSub) + fSub[MN] - wSub[MN] * vSub + wCD[BN] * wCD[BN];
          wCD[E] = fdiv(cspeed * wSub[E] - fSub[E] + lambdaCD * pCD - vdotb * wCD[BN], den);
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) {
            if (iw == BN || (P::GLM && iw == P::PSI)) fC[iw] = R(0);
            else fC[iw] = fSub[iw] + cspeed * (wCD[iw] - wSub[iw]);
          }
        } else {
          const R csmls = frcp(cspeed - lambdaCD);
          wCD[RHO] = wSub[RHO] * (cspeed - vSub) * csmls;
#pragma unroll
          for (int d = 0; d < P::NDIR; ++d) {
            if (d != ID) wCD[P::MOM + d] = (cspeed * wSub[P::MOM + d] - fSub[P::MOM + d]) * csmls;
            else wCD[P::MOM + d] = wCD[RHO] * lambdaCD;
          }
          const R pCD = wSub[RHO] * (cspeed - vSub) * (lambdaCD - vSub) + fSub[MN] - wSub[MN] * vSub;
          wCD[E] = (cspeed * wSub[E] - fSub[E] + lambdaCD * pCD) * csmls;
#pragma unroll
          for (int iw = 0; iw < NW; ++iw) fC[iw] = fSub[iw] + cspeed * (wCD[iw] - wSub[iw]);
        }
      } else if (patchf == -2) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) fC[iw] = fL[iw];
      } else if (patchf == 2) {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) fC[iw] = fR[iw];
      } else {
#pragma unroll
        for (int iw = 0; iw < NW; ++iw) fC[iw] = Fhll[iw];
      }
      // flux_type: the normal field takes Lax-Friedrichs dissipation only
      // without GLM (src/mod_finite_volume.t:477-479)
      if constexpr (BTV >= 0)
        fC[BTV] = -R(fp.tvdlfeps) * R(0.5) * mx2(cmax, R(fabs(cmin))) * (wRC[BTV] - wLC[BTV]);
    }
  }
}

// cell-centred cmax for dt: hd_get_cmax src/hd/mod_hd_phys.t:787-811,
// mhd_get_cmax_origin src/mhd/mod_mhd_phys.t:2482-2550 (FP64, primitive
// w), written out without contraction (see cfl_cell)
template <class P, int ID>
__device__ __forceinline__ double cfl_cmax(const double* w, const FVParams& fp) {
#pragma clang fp contract(off)
  if constexpr (!P::MHD) {
    return fabs(w[P::MOM + ID]) + sqrt(fp.gamma * w[P::E] / w[P::RHO]);
  } else {
    // mhd_get_csound_prim, as fast_speed
    const double inv_rho = frcp(w[P::RHO]);
    const double cs = fp.gamma * w[P::E] * inv_rho;
    double b2 = w[P::MAG] * w[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) b2 = b2 + w[P::MAG + d] * w[P::MAG + d];
    const double cfast2 = b2 * inv_rho + cs;
    const double bn = w[P::MAG + ID];
    double av = cfast2 * cfast2 - 4.0 * cs * (bn * bn) * inv_rho;
    if (av < 0.0) av = 0.0;
    return fabs(w[P::MOM + ID]) + fsqrt(0.5 * (cfast2 + fsqrt(av)));
  }
}

// getdt_courant (maxsum, slab_uniform; src/mod_dt.t:160-287) of one
// conserved cell folded into the running maxima maxtot (of sum_d
// cmax_d/dx_d) and cmaxb (of cmax).  Every operation is written here
// without contraction whatever the translation unit's setting, so that the
// dt kernel and the stage kernel's fused CFL partials round alike.
template <class P>
__device__ __forceinline__ void cfl_cell(const double* cons, const FVParams& fp, double& maxtot, double& cmaxb) {
#pragma clang fp contract(off)
  double w[P::NW];
#pragma unroll
  for (int iw = 0; iw < P::NW; ++iw) w[iw] = cons[iw];
  // to_prim (hd_to_primitive / mhd_to_primitive_origin)
  const double inv_rho = 1.0 / w[P::RHO];
#pragma unroll
  for (int d = 0; d < P::NDIR; ++d) w[P::MOM + d] = w[P::MOM + d] * inv_rho;
  double vsq = w[P::MOM] * w[P::MOM];
#pragma unroll
  for (int d = 1; d < P::NDIR; ++d) vsq = vsq + w[P::MOM + d] * w[P::MOM + d];
  if constexpr (!P::MHD) {
    w[P::E] = (fp.gamma - 1.0) * (w[P::E] - 0.5 * w[P::RHO] * vsq);
  } else {
    double bsq = w[P::MAG] * w[P::MAG];
#pragma unroll
    for (int d = 1; d < P::NDIR; ++d) bsq = bsq + w[P::MAG + d] * w[P::MAG + d];
    w[P::E] = fp.gamma_1 * (w[P::E] - 0.5 * (w[P::RHO] * vsq + bsq));
  }
  double cm = cfl_cmax<P, 0>(w, fp);
  cmaxb = fmax(cmaxb, cm);
  double tot = cm * fp.dxinv[0];
  if constexpr (P::NDIM > 1) {
    cm = cfl_cmax<P, 1>(w, fp);
    cmaxb = fmax(cmaxb, cm);
    tot = tot + cm * fp.dxinv[1];
  }
  if constexpr (P::NDIM > 2) {
    cm = cfl_cmax<P, 2>(w, fp);
    cmaxb = fmax(cmaxb, cm);
    tot = tot + cm * fp.dxinv[2];
  }
  maxtot = fmax(maxtot, tot);
}

}  // namespace FV_NS
