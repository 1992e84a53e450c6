/*
 * amrvac_oracle.c -- FP64 CPU restatement of MPI-AMRVAC's finite-volume
 * block update.  TEST INFRASTRUCTURE ONLY: it is the parity checker for the
 * HIP path (tests/, __graft_entry__.smoke) and the CPU baseline of bench.py.
 * The product never links it.
 *
 * Every routine restates one routine of the reference, cited by file:line
 * (paths relative to the reference root).  Expression order follows the
 * reference's Fortran so that, compiled without FMA contraction
 * (-ffp-contract=off), each operation rounds the way the reference's does.
 *
 * Pinning (what checks each function; details in tests/):
 *  - golden regression logs of the reference's own test suite
 *    (tests/mhd/Alfven_wave_2.5D, tests/mhd/field_loop_2.5D): MHD fluxes,
 *    HLL/TVDLF, cada3/minmod, Hancock predictor, Linde source, cmaxmean
 *    bounds, ssprk3 / predictor-corrector, maxsum dt;
 *  - closed forms and invariants (tests/test_oracle_math.py): exact Riemann
 *    solution (Sod), HLLC stationary-contact exactness, flux consistency,
 *    limiter identities, WENO5 polynomial exactness, GLM psi decay,
 *    conservation, uniform-state preservation, Galilean shift.
 * Parity unpinned: none of the golden logs exercises GLM together with
 * Einfeldt bounds, so the GLM path is pinned only by the closed-form checks.
 */
#include "amrvac_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXW 16
#define MAXLINE 64
#define MIN2(a, b) ((a) < (b) ? (a) : (b))
#define MAX2(a, b) ((a) > (b) ? (a) : (b))

/* src/mod_constants.t:8-18 */
static const double smalldouble = 1.0e-12;
static const double bigdouble = 1.0e+99;
static const double half = 0.5, one = 1.0, two = 2.0, zero = 0.0;
static const double quarter = 0.25;
static const double third = 1.0 / 3.0;

enum { FT_DEFAULT = 0, FT_TVDLF = 1, FT_SPECIAL = 2 };

struct orc_sim {
  orc_params p;
  int nw;
  int nws, nwt;    /* face-field components (constrained transport) and nw + nws */
  int ct;          /* type_ct, 0 without constrained transport */
  int rho_, mom[3], e_, mag[3], psi_;
  int ng, bnx[3], nx[3];
  long ncell;
  int nb;
  double dx[3];
  int nnb;
  int* neighbor;   /* [nb][nnb] same-level neighbour, -1 otherwise */
  double* xlohi;
  double* bdx;     /* [nb][3] cell size of each block */
  int amr;         /* level jumps present: AMR ghost cells and refluxing */
  int fix_conserve;
  int* ntype;      /* [nb][nnb]: 1 boundary, 2 coarse, 3 sibling, 4 fine */
  int* nbr;        /* [nb][nnb] sibling or coarse leaf */
  int* nchild;     /* [nb][nnb][nch] children of a fine neighbour */
  int* icpos;      /* [nb][ndim] position within the parent, 1 or 2 */
  int nch;
  int nxc[3];      /* coarse representation (psc) with ghost cells */
  long ncellc;
  double* psc;     /* [nb][nw][ncellc] */
  int nface;       /* transverse cells of one block face */
  double* bflux;   /* [nb][ndim][2][nw][nface] scaled boundary face fluxes */
  int bc[6];
  double* w[4];  /* ps, ps1, ps2, ps3 */
  int cur_ib;      /* block of a per-block copy inside advect1 (-1 otherwise) */
  int crash;
  int ftype[3][MAXW];
  double inv_gamma_1, gamma_1;
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

int orc_nw(const orc_params* p) {
  if (p->phys == ORC_PHYS_HD) return p->ndir + 2;
  return 2 * p->ndir + 2 + (p->divb == ORC_DIVB_GLM ? 1 : 0);
}

static int dir_nidx(const orc_sim* s, const int* i);

static inline long IX(const orc_sim* s, int i, int j, int k) {
  return (long)i + (long)s->nx[0] * ((long)j + (long)s->nx[1] * (long)k);
}
static inline double* BLK(const orc_sim* s, int slot, int ib) {
  return s->w[slot] + (long)ib * s->nwt * s->ncell;
}

/* ------------------------------------------------------------------ */
/* physics: conversions                                                */
/* ------------------------------------------------------------------ */

/* hd_to_primitive src/hd/mod_hd_phys.t:703-734;
   mhd_to_primitive_origin src/mhd/mod_mhd_phys.t:1741-1765 */
static void to_primitive_cell(const orc_sim* s, double* w /* [nw] */) {
  const int nd = s->p.ndir;
  double inv_rho = 1.0 / w[s->rho_];
  for (int c = 0; c < nd; ++c) w[s->mom[c]] = w[s->mom[c]] * inv_rho;
  double vsq = w[s->mom[0]] * w[s->mom[0]];
  for (int c = 1; c < nd; ++c) vsq = vsq + w[s->mom[c]] * w[s->mom[c]];
  if (s->p.phys == ORC_PHYS_HD) {
    w[s->e_] = (s->p.gamma - 1.0) * (w[s->e_] - half * w[s->rho_] * vsq);
  } else {
    double bsq = w[s->mag[0]] * w[s->mag[0]];
    for (int c = 1; c < nd; ++c) bsq = bsq + w[s->mag[c]] * w[s->mag[c]];
    w[s->e_] = s->gamma_1 * (w[s->e_] - half * (w[s->rho_] * vsq + bsq));
  }
}

/* hd_to_conserved src/hd/mod_hd_phys.t:677-700;
   mhd_to_conserved_origin src/mhd/mod_mhd_phys.t:1562-1579 */
static void to_conserved_cell(const orc_sim* s, double* w) {
  const int nd = s->p.ndir;
  double vsq = w[s->mom[0]] * w[s->mom[0]];
  for (int c = 1; c < nd; ++c) vsq = vsq + w[s->mom[c]] * w[s->mom[c]];
  if (s->p.phys == ORC_PHYS_HD) {
    w[s->e_] = w[s->e_] * s->inv_gamma_1 + half * vsq * w[s->rho_];
  } else {
    double bsq = w[s->mag[0]] * w[s->mag[0]];
    for (int c = 1; c < nd; ++c) bsq = bsq + w[s->mag[c]] * w[s->mag[c]];
    w[s->e_] = w[s->e_] * s->inv_gamma_1 + half * (vsq * w[s->rho_] + bsq);
  }
  for (int c = 0; c < nd; ++c) w[s->mom[c]] = w[s->rho_] * w[s->mom[c]];
}

/* hd_get_flux src/hd/mod_hd_phys.t:1097-1146;
   mhd_get_flux src/mhd/mod_mhd_phys.t:3810-3889 (GLM :3868-3874) */
static void get_flux(const orc_sim* s, const double* wC, const double* w,
                     int idim, double cmax_global, double* f) {
  const int nd = s->p.ndir;
  const double vn = w[s->mom[idim]];
  if (s->p.phys == ORC_PHYS_HD) {
    f[s->rho_] = vn * w[s->rho_];
    for (int c = 0; c < nd; ++c) f[s->mom[c]] = vn * wC[s->mom[c]];
    f[s->mom[idim]] = f[s->mom[idim]] + w[s->e_];
    f[s->e_] = vn * (wC[s->e_] + w[s->e_]);
    return;
  }
  const double bn = w[s->mag[idim]];
  f[s->rho_] = vn * w[s->rho_];
  for (int c = 0; c < nd; ++c)
    f[s->mom[c]] = wC[s->mom[idim]] * w[s->mom[c]] - bn * w[s->mag[c]];
  double bsq = w[s->mag[0]] * w[s->mag[0]];
  for (int c = 1; c < nd; ++c) bsq = bsq + w[s->mag[c]] * w[s->mag[c]];
  double ptotal = w[s->e_] + half * bsq;
  f[s->mom[idim]] = f[s->mom[idim]] + ptotal;
  double bv = w[s->mag[0]] * w[s->mom[0]];
  for (int c = 1; c < nd; ++c) bv = bv + w[s->mag[c]] * w[s->mom[c]];
  f[s->e_] = vn * (wC[s->e_] + ptotal) - bn * bv;
  for (int c = 0; c < nd; ++c)
    f[s->mag[c]] = vn * w[s->mag[c]] - bn * w[s->mom[c]];
  if (s->p.divb == ORC_DIVB_GLM) {
    f[s->mag[idim]] = w[s->psi_];
    f[s->psi_] = cmax_global * cmax_global * w[s->mag[idim]];
  }
}

/* mhd_get_csound_prim src/mhd/mod_mhd_phys.t:3329-3379 (no B0, no Hall) */
static double mhd_csound_prim(const orc_sim* s, const double* w, int idim) {
  const int nd = s->p.ndir;
  double inv_rho = 1.0 / w[s->rho_];
  double cs = s->p.gamma * w[s->e_] * inv_rho;
  double b2 = w[s->mag[0]] * w[s->mag[0]];
  for (int c = 1; c < nd; ++c) b2 = b2 + w[s->mag[c]] * w[s->mag[c]];
  double cfast2 = b2 * inv_rho + cs;
  double bn = w[s->mag[idim]];
  double avmincs2 = cfast2 * cfast2 - 4.0 * cs * (bn * bn) * inv_rho;
  if (avmincs2 < zero) avmincs2 = zero;
  return sqrt(half * (cfast2 + sqrt(avmincs2)));
}

/* hd_get_cmax src/hd/mod_hd_phys.t:787-811;
   mhd_get_cmax_origin src/mhd/mod_mhd_phys.t:2482-2550 (w primitive) */
static double get_cmax(const orc_sim* s, const double* w, int idim) {
  if (s->p.phys == ORC_PHYS_HD)
    return fabs(w[s->mom[idim]]) + sqrt(s->p.gamma * w[s->e_] / w[s->rho_]);
  return fabs(w[s->mom[idim]]) + mhd_csound_prim(s, w, idim);
}

/* hd_get_pthermal src/hd/mod_hd_phys.t:1021-1064 with hd_kin_en :1396-1408
   (conserved input): clipped to small_pressure with fix_small_values, else
   sets crash when below small_pressure */
static double hd_pthermal(const orc_sim* s, const double* w, int* crash) {
  const int nd = s->p.ndir;
  double msq = 0.0;
  for (int c = 0; c < nd; ++c) msq = msq + w[s->mom[c]] * w[s->mom[c]];
  double ke = 0.5 * msq / w[s->rho_];
  double pth = (s->p.gamma - 1.0) * (w[s->e_] - ke);
  if (s->p.fix_small_values) {
    if (pth < s->p.small_pressure) pth = s->p.small_pressure;
  } else if (s->p.check_small_values && pth < s->p.small_pressure) {
    *crash = 1;
  }
  return pth;
}

/* phys_check_w on a conserved cell: hd_check_w src/hd/mod_hd_phys.t:648-674,
   mhd_check_w_origin src/mhd/mod_mhd_phys.t:1452-1473 (small_e =
   small_pressure/(gamma-1), :1214).  With fix_small_values and the default
   small_values_method 'error' a flagged cell sets crash
   (small_values_error src/mod_small_values.t:25-87). */
static int small_cons_fails(const orc_sim* s, const double* w) {
  const int nd = s->p.ndir;
  if (w[s->rho_] < s->p.small_density) return 1;
  double msq = 0.0;
  for (int c = 0; c < nd; ++c) msq = msq + w[s->mom[c]] * w[s->mom[c]];
  if (s->p.phys == ORC_PHYS_HD) {
    const double tmp = (s->p.gamma - 1.0) * (w[s->e_] - half * msq / w[s->rho_]);
    return tmp < s->p.small_pressure;
  }
  double bsq = 0.0;
  for (int c = 0; c < nd; ++c) bsq = bsq + w[s->mag[c]] * w[s->mag[c]];
  const double small_e = s->p.small_pressure * s->inv_gamma_1;
  return w[s->e_] - half * (msq / w[s->rho_] + bsq) < small_e;
}

/* phys_handle_small_values(.false., w, ixO) with small_values_method
   'error' over the interior of a block (conserved variables) */
static int small_values_interior(const orc_sim* s, const double* w) {
  const long n = s->ncell;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < s->p.ndim ? s->ng : 0;
    hi[d] = d < s->p.ndim ? s->ng + s->bnx[d] - 1 : 0;
  }
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        double cell[MAXW];
        for (int iw = 0; iw < s->nw; ++iw) cell[iw] = w[iw * n + IX(s, i, j, k)];
        if (small_cons_fails(s, cell)) return 1;
      }
  return 0;
}

/* hd_get_cbounds src/hd/mod_hd_phys.t:891-1001;
   mhd_get_cbounds src/mhd/mod_mhd_phys.t:3037-3127 */
static void get_cbounds(const orc_sim* s, const double* wLC, const double* wRC,
                        const double* wLp, const double* wRp, int idim,
                        int want_cmin, double* cmax, double* cmin, int* crash) {
  const int rho = s->rho_, mn = s->mom[idim];
  const double g = s->p.gamma;
  if (s->p.phys == ORC_PHYS_HD) {
    switch (s->p.boundspeed) {
      case 1: {
        double tmp1 = sqrt(wLp[rho]);
        double tmp2 = sqrt(wRp[rho]);
        double tmp3 = 1.0 / (tmp1 + tmp2);
        double umean = (wLp[mn] * tmp1 + wRp[mn] * tmp2) * tmp3;
        double csL = g * wLp[s->e_] / wLp[rho];
        double csR = g * wRp[s->e_] / wRp[rho];
        double dv = wRp[mn] - wLp[mn];
        double dmean = (tmp1 * csL + tmp2 * csR) * tmp3 +
                       0.5 * tmp1 * tmp2 * (tmp3 * tmp3) * (dv * dv);
        dmean = sqrt(dmean);
        if (want_cmin) {
          *cmin = umean - dmean;
          *cmax = umean + dmean;
        } else {
          *cmax = fabs(umean) + dmean;
        }
        return;
      }
      case 2: {
        double wm[MAXW];
        for (int iw = 0; iw < s->nw; ++iw) wm[iw] = 0.5 * (wLC[iw] + wRC[iw]);
        double tmp1 = wm[mn] / wm[rho];
        double cs = g * hd_pthermal(s, wm, crash) / wm[rho];
        cs = sqrt(cs);
        if (want_cmin) {
          *cmax = MAX2(tmp1 + cs, zero);
          *cmin = MIN2(tmp1 - cs, zero);
        } else {
          *cmax = fabs(tmp1) + cs;
        }
        return;
      }
      default: {
        double csL = g * wLp[s->e_] / wLp[rho];
        double csR = g * wRp[s->e_] / wRp[rho];
        double cs = MAX2(sqrt(csL), sqrt(csR));
        if (want_cmin) {
          *cmin = MIN2(wLp[mn], wRp[mn]) - cs;
          *cmax = MAX2(wLp[mn], wRp[mn]) + cs;
        } else {
          *cmax = MAX2(wLp[mn], wRp[mn]) + cs;
        }
        return;
      }
    }
  }
  switch (s->p.boundspeed) {
    case 1: {
      double csL = mhd_csound_prim(s, wLp, idim);
      double csR = mhd_csound_prim(s, wRp, idim);
      double tmp1 = sqrt(wLp[rho]);
      double tmp2 = sqrt(wRp[rho]);
      double tmp3 = 1.0 / (tmp1 + tmp2);
      double umean = (wLp[mn] * tmp1 + wRp[mn] * tmp2) * tmp3;
      double dv = wRp[mn] - wLp[mn];
      double dmean = sqrt((tmp1 * (csL * csL) + tmp2 * (csR * csR)) * tmp3 +
                          half * tmp1 * tmp2 * (tmp3 * tmp3) * (dv * dv));
      if (want_cmin) {
        *cmin = umean - dmean;
        *cmax = umean + dmean;
      } else {
        *cmax = fabs(umean) + dmean;
      }
      return;
    }
    case 2: {
      double wm[MAXW];
      for (int iw = 0; iw < s->nw; ++iw) wm[iw] = 0.5 * (wLp[iw] + wRp[iw]);
      double cs = mhd_csound_prim(s, wm, idim);
      if (want_cmin) {
        *cmax = MAX2(wm[mn] + cs, zero);
        *cmin = MIN2(wm[mn] - cs, zero);
      } else {
        *cmax = fabs(wm[mn]) + cs;
      }
      return;
    }
    default: {
      double csL = mhd_csound_prim(s, wLp, idim);
      double csR = mhd_csound_prim(s, wRp, idim);
      double cs = MAX2(csL, csR);
      if (want_cmin) {
        *cmin = MIN2(wLp[mn], wRp[mn]) - cs;
        *cmax = MAX2(wLp[mn], wRp[mn]) + cs;
      } else {
        *cmax = MAX2(wLp[mn], wRp[mn]) + cs;
      }
      return;
    }
  }
}

/* mhd_modify_wLR src/mhd/mod_mhd_phys.t:6326-6377 (GLM branch) */
static void modify_wLR_glm(const orc_sim* s, double* wLC, double* wRC,
                           double* wLp, double* wRp, int idir, double ch) {
  const int b = s->mag[idir], psi = s->psi_, e = s->e_;
  double dB = wRp[b] - wLp[b];
  double dPsi = wRp[psi] - wLp[psi];
  wLp[b] = half * (wRp[b] + wLp[b] - dPsi / ch);
  wLp[psi] = half * (wRp[psi] + wLp[psi] - dB * ch);
  wRp[b] = wLp[b];
  wRp[psi] = wLp[psi];
  wRC[e] = wRC[e] - half * (wRC[b] * wRC[b]);
  wLC[e] = wLC[e] - half * (wLC[b] * wLC[b]);
  wRC[b] = wLp[b];
  wRC[psi] = wLp[psi];
  wLC[b] = wLp[b];
  wLC[psi] = wLp[psi];
  wRC[e] = wRC[e] + half * (wRC[b] * wRC[b]);
  wLC[e] = wLC[e] + half * (wLC[b] * wLC[b]);
}

/* ------------------------------------------------------------------ */
/* limiters                                                            */
/* ------------------------------------------------------------------ */

/* dwlimiter2 src/limiter/mod_limiter.t:128-334, one cell: a = dwC(i-1),
   b = dwC(i); returns ldw(i) and rdw(i) */
static void dwlimiter_cell(const orc_sim* s, int idim, double a, double b,
                           double* ldw, double* rdw) {
  const double qsmall = 1.0e-12, qsmall2 = 2.0e-12;
  const double eps = sqrt(2.220446049250313e-16); /* sqrt(epsilon(1d0)) */
  const double c_mcbeta = 1.4;
  const double cadalfa = 0.5, cadbeta = 2.0, cadgamma = 1.6;
  const double cadepsilon = 1.0e-14, invcadepsilon = 1.0e14;
  double tmp, tmp2;
  switch (s->p.limiter) {
    case ORC_LIM_MINMOD: /* :173-179 */
      tmp = copysign(one, b);
      tmp = tmp * MAX2(zero, MIN2(fabs(b), tmp * a));
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_WOODWARD: /* :180-187 */
      tmp = copysign(one, b);
      tmp = 2 * tmp *
            MAX2(zero, MIN2(MIN2(fabs(b), tmp * a), tmp * quarter * (a + b)));
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_MCBETA: /* :188-195 */
      tmp = copysign(one, b);
      tmp = tmp * MAX2(zero, MIN2(MIN2(c_mcbeta * fabs(b), c_mcbeta * tmp * a),
                                  tmp * half * (a + b)));
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_SUPERBEE: /* :196-203 */
      tmp = copysign(one, b);
      tmp = tmp * MAX2(MAX2(zero, MIN2(2 * fabs(b), tmp * a)),
                       MIN2(fabs(b), 2 * tmp * a));
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_VANLEER: /* :204-209 */
      tmp = 2 * MAX2(a * b, zero) / (b + a + qsmall);
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_ALBADA: /* :210-216 */
      tmp = (a * (b * b + qsmall) + b * (a * a + qsmall)) /
            (b * b + a * a + qsmall2);
      *ldw = tmp;
      *rdw = tmp;
      return;
    case ORC_LIM_KOREN: /* :217-230 */
      tmp = copysign(one, b);
      tmp2 = MIN2(2 * fabs(b), 2 * tmp * a);
      *ldw = tmp * MAX2(zero, MIN2(tmp2, (a * tmp + 2 * fabs(b)) * third));
      *rdw = tmp * MAX2(zero, MIN2(tmp2, (2 * a * tmp + fabs(b)) * third));
      return;
    case ORC_LIM_CADA: { /* :231-252 */
      double A;
      tmp = a / (b + copysign(eps, b));
      A = (two + tmp) * third;
      *ldw = MAX2(zero, MIN2(A, MAX2(-cadalfa * tmp, MIN2(MIN2(cadbeta * tmp, A), cadgamma)))) * b;
      tmp = b / (a + copysign(eps, a));
      A = (two + tmp) * third;
      *rdw = MAX2(zero, MIN2(A, MAX2(-cadalfa * tmp, MIN2(MIN2(cadbeta * tmp, A), cadgamma)))) * a;
      return;
    }
    case ORC_LIM_CADA3: { /* :253-290 */
      double r = s->p.cada3_radius * s->dx[idim];
      double rdelinv = one / (r * r);
      double tmpeta = (b * b + a * a) * rdelinv;
      double A, lim;
      /* left */
      tmp = a / (b + copysign(eps, b));
      A = (two + tmp) * third;
      lim = MAX2(zero, MIN2(A, MAX2(-cadalfa * tmp,
                                    MIN2(MIN2(cadbeta * tmp, A), cadgamma))));
      if (tmpeta <= one - cadepsilon) {
        *ldw = A;
      } else if (tmpeta >= one + cadepsilon) {
        *ldw = lim;
      } else {
        tmp2 = (tmpeta - one) * invcadepsilon;
        *ldw = half * ((one - tmp2) * A + (one + tmp2) * lim);
      }
      *ldw = *ldw * b;
      /* right */
      tmp = b / (a + copysign(eps, a));
      A = (two + tmp) * third;
      lim = MAX2(zero, MIN2(A, MAX2(-cadalfa * tmp,
                                    MIN2(MIN2(cadbeta * tmp, A), cadgamma))));
      if (tmpeta <= one - cadepsilon) {
        *rdw = A;
      } else if (tmpeta >= one + cadepsilon) {
        *rdw = lim;
      } else {
        tmp2 = (tmpeta - one) * invcadepsilon;
        *rdw = half * ((one - tmp2) * A + (one + tmp2) * lim);
      }
      *rdw = *rdw * a;
      return;
    }
    default:
      *ldw = 0.0;
      *rdw = 0.0;
      return;
  }
}

/* primitive-state check used by fix_limiter: hd_check_w
   src/hd/mod_hd_phys.t:648-674, mhd_check_w_origin
   src/mhd/mod_mhd_phys.t:1452-1473 (primitive=.true.) */
static int check_w_prim_fails(const orc_sim* s, const double* w) {
  if (w[s->rho_] < s->p.small_density) return 1;
  if (w[s->e_] < s->p.small_pressure) return 1;
  return 0;
}

/* WENO5limiter var=1 src/limiter/mod_weno.t:254-377 for face i (between
   line cells i and i+1), one variable; returns the unfixed L and R values */
static void weno5_face(const double* w, int i, double* wl, double* wr) {
  const double bc1 = 1.0833333333333333, bc2 = 0.25;
  const double d1 = 1.0 / 10.0, d2 = 3.0 / 5.0, d3 = 3.0 / 10.0;
  const double u11 = 1.0 / 3.0, u12 = -7.0 / 6.0, u13 = 11.0 / 6.0;
  const double u21 = -1.0 / 6.0, u22 = 5.0 / 6.0, u23 = 1.0 / 3.0;
  const double u31 = 1.0 / 3.0, u32 = 5.0 / 6.0, u33 = -1.0 / 6.0;
  const double weps = 1.0e-18;
  double f1, f2, f3, b1, b2, b3, a1, a2, a3, as, t, q;
  /* left side */
  {
    double wmm = w[i - 2], wm = w[i - 1], w0 = w[i], wp = w[i + 1], wpp = w[i + 2];
    f1 = u11 * wmm + u12 * wm + u13 * w0;
    f2 = u21 * wm + u22 * w0 + u23 * wp;
    f3 = u31 * w0 + u32 * wp + u33 * wpp;
    t = wmm + w0 - 2.0 * wm; q = wmm - 4.0 * wm + 3.0 * w0;
    b1 = bc1 * (t * t) + bc2 * (q * q);
    t = wm + wp - 2.0 * w0; q = wm - wp;
    b2 = bc1 * (t * t) + bc2 * (q * q);
    t = w0 + wpp - 2.0 * wp; q = 3.0 * w0 - 4.0 * wp + wpp;
    b3 = bc1 * (t * t) + bc2 * (q * q);
    t = b1 + weps; a1 = d1 / (t * t);
    t = b2 + weps; a2 = d2 / (t * t);
    t = b3 + weps; a3 = d3 / (t * t);
    as = 0.0; as = as + a1; as = as + a2; as = as + a3;
    double fl = 0.0;
    fl = fl + f1 * a1 / as;
    fl = fl + f2 * a2 / as;
    fl = fl + f3 * a3 / as;
    *wl = fl;
  }
  /* right side */
  {
    double wppp = w[i + 3], wpp = w[i + 2], wp = w[i + 1], w0 = w[i], wm = w[i - 1];
    f1 = u11 * wppp + u12 * wpp + u13 * wp;
    f2 = u21 * wpp + u22 * wp + u23 * w0;
    f3 = u31 * wp + u32 * w0 + u33 * wm;
    t = wppp + wp - 2.0 * wpp; q = wppp - 4.0 * wpp + 3.0 * wp;
    b1 = bc1 * (t * t) + bc2 * (q * q);
    t = wpp + w0 - 2.0 * wp; q = wpp - w0;
    b2 = bc1 * (t * t) + bc2 * (q * q);
    t = wp + wm - 2.0 * w0; q = 3.0 * wp - 4.0 * w0 + wm;
    b3 = bc1 * (t * t) + bc2 * (q * q);
    t = b1 + weps; a1 = d1 / (t * t);
    t = b2 + weps; a2 = d2 / (t * t);
    t = b3 + weps; a3 = d3 / (t * t);
    as = 0.0; as = as + a1; as = as + a2; as = as + a3;
    double fr = 0.0;
    fr = fr + f1 * a1 / as;
    fr = fr + f2 * a2 / as;
    fr = fr + f3 * a3 / as;
    *wr = fr;
  }
}

/* ------------------------------------------------------------------ */
/* Riemann solvers (one face)                                          */
/* ------------------------------------------------------------------ */

/* get_Riemann_flux_tvdlf src/mod_finite_volume.t:348-379 */
static void riemann_tvdlf(const orc_sim* s, const double* wLC, const double* wRC,
                          const double* fL, const double* fR, double cmax,
                          double* fC) {
  double fac = -0.5 * s->p.tvdlfeps * cmax;
  for (int iw = 0; iw < s->nw; ++iw) {
    fC[iw] = 0.5 * (fL[iw] + fR[iw]);
    fC[iw] = fC[iw] + fac * (wRC[iw] - wLC[iw]);
  }
}

/* get_Riemann_flux_hll src/mod_finite_volume.t:381-437 */
static void riemann_hll(const orc_sim* s, int idim, const double* wLC,
                        const double* wRC, const double* fL, const double* fR,
                        double cmax, double cmin, double* fC) {
  for (int iw = 0; iw < s->nw; ++iw) {
    if (s->ftype[idim][iw] == FT_TVDLF) {
      /* constrained transport: no normal-field flux (:389-392) */
      fC[iw] = s->ct ? zero : -s->p.tvdlfeps * half * MAX2(cmax, fabs(cmin)) * (wRC[iw] - wLC[iw]);
    } else if (cmin >= zero) {
      fC[iw] = fL[iw];
    } else if (cmax <= zero) {
      fC[iw] = fR[iw];
    } else {
      fC[iw] = (cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * (wRC[iw] - wLC[iw])) /
               (cmax - cmin);
    }
  }
}

/* get_Riemann_flux_hlld src/mod_finite_volume.t:505-718 (Miyoshi & Kusano
   2005; no B0, no stagger grid, no equilibrium variables, ndir = 3): the
   four star states w1L, w1R, w2L, w2R of one face, then the region of the
   fan containing the face picks the flux per variable.  wLp/wRp are the
   primitive states after modify_wLR, wLC/wRC the conserved ones. */
static void riemann_hlld(const orc_sim* s, int idim, const double* wLC, const double* wRC,
                         const double* wLp, const double* wRp, const double* fL,
                         const double* fR, double cmax, double cmin, const double* bx_ct,
                         double* fC) {
  const double sR = cmax, sL = cmin;
  const int rho_ = s->rho_, e_ = s->e_, p_ = s->e_;
  const int ip1 = idim, ip2 = (idim + 1) % 3, ip3 = (idim + 2) % 3;
  const int m1 = s->mom[ip1], m2 = s->mom[ip2], m3 = s->mom[ip3];
  const int b1 = s->mag[ip1], b2 = s->mag[ip2], b3 = s->mag[ip3];
  double w1R[MAXW], w1L[MAXW], w2R[MAXW], w2L[MAXW];
  for (int iw = 0; iw < s->nw; ++iw) { w1R[iw] = w1L[iw] = w2R[iw] = w2L[iw] = 0.0; }
  const double* BR = wRC;  /* magnetic field of the reconstructions, by index */
  const double* BL = wLC;
  /* HLL estimate of the normal field, Li 2005 eq. (33); with constrained
     transport the face field of block (ps, whatever state the stage
     reads, :536-537) */
  const double Bx = bx_ct ? *bx_ct : (sR * BR[b1] - sL * BL[b1]) / (sR - sL);
  const double ptR = wRp[p_] + 0.5 * (BR[s->mag[0]] * BR[s->mag[0]] + BR[s->mag[1]] * BR[s->mag[1]] +
                                      BR[s->mag[2]] * BR[s->mag[2]]);
  const double ptL = wLp[p_] + 0.5 * (BL[s->mag[0]] * BL[s->mag[0]] + BL[s->mag[1]] * BL[s->mag[1]] +
                                      BL[s->mag[2]] * BL[s->mag[2]]);
  const double suR = (sR - wRp[m1]) * wRC[rho_];
  const double suL = (sL - wLp[m1]) * wLC[rho_];
  const double sm = (suR * wRp[m1] - suL * wLp[m1] - ptR + ptL) / (suR - suL);
  w1R[m1] = sm; w1L[m1] = sm; w2R[m1] = sm; w2L[m1] = sm;
  w1R[b1] = Bx; w1L[b1] = Bx;
  w1R[rho_] = suR / (sR - sm);
  w1L[rho_] = suL / (sL - sm);
  double r1R = suR * (sR - sm) - Bx * Bx;
  if (r1R != 0.0) r1R = 1.0 / r1R;
  double r1L = suL * (sL - sm) - Bx * Bx;
  if (r1L != 0.0) r1L = 1.0 / r1L;
  w1R[m2] = wRp[m2] - Bx * BR[b2] * (sm - wRp[m1]) * r1R;
  w1L[m2] = wLp[m2] - Bx * BL[b2] * (sm - wLp[m1]) * r1L;
  w1R[b2] = (suR * (sR - wRp[m1]) - Bx * Bx) * r1R;
  w1L[b2] = (suL * (sL - wLp[m1]) - Bx * Bx) * r1L;
  w1R[m3] = wRp[m3] - Bx * BR[b3] * (sm - wRp[m1]) * r1R;
  w1L[m3] = wLp[m3] - Bx * BL[b3] * (sm - wLp[m1]) * r1L;
  w1R[b3] = BR[b3] * w1R[b2];
  w1L[b3] = BL[b3] * w1L[b2];
  w1R[b2] = BR[b2] * w1R[b2];
  w1L[b2] = BL[b2] * w1L[b2];
  /* total pressure of the star region (Guo eq. 25), kept in the energy slot
     until the energy is formed (p_ = e_) */
  w1R[p_] = suR * (sm - wRp[m1]) + ptR;
  w1L[p_] = w1R[p_];
  {
    const int* mo = s->mom; const int* ma = s->mag;
    const double vbR = wRp[mo[0]] * wRC[ma[0]] + wRp[mo[1]] * wRC[ma[1]] + wRp[mo[2]] * wRC[ma[2]];
    const double vb1R = w1R[mo[0]] * w1R[ma[0]] + w1R[mo[1]] * w1R[ma[1]] + w1R[mo[2]] * w1R[ma[2]];
    w1R[e_] = ((sR - wRp[m1]) * wRC[e_] - ptR * wRp[m1] + w1R[p_] * sm + Bx * (vbR - vb1R)) / (sR - sm);
    const double vbL = wLp[mo[0]] * wLC[ma[0]] + wLp[mo[1]] * wLC[ma[1]] + wLp[mo[2]] * wLC[ma[2]];
    const double vb1L = w1L[mo[0]] * w1L[ma[0]] + w1L[mo[1]] * w1L[ma[1]] + w1L[mo[2]] * w1L[ma[2]];
    w1L[e_] = ((sL - wLp[m1]) * wLC[e_] - ptL * wLp[m1] + w1L[p_] * sm + Bx * (vbL - vb1L)) / (sL - sm);
  }
  /* the Alfven-wave (double star) states */
  w2R[rho_] = w1R[rho_]; w2L[rho_] = w1L[rho_];
  w2R[b1] = w1R[b1]; w2L[b1] = w1L[b1];
  const double sR1 = sqrt(w1R[rho_]);
  const double sL1 = sqrt(w1L[rho_]);
  const double tmp = 1.0 / (sR1 + sL1);
  const double signBx = copysign(1.0, Bx);
  const double s1R = sm + fabs(Bx) / sR1;
  const double s1L = sm - fabs(Bx) / sL1;
  w2R[m2] = (sL1 * w1L[m2] + sR1 * w1R[m2] + (w1R[b2] - w1L[b2]) * signBx) * tmp;
  w2L[m2] = w2R[m2];
  w2R[b2] = (sL1 * w1R[b2] + sR1 * w1L[b2] + sL1 * sR1 * (w1R[m2] - w1L[m2]) * signBx) * tmp;
  w2L[b2] = w2R[b2];
  w2R[m3] = (sL1 * w1L[m3] + sR1 * w1R[m3] + (w1R[b3] - w1L[b3]) * signBx) * tmp;
  w2L[m3] = w2R[m3];
  w2R[b3] = (sL1 * w1R[b3] + sR1 * w1L[b3] + sL1 * sR1 * (w1R[m3] - w1L[m3]) * signBx) * tmp;
  w2L[b3] = w2R[b3];
  {
    const int* mo = s->mom; const int* ma = s->mag;
    const double d1R = w1R[mo[0]] * w1R[ma[0]] + w1R[mo[1]] * w1R[ma[1]] + w1R[mo[2]] * w1R[ma[2]];
    const double d2R = w2R[mo[0]] * w2R[ma[0]] + w2R[mo[1]] * w2R[ma[1]] + w2R[mo[2]] * w2R[ma[2]];
    w2R[e_] = w1R[e_] + sR1 * (d1R - d2R) * signBx;
    const double d1L = w1L[mo[0]] * w1L[ma[0]] + w1L[mo[1]] * w1L[ma[1]] + w1L[mo[2]] * w1L[ma[2]];
    const double d2L = w2L[mo[0]] * w2L[ma[0]] + w2L[mo[1]] * w2L[ma[1]] + w2L[mo[2]] * w2L[ma[2]];
    w2L[e_] = w1L[e_] - sL1 * (d1L - d2L) * signBx;
  }
  /* velocities to momenta */
  for (int c = 0; c < 3; ++c) {
    const int m = s->mom[c];
    w1R[m] = w1R[m] * w1R[rho_];
    w1L[m] = w1L[m] * w1L[rho_];
    w2R[m] = w2R[m] * w2R[rho_];
    w2L[m] = w2L[m] * w2L[rho_];
  }
  for (int iw = 0; iw < s->nw; ++iw) {
    if (s->ftype[idim][iw] == FT_SPECIAL) {
      fC[iw] = fL[iw];
    } else if (sL > 0.0) {
      fC[iw] = fL[iw];
    } else if (s1L >= 0.0) {
      fC[iw] = fL[iw] + sL * (w1L[iw] - wLC[iw]);
    } else if (sm >= 0.0) {
      fC[iw] = fL[iw] + sL * (w1L[iw] - wLC[iw]) + s1L * (w2L[iw] - w1L[iw]);
    } else if (s1R >= 0.0) {
      fC[iw] = fR[iw] + sR * (w1R[iw] - wRC[iw]) + s1R * (w2R[iw] - w1R[iw]);
    } else if (sR >= 0.0) {
      fC[iw] = fR[iw] + sR * (w1R[iw] - wRC[iw]);
    } else {
      fC[iw] = fR[iw];
    }
  }
}

/* get_Riemann_flux_hllc src/mod_finite_volume.t:439-504 with hd_get_lCD
   src/hd/mod_hd_hllc.t:56-152 and hd_get_wCD :154-257, or mhd_get_lCD
   src/mhd/mod_mhd_hllc.t:55-145 (the same contact speed and patch logic,
   flathllc off) and mhd_get_wCD :147-240 */
static void riemann_hllc(const orc_sim* s, int idim, const double* wLC,
                         const double* wRC, const double* fL, const double* fR,
                         double cmax, double cmin, double* fC) {
  const int nw = s->nw, rho = s->rho_, mn = s->mom[idim], e = s->e_;
  double whll[MAXW], Fhll[MAXW], fCD[MAXW];
  double lambdaCD = 0.0;
  int patchf = 1;
  if (cmin >= zero) patchf = -2;
  else if (cmax <= zero) patchf = 2;
  if (patchf == 1) {
    for (int iw = 0; iw < nw; ++iw) {
      Fhll[iw] = (cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * (wRC[iw] - wLC[iw])) /
                 (cmax - cmin);
      whll[iw] = (cmax * wRC[iw] - cmin * wLC[iw] + fL[iw] - fR[iw]) / (cmax - cmin);
    }
    lambdaCD = whll[mn] / whll[rho];
    int cond = 1;
    if (cmin < zero && lambdaCD > zero && lambdaCD < cmax) {
      patchf = -1;
    } else if (cmax > zero && lambdaCD < zero && lambdaCD > cmin) {
      patchf = 1;
    } else if (lambdaCD >= cmax || lambdaCD <= cmin) {
      lambdaCD = zero;
      patchf = 3;
    }
    if (patchf == 3) cond = 0;
    if (cond && lambdaCD == zero) patchf = (-cmin >= cmax) ? 1 : -1;
  }
  if (patchf == 1 || patchf == -1) {
    const double* wSub;
    const double* fSub;
    double cspeed, vSub;
    if (patchf == 1) {
      cspeed = cmax;
      vSub = wRC[mn] / wRC[rho];
      wSub = wRC;
      fSub = fR;
    } else {
      cspeed = cmin;
      vSub = wLC[mn] / wLC[rho];
      wSub = wLC;
      fSub = fL;
    }
    double wCD[MAXW];
    if (s->p.phys == ORC_PHYS_MHD) {
      /* mhd_get_wCD src/mhd/mod_mhd_hllc.t:147-240 (Li 2005; no tracers):
         the HLL magnetic field, eqs. 17, 20-22 and 31; the normal field
         and psi take no flux in the contact region */
      const double den = cspeed - lambdaCD;
      const int bn = s->mag[idim];
      wCD[rho] = wSub[rho] * (cspeed - vSub) / den;
      for (int c = 0; c < s->p.ndir; ++c) wCD[s->mag[c]] = whll[s->mag[c]];
      for (int c = 0; c < s->p.ndir; ++c) {
        const int m = s->mom[c];
        if (c != idim)
          wCD[m] = (cspeed * wSub[m] - fSub[m] - wCD[bn] * wCD[s->mag[c]]) / den;
        else
          wCD[m] = wCD[rho] * lambdaCD;
      }
      double vb = whll[s->mom[0]] * whll[s->mag[0]];
      for (int c = 1; c < s->p.ndir; ++c) vb = vb + whll[s->mom[c]] * whll[s->mag[c]];
      const double VdotBCD = vb / whll[rho];
      const double pCD = wSub[rho] * (cspeed - vSub) * (lambdaCD - vSub) + fSub[mn] - wSub[mn] * vSub +
                         wCD[bn] * wCD[bn];
      wCD[e] = (cspeed * wSub[e] - fSub[e] + lambdaCD * pCD - VdotBCD * wCD[bn]) / den;
      for (int iw = 0; iw < nw; ++iw)
        fCD[iw] = (iw == bn || iw == s->psi_) ? zero : fSub[iw] + cspeed * (wCD[iw] - wSub[iw]);
    } else {
      double csmls = one / (cspeed - lambdaCD);
      wCD[rho] = wSub[rho] * (cspeed - vSub) * csmls;
      for (int c = 0; c < s->p.ndir; ++c) {
        if (c != idim)
          wCD[s->mom[c]] = (cspeed * wSub[s->mom[c]] - fSub[s->mom[c]]) * csmls;
        else
          wCD[s->mom[c]] = wCD[rho] * lambdaCD;
      }
      double pCD = wSub[rho] * (cspeed - vSub) * (lambdaCD - vSub) + fSub[mn] -
                   wSub[mn] * vSub;
      wCD[e] = (cspeed * wSub[e] - fSub[e] + lambdaCD * pCD) * csmls;
      for (int iw = 0; iw < nw; ++iw)
        fCD[iw] = fSub[iw] + cspeed * (wCD[iw] - wSub[iw]);
    }
  }
  for (int iw = 0; iw < nw; ++iw) {
    if (s->ftype[idim][iw] == FT_TVDLF) {
      fC[iw] = -s->p.tvdlfeps * half * MAX2(cmax, fabs(cmin)) * (wRC[iw] - wLC[iw]);
      continue;
    }
    switch (patchf) {
      case -2: fC[iw] = fL[iw]; break;
      case -1:
      case 1: fC[iw] = fCD[iw]; break;
      case 2: fC[iw] = fR[iw]; break;
      default: fC[iw] = Fhll[iw]; break;
    }
  }
}

/* ------------------------------------------------------------------ */
/* per-block operators                                                 */
/* ------------------------------------------------------------------ */

/* phys_to_primitive over the ghosted block (ixI); with fix_small_values
   hd_to_primitive / mhd_to_primitive_origin first check every cell
   (src/hd/mod_hd_phys.t:713-715, src/mhd/mod_mhd_phys.t:1750-1753) */
static void block_to_primitive(const orc_sim* s, const double* wc, double* wp, int* crash) {
  const long n = s->ncell;
  double cell[MAXW];
  for (long c = 0; c < n; ++c) {
    for (int iw = 0; iw < s->nw; ++iw) cell[iw] = wc[iw * n + c];
    if (s->p.fix_small_values && crash && small_cons_fails(s, cell)) *crash = 1;
    to_primitive_cell(s, cell);
    for (int iw = 0; iw < s->nw; ++iw) wp[iw * n + c] = cell[iw];
  }
}

/* reconstruct_LR src/mod_finite_volume.t:1110-1228 on one line along idim:
   lw[iw][i] holds primitive values along the line; on return wLp/wRp hold
   the primitive face states of face i (between i and i+1) for
   i in [ilo, ihi]. */
static void reconstruct_line(const orc_sim* s, int idim, double lw[][MAXLINE],
                             int ilo, int ihi, double wLp[][MAXW],
                             double wRp[][MAXW]) {
  const int nw = s->nw;
  for (int i = ilo; i <= ihi; ++i)
    for (int iw = 0; iw < nw; ++iw) {
      wLp[i][iw] = lw[iw][i];
      wRp[i][iw] = lw[iw][i + 1];
    }
  if (s->p.limiter == ORC_LIM_WENO5) {
    double tl[MAXLINE][MAXW], tr[MAXLINE][MAXW];
    for (int iw = 0; iw < nw; ++iw)
      for (int i = ilo; i <= ihi; ++i) weno5_face(lw[iw], i, &tl[i][iw], &tr[i][iw]);
    /* fix_limiter src/limiter/mod_weno.t:100-133 */
    for (int i = ilo; i <= ihi; ++i) {
      if (check_w_prim_fails(s, tl[i]) || check_w_prim_fails(s, tr[i])) continue;
      for (int iw = 0; iw < nw; ++iw) {
        wLp[i][iw] = tl[i][iw];
        wRp[i][iw] = tr[i][iw];
      }
    }
    return;
  }
  /* TVD limiters: dwC(i) = w(i+1)-w(i) on [ilo-1, ihi+1] */
  for (int iw = 0; iw < nw; ++iw) {
    double dwC[MAXLINE], ldw[MAXLINE], rdw[MAXLINE];
    for (int i = ilo - 1; i <= ihi + 1; ++i) dwC[i] = lw[iw][i + 1] - lw[iw][i];
    for (int i = ilo; i <= ihi + 1; ++i)
      dwlimiter_cell(s, idim, dwC[i - 1], dwC[i], &ldw[i], &rdw[i]);
    for (int i = ilo; i <= ihi; ++i) {
      wLp[i][iw] = wLp[i][iw] + half * ldw[i];
      wRp[i][iw] = wRp[i][iw] - half * rdw[i + 1];
    }
  }
}

/* line addressing: a line along idim at transverse position (a, b), where a
   indexes the lower and b the higher of the two other dimensions */
static void trans_dims(int idim, int* ta, int* tb) {
  if (idim == 0) { *ta = 1; *tb = 2; }
  else if (idim == 1) { *ta = 0; *tb = 2; }
  else { *ta = 0; *tb = 1; }
}

static long line_cell(const orc_sim* s, int idim, int a, int b, int i) {
  int c[3], ta, tb;
  trans_dims(idim, &ta, &tb);
  c[idim] = i; c[ta] = a; c[tb] = b;
  return IX(s, c[0], c[1], c[2]);
}

static void get_line(const orc_sim* s, const double* arr, int nvar, int idim,
                     int a, int b, double out[][MAXLINE]) {
  const int nl = s->nx[idim];
  for (int iw = 0; iw < nvar; ++iw)
    for (int i = 0; i < nl; ++i)
      out[iw][i] = arr[iw * s->ncell + line_cell(s, idim, a, b, i)];
}

/* transverse ranges: interior in active dims, single index in unused dims */
static void transverse_ranges(const orc_sim* s, int idim, int* alo, int* ahi,
                              int* blo, int* bhi) {
  int ta, tb;
  trans_dims(idim, &ta, &tb);
  *alo = (ta < s->p.ndim) ? s->ng : 0;
  *ahi = (ta < s->p.ndim) ? s->ng + s->bnx[ta] - 1 : 0;
  *blo = (tb < s->p.ndim) ? s->ng : 0;
  *bhi = (tb < s->p.ndim) ? s->ng + s->bnx[tb] - 1 : 0;
}

/* central-difference helpers of src/mod_geometry.t: gradient :329-403
   (nth=1, Cartesian) and divvector :582-655 (nth=1, slab_uniform) */
static double gradient_c(const orc_sim* s, const double* q, int i, int j, int k,
                         int idir) {
  int h[3] = {i, j, k}, p[3] = {i, j, k};
  h[idir] -= 1; p[idir] += 1;
  double g = (q[IX(s, p[0], p[1], p[2])] - q[IX(s, h[0], h[1], h[2])]) * half;
  return g / s->dx[idir];
}

static double divb_c(const orc_sim* s, const double* w, int i, int j, int k) {
  double divq = 0.0;
  for (int d = 0; d < s->p.ndim; ++d) {
    const double* q = w + (long)s->mag[d] * s->ncell;
    int h[3] = {i, j, k}, p[3] = {i, j, k};
    h[d] -= 1; p[d] += 1;
    divq = divq + half * (q[IX(s, p[0], p[1], p[2])] - q[IX(s, h[0], h[1], h[2])]) *
                      (1.0 / s->dx[d]);
  }
  return divq;
}

/* add_source_glm src/mhd/mod_mhd_phys.t:5472-5531 */
static void add_source_glm(const orc_sim* s, double qdt, const double* wCT,
                           double* w, double ch) {
  const long n = s->ncell;
  double mindx = s->dx[0];
  for (int d = 1; d < s->p.ndim; ++d) mindx = MIN2(mindx, s->dx[d]);
  double decay = exp(-qdt * ch * s->p.glm_alpha / mindx);
  int klo = s->p.ndim > 2 ? s->ng : 0, khi = s->p.ndim > 2 ? s->ng + s->bnx[2] - 1 : 0;
  int jlo = s->p.ndim > 1 ? s->ng : 0, jhi = s->p.ndim > 1 ? s->ng + s->bnx[1] - 1 : 0;
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = s->ng; i < s->ng + s->bnx[0]; ++i) {
        long c = IX(s, i, j, k);
        w[s->psi_ * n + c] = decay * w[s->psi_ * n + c];
      }
  if (!s->p.glm_extended) return;
  const double* psi = wCT + (long)s->psi_ * n;
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = s->ng; i < s->ng + s->bnx[0]; ++i) {
        long c = IX(s, i, j, k);
        for (int d = 0; d < s->p.ndim; ++d) {
          double gp = gradient_c(s, psi, i, j, k, d);
          w[s->e_ * n + c] = w[s->e_ * n + c] - qdt * wCT[s->mag[d] * n + c] * gp;
        }
        double divb = divb_c(s, wCT, i, j, k);
        for (int d = 0; d < s->p.ndir; ++d)
          w[s->mom[d] * n + c] = w[s->mom[d] * n + c] - qdt * wCT[s->mag[d] * n + c] * divb;
      }
}

/* add_source_linde src/mhd/mod_mhd_phys.t:5599-5664.  div B on ixO^LADD1;
   the update range ixp is ixO, retreated by one cell layer from every face
   whose neighbour is coarser or finer (the level-jump rule :5616-5641) */
static void add_source_linde(const orc_sim* s, double qdt, const double* wCT,
                             double* w) {
  (void)qdt;
  const long n = s->ncell;
  double* divb = (double*)calloc((size_t)n, sizeof(double));
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    if (d < s->p.ndim) { lo[d] = s->ng; hi[d] = s->ng + s->bnx[d] - 1; }
    else { lo[d] = 0; hi[d] = 0; }
  }
  /* divb on ixO^LADD1 */
  for (int k = (s->p.ndim > 2 ? lo[2] - 1 : 0); k <= (s->p.ndim > 2 ? hi[2] + 1 : 0); ++k)
    for (int j = (s->p.ndim > 1 ? lo[1] - 1 : 0); j <= (s->p.ndim > 1 ? hi[1] + 1 : 0); ++j)
      for (int i = lo[0] - 1; i <= hi[0] + 1; ++i) divb[IX(s, i, j, k)] = divb_c(s, wCT, i, j, k);
  if (s->cur_ib >= 0)
    for (int d = 0; d < s->p.ndim; ++d)
      for (int side = 0; side < 2; ++side) {
        int v[3] = {0, 0, 0};
        v[d] = side == 0 ? -1 : 1;
        const int t = s->ntype[(long)s->cur_ib * s->nnb + dir_nidx(s, v)];
        if (t == 2 || t == 4) {
          if (side == 0) lo[d] = lo[d] + 1;
          else hi[d] = hi[d] - 1;
        }
      }
  double invds2 = 0.0;
  for (int d = 0; d < s->p.ndim; ++d) invds2 = invds2 + 1.0 / (s->dx[d] * s->dx[d]);
  for (int idim = 0; idim < s->p.ndim; ++idim) {
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          long c = IX(s, i, j, k);
          double g = gradient_c(s, divb, i, j, k, idim);
          g = g * s->p.divbdiff / invds2;
          w[s->mag[idim] * n + c] = w[s->mag[idim] * n + c] + g;
          w[s->e_ * n + c] = w[s->e_ * n + c] + wCT[s->mag[idim] * n + c] * g;
        }
  }
  free(divb);
}

/* addsource2 src/mod_source.t:130-160 -> hd_add_source (no active terms for
   these configs) / mhd_add_source src/mhd/mod_mhd_phys.t:4565-4695 */
static void add_sources(const orc_sim* s, double qdt, const double* wCT,
                        double* w, double ch) {
  if (s->p.phys != ORC_PHYS_MHD || s->p.ndim < 2) return;
  if (s->p.divb == ORC_DIVB_GLM) add_source_glm(s, qdt, wCT, w, ch);
  else if (s->p.divb == ORC_DIVB_LINDE) add_source_linde(s, qdt, wCT, w);
}

/* ------------------------------------------------------------------ */
/* constrained transport (stagger_grid, uniform Cartesian blocks)        */
/* ------------------------------------------------------------------ */

/* Levi-Civita symbol lvc of src/mod_global_parameters.t, 0-based */
static int lvc0(int a, int b, int c) {
  if (a == b || b == c || a == c) return 0;
  return ((a + 1) % 3 == b) ? 1 : -1;
}

/* per-block face data of one stage for the edge electric fields */
typedef struct {
  double* fm;     /* [ndim][ndir][ncell] unscaled flux of mag(c) through the upper idim face */
  double* vnorm;  /* [ndim][ncell] uct_contact: mean normal velocity of the face */
  double* cbmin;  /* [ndim][ncell] uct_hll: max(-cmin, 0) */
  double* cbmax;  /* [ndim][ncell] uct_hll: max(cmax, 0) */
  double* vbar;   /* [ndim][2][ncell] uct_hll: vbarC */
} ct_faces;

static double surfaceC(const orc_sim* s, int idim) {
  /* get_surface_area src/mod_geometry.t:196-211 (Cartesian) */
  if (s->p.ndim == 2) return idim == 0 ? s->dx[1] : s->dx[0];
  if (idim == 0) return s->dx[1] * s->dx[2];
  if (idim == 1) return s->dx[0] * s->dx[2];
  return s->dx[0] * s->dx[1];
}

static double dsC(const orc_sim* s, int idir) {
  /* src/amr/mod_amr_solution_node.t:462-467: dx along the edge, 1 for an
     ignorable direction */
  return idir < s->p.ndim ? s->dx[idir] : one;
}

static inline long SH(const orc_sim* s, long c, int d, int by) {
  return c + (d == 0 ? by : d == 1 ? (long)by * s->nx[0] : (long)by * s->nx[0] * s->nx[1]);
}

/* the box lo..hi of cell (or face, edge) indices: interior shifted by
   lo_off/hi_off per dimension */
static void ct_box(const orc_sim* s, const int* lo_off, const int* hi_off, int* lo, int* hi) {
  for (int d = 0; d < 3; ++d) {
    if (d < s->p.ndim) {
      lo[d] = s->ng + lo_off[d];
      hi[d] = s->ng + s->bnx[d] - 1 + hi_off[d];
    } else {
      lo[d] = 0;
      hi[d] = 0;
    }
  }
}

/* reconstruct src/mod_constrained_transport.t:177-237 (TVD limiters): the
   value of face quantity q on both sides of the edges of ixC along idir.
   qL/qR are indexed like q. */
static void ct_reconstruct(const orc_sim* s, const int* clo, const int* chi, int idir, const double* q,
                           double* qL, double* qR) {
  int glo[3], ghi[3];
  for (int d = 0; d < 3; ++d) { glo[d] = clo[d]; ghi[d] = chi[d]; }
  glo[idir] -= 1;
  ghi[idir] += 1;  /* gxC = [ixCmin - 1, jxCmax] */
  const long n = s->ncell;
  double* dq = (double*)malloc(sizeof(double) * (size_t)n);
  for (int k = glo[2]; k <= ghi[2]; ++k)
    for (int j = glo[1]; j <= ghi[1]; ++j)
      for (int i = glo[0]; i <= ghi[0]; ++i) {
        const long c = IX(s, i, j, k);
        qR[c] = q[SH(s, c, idir, 1)];
        qL[c] = q[c];
        dq[c] = qR[c] - qL[c];
      }
  for (int k = clo[2]; k <= chi[2]; ++k)
    for (int j = clo[1]; j <= chi[1]; ++j)
      for (int i = clo[0]; i <= chi[0]; ++i) {
        const long c = IX(s, i, j, k);
        double ldq, rdq_dummy, ldq_dummy, rdq;
        /* ldq(ix) from dq(ix-1), dq(ix); rdq(ix+1) from dq(ix), dq(ix+1) */
        dwlimiter_cell(s, idir, dq[SH(s, c, idir, -1)], dq[c], &ldq, &rdq_dummy);
        dwlimiter_cell(s, idir, dq[c], dq[SH(s, c, idir, 1)], &ldq_dummy, &rdq);
        qL[c] = qL[c] + half * ldq;
        qR[c] = qR[c] - half * rdq;
      }
  free(dq);
}

/* mhd_update_faces_average / _contact / _hll src/mhd/mod_mhd_phys.t:
   6951-7479 (no resistivity, no ambipolar term, no user electric field):
   edge electric fields times qdt and edge length, then the circulation
   around every face updates the face field bfaces of the new state.
   wp: primitive stage input; bfacesCT: face field of the stage input. */
static void ct_update_faces(const orc_sim* s, double qdt, const double* wp, const ct_faces* F,
                            const double* bfacesCT, double* bfaces) {
  const int nd = s->p.ndim, ndir = s->p.ndir;
  const long n = s->ncell;
  const int sdim = nd == 3 ? 0 : 2;
  double* fE = (double*)calloc((size_t)3 * n, sizeof(double));
  const double* fm = F->fm;
#define FM(c, iw_c, face_d) fm[((long)(face_d) * ndir + (iw_c)) * n + (c)]
  if (s->ct == ORC_CT_AVERAGE || s->ct == ORC_CT_CONTACT) {
    double* ECC = NULL;
    double* EL = NULL;
    double* ER = NULL;
    if (s->ct == ORC_CT_CONTACT) {
      /* ECC = b x v at cell centres over the whole block */
      ECC = (double*)calloc((size_t)3 * n, sizeof(double));
      EL = (double*)calloc((size_t)n, sizeof(double));
      ER = (double*)calloc((size_t)n, sizeof(double));
      const int b1 = s->mag[0], b2 = s->mag[1], m1 = s->mom[0], m2 = s->mom[1];
      for (long c = 0; c < n; ++c) {
        if (nd == 3) {
          const int b3 = s->mag[2], m3 = s->mom[2];
          ECC[0 * n + c] = wp[b2 * n + c] * wp[m3 * n + c] - wp[b3 * n + c] * wp[m2 * n + c];
          ECC[1 * n + c] = wp[b3 * n + c] * wp[m1 * n + c] - wp[b1 * n + c] * wp[m3 * n + c];
        }
        ECC[2 * n + c] = wp[b1 * n + c] * wp[m2 * n + c] - wp[b2 * n + c] * wp[m1 * n + c];
      }
    }
    for (int idim1 = 0; idim1 < nd; ++idim1)
      for (int idim2 = 0; idim2 < nd; ++idim2)
        for (int idir = sdim; idir < 3; ++idir) {
          if (lvc0(idim1, idim2, idir) != 1) continue;
          int lo_off[3], hi_off[3], lo[3], hi[3];
          for (int d = 0; d < 3; ++d) { lo_off[d] = (d == idir) ? 0 : -1; hi_off[d] = 0; }
          ct_box(s, lo_off, hi_off, lo, hi);
          const int c1 = idim1, c2 = idim2;  /* mag(idim1), mag(idim2) as component index */
          for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
              for (int i = lo[0]; i <= hi[0]; ++i) {
                const long c = IX(s, i, j, k);
                fE[idir * n + c] = quarter * (FM(c, c1, idim2) + FM(SH(s, c, idim1, 1), c1, idim2) -
                                              FM(c, c2, idim1) - FM(SH(s, c, idim2, 1), c2, idim1));
                if (s->ct == ORC_CT_AVERAGE) fE[idir * n + c] = fE[idir * n + c] * qdt * dsC(s, idir);
              }
          if (s->ct != ORC_CT_CONTACT) continue;
          const double* E = ECC + (long)idir * n;
          const double* vn1 = F->vnorm + (long)idim1 * n;
          const double* vn2 = F->vnorm + (long)idim2 * n;
          /* first pair: ixA = [ixCmin, ixCmax + e(idim1)] */
          for (int k = lo[2]; k <= hi[2] + (idim1 == 2); ++k)
            for (int j = lo[1]; j <= hi[1] + (idim1 == 1); ++j)
              for (int i = lo[0]; i <= hi[0] + (idim1 == 0); ++i) {
                const long c = IX(s, i, j, k);
                EL[c] = FM(c, c1, idim2) - E[c];
                ER[c] = FM(c, c1, idim2) - E[SH(s, c, idim2, 1)];
              }
          for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
              for (int i = lo[0]; i <= hi[0]; ++i) {
                const long c = IX(s, i, j, k), c1p = SH(s, c, idim1, 1), c2p = SH(s, c, idim2, 1);
                double ELC, ERC;
                if (vn1[c] > 0.0) ELC = EL[c];
                else if (vn1[c] < 0.0) ELC = EL[c1p];
                else ELC = 0.5 * (EL[c] + EL[c1p]);
                if (vn1[c2p] > 0.0) ERC = ER[c];
                else if (vn1[c2p] < 0.0) ERC = ER[c1p];
                else ERC = 0.5 * (ER[c] + ER[c1p]);
                fE[idir * n + c] = fE[idir * n + c] + 0.25 * (ELC + ERC);
              }
          /* second pair: ixA = [ixCmin, ixCmax + e(idim2)] */
          for (int k = lo[2]; k <= hi[2] + (idim2 == 2); ++k)
            for (int j = lo[1]; j <= hi[1] + (idim2 == 1); ++j)
              for (int i = lo[0]; i <= hi[0] + (idim2 == 0); ++i) {
                const long c = IX(s, i, j, k);
                EL[c] = -FM(c, c2, idim1) - E[c];
                ER[c] = -FM(c, c2, idim1) - E[SH(s, c, idim1, 1)];
              }
          for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
              for (int i = lo[0]; i <= hi[0]; ++i) {
                const long c = IX(s, i, j, k), c1p = SH(s, c, idim1, 1), c2p = SH(s, c, idim2, 1);
                double ELC, ERC;
                if (vn2[c] > 0.0) ELC = EL[c];
                else if (vn2[c] < 0.0) ELC = EL[c2p];
                else ELC = 0.5 * (EL[c] + EL[c2p]);
                if (vn2[c1p] > 0.0) ERC = ER[c];
                else if (vn2[c1p] < 0.0) ERC = ER[c2p];
                else ERC = 0.5 * (ER[c] + ER[c2p]);
                fE[idir * n + c] = fE[idir * n + c] + 0.25 * (ELC + ERC);
                fE[idir * n + c] = fE[idir * n + c] * qdt * dsC(s, idir);
              }
        }
    free(ECC);
    free(EL);
    free(ER);
  } else { /* uct_hll */
    double* vtilL[2];
    double* vtilR[2];
    double* btilL[3] = {NULL, NULL, NULL};
    double* btilR[3] = {NULL, NULL, NULL};
    for (int q = 0; q < 2; ++q) {
      vtilL[q] = (double*)calloc((size_t)n, sizeof(double));
      vtilR[q] = (double*)calloc((size_t)n, sizeof(double));
    }
    for (int idir = sdim; idir < 3; ++idir) {
      int lo_off[3], hi_off[3], lo[3], hi[3];
      for (int d = 0; d < 3; ++d) { lo_off[d] = (d == idir) ? 0 : -1; hi_off[d] = 0; }
      ct_box(s, lo_off, hi_off, lo, hi);
      const int idim1 = (idir + 1) % 3, idim2 = (idir + 2) % 3;
      for (int q = 0; q < 3; ++q) {
        if (q != idim1 && q != idim2) continue;
        if (!btilL[q]) {
          btilL[q] = (double*)calloc((size_t)n, sizeof(double));
          btilR[q] = (double*)calloc((size_t)n, sizeof(double));
        }
      }
      ct_reconstruct(s, lo, hi, idim2, F->vbar + ((long)idim1 * 2 + 0) * n, vtilL[1], vtilR[1]);
      ct_reconstruct(s, lo, hi, idim1, F->vbar + ((long)idim2 * 2 + 1) * n, vtilL[0], vtilR[0]);
      ct_reconstruct(s, lo, hi, idim2, bfacesCT + (long)idim1 * n, btilL[idim1], btilR[idim1]);
      ct_reconstruct(s, lo, hi, idim1, bfacesCT + (long)idim2 * n, btilL[idim2], btilR[idim2]);
      const double* cmn1 = F->cbmin + (long)idim1 * n;
      const double* cmx1 = F->cbmax + (long)idim1 * n;
      const double* cmn2 = F->cbmin + (long)idim2 * n;
      const double* cmx2 = F->cbmax + (long)idim2 * n;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) {
            const long c = IX(s, i, j, k), cp = SH(s, c, idim2, 1), jc = SH(s, c, idim1, 1);
            const double cm1 = MAX2(cmn1[cp], cmn1[c]);
            const double cp1 = MAX2(cmx1[cp], cmx1[c]);
            const double cm2 = MAX2(cmn2[jc], cmn2[c]);
            const double cp2 = MAX2(cmx2[jc], cmx2[c]);
            double e = -(cp1 * vtilL[0][c] * btilL[idim2][c] + cm1 * vtilR[0][c] * btilR[idim2][c] -
                         cp1 * cm1 * (btilR[idim2][c] - btilL[idim2][c])) /
                           (cp1 + cm1) +
                       (cp2 * vtilL[1][c] * btilL[idim1][c] + cm2 * vtilR[1][c] * btilR[idim1][c] -
                        cp2 * cm2 * (btilR[idim1][c] - btilL[idim1][c])) /
                           (cp2 + cm2);
            fE[idir * n + c] = qdt * dsC(s, idir) * e;
          }
    }
    for (int q = 0; q < 2; ++q) { free(vtilL[q]); free(vtilR[q]); }
    for (int q = 0; q < 3; ++q) { free(btilL[q]); free(btilR[q]); }
  }
#undef FM
  /* circulation around each face, then B_face -= circ / surfaceC */
  for (int idim1 = 0; idim1 < nd; ++idim1) {
    int lo_off[3], hi_off[3], lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo_off[d] = (d == idim1) ? -1 : 0; hi_off[d] = 0; }
    ct_box(s, lo_off, hi_off, lo, hi);
    const double sC = surfaceC(s, idim1);
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          const long c = IX(s, i, j, k);
          double circ = zero;
          for (int idim2 = 0; idim2 < nd; ++idim2)
            for (int idir = sdim; idir < 3; ++idir) {
              const int l = lvc0(idim1, idim2, idir);
              if (l == 0) continue;
              const double dE = fE[idir * n + c] - fE[idir * n + SH(s, c, idim2, -1)];
              if (l == 1) circ = circ + dE;
              else circ = circ - dE;
            }
          if (sC > smalldouble) bfaces[idim1 * n + c] = bfaces[idim1 * n + c] - circ / sC;
        }
  }
  free(fE);
}

/* mhd_face_to_center src/mhd/mod_mhd_phys.t:7586-7648 (arithmetic mean of
   the two faces, area weighted) on the interior cells */
static void ct_face_to_center(const orc_sim* s, double* w) {
  const long n = s->ncell;
  const double* ws = w + (long)s->nw * n;
  int lo_off[3] = {0, 0, 0}, hi_off[3] = {0, 0, 0}, lo[3], hi[3];
  ct_box(s, lo_off, hi_off, lo, hi);
  for (int d = 0; d < s->p.ndim; ++d) {
    const double sC = surfaceC(s, d), sf = sC; /* surface = surfaceC on Cartesian blocks */
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          const long c = IX(s, i, j, k);
          w[s->mag[d] * n + c] = half / sf * (ws[d * n + c] * sC + ws[d * n + SH(s, c, d, -1)] * sC);
        }
  }
}

/* finite_volume src/mod_finite_volume.t:116-339 (slab_uniform, no local
   timestep, fix_small_values off).  With constrained transport the fluxes
   are also computed on transverse_ghost_cells extra lines (:183-190), the
   normal field of the reconstructions is the face field of the stage input
   (mhd_modify_wLR src/mhd/mod_mhd_phys.t:6338-6345), and the face field of
   wnew is advanced by the edge electric fields before the cell field is
   reset from it (:249, :323).  ws_block: the face field of ps (block, which
   the HLLD solver reads, :536-537). */
static void finite_volume(const orc_sim* s, int method, double qdt,
                          const double* wCT, double* wnew, double ch, int* crash,
                          double* bf, const double* ws_block) {
  const int nw = s->nw;
  const long n = s->ncell;
  double* wprim = (double*)malloc(sizeof(double) * (size_t)(nw * n));
  block_to_primitive(s, wCT, wprim, crash);
  const int ct = s->ct;
  const int tgc = ct == ORC_CT_HLL ? 2 : ct ? 1 : 0;
  const double* wsCT = wCT + (long)nw * n;
  ct_faces F = {NULL, NULL, NULL, NULL, NULL};
  if (ct) {
    F.fm = (double*)calloc((size_t)s->p.ndim * s->p.ndir * n, sizeof(double));
    if (ct == ORC_CT_CONTACT) F.vnorm = (double*)calloc((size_t)s->p.ndim * n, sizeof(double));
    if (ct == ORC_CT_HLL) {
      F.cbmin = (double*)calloc((size_t)s->p.ndim * n, sizeof(double));
      F.cbmax = (double*)calloc((size_t)s->p.ndim * n, sizeof(double));
      F.vbar = (double*)calloc((size_t)s->p.ndim * 2 * n, sizeof(double));
    }
  }
  for (int idim = 0; idim < s->p.ndim; ++idim) {
    int alo, ahi, blo, bhi, ta, tb;
    transverse_ranges(s, idim, &alo, &ahi, &blo, &bhi);
    trans_dims(idim, &ta, &tb);
    const int ea = ta < s->p.ndim ? tgc : 0, eb = tb < s->p.ndim ? tgc : 0;
    const int ilo = s->ng - 1, ihi = s->ng + s->bnx[idim] - 1; /* ixC */
    const double dxinv = -qdt / s->dx[idim];
    for (int b = blo - eb; b <= bhi + eb; ++b)
      for (int a = alo - ea; a <= ahi + ea; ++a) {
        const int inner = a >= alo && a <= ahi && b >= blo && b <= bhi;
        double lw[MAXW][MAXLINE];
        double wLp[MAXLINE][MAXW], wRp[MAXLINE][MAXW];
        double fC[MAXLINE][MAXW];
        get_line(s, wprim, nw, idim, a, b, lw);
        reconstruct_line(s, idim, lw, ilo, ihi, wLp, wRp);
        if (s->p.fix_small_values && s->p.limiter != ORC_LIM_WENO5)
          for (int i = ilo; i <= ihi; ++i)  /* 'reconstruct left / right' :1213-1216 */
            if (check_w_prim_fails(s, wLp[i]) || check_w_prim_fails(s, wRp[i])) *crash = 1;
        for (int i = ilo; i <= ihi; ++i) {
          double wLC[MAXW], wRC[MAXW], fL[MAXW], fR[MAXW];
          double cmax = 0.0, cmin = 0.0;
          const long cf = line_cell(s, idim, a, b, i);
          for (int iw = 0; iw < nw; ++iw) { wLC[iw] = wLp[i][iw]; wRC[iw] = wRp[i][iw]; }
          to_conserved_cell(s, wLC);
          to_conserved_cell(s, wRC);
          if (s->p.phys == ORC_PHYS_MHD && s->p.divb == ORC_DIVB_GLM)
            modify_wLR_glm(s, wLC, wRC, wLp[i], wRp[i], idim, ch);
          if (ct) {
            const double bn = wsCT[(long)idim * n + cf];
            const int bi = s->mag[idim];
            wLC[bi] = bn; wRC[bi] = bn; wLp[i][bi] = bn; wRp[i][bi] = bn;
          }
          get_flux(s, wLC, wLp[i], idim, ch, fL);
          get_flux(s, wRC, wRp[i], idim, ch, fR);
          if (method == ORC_FS_TVDLF) {
            get_cbounds(s, wLC, wRC, wLp[i], wRp[i], idim, 0, &cmax, &cmin, crash);
            riemann_tvdlf(s, wLC, wRC, fL, fR, cmax, fC[i]);
          } else {
            get_cbounds(s, wLC, wRC, wLp[i], wRp[i], idim, 1, &cmax, &cmin, crash);
            if (method == ORC_FS_HLLC)
              riemann_hllc(s, idim, wLC, wRC, fL, fR, cmax, cmin, fC[i]);
            else if (method == ORC_FS_HLLD)
              riemann_hlld(s, idim, wLC, wRC, wLp[i], wRp[i], fL, fR, cmax, cmin,
                           ct ? ws_block + (long)idim * n + cf : NULL, fC[i]);
            else
              riemann_hll(s, idim, wLC, wRC, fL, fR, cmax, cmin, fC[i]);
          }
          if (ct) {
            /* face data of the edge electric fields: unscaled fluxes of the
               field components and the ct velocities
               (mhd_get_ct_velocity_* src/mhd/mod_mhd_phys.t:3261-3326) */
            for (int m = 0; m < s->p.ndir; ++m)
              F.fm[((long)idim * s->p.ndir + m) * n + cf] = fC[i][s->mag[m]];
            if (ct == ORC_CT_CONTACT)
              F.vnorm[(long)idim * n + cf] = 0.5 * (wLp[i][s->mom[idim]] + wRp[i][s->mom[idim]]);
            if (ct == ORC_CT_HLL) {
              double cbmax = MAX2(cmax, zero), cbmin;
              cbmin = (method == ORC_FS_TVDLF) ? cbmax : MAX2(-cmin, zero);
              F.cbmin[(long)idim * n + cf] = cbmin;
              F.cbmax[(long)idim * n + cf] = cbmax;
              const int iN = (idim + 1) % s->p.ndir, iE = (idim + 2) % s->p.ndir;
              const double l1 = wLp[i][s->mom[iN]], r1 = wRp[i][s->mom[iN]], l2 = wLp[i][s->mom[iE]];
              F.vbar[((long)idim * 2 + 0) * n + cf] = (cbmax * l1 + cbmin * r1) / (cbmax + cbmin);
              /* the second average takes vbarRC(:,idim,1), as the reference does */
              F.vbar[((long)idim * 2 + 1) * n + cf] = (cbmax * l2 + cbmin * r1) / (cbmax + cbmin);
            }
          }
        }
        if (!inner) continue;
        /* flux-divergence update :269-284 */
        for (int iw = 0; iw < nw; ++iw) {
          for (int i = ilo; i <= ihi; ++i) fC[i][iw] = dxinv * fC[i][iw];
          if (bf) { /* the block-boundary faces, for store_flux */
            const long t = (long)(a - alo) + (long)(ahi - alo + 1) * (b - blo);
            bf[((long)(idim * 2 + 0) * nw + iw) * s->nface + t] = fC[ilo][iw];
            bf[((long)(idim * 2 + 1) * nw + iw) * s->nface + t] = fC[ihi][iw];
          }
          for (int i = ilo + 1; i <= ihi; ++i) {
            long c = line_cell(s, idim, a, b, i);
            wnew[iw * n + c] = wnew[iw * n + c] + fC[i][iw] - fC[i - 1][iw];
          }
        }
      }
  }
  if (ct) {
    ct_update_faces(s, qdt, wprim, &F, wsCT, wnew + (long)nw * n);
    ct_face_to_center(s, wnew);
    free(F.fm); free(F.vnorm); free(F.cbmin); free(F.cbmax); free(F.vbar);
  }
  /* check unphysical values (:325-333; small_values_method 'error' only
     sets crash, after which the run stops) */
  if (s->p.fix_small_values && small_values_interior(s, wnew)) *crash = 1;
  add_sources(s, qdt, wCT, wnew, ch);
  /* the Linde / GLM sources check again (src/mhd/mod_mhd_phys.t:5529, 5662) */
  if (s->p.fix_small_values && s->p.phys == ORC_PHYS_MHD && s->p.ndim > 1 &&
      (s->p.divb == ORC_DIVB_GLM || s->p.divb == ORC_DIVB_LINDE) && small_values_interior(s, wnew))
    *crash = 1;
  free(wprim);
}

/* hancock src/mod_finite_volume.t:19-113 */
static void hancock(const orc_sim* s, double qdt, const double* wCT,
                    double* wnew, double ch, int* crash) {
  const int nw = s->nw;
  const long n = s->ncell;
  double* wprim = (double*)malloc(sizeof(double) * (size_t)(nw * n));
  block_to_primitive(s, wCT, wprim, crash);
  for (int idim = 0; idim < s->p.ndim; ++idim) {
    int alo, ahi, blo, bhi;
    transverse_ranges(s, idim, &alo, &ahi, &blo, &bhi);
    const int olo = s->ng, ohi = s->ng + s->bnx[idim] - 1; /* ixO */
    const double dxinv = -qdt / s->dx[idim];
    for (int b = blo; b <= bhi; ++b)
      for (int a = alo; a <= ahi; ++a) {
        double lw[MAXW][MAXLINE];
        get_line(s, wprim, nw, idim, a, b, lw);
        for (int i = olo; i <= ohi; ++i) {
          /* wLp(i) = w(i) + ldw(i)/2 ; wRp(i-1) = w(i) - rdw(i)/2 */
          double wL[MAXW], wR[MAXW], wLC[MAXW], wRC[MAXW], fL[MAXW], fR[MAXW];
          for (int iw = 0; iw < nw; ++iw) {
            double ldw, rdw;
            double am = lw[iw][i] - lw[iw][i - 1];
            double bp = lw[iw][i + 1] - lw[iw][i];
            dwlimiter_cell(s, idim, am, bp, &ldw, &rdw);
            wL[iw] = lw[iw][i] + half * ldw;
            wR[iw] = lw[iw][i] - half * rdw;
          }
          if (s->p.fix_small_values && (check_w_prim_fails(s, wL) || check_w_prim_fails(s, wR))) *crash = 1;
          for (int iw = 0; iw < nw; ++iw) { wLC[iw] = wL[iw]; wRC[iw] = wR[iw]; }
          to_conserved_cell(s, wLC);
          to_conserved_cell(s, wRC);
          get_flux(s, wRC, wR, idim, ch, fR);
          get_flux(s, wLC, wL, idim, ch, fL);
          long c = line_cell(s, idim, a, b, i);
          for (int iw = 0; iw < nw; ++iw)
            wnew[iw * n + c] = wnew[iw * n + c] + dxinv * (fL[iw] - fR[iw]);
        }
      }
  }
  add_sources(s, qdt, wCT, wnew, ch);
  /* :109-111 */
  if (s->p.fix_small_values && small_values_interior(s, wnew)) *crash = 1;
  free(wprim);
}

/* ------------------------------------------------------------------ */
/* ghost cells                                                          */
/* ------------------------------------------------------------------ */

/* bc_fill_srl src/mod_ghostcells_update.t:812-852 (same-level copy) and
   fill_boundary_after_gc :579-615 -> bc_phys src/mod_boundary_conditions.t:12-120 */
static void getbc_phys(orc_sim* s, int slot, int ib);
static void getbc_block_srl(orc_sim* s, int slot, int ib);

static void getbc_block(orc_sim* s, int slot, int ib) {
  getbc_block_srl(s, slot, ib);
  getbc_phys(s, slot, ib);
}

/* bc_fill_srl: ghost cells from same-level neighbours */
static void getbc_block_srl(orc_sim* s, int slot, int ib) {
  const int nd = s->p.ndim, ng = s->ng;
  const long n = s->ncell;
  double* w = BLK(s, slot, ib);
  int dirs[3];
  for (int nidx = 0; nidx < s->nnb; ++nidx) {
    int t = nidx;
    int allzero = 1;
    for (int d = 0; d < 3; ++d) {
      if (d < nd) { dirs[d] = t % 3 - 1; t /= 3; } else dirs[d] = 0;
      if (dirs[d] != 0) allzero = 0;
    }
    if (allzero) continue;
    int nb = s->neighbor[(long)ib * s->nnb + nidx];
    if (nb < 0) continue;
    const double* wn = BLK(s, slot, nb);
    int lo[3], hi[3], sh[3];
    for (int d = 0; d < 3; ++d) {
      if (d >= nd) { lo[d] = 0; hi[d] = 0; sh[d] = 0; continue; }
      if (dirs[d] == -1) { lo[d] = 0; hi[d] = ng - 1; sh[d] = s->bnx[d]; }
      else if (dirs[d] == 0) { lo[d] = ng; hi[d] = ng + s->bnx[d] - 1; sh[d] = 0; }
      else { lo[d] = ng + s->bnx[d]; hi[d] = s->nx[d] - 1; sh[d] = -s->bnx[d]; }
    }
    for (int iw = 0; iw < s->nwt; ++iw)
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i)
            w[iw * n + IX(s, i, j, k)] = wn[iw * n + IX(s, i + sh[0], j + sh[1], k + sh[2])];
  }
}

/* fill_boundary_after_gc src/mod_ghostcells_update.t:579-615: bc_phys on
   every physical side, dimension by dimension */
static void getbc_phys(orc_sim* s, int slot, int ib) {
  const int nd = s->p.ndim, ng = s->ng;
  const long n = s->ncell;
  double* w = BLK(s, slot, ib);
  for (int idim = 0; idim < nd; ++idim) {
    int bmin[3], bmax[3];
    for (int d = 0; d < 3; ++d) {
      bmin[d] = 0;
      bmax[d] = (d < nd) ? s->nx[d] - 1 : 0;
      if (d < nd && d > idim) {
        int face_lo = 0, face_hi = 0;
        for (int side = 0; side < 2; ++side) {
          int nidx = 0, mul = 1;
          for (int dd = 0; dd < nd; ++dd) {
            int v = (dd == d) ? (side == 0 ? -1 : 1) : 0;
            nidx += (v + 1) * mul;
            mul *= 3;
          }
          int isb = s->ntype[(long)ib * s->nnb + nidx] == 1;
          if (side == 0) face_lo = isb; else face_hi = isb;
        }
        bmin[d] = face_lo ? ng : 0;
        bmax[d] = face_hi ? s->nx[d] - 1 - ng : s->nx[d] - 1;
      }
    }
    for (int side = 0; side < 2; ++side) {
      int nidx = 0, mul = 1;
      for (int dd = 0; dd < nd; ++dd) {
        int v = (dd == idim) ? (side == 0 ? -1 : 1) : 0;
        nidx += (v + 1) * mul;
        mul *= 3;
      }
      if (s->ntype[(long)ib * s->nnb + nidx] != 1) continue;
      int type = s->bc[2 * idim + side];
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) { lo[d] = bmin[d]; hi[d] = bmax[d]; }
      if (side == 0) { lo[idim] = 0; hi[idim] = ng - 1; }
      else { lo[idim] = s->nx[idim] - ng; hi[idim] = s->nx[idim] - 1; }
      for (int iw = 0; iw < s->nw; ++iw)
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              int c[3] = {i, j, k}, src[3] = {i, j, k};
              if (type == ORC_BC_CONT) {
                src[idim] = (side == 0) ? ng : s->nx[idim] - ng - 1;
              } else { /* symm / asymm: mirror */
                if (side == 0) src[idim] = 2 * ng - 1 - c[idim];
                else src[idim] = 2 * (s->nx[idim] - ng) - 1 - c[idim];
              }
              double v = w[iw * n + IX(s, src[0], src[1], src[2])];
              if (type == ORC_BC_ASYMM) v = -v;
              w[iw * n + IX(s, i, j, k)] = v;
            }
    }
  }
}


/* ------------------------------------------------------------------ */
/* AMR: ghost cells across level jumps, refluxing                       */
/* ------------------------------------------------------------------ */

/* Fortran-style 1-based cell access of a block array (nx cells per dim) */
static inline long IX1(const int* nx, int i, int j, int k) {
  return (long)(i - 1) + (long)nx[0] * ((long)(j - 1) + (long)nx[1] * (long)(k - 1));
}

/* index ranges of init_bc, src/mod_ghostcells_update.t:116-180, per
   dimension d (1-based; interpolation_order 2) */
typedef struct { int lo, hi; } orc_rng;

static orc_rng rng_R_srl(const orc_sim* s, int d, int i) {
  const int ng = s->ng, mlo = ng + 1, mhi = ng + s->bnx[d];
  orc_rng r;
  if (i == -1) { r.lo = 1; r.hi = ng; }
  else if (i == 0) { r.lo = mlo; r.hi = mhi; }
  else { r.lo = mhi + 1; r.hi = s->nx[d]; }
  return r;
}
static orc_rng rng_S_r(const orc_sim* s, int d, int i) {
  const int ng = s->ng, clo = ng + 1, chi = ng + s->bnx[d] / 2;
  orc_rng r;
  if (i == -1) { r.lo = clo; r.hi = clo - 1 + ng; }
  else if (i == 0) { r.lo = clo; r.hi = chi; }
  else { r.lo = chi + 1 - ng; r.hi = chi; }
  return r;
}
static orc_rng rng_R_r(const orc_sim* s, int d, int inc) {
  const int ng = s->ng, mlo = ng + 1, mhi = ng + s->bnx[d], nxco = s->bnx[d] / 2;
  orc_rng r;
  if (inc == 0) { r.lo = 1; r.hi = ng; }
  else if (inc == 1) { r.lo = mlo; r.hi = mlo - 1 + nxco; }
  else if (inc == 2) { r.lo = mlo + nxco; r.hi = mhi; }
  else { r.lo = mhi + 1; r.hi = s->nx[d]; }
  return r;
}
static orc_rng rng_S_p(const orc_sim* s, int d, int inc) {
  const int ng = s->ng, mlo = ng + 1, mhi = ng + s->bnx[d], nxco = s->bnx[d] / 2;
  const int ngco = (ng + 1) / 2;
  orc_rng r;
  if (inc == 0) { r.lo = mlo - 1; r.hi = mlo - 1 + ngco + 1; }
  else if (inc == 1) { r.lo = mlo - 1; r.hi = mlo - 1 + nxco + ngco + 1; }
  else if (inc == 2) { r.lo = mlo + nxco - ngco - 1; r.hi = mhi + 1; }
  else { r.lo = mhi + 1 - ngco - 1; r.hi = mhi + 1; }
  return r;
}
static orc_rng rng_R_p(const orc_sim* s, int d, int inc) {
  const int ng = s->ng, clo = ng + 1, chi = ng + s->bnx[d] / 2;
  const int ngco = (ng + 1) / 2;
  orc_rng r;
  if (inc == 0) { r.lo = clo - ngco - 1; r.hi = ng + 1; }
  else if (inc == 1) { r.lo = clo - 1; r.hi = chi + ngco + 1; }
  else if (inc == 2) { r.lo = clo - ngco - 1; r.hi = chi + 1; }
  else { r.lo = chi + 1 - 1; r.hi = chi + ngco + 1; }
  return r;
}

static int dir_nidx(const orc_sim* s, const int* i) {
  int k = 0, m = 1;
  for (int d = 0; d < s->p.ndim; ++d) { k += (i[d] + 1) * m; m *= 3; }
  return k;
}

static double* PSC(const orc_sim* s, int ib) { return s->psc + (long)ib * s->nw * s->ncellc; }

/* coarsen_grid src/amr/mod_coarsen.t:12-79 (slab_uniform): mean of the
   2^ndim fine cells, summed in array-element order; fine interior ixM of
   wFi (layout nxf) into coarse cells [clo, chi] of wCo (layout nxc), fine
   cell of coarse cell c: 2*(c - clo) + flo */
static void coarsen_cells(const orc_sim* s, const double* wFi, const int* nxf, long ncf, int flo_[3],
                          double* wCo, const int* nxc, long ncc, const int* clo, const int* chi) {
  const int nd = s->p.ndim;
  const double cofi = one / (double)(1 << nd);
  for (int iw = 0; iw < s->nw; ++iw)
    for (int kc = clo[2]; kc <= chi[2]; ++kc)
      for (int jc = clo[1]; jc <= chi[1]; ++jc)
        for (int ic = clo[0]; ic <= chi[0]; ++ic) {
          int f[3] = {2 * (ic - clo[0]) + flo_[0], 2 * (jc - clo[1]) + flo_[1], 2 * (kc - clo[2]) + flo_[2]};
          double sum = 0.0;
          int first = 1;
          for (int dk = 0; dk <= (nd > 2); ++dk)
            for (int dj = 0; dj <= (nd > 1); ++dj)
              for (int di = 0; di <= 1; ++di) {
                double v = wFi[iw * ncf + IX1(nxf, f[0] + di, f[1] + dj, f[2] + dk)];
                if (first) { sum = v; first = 0; } else sum = sum + v;
              }
          wCo[iw * ncc + IX1(nxc, ic, jc, kc)] = sum * cofi;
        }
}

/* bc_phys src/mod_boundary_conditions.t:12-120 on an arbitrary array
   (layout nxa), side iside (1 low, 2 high) of dimension idim over the
   1-based range [blo, bhi] */
static void bc_phys_range(const orc_sim* s, double* w, const int* nxa, long nca, int iside, int idim,
                          const int* blo, const int* bhi) {
  const int ng = s->ng;
  int olo[3], ohi[3];
  for (int d = 0; d < 3; ++d) { olo[d] = blo[d]; ohi[d] = bhi[d]; }
  if (iside == 2) olo[idim] = bhi[idim] + 1 - ng;
  else ohi[idim] = blo[idim] - 1 + ng;
  const int type = s->bc[2 * idim + (iside - 1)];
  if (type == ORC_BC_PERIODIC) return;
  for (int iw = 0; iw < s->nw; ++iw)
    for (int k = olo[2]; k <= ohi[2]; ++k)
      for (int j = olo[1]; j <= ohi[1]; ++j)
        for (int i = olo[0]; i <= ohi[0]; ++i) {
          int c[3] = {i, j, k}, src[3] = {i, j, k};
          if (type == ORC_BC_CONT) {
            src[idim] = (iside == 2) ? olo[idim] - 1 : ohi[idim] + 1;
          } else if (iside == 2) {
            src[idim] = 2 * olo[idim] - 1 - c[idim]; /* olo-1, olo-2, ... */
          } else {
            src[idim] = 2 * ohi[idim] + 1 - c[idim];
          }
          double v = w[iw * nca + IX1(nxa, src[0], src[1], src[2])];
          if (type == ORC_BC_ASYMM) v = -v;
          w[iw * nca + IX1(nxa, c[0], c[1], c[2])] = v;
        }
}

/* interpolation_linear src/mod_ghostcells_update.t:1381-1525 (slab_uniform,
   conserved variables): fine ghost cells [flo, fhi] of block ib from its
   coarse representation */
static void interpolation_linear(orc_sim* s, int slot, int ib, const int* flo, const int* fhi) {
  const int nd = s->p.ndim, ng = s->ng;
  double* w = BLK(s, slot, ib);
  const double* pc = PSC(s, ib);
  double dxFi[3], dxCo[3], invdxCo[3], xFimin[3], xComin[3];
  for (int d = 0; d < nd; ++d) {
    dxFi[d] = s->bdx[ib * 3 + d];
    dxCo[d] = two * dxFi[d];
    invdxCo[d] = 1.0 / dxCo[d];
    const double rpxmin = s->xlohi[((long)ib * nd + d) * 2];
    xFimin[d] = rpxmin - (double)ng * dxFi[d];
    xComin[d] = rpxmin - (double)ng * dxCo[d];
  }
  for (int k = flo[2]; k <= fhi[2]; ++k)
    for (int j = flo[1]; j <= fhi[1]; ++j)
      for (int i = flo[0]; i <= fhi[0]; ++i) {
        int f[3] = {i, j, k}, co[3] = {1, 1, 1};
        double eta[3] = {0, 0, 0};
        for (int d = 0; d < nd; ++d) {
          const double xFi = xFimin[d] + ((double)f[d] - half) * dxFi[d];
          co[d] = (int)((xFi - xComin[d]) * invdxCo[d]) + 1;
          const double xCo = xComin[d] + ((double)co[d] - half) * dxCo[d];
          eta[d] = (xFi - xCo) * invdxCo[d];
        }
        for (int iw = 0; iw < s->nw; ++iw) {
          const double* q = pc + iw * s->ncellc;
          const double c0 = q[IX1(s->nxc, co[0], co[1], co[2])];
          double val = c0;
          for (int d = 0; d < nd; ++d) {
            int h[3] = {co[0], co[1], co[2]}, p[3] = {co[0], co[1], co[2]};
            h[d] -= 1; p[d] += 1;
            const double slopeL = c0 - q[IX1(s->nxc, h[0], h[1], h[2])];
            const double slopeR = q[IX1(s->nxc, p[0], p[1], p[2])] - c0;
            const double slopeC = half * (slopeR + slopeL);
            const double signC = copysign(one, slopeC);
            const double slope = signC * MAX2(zero, MIN2(MIN2(fabs(slopeC), signC * slopeL), signC * slopeR));
            val = val + slope * eta[d];
          }
          w[iw * s->ncell + IX1(s->nx, f[0], f[1], f[2])] = val;
        }
      }
}

/* bc_prolong src/mod_ghostcells_update.t:1261-1339 (no stretching, no
   primitive prolongation) */
static void bc_prolong(orc_sim* s, int slot, int ib, const int* i) {
  const int nd = s->p.ndim, ng = s->ng;
  int flo[3] = {1, 1, 1}, fhi[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) {
    orc_rng r = rng_R_srl(s, d, i[d]);
    flo[d] = r.lo; fhi[d] = r.hi;
  }
  int phybound = 0;
  for (int k = 0; k < s->nnb; ++k)
    if (s->ntype[(long)ib * s->nnb + k] == 1) phybound = 1;
  if (phybound) {
    double dxFi[3], invdxCo[3], xFimin[3], xComin[3];
    for (int d = 0; d < nd; ++d) {
      dxFi[d] = s->bdx[ib * 3 + d];
      const double dxCo = two * dxFi[d];
      invdxCo[d] = 1.0 / dxCo;
      const double rpxmin = s->xlohi[((long)ib * nd + d) * 2];
      xFimin[d] = rpxmin - (double)ng * dxFi[d];
      xComin[d] = rpxmin - (double)ng * dxCo;
    }
    for (int idims = 0; idims < nd; ++idims) {
      int clo[3] = {1, 1, 1}, chi[3] = {1, 1, 1};
      for (int d = 0; d < nd; ++d) {
        clo[d] = (int)((xFimin[d] + ((double)flo[d] - half) * dxFi[d] - xComin[d]) * invdxCo[d]) + 1 - 1;
        chi[d] = (int)((xFimin[d] + ((double)fhi[d] - half) * dxFi[d] - xComin[d]) * invdxCo[d]) + 1 + 1;
      }
      if (nd == 3) {
        /* avoid undetermined ghost cells at physical boundary edges */
        const int cmlo = ng + 1, cmhi = ng + s->bnx[0] / 2;
        int face[3][2];
        for (int d = 0; d < 3; ++d)
          for (int sd = 0; sd < 2; ++sd) {
            int v[3] = {0, 0, 0};
            v[d] = sd == 0 ? -1 : 1;
            face[d][sd] = s->ntype[(long)ib * s->nnb + dir_nidx(s, v)] == 1;
          }
        if (idims == 0 && (face[0][0] || face[0][1])) {
          if (face[1][0]) clo[1] = cmlo;
          if (face[2][0]) clo[2] = cmlo;
          if (face[1][1]) chi[1] = ng + s->bnx[1] / 2;
          if (face[2][1]) chi[2] = ng + s->bnx[2] / 2;
        } else if (idims == 1 && (face[1][0] || face[1][1])) {
          if (face[2][0]) clo[2] = cmlo;
          if (face[2][1]) chi[2] = ng + s->bnx[2] / 2;
        }
        (void)cmhi;
      }
      for (int iside = 1; iside <= 2; ++iside) {
        int v[3] = {0, 0, 0};
        v[idims] = 2 * iside - 3;
        if (s->ntype[(long)ib * s->nnb + dir_nidx(s, v)] != 1) continue;
        if ((iside == 1 && clo[idims] < 1 + ng) || (iside == 2 && chi[idims] > s->nxc[idims] - ng)) {
          int blo[3], bhi[3];
          for (int d = 0; d < 3; ++d) { blo[d] = clo[d]; bhi[d] = chi[d]; }
          blo[idims] = 1;
          bhi[idims] = s->nxc[idims];
          bc_phys_range(s, PSC(s, ib), s->nxc, s->ncellc, iside, idims, blo, bhi);
        }
      }
    }
  }
  interpolation_linear(s, slot, ib, flo, fhi);
}

/* getbc src/mod_ghostcells_update.t:350-566 with level jumps, one rank:
   coarse representations, sibling copies and restriction to coarser
   neighbours, coarse values for finer neighbours, prolongation, then the
   physical boundaries */
static void getbc_amr(orc_sim* s, int slot) {
  const int nd = s->p.ndim, ng = s->ng;
  const int nb = s->nb;
  /* coarse values of blocks with a coarser neighbour */
  for (int ib = 0; ib < nb; ++ib) {
    int has_coarse = 0;
    for (int k = 0; k < s->nnb; ++k)
      if (s->ntype[(long)ib * s->nnb + k] == 2) has_coarse = 1;
    if (!has_coarse) continue;
    int flo[3] = {1, 1, 1}, clo[3] = {1, 1, 1}, chi[3] = {1, 1, 1};
    for (int d = 0; d < nd; ++d) { flo[d] = ng + 1; clo[d] = ng + 1; chi[d] = ng + s->bnx[d] / 2; }
    coarsen_cells(s, BLK(s, slot, ib), s->nx, s->ncell, flo, PSC(s, ib), s->nxc, s->ncellc, clo, chi);
  }
  /* sibling copies (bc_fill_srl) and restriction (bc_fill_restrict) */
  for (int ib = 0; ib < nb; ++ib) getbc_block_srl(s, slot, ib);
  for (int ib = 0; ib < nb; ++ib) {
    for (int k = 0; k < s->nnb; ++k) {
      if (s->ntype[(long)ib * s->nnb + k] != 2) continue;
      int i[3] = {0, 0, 0}, t = k;
      for (int d = 0; d < nd; ++d) { i[d] = t % 3 - 1; t /= 3; }
      const int* ic = s->icpos + (long)ib * nd;
      int ok = 1;
      for (int d = 0; d < nd; ++d)
        if (!(i[d] == 0 || i[d] == 2 * ic[d] - 3)) ok = 0;
      if (!ok) continue;
      const int C = s->nbr[(long)ib * s->nnb + k];
      orc_rng S[3], R[3];
      for (int d = 0; d < 3; ++d) {
        if (d < nd) { S[d] = rng_S_r(s, d, i[d]); R[d] = rng_R_r(s, d, -2 * i[d] + ic[d]); }
        else { S[d].lo = S[d].hi = 1; R[d].lo = R[d].hi = 1; }
      }
      const double* src = PSC(s, ib);
      double* dst = BLK(s, slot, C);
      for (int iw = 0; iw < s->nw; ++iw)
        for (int kk = 0; kk <= S[2].hi - S[2].lo; ++kk)
          for (int jj = 0; jj <= S[1].hi - S[1].lo; ++jj)
            for (int ii = 0; ii <= S[0].hi - S[0].lo; ++ii)
              dst[iw * s->ncell + IX1(s->nx, R[0].lo + ii, R[1].lo + jj, R[2].lo + kk)] =
                  src[iw * s->ncellc + IX1(s->nxc, S[0].lo + ii, S[1].lo + jj, S[2].lo + kk)];
    }
  }
  /* coarse values for finer neighbours (bc_fill_prolong) */
  for (int ib = 0; ib < nb; ++ib) {
    for (int k = 0; k < s->nnb; ++k) {
      if (s->ntype[(long)ib * s->nnb + k] != 4) continue;
      int i[3] = {0, 0, 0}, t = k;
      for (int d = 0; d < nd; ++d) { i[d] = t % 3 - 1; t /= 3; }
      int iclo[3] = {1, 1, 1}, ichi[3] = {1, 1, 1};
      for (int d = 0; d < nd; ++d) {
        iclo[d] = i[d] == -1 ? 2 : 1;
        ichi[d] = i[d] == 1 ? 1 : 2;
      }
      for (int c2 = iclo[2]; c2 <= ichi[2]; ++c2)
        for (int c1 = iclo[1]; c1 <= ichi[1]; ++c1)
          for (int c0 = iclo[0]; c0 <= ichi[0]; ++c0) {
            int ic[3] = {c0, c1, c2};
            int cidx = 0;
            for (int d = 0; d < nd; ++d) cidx += (ic[d] - 1) << d;
            const int F = s->nchild[((long)ib * s->nnb + k) * s->nch + cidx];
            orc_rng S[3], R[3];
            for (int d = 0; d < 3; ++d) {
              if (d < nd) { S[d] = rng_S_p(s, d, 2 * i[d] + ic[d]); R[d] = rng_R_p(s, d, ic[d] - i[d]); }
              else { S[d].lo = S[d].hi = 1; R[d].lo = R[d].hi = 1; }
            }
            const double* src = BLK(s, slot, ib);
            double* dst = PSC(s, F);
            for (int iw = 0; iw < s->nw; ++iw)
              for (int kk = 0; kk <= S[2].hi - S[2].lo; ++kk)
                for (int jj = 0; jj <= S[1].hi - S[1].lo; ++jj)
                  for (int ii = 0; ii <= S[0].hi - S[0].lo; ++ii)
                    dst[iw * s->ncellc + IX1(s->nxc, R[0].lo + ii, R[1].lo + jj, R[2].lo + kk)] =
                        src[iw * s->ncell + IX1(s->nx, S[0].lo + ii, S[1].lo + jj, S[2].lo + kk)];
          }
    }
  }
  /* prolongation (gc_prolong: directions with dimension 1 fastest) */
  for (int ib = 0; ib < nb; ++ib)
    for (int k = 0; k < s->nnb; ++k) {
      if (s->ntype[(long)ib * s->nnb + k] != 2) continue;
      int i[3] = {0, 0, 0}, t = k;
      for (int d = 0; d < nd; ++d) { i[d] = t % 3 - 1; t /= 3; }
      bc_prolong(s, slot, ib, i);
    }
  /* physical boundaries */
  for (int ib = 0; ib < nb; ++ib) getbc_phys(s, slot, ib);
}

/* fix_conserve src/mod_fix_conserve.t:465-595 with store_flux :597-648
   (slab_uniform, one rank): the coarse cells next to a finer neighbour get
   the sum of the fine fluxes instead of their own */
static void fix_conserve(orc_sim* s, int slot) {
  const int nd = s->p.ndim, ng = s->ng, nw = s->nw;
  const double cofi = one / (double)(1 << nd);
  const long bstride = (long)nd * 2 * nw * s->nface;
  for (int ib = 0; ib < s->nb; ++ib) {
    double* w = BLK(s, slot, ib);
    for (int idims = 0; idims < nd; ++idims) {
      int ta = -1, tb = -1;
      for (int d = 0; d < nd; ++d) {
        if (d == idims) continue;
        if (ta < 0) ta = d; else tb = d;
      }
      const int na = ta >= 0 ? s->bnx[ta] : 1, nbt = tb >= 0 ? s->bnx[tb] : 1;
      for (int iside = 1; iside <= 2; ++iside) {
        int v[3] = {0, 0, 0};
        v[idims] = 2 * iside - 3;
        const int k = dir_nidx(s, v);
        if (s->ntype[(long)ib * s->nnb + k] != 4) continue;
        const int ix = iside == 1 ? ng + 1 : ng + s->bnx[idims];
        /* own flux: pflux(1) = -fC(low face), pflux(2) = fC(high face) */
        const double* own = s->bflux + ib * bstride + (long)(idims * 2 + (iside - 1)) * nw * s->nface;
        for (int iw = 0; iw < nw; ++iw)
          for (int b2 = 0; b2 < nbt; ++b2)
            for (int a2 = 0; a2 < na; ++a2) {
              int c[3] = {1, 1, 1};
              c[idims] = ix;
              if (ta >= 0) c[ta] = ng + 1 + a2;
              if (tb >= 0) c[tb] = ng + 1 + b2;
              const double f = own[(long)iw * s->nface + a2 + (long)na * b2];
              const double pflux = iside == 1 ? -f : f;
              long cc = iw * s->ncell + IX1(s->nx, c[0], c[1], c[2]);
              w[cc] = w[cc] - pflux;
            }
        /* children across the face */
        int iclo[3] = {1, 1, 1}, ichi[3] = {1, 1, 1};
        for (int d = 0; d < nd; ++d) {
          iclo[d] = v[d] == -1 ? 2 : 1;
          ichi[d] = v[d] == 1 ? 1 : 2;
        }
        const int iother = 3 - iside;
        for (int c2 = iclo[2]; c2 <= ichi[2]; ++c2)
          for (int c1 = iclo[1]; c1 <= ichi[1]; ++c1)
            for (int c0 = iclo[0]; c0 <= ichi[0]; ++c0) {
              int ic[3] = {c0, c1, c2};
              int cidx = 0;
              for (int d = 0; d < nd; ++d) cidx += (ic[d] - 1) << d;
              const int F = s->nchild[((long)ib * s->nnb + k) * s->nch + cidx];
              const double* ff = s->bflux + F * bstride + (long)(idims * 2 + (iother - 1)) * nw * s->nface;
              const int nca = na / 2, ncb = tb >= 0 ? nbt / 2 : 1;
              for (int iw = 0; iw < nw; ++iw)
                for (int b2 = 0; b2 < ncb; ++b2)
                  for (int a2 = 0; a2 < nca; ++a2) {
                    /* store_flux of the fine block: sum over the 2 (2x2)
                       fine faces in array-element order */
                    const int fa = 2 * a2, fb = tb >= 0 ? 2 * b2 : 0;
                    const double* q = ff + (long)iw * s->nface;
                    double sum = q[fa + (long)na * fb];
                    sum = sum + q[fa + 1 + (long)na * fb];
                    if (tb >= 0) {
                      sum = sum + q[fa + (long)na * (fb + 1)];
                      sum = sum + q[fa + 1 + (long)na * (fb + 1)];
                    }
                    const double pf = iother == 1 ? sum : -sum;
                    int c[3] = {1, 1, 1};
                    c[idims] = ix;
                    if (ta >= 0) c[ta] = ng + 1 + (ic[ta] - 1) * nca + a2;
                    if (tb >= 0) c[tb] = ng + 1 + (ic[tb] - 1) * ncb + b2;
                    long cc = iw * s->ncell + IX1(s->nx, c[0], c[1], c[2]);
                    w[cc] = w[cc] + pf * cofi;
                  }
            }
      }
    }
  }
}

int orc_getbc(orc_sim* s, int slot) {
  if (s->amr) {
    getbc_amr(s, slot);
    return 0;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int ib = 0; ib < s->nb; ++ib) getbc_block(s, slot, ib);
  return 0;
}

/* ------------------------------------------------------------------ */
/* public API                                                          */
/* ------------------------------------------------------------------ */

int orc_create(const orc_params* p, int nblocks, const double* dx,
               const int* neighbor, const double* xlohi, const int* bc,
               orc_sim** out) {
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  s->cur_ib = -1;
  if (!s) return 1;
  s->p = *p;
  if (p->ndim < 1 || p->ndim > 3 || p->ndir < p->ndim || p->ndir > 3) { free(s); return 2; }
  s->nw = orc_nw(p);
  if (s->nw > MAXW) { free(s); return 3; }
  s->nws = 0;
  s->ct = 0;
  if (p->phys == ORC_PHYS_MHD && p->divb == ORC_DIVB_CT) {
    /* constrained transport: uniform periodic 2-D and 3-D meshes (the 2-D
       form is checked against the reference's logs, the 3-D one against
       the 2-D one on z-invariant states and by axis permutation) */
    if (p->ndim < 2 || p->type_ct < ORC_CT_AVERAGE || p->type_ct > ORC_CT_HLL) { free(s); return 6; }
    for (int i = 0; i < 2 * p->ndim; ++i)
      if (bc[i] != ORC_BC_PERIODIC) { free(s); return 6; }
    s->nws = p->ndim;
    s->ct = p->type_ct;
  }
  s->nwt = s->nw + s->nws;
  s->rho_ = 0;
  for (int c = 0; c < p->ndir; ++c) s->mom[c] = 1 + c;
  s->e_ = p->ndir + 1;
  if (p->phys == ORC_PHYS_MHD) {
    for (int c = 0; c < p->ndir; ++c) s->mag[c] = p->ndir + 2 + c;
    s->psi_ = (p->divb == ORC_DIVB_GLM) ? 2 * p->ndir + 2 : -1;
  }
  s->gamma_1 = p->gamma - 1.0;
  s->inv_gamma_1 = 1.0 / (p->gamma - 1.0);
  /* flux_type table: src/mhd/mod_mhd_phys.t:565-589 */
  for (int d = 0; d < 3; ++d)
    for (int iw = 0; iw < MAXW; ++iw) s->ftype[d][iw] = FT_DEFAULT;
  if (p->phys == ORC_PHYS_MHD && p->ndim > 1) {
    if (p->divb == ORC_DIVB_GLM) {
      for (int d = 0; d < p->ndir; ++d) {
        s->ftype[d][s->psi_] = FT_SPECIAL;
        s->ftype[d][s->mag[d]] = FT_SPECIAL;
      }
    } else {
      for (int d = 0; d < p->ndir; ++d) s->ftype[d][s->mag[d]] = FT_TVDLF;
    }
  }
  s->ng = p->nghost;
  s->ncell = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < p->ndim) {
      s->bnx[d] = p->block_nx[d];
      s->nx[d] = p->block_nx[d] + 2 * p->nghost;
      s->dx[d] = dx[d];
    } else {
      s->bnx[d] = 1;
      s->nx[d] = 1;
      s->dx[d] = 1.0;
    }
    if (s->nx[d] > MAXLINE) { free(s); return 4; }
    s->ncell *= s->nx[d];
  }
  s->nb = nblocks;
  s->nnb = 1;
  for (int d = 0; d < p->ndim; ++d) s->nnb *= 3;
  s->neighbor = (int*)malloc(sizeof(int) * (size_t)nblocks * s->nnb);
  memcpy(s->neighbor, neighbor, sizeof(int) * (size_t)nblocks * s->nnb);
  s->xlohi = (double*)malloc(sizeof(double) * (size_t)nblocks * 2 * p->ndim);
  memcpy(s->xlohi, xlohi, sizeof(double) * (size_t)nblocks * 2 * p->ndim);
  s->bdx = (double*)malloc(sizeof(double) * (size_t)nblocks * 3);
  s->ntype = (int*)malloc(sizeof(int) * (size_t)nblocks * s->nnb);
  s->nbr = (int*)malloc(sizeof(int) * (size_t)nblocks * s->nnb);
  for (int ib = 0; ib < nblocks; ++ib) {
    for (int d = 0; d < 3; ++d) s->bdx[ib * 3 + d] = s->dx[d];
    for (int k = 0; k < s->nnb; ++k) {
      const int nb = neighbor[(long)ib * s->nnb + k];
      s->nbr[(long)ib * s->nnb + k] = nb;
      s->ntype[(long)ib * s->nnb + k] = (k == (s->nnb - 1) / 2) ? 0 : (nb >= 0 ? 3 : 1);
    }
  }
  s->nch = 1 << p->ndim;
  s->nface = 1;
  for (int d = 1; d < p->ndim; ++d) s->nface *= p->block_nx[d] > p->block_nx[0] ? p->block_nx[d] : p->block_nx[0];
  for (int i = 0; i < 6; ++i) s->bc[i] = (i < 2 * p->ndim) ? bc[i] : ORC_BC_PERIODIC;
  for (int sl = 0; sl < 4; ++sl) {
    s->w[sl] = (double*)calloc((size_t)nblocks * s->nwt * s->ncell, sizeof(double));
    if (!s->w[sl]) { orc_destroy(s); return 5; }
  }
  *out = s;
  return 0;
}

void orc_destroy(orc_sim* s) {
  if (!s) return;
  for (int sl = 0; sl < 4; ++sl) free(s->w[sl]);
  free(s->neighbor);
  free(s->xlohi);
  free(s->bdx);
  free(s->ntype);
  free(s->nbr);
  free(s->nchild);
  free(s->icpos);
  free(s->psc);
  free(s->bflux);
  free(s);
}

int orc_set_w(orc_sim* s, int slot, const double* w) {
  if (slot < 0 || slot > 3) return 1;
  memcpy(s->w[slot], w, sizeof(double) * (size_t)s->nb * s->nwt * s->ncell);
  return 0;
}

int orc_get_w(const orc_sim* s, int slot, double* w) {
  if (slot < 0 || slot > 3) return 1;
  memcpy(w, s->w[slot], sizeof(double) * (size_t)s->nb * s->nwt * s->ncell);
  return 0;
}

/* src/amr/mod_amr_solution_node.t:176-186 */
int orc_get_x(const orc_sim* s, double* x) {
  const long n = s->ncell;
  for (int ib = 0; ib < s->nb; ++ib) {
    double* xb = x + (long)ib * s->p.ndim * n;
    for (int d = 0; d < s->p.ndim; ++d) {
      double xlo = s->xlohi[((long)ib * s->p.ndim + d) * 2];
      double xhi = s->xlohi[((long)ib * s->p.ndim + d) * 2 + 1];
      for (int k = 0; k < s->nx[2]; ++k)
        for (int j = 0; j < s->nx[1]; ++j)
          for (int i = 0; i < s->nx[0]; ++i) {
            int c[3] = {i, j, k};
            int ix = c[d] + 1; /* Fortran index */
            double v;
            const double dxb = s->bdx[ib * 3 + d];
            if (ix <= s->bnx[d])
              v = xlo + ((double)(ix - s->ng) - half) * dxb;
            else
              v = xhi + ((double)(ix - (s->bnx[d] + s->ng)) - half) * dxb;
            xb[d * n + IX(s, i, j, k)] = v;
          }
    }
  }
  return 0;
}

/* getdt_courant, maxsum + slab_uniform branch, src/mod_dt.t:160-287 */
int orc_getdt_courant(orc_sim* s, int slot, double courantpar,
                      double* dtmin_out, double* cmax_out) {
  double dtmin = bigdouble, cmax_all = 0.0;
  int crash_dt = 0;
#pragma omp parallel for schedule(static) reduction(min : dtmin) reduction(max : cmax_all) reduction(| : crash_dt) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int ib = 0; ib < s->nb; ++ib) {
    const double* w = BLK(s, slot, ib);
    const long n = s->ncell;
    double dtnew = bigdouble;
    double maxtot = -bigdouble, cmaxb = 0.0;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = (d < s->p.ndim) ? s->ng : 0;
      hi[d] = (d < s->p.ndim) ? s->ng + s->bnx[d] - 1 : 0;
    }
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          double cell[MAXW];
          long c = IX(s, i, j, k);
          for (int iw = 0; iw < s->nw; ++iw) cell[iw] = w[iw * n + c];
          /* getdt_courant converts with phys_to_primitive (src/mod_dt.t:195) */
          if (s->p.fix_small_values && small_cons_fails(s, cell)) crash_dt = 1;
          to_primitive_cell(s, cell);
          double tot = 0.0;
          for (int d = 0; d < s->p.ndim; ++d) {
            double cm = get_cmax(s, cell, d);
            cmaxb = MAX2(cmaxb, cm);
            double dxinv = one / s->bdx[ib * 3 + d];
            if (d == 0) tot = cm * dxinv;
            else tot = tot + cm * dxinv;
          }
          maxtot = MAX2(maxtot, tot);
        }
    if (maxtot > smalldouble) dtnew = MIN2(dtnew, courantpar / maxtot);
    dtmin = MIN2(dtmin, dtnew);
    cmax_all = MAX2(cmax_all, cmaxb);
  }
  if (crash_dt) s->crash = 1;
  *dtmin_out = dtmin;
  *cmax_out = cmax_all;
  return 0;
}

static void fix_conserve(orc_sim* s, int slot);

int orc_advect1(orc_sim* s, int method, double dtfactor, double dt,
                double cmax_global, int psa, int psb) {
  if (psa < 0 || psa > 3 || psb < 0 || psb > 3 || psa == psb) return 1;
  /* with constrained transport hancock advances the cells only; the face
     field of psb stays as the stage's initial combination left it */
  const double qdt = dtfactor * dt;
  int crash = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : crash) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int ib = 0; ib < s->nb; ++ib) {
    const double* wCT = BLK(s, psa, ib);
    double* wnew = BLK(s, psb, ib);
    int c = 0;
    /* the block's own cell size (dxlevel of its level) */
    orc_sim sb = *s;
    for (int d = 0; d < 3; ++d) sb.dx[d] = s->bdx[ib * 3 + d];
    sb.cur_ib = ib;
    double* bf = (s->amr && s->fix_conserve && method != ORC_FS_HANCOCK)
                     ? s->bflux + (long)ib * s->p.ndim * 2 * s->nw * s->nface
                     : NULL;  /* hancock stores no fluxes */
    if (method == ORC_FS_HANCOCK)
      hancock(&sb, qdt, wCT, wnew, cmax_global, &c);
    else
      finite_volume(&sb, method, qdt, wCT, wnew, cmax_global, &c, bf,
                    s->ct ? BLK(s, 0, ib) + (long)s->nw * s->ncell : NULL);
    crash |= c;
  }
  if (crash) s->crash = 1;
  if (s->amr && s->fix_conserve) fix_conserve(s, psb);
  orc_getbc(s, psb);
  return 0;
}

void orc_set_fix_conserve(orc_sim* s, int on) { s->fix_conserve = on; }

int orc_axpby(orc_sim* s, int dst, double a, int x, double b, int y) {
  const long tot = (long)s->nb * s->nwt * s->ncell;
  double* d = s->w[dst];
  const double* xv = s->w[x];
  const double* yv = s->w[y];
#pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (long i = 0; i < tot; ++i) d[i] = a * xv[i] + b * yv[i];
  return 0;
}

/* ps(dst) = a state combination given as a postfix program, over every
   cell of every block, one rounding per operation: the update loops of
   advect between the stages (src/mod_advance.t:113-133, 269-291, 485-507,
   515-557).  ops: 0..3 push slot ps..ps3, 8 push the next constant, 9 add,
   10 subtract, 11 multiply, 12 divide, 13 negate; the program is checked by
   the caller (at most 8 deep). */
int orc_combine(orc_sim* s, int dst, int nops, const int* ops, const double* k) {
  const long tot = (long)s->nb * s->nwt * s->ncell;
  double* d = s->w[dst];
#pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (long i = 0; i < tot; ++i) {
    double st[8];
    int sp = 0, kc = 0;
    for (int q = 0; q < nops; ++q) {
      const int op = ops[q];
      if (op >= 0 && op <= 3) {
        st[sp++] = s->w[op][i];
      } else if (op == 8) {
        st[sp++] = k[kc++];
      } else if (op == 13) {
        st[sp - 1] = -st[sp - 1];
      } else {
        const double a = st[sp - 2], b = st[sp - 1];
        --sp;
        st[sp - 1] = op == 9 ? a + b : op == 10 ? a - b : op == 11 ? a * b : a / b;
      }
    }
    d[i] = st[0];
  }
  return 0;
}

int orc_copy(orc_sim* s, int dst, int src) {
  memcpy(s->w[dst], s->w[src], sizeof(double) * (size_t)s->nb * s->nwt * s->ncell);
  return 0;
}

/* get_volume_average src/io/mod_input_output.t:2694-2729 (slab uniform:
   dvolume = prod(dx)) */
int orc_volume_average(const orc_sim* s, int slot, double* modes) {
  const long n = s->ncell;
  for (int power = 1; power <= 2; ++power) {
    double wsum[MAXW + 1];
    for (int iw = 0; iw <= s->nw; ++iw) wsum[iw] = 0.0;
    for (int ib = 0; ib < s->nb; ++ib) {
      const double* w = BLK(s, slot, ib);
      double dvol = 1.0;
      for (int d = 0; d < s->p.ndim; ++d) dvol = dvol * s->bdx[ib * 3 + d];
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = (d < s->p.ndim) ? s->ng : 0;
        hi[d] = (d < s->p.ndim) ? s->ng + s->bnx[d] - 1 : 0;
      }
      double vs = 0.0;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) vs = vs + dvol;
      wsum[s->nw] = wsum[s->nw] + vs;
      for (int iw = 0; iw < s->nw; ++iw) {
        double acc = 0.0;
        for (int k = lo[2]; k <= hi[2]; ++k)
          for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
              double v = w[iw * n + IX(s, i, j, k)];
              acc = acc + dvol * (power == 1 ? v : v * v);
            }
        wsum[iw] = wsum[iw] + acc;
      }
    }
    for (int iw = 0; iw < s->nw; ++iw)
      modes[(power - 1) * s->nw + iw] = wsum[iw] / wsum[s->nw];
  }
  return 0;
}

int orc_crash_flag(const orc_sim* s) { return s->crash; }

/* a simulation over a block-structured AMR level set: dxb[nb][ndim] cell
   size per block, ntype/nbr[nb][3^ndim] neighbour type and sibling / coarse
   leaf, nchild[nb][3^ndim][2^ndim] children of fine neighbours,
   icpos[nb][ndim] position within the parent */
int orc_create_amr(const orc_params* p, int nblocks, const double* dxb, const int* ntype,
                   const int* nbr, const int* nchild, const int* icpos, const double* xlohi,
                   const int* bc, orc_sim** out) {
  if (p->phys == ORC_PHYS_MHD && p->divb == ORC_DIVB_CT) return 6; /* no staggered AMR */
  int nnb = 1, nch = 1;
  for (int d = 0; d < p->ndim; ++d) { nnb *= 3; nch *= 2; }
  int* srl = (int*)malloc(sizeof(int) * (size_t)nblocks * nnb);
  for (long k = 0; k < (long)nblocks * nnb; ++k) srl[k] = (ntype[k] == 3 || ntype[k] == 0) ? nbr[k] : -1;
  double dx1[3] = {dxb[0], p->ndim > 1 ? dxb[1] : 1.0, p->ndim > 2 ? dxb[2] : 1.0};
  int rc = orc_create(p, nblocks, dx1, srl, xlohi, bc, out);
  free(srl);
  if (rc) return rc;
  orc_sim* s = *out;
  for (int ib = 0; ib < nblocks; ++ib)
    for (int d = 0; d < p->ndim; ++d) s->bdx[ib * 3 + d] = dxb[ib * p->ndim + d];
  memcpy(s->ntype, ntype, sizeof(int) * (size_t)nblocks * nnb);
  memcpy(s->nbr, nbr, sizeof(int) * (size_t)nblocks * nnb);
  s->nchild = (int*)malloc(sizeof(int) * (size_t)nblocks * nnb * nch);
  memcpy(s->nchild, nchild, sizeof(int) * (size_t)nblocks * nnb * nch);
  s->icpos = (int*)malloc(sizeof(int) * (size_t)nblocks * p->ndim);
  memcpy(s->icpos, icpos, sizeof(int) * (size_t)nblocks * p->ndim);
  s->amr = 1;
  s->ncellc = 1;
  for (int d = 0; d < 3; ++d) {
    s->nxc[d] = d < p->ndim ? p->block_nx[d] / 2 + 2 * p->nghost : 1;
    s->ncellc *= s->nxc[d];
  }
  s->psc = (double*)calloc((size_t)nblocks * s->nw * s->ncellc, sizeof(double));
  s->bflux = (double*)calloc((size_t)nblocks * p->ndim * 2 * s->nw * s->nface, sizeof(double));
  if (!s->psc || !s->bflux) { orc_destroy(s); return 5; }
  return 0;
}

/* regrid helpers on single ghosted blocks of a simulation's layout */

/* coarsen_grid_siblings src/amr/mod_coarsen_refine.t:322-400: child ic
   (1 or 2 per dimension) into its quadrant of the parent */
int orc_coarsen_child(const orc_sim* s, const double* wFi, const int* ic, double* wCo) {
  const int nd = s->p.ndim, ng = s->ng;
  int flo[3] = {1, 1, 1}, clo[3] = {1, 1, 1}, chi[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) {
    flo[d] = ng + 1;
    clo[d] = ng + 1 + (ic[d] - 1) * s->bnx[d] / 2;
    chi[d] = ng + s->bnx[d] + (ic[d] - 2) * s->bnx[d] / 2;
  }
  coarsen_cells(s, wFi, s->nx, s->ncell, flo, wCo, s->nx, s->ncell, clo, chi);
  return 0;
}

/* prolong_grid / prolong_2nd src/amr/mod_refine.t:43-167 (slab_uniform,
   conserved variables): child ic of the parent wCo (ghost cells filled) */
int orc_prolong_2nd(const orc_sim* s, const double* wCo, const int* ic, double* wFi) {
  const int nd = s->p.ndim, ng = s->ng;
  int clo[3] = {1, 1, 1}, chi[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) {
    clo[d] = ng + 1 + (ic[d] - 1) * s->bnx[d] / 2;
    chi[d] = ng + s->bnx[d] + (ic[d] - 2) * s->bnx[d] / 2;
  }
  for (int kc = clo[2]; kc <= chi[2]; ++kc)
    for (int jc = clo[1]; jc <= chi[1]; ++jc)
      for (int icc = clo[0]; icc <= chi[0]; ++icc) {
        int co[3] = {icc, jc, kc}, fi[3] = {1, 1, 1};
        for (int d = 0; d < nd; ++d) fi[d] = 2 * (co[d] - clo[d]) + ng + 1;
        double slope[MAXW][3];
        for (int d = 0; d < nd; ++d) {
          int h[3] = {co[0], co[1], co[2]}, p[3] = {co[0], co[1], co[2]};
          h[d] -= 1; p[d] += 1;
          for (int iw = 0; iw < s->nw; ++iw) {
            const double* q = wCo + iw * s->ncell;
            const double c0 = q[IX1(s->nx, co[0], co[1], co[2])];
            const double slopeL = c0 - q[IX1(s->nx, h[0], h[1], h[2])];
            const double slopeR = q[IX1(s->nx, p[0], p[1], p[2])] - c0;
            const double slopeC = half * (slopeR + slopeL);
            const double signC = copysign(one, slopeC);
            slope[iw][d] = signC * MAX2(zero, MIN2(MIN2(fabs(slopeC), signC * slopeL), signC * slopeR));
          }
        }
        for (int dk = 0; dk <= (nd > 2); ++dk)
          for (int dj = 0; dj <= (nd > 1); ++dj)
            for (int di = 0; di <= 1; ++di) {
              int f[3] = {fi[0] + di, nd > 1 ? fi[1] + dj : 1, nd > 2 ? fi[2] + dk : 1};
              int off[3] = {di, dj, dk};
              double eta[3];
              for (int d = 0; d < nd; ++d) eta[d] = 0.5 * ((double)off[d] - 0.5);
              for (int iw = 0; iw < s->nw; ++iw) {
                double v = wCo[iw * s->ncell + IX1(s->nx, co[0], co[1], co[2])];
                for (int d = 0; d < nd; ++d) v = v + slope[iw][d] * eta[d];
                wFi[iw * s->ncell + IX1(s->nx, f[0], f[1], f[2])] = v;
              }
            }
      }
  return 0;
}

/* lohner_grid src/amr/mod_errest.t:63-199 (Cartesian, no log flags, no B0):
   error estimate of every interior cell of one ghosted block; weight[nw] */
int orc_lohner_error(const orc_sim* s, const double* w, const double* weight, double wavefilter,
                     double* err) {
  const int nd = s->p.ndim, ng = s->ng;
  const double epsilon = 1.0e-6;
  int mlo[3] = {1, 1, 1}, mhi[3] = {1, 1, 1};
  for (int d = 0; d < nd; ++d) { mlo[d] = ng + 1; mhi[d] = ng + s->bnx[d]; }
  int nb0 = mhi[0] - mlo[0] + 1, nb1 = mhi[1] - mlo[1] + 1;
  long nint = (long)nb0 * nb1 * (mhi[2] - mlo[2] + 1);
  for (long c = 0; c < nint; ++c) err[c] = 0.0;
  double* tmp = (double*)malloc(sizeof(double) * s->ncell);
  double* tmp2 = (double*)malloc(sizeof(double) * s->ncell);
  double* num = (double*)malloc(sizeof(double) * nint);
  double* den = (double*)malloc(sizeof(double) * nint);
  int xlo[3], xhi[3]; /* ix = ixM extended by one */
  for (int d = 0; d < 3; ++d) { xlo[d] = d < nd ? mlo[d] - 1 : 1; xhi[d] = d < nd ? mhi[d] + 1 : 1; }
#define LIDX(i, j, k) ((long)((i) - mlo[0]) + (long)nb0 * ((long)((j) - mlo[1]) + (long)nb1 * ((k) - mlo[2])))
  for (int iflag = 0; iflag < s->nw; ++iflag) {
    if (weight[iflag] == 0.0) continue;
    const double* q = w + iflag * s->ncell;
    for (long c = 0; c < nint; ++c) num[c] = 0.0;
    for (int idims = 0; idims < nd; ++idims) {
      for (int k = xlo[2]; k <= xhi[2]; ++k)
        for (int j = xlo[1]; j <= xhi[1]; ++j)
          for (int i = xlo[0]; i <= xhi[0]; ++i) {
            int p[3] = {i, j, k}, h[3] = {i, j, k};
            p[idims] += 1; h[idims] -= 1;
            tmp[IX1(s->nx, i, j, k)] = q[IX1(s->nx, p[0], p[1], p[2])] - q[IX1(s->nx, h[0], h[1], h[2])];
          }
      for (int idims2 = 0; idims2 < nd; ++idims2)
        for (int k = mlo[2]; k <= mhi[2]; ++k)
          for (int j = mlo[1]; j <= mhi[1]; ++j)
            for (int i = mlo[0]; i <= mhi[0]; ++i) {
              int p[3] = {i, j, k}, h[3] = {i, j, k};
              p[idims2] += 1; h[idims2] -= 1;
              const double dd = tmp[IX1(s->nx, p[0], p[1], p[2])] - tmp[IX1(s->nx, h[0], h[1], h[2])];
              num[LIDX(i, j, k)] = num[LIDX(i, j, k)] + dd * dd;
            }
    }
    for (long c = 0; c < nint; ++c) den[c] = 0.0;
    for (int idims = 0; idims < nd; ++idims) {
      for (long c = 0; c < s->ncell; ++c) tmp[c] = fabs(q[c]);
      for (int k = xlo[2]; k <= xhi[2]; ++k)
        for (int j = xlo[1]; j <= xhi[1]; ++j)
          for (int i = xlo[0]; i <= xhi[0]; ++i) {
            int p[3] = {i, j, k}, h[3] = {i, j, k};
            p[idims] += 1; h[idims] -= 1;
            tmp2[IX1(s->nx, i, j, k)] = tmp[IX1(s->nx, p[0], p[1], p[2])] + tmp[IX1(s->nx, h[0], h[1], h[2])];
          }
      for (int k = mlo[2]; k <= mhi[2]; ++k)
        for (int j = mlo[1]; j <= mhi[1]; ++j)
          for (int i = mlo[0]; i <= mhi[0]; ++i) {
            int p[3] = {i, j, k}, h[3] = {i, j, k};
            p[idims] += 2; h[idims] -= 2;
            const double c0 = q[IX1(s->nx, i, j, k)];
            tmp[IX1(s->nx, i, j, k)] = fabs(q[IX1(s->nx, p[0], p[1], p[2])] - c0) +
                                       fabs(c0 - q[IX1(s->nx, h[0], h[1], h[2])]);
          }
      for (int idims2 = 0; idims2 < nd; ++idims2)
        for (int k = mlo[2]; k <= mhi[2]; ++k)
          for (int j = mlo[1]; j <= mhi[1]; ++j)
            for (int i = mlo[0]; i <= mhi[0]; ++i) {
              int p[3] = {i, j, k}, h[3] = {i, j, k};
              p[idims2] += 1; h[idims2] -= 1;
              const double v = tmp[IX1(s->nx, i, j, k)] +
                               wavefilter * (tmp2[IX1(s->nx, p[0], p[1], p[2])] + tmp2[IX1(s->nx, h[0], h[1], h[2])]);
              den[LIDX(i, j, k)] = den[LIDX(i, j, k)] + v * v;
            }
    }
    for (long c = 0; c < nint; ++c) err[c] = err[c] + weight[iflag] * sqrt(num[c] / MAX2(den[c], epsilon));
  }
#undef LIDX
  free(tmp); free(tmp2); free(num); free(den);
  return 0;
}


/* ------------------------------------------------------------------ */
/* unit-test hooks (single face / single cell)                         */
/* ------------------------------------------------------------------ */

int orc_dwlimiter(const orc_sim* s, int idim, double a, double b, double* ldw, double* rdw) {
  dwlimiter_cell(s, idim, a, b, ldw, rdw);
  return 0;
}

int orc_weno5_face(const double* w6, double* wl, double* wr) {
  /* w6 = w(i-2) .. w(i+3); face between i and i+1 */
  weno5_face(w6 + 2, 0, wl, wr);
  return 0;
}

/* one face of finite_volume: lw[iw*nl + k] primitive values of the nl
   cells around the face (face between cells nl/2-1 and nl/2) */
int orc_face_flux(orc_sim* s, int method, int idim, int nl, const double* lw_in, double ch,
                  double* fC, double* cmax_out, double* cmin_out) {
  double lw[MAXW][MAXLINE];
  double wLp[MAXLINE][MAXW], wRp[MAXLINE][MAXW];
  if (nl > MAXLINE || nl < 4) return 1;
  for (int iw = 0; iw < s->nw; ++iw)
    for (int k = 0; k < nl; ++k) lw[iw][k] = lw_in[iw * nl + k];
  const int i = nl / 2 - 1;
  reconstruct_line(s, idim, lw, i, i, wLp, wRp);
  double wLC[MAXW], wRC[MAXW], fL[MAXW], fR[MAXW], cmax = 0.0, cmin = 0.0;
  int crash = 0;
  for (int iw = 0; iw < s->nw; ++iw) { wLC[iw] = wLp[i][iw]; wRC[iw] = wRp[i][iw]; }
  to_conserved_cell(s, wLC);
  to_conserved_cell(s, wRC);
  if (s->p.phys == ORC_PHYS_MHD && s->p.divb == ORC_DIVB_GLM)
    modify_wLR_glm(s, wLC, wRC, wLp[i], wRp[i], idim, ch);
  get_flux(s, wLC, wLp[i], idim, ch, fL);
  get_flux(s, wRC, wRp[i], idim, ch, fR);
  if (method == ORC_FS_TVDLF) {
    get_cbounds(s, wLC, wRC, wLp[i], wRp[i], idim, 0, &cmax, &cmin, &crash);
    riemann_tvdlf(s, wLC, wRC, fL, fR, cmax, fC);
  } else {
    get_cbounds(s, wLC, wRC, wLp[i], wRp[i], idim, 1, &cmax, &cmin, &crash);
    if (method == ORC_FS_HLLC) riemann_hllc(s, idim, wLC, wRC, fL, fR, cmax, cmin, fC);
    else if (method == ORC_FS_HLLD) riemann_hlld(s, idim, wLC, wRC, wLp[i], wRp[i], fL, fR, cmax, cmin, NULL, fC);
    else riemann_hll(s, idim, wLC, wRC, fL, fR, cmax, cmin, fC);
  }
  if (cmax_out) *cmax_out = cmax;
  if (cmin_out) *cmin_out = cmin;
  return 0;
}

/* cell cmax of a conserved state (getdt_courant per cell) */
double orc_cell_cmax(const orc_sim* s, const double* wcons, int idim) {
  double cell[MAXW];
  for (int iw = 0; iw < s->nw; ++iw) cell[iw] = wcons[iw];
  to_primitive_cell(s, cell);
  return get_cmax(s, cell, idim);
}
