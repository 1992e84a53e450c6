// fv_adjoint.h -- reverse mode (adjoint) of the finite-volume face flux for
// hydrodynamics: conserved stencil -> primitive (hd_to_primitive) -> TVD
// reconstruction (minmod / vanleer, dwlimiter2) -> hd_to_conserved ->
// hd_get_flux -> Einfeldt bounds -> TVDLF / HLL / HLLC
// (src/mod_finite_volume.t:116-339, :348-504; src/hd/mod_hd_phys.t;
// src/hd/mod_hd_hllc.t).  Every branch (limiter cases, Riemann patches) is
// differentiated along the branch the forward evaluation takes.
//
// face_adj_prim recomputes the forward face from the primitive states of its
// four cells and returns their primitive adjoints; prim_rev carries a cell's
// summed primitive adjoint back to its conserved state (the reverse of the
// conversion is linear, so it is applied once per cell rather than per face).
#pragma once

#include "fv_device.h"

namespace FV_NS {

// limited slope t(a, b) of dwlimiter2 for the symmetric TVD limiters and
// its partial derivatives
template <class R, int LIM>
__device__ __forceinline__ R lim_grad(R a, R b, R& da, R& db) {
  if constexpr (LIM == LIM_MINMOD) {
    const R s = copysign(R(1), b);
    const R sa = s * a, ab = fabs(b);
    if (!(sa > R(0))) { da = R(0); db = R(0); return R(0); }
    if (ab <= sa) { da = R(0); db = R(1); return s * ab; }
    da = R(1); db = R(0); return a;
  } else {  // vanleer: 2 max(ab, 0) / (a + b + 1e-12)
    const R ab = a * b;
    if (!(ab > R(0))) { da = R(0); db = R(0); return R(0); }
    const R q = b + a + R(1.0e-12);
    const R iq = frcp(q);
    const R t = R(2) * ab * iq;
    da = R(2) * b * (q - a) * iq * iq;
    db = R(2) * a * (q - b) * iq * iq;
    return t;
  }
}

// cada_limited (the limiter_cada branch of dwlimiter2, src/limiter/
// mod_limiter.t:233-251) of theta and its derivative, along the branch the
// forward min / max take
template <class R>
__device__ __forceinline__ R cada_limited_grad(R th, R& dth) {
  const R cadalfa = R(0.5), cadbeta = R(2), cadgamma = R(1.6);
  const R A = (R(2) + th) * R(1.0 / 3.0);
  R v = cadbeta * th, dv = cadbeta;                  // min(beta theta, A)
  if (A < v) { v = A; dv = R(1.0 / 3.0); }
  if (cadgamma < v) { v = cadgamma; dv = R(0); }     // min(., gamma)
  if (-cadalfa * th > v) { v = -cadalfa * th; dv = -cadalfa; }  // max(-alfa theta, .)
  if (A < v) { v = A; dv = R(1.0 / 3.0); }           // min(A, .)
  if (R(0) > v) { v = R(0); dv = R(0); }             // max(0, .)
  dth = dv;
  return v;
}

// One face side of cada3 (limiter_cada3, src/limiter/mod_limiter.t:253-290,
// as face_limit evaluates it): F(x, d0) = phi(x / (d0 + sign(eps, d0)),
// (d0^2 + x^2) rdelinv) * d0, x the outer difference (dwC(i-1) for the
// left state, dwC(i+1) for the right one) and d0 = dwC(i); phi = cada3_side.
// Returns F and its partials in x and d0.
template <class R>
__device__ __forceinline__ R cada3_grad(R x, R d0, R rdel, R& dx, R& dd0) {
  const R eps = R(1.4901161193847656e-08);
  const R q = frcp(d0 + copysign(eps, d0));
  const R th = x * q;
  const R eta = (d0 * d0 + x * x) * rdel;
  const R A = (R(2) + th) * R(1.0 / 3.0);
  R lth;
  const R lim = cada_limited_grad(th, lth);
  R phi, pth, peta;
  if (eta <= R(1) - R(1.0e-14)) {
    phi = A; pth = R(1.0 / 3.0); peta = R(0);
  } else if (eta >= R(1) + R(1.0e-14)) {
    phi = lim; pth = lth; peta = R(0);
  } else {  // the narrow blend of the two, linear in eta
    const R t2 = (eta - R(1)) * R(1.0e14);
    phi = R(0.5) * ((R(1) - t2) * A + (R(1) + t2) * lim);
    pth = R(0.5) * ((R(1) - t2) * R(1.0 / 3.0) + (R(1) + t2) * lth);
    peta = R(0.5e14) * (lim - A);
  }
  dx = d0 * (pth * q + peta * R(2) * x * rdel);
  dd0 = phi + d0 * (-pth * x * (q * q) + peta * R(2) * d0 * rdel);
  return phi * d0;
}

// Limited slopes of the two states of a face and their partials: t1 = ldw
// of the left cell from (dm, d0), t2 = rdw of the right cell from (d0, dp);
// d1a, d1b = dt1/d(dm, d0); d2a, d2b = dt2/d(d0, dp)
template <class R, int LIM>
__device__ __forceinline__ void face_slopes(R dm, R d0, R dp, R rdel, R& t1, R& t2, R& d1a, R& d1b, R& d2a,
                                            R& d2b) {
  if constexpr (LIM == LIM_CADA3) {
    t1 = cada3_grad(dm, d0, rdel, d1a, d1b);
    t2 = cada3_grad(dp, d0, rdel, d2b, d2a);
  } else {
    t1 = lim_grad<R, LIM>(dm, d0, d1a, d1b);
    t2 = lim_grad<R, LIM>(d0, dp, d2a, d2b);
  }
}

// Reverse of the face flux in primitive form: from the primitive states w of
// the face's four cells (as the sweep rounds them to R) and the flux adjoint
// fbar, writes the primitive adjoints wb[c] = (d fC / d w_c)^T fbar (or adds
// them to wb[c] when ACC).
template <class P, class R, int RIEM, int ID, int LIM, bool ACC = false, class W = double>
__device__ void face_adj_prim_hd(const R (*w)[P::NW], const FVParams& fp, R ch, const R* fbar, W (*wb)[P::NW]) {
  static_assert(!P::MHD, "hydrodynamics");
  constexpr int NW = P::NW, ND = P::NDIR, RHO = P::RHO, E = P::E, MN = P::MOM + ID;
  const R gam = R(fp.gamma), inv_g1 = R(fp.inv_gamma_1);
  // ---------------- forward ----------------
  const R rdel = R(fp.rdelinv[ID]);
  R wL[NW], wR[NW], d1a[NW], d1b[NW], d2a[NW], d2b[NW];
  for (int iw = 0; iw < NW; ++iw) {
    const R dm = w[1][iw] - w[0][iw], d0 = w[2][iw] - w[1][iw], dp = w[3][iw] - w[2][iw];
    R t1, t2;
    face_slopes<R, LIM>(dm, d0, dp, rdel, t1, t2, d1a[iw], d1b[iw], d2a[iw], d2b[iw]);
    wL[iw] = w[1][iw] + R(0.5) * t1;
    wR[iw] = w[2][iw] - R(0.5) * t2;
  }
  // conserved face states
  R wLC[NW], wRC[NW], vsqL = R(0), vsqR = R(0);
  for (int k = 0; k < ND; ++k) {
    vsqL = vsqL + wL[P::MOM + k] * wL[P::MOM + k];
    vsqR = vsqR + wR[P::MOM + k] * wR[P::MOM + k];
  }
  wLC[RHO] = wL[RHO];
  wRC[RHO] = wR[RHO];
  wLC[E] = wL[E] * inv_g1 + R(0.5) * vsqL * wL[RHO];
  wRC[E] = wR[E] * inv_g1 + R(0.5) * vsqR * wR[RHO];
  for (int k = 0; k < ND; ++k) {
    wLC[P::MOM + k] = wL[RHO] * wL[P::MOM + k];
    wRC[P::MOM + k] = wR[RHO] * wR[P::MOM + k];
  }
  R fL[NW], fR[NW];
  phys_flux<P, R, ID>(wLC, wL, R(0), fL);
  phys_flux<P, R, ID>(wRC, wR, R(0), fR);
  // Einfeldt bounds (hd_get_cbounds, typeboundspeed 'Einfeldt')
  const R t1 = fsqrt(wL[RHO]), t2 = fsqrt(wR[RHO]);
  const R t3 = frcp(t1 + t2);
  const R um = (wL[MN] * t1 + wR[MN] * t2) * t3;
  const R csL = fdiv(gam * wL[E], wL[RHO]), csR = fdiv(gam * wR[E], wR[RHO]);
  const R dv = wR[MN] - wL[MN];
  const R A = t1 * csL + t2 * csR;
  const R D2 = A * t3 + R(0.5) * t1 * t2 * (t3 * t3) * (dv * dv);
  const R dmean = fsqrt(D2);
  R cmax, cmin;
  if constexpr (RIEM == RIEM_TVDLF) {
    cmax = fabs(um) + dmean;
    cmin = R(0);
  } else {
    cmin = um - dmean;
    cmax = um + dmean;
  }
  // ---------------- reverse ----------------
  R fLb[NW], fRb[NW], wLCb[NW], wRCb[NW];
  for (int iw = 0; iw < NW; ++iw) { fLb[iw] = R(0); fRb[iw] = R(0); wLCb[iw] = R(0); wRCb[iw] = R(0); }
  R cmaxb = R(0), cminb = R(0);
  if constexpr (RIEM == RIEM_TVDLF) {
    const R fac = R(-0.5) * R(fp.tvdlfeps) * cmax;
    R s = R(0);
    for (int iw = 0; iw < NW; ++iw) {
      fLb[iw] = R(0.5) * fbar[iw];
      fRb[iw] = R(0.5) * fbar[iw];
      wRCb[iw] = fac * fbar[iw];
      wLCb[iw] = -fac * fbar[iw];
      s = s + fbar[iw] * (wRC[iw] - wLC[iw]);
    }
    cmaxb = R(-0.5) * R(fp.tvdlfeps) * s;
  } else if constexpr (RIEM == RIEM_HLL) {
    if (cmin >= R(0)) {
      for (int iw = 0; iw < NW; ++iw) fLb[iw] = fbar[iw];
    } else if (cmax <= R(0)) {
      for (int iw = 0; iw < NW; ++iw) fRb[iw] = fbar[iw];
    } else {
      const R iD = frcp(cmax - cmin);
      R Db = R(0);
      for (int iw = 0; iw < NW; ++iw) {
        const R dw = wRC[iw] - wLC[iw];
        const R f = (cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * dw) * iD;
        const R Nb = fbar[iw] * iD;
        Db = Db - fbar[iw] * f * iD;
        fLb[iw] = Nb * cmax;
        fRb[iw] = -Nb * cmin;
        wRCb[iw] = Nb * cmin * cmax;
        wLCb[iw] = -Nb * cmin * cmax;
        cmaxb = cmaxb + Nb * (fL[iw] + cmin * dw);
        cminb = cminb + Nb * (-fR[iw] + cmax * dw);
      }
      cmaxb = cmaxb + Db;
      cminb = cminb - Db;
    }
  } else {  // HLLC, src/mod_finite_volume.t:439-504, hd_get_lCD / hd_get_wCD
    int patchf = 1;
    if (cmin >= R(0)) patchf = -2;
    else if (cmax <= R(0)) patchf = 2;
    R lam = R(0), iD = R(0), whr = R(0), whm = R(0);
    if (patchf == 1) {
      iD = frcp(cmax - cmin);
      whr = (cmax * wRC[RHO] - cmin * wLC[RHO] + fL[RHO] - fR[RHO]) * iD;
      whm = (cmax * wRC[MN] - cmin * wLC[MN] + fL[MN] - fR[MN]) * iD;
      lam = fdiv(whm, whr);
      bool cond = true;
      if (cmin < R(0) && lam > R(0) && lam < cmax) {
        patchf = -1;
      } else if (cmax > R(0) && lam < R(0) && lam > cmin) {
        patchf = 1;
      } else if (lam >= cmax || lam <= cmin) {
        lam = R(0);
        patchf = 3;
        cond = false;
      }
      if (cond && lam == R(0)) patchf = (-cmin >= cmax) ? 1 : -1;
    }
    if (patchf == -2) {
      for (int iw = 0; iw < NW; ++iw) fLb[iw] = fbar[iw];
    } else if (patchf == 2) {
      for (int iw = 0; iw < NW; ++iw) fRb[iw] = fbar[iw];
    } else if (patchf == 3) {
      R Db = R(0);
      for (int iw = 0; iw < NW; ++iw) {
        const R dw = wRC[iw] - wLC[iw];
        const R f = (cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * dw) * iD;
        const R Nb = fbar[iw] * iD;
        Db = Db - fbar[iw] * f * iD;
        fLb[iw] = Nb * cmax;
        fRb[iw] = -Nb * cmin;
        wRCb[iw] = Nb * cmin * cmax;
        wLCb[iw] = -Nb * cmin * cmax;
        cmaxb = cmaxb + Nb * (fL[iw] + cmin * dw);
        cminb = cminb + Nb * (-fR[iw] + cmax * dw);
      }
      cmaxb = cmaxb + Db;
      cminb = cminb - Db;
    } else {
      const bool right = patchf == 1;
      const R cs = right ? cmax : cmin;
      const R* wS = right ? wRC : wLC;
      const R* fS = right ? fR : fL;
      R* wSb = right ? wRCb : wLCb;
      R* fSb = right ? fRb : fLb;
      const R vS = fdiv(wS[MN], wS[RHO]);
      const R csl = cs - lam;
      const R csmls = frcp(csl);
      const R Z = wS[RHO] * (cs - vS);
      R wCD[NW];
      wCD[RHO] = Z * csmls;
      for (int k = 0; k < ND; ++k)
        if (k != ID) wCD[P::MOM + k] = (cs * wS[P::MOM + k] - fS[P::MOM + k]) * csmls;
      wCD[MN] = wCD[RHO] * lam;
      const R pCD = wS[RHO] * (cs - vS) * (lam - vS) + fS[MN] - wS[MN] * vS;
      const R XE = cs * wS[E] - fS[E] + lam * pCD;
      wCD[E] = XE * csmls;
      // fCD = fS + cs (wCD - wS)
      R csb = R(0), lamb = R(0), csmlsb = R(0), vSb = R(0), wCDb[NW];
      for (int iw = 0; iw < NW; ++iw) {
        fSb[iw] = fSb[iw] + fbar[iw];
        csb = csb + fbar[iw] * (wCD[iw] - wS[iw]);
        wCDb[iw] = cs * fbar[iw];
        wSb[iw] = wSb[iw] - cs * fbar[iw];
      }
      // wCD[E]
      {
        const R XEb = wCDb[E] * csmls;
        csmlsb = csmlsb + wCDb[E] * XE;
        csb = csb + XEb * wS[E];
        wSb[E] = wSb[E] + XEb * cs;
        fSb[E] = fSb[E] - XEb;
        lamb = lamb + XEb * pCD;
        const R pCDb = XEb * lam;
        wSb[RHO] = wSb[RHO] + pCDb * (cs - vS) * (lam - vS);
        csb = csb + pCDb * wS[RHO] * (lam - vS);
        lamb = lamb + pCDb * wS[RHO] * (cs - vS);
        vSb = vSb + pCDb * (-wS[RHO] * (lam - vS) - wS[RHO] * (cs - vS) - wS[MN]);
        fSb[MN] = fSb[MN] + pCDb;
        wSb[MN] = wSb[MN] - pCDb * vS;
      }
      // wCD[MN] = wCD[RHO] * lam
      R rhoCDb = wCDb[RHO] + wCDb[MN] * lam;
      lamb = lamb + wCDb[MN] * wCD[RHO];
      // transverse momenta
      for (int k = 0; k < ND; ++k) {
        if (k == ID) continue;
        const int m = P::MOM + k;
        const R Y = cs * wS[m] - fS[m];
        const R Yb = wCDb[m] * csmls;
        csmlsb = csmlsb + wCDb[m] * Y;
        csb = csb + Yb * wS[m];
        wSb[m] = wSb[m] + Yb * cs;
        fSb[m] = fSb[m] - Yb;
      }
      // wCD[RHO] = Z csmls, Z = wS[RHO] (cs - vS)
      {
        const R Zb = rhoCDb * csmls;
        csmlsb = csmlsb + rhoCDb * Z;
        wSb[RHO] = wSb[RHO] + Zb * (cs - vS);
        csb = csb + Zb * wS[RHO];
        vSb = vSb - Zb * wS[RHO];
      }
      // csmls = 1/(cs - lam)
      const R cslb = -csmlsb * csmls * csmls;
      csb = csb + cslb;
      lamb = lamb - cslb;
      // vS = wS[MN] / wS[RHO]
      const R irS = frcp(wS[RHO]);
      wSb[MN] = wSb[MN] + vSb * irS;
      wSb[RHO] = wSb[RHO] - vSb * vS * irS;
      // lam = whm / whr
      const R irh = frcp(whr);
      const R whmb = lamb * irh;
      const R whrb = -lamb * lam * irh;
      // whll = (cmax wRC - cmin wLC + fL - fR) iD for RHO and MN
      R iDb = R(0);
      {
        const R Qr = fdiv(whr, iD), Qm = fdiv(whm, iD);  // numerators
        (void)Qr; (void)Qm;
        const int vars[2] = {RHO, MN};
        const R wb[2] = {whrb, whmb};
        const R wh[2] = {whr, whm};
        for (int j = 0; j < 2; ++j) {
          const int iw = vars[j];
          const R Qb = wb[j] * iD;
          iDb = iDb + wb[j] * (cmax * wRC[iw] - cmin * wLC[iw] + fL[iw] - fR[iw]);
          cmaxb = cmaxb + Qb * wRC[iw];
          cminb = cminb - Qb * wLC[iw];
          wRCb[iw] = wRCb[iw] + Qb * cmax;
          wLCb[iw] = wLCb[iw] - Qb * cmin;
          fLb[iw] = fLb[iw] + Qb;
          fRb[iw] = fRb[iw] - Qb;
          (void)wh;
        }
      }
      const R Db = -iDb * iD * iD;
      cmaxb = cmaxb + Db;
      cminb = cminb - Db;
      if (right) cmaxb = cmaxb + csb;
      else cminb = cminb + csb;
    }
  }
  // bounds -> face primitive states
  R wLb[NW], wRb[NW];
  for (int iw = 0; iw < NW; ++iw) { wLb[iw] = R(0); wRb[iw] = R(0); }
  {
    R umb, dmb;
    if constexpr (RIEM == RIEM_TVDLF) {
      umb = copysign(R(1), um) * cmaxb;
      dmb = cmaxb;
    } else {
      umb = cmaxb + cminb;
      dmb = cmaxb - cminb;
    }
    const R D2b = fdiv(dmb * R(0.5), dmean);
    const R Ab = D2b * t3;
    R t3b = D2b * A;
    const R Bb = D2b;
    R t1b = Bb * R(0.5) * t2 * (t3 * t3) * (dv * dv);
    R t2b = Bb * R(0.5) * t1 * (t3 * t3) * (dv * dv);
    t3b = t3b + Bb * t1 * t2 * t3 * (dv * dv);
    const R dvb = Bb * t1 * t2 * (t3 * t3) * dv;
    t1b = t1b + Ab * csL;
    t2b = t2b + Ab * csR;
    const R csLb = Ab * t1, csRb = Ab * t2;
    const R irL = frcp(wL[RHO]), irR = frcp(wR[RHO]);
    wLb[E] = wLb[E] + csLb * gam * irL;
    wLb[RHO] = wLb[RHO] - csLb * csL * irL;
    wRb[E] = wRb[E] + csRb * gam * irR;
    wRb[RHO] = wRb[RHO] - csRb * csR * irR;
    wRb[MN] = wRb[MN] + dvb;
    wLb[MN] = wLb[MN] - dvb;
    wLb[MN] = wLb[MN] + umb * t1 * t3;
    wRb[MN] = wRb[MN] + umb * t2 * t3;
    t1b = t1b + umb * wL[MN] * t3;
    t2b = t2b + umb * wR[MN] * t3;
    t3b = t3b + umb * (wL[MN] * t1 + wR[MN] * t2);
    t1b = t1b - t3b * t3 * t3;
    t2b = t2b - t3b * t3 * t3;
    wLb[RHO] = wLb[RHO] + fdiv(t1b * R(0.5), t1);
    wRb[RHO] = wRb[RHO] + fdiv(t2b * R(0.5), t2);
  }
  // fluxes: f_rho = vn rho, f_mk = vn m_k, f_mn += p, f_E = vn (E + p)
  auto flux_rev = [&](const R* wX, const R* wXC, const R* fXb, R* wXb, R* wXCb) {
    const R vn = wX[MN];
    R vnb = fXb[RHO] * wX[RHO] + fXb[E] * (wXC[E] + wX[E]);
    wXb[RHO] = wXb[RHO] + fXb[RHO] * vn;
    for (int k = 0; k < ND; ++k) {
      vnb = vnb + fXb[P::MOM + k] * wXC[P::MOM + k];
      wXCb[P::MOM + k] = wXCb[P::MOM + k] + fXb[P::MOM + k] * vn;
    }
    wXb[E] = wXb[E] + fXb[MN] + fXb[E] * vn;
    wXCb[E] = wXCb[E] + fXb[E] * vn;
    wXb[MN] = wXb[MN] + vnb;
  };
  flux_rev(wL, wLC, fLb, wLb, wLCb);
  flux_rev(wR, wRC, fRb, wRb, wRCb);
  // to_cons: E = p inv_g1 + 0.5 vsq rho, m_k = rho v_k, rho = rho
  auto cons_rev = [&](const R* wX, R vsq, const R* wXCb, R* wXb) {
    wXb[RHO] = wXb[RHO] + wXCb[RHO];
    wXb[E] = wXb[E] + inv_g1 * wXCb[E];
    const R vsqb = R(0.5) * wX[RHO] * wXCb[E];
    wXb[RHO] = wXb[RHO] + R(0.5) * vsq * wXCb[E];
    for (int k = 0; k < ND; ++k) {
      const int m = P::MOM + k;
      wXb[RHO] = wXb[RHO] + wXCb[m] * wX[m];
      wXb[m] = wXb[m] + R(2) * wX[m] * vsqb + wX[RHO] * wXCb[m];
    }
  };
  cons_rev(wL, vsqL, wLCb, wLb);
  cons_rev(wR, vsqR, wRCb, wRb);
  // reconstruction: wL = w1 + t1/2, wR = w2 - t2/2
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const R t1b = R(0.5) * wLb[iw], t2b = R(-0.5) * wRb[iw];
    const R dmb = t1b * d1a[iw];
    const R d0b = t1b * d1b[iw] + t2b * d2a[iw];
    const R dpb = t2b * d2b[iw];
    if constexpr (ACC) {
      wb[0][iw] += W(-dmb);
      wb[1][iw] += W(wLb[iw] + dmb - d0b);
      wb[2][iw] += W(wRb[iw] + d0b - dpb);
      wb[3][iw] += W(dpb);
    } else {
      wb[0][iw] = W(-dmb);
      wb[1][iw] = W(wLb[iw] + dmb - d0b);
      wb[2][iw] = W(wRb[iw] + d0b - dpb);
      wb[3][iw] = W(dpb);
    }
  }
}

// Ideal MHD without divergence cleaning sources in the face (typedivbfix
// 'none' or 'linde'): TVD reconstruction -> mhd_to_conserved ->
// mhd_get_flux (src/mhd/mod_mhd_phys.t:3810-3889) -> Einfeldt bounds with
// the fast speed (mhd_get_cbounds :3037-3127, mhd_get_csound_prim
// :3329-3379) -> HLL or TVDLF; the normal field takes Lax-Friedrichs
// dissipation under HLL (flux_type, :580-588).
template <class P, class R, int RIEM, int ID, int LIM, bool ACC = false, class W = double>
__device__ void face_adj_prim_mhd(const R (*w)[P::NW], const FVParams& fp, R ch, const R* fbar,
                                  W (*wb)[P::NW]) {
  static_assert(P::MHD && !P::LINDE, "ideal MHD, without divergence control in the face or with GLM");
  static_assert(RIEM == RIEM_HLL || RIEM == RIEM_TVDLF, "MHD faces take hll or tvdlf");
  constexpr int NW = P::NW, ND = P::NDIR, RHO = P::RHO, E = P::E, MN = P::MOM + ID, BN = P::MAG + ID;
  // flux_type: the normal field takes Lax-Friedrichs dissipation under HLL
  // only without GLM (src/mhd/mod_mhd_phys.t:580-588)
  constexpr int BTV = (RIEM == RIEM_HLL && P::NDIM > 1 && !P::GLM) ? BN : -1;
  const R gam = R(fp.gamma), inv_g1 = R(fp.inv_gamma_1), eps = R(fp.tvdlfeps);
  // ---------------- forward ----------------
  const R rdel = R(fp.rdelinv[ID]);
  R wL[NW], wR[NW], d1a[NW], d1b[NW], d2a[NW], d2b[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const R dm = w[1][iw] - w[0][iw], d0 = w[2][iw] - w[1][iw], dp = w[3][iw] - w[2][iw];
    R t1, t2;
    face_slopes<R, LIM>(dm, d0, dp, rdel, t1, t2, d1a[iw], d1b[iw], d2a[iw], d2b[iw]);
    wL[iw] = w[1][iw] + R(0.5) * t1;
    wR[iw] = w[2][iw] - R(0.5) * t2;
  }
  // GLM (mhd_modify_wLR, src/mhd/mod_mhd_phys.t:6326-6377): both states
  // take the same normal field and psi; the conserved states below are
  // those of the modified primitive ones, as the forward's energy
  // correction makes them
  if constexpr (P::GLM) {
    constexpr int S = P::PSI;
    const R dB = wR[BN] - wL[BN], dPsi = wR[S] - wL[S];
    const R bs = R(0.5) * (wR[BN] + wL[BN] - fdiv(dPsi, ch));
    const R ps = R(0.5) * (wR[S] + wL[S] - dB * ch);
    wL[BN] = bs; wR[BN] = bs;
    wL[S] = ps; wR[S] = ps;
  }
  R wLC[NW], wRC[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) { wLC[iw] = wL[iw]; wRC[iw] = wR[iw]; }
  to_cons<P, R>(wLC, inv_g1);
  to_cons<P, R>(wRC, inv_g1);
  R fL[NW], fR[NW];
  phys_flux<P, R, ID>(wLC, wL, ch, fL);
  phys_flux<P, R, ID>(wRC, wR, ch, fR);
  // fast speeds, kept with their intermediates for the reverse pass
  struct Fast { R ir, cs, b2, cf2, av, sq, c; };
  auto fast = [&](const R* x) {
    Fast q;
    q.ir = frcp(x[RHO]);
    q.cs = gam * x[E] * q.ir;
    R b2 = x[P::MAG] * x[P::MAG];
#pragma unroll
    for (int k = 1; k < ND; ++k) b2 = b2 + x[P::MAG + k] * x[P::MAG + k];
    q.b2 = b2;
    q.cf2 = b2 * q.ir + q.cs;
    q.av = q.cf2 * q.cf2 - R(4) * q.cs * (x[BN] * x[BN]) * q.ir;
    q.sq = q.av > R(0) ? fsqrt(q.av) : R(0);
    q.c = fsqrt(R(0.5) * (q.cf2 + q.sq));
    return q;
  };
  const Fast qL = fast(wL), qR = fast(wR);
  const R t1 = fsqrt(wL[RHO]), t2 = fsqrt(wR[RHO]);
  const R t3 = frcp(t1 + t2);
  const R um = (wL[MN] * t1 + wR[MN] * t2) * t3;
  const R dv = wR[MN] - wL[MN];
  const R A = t1 * (qL.c * qL.c) + t2 * (qR.c * qR.c);
  const R D2 = A * t3 + R(0.5) * t1 * t2 * (t3 * t3) * (dv * dv);
  const R dmean = fsqrt(D2);
  R cmax, cmin;
  if constexpr (RIEM == RIEM_TVDLF) {
    cmax = fabs(um) + dmean;
    cmin = R(0);
  } else {
    cmin = um - dmean;
    cmax = um + dmean;
  }
  // ---------------- reverse ----------------
  R fLb[NW], fRb[NW], wLCb[NW], wRCb[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) { fLb[iw] = R(0); fRb[iw] = R(0); wLCb[iw] = R(0); wRCb[iw] = R(0); }
  R cmaxb = R(0), cminb = R(0);
  if constexpr (RIEM == RIEM_TVDLF) {
    const R fac = R(-0.5) * eps * cmax;
    R s = R(0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      fLb[iw] = R(0.5) * fbar[iw];
      fRb[iw] = R(0.5) * fbar[iw];
      wRCb[iw] = fac * fbar[iw];
      wLCb[iw] = -fac * fbar[iw];
      s = s + fbar[iw] * (wRC[iw] - wLC[iw]);
    }
    cmaxb = R(-0.5) * eps * s;
  } else {
    const bool mid = !(cmin >= R(0)) && !(cmax <= R(0));
    const R iD = mid ? frcp(cmax - cmin) : R(0);
    R Db = R(0);
#pragma unroll
    for (int iw = 0; iw < NW; ++iw) {
      if (iw == BTV) {
        // fC = -eps/2 max(cmax, |cmin|) (wRC - wLC)
        const bool use_max = cmax >= fabs(cmin);
        const R M = use_max ? cmax : fabs(cmin);
        const R dw = wRC[iw] - wLC[iw];
        wRCb[iw] = R(-0.5) * eps * M * fbar[iw];
        wLCb[iw] = R(0.5) * eps * M * fbar[iw];
        const R Mb = R(-0.5) * eps * dw * fbar[iw];
        if (use_max) cmaxb = cmaxb + Mb;
        else cminb = cminb + copysign(R(1), cmin) * Mb;
      } else if (cmin >= R(0)) {
        fLb[iw] = fbar[iw];
      } else if (cmax <= R(0)) {
        fRb[iw] = fbar[iw];
      } else {
        const R dw = wRC[iw] - wLC[iw];
        const R f = (cmax * fL[iw] - cmin * fR[iw] + cmin * cmax * dw) * iD;
        const R Nb = fbar[iw] * iD;
        Db = Db - fbar[iw] * f * iD;
        fLb[iw] = Nb * cmax;
        fRb[iw] = -Nb * cmin;
        wRCb[iw] = Nb * cmin * cmax;
        wLCb[iw] = -Nb * cmin * cmax;
        cmaxb = cmaxb + Nb * (fL[iw] + cmin * dw);
        cminb = cminb + Nb * (-fR[iw] + cmax * dw);
      }
    }
    cmaxb = cmaxb + Db;
    cminb = cminb - Db;
  }
  R wLb[NW], wRb[NW];
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) { wLb[iw] = R(0); wRb[iw] = R(0); }
  // bounds -> face primitive states
  {
    R umb, dmb;
    if constexpr (RIEM == RIEM_TVDLF) {
      umb = copysign(R(1), um) * cmaxb;
      dmb = cmaxb;
    } else {
      umb = cmaxb + cminb;
      dmb = cmaxb - cminb;
    }
    const R D2b = fdiv(dmb * R(0.5), dmean);
    const R Ab = D2b * t3;
    R t3b = D2b * A;
    R t1b = D2b * R(0.5) * t2 * (t3 * t3) * (dv * dv);
    R t2b = D2b * R(0.5) * t1 * (t3 * t3) * (dv * dv);
    t3b = t3b + D2b * t1 * t2 * t3 * (dv * dv);
    const R dvb = D2b * t1 * t2 * (t3 * t3) * dv;
    t1b = t1b + Ab * (qL.c * qL.c);
    t2b = t2b + Ab * (qR.c * qR.c);
    const R cLb = R(2) * qL.c * Ab * t1, cRb = R(2) * qR.c * Ab * t2;
    wRb[MN] = wRb[MN] + dvb;
    wLb[MN] = wLb[MN] - dvb;
    wLb[MN] = wLb[MN] + umb * t1 * t3;
    wRb[MN] = wRb[MN] + umb * t2 * t3;
    t1b = t1b + umb * wL[MN] * t3;
    t2b = t2b + umb * wR[MN] * t3;
    t3b = t3b + umb * (wL[MN] * t1 + wR[MN] * t2);
    t1b = t1b - t3b * t3 * t3;
    t2b = t2b - t3b * t3 * t3;
    wLb[RHO] = wLb[RHO] + fdiv(t1b * R(0.5), t1);
    wRb[RHO] = wRb[RHO] + fdiv(t2b * R(0.5), t2);
    // fast speed reverse: c = sqrt((cf2 + sqrt(av)) / 2)
    auto fast_rev = [&](const R* x, const Fast& q, R cb, R* xb) {
      const R hb = fdiv(cb * R(0.5), q.c);
      R cf2b = R(0.5) * hb;
      const R sqb = R(0.5) * hb;
      const R avb = q.av > R(0) ? fdiv(sqb * R(0.5), q.sq) : R(0);
      const R bn = x[BN];
      cf2b = cf2b + R(2) * q.cf2 * avb;
      R csb = R(-4) * (bn * bn) * q.ir * avb;
      R irb = R(-4) * q.cs * (bn * bn) * avb;
      const R bnb = R(-8) * q.cs * bn * q.ir * avb;
      const R b2b = cf2b * q.ir;
      irb = irb + cf2b * q.b2;
      csb = csb + cf2b;
      xb[E] = xb[E] + csb * gam * q.ir;
      irb = irb + csb * gam * x[E];
#pragma unroll
      for (int k = 0; k < ND; ++k) xb[P::MAG + k] = xb[P::MAG + k] + R(2) * x[P::MAG + k] * b2b;
      xb[BN] = xb[BN] + bnb;
      xb[RHO] = xb[RHO] - irb * q.ir * q.ir;
    };
    fast_rev(wL, qL, cLb, wLb);
    fast_rev(wR, qR, cRb, wRb);
  }
  // fluxes (mhd_get_flux without B0 / GLM)
  auto flux_rev = [&](const R* x, const R* xC, const R* fb, R* xb, R* xCb) {
    const R vn = x[MN], bn = x[BN];
    R bsq = R(0), bv = R(0);
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      bsq = bsq + x[P::MAG + k] * x[P::MAG + k];
      bv = bv + x[P::MAG + k] * x[P::MOM + k];
    }
    const R ptot = x[E] + R(0.5) * bsq;
    R vnb = fb[RHO] * x[RHO] + fb[E] * (xC[E] + ptot);
    R bnb = -fb[E] * bv;
    R ptotb = fb[MN] + fb[E] * vn;
    const R bvb = -fb[E] * bn;
    xb[RHO] = xb[RHO] + fb[RHO] * vn;
    xCb[E] = xCb[E] + fb[E] * vn;
    R cmnb = R(0);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const int m = P::MOM + d, b = P::MAG + d;
      // the induction row of the normal field is GLM's psi flux instead
      const R fbb = (P::GLM && d == ID) ? R(0) : fb[b];
      // f[m] = xC[MN] v_d - bn B_d
      cmnb = cmnb + fb[m] * x[m];
      xb[m] = xb[m] + fb[m] * xC[MN] + bvb * x[b] - fbb * bn;
      bnb = bnb - fb[m] * x[b] - fbb * x[m];
      // f[b] = vn B_d - bn v_d; bv; ptot
      vnb = vnb + fbb * x[b];
      xb[b] = xb[b] - fb[m] * bn + bvb * x[m] + fbb * vn + ptotb * x[b];
    }
    xCb[MN] = xCb[MN] + cmnb;
    xb[E] = xb[E] + ptotb;
    xb[MN] = xb[MN] + vnb;
    xb[BN] = xb[BN] + bnb;
    if constexpr (P::GLM) {  // f[BN] = psi, f[psi] = ch^2 B_n
      xb[P::PSI] = xb[P::PSI] + fb[BN];
      xb[BN] = xb[BN] + ch * ch * fb[P::PSI];
    }
  };
  flux_rev(wL, wLC, fLb, wLb, wLCb);
  flux_rev(wR, wRC, fRb, wRb, wRCb);
  // to_cons: E = p inv_g1 + (vsq rho + bsq)/2, m_k = rho v_k
  auto cons_rev = [&](const R* x, const R* xCb, R* xb) {
    R vsq = R(0);
#pragma unroll
    for (int k = 0; k < ND; ++k) vsq = vsq + x[P::MOM + k] * x[P::MOM + k];
    xb[RHO] = xb[RHO] + xCb[RHO] + R(0.5) * vsq * xCb[E];
    xb[E] = xb[E] + inv_g1 * xCb[E];
    const R vsqb = R(0.5) * x[RHO] * xCb[E];
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      const int m = P::MOM + k, b = P::MAG + k;
      xb[RHO] = xb[RHO] + xCb[m] * x[m];
      xb[m] = xb[m] + R(2) * x[m] * vsqb + x[RHO] * xCb[m];
      xb[b] = xb[b] + xCb[b] + x[b] * xCb[E];
    }
    if constexpr (P::GLM) xb[P::PSI] = xb[P::PSI] + xCb[P::PSI];
  };
  cons_rev(wL, wLCb, wLb);
  cons_rev(wR, wRCb, wRb);
  if constexpr (P::GLM) {
    // reverse of the shared interface state (B_n*, psi*) of both sides
    constexpr int S = P::PSI;
    const R bsb = wLb[BN] + wRb[BN], psb = wLb[S] + wRb[S];
    const R bq = fdiv(R(0.5) * bsb, ch), pq = R(0.5) * ch * psb;
    wLb[BN] = R(0.5) * bsb + pq;
    wRb[BN] = R(0.5) * bsb - pq;
    wLb[S] = bq + R(0.5) * psb;
    wRb[S] = -bq + R(0.5) * psb;
  }
#pragma unroll
  for (int iw = 0; iw < NW; ++iw) {
    const R t1b = R(0.5) * wLb[iw], t2b = R(-0.5) * wRb[iw];
    const R dmb = t1b * d1a[iw];
    const R d0b = t1b * d1b[iw] + t2b * d2a[iw];
    const R dpb = t2b * d2b[iw];
    if constexpr (ACC) {
      wb[0][iw] += W(-dmb);
      wb[1][iw] += W(wLb[iw] + dmb - d0b);
      wb[2][iw] += W(wRb[iw] + d0b - dpb);
      wb[3][iw] += W(dpb);
    } else {
      wb[0][iw] = W(-dmb);
      wb[1][iw] = W(wLb[iw] + dmb - d0b);
      wb[2][iw] = W(wRb[iw] + d0b - dpb);
      wb[3][iw] = W(dpb);
    }
  }
}

template <class P, class R, int RIEM, int ID, int LIM, bool ACC = false, class W = double>
__device__ __forceinline__ void face_adj_prim(const R (*w)[P::NW], const FVParams& fp, R ch, const R* fbar,
                                              W (*wb)[P::NW]) {
  if constexpr (P::MHD) face_adj_prim_mhd<P, R, RIEM, ID, LIM, ACC, W>(w, fp, ch, fbar, wb);
  else face_adj_prim_hd<P, R, RIEM, ID, LIM, ACC, W>(w, fp, ch, fbar, wb);
}

// Reverse of hd_to_primitive (FP64) for one cell: cbar += (d w / d cons)^T wb
// with rho = rho, v_k = m_k / rho, p = (gamma-1)(e - rho vsq / 2).
template <class P>
__device__ __forceinline__ void prim_rev(const double* cons, const double* wb, const FVParams& fp, double* cbar) {
  constexpr int ND = P::NDIR, RHO = P::RHO, E = P::E;
  const double rho = cons[RHO];
  const double inv = frcp(rho);
  double vv[ND], vsq = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    vv[k] = cons[P::MOM + k] * inv;
    vsq = k == 0 ? vv[k] * vv[k] : vsq + vv[k] * vv[k];
  }
  const double g1 = fp.gamma - 1.0;
  const double pb = wb[E];
  cbar[E] += g1 * pb;
  const double tt = -0.5 * g1 * pb;
  if constexpr (P::MHD) {  // p = (gamma-1)(e - (rho vsq + bsq)/2)
#pragma unroll
    for (int k = 0; k < ND; ++k) cbar[P::MAG + k] += wb[P::MAG + k] + 2.0 * tt * cons[P::MAG + k];
    if constexpr (P::GLM) cbar[P::PSI] += wb[P::PSI];
  }
  double rhob = wb[RHO] + tt * vsq;
  const double vsqb = tt * rho;
  double invb = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double vb = wb[P::MOM + k] + 2.0 * vv[k] * vsqb;
    cbar[P::MOM + k] += vb * inv;
    invb += vb * cons[P::MOM + k];
  }
  rhob -= invb * inv * inv;
  cbar[RHO] += rhob;
}

// Conserved form: recomputes the forward face from the conserved values of
// its four cells and accumulates (d fC / d cons)^T fbar into cbar.
template <class P, class R, int RIEM, int ID, int LIM>
__device__ void face_flux_adj(const double (*cons)[P::NW], const FVParams& fp, R ch, const R* fbar,
                              double (*cbar)[P::NW]) {
  constexpr int NW = P::NW;
  R w[4][NW];
  for (int c = 0; c < 4; ++c) {
    double t[NW];
    to_prim<P>(cons[c], t, fp);
    for (int iw = 0; iw < NW; ++iw) w[c][iw] = R(t[iw]);
  }
  double wb[4][NW];
  face_adj_prim<P, R, RIEM, ID, LIM>(w, fp, ch, fbar, wb);
  for (int c = 0; c < 4; ++c) prim_rev<P>(cons[c], wb[c], fp, cbar[c]);
}

}  // namespace FV_NS
