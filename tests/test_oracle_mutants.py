"""The closed-form checks of test_oracle_math.py must be able to fail.

Each case plants one plausible mistake (a flipped sign, a dropped factor, a
swapped operand) in a copy of the oracle's C source, builds it in a temporary
directory and runs the check that is meant to catch it.  A check that still
passes on the planted mistake pins nothing, and this test fails.
"""
from __future__ import annotations

import os
import subprocess

import pytest

from test_oracle_ct3d import CT3D_CHECKS
from test_oracle_math import ALL_CHECKS

CHECKS = {**ALL_CHECKS, **CT3D_CHECKS}

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# name: (text in amrvac_oracle.c, its replacement, occurrence to replace
# (0-based), check that must fail)
MUTANTS = {
    # constrained transport
    "ct_vbar_second_average": ("(cbmax * l2 + cbmin * r1) / (cbmax + cbmin);",
                               "(cbmax * l2 + cbmin * wRp[i][s->mom[iE]]) / (cbmax + cbmin);", 0, "ct_golden_hll"),
    "ct_average_factor": ("fE[idir * n + c] = quarter * (FM(c, c1, idim2)", "fE[idir * n + c] = half * (FM(c, c1, idim2)", 0,
                          "ct_golden_average"),
    "ct_contact_upwind": ("if (vn1[c] > 0.0) ELC = EL[c];", "if (vn1[c] > 0.0) ELC = EL[c1p];", 0, "ct_golden_contact"),
    "ct_circulation_sign": ("if (l == 1) circ = circ + dE;", "if (l == 1) circ = circ - dE;", 0, "ct_divb"),
    "ct_face_to_center": ("(ws[d * n + c] * sC + ws[d * n + SH(s, c, d, -1)] * sC)", "(ws[d * n + c] * sC + ws[d * n + c] * sC)",
                          0, "ct_golden_contact"),
    "ct_normal_field_of_prims": ("wLC[bi] = bn; wRC[bi] = bn; wLp[i][bi] = bn; wRp[i][bi] = bn;",
                                 "wLC[bi] = bn; wRC[bi] = bn;", 0, "ct_golden_contact"),
    # an edge field taken one cell off along its own edge (3-D only); the
    # extrusions see it only where the shift reaches the block edge, the
    # mirror check everywhere
    "ct3d_edge_offset_along_edge": (
        "const double dE = fE[idir * n + c] - fE[idir * n + SH(s, c, idim2, -1)];",
        "const long ce = nd == 3 ? SH(s, c, idir, 1) : c;\n"
        "              const double dE = fE[idir * n + ce] - fE[idir * n + SH(s, ce, idim2, -1)];", 0, "ct3d_mirror"),
    "ct3d_contact_cell_efield_sign": (
        "ECC[0 * n + c] = wp[b2 * n + c] * wp[m3 * n + c] - wp[b3 * n + c] * wp[m2 * n + c];",
        "ECC[0 * n + c] = wp[b3 * n + c] * wp[m2 * n + c] - wp[b2 * n + c] * wp[m3 * n + c];", 0, "ct3d_extruded"),
    # the cell field re-derived from the faces after the Hancock predictor
    "hancock_ct_face_to_center": ("      hancock(&sb, qdt, wCT, wnew, cmax_global, &c);",
                                  "      { hancock(&sb, qdt, wCT, wnew, cmax_global, &c);\n"
                                  "        if (s->ct) ct_face_to_center(&sb, wnew); }", 0, "hancock_ct"),
    "hll_dissipation_sign": ("+ cmin * cmax * (wRC[iw] - wLC[iw])) /\n               (cmax - cmin);\n    }\n  }\n}",
                             "- cmin * cmax * (wRC[iw] - wLC[iw])) /\n               (cmax - cmin);\n    }\n  }\n}",
                             0, "sod_hll_cada3"),
    "hllc_energy_pressure_work": ("+ lambdaCD * pCD) * csmls", "- lambdaCD * pCD) * csmls", 0, "sod_hllc_vanleer"),
    "hllc_contact_speed": ("lambdaCD = whll[mn] / whll[rho];", "lambdaCD = whll[mn] / whll[e];", 0,
                           "smooth_hllc_vanleer"),
    "hd_energy_flux_no_pressure": ("f[s->e_] = vn * (wC[s->e_] + w[s->e_]);", "f[s->e_] = vn * wC[s->e_];", 0,
                                   "flux_consistency"),
    "hd_momentum_flux_no_pressure": ("f[s->mom[idim]] = f[s->mom[idim]] + w[s->e_];",
                                     "f[s->mom[idim]] = f[s->mom[idim]];", 0, "flux_consistency"),
    "hlld_contact_speed_pressure_sign": ("- ptR + ptL) / (suR - suL);", "+ ptR - ptL) / (suR - suL);", 0,
                                         "hlld_discontinuities"),
    "hlld_alfven_state_sign": ("(w1R[b2] - w1L[b2]) * signBx) * tmp;", "(w1R[b2] - w1L[b2]) * Bx) * tmp;", 0,
                               "hlld_paper"),
    "hlld_tangential_momentum_term": ("w1R[m2] = wRp[m2] - Bx * BR[b2] * (sm - wRp[m1]) * r1R;",
                                      "w1R[m2] = wRp[m2] - Bx * BR[b2] * (sm - wRp[m1]);", 0,
                                      "hlld_paper"),
    "hlld_star_energy_work": ("+ w1R[p_] * sm + Bx * (vbR - vb1R)) / (sR - sm);",
                              "+ w1R[p_] * sm - Bx * (vbR - vb1R)) / (sR - sm);", 0, "hlld_paper"),
    "mhd_hllc_pressure_bn_term": ("- wSub[mn] * vSub +\n                         wCD[bn] * wCD[bn];",
                                  "- wSub[mn] * vSub -\n                         wCD[bn] * wCD[bn];", 0, "flux_consistency"),
    "mhd_hllc_tangential_momentum_field": ("fSub[m] - wCD[bn] * wCD[s->mag[c]]) / den;",
                                           "fSub[m] + wCD[bn] * wCD[s->mag[c]]) / den;", 0, "flux_consistency"),
    "mhd_hllc_energy_vdotb": ("lambdaCD * pCD - VdotBCD * wCD[bn]) / den;", "lambdaCD * pCD) / den;", 0,
                              "flux_consistency"),
    "mhd_energy_flux_sign": ("- bn * bv;", "+ bn * bv;", 0, "flux_consistency"),
    "mhd_induction_sign": ("vn * w[s->mag[c]] - bn * w[s->mom[c]]", "vn * w[s->mag[c]] + bn * w[s->mom[c]]", 0,
                           "flux_consistency"),
    "mhd_slow_instead_of_fast": ("sqrt(half * (cfast2 + sqrt(avmincs2)))", "sqrt(half * (cfast2 - sqrt(avmincs2)))",
                                 0, "signal_speeds"),
    "mhd_fast_speed_bn_unsquared": ("4.0 * cs * (bn * bn) * inv_rho", "4.0 * cs * fabs(bn) * inv_rho", 0,
                                    "signal_speeds"),
    "einfeldt_missing_sqrt": ("dmean = sqrt(dmean);", "dmean = dmean;", 0, "signal_speeds"),
    "glm_bn_characteristic_sign": ("wLp[b] = half * (wRp[b] + wLp[b] - dPsi / ch);",
                                   "wLp[b] = half * (wRp[b] + wLp[b] + dPsi / ch);", 0, "glm_interface"),
    "glm_psi_characteristic_sign": ("wLp[psi] = half * (wRp[psi] + wLp[psi] - dB * ch);",
                                    "wLp[psi] = half * (wRp[psi] + wLp[psi] + dB * ch);", 0, "glm_interface"),
    "glm_decay_dx_inverted": ("exp(-qdt * ch * s->p.glm_alpha / mindx)", "exp(-qdt * ch * s->p.glm_alpha * mindx)", 0,
                              "glm_decay"),
    "glm_psi_flux_ch_unsquared": ("cmax_global * cmax_global * w[s->mag[idim]]", "cmax_global * w[s->mag[idim]]", 0,
                                  "glm_interface"),
    "minmod_max": ("MAX2(zero, MIN2(fabs(b), tmp * a))", "MAX2(zero, MAX2(fabs(b), tmp * a))", 0, "limiters"),
    "vanleer_half": ("tmp = 2 * MAX2(a * b, zero) / (b + a + qsmall);", "tmp = MAX2(a * b, zero) / (b + a + qsmall);",
                     0, "limiters"),
    "cada3_left_slope": ("A = (two + tmp) * third;", "A = (one + tmp) * third;", 0, "limiters"),
    "cada3_right_slope": ("A = (two + tmp) * third;", "A = (two + tmp) * half;", 1, "limiters"),
    "weno5_candidate_coefficient": ("u22 = 5.0 / 6.0", "u22 = 4.0 / 6.0", 0, "weno5_polynomials"),
    "weno5_smoothness_indicator": ("b2 = bc1 * (t * t) + bc2 * (q * q);", "b2 = bc1 * (t * t) + bc2 * (t * q);", 0,
                                   "smooth_tvdlf_weno5"),
    "cada3_smooth_branch_dropped": ("if (tmpeta <= one - cadepsilon) {", "if (tmpeta <= -one) {", 0,
                                    "smooth_hll_cada3"),
    "weno5_right_stencil_shift": ("f3 = u31 * wp + u32 * w0 + u33 * wm;", "f3 = u31 * wp + u32 * w0 + u33 * w0;", 0,
                                  "weno5_polynomials"),
    "flux_divergence_sign": ("wnew[iw * n + c] + fC[i][iw] - fC[i - 1][iw]",
                             "wnew[iw * n + c] - fC[i][iw] + fC[i - 1][iw]", 0, "sod_hllc_vanleer"),
    "flux_divergence_offset": ("wnew[iw * n + c] + fC[i][iw] - fC[i - 1][iw]",
                               "wnew[iw * n + c] + fC[i][iw] - fC[i][iw]", 0, "sod_hll_cada3"),
    "hancock_predictor_sign": ("dxinv * (fL[iw] - fR[iw])", "dxinv * (fR[iw] - fL[iw])", 0,
                               "smooth_twostep_vanleer"),
    "weno5_fix_limiter_dropped": ("if (check_w_prim_fails(s, tl[i]) || check_w_prim_fails(s, tr[i])) continue;",
                                  "if (check_w_prim_fails(s, tl[i]) || check_w_prim_fails(s, tr[i])) {}", 0,
                                  "weno5_fix_limiter"),
    "weno5_fix_limiter_left_only": ("if (check_w_prim_fails(s, tl[i]) || check_w_prim_fails(s, tr[i])) continue;",
                                    "if (check_w_prim_fails(s, tl[i])) continue;", 0, "weno5_fix_limiter"),
    "weno5_fix_limiter_pressure_unchecked": ("if (w[s->e_] < s->p.small_pressure) return 1;",
                                             "if (w[s->e_] < -1.0e300) return 1;", 0, "weno5_fix_limiter"),
    "small_values_after_update_dropped": ("  if (s->p.fix_small_values && small_values_interior(s, wnew)) *crash = 1;\n  add_sources",
                                          "  add_sources", 0, "small_values_crash"),
    "small_values_mhd_magnetic_energy_dropped": ("return w[s->e_] - half * (msq / w[s->rho_] + bsq) < small_e;",
                                                 "return w[s->e_] - half * (msq / w[s->rho_]) < small_e;", 0,
                                                 "small_values_crash"),
    "linde_level_jump_rule_dropped": ("if (t == 2 || t == 4) {", "if (t == -7) {", 0, "linde_level_jump"),
    "tvdlf_no_dissipation": ("double fac = -0.5 * s->p.tvdlfeps * cmax;", "double fac = 0.0 * s->p.tvdlfeps * cmax;",
                             0, "sod_tvdlf_weno5"),
}


def _replace_nth(src: str, old: str, new: str, nth: int) -> str:
    pos = -1
    for _ in range(nth + 1):
        pos = src.find(old, pos + 1)
        if pos < 0:
            raise AssertionError(f"mutation site not found: {old!r} (occurrence {nth})")
    return src[:pos] + new + src[pos + len(old):]


@pytest.fixture(scope="module")
def oracle_source():
    with open(os.path.join(ORACLE_DIR, "amrvac_oracle.c")) as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(name, oracle_source, tmp_path):
    old, new, nth, check = MUTANTS[name]
    src = _replace_nth(oracle_source, old, new, nth)
    assert src != oracle_source
    c_path = tmp_path / "amrvac_oracle.c"
    c_path.write_text(src)
    so_path = tmp_path / f"liborc_{name}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-I", ORACLE_DIR,
                    "-shared", "-o", str(so_path), str(c_path), "-lm"], check=True)
    import numpy as np
    with np.errstate(all="ignore"):
        try:
            msg = CHECKS[check](str(so_path))
        except RuntimeError as e:  # the run crashed: caught as well
            msg = str(e)
    assert msg != "", f"check {check!r} passes with the planted mistake {name!r}"
