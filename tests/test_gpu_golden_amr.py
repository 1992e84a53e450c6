"""The device against the oracle and the reference's maintained AMR
regression logs (tests/golden_setups.py): 3-D MHD in 10^3 and 8^3 blocks
with Linde cleaning across level jumps or GLM, 2-D MHD with two vector
components, 2-D HD, through the host's regrid every step with the device
Löhner estimate and remap.  The contraction-free build must follow the
oracle bit for bit (leaves, dt sequence, state); the production build must
reproduce the reference's log at the compare_logs tolerance."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd.amr import AMRIntegrator
from amrvac_amd.host import logs_match
from oracle.oracle import OracleFV

from golden_setups import golden, setup

pytestmark = pytest.mark.gpu

# (setup, scheme, time_max for the bitwise comparison with the oracle)
EXACT = [
    ("kh3d", "3step_hll_cada", 3.0e-3),
    ("kh3d", "3step_hlld_cada", 3.0e-3),
    ("bw3d", "3step_hll_cada", 4.0e-3),
    ("bw3d", "2step_tvdlf_mm", 4.0e-3),
    ("bw3d", "4step_hll_mc", 4.0e-3),
    ("ot2d", "3step_hll_cada", 2.0e-2),
    ("ot2d", "2step_tvdlf_mm", 2.0e-2),
    ("kh2d", "4step_hllc_ko", 1.0e-2),
]


def _device(precision):
    from amrvac_amd.fv import DeviceFV
    return lambda p, f: DeviceFV(p, f, precision=precision)


@pytest.mark.parametrize("case,scheme,tmax", EXACT)
def test_maintained_setup_exact_matches_oracle_bitwise(case, scheme, tmax):
    print(f"[golden amr] {case} {scheme} fp64_exact vs oracle to t={tmax}", flush=True)
    res = {}
    for name, mk in (("device", _device("fp64_exact")), ("oracle", lambda p, f: OracleFV(p, f, threads=16))):
        par, forest, ap, init = setup(case, scheme, tmax)
        ti = AMRIntegrator(par, forest, ap, mk, init)
        st = ti.run()
        res[name] = ([(n.level, n.ig) for n in forest.leaves], st.dts, forest.interior(ti.be.get_state(0)),
                     np.array(st.log), ti.regrids)
    d, o = res["device"], res["oracle"]
    assert d[0] == o[0]
    assert d[1] == o[1]
    assert np.array_equal(d[2], o[2])
    # the log's volume sums reduce in another order on the device
    assert np.allclose(d[3], o[3], rtol=1e-12, atol=1e-14)


LOGS = [
    ("kh3d", "3step_hll_cada"), ("kh3d", "3step_hlld_cada"),
    ("bw3d", "3step_hll_cada"), ("bw3d", "3step_hlld_cada"), ("bw3d", "4step_hll_mc"), ("bw3d", "2step_tvdlf_mm"),
    ("ot2d", "3step_hll_cada"), ("ot2d", "2step_tvdlf_mm"),
    ("kh2d", "3step_hll_cada"), ("kh2d", "4step_hllc_ko"),
]


@pytest.mark.parametrize("case,scheme", LOGS)
def test_maintained_setup_fp64_reproduces_reference_log(case, scheme):
    print(f"[golden amr] {case} {scheme} fp64 vs the reference's log", flush=True)
    par, forest, ap, init = setup(case, scheme)
    ti = AMRIntegrator(par, forest, ap, _device("fp64"), init)
    got = np.array(ti.run().log)
    ref = golden(case, scheme)
    assert got.shape == ref.shape
    assert logs_match(got, ref)
