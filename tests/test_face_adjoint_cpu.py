"""The hand-written reverse-mode face flux (fv_adjoint.h) against central
finite differences of the forward face flux, compiled for the CPU from the
same device headers in the contraction-free (FV_EXACT) form: hydrodynamics
and ideal MHD, without divergence control or with GLM, minmod / vanleer /
cada3, HLLC / HLL / TVDLF, every direction."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_cpu")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_face_flux_adjoint_vs_finite_differences(tmp_path):
    exe = str(tmp_path / "face_adj_check")
    cmd = ["g++", "-O1", "-std=c++17", "-DFV_EXACT", "-ffp-contract=off", "-I", ".", "-I", "../../include",
           "face_adj_check.cpp", "-o", exe]
    subprocess.run(cmd, cwd=HERE, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("worst rel err") == 17
