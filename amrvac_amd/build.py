"""Build the HIP library libamrvac_fv.so in-tree for gfx950, and the host
library libamrvac_amr.so (the native AMR tree routines).

Every stage-kernel translation unit is compiled twice: once with the default
FMA contraction (the production FP64 and mixed kernels) and once with
``-DFV_EXACT -ffp-contract=off`` (FP64 kernels whose every operation rounds
like the CPU restatement's, for bit-level parity checks).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libamrvac_fv.so")
AMR_LIB = os.path.join(HERE, "libamrvac_amr.so")

ARCH = os.environ.get("AMRVAC_FV_ARCH", "gfx950")
KERNEL_UNITS = ["fv_k2_hd", "fv_k2_mhd", "fv_k2_mhd_ndir2", "fv_k3_hd", "fv_k3_mhd", "fv_ct", "fv_ct3"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 arithmetic stays off (v_pk_* forms are avoided)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.abspath(os.path.join(HERE, "..", "include"))]


def _compile(src: str, obj: str, extra: list[str]) -> tuple[str, int, str]:
    cmd = [hipcc()] + common_flags() + extra + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return obj, p.returncode, p.stdout + p.stderr


def build(verbose: bool = False, jobs: int | None = None, csrc: str = CSRC, build_dir: str = BUILD,
          lib: str = LIB) -> str:
    """Compile every unit under ``csrc`` and link ``lib`` (defaults: the
    package's own sources and in-tree library)."""
    os.makedirs(build_dir, exist_ok=True)
    tasks = []
    for u in KERNEL_UNITS:
        src = os.path.join(csrc, u + ".hip")
        tasks.append((src, os.path.join(build_dir, u + ".o"), []))
        tasks.append((src, os.path.join(build_dir, u + "_x.o"), ["-DFV_EXACT", "-ffp-contract=off"]))
    # the dt / volume-average / ghost-fill kernels are memory bound; they are
    # built without contraction so that dt and cmax_global round like the
    # CPU restatement's in every precision mode
    tasks.append((os.path.join(csrc, "fv_capi.hip"), os.path.join(build_dir, "fv_capi.o"), ["-ffp-contract=off"]))
    # adjoint stage kernels (production precision forms only)
    tasks.append((os.path.join(csrc, "fv_adj.hip"), os.path.join(build_dir, "fv_adj.o"), []))
    jobs = jobs or min(len(tasks), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda t: _compile(*t), tasks))
    failed = [r for r in results if r[1] != 0]
    for obj, rc, out in results:
        if verbose or rc != 0:
            if out.strip():
                sys.stderr.write(f"--- {os.path.basename(obj)} (rc={rc})\n{out}\n")
    if failed:
        raise RuntimeError("HIP compilation failed: " + ", ".join(os.path.basename(f[0]) for f in failed))
    objs = [t[1] for t in tasks]
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, lib)
    return lib


def build_host(lib: str = AMR_LIB, cxx: str | None = None) -> str:
    """Compile the native AMR tree routines (csrc/amr_tree.cpp, host code
    only) into ``lib`` with the host C++ compiler."""
    cxx = cxx or os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = lib + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I",
           os.path.abspath(os.path.join(HERE, "..", "include")), os.path.join(CSRC, "amr_tree.cpp"), "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("AMR tree library build failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host())
