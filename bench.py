#!/usr/bin/env python3
"""Throughput of the MI355X finite-volume block update on the 3-D ideal MHD
benchmark (Orszag-Tang + GLM, 256^3 in 16^3 blocks, ssprk3 / HLL / cada3,
CFL 0.6), in Mcell-updates/s: one cell-update = one interior cell advanced by
one RK stage (the reference's own counter, src/amrvac.t:351,375).

One "step" = setdt (device CFL reduction) + the three ssprk3 stages, on a
level resident in HBM.  Prints one JSON line; see --help.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def workload_par(precision: str):
    from amrvac_amd.host import Par
    return Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
               flux_scheme="hll", limiter="cada3", typeboundspeed="Einfeldt", courantpar=0.6,
               precision=precision)


def workload_mesh(n: int):
    from amrvac_amd.mesh import UniformMesh
    return UniformMesh(3, (n, n, n), (16, 16, 16), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), ("periodic",) * 6)


def run_device(par, mesh, w0, steps, warmup, label):
    from amrvac_amd.fv import DeviceFV
    from amrvac_amd.host import TimeIntegrator
    be = DeviceFV(par, mesh)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    for _ in range(warmup):
        ti.step()
    be.sync()
    be.sweep_timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ti.step()
    be.sync()
    t1 = time.perf_counter()
    avg_ms, nlaunch = be.sweep_timing()
    print(f"[bench] {label}: {steps} steps in {t1 - t0:.4f} s, sweep avg {avg_ms:.4f} ms x {nlaunch}",
          file=sys.stderr, flush=True)
    return be, ti, t1 - t0, avg_ms, nlaunch


def cpu_baseline(seconds: float, n: int = 64):
    """The FP64 CPU oracle on a bounded sample of the same workload."""
    from oracle.oracle import OracleFV
    from amrvac_amd import problems
    from amrvac_amd.host import TimeIntegrator
    threads = os.cpu_count() or 1
    par = workload_par("fp64")
    mesh = workload_mesh(n)
    be = OracleFV(par, mesh, threads=threads)
    be.set_state(0, problems.orszag_tang_3d(mesh))
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.step()  # warm-up step (page faults, thread start)
    t0 = time.perf_counter()
    k = 0
    while True:
        ti.step()
        k += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    cells = n ** 3
    return {"value": cells * 3 * k / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{k} ssprk3 steps of the same 3-D MHD-GLM setup at {n}^3 in 16^3 blocks, "
                      f"FP64 CPU restatement, OpenMP over {threads} threads"}


def adjoint_cost(kind: str, reps: int = 2):
    """Cost of the gradient against the forward run it differentiates, on the
    differentiable workloads: ``kh2d`` = 2-D HD Kelvin-Helmholtz 128^2,
    threestep / HLLC / vanleer, 200 steps, loss on the final density; ``ot3d`` = 3-D ideal MHD Orszag-Tang 128^3, threestep /
    HLL / vanleer, 100 steps, loss = final magnetic energy, without divergence control (``ot3d_glm``: with GLM;
    ``ot3d_glm_cada3``: GLM and cada3).  Both mixed precision (FP32 faces, FP64 state).  The run keeps
    the stage states of every step on a device tape when they fit in half the free device memory
    (every workload here), so the backward pass recomputes nothing.  Returns forward
    time (CFL dt per step, as a plain run), forward+backward time and
    backward / forward."""
    import torch
    from amrvac_amd import adjoint
    from amrvac_amd.host import Par
    from amrvac_amd.mesh import UniformMesh
    dev = torch.device("cuda")
    if kind == "kh2d":
        n, nsteps = 128, 200
        mesh = UniformMesh(2, (n, n), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
        par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
                  limiter="vanleer", courantpar=0.8)
        idx = torch.as_tensor(adjoint.block_index(mesh), device=dev)
        run = adjoint.SSPRK3Run(par, mesh, nsteps, precision="mixed")
        rho, vx, vy, p = adjoint.kh_fields(mesh, 20261202, device=dev, rho_amp=1.0e-3)
        rt, vxt, vyt, pt = adjoint.kh_fields(mesh, 20261203, device=dev, rho_amp=1.0e-3)
        with torch.no_grad():
            target = run.forward(adjoint.to_blocks(adjoint.hd_conserved(rt, [vxt, vyt], pt, 1.4), idx))[:, 0].clone()
        xs = (rho, vx, vy)

        def loss(xs):
            u0 = adjoint.to_blocks(adjoint.hd_conserved(xs[0], [xs[1], xs[2]], p, 1.4), idx)
            return ((adjoint.run_differentiable(u0, run)[:, 0] - target) ** 2).mean()
        workload = "2-D HD Kelvin-Helmholtz 128^2 (16^2 blocks), threestep/HLLC/vanleer, 200 steps"
    else:
        # ot3d: the differentiable 3-D workload (vanleer, no divergence
        # control); ot3d_glm: the same with the GLM cleaning of the 256^3
        # setup it takes its initial condition from; ot3d_glm_cada3: the
        # 256^3 setup's scheme (GLM, cada3)
        glm = kind in ("ot3d_glm", "ot3d_glm_cada3")
        cada3 = kind == "ot3d_glm_cada3"
        n, nsteps = 128, 100
        mesh = UniformMesh(3, (n, n, n), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
        par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm" if glm else "none",
                  time_stepper="threestep", flux_scheme="hll", limiter="cada3" if cada3 else "vanleer",
                  courantpar=0.6)
        idx = torch.as_tensor(adjoint.block_index(mesh), device=dev)
        run = adjoint.SSPRK3Run(par, mesh, nsteps, precision="mixed")
        rho, v, p, bs = adjoint.orszag_tang_fields(mesh, 20261202, device=dev)
        psi = torch.zeros_like(rho) if glm else None
        xs = (torch.tensor(1.0, dtype=torch.float64, device=dev), rho)

        def loss(xs):
            u0 = adjoint.to_blocks(adjoint.mhd_conserved(xs[1], v, p, [xs[0] * c for c in bs], 5.0 / 3.0, psi=psi),
                                   idx)
            bf = adjoint.run_differentiable(u0, run)[:, 5:8]
            return 0.5 * (bf * bf).sum() / n ** 3
        workload = (f"3-D ideal MHD Orszag-Tang 128^3 (16^3 blocks), threestep/HLL/"
                    f"{'cada3' if cada3 else 'vanleer'}{' + GLM' if glm else ''}, 100 steps")
    tf = tg = float("inf")
    for _ in range(reps + 1):  # the first pass warms up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            loss(xs)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        xg = [x.clone().requires_grad_(True) for x in xs]
        torch.autograd.grad(loss(xg), xg)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf, tg = min(tf, t1 - t0), min(tg, t2 - t1)
    print(f"[bench] adjoint {kind}: forward {tf * 1e3:.1f} ms, forward+backward {tg * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    tape = "stage tape of the whole run (no recompute)" if run.every >= nsteps else f"checkpoints every {run.every}"
    return {"workload": workload + ", " + tape, "precision": "mixed", "forward_ms": round(tf * 1e3, 2),
            "forward_backward_ms": round(tg * 1e3, 2), "backward_over_forward": round((tg - tf) / tf, 3)}


def amr_blast(precision: str, steps: int, noise: float = 0.01):
    """The block-structured AMR workload: 2-D HD blast wave (rho = 1 with
    seeded 1% noise, p = 1e-2, p = 10 for r < 0.05 at the centre), base
    grid 64^2 in 16^2 blocks, 4 levels (effective 512^2), Löhner on rho and
    e (weights 0.5/0.5, tol 0.1) with a regrid every step, twostep / TVDLF
    / minmod / cmaxmean, CFL 0.9.  Estimation and data movement of the
    regrid run on the device; the tree stays on the host.  Cell-updates
    count the leaf cells of each step times its two stages.  The 1% noise
    puts every block at the finest level (the estimator sees it as
    structure); noise=0 gives a tree that follows the blast."""
    from amrvac_amd import problems
    from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
    from amrvac_amd.fv import DeviceFV
    from amrvac_amd.host import Par
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
              limiter="minmod", typeboundspeed="cmaxmean", courantpar=0.9, precision=precision)
    forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("cont",) * 4, max_level=4)
    ap = AMRPar(refine_max_level=4, w_refine_weight={0: 0.5, 3: 0.5})

    def init(f, blocks):
        return problems.blast_wave_blocks(f, blocks, rho0=1.0, p_out=1.0e-2, p_in=10.0, radius=0.05, noise=noise)
    ti = AMRIntegrator(par, forest, ap, lambda p, f: DeviceFV(p, f), init)
    ti.run(nsteps=2)
    ti.be.sync()
    regrids0 = ti.regrids
    # leaves of each step as it is advanced (the tree changes between steps)
    leaves = []
    advect = ti.advect

    def counted_advect():
        leaves.append(forest.nblocks)
        advect()
    ti.advect = counted_advect
    t0 = time.perf_counter()
    ti.run(nsteps=steps)
    ti.be.sync()
    dt = time.perf_counter() - t0
    ti.advect = advect
    updates = sum(leaves) * 256 * 2
    steps = len(leaves)
    print(f"[bench] amr blast {precision}: {steps} steps in {dt:.3f} s, {forest.nblocks} leaves, "
          f"levels {forest.levmin}-{forest.levmax}, {ti.regrids - regrids0} regrids", file=sys.stderr, flush=True)
    return {"value": round(updates / dt / 1e6, 2), "unit": "Mcell-updates/s", "ms_per_step": round(dt / steps * 1e3, 3),
            "noise": noise,
            "steps": steps, "leaves": forest.nblocks, "levmax": forest.levmax,
            "regrids": ti.regrids - regrids0}


def ct_field_loop(steps: int, warmup: int = 2, n: int = 1024):
    """Constrained transport: the 2.5-D advected field loop
    (tests/mhd/field_loop_2.5D, staggered face field, type_ct uct_contact)
    on n x n/2 cells in 16^2 blocks, threestep / HLL / cada3 / cmaxmean,
    CFL 0.9, FP64 (and the mixed form for uct_contact); cell-updates per RK
    stage as for the main workload."""
    from amrvac_amd import problems
    from amrvac_amd.fv import DeviceFV
    from amrvac_amd.host import Par, TimeIntegrator
    from amrvac_amd.mesh import UniformMesh
    out = {"workload": f"2.5-D MHD field loop with constrained transport, {n}x{n // 2} in 16^2 blocks, "
                       "threestep/HLL/cada3/cmaxmean, FP64 (uct_contact_mixed: the mixed form)"}
    for ct, prec in (("uct_contact", "fp64"), ("average", "fp64"), ("uct_hll", "fp64"), ("uct_contact", "mixed")):
        par = Par(physics="mhd", ndim=2, ndir=3, gamma=5.0 / 3.0, typedivbfix="ct", type_ct=ct,
                  time_stepper="threestep", flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean",
                  courantpar=0.9, precision=prec)
        mesh = UniformMesh(2, (n, n // 2), (16, 16), (-1.0, -0.5), (1.0, 0.5), ("periodic",) * 4)
        be = DeviceFV(par, mesh)
        be.set_state(0, problems.field_loop_ct(mesh))
        ti = TimeIntegrator(par, mesh, be)
        for _ in range(warmup):
            ti.step()
        be.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ti.step()
        be.sync()
        dt = time.perf_counter() - t0
        cells = n * (n // 2)
        print(f"[bench] ct {ct} {prec}: {steps} steps in {dt:.4f} s", file=sys.stderr, flush=True)
        out[ct if prec == "fp64" else f"{ct}_mixed"] = {"value": round(cells * 3 * steps / dt / 1e6, 2), "unit": "Mcell-updates/s",
                   "ms_per_step": round(dt / steps * 1e3, 4), "steps": steps}
        del be
    return out


def ct3d_field_loop(steps: int, warmup: int = 2, n: int = 128):
    """Constrained transport in 3-D: the Gardiner & Stone tilted field loop
    on n x n x 2n cells in 16^3 blocks, threestep / HLL / cada3 / cmaxmean,
    CFL 0.4, FP64 (and the mixed form for uct_contact); cell-updates per RK
    stage as for the main workload."""
    from amrvac_amd import problems
    from amrvac_amd.fv import DeviceFV
    from amrvac_amd.host import Par, TimeIntegrator
    from amrvac_amd.mesh import UniformMesh
    out = {"workload": f"3-D MHD field loop (Gardiner & Stone 2008) with constrained transport, "
                       f"{n}x{n}x{2 * n} in 16^3 blocks, threestep/HLL/cada3/cmaxmean, FP64 "
                       "(uct_contact_mixed: the mixed form)"}
    mesh = UniformMesh(3, (n, n, 2 * n), (16, 16, 16), (-0.5, -0.5, -1.0), (0.5, 0.5, 1.0), ("periodic",) * 6)
    w0 = problems.field_loop_3d_ct(mesh)
    for ct, prec in (("uct_contact", "fp64"), ("average", "fp64"), ("uct_hll", "fp64"), ("uct_contact", "mixed")):
        par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="ct", type_ct=ct,
                  time_stepper="threestep", flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean",
                  courantpar=0.4, precision=prec)
        be = DeviceFV(par, mesh)
        be.set_state(0, w0)
        ti = TimeIntegrator(par, mesh, be)
        for _ in range(warmup):
            ti.step()
        be.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ti.step()
        be.sync()
        dt = time.perf_counter() - t0
        cells = n * n * 2 * n
        print(f"[bench] ct3d {ct} {prec}: {steps} steps in {dt:.4f} s", file=sys.stderr, flush=True)
        out[ct if prec == "fp64" else f"{ct}_mixed"] = {"value": round(cells * 3 * steps / dt / 1e6, 2),
                                                        "unit": "Mcell-updates/s",
                                                        "ms_per_step": round(dt / steps * 1e3, 4), "steps": steps}
        del be
    return out


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            return json.load(fh).get(name)
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256, help="cells per side (multiple of 16)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost legs")
    ap.add_argument("--no-amr", action="store_true", help="skip the AMR blast-wave leg")
    ap.add_argument("--amr-steps", type=int, default=200)
    ap.add_argument("--no-ct", action="store_true", help="skip the constrained-transport leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, copy one state slot device-to-device (a plain copy of "
                         "known size, the reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the block update runs on a single MI355X in this build "
              "(multi-GPU sharding is a design note, not an implemented path).", file=sys.stderr)
        return 2
    from amrvac_amd import problems
    n = args.size
    mesh = workload_mesh(n)
    t0 = time.perf_counter()
    w0 = problems.orszag_tang_3d(mesh)
    print(f"[bench] initial condition {n}^3: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cells = n ** 3
    stages = 3
    nw = 9
    # algorithmic HBM bytes per cell and RK stage, FP64 state in both forms:
    # stage input + RK base + output, 3 * nw * 8 = 216 B; stage 1's RK base
    # is its stage input, read once, so the launches average (2 + 3 + 3) / 3
    # slot passes = 192 B, reported beside it
    b_alg = 3 * nw * 8
    b_alg_min = (2 + 3 + 3) * nw * 8 / 3
    var_names = ["rho", "m1", "m2", "m3", "e", "b1", "b2", "b3", "psi"]

    par64 = workload_par("fp64")
    be64, ti64, t64, ms64, _ = run_device(par64, mesh, w0, args.steps, args.warmup, "fp64")
    value64 = cells * stages * args.steps / t64 / 1e6
    if args.copy_probe:
        # reference points for the HBM byte counters, each of one slot's
        # size: the runtime's device-to-device copy, an elementwise FP64
        # kernel (read and write, 8-byte loads as the stage kernel's) and an
        # FP64 reduction (read only)
        be64.copy(1, 0)
        be64.sync()
        import torch
        nslot = mesh.nblocks * nw * 4096
        xa = torch.ones(nslot, dtype=torch.float64, device="cuda")
        xb = torch.empty_like(xa)
        torch.cuda.synchronize()
        torch.add(xa, 1.0, out=xb)
        torch.cuda.synchronize()
        s = float(xb.sum())
        del xa, xb
        print(f"[bench] copy probes: {nslot * 8} bytes read and written (copy, elementwise add), "
              f"{nslot * 8} bytes read (sum {s:.0f})", file=sys.stderr)
    achieved = b_alg * cells / (ms64 * 1e-3) / 1e9
    result = {
        "metric": "Mcell-updates/s 3D MHD 256^3, FP64 vs mixed @L1<=1e-5; adjoint/forward cost",
        "value": round(value64, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Orszag-Tang initial condition)",
        "config": {"workload": f"3-D ideal MHD + GLM, Orszag-Tang, {n}^3 in 16^3 blocks "
                               f"({mesh.nblocks} blocks), ssprk3 / HLL / cada3 / Einfeldt, CFL 0.6",
                   "cells": cells, "stages_per_step": stages},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("fp64"),
                     "kernel": "fv::sweep_kernel<MHD-GLM 3-D, double, HLL>", "avg_launch_ms": round(ms64, 4),
                     "alg_bytes_per_cell": b_alg, "alg_bytes_per_launch": int(b_alg * cells),
                     "frac_at_192B_per_cell": round(b_alg_min * cells / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "target_frac": 0.6},
    }
    if not args.no_mixed:
        w64 = mesh.interior(be64.get_state(0))
        del be64
        parmx = workload_par("mixed")
        bemx, timx, tmx, msmx, _ = run_device(parmx, mesh, w0, args.steps, args.warmup, "mixed")
        wmx = mesh.interior(bemx.get_state(0))
        del bemx
        l1 = [float(np.abs(wmx[:, iw] - w64[:, iw]).sum() / max(np.abs(w64[:, iw]).sum(), 1e-300))
              for iw in range(nw)]
        valmx = cells * stages * args.steps / tmx / 1e6
        achmx = b_alg * cells / (msmx * 1e-3) / 1e9
        result["mixed"] = {"value": round(valmx, 2), "speedup_vs_fp64": round(valmx / value64, 3),
                           "speedup_target": 1.6,
                           "rel_l1_vs_fp64": {var_names[iw]: float(f"{l1[iw]:.3e}") for iw in range(nw)},
                           "density_l1_target": 1.0e-5,
                           "mass_drift_vs_fp64": float(f"{abs(wmx[:, 0].sum() - w64[:, 0].sum()) / w64[:, 0].sum():.3e}"),
                           "energy_drift_vs_fp64": float(f"{abs(wmx[:, 4].sum() - w64[:, 4].sum()) / w64[:, 4].sum():.3e}"),
                           "roofline_frac": round(achmx / HBM_PEAK_GBS, 4),
                           "roofline_frac_at_192B_per_cell": round(b_alg_min * cells / (msmx * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "roofline_target_frac": 0.6, "avg_launch_ms": round(msmx, 4),
                           "traffic": load_traffic("mixed")}
        del w64, wmx
    if not args.no_adjoint:
        result["adjoint_forward_cost"] = {k: adjoint_cost(k) for k in ("kh2d", "ot3d", "ot3d_glm", "ot3d_glm_cada3")}
    if not args.no_amr:
        result["amr_blast"] = {"workload": "2-D HD blast wave, 64^2 base in 16^2 blocks, 4 levels (512^2 effective), "
                                           "Loehner regrid every step, twostep/TVDLF/minmod",
                               **{p: amr_blast(p, args.amr_steps) for p in ("fp64", "mixed")},
                               "fp64_noise_free": amr_blast("fp64", args.amr_steps, noise=0.0)}
    if not args.no_ct:
        result["ct_field_loop"] = ct_field_loop(20)
        result["ct3d_field_loop"] = ct3d_field_loop(10)
    if args.cpu_seconds > 0:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        except Exception as exc:  # the oracle library is optional on the measuring machine
            result["cpu_baseline"] = {"value": None, "error": str(exc)}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
