"""Snapshot files (.dat, format version 5): the byte layout of
doc/fileformat.md and write_snapshot (src/io/mod_input_output.t:2004-2172),
and restarts from them on uniform and AMR meshes (the oracle backend here;
the device backend in tests/test_gpu_datfile.py)."""
from __future__ import annotations

import os
import struct

import numpy as np

from amrvac_amd import datfile, problems
from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV


def _kh():
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="vanleer", courantpar=0.8)
    mesh = UniformMesh(2, (64, 32), (16, 16), (0.0, 0.0), (2.0, 1.0), ("periodic",) * 4)
    return par, mesh


def test_uniform_snapshot_layout_and_round_trip(tmp_path):
    par, mesh = _kh()
    be = OracleFV(par, mesh)
    be.set_state(0, problems.kelvin_helmholtz_2d(mesh))
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    w = be.get_state(0)
    path = os.path.join(tmp_path, "kh0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, w, it=ti.state.it, global_time=ti.state.global_time,
                                  snapshotnext=4)
    raw = open(path, "rb").read()
    # header fields at the documented positions
    assert struct.unpack_from("<10i", raw, 0) == (5, snap.offset_tree, snap.offset_block, 4, 2, 2, 1, 8, 0, 3)
    assert struct.unpack_from("<d", raw, 40)[0] == ti.state.global_time
    assert raw[40 + 8 + 4 * 8 + 4 * 4 + 2 * 4:][:16] == b"Cartesian_2D    "
    assert snap.w_names == ["rho", "m1", "m2", "e"] and snap.physics_type == "hd"
    assert snap.params == {"gamma": 1.4} and snap.snapshotnext == 5
    assert snap.periodic == (True, True) and snap.domain_nx == (64, 32) and snap.block_nx == (16, 16)
    # the reference's own check of the tree section length (write_snapshot)
    assert snap.offset_block - snap.offset_tree == (snap.nleafs + snap.nparents) * 4 + snap.nleafs * ((1 + 2) * 4 + 2 * 4)
    # blocks follow each other: 2*ndim ghost counts + nw * 16 * 16 doubles
    assert np.all(np.diff(snap.block_offsets) == 2 * 2 * 4 + 4 * 16 * 16 * 8)
    assert snap.block_offsets[-1] + 2 * 2 * 4 + 4 * 16 * 16 * 8 == len(raw)
    # the data: interior cells in Morton order, Fortran order per block
    assert np.array_equal(snap.w, mesh.interior(w))
    assert np.array_equal(snap.ig, mesh.ig) and np.all(snap.level == 1) and np.all(snap.leaf)
    b = 5
    first = struct.unpack_from("<4i3d", raw, int(snap.block_offsets[b]))
    assert first[:4] == (0, 0, 0, 0)
    ng = mesh.nghost
    assert first[4:] == tuple(w[b, 0, ng, ng:ng + 3])  # w(1:3, 1, rho): x fastest


def test_uniform_restart_is_exact(tmp_path):
    par, mesh = _kh()
    w0 = problems.kelvin_helmholtz_2d(mesh)
    ref = OracleFV(par, mesh)
    ref.set_state(0, w0)
    ref.getbc(0)
    TimeIntegrator(par, mesh, ref).run(nsteps=6)
    be = OracleFV(par, mesh)
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    path = os.path.join(tmp_path, "kh0001.dat")
    datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it, global_time=ti.state.global_time)
    snap = datfile.read_snapshot(path)
    be2 = OracleFV(par, mesh)
    be2.set_state(0, datfile.state_from_snapshot(snap, mesh))
    be2.getbc(0)
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=3)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))


def test_amr_snapshot_rebuilds_the_forest(tmp_path):
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="threestep", flux_scheme="hll",
              limiter="cada3", typeboundspeed="cmaxmean", courantpar=0.9)
    forest = AMRForest(2, (32, 32), (16, 16), (0.0, 0.0), (2.0, 2.0), ("cont",) * 4, max_level=3)
    ap = AMRPar(refine_max_level=3, w_refine_weight={0: 0.5, 3: 0.5})
    init = lambda f, b: problems.blast_wave_blocks(f, b)
    ti = AMRIntegrator(par, forest, ap, lambda p, f: OracleFV(p, f, threads=4), init)
    ti.run(nsteps=4)
    assert forest.levmax == 3 and forest.levmin < 3
    path = os.path.join(tmp_path, "bw0000.dat")
    snap = datfile.write_snapshot(path, par, forest, ti.be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    assert snap.nparents > 0 and snap.levmax == 3
    assert snap.offset_block - snap.offset_tree == (snap.nleafs + snap.nparents) * 4 + snap.nleafs * ((1 + 2) * 4 + 2 * 4)
    f2 = datfile.forest_from_snapshot(snap, ("cont",) * 4, max_level=3)
    assert [(n.level, n.ig) for n in f2.leaves] == [(n.level, n.ig) for n in forest.leaves]
    be2 = OracleFV(par, f2, threads=4)
    be2.set_state(0, datfile.state_from_snapshot(snap, f2))
    be2.getbc(0)
    # the restarted state with its ghost cells equals the running one
    assert np.array_equal(be2.get_state(0), ti.be.get_state(0))


def test_amr_restart_continues_exactly(tmp_path):
    """An AMR run restarted from its snapshot (the forest rebuilt from the
    leaf flags, no initial refinement) continues bit for bit, regrids
    included."""
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
              limiter="minmod", typeboundspeed="cmaxmean", courantpar=0.9)
    ap = AMRPar(refine_max_level=3, w_refine_weight={0: 0.5, 3: 0.5})
    init = lambda f, b: problems.blast_wave_blocks(f, b)
    mk = lambda p, f: OracleFV(p, f, threads=4)

    def forest():
        return AMRForest(2, (32, 32), (16, 16), (0.0, 0.0), (2.0, 2.0), ("cont",) * 4, max_level=3)

    f_ref = forest()
    ref = AMRIntegrator(par, f_ref, ap, mk, init)
    ref.run(nsteps=8)
    f1 = forest()
    ti = AMRIntegrator(par, f1, ap, mk, init)
    ti.run(nsteps=4)
    path = os.path.join(tmp_path, "bw0001.dat")
    snap = datfile.write_snapshot(path, par, f1, ti.be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    f2 = datfile.forest_from_snapshot(datfile.read_snapshot(path), ("cont",) * 4, max_level=3)
    ti2 = AMRIntegrator(par, f2, ap, mk, lambda f, b: datfile.state_from_snapshot(snap, f)[b], restart=True)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=4)
    assert ti2.state.dts == ref.state.dts[4:]
    assert [(n.level, n.ig) for n in f2.leaves] == [(n.level, n.ig) for n in f_ref.leaves]
    assert np.array_equal(f2.interior(ti2.be.get_state(0)), f_ref.interior(ref.be.get_state(0)))


def test_staggered_snapshot_restart_is_exact(tmp_path):
    """Constrained transport: the face field ws(ixM-1:ixM, 1:ndim) follows
    each block's cell data and the header's staggered flag is set; a run
    restarted from the snapshot continues bit for bit."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_golden import field_loop_setup
    par, mesh = field_loop_setup("3step_hll_cada_ct")
    w0 = problems.field_loop_ct(mesh)
    ref = OracleFV(par, mesh)
    ref.set_state(0, w0)
    ref.getbc(0)
    TimeIntegrator(par, mesh, ref).run(nsteps=6)
    be = OracleFV(par, mesh)
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    path = os.path.join(tmp_path, "fl0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it, global_time=ti.state.global_time)
    raw = open(path, "rb").read()
    assert snap.staggered and snap.ws.shape == (mesh.nblocks, 2, 17, 17)
    assert np.all(np.diff(snap.block_offsets) == 2 * 2 * 4 + 8 * 16 * 16 * 8 + 2 * 17 * 17 * 8)
    assert snap.block_offsets[-1] + 2 * 2 * 4 + 8 * 16 * 16 * 8 + 2 * 17 * 17 * 8 == len(raw)
    # the first face value of block 0: ws(ixMlo1-1, ixMlo2-1, 1)
    b0 = int(snap.block_offsets[0]) + 2 * 2 * 4 + 8 * 16 * 16 * 8
    ng = mesh.nghost
    assert struct.unpack_from("<d", raw, b0)[0] == be.get_state(0)[0, 8, ng - 1, ng - 1]
    be2 = OracleFV(par, mesh)
    be2.set_state(0, datfile.state_from_snapshot(datfile.read_snapshot(path), mesh))
    be2.getbc(0)
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=3)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))


def test_staggered_snapshot_restart_is_exact_in_3d(tmp_path):
    """Constrained transport in 3-D (the Gardiner & Stone field loop): three
    face components of 17^3 faces follow each block's cell data, and a run
    restarted from the snapshot continues bit for bit."""
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="ct", type_ct="uct_hll",
              time_stepper="threestep", flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean",
              courantpar=0.4)
    mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (-0.5, -0.5, -1.0), (0.5, 0.5, 1.0), ("periodic",) * 6)
    w0 = problems.field_loop_3d_ct(mesh)
    ref = OracleFV(par, mesh)
    ref.set_state(0, w0)
    ref.getbc(0)
    TimeIntegrator(par, mesh, ref).run(nsteps=4)
    be = OracleFV(par, mesh)
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=2)
    path = os.path.join(tmp_path, "fl3d0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it, global_time=ti.state.global_time)
    raw = open(path, "rb").read()
    per_block = 2 * 3 * 4 + 8 * 16 ** 3 * 8 + 3 * 17 ** 3 * 8
    assert snap.staggered and snap.ws.shape == (mesh.nblocks, 3, 17, 17, 17)
    assert np.all(np.diff(snap.block_offsets) == per_block)
    assert snap.block_offsets[-1] + per_block == len(raw)
    # the first face value of block 1: ws(ixMlo1-1, ixMlo2-1, ixMlo3-1, 1)
    b1 = int(snap.block_offsets[1]) + 2 * 3 * 4 + 8 * 16 ** 3 * 8
    ng = mesh.nghost
    assert struct.unpack_from("<d", raw, b1)[0] == be.get_state(0)[1, 8, ng - 1, ng - 1, ng - 1]
    be2 = OracleFV(par, mesh)
    be2.set_state(0, datfile.state_from_snapshot(datfile.read_snapshot(path), mesh))
    be2.getbc(0)
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=2)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))
