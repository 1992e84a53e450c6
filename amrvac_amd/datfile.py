"""MPI-AMRVAC snapshot files (``.dat``, format version 5) from and to the
state of a finite-volume backend.

The byte layout is the reference's (doc/fileformat.md):
  * the header of snapshot_write_header1 (src/io/mod_input_output_helper.t:
    128-201) with the physics block of hd_write_info / mhd_write_info
    (src/hd/mod_hd_phys.t:141-156, src/mhd/mod_mhd_phys.t:278-294);
  * the tree of write_snapshot (src/io/mod_input_output.t:2004-2172): the
    leaf flags of a depth-first walk from every level-1 block in Morton
    order, children with ic(1) fastest (write_forest,
    src/amr/mod_functions_forest.t:300-333), then the level and spatial
    index of every leaf and the byte offset of its block;
  * per leaf in Morton order: the ghost-cell counts (0, as with the default
    save_physical_boundary = .false., block_shape_io
    src/io/mod_input_output_helper.t:54-84) and w(ixM^S, 1:nw) of the
    conserved state, Fortran order; with a staggered face field
    (constrained transport) followed by ws(ixM^L-1:ixM^L, 1:nws)
    (src/io/mod_input_output.t:2100-2106), the header's staggered flag set.
Integers are 4-byte, logicals 4-byte (0/1), offsets 8-byte, reals 8-byte,
little-endian; names are blank-padded to 16 characters.

The state comes from (and goes back to) a backend slot through get_state /
set_state, so a device run writes its snapshot straight from HBM and
restarts from one.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

VERSION = 5          # version_number, src/io/mod_input_output_helper.t:15
NAME_LEN = 16        # name_len, src/mod_basic_types.t:22


def variable_names(par) -> list[str]:
    """cons_wnames of hd_phys_init / mhd_phys_init (src/mod_variables.t)."""
    names = ["rho"] + [f"m{d + 1}" for d in range(par.ndir)] + ["e"]
    if par.physics == "mhd":
        names += [f"b{d + 1}" for d in range(par.ndir)]
        if par.typedivbfix == "glm":
            names.append("psi")
    return names


def geometry_name(par) -> str:
    """The Cartesian names of set_coordinate_system (src/mod_geometry.t)."""
    if par.ndir > par.ndim:
        return f"Cartesian_{par.ndim}.5D" if par.ndir == par.ndim + 1 else f"Cartesian_{par.ndim}.75D"
    return f"Cartesian_{par.ndim}D"


def _name(s: str) -> bytes:
    b = s.encode("ascii")[:NAME_LEN]
    return b + b" " * (NAME_LEN - len(b))


@dataclass
class Snapshot:
    """Contents of a snapshot file."""

    nw: int
    ndir: int
    ndim: int
    levmax: int
    it: int
    global_time: float
    xprobmin: tuple
    xprobmax: tuple
    domain_nx: tuple
    block_nx: tuple
    periodic: tuple
    geometry: str
    staggered: bool
    w_names: list
    physics_type: str
    params: dict
    snapshotnext: int
    slicenext: int
    collapsenext: int
    leaf: np.ndarray = field(repr=False)        # depth-first leaf flags
    level: np.ndarray = field(repr=False)       # [nleafs]
    ig: np.ndarray = field(repr=False)          # [nleafs, ndim], 1-based
    w: np.ndarray = field(repr=False)           # [nleafs, nw, (k,) j, i] interior cells
    ws: np.ndarray = field(default=None, repr=False)  # [nleafs, nws, (k,) j, i] faces ixM-1..ixM (staggered)
    offset_tree: int = 0
    offset_block: int = 0
    block_offsets: np.ndarray = field(default=None, repr=False)

    @property
    def nleafs(self) -> int:
        return int(self.level.shape[0])

    @property
    def nparents(self) -> int:
        return int(self.leaf.shape[0]) - self.nleafs


def _tree(mesh):
    """Depth-first leaf flags, levels and spatial indices of the leaves."""
    if hasattr(mesh, "roots"):  # AMR forest
        flags: list[bool] = []

        def walk(n) -> None:
            flags.append(n.leaf)
            if not n.leaf:
                for c in n.children:  # ic(1) fastest
                    walk(c)

        for r in mesh.roots:
            walk(r)
        level = np.array([n.level for n in mesh.leaves], dtype=np.int32)
        ig = np.array([n.ig for n in mesh.leaves], dtype=np.int32).reshape(-1, mesh.ndim)
        return np.array(flags, dtype=bool), level, ig
    nb = mesh.nblocks
    return np.ones(nb, dtype=bool), np.ones(nb, dtype=np.int32), np.asarray(mesh.ig, dtype=np.int32)


def write_snapshot(path: str, par, mesh, w: np.ndarray, it: int = 0, global_time: float = 0.0,
                   snapshotnext: int = 0, slicenext: int = 0, collapsenext: int = 0) -> Snapshot:
    """Write a version-5 snapshot.  w: [nb, nw, (k,) j, i] in the backend's
    host layout with ghost cells (get_state), conserved variables; only the
    interior cells are written.  snapshotnext is written +1 as at a regular
    output (the reference increments it after writing)."""
    nd = par.ndim
    leaf, level, ig = _tree(mesh)
    nleafs = int(level.shape[0])
    nparents = int(leaf.shape[0]) - nleafs
    w = np.asarray(w, dtype=np.float64)
    nws = getattr(par, "nws", 0)
    if w.shape[0] != nleafs or w.shape[1] != par.nw + nws:
        raise ValueError(f"state shape {w.shape} does not match {nleafs} leaves x {par.nw + nws} variables")
    inner = np.ascontiguousarray(mesh.interior(w[:, :par.nw]))
    faces = None
    if nws:
        # faces ixM-1..ixM in every dimension: the upper faces of the cells
        # ng-1 .. ng+block_nx-1 (ghost faces as getbc left them)
        ng = mesh.nghost
        sl = tuple(slice(ng - 1, ng + mesh.block_nx[d]) for d in reversed(range(nd)))
        faces = np.ascontiguousarray(w[(slice(None), slice(par.nw, par.nw + nws)) + sl])
    names = variable_names(par)
    periodic = tuple(mesh.bc[2 * d] == "periodic" for d in range(nd))

    def header(off_tree: int, off_block: int) -> bytes:
        h = struct.pack("<10i", VERSION, off_tree, off_block, par.nw, par.ndir, nd, int(level.max()), nleafs,
                        nparents, it)
        h += struct.pack("<d", global_time)
        h += struct.pack(f"<{nd}d", *mesh.xprobmin) + struct.pack(f"<{nd}d", *mesh.xprobmax)
        h += struct.pack(f"<{nd}i", *mesh.domain_nx) + struct.pack(f"<{nd}i", *mesh.block_nx)
        h += struct.pack(f"<{nd}i", *[1 if p else 0 for p in periodic])
        h += _name(geometry_name(par))
        h += struct.pack("<i", 1 if nws else 0)  # staggered
        h += b"".join(_name(n) for n in names)
        h += _name(par.physics)
        h += struct.pack("<i", 1) + struct.pack("<d", par.gamma) + _name("gamma")
        h += struct.pack("<3i", snapshotnext + 1, slicenext, collapsenext)
        return h

    off_tree = len(header(0, 0))
    tree = leaf.astype("<i4").tobytes() + level.astype("<i4").tobytes() + ig.astype("<i4").tobytes()
    off_block = off_tree + len(tree) + 8 * nleafs
    per_block = 2 * nd * 4 + inner[0].size * 8 + (faces[0].size * 8 if nws else 0)
    offsets = off_block + per_block * np.arange(nleafs, dtype=np.int64)
    with open(path, "wb") as fh:
        fh.write(header(off_tree, off_block))
        fh.write(tree)
        fh.write(offsets.astype("<i8").tobytes())
        zeros = struct.pack(f"<{2 * nd}i", *([0] * (2 * nd)))
        for b in range(nleafs):
            fh.write(zeros)
            fh.write(inner[b].astype("<f8").tobytes())  # (iw, (k,) j, i) C order = Fortran w(i, j, (k,) iw)
            if nws:
                fh.write(faces[b].astype("<f8").tobytes())
    return read_snapshot(path)


def read_snapshot(path: str) -> Snapshot:
    """Read a version-5 snapshot (snapshot_read_header
    src/io/mod_input_output.t:1833-2002, read_snapshot :2176-2379)."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def take(fmt: str):
        nonlocal pos
        vals = struct.unpack_from("<" + fmt, data, pos)
        pos += struct.calcsize("<" + fmt)
        return vals

    def name() -> str:
        nonlocal pos
        s = data[pos:pos + NAME_LEN].decode("ascii").rstrip()
        pos += NAME_LEN
        return s

    version, off_tree, off_block, nw, ndir, nd, levmax, nleafs, nparents, it = take("10i")
    if version != VERSION:
        raise ValueError(f"snapshot version {version}, expected {VERSION}")
    (global_time,) = take("d")
    xmin, xmax = take(f"{nd}d"), take(f"{nd}d")
    dnx, bnx = take(f"{nd}i"), take(f"{nd}i")
    periodic = tuple(bool(v) for v in take(f"{nd}i"))
    geom = name()
    (stag,) = take("i")
    w_names = [name() for _ in range(nw)]
    phys = name()
    (npar,) = take("i")
    pvals = take(f"{npar}d")
    pnames = [name() for _ in range(npar)]
    snext, slnext, cnext = take("3i")
    if pos != off_tree:
        raise ValueError(f"tree offset {off_tree} does not follow the header ({pos})")
    leaf = np.frombuffer(data, "<i4", nleafs + nparents, off_tree).astype(bool)
    p = off_tree + 4 * (nleafs + nparents)
    level = np.frombuffer(data, "<i4", nleafs, p).astype(np.int32)
    p += 4 * nleafs
    ig = np.frombuffer(data, "<i4", nd * nleafs, p).astype(np.int32).reshape(nleafs, nd)
    p += 4 * nd * nleafs
    offsets = np.frombuffer(data, "<i8", nleafs, p).astype(np.int64)
    p += 8 * nleafs
    if p != off_block:
        raise ValueError(f"block offset {off_block} does not follow the tree ({p})")
    blocks, fblocks = [], []
    nws = nd if stag else 0  # nws = ndim (src/mhd/mod_mhd_phys.t:558)
    for b in range(nleafs):
        q = int(offsets[b])
        ng = np.frombuffer(data, "<i4", 2 * nd, q)
        shape = [bnx[d] + int(ng[d]) + int(ng[nd + d]) for d in range(nd)]
        n = nw * int(np.prod(shape))
        arr = np.frombuffer(data, "<f8", n, q + 8 * nd).reshape([nw] + shape[::-1])
        blocks.append(np.array(arr))
        if nws:
            fshape = [s_ + 1 for s_ in shape]
            nf = nws * int(np.prod(fshape))
            farr = np.frombuffer(data, "<f8", nf, q + 8 * nd + 8 * n).reshape([nws] + fshape[::-1])
            fblocks.append(np.array(farr))
    return Snapshot(nw=nw, ndir=ndir, ndim=nd, levmax=levmax, it=it, global_time=global_time,
                    xprobmin=tuple(xmin), xprobmax=tuple(xmax), domain_nx=tuple(dnx), block_nx=tuple(bnx),
                    periodic=periodic, geometry=geom, staggered=bool(stag), w_names=w_names,
                    physics_type=phys, params=dict(zip(pnames, pvals)), snapshotnext=snext, slicenext=slnext,
                    collapsenext=cnext, leaf=leaf, level=level, ig=ig, w=np.stack(blocks),
                    ws=np.stack(fblocks) if nws else None,
                    offset_tree=off_tree, offset_block=off_block, block_offsets=offsets)


def forest_from_snapshot(snap: Snapshot, bc, max_level: int, nghost: int = 2):
    """The AMR forest a snapshot describes (read_forest: the depth-first
    leaf flags refine the level-1 blocks in Morton order)."""
    from .amr import AMRForest
    f = AMRForest(snap.ndim, snap.domain_nx, snap.block_nx, snap.xprobmin, snap.xprobmax, tuple(bc),
                  max_level=max_level, nghost=nghost)
    flags = iter(snap.leaf)

    def build(n) -> None:
        if not next(flags):
            for c in f.refine_node(n):
                build(c)

    for r in list(f.roots):
        build(r)
    f.update()
    if [n.level for n in f.leaves] != list(snap.level) or \
            [tuple(n.ig) for n in f.leaves] != [tuple(v) for v in snap.ig]:
        raise ValueError("snapshot leaves do not match the tree of its leaf flags")
    return f


def state_from_snapshot(snap: Snapshot, mesh) -> np.ndarray:
    """The snapshot's interior data in the backend's host layout with ghost
    cells ([nb, nw (+ nws), (k,) j, i]; ghost cells zero, to be filled by
    getbc; a staggered snapshot's faces ixM-1..ixM after the cell
    variables)."""
    nws = 0 if snap.ws is None else snap.ws.shape[1]
    shape = (snap.nleafs, snap.nw + nws) + tuple(mesh.nxg[::-1])
    w = np.zeros(shape, dtype=np.float64)
    mesh.interior(w[:, :snap.nw])[...] = snap.w
    if nws:
        ng = mesh.nghost
        sl = tuple(slice(ng - 1, ng + mesh.block_nx[d]) for d in reversed(range(snap.ndim)))
        w[(slice(None), slice(snap.nw, snap.nw + nws)) + sl] = snap.ws
    return w
