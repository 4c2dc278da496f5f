"""The reference's run-state files, read and written byte for byte:

  sav.mspms / old.mspms  checkpoint (saveit / loadit, src/loadsave.c:19-128):
      xx, yy, zz, vx, vy, vz [natom] FP64 (reduced units) | qq ps gg ss qqs pss
      ggs sss | vts rts vbs | boxlx boxly boxlz boxv (FP64) | bEquilibrium
      (C bool, 1 byte) | nframe, istep (int32) | counts[40] (int32) |
      accum[40] (FP64)
  trj.mspms  trajectory (trajectory, src/dumpxyz.c:36-66): on the first frame
      nspecie, nmole, natom (int32), per species nmole_per_specie, natom
      (int32) and per site its name (5 bytes) and aw (FP64); then per frame
      nmole_per_specie[nspecie] (int32) and xx, yy, zz [natom] (FP64)
  ss.mspms   snapshot (snapshot, src/dumpxyz.c:19-34): natom / "istep boxlx
      boxly boxlz" / "name  x  y  z" lines, printf %lf (reduced units)

Native byte order and the x86-64 C layout (int 4 bytes, bool 1 byte).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

NCOUNTS_MAX = 40  # src/defs.h:77
SAVE, LOAD, TRAJECTORY, SNAPSHOT = "sav.mspms", "old.mspms", "trj.mspms", "ss.mspms"


@dataclass
class Checkpoint:
    x: np.ndarray                     # (3, natom) reduced
    v: np.ndarray                     # (3, natom) reduced
    qq: float = 0.0
    ps: float = 0.0
    gg: float = 0.0
    ss: float = 0.0
    qqs: float = 0.0
    pss: float = 0.0
    ggs: float = 0.0
    sss: float = 0.0
    vts: float = 0.0
    rts: float = 0.0
    vbs: float = 0.0
    box: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    boxv: float = 0.0
    equilibrium: bool = False
    nframe: int = 0
    istep: int = 0
    counts: np.ndarray = field(default_factory=lambda: np.zeros(NCOUNTS_MAX, dtype=np.int32))
    accum: np.ndarray = field(default_factory=lambda: np.zeros(NCOUNTS_MAX))

    @property
    def natom(self) -> int:
        return int(self.x.shape[1])


_SCALARS = ("qq", "ps", "gg", "ss", "qqs", "pss", "ggs", "sss", "vts", "rts", "vbs")


def write_checkpoint(path: str, ck: Checkpoint) -> None:
    """saveit (src/loadsave.c:19-58)"""
    with open(path, "wb") as f:
        for a in (ck.x, ck.v):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(struct.pack("<11d", *(float(getattr(ck, k)) for k in _SCALARS)))
        f.write(struct.pack("<4d", *(float(b) for b in ck.box), float(ck.boxv)))
        f.write(struct.pack("<?ii", bool(ck.equilibrium), int(ck.nframe), int(ck.istep)))
        f.write(np.asarray(ck.counts, dtype="<i4").reshape(NCOUNTS_MAX).tobytes())
        f.write(np.asarray(ck.accum, dtype="<f8").reshape(NCOUNTS_MAX).tobytes())


def read_checkpoint(path: str, natom: int) -> Checkpoint:
    """loadit's reads (src/loadsave.c:60-128); the counters are read as
    written, whatever the starting option."""
    with open(path, "rb") as f:
        raw = f.read()
    need = 6 * 8 * natom + 15 * 8 + 9 + NCOUNTS_MAX * 12
    if len(raw) != need:
        raise ValueError(f"{path}: {len(raw)} bytes, a checkpoint of {natom} atoms has {need}")
    arr = np.frombuffer(raw[:48 * natom], dtype="<f8").reshape(6, natom)
    off = 48 * natom
    sc = struct.unpack_from("<11d", raw, off)
    off += 88
    bx = struct.unpack_from("<4d", raw, off)
    off += 32
    eq, nframe, istep = struct.unpack_from("<?ii", raw, off)
    off += 9
    counts = np.frombuffer(raw[off:off + 4 * NCOUNTS_MAX], dtype="<i4").copy()
    off += 4 * NCOUNTS_MAX
    accum = np.frombuffer(raw[off:off + 8 * NCOUNTS_MAX], dtype="<f8").copy()
    ck = Checkpoint(x=arr[0:3].copy(), v=arr[3:6].copy(), box=tuple(bx[:3]), boxv=bx[3],
                    equilibrium=bool(eq), nframe=nframe, istep=istep, counts=counts,
                    accum=accum)
    for k, v in zip(_SCALARS, sc):
        setattr(ck, k, v)
    return ck


@dataclass
class TrajectoryHeader:
    nmole_per_specie: List[int]
    natom_per_specie: List[int]
    names: List[List[str]]            # per species, per site
    aw: List[List[float]]             # per species, per site (reduced)

    @property
    def natom(self) -> int:
        return sum(n * a for n, a in zip(self.nmole_per_specie, self.natom_per_specie))


def _name5(s: str) -> bytes:
    b = s.encode()[:4]
    return b + b"\0" * (5 - len(b))


class TrajectoryWriter:
    """Appends frames to trj.mspms as trajectory() does (the header with the
    first frame)."""

    def __init__(self, path: str, header: TrajectoryHeader, nframe: int = 0):
        self.path, self.h, self.nframe = path, header, nframe
        if nframe == 0:
            open(path, "wb").close()

    def append(self, x: np.ndarray, nmole_per_specie: Sequence[int] = None) -> None:
        h = self.h
        with open(self.path, "ab") as f:
            if self.nframe == 0:
                ns = len(h.nmole_per_specie)
                f.write(struct.pack("<iii", ns, sum(h.nmole_per_specie), h.natom))
                for k in range(ns):
                    f.write(struct.pack("<ii", h.nmole_per_specie[k], h.natom_per_specie[k]))
                    for nm, w in zip(h.names[k], h.aw[k]):
                        f.write(_name5(nm))
                        f.write(struct.pack("<d", w))
            self.nframe += 1
            nm = h.nmole_per_specie if nmole_per_specie is None else nmole_per_specie
            f.write(np.asarray(nm, dtype="<i4").tobytes())
            f.write(np.ascontiguousarray(x, dtype="<f8").reshape(3, -1).tobytes())


def read_trajectory(path: str) -> Tuple[TrajectoryHeader, np.ndarray]:
    """Header and frames (nframe, 3, natom) of a trj.mspms."""
    with open(path, "rb") as f:
        raw = f.read()
    ns, _nmole, natom = struct.unpack_from("<iii", raw, 0)
    off = 12
    nmp, nap, names, aw = [], [], [], []
    for _ in range(ns):
        nm, na = struct.unpack_from("<ii", raw, off)
        off += 8
        nmp.append(nm)
        nap.append(na)
        nn, ww = [], []
        for _ in range(na):
            nn.append(raw[off:off + 5].split(b"\0")[0].decode())
            ww.append(struct.unpack_from("<d", raw, off + 5)[0])
            off += 13
        names.append(nn)
        aw.append(ww)
    frames = []
    step = 4 * ns + 24 * natom
    while off + step <= len(raw):
        off += 4 * ns
        frames.append(np.frombuffer(raw[off:off + 24 * natom], dtype="<f8").reshape(3, natom))
        off += 24 * natom
    return TrajectoryHeader(nmp, nap, names, aw), np.array(frames)


def write_snapshot(path: str, istep: int, box: Sequence[float], names: Sequence[str],
                   x: np.ndarray) -> None:
    """snapshot (src/dumpxyz.c:19-34), printf %lf formatting"""
    out = [f"{len(names)}", f"{istep} {box[0]:f} {box[1]:f} {box[2]:f}"]
    out += [f"{n}  {a:f}  {b:f}  {c:f}" for n, a, b, c in zip(names, x[0], x[1], x[2])]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_snapshot(path: str):
    """(istep, box, names, x (3, natom)) of an ss.mspms"""
    with open(path) as f:
        lines = f.read().split("\n")
    n = int(lines[0].split()[0])
    t = lines[1].split()
    rows = [ln.split() for ln in lines[2:2 + n]]
    x = np.array([[float(r[1]), float(r[2]), float(r[3])] for r in rows]).T
    return int(t[0]), tuple(float(v) for v in t[1:4]), [r[0] for r in rows], x


def trajectory_header(rundir, system) -> TrajectoryHeader:
    """The trj.mspms header of a run: site names from the run directory,
    masses in reduced units (the system's)."""
    return TrajectoryHeader(
        nmole_per_specie=[sp.nmole for sp in system.species],
        natom_per_specie=[sp.natom for sp in system.species],
        names=[[site.name for site in sp.sites] for sp in rundir.species],
        aw=[[float(w) for w in sp.aw] for sp in system.species])
