"""The "myinterp" solid-fluid model: 32-coefficient cube polynomials.

The reference reads them (init_my_interp, src/sffrc.c:228-748) and
evaluates them per site (get_values_from_grid, src/sffrc.c:951-1089) but
ships no program that writes them; `generate` does, on the GPU.

Cube file (native binary, src/sffrc.c:296-737):
    uclx ucly uclz xcenter ycenter zcenter (FP64) | ngrid_total ngrid_x
    ngrid_y ngrid_z (int32) | grid_itvl_x _y _z (FP64) | 128 tables of
    ncube FP64, ncube = (ngrid_x-1)(ngrid_y-1)(ngrid_z-1), in the order
    ene0..ene31, fxa0..fxa31, fya0..fya31, fza0..fza31; cube =
    (ix ncube_y + iy) ncube_z + iz.  Coefficient m of each polynomial
    multiplies monomial m of Amatrix_ele_assign_value (src/sffrc.c:771-802)
    in absolute Angstrom coordinates; energy in J/mol, force in J/mol/A.

Run-input section (src/sffrc.c:262-297), read when iSF_type = 4:
    MYINTERP
    uclx ucly uclz
    xcenter ycenter zcenter
    nunique_atom
    one file name per line (table type t = line t, used by interp_type t+1)
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import abi
from .engine import _check, lib
from .sfgridgen import GridInput, dims

NQ = 7   # node data per polynomial: value, d/dx, d/dy, d/dz, d2/dxdy, d2/dxdz, d2/dydz


@dataclass
class CubeFile:
    ucl: Tuple[float, float, float]
    center: Tuple[float, float, float]
    ngrid: Tuple[int, int, int]
    h: Tuple[float, float, float]
    coef: np.ndarray                  # (128, ncube)

    @property
    def ncube(self) -> Tuple[int, int, int]:
        return tuple(n - 1 for n in self.ngrid)


def write_cubes(path: str, cf: CubeFile) -> None:
    nx, ny, nz = cf.ngrid
    ncube = (nx - 1) * (ny - 1) * (nz - 1)
    if cf.coef.shape != (128, ncube):
        raise ValueError(f"coefficients {cf.coef.shape}, expected (128, {ncube})")
    with open(path, "wb") as f:
        f.write(struct.pack("<6d", *cf.ucl, *cf.center))
        f.write(struct.pack("<4i", nx * ny * nz, nx, ny, nz))
        f.write(struct.pack("<3d", *cf.h))
        f.write(np.ascontiguousarray(cf.coef, dtype="<f8").tobytes())


def read_cubes(path: str) -> CubeFile:
    with open(path, "rb") as f:
        raw = f.read()
    uc = struct.unpack_from("<6d", raw, 0)
    _ntot, nx, ny, nz = struct.unpack_from("<4i", raw, 48)
    h = struct.unpack_from("<3d", raw, 64)
    ncube = (nx - 1) * (ny - 1) * (nz - 1)
    need = 88 + 128 * 8 * ncube
    if len(raw) < need:
        raise ValueError(f"{path}: {len(raw)} bytes, {need} expected for {ncube} cubes")
    coef = np.frombuffer(raw[88:need], dtype="<f8").reshape(128, ncube).copy()
    return CubeFile(ucl=uc[:3], center=uc[3:], ngrid=(nx, ny, nz), h=h, coef=coef)


def node_data(g: GridInput, device: int = 0) -> np.ndarray:
    """(4, 7, nx, ny, nz): the LJ sum and its derivatives at the nodes of
    the grid generator for polynomials (E, Fx, Fy, Fz), on the GPU."""
    from .sfgridgen import _spec
    nx, ny, nz = dims(g)
    out = np.empty((4, NQ, nx, ny, nz))
    s, keep = _spec(g)
    L = lib()
    L.mspms_myinterp_nodes.argtypes = [C.c_int, C.POINTER(abi.SfGridGenSpec),
                                       C.POINTER(C.c_double)]
    _check(L.mspms_myinterp_nodes(int(device), C.byref(s), abi.dptr(out)),
           "mspms_myinterp_nodes")
    return out


def fit(nodes: np.ndarray, h: Sequence[float], origin: Sequence[float] = (0.0, 0.0, 0.0),
        device: int = 0) -> np.ndarray:
    """(128, ncube) coefficients from node data (4, 7, nx, ny, nz), on the GPU."""
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    _, nq, nx, ny, nz = nodes.shape
    if nq != NQ or nodes.shape[0] != 4:
        raise ValueError(f"node data {nodes.shape}, expected (4, 7, nx, ny, nz)")
    coef = np.empty((128, (nx - 1) * (ny - 1) * (nz - 1)))
    L = lib()
    L.mspms_myinterp_fit.argtypes = [C.c_int] * 4 + [C.c_double] * 6 + [C.POINTER(C.c_double)] * 2
    _check(L.mspms_myinterp_fit(int(device), nx, ny, nz, *(float(v) for v in origin),
                                *(float(v) for v in h), abi.dptr(nodes), abi.dptr(coef)),
           "mspms_myinterp_fit")
    return coef


def generate(g: GridInput, device: int = 0) -> CubeFile:
    """The cube file for one fluid site type over the solid of `g` (the grid
    generator's nodes: origin 0, the cell centred at (uclx/2, ucly/2))."""
    h = (g.grid_itvl,) * 3
    coef = fit(node_data(g, device), h, (0.0, 0.0, 0.0), device)
    return CubeFile(ucl=(g.uclx, g.ucly, g.uclz), center=(g.uclx / 2.0, g.ucly / 2.0, 0.0),
                    ngrid=dims(g), h=h, coef=coef)


def engine_tables(ucl: Sequence[float], center: Sequence[float],
                  files: List[CubeFile]) -> Dict:
    """The mspms_sf_cubes fields for a run: the cell of the run input
    (init_my_interp: xmin = xcenter - uclx/2, ymin likewise, zmax = uclz),
    the files' lattice (all files share one).  A file whose cell differs
    from the input's is refused, as init_my_interp refuses it."""
    if not 1 <= len(files) <= 5:
        raise ValueError("1..5 cube files (NUNIQUE_ATOM_MAX)")
    for k, f in enumerate(files):
        if tuple(f.ucl) != tuple(ucl):
            raise ValueError(f"unit cell mismatch in cube file {k}: {f.ucl} vs {tuple(ucl)}")
        if f.ngrid != files[0].ngrid or f.h != files[0].h:
            raise ValueError("cube files must share one lattice")
    ncx, ncy, ncz = files[0].ncube
    return dict(ntype=len(files), ncx=ncx, ncy=ncy, ncz=ncz,
                xmin=center[0] - ucl[0] / 2.0, ymin=center[1] - ucl[1] / 2.0, zmax=ucl[2],
                hx=files[0].h[0], hy=files[0].h[1], hz=files[0].h[2],
                coef=[np.ascontiguousarray(f.coef) for f in files])


def c_cubes(d: Dict):
    """(mspms_sf_cubes, keep-alive list) from engine_tables' dict."""
    q = abi.SFCubes()
    for k in ("ntype", "ncx", "ncy", "ncz", "xmin", "ymin", "zmax", "hx", "hy", "hz"):
        setattr(q, k, d[k])
    ptrs = (C.POINTER(C.c_double) * (128 * d["ntype"]))()
    for t, coef in enumerate(d["coef"]):
        for m in range(128):
            ptrs[t * 128 + m] = abi.dptr(coef[m])
    q.tables = ptrs
    return q, [ptrs, d["coef"]]
