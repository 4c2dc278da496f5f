"""Solid-fluid grid generation on the GPU: the reference's offline
generator tools/sfgridgen.c (energy, gradient and force Jacobian of the LJ
potential of a z-periodic solid over a grid of one unit cell), with its
input and output file formats.

Input (tools/sfgridgen.c:104-117), free-format text:
    fluid_name fluid_sigma(A) fluid_epsilon(K)
    grid_interval(A) rcuton(A) rcutoff(A) isLJswitch
    s_natom uclx(A) ucly(A) uclz(A)
    solid_sigma(A) solid_epsilon(K)
    name x y z                     (s_natom lines, Angstrom)
Output out.sfgridgen (tools/sfgridgen.c:318-352), native binary:
    uclx ucly uclz xcenter ycenter zcenter (FP64) | ngrid_total ngrid_x
    ngrid_y ngrid_z (int32) | grid_interval x3 (FP64) | 13 tables of
    ngrid_total FP64: energy (J/mol), dE/dx, dE/dy, dE/dz, then the Jacobian
    of the force F = -grad E, dFx/dx dFx/dy dFx/dz dFy/dx .. dFz/dz (minus the
    Hessian of E); node = (i ngrid_y + j) ngrid_z + k
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass
from typing import List

import numpy as np

from . import abi
from .engine import _check, lib

FIELDS = ("E", "dEx", "dEy", "dEz", "dFxx", "dFxy", "dFxz", "dFyx", "dFyy", "dFyz", "dFzx",
          "dFzy", "dFzz")


@dataclass
class GridInput:
    f_name: str
    f_sigma: float
    f_epsilon: float
    grid_itvl: float
    rcuton: float
    rcutoff: float
    isLJswitch: int
    uclx: float
    ucly: float
    uclz: float
    s_sigma: float
    s_epsilon: float
    s_names: List[str]
    s_xyz: np.ndarray                   # (3, s_natom) Angstrom

    @property
    def s_natom(self) -> int:
        return int(self.s_xyz.shape[1])


def read_input(path: str) -> GridInput:
    with open(path) as f:
        tok = f.read().split()
    it = iter(tok)
    nxt = lambda: next(it)  # noqa: E731
    fn, fs, fe = nxt(), float(nxt()), float(nxt())
    gi, ron, roff, sw = float(nxt()), float(nxt()), float(nxt()), int(nxt())
    ns, ux, uy, uz = int(nxt()), float(nxt()), float(nxt()), float(nxt())
    ss, se = float(nxt()), float(nxt())
    names, xyz = [], np.zeros((3, ns))
    for k in range(ns):
        names.append(nxt())
        xyz[:, k] = (float(nxt()), float(nxt()), float(nxt()))
    return GridInput(fn, fs, fe, gi, ron, roff, sw, ux, uy, uz, ss, se, names, xyz)


def _spec(g: GridInput):
    xyz = np.ascontiguousarray(g.s_xyz, dtype=np.float64)
    s = abi.SfGridGenSpec(g.f_sigma, g.f_epsilon, g.grid_itvl, g.rcuton, g.rcutoff, int(g.isLJswitch),
              g.s_natom, g.uclx, g.ucly, g.uclz, g.s_sigma, g.s_epsilon,
              abi.dptr(xyz[0]), abi.dptr(xyz[1]), abi.dptr(xyz[2]))
    return s, xyz


def dims(g: GridInput):
    """(nx, ny, nz) = floor(ucl / interval) + 2 (tools/sfgridgen.c:129-131)"""
    s, keep = _spec(g)
    n = [C.c_int(), C.c_int(), C.c_int()]
    L = lib()
    L.mspms_sfgridgen_dims.argtypes = [C.POINTER(abi.SfGridGenSpec)] + [C.POINTER(C.c_int)] * 3
    _check(L.mspms_sfgridgen_dims(C.byref(s), *(C.byref(v) for v in n)), "mspms_sfgridgen_dims")
    return tuple(v.value for v in n)


def generate(g: GridInput, device: int = 0) -> np.ndarray:
    """The 13 tables, shape (13, nx, ny, nz), computed on the GPU."""
    nx, ny, nz = dims(g)
    out = np.empty((13, nx, ny, nz))
    s, keep = _spec(g)
    L = lib()
    L.mspms_sfgridgen.argtypes = [C.c_int, C.POINTER(abi.SfGridGenSpec), C.POINTER(C.c_double)]
    _check(L.mspms_sfgridgen(int(device), C.byref(s), abi.dptr(out)), "mspms_sfgridgen")
    return out


def write_output(path: str, g: GridInput, tables: np.ndarray) -> None:
    """out.sfgridgen as the generator writes it"""
    _, nx, ny, nz = tables.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<6d", g.uclx, g.ucly, g.uclz, g.uclx / 2.0, g.ucly / 2.0, 0.0))
        f.write(struct.pack("<4i", nx * ny * nz, nx, ny, nz))
        f.write(struct.pack("<3d", g.grid_itvl, g.grid_itvl, g.grid_itvl))
        f.write(np.ascontiguousarray(tables, dtype="<f8").tobytes())


def read_output(path: str):
    """(header dict, tables (13, nx, ny, nz)) of an out.sfgridgen"""
    with open(path, "rb") as f:
        raw = f.read()
    ucl_c = struct.unpack_from("<6d", raw, 0)
    ntot, nx, ny, nz = struct.unpack_from("<4i", raw, 48)
    h = struct.unpack_from("<3d", raw, 64)
    tab = np.frombuffer(raw[88:88 + 13 * 8 * ntot], dtype="<f8").reshape(13, nx, ny, nz)
    return dict(ucl=ucl_c[:3], center=ucl_c[3:], ntot=ntot, n=(nx, ny, nz), h=h), tab
