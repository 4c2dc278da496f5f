"""ctypes driver of the CPU oracle and of the reference build.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

  Oracle(system).frclong() / .frcshort()   the FP64 C restatement
  run_reference(rundir)                    the reference's own frclong /
                                           frcshort via oracle/_ref
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from typing import Optional

import numpy as np

from mspms2_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "mspms_ref_harness")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        sp, pp = C.POINTER(abi.Species), C.POINTER(abi.Params)
        dp = C.POINTER(C.c_double)
        _lib.oracle_frclong.argtypes = [C.c_int, sp, pp, dp, dp, dp, dp, dp, dp,
                                        C.POINTER(abi.LongTerms)]
        _lib.oracle_frcshort.argtypes = [C.c_int, sp, pp, dp, dp, dp, dp, dp, dp,
                                         C.POINTER(abi.ShortTerms)]
        _lib.oracle_derive_params.argtypes = [pp]
        _lib.oracle_check_uniques.argtypes = [sp, C.POINTER(C.c_uint8),
                                              C.POINTER(C.c_uint8)]
        _lib.oracle_hermite_site.argtypes = [C.POINTER(abi.SFGrid), C.c_int, C.c_double,
                                             C.c_double, C.c_double, C.c_double, dp, dp]
        _lib.oracle_grid_min_pe.argtypes = [C.POINTER(abi.SFGrid), C.c_int]
        _lib.oracle_grid_min_pe.restype = C.c_double
        _lib.oracle_rmarin.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.oracle_ranmar_next.argtypes = [C.c_void_p, dp, C.c_int]
    return _lib


class Oracle:
    """FP64 CPU restatement of frclong / frcshort for one System."""

    def __init__(self, system):
        self.s = system
        self.sp, self._keep_sp = system.c_species()
        self.p = system.c_params()
        rc = lib().oracle_derive_params(C.byref(self.p))
        if rc:
            raise RuntimeError(f"oracle_derive_params: {abi.ERROR_NAMES.get(rc, rc)}")

    def frclong(self, xx=None, yy=None, zz=None):
        s = self.s
        xx = s.xx if xx is None else np.ascontiguousarray(xx, dtype=np.float64)
        yy = s.yy if yy is None else np.ascontiguousarray(yy, dtype=np.float64)
        zz = s.zz if zz is None else np.ascontiguousarray(zz, dtype=np.float64)
        n = s.natom
        f = np.zeros((3, n))
        t = abi.LongTerms()
        rc = lib().oracle_frclong(len(s.species), self.sp, C.byref(self.p), abi.dptr(xx),
                                  abi.dptr(yy), abi.dptr(zz), abi.dptr(f[0]), abi.dptr(f[1]),
                                  abi.dptr(f[2]), C.byref(t))
        if rc:
            raise RuntimeError(f"oracle_frclong: {abi.ERROR_NAMES.get(rc, rc)}")
        return f, t.as_dict()

    def frcshort(self, xx=None, yy=None, zz=None):
        s = self.s
        xx = s.xx if xx is None else np.ascontiguousarray(xx, dtype=np.float64)
        yy = s.yy if yy is None else np.ascontiguousarray(yy, dtype=np.float64)
        zz = s.zz if zz is None else np.ascontiguousarray(zz, dtype=np.float64)
        n = s.natom
        f = np.zeros((3, n))
        t = abi.ShortTerms()
        rc = lib().oracle_frcshort(len(s.species), self.sp, C.byref(self.p), abi.dptr(xx),
                                   abi.dptr(yy), abi.dptr(zz), abi.dptr(f[0]),
                                   abi.dptr(f[1]), abi.dptr(f[2]), C.byref(t))
        if rc:
            raise RuntimeError(f"oracle_frcshort: {abi.ERROR_NAMES.get(rc, rc)}")
        return f, t.as_dict()

    def derived(self) -> dict:
        return {n: getattr(self.p, n) for n, _ in abi.Params._fields_
                if isinstance(getattr(self.p, n), (int, float))}


def check_uniques(species_c) -> tuple:
    au = np.zeros(max(1, species_c.nangle), dtype=np.uint8)
    du = np.zeros(max(1, species_c.ndih), dtype=np.uint8)
    lib().oracle_check_uniques(C.byref(species_c), abi.u8ptr(au), abi.u8ptr(du))
    return au[:species_c.nangle], du[:species_c.ndih]


class Ranmar:
    """RANMAR stream (src/random.c:14-109)."""

    def __init__(self, ij: int, kl: int):
        self._state = (C.c_byte * 1024)()
        lib().oracle_rmarin(self._state, ij, kl)

    def next(self, n: int) -> np.ndarray:
        out = np.empty(n)
        lib().oracle_ranmar_next(self._state, abi.dptr(out), n)
        return out


# --------------------------------------------------------------------------- #
REF_SCALARS = [
    "uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uvacuum", "uelec",
    "usflj", "uotherff", "uinter", "virial_inter",
    "ubond", "uangle", "udih", "uimp", "uintra", "virial_intra",
    "boxlx", "boxly", "boxlz", "rcutoff", "rcutoffsq", "rcutoffelec", "rcutoffelecsq",
    "rcuton", "rcutonsq", "roff2_minus_ron2_cube", "shift1", "shift4",
    "kappa", "coulomb_prefactor", "twopi_lx", "twopi_ly", "twopi_lz",
    "bfactor_ewald", "vfactor_ewald", "twopi_over_3v",
    "wolfvcon1", "wolfvcon2", "wolfvcon3", "wolffcon1", "wolffcon2",
    "f0", "sigma_base", "epsilon_base", "uljlrc", "pljlrc", "time_base", "delt",
    "delts", "const_3pisq_theta"]


# the md[] block of a version-2 dump (oracle/ref_harness.c, --md-steps)
MD_SCALARS = [
    "nmd", "what_ensemble", "nfree", "nstep_inner", "treq", "preq", "delt", "delts", "qq",
    "qqs", "gg", "ggs", "ss", "ps", "sss", "pss", "unhts", "unhtss", "ukin", "tinst", "Qts",
    "Qbs", "vts", "vbs", "rts", "utsbs", "boxv", "pinst"]


def read_ref_dump(path: str) -> dict:
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != b"MSPMSREF":
        raise ValueError("not a reference dump")
    ver, natom, nsc = np.frombuffer(raw[8:20], dtype=np.int32)
    off = 20
    sc = np.frombuffer(raw[off:off + 8 * nsc], dtype=np.float64)
    off += 8 * nsc
    arr = np.frombuffer(raw[off:off + 8 * 9 * natom], dtype=np.float64).reshape(9, natom)
    out = {n: float(v) for n, v in zip(REF_SCALARS, sc)}
    out.update(xx=arr[0].copy(), yy=arr[1].copy(), zz=arr[2].copy(),
               fl=arr[3:6].copy(), fs=arr[6:9].copy(), natom=int(natom))
    if ver >= 2:
        off += 8 * 9 * natom
        (nmd,) = np.frombuffer(raw[off:off + 4], dtype=np.int32)
        off += 4
        md = np.frombuffer(raw[off:off + 8 * nmd], dtype=np.float64)
        off += 8 * nmd
        out["md"] = {n: float(v) for n, v in zip(MD_SCALARS, md)}
        v = np.frombuffer(raw[off:off + 8 * 6 * natom], dtype=np.float64).reshape(6, natom)
        out.update(v0=v[0:3].copy(), v=v[3:6].copy())
    return out


def reference_available() -> bool:
    return os.path.exists(REF_HARNESS)


def run_reference(rundir: str, repeat: int = 0, timeout: Optional[float] = None,
                  md_steps: int = 0, files_out: Optional[str] = None,
                  siman: bool = False, hmc: int = 0) -> dict:
    """Runs the reference's start-up path + frclong + frcshort on a run
    directory (in a scratch copy) and returns its dump; md_steps > 0 then
    runs that many steps of the input's integrator (vver / vver_nh_3 /
    npt_respa) and adds the velocities and the thermostat state.  With
    files_out, the reference's own trajectory frame, snapshot and checkpoint
    of the final state (trj.mspms, ss.mspms, sav.mspms) are copied there;
    an old.mspms in the run directory is passed along (restart inputs), and
    so are the cube files a MYINTERP section names.
    siman runs the reference's simulated annealing driver instead; hmc > 0
    runs that many moves of its HMC driver and adds the per-move record
    (d["hmc"]: accept, dH, delt, uinter, uintra, ukin per move)."""
    if not reference_available():
        raise FileNotFoundError(REF_HARNESS)
    with tempfile.TemporaryDirectory(prefix="mspms_ref_") as tmp:
        extra = []
        try:
            from mspms2_amd.runio import read_in
            extra = list(read_in(os.path.join(rundir, "in.mspms")).mi_files)
        except (OSError, ValueError, IndexError):
            pass
        for fn in ["in.mspms", "cfg.mspms", "xyz.mspms", "old.mspms"] + extra:
            if fn not in ("in.mspms", "cfg.mspms", "xyz.mspms") and \
                    not os.path.exists(os.path.join(rundir, fn)):
                continue
            with open(os.path.join(rundir, fn), "rb") as src, \
                    open(os.path.join(tmp, fn), "wb") as dst:
                dst.write(src.read())
        out = os.path.join(tmp, "dump.bin")
        cmd = [REF_HARNESS, tmp, out] + (["--repeat", str(repeat)] if repeat else []) + \
            (["--md-steps", str(md_steps)] if md_steps else []) + \
            (["--write-files"] if files_out else []) + (["--siman"] if siman else []) + \
            (["--hmc", str(hmc)] if hmc else [])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed ({r.returncode}): {r.stderr[-2000:]}")
        d = read_ref_dump(out)
        for ln in r.stdout.splitlines():
            if ln.startswith("frclong_seconds_per_call"):
                d["frclong_seconds_per_call"] = float(ln.split()[1])
            if ln.startswith("md_seconds_per_step"):
                d["md_seconds_per_step"] = float(ln.split()[1])
            if ln.startswith("hmc_move"):
                t = ln.split()
                d.setdefault("hmc", []).append(
                    dict(accept=int(t[2]), dH=float(t[3]), delt=float(t[4]),
                         uinter=float(t[5]), uintra=float(t[6]), ukin=float(t[7])))
        if files_out:
            os.makedirs(files_out, exist_ok=True)
            for fn in ("trj.mspms", "ss.mspms", "sav.mspms"):
                with open(os.path.join(tmp, fn), "rb") as src, \
                        open(os.path.join(files_out, fn), "wb") as dst:
                    dst.write(src.read())
        return d
