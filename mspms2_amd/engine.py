"""Python binding of libmspms_gpu.so (the C ABI of include/mspms_gpu.h).

`Engine(system)` evaluates frclong / frcshort of an mspms2 system on the
GPU.  There is no CPU fallback: if the HIP library is not built, or no GPU
is visible, construction fails with an error naming the cause.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmspms_gpu.so")

_lib = None

EXPORTS = [
    "mspms_derive_params", "mspms_check_uniques", "mspms_last_error", "mspms_create",
    "mspms_destroy", "mspms_set_system", "mspms_set_params", "mspms_set_replicas",
    "mspms_natom", "mspms_nreplica", "mspms_frclong_host", "mspms_frcshort_host",
    "mspms_frclong_device", "mspms_frcshort_device", "mspms_energy_param_grad",
    "mspms_sync", "mspms_stream", "mspms_set_profiling", "mspms_last_timings",
    "mspms_vver_step_device",
]


class EngineError(RuntimeError):
    pass


def lib():
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError(f"{LIB_PATH} is missing: build it with `python -m mspms2_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.mspms_last_error.restype = C.c_char_p
    L.mspms_derive_params.argtypes = [C.POINTER(abi.Params)]
    L.mspms_check_uniques.argtypes = [C.POINTER(abi.Species), C.POINTER(C.c_uint8),
                                      C.POINTER(C.c_uint8)]
    L.mspms_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
    L.mspms_destroy.argtypes = [vp]
    L.mspms_set_system.argtypes = [vp, C.c_int, C.POINTER(abi.Species), C.POINTER(abi.Params)]
    L.mspms_set_params.argtypes = [vp, C.POINTER(abi.Params)]
    L.mspms_set_replicas.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(abi.Params)]
    L.mspms_natom.argtypes = [vp]
    L.mspms_nreplica.argtypes = [vp]
    L.mspms_frclong_host.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.POINTER(abi.LongTerms)]
    L.mspms_frcshort_host.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.POINTER(abi.ShortTerms)]
    L.mspms_frclong_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.mspms_frcshort_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.mspms_vver_step_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, vp, vp, vp,
                                         vp]
    L.mspms_sync.argtypes = [vp]
    L.mspms_stream.argtypes = [vp]
    L.mspms_stream.restype = vp
    L.mspms_set_profiling.argtypes = [vp, C.c_int]
    L.mspms_last_timings.argtypes = [vp, dp, C.c_int]
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().mspms_last_error().decode(errors="replace")
        raise EngineError(f"{what}: {abi.ERROR_NAMES.get(rc, rc)}: {msg}")


def derive_params(p: abi.Params) -> abi.Params:
    _check(lib().mspms_derive_params(C.byref(p)), "mspms_derive_params")
    return p


PROFILE_GROUPS = ("cluster", "pair", "intra", "recip_s", "recip_f", "finalize", "bonded", "sf")


class Engine:
    """GPU force engine for one mspms2 system (or a batch of replicas)."""

    def __init__(self, system, precision: str = "mixed", device: int = 0):
        self.system = system
        self.precision = {"fp64": abi.PREC_FP64, "mixed": abi.PREC_MIXED}[precision]
        L = lib()
        h = C.c_void_p()
        _check(L.mspms_create(C.byref(h), device, self.precision), "mspms_create")
        self._h = h
        self._sp, self._keep = system.c_species()
        self.params = system.c_params()
        derive_params(self.params)
        _check(L.mspms_set_system(self._h, len(system.species), self._sp, C.byref(self.params)),
               "mspms_set_system")
        self.nrep = 1

    def close(self):
        if getattr(self, "_h", None):
            lib().mspms_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def natom(self) -> int:
        return lib().mspms_natom(self._h)

    def set_replicas(self, nmole: Sequence[Sequence[int]], params: Sequence[abi.Params]):
        n = len(params)
        arr = (abi.Params * n)(*params)
        for k in range(n):
            derive_params(arr[k])
        nm = np.ascontiguousarray(np.asarray(nmole, dtype=np.int32).reshape(n, -1))
        self._rep_keep = (arr, nm, params)
        _check(lib().mspms_set_replicas(self._h, n, nm.ctypes.data_as(C.POINTER(C.c_int)), arr),
               "mspms_set_replicas")
        self.nrep = n

    # ---- host-buffer evaluation --------------------------------------------
    def frclong(self, xx=None, yy=None, zz=None):
        s = self.system
        xx = np.ascontiguousarray(s.xx if xx is None else xx, dtype=np.float64)
        yy = np.ascontiguousarray(s.yy if yy is None else yy, dtype=np.float64)
        zz = np.ascontiguousarray(s.zz if zz is None else zz, dtype=np.float64)
        n = self.natom
        f = np.empty((3, n))
        t = (abi.LongTerms * self.nrep)()
        _check(lib().mspms_frclong_host(self._h, abi.dptr(xx), abi.dptr(yy), abi.dptr(zz),
                                        abi.dptr(f[0]), abi.dptr(f[1]), abi.dptr(f[2]), t),
               "mspms_frclong_host")
        terms = [x.as_dict() for x in t]
        return f, (terms[0] if self.nrep == 1 else terms)

    def frcshort(self, xx=None, yy=None, zz=None):
        s = self.system
        xx = np.ascontiguousarray(s.xx if xx is None else xx, dtype=np.float64)
        yy = np.ascontiguousarray(s.yy if yy is None else yy, dtype=np.float64)
        zz = np.ascontiguousarray(s.zz if zz is None else zz, dtype=np.float64)
        n = self.natom
        f = np.empty((3, n))
        t = (abi.ShortTerms * self.nrep)()
        _check(lib().mspms_frcshort_host(self._h, abi.dptr(xx), abi.dptr(yy), abi.dptr(zz),
                                         abi.dptr(f[0]), abi.dptr(f[1]), abi.dptr(f[2]), t),
               "mspms_frcshort_host")
        terms = [x.as_dict() for x in t]
        return f, (terms[0] if self.nrep == 1 else terms)

    # ---- device-buffer evaluation (torch tensors or raw pointers) -----------
    @staticmethod
    def _ptr(t):
        return t if isinstance(t, int) else t.data_ptr()

    def frclong_device(self, pos, force, terms, stream=None):
        """pos, force: (3, N) float64 device tensors; terms: device buffer of
        nrep * sizeof(mspms_long_terms) bytes.  Asynchronous."""
        n = self.natom
        px, fx = self._ptr(pos), self._ptr(force)
        _check(lib().mspms_frclong_device(self._h, px, px + 8 * n, px + 16 * n, fx, fx + 8 * n,
                                          fx + 16 * n, self._ptr(terms), stream),
               "mspms_frclong_device")

    def frcshort_device(self, pos, force, terms, stream=None):
        n = self.natom
        px, fx = self._ptr(pos), self._ptr(force)
        _check(lib().mspms_frcshort_device(self._h, px, px + 8 * n, px + 16 * n, fx, fx + 8 * n,
                                           fx + 16 * n, self._ptr(terms), stream),
               "mspms_frcshort_device")

    def vver_step_device(self, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin,
                         stream=None):
        """One RESPA velocity-Verlet outer step on device tensors (src/md.c
        vver); delt in reduced time units."""
        _check(lib().mspms_vver_step_device(self._h, self._ptr(pos), self._ptr(vel),
                                            self._ptr(fl), self._ptr(fs), float(delt),
                                            int(nstep_inner), self._ptr(lterms),
                                            self._ptr(sterms), self._ptr(ukin), stream),
               "mspms_vver_step_device")

    def sync(self):
        _check(lib().mspms_sync(self._h), "mspms_sync")

    @property
    def stream(self) -> int:
        return lib().mspms_stream(self._h) or 0

    def set_profiling(self, on: bool = True):
        _check(lib().mspms_set_profiling(self._h, 1 if on else 0), "mspms_set_profiling")

    def timings(self) -> dict:
        out = np.zeros(len(PROFILE_GROUPS))
        n = lib().mspms_last_timings(self._h, abi.dptr(out), len(PROFILE_GROUPS))
        return {k: float(out[i]) for i, k in enumerate(PROFILE_GROUPS[:n])}


LONG_TERMS_BYTES = C.sizeof(abi.LongTerms)
SHORT_TERMS_BYTES = C.sizeof(abi.ShortTerms)
