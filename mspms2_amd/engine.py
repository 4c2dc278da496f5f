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
    "mspms_vver_step_device", "mspms_vver_run_device", "mspms_term_pair", "mspms_term_bonded",
    "mspms_autotune", "mspms_set_launch_shape", "mspms_hmc_init", "mspms_hmc_run_device",
    "mspms_hmc_stats", "mspms_hmc_draw", "mspms_set_list_skin", "mspms_list_builds",
    "mspms_nvt_set", "mspms_nvt_get", "mspms_nvt_run_device", "mspms_npt_set", "mspms_npt_get",
    "mspms_npt_run_device", "mspms_siman_run_device",
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
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.
    # Loaded first, it satisfies this library's dependency of that name; the
    # other order loads a second runtime that leaves torch without a GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
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
    L.mspms_vver_run_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp,
                                        vp, vp, vp, vp, vp, C.c_int, vp]
    L.mspms_energy_param_grad.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.POINTER(abi.LongTerms),
                                          C.POINTER(abi.ParamGrad)]
    L.mspms_hmc_init.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp,
                                 C.POINTER(C.c_int), vp]
    L.mspms_hmc_run_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int,
                                       vp, vp, vp, C.c_int, vp]
    L.mspms_hmc_set_delt.argtypes = [vp, dp, C.POINTER(C.c_int), dp]
    L.mspms_hmc_get_delt.argtypes = [vp, dp]
    L.mspms_hmc_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]
    L.mspms_hmc_draw.argtypes = [vp, C.c_int, C.c_int, dp]
    L.mspms_autotune.argtypes = [vp, vp, vp, vp, C.c_int, C.POINTER(abi.TuneResult)]
    L.mspms_set_launch_shape.argtypes = [vp, C.c_double, C.c_int]
    L.mspms_set_list_skin.argtypes = [vp, C.c_double]
    L.mspms_set_sf_split.argtypes = [vp, C.c_int]
    L.mspms_set_list_capacity.argtypes = [vp, C.c_int]
    L.mspms_nvt_set.argtypes = [vp, C.POINTER(abi.NvtState)]
    L.mspms_nvt_get.argtypes = [vp, C.POINTER(abi.NvtState)]
    L.mspms_nvt_run_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp,
                                       vp, C.c_int, vp]
    L.mspms_list_builds.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.mspms_npt_set.argtypes = [vp, C.POINTER(abi.NptState)]
    L.mspms_siman_run_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, dp, dp,
                                         dp, C.c_int, C.POINTER(C.c_int), dp, vp, vp, vp, vp]
    L.mspms_npt_get.argtypes = [vp, C.POINTER(abi.NptState)]
    L.mspms_npt_run_device.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp,
                                       vp, C.c_int, vp]
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

    @classmethod
    def replicas(cls, systems, precision: str = "mixed", device: int = 0) -> "Engine":
        """One engine over a packed table of independent replicas (same
        species templates, any molecule counts, boxes and cutoffs).  Atoms
        are packed replica after replica; `rep_atom0[r]` is the first atom
        of replica r.  Positions of the packed system are the replicas'
        positions concatenated (`pack`)."""
        s0 = systems[0]
        for s in systems[1:]:
            if len(s.species) != len(s0.species) or any(
                    a.natom != b.natom for a, b in zip(s.species, s0.species)):
                raise EngineError("replicas must share the species templates")
        eng = cls(s0, precision, device)
        eng.set_replicas([[sp.nmole for sp in s.species] for s in systems],
                         [s.c_params() for s in systems])
        eng.replica_systems = list(systems)
        eng.rep_atom0 = np.cumsum([0] + [s.natom for s in systems])
        return eng

    def pack(self, arrays):
        """Concatenates per-replica (3, n_r) arrays into the packed (3, N)."""
        return np.concatenate([np.asarray(a, dtype=np.float64) for a in arrays], axis=1)

    def unpack(self, f):
        """Splits a packed (3, N) array into per-replica views."""
        a = self.rep_atom0
        return [f[:, a[r]:a[r + 1]] for r in range(len(a) - 1)]

    def packed_positions(self):
        return self.pack([np.stack([s.xx, s.yy, s.zz]) for s in self.replica_systems])

    # ---- host-buffer evaluation --------------------------------------------
    def _xyz(self, xx, yy, zz):
        if xx is None:
            if getattr(self, "replica_systems", None):
                xx, yy, zz = self.packed_positions()
            else:
                xx, yy, zz = self.system.xx, self.system.yy, self.system.zz
        return tuple(np.ascontiguousarray(v, dtype=np.float64) for v in (xx, yy, zz))

    def frclong(self, xx=None, yy=None, zz=None):
        xx, yy, zz = self._xyz(xx, yy, zz)
        n = self.natom
        f = np.empty((3, n))
        t = (abi.LongTerms * self.nrep)()
        _check(lib().mspms_frclong_host(self._h, abi.dptr(xx), abi.dptr(yy), abi.dptr(zz),
                                        abi.dptr(f[0]), abi.dptr(f[1]), abi.dptr(f[2]), t),
               "mspms_frclong_host")
        terms = [x.as_dict() for x in t]
        return f, (terms[0] if self.nrep == 1 else terms)

    def frcshort(self, xx=None, yy=None, zz=None):
        xx, yy, zz = self._xyz(xx, yy, zz)
        n = self.natom
        f = np.empty((3, n))
        t = (abi.ShortTerms * self.nrep)()
        _check(lib().mspms_frcshort_host(self._h, abi.dptr(xx), abi.dptr(yy), abi.dptr(zz),
                                         abi.dptr(f[0]), abi.dptr(f[1]), abi.dptr(f[2]), t),
               "mspms_frcshort_host")
        terms = [x.as_dict() for x in t]
        return f, (terms[0] if self.nrep == 1 else terms)

    def param_grad(self, xx=None, yy=None, zz=None):
        """FP64 frclong plus dE/d(epsilon, sigma, charge) of uinter per
        replica and global site id: returns (forces, terms, grads) with
        grads[name] of shape (nrep, nsite)."""
        xx, yy, zz = self._xyz(xx, yy, zz)
        n = self.natom
        nsite = sum(s.natom for s in self.system.species)
        f = np.empty((3, n))
        t = (abi.LongTerms * self.nrep)()
        g = {k: np.zeros((self.nrep, nsite)) for k in ("epsilon", "sigma", "charge")}
        pg = abi.ParamGrad(abi.dptr(g["epsilon"]), abi.dptr(g["sigma"]), abi.dptr(g["charge"]))
        _check(lib().mspms_energy_param_grad(self._h, abi.dptr(xx), abi.dptr(yy), abi.dptr(zz),
                                             abi.dptr(f[0]), abi.dptr(f[1]), abi.dptr(f[2]), t,
                                             C.byref(pg)), "mspms_energy_param_grad")
        terms = [x.as_dict() for x in t]
        return f, (terms[0] if self.nrep == 1 else terms), g

    def set_site_params(self, epsilon=None, sigma=None, charge=None):
        """Replaces the per-site template values (arrays over global site
        ids, reduced units) and rebuilds the device tables."""
        off = 0
        for sp in self.system.species:
            sl = slice(off, off + sp.natom)
            for name, v in (("epsilon", epsilon), ("sigma", sigma), ("charge", charge)):
                if v is not None:
                    getattr(sp, name)[:] = np.asarray(v, dtype=np.float64)[sl]
            off += sp.natom
        reps = getattr(self, "replica_systems", None)
        self._sp, self._keep = self.system.c_species()
        _check(lib().mspms_set_system(self._h, len(self.system.species), self._sp,
                                      C.byref(self.params)), "mspms_set_system")
        if reps:
            arr, nm, params = self._rep_keep
            _check(lib().mspms_set_replicas(self._h, len(params),
                                            nm.ctypes.data_as(C.POINTER(C.c_int)), arr),
                   "mspms_set_replicas")

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

    def vver_run_device(self, pos, vel, fl, fs, delt, nstep_inner, nsteps, lterms, sterms, ukin,
                        lhist=None, shist=None, ukin_hist=None, graph=True, stream=None):
        """nsteps RESPA outer steps; graph=True replays one captured step."""
        p = self._ptr
        _check(lib().mspms_vver_run_device(
            self._h, p(pos), p(vel), p(fl), p(fs), float(delt), int(nstep_inner), int(nsteps),
            p(lterms), p(sterms), p(ukin), None if lhist is None else p(lhist),
            None if shist is None else p(shist), None if ukin_hist is None else p(ukin_hist),
            1 if graph else 0, stream), "mspms_vver_run_device")

    # ---- batched HMC (one chain per replica) ---------------------------------
    def hmc_init(self, ij, kl, treq, nconstraint, vel):
        """Seeds the replicas' RANMAR streams and draws initial velocities
        into the (3, N) device tensor `vel`."""
        n = self.nrep
        ij = np.ascontiguousarray(np.broadcast_to(ij, n), dtype=np.int32)
        kl = np.ascontiguousarray(np.broadcast_to(kl, n), dtype=np.int32)
        tq = np.ascontiguousarray(np.broadcast_to(treq, n), dtype=np.float64)
        nc = np.ascontiguousarray(np.broadcast_to(nconstraint, n), dtype=np.int32)
        _check(lib().mspms_hmc_init(self._h, abi.iptr(ij), abi.iptr(kl), abi.dptr(tq),
                                    abi.iptr(nc), self._ptr(vel)), "mspms_hmc_init")

    def hmc_set_delt(self, delt, adj_cycle=0, rreq_disp=0.0):
        """Per-chain outer steps (reduced) and the equilibration-phase delt
        adaptation: every adj_cycle moves, delt *= 1 - rreq_disp +
        accepted / adj_cycle (adj_cycle 0: fixed); scalars or per-replica
        sequences."""
        n = self.nrep
        d = np.ascontiguousarray(np.broadcast_to(delt, n), dtype=np.float64)
        cy = np.ascontiguousarray(np.broadcast_to(adj_cycle, n), dtype=np.int32)
        rr = np.ascontiguousarray(np.broadcast_to(rreq_disp, n), dtype=np.float64)
        _check(lib().mspms_hmc_set_delt(self._h, abi.dptr(d), abi.iptr(cy), abi.dptr(rr)),
               "mspms_hmc_set_delt")

    def hmc_get_delt(self):
        out = np.zeros(self.nrep)
        _check(lib().mspms_hmc_get_delt(self._h, abi.dptr(out)), "mspms_hmc_get_delt")
        return out

    def hmc_run_device(self, pos, vel, fl, fs, delt, nstep_inner, nmd, nmove, lterms, sterms,
                       ukin, graph=True, stream=None):
        p = self._ptr
        _check(lib().mspms_hmc_run_device(self._h, p(pos), p(vel), p(fl), p(fs), float(delt),
                                          int(nstep_inner), int(nmd), int(nmove), p(lterms),
                                          p(sterms), p(ukin), 1 if graph else 0, stream),
               "mspms_hmc_run_device")

    def nvt_set(self, qq, qqs, treq, nfree, ss=0.0, ps=0.0, sss=0.0, pss=0.0):
        """Nose-Hoover thermostats of every replica (init_nvt's start when
        the positions and momenta are left at zero); scalars or per-replica
        sequences; gg = ggs = nfree."""
        n = self.nrep
        arr = (abi.NvtState * n)()
        b = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), n)  # noqa: E731
        cols = dict(qq=b(qq), qqs=b(qqs), treq=b(treq), ss=b(ss), ps=b(ps), sss=b(sss),
                    pss=b(pss), nf=np.broadcast_to(np.asarray(nfree), n))
        for r in range(n):
            st = arr[r]
            st.qq, st.qqs, st.treq = cols["qq"][r], cols["qqs"][r], cols["treq"][r]
            st.gg = st.ggs = float(cols["nf"][r])
            st.nfree = int(cols["nf"][r])
            st.ss, st.ps, st.sss, st.pss = cols["ss"][r], cols["ps"][r], cols["sss"][r], \
                cols["pss"][r]
        _check(lib().mspms_nvt_set(self._h, arr), "mspms_nvt_set")

    def nvt_get(self):
        arr = (abi.NvtState * self.nrep)()
        _check(lib().mspms_nvt_get(self._h, arr), "mspms_nvt_get")
        return [a.as_dict() for a in arr]

    def nvt_run_device(self, pos, vel, fl, fs, delt, nstep_inner, nsteps, lterms, sterms, ukin,
                       graph=True, stream=None):
        """nsteps vver_nh_3 steps of every replica on (3, N) device tensors;
        fl / fs must hold the forces of pos."""
        p = self._ptr
        _check(lib().mspms_nvt_run_device(self._h, p(pos), p(vel), p(fl), p(fs), float(delt),
                                          int(nstep_inner), int(nsteps), p(lterms), p(sterms),
                                          p(ukin), 1 if graph else 0, stream),
               "mspms_nvt_run_device")

    def siman_run_device(self, pos, vel, fl, fs, delt, nstep_inner, n_tries, t_initial, t_min,
                         mu_t, lterms, sterms, ukin, max_cycles=1000, stream=None):
        """Simulated annealing of every replica (reduced temperatures; scalars
        or per-replica sequences); returns (cycles, t_final) per replica."""
        n = self.nrep
        b = lambda v: np.ascontiguousarray(np.broadcast_to(v, n), dtype=np.float64)  # noqa
        ti, tm, mu = b(t_initial), b(t_min), b(mu_t)
        cyc = np.zeros(n, dtype=np.int32)
        tf = np.zeros(n)
        p = self._ptr
        _check(lib().mspms_siman_run_device(
            self._h, p(pos), p(vel), p(fl), p(fs), float(delt), int(nstep_inner), int(n_tries),
            abi.dptr(ti), abi.dptr(tm), abi.dptr(mu), int(max_cycles),
            cyc.ctypes.data_as(C.POINTER(C.c_int)), abi.dptr(tf), p(lterms), p(sterms),
            p(ukin), stream), "mspms_siman_run_device")
        return cyc, tf

    def npt_set(self, Qts, Qbs, treq, preq, nfree, lj_tail=False, vts=None, vbs=None,
                rts=0.0):
        """NPT thermostat / barostat of every replica; vts / vbs default to
        init_npt_respa's sqrt((nfree + 1) / Q) (src/initvars.c:379-381).
        Resets every replica to the box of its parameters."""
        n = self.nrep
        b = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), n)  # noqa: E731
        qt, qb, tq, pq, rt = b(Qts), b(Qbs), b(treq), b(preq), b(rts)
        nf = np.broadcast_to(np.asarray(nfree), n)
        lt = np.broadcast_to(np.asarray(lj_tail), n)
        arr = (abi.NptState * n)()
        for r in range(n):
            st = arr[r]
            st.Qts, st.Qbs, st.treq, st.preq, st.rts = qt[r], qb[r], tq[r], pq[r], rt[r]
            st.nfree = int(nf[r])
            st.lj_tail = 1 if lt[r] else 0
            st.vts = float(b(vts)[r]) if vts is not None else np.sqrt((nf[r] + 1.0) / qt[r])
            st.vbs = float(b(vbs)[r]) if vbs is not None else np.sqrt((nf[r] + 1.0) / qb[r])
        _check(lib().mspms_npt_set(self._h, arr), "mspms_npt_set")

    def npt_get(self):
        arr = (abi.NptState * self.nrep)()
        _check(lib().mspms_npt_get(self._h, arr), "mspms_npt_get")
        return [a.as_dict() for a in arr]

    def npt_run_device(self, pos, vel, fl, fs, delt, nstep_inner, nsteps, lterms, sterms, ukin,
                       graph=True, stream=None):
        """nsteps npt_respa steps of every replica on (3, N) device tensors;
        fl / fs and lterms / sterms must hold the forces and terms of pos."""
        p = self._ptr
        _check(lib().mspms_npt_run_device(self._h, p(pos), p(vel), p(fl), p(fs), float(delt),
                                          int(nstep_inner), int(nsteps), p(lterms), p(sterms),
                                          p(ukin), 1 if graph else 0, stream),
               "mspms_npt_run_device")

    def hmc_stats(self):
        n = self.nrep
        acc = np.zeros(n, dtype=np.int64)
        mov = np.zeros(n, dtype=np.int64)
        dh = np.zeros(n)
        _check(lib().mspms_hmc_stats(self._h, acc.ctypes.data_as(C.POINTER(C.c_longlong)),
                                     mov.ctypes.data_as(C.POINTER(C.c_longlong)), abi.dptr(dh)),
               "mspms_hmc_stats")
        return acc, mov, dh

    def hmc_draw(self, replica, n):
        out = np.zeros(n)
        _check(lib().mspms_hmc_draw(self._h, int(replica), int(n), abi.dptr(out)),
               "mspms_hmc_draw")
        return out

    def autotune(self, pos, repeats=3) -> dict:
        """Picks the fastest launch shape for frclong on the (3, N) device
        positions `pos`; returns the chosen shape and every candidate's time."""
        p = self._ptr(pos)
        n = self.natom
        r = abi.TuneResult()
        _check(lib().mspms_autotune(self._h, p, p + 8 * n, p + 16 * n, int(repeats), C.byref(r)),
               "mspms_autotune")
        return r.as_dict()

    def set_launch_shape(self, cell_atoms=8.0, waves_per_cluster=4):
        _check(lib().mspms_set_launch_shape(self._h, float(cell_atoms), int(waves_per_cluster)),
               "mspms_set_launch_shape")

    def set_list_skin(self, skin: float):
        """Pair-list buffer: lists are built with the cutoff + skin and reused
        while every site stays within skin / 2 of its position at the build."""
        _check(lib().mspms_set_list_skin(self._h, C.c_double(float(skin))), "mspms_set_list_skin")

    def set_list_capacity(self, max_entries: int):
        """Bound on the pair-list row capacity (0: none)."""
        _check(lib().mspms_set_list_capacity(self._h, int(max_entries)),
               "mspms_set_list_capacity")

    def set_sf_split(self, on: bool = True):
        """RESPA split of the solid-fluid wall: its forces move from frclong
        to frcshort (the inner loop of a device RESPA step)."""
        _check(lib().mspms_set_sf_split(self._h, 1 if on else 0), "mspms_set_sf_split")

    def list_builds(self) -> int:
        n = C.c_longlong(0)
        _check(lib().mspms_list_builds(self._h, C.byref(n)), "mspms_list_builds")
        return int(n.value)

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
