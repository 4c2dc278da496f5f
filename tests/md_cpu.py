"""CPU restatement of the reference's Nose-Hoover NVT RESPA step (test
infrastructure): vver_nh_3 (src/nvtnh.c:11-221), thermostat set-up of
init_nvt (src/initvars.c:298-347), over the oracle's frclong / frcshort.

Expression order as in the reference; every sum over atoms is taken left
to right (np.cumsum), in the reference's x, y, z order per atom.  The
implicit velocity update is the reference's Newton iteration: at most 100
passes, ended when |(p_new - p_old) / p_new| <= 1e-10.  The reference also
tests the atom velocities, but with the loop index left at natom
(src/nvtnh.c:125-139, :225-245): it reads the unused element natom, which is
zero, so 0/0 compares false and only the thermostat test decides."""
from __future__ import annotations

import numpy as np

from oracle.oracle import Oracle

ERR = 1.0e-10  # src/nvtnh.c:19


def seqsum(a: np.ndarray) -> float:
    """Left-to-right FP64 sum (the reference's loop order)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


class NvtCpu:
    """One Nose-Hoover NVT chain of vver_nh_3 steps on the oracle."""

    def __init__(self, system, v0, qq: float, qqs: float, nconstraint: int = 3):
        s = system
        self.s = s
        self.o = Oracle(s)
        self.rm = 1.0 / s.masses()             # r_mass = 1 / aw
        self.treq = s.treq
        self.nfree = 3 * s.natom - nconstraint  # src/initvars.c:588
        self.qq, self.qqs = qq, qqs
        self.gg = self.ggs = float(self.nfree)  # src/initvars.c:332-335
        self.ss = self.ps = self.sss = self.pss = 0.0
        self.unhts = self.unhtss = 0.0
        self.ukin = self.tinst = 0.0
        delt = s.delt
        nin = s.run.nstep_inner
        self.nstep_inner = nin
        # src/initvars.c:604-607, :329-330
        self.delt, self.deltby2 = delt, delt / 2.0
        self.delts = delt / nin
        self.deltsby2 = self.delts / 2.0
        self.delt_sqby2 = delt * delt / 2.0
        self.delts_sqby2 = self.delts * self.delts / 2.0
        self.x = np.stack([s.xx, s.yy, s.zz]).astype(np.float64)
        self.v = np.array(v0, dtype=np.float64)
        self.fl, _ = self.o.frclong(*self.x)
        self.fs, _ = self.o.frcshort(*self.x)

    def _mv2(self, v):
        return seqsum((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / self.rm)

    def _solve(self, v, f, p, q, dt, dtby2, g):
        """Newton iteration for the velocities and thermostat momentum at the
        end of a (sub)step (src/nvtnh.c:64-156 inner, :167-250 outer)."""
        rm, treq = self.rm, self.treq
        vn = v.copy()
        sumv2 = self._mv2(vn)
        pn = p
        it = 0
        ready = False
        while not ready and it < 100:
            it += 1
            po = pn
            vo = vn.copy()
            b = -dtby2 * (f * rm - po * vo) - (v - vo)
            ri = vo * dt / q / rm
            delp = seqsum((ri * b).T)  # atom by atom, x then y then z
            di = -(po * dtby2 + 1.0)
            delp = delp - di * ((-sumv2 + g * treq) * dtby2 / q - (p - po))
            delp = delp / (-dt * dtby2 * sumv2 / q + di)
            vn = vn + (b + dtby2 * vo * delp) / di
            sumv2 = self._mv2(vn)
            pn = po + delp
            with np.errstate(divide="ignore", invalid="ignore"):
                ready = not (abs((pn - po) / pn) > ERR)
        return vn, pn, sumv2

    def step(self):
        rm, treq = self.rm, self.treq
        v = self.v + self.deltby2 * (self.fl * rm - self.ps * self.v)
        sumv2 = 0.0
        for _ in range(self.nstep_inner):
            sumv2 = self._mv2(v)  # before the kick (src/nvtnh.c:40-49)
            v = v + self.deltsby2 * (self.fs * rm - self.pss * v)
            self.x = self.x + self.delts * v
            self.sss = self.sss + self.pss * self.delts + \
                (sumv2 - self.ggs * treq) * self.delts_sqby2 / self.qqs
            self.pss = self.pss + (sumv2 - self.ggs * treq) * self.deltsby2 / self.qqs
            self.fs, _ = self.o.frcshort(*self.x)
            v, self.pss, sumv2 = self._solve(v, self.fs, self.pss, self.qqs, self.delts,
                                             self.deltsby2, self.ggs)
            self.unhtss = (self.pss * self.pss * self.qqs) / 2.0 + self.ggs * treq * self.sss
        self.ss = self.ss + self.ps * self.delt + \
            (sumv2 - self.gg * treq) * self.delt_sqby2 / self.qq
        self.ps = self.ps + (sumv2 - self.gg * treq) * self.deltby2 / self.qq
        self.fl, self.tl = self.o.frclong(*self.x)
        v, self.ps, sumv2 = self._solve(v, self.fl, self.ps, self.qq, self.delt, self.deltby2,
                                        self.gg)
        self.v = v
        self.ukin = sumv2 / 2.0
        self.unhts = (self.ps * self.ps * self.qq) / 2.0 + self.gg * treq * self.ss
        self.tinst = 2.0 * self.ukin / self.nfree


class NveCpu:
    """The reference's NVE RESPA step, vver (src/md.c:22-83), on the oracle."""

    def __init__(self, system, v0, nconstraint: int = 3):
        s = system
        self.o = Oracle(s)
        self.rm = 1.0 / s.masses()
        self.nfree = 3 * s.natom - nconstraint
        delt, nin = s.delt, s.run.nstep_inner
        self.nstep_inner = nin
        self.deltby2 = delt / 2.0
        self.delts = delt / nin
        self.deltsby2 = self.delts / 2.0
        self.x = np.stack([s.xx, s.yy, s.zz]).astype(np.float64)
        self.v = np.array(v0, dtype=np.float64)
        self.fl, _ = self.o.frclong(*self.x)
        self.fs, _ = self.o.frcshort(*self.x)
        self.ukin = self.tinst = 0.0

    def step(self):
        rm = self.rm
        v = self.v + self.deltby2 * self.fl * rm
        for _ in range(self.nstep_inner):
            v = v + self.deltsby2 * self.fs * rm
            self.x = self.x + self.delts * v
            self.fs, _ = self.o.frcshort(*self.x)
            v = v + self.deltsby2 * self.fs * rm
        self.fl, self.tl = self.o.frclong(*self.x)
        v = v + self.deltby2 * self.fl * rm
        self.v = v
        self.ukin = seqsum((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / rm) * 0.5
        self.tinst = 2.0 * self.ukin / self.nfree
