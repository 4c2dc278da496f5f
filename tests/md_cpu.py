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

import math

import numpy as np

from oracle.oracle import Oracle

ERR = 1.0e-10  # src/nvtnh.c:19


def seqsum(a: np.ndarray) -> float:
    """Left-to-right FP64 sum (the reference's loop order)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


class NvtCpu:
    """One Nose-Hoover NVT chain of vver_nh_3 steps on the oracle."""

    def __init__(self, system, v0, qq: float, qqs: float, nconstraint: int = 3,
                 state: dict = None, x0=None):
        s = system
        self.s = s
        self.o = Oracle(s)
        self.rm = 1.0 / s.masses()             # r_mass = 1 / aw
        self.treq = s.treq
        self.nfree = 3 * s.natom - nconstraint  # src/initvars.c:588
        self.qq, self.qqs = qq, qqs
        self.gg = self.ggs = float(self.nfree)  # src/initvars.c:332-335
        self.ss = self.ps = self.sss = self.pss = 0.0
        # a restart's thermostats (loadit, src/loadsave.c:86-93)
        for k, v in (state or {}).items():
            setattr(self, k, v)
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
        self.x = (np.stack([s.xx, s.yy, s.zz]) if x0 is None else np.array(x0)).astype(np.float64)
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


PI = 3.141592653589793  # src/defs.h:98


def ljlrc_pressure_terms(system) -> np.ndarray:
    """pljlrc_term[m][n] of InitLJlrcCommonTerms (src/initvars.c:474-521)."""
    s = system
    rc = s.params["rcutoff"]
    t1 = (32.0 / 9.0) * PI * rc ** -9.0
    t2 = -(16.0 / 3.0) * PI * rc ** -3.0
    ns = len(s.species)
    out = np.zeros((ns, ns))
    for m, a in enumerate(s.species):
        for n, b in enumerate(s.species):
            acc = 0.0
            for i in range(a.natom):
                for j in range(b.natom):
                    sij = 0.5 * (a.sigma[i] + b.sigma[j])
                    eij = math.sqrt(a.epsilon[i] * b.epsilon[j])
                    temp2 = eij * sij ** 3.0
                    temp3 = sij ** 9.0 * t1 + sij ** 3.0 * t2
                    acc += temp3 * temp2
            out[m, n] = acc
    return out


class NptCpu:
    """The reference's NPT RESPA step, npt_respa with npt_nh_operator
    (src/nptnh.c:10-234; set-up init_npt_respa, src/initvars.c:351-396),
    on the oracle, whose box-dependent parameters follow the box as the
    reference's globals do (src/nptnh.c:180-199)."""

    def __init__(self, system, v0, Qts: float, Qbs: float, nconstraint: int = 3):
        s = system
        self.s = s
        self.o = Oracle(s)
        self.rm = 1.0 / s.masses()
        self.treq, self.preq = s.treq, s.preq
        self.natom = s.natom
        self.nfree = 3 * s.natom - nconstraint
        self.Qts, self.Qbs = Qts, Qbs
        self.vts = math.sqrt((self.nfree + 1.0) / Qts)
        self.vbs = math.sqrt((self.nfree + 1.0) / Qbs)
        self.rts = 0.0
        self.box = [s.params["boxlx"], s.params["boxly"], s.params["boxlz"]]
        self.boxv = self.box[0] * self.box[1] * self.box[2]
        self.nmole = [sp.nmole for sp in s.species]
        self.pterm = ljlrc_pressure_terms(s) if s.run.isLJlrcOn else np.zeros(
            (len(s.species), len(s.species)))
        self.pljlrc = self._ljlrc()
        delt, nin = s.delt, s.run.nstep_inner
        self.nstep_inner = nin
        self.deltby2 = delt / 2.0
        self.delts = delt / nin
        self.deltsby2 = self.delts / 2.0
        self.dt_outer2, self.dt_outer4, self.dt_outer8 = self.deltby2, delt / 4.0, delt / 8.0
        self.ewald = s.params["iChargeType"] == 1
        self.x = np.stack([s.xx, s.yy, s.zz]).astype(np.float64)
        self.v = np.array(v0, dtype=np.float64)
        self.fl, self.tl = self.o.frclong(*self.x)
        self.fs, self.ts = self.o.frcshort(*self.x)
        self.utsbs = self.ukin = self.tinst = self.pinst = 0.0

    def _ljlrc(self):
        """calculate_ljlrc's pressure part (src/misc.c:131-148)"""
        p = 0.0
        for m in range(len(self.nmole)):
            for n in range(len(self.nmole)):
                p += self.pterm[m, n] * self.nmole[m] * self.nmole[n] / self.boxv / self.boxv
        return p

    def op(self):
        """npt_nh_operator (src/nptnh.c:10-130)"""
        rm, treq, nfree = self.rm, self.treq, self.nfree
        v = self.v
        N_plus1RT = (nfree + 1) * treq
        one_3N = 1.0 + 3.0 / nfree
        mvsq = seqsum((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / rm)
        pdiff = self.tl["virial_inter"] + self.ts["virial_intra"] + \
            self.pljlrc * self.boxv * 3.0 - self.preq * self.boxv * 3.0
        Gts = (mvsq + self.Qbs * self.vbs * self.vbs - N_plus1RT) / self.Qts
        self.vts = self.vts + self.dt_outer4 * Gts
        BB = math.exp(-self.vts * self.dt_outer8)
        self.vbs = self.vbs * BB
        Gbs = (one_3N * mvsq + pdiff) / self.Qbs
        self.vbs = self.vbs + self.dt_outer4 * Gbs
        self.vbs = self.vbs * BB
        self.rts = self.rts + self.deltby2 * self.vts
        AA = math.exp(-self.dt_outer2 * (self.vts + one_3N * self.vbs))
        self.v = v * AA
        self.vbs = self.vbs * BB
        mvsq = mvsq * AA * AA
        Gbs = (one_3N * mvsq + pdiff) / self.Qbs
        self.vbs = self.vbs + self.dt_outer4 * Gbs
        self.vbs = self.vbs * BB
        Gts = (mvsq + self.Qbs * self.vbs * self.vbs - N_plus1RT) / self.Qts
        self.vts = self.vts + self.dt_outer4 * Gts
        self.utsbs = 0.5 * self.Qbs * self.vbs * self.vbs + 0.5 * self.Qts * self.vts * \
            self.vts + (nfree + 1) * treq * self.rts + self.preq * self.boxv
        self.ukin = mvsq / 2.0
        self.tinst = 2.0 * self.ukin / nfree
        pideal = self.natom / self.boxv * self.tinst
        self.pinst = pideal + (self.tl["virial_inter"] + self.ts["virial_intra"]) / 3.0 / \
            self.boxv + self.pljlrc

    def step(self):
        rm = self.rm
        self.op()
        v = self.v + self.deltby2 * self.fl * rm
        for _ in range(self.nstep_inner):
            AA = math.exp(self.deltsby2 * self.vbs)
            v = v + self.deltsby2 * self.fs * rm
            x = self.x * AA
            x = x + self.delts * v
            self.x = x * AA
            expfactor = math.exp(self.delts * self.vbs)
            self.box = [b * expfactor for b in self.box]
            self.boxv = self.box[0] * self.box[1] * self.box[2]
            self.pljlrc = self._ljlrc()
            p = self.o.p
            p.boxlx, p.boxly, p.boxlz = self.box
            if self.ewald:
                p.vfactor_ewald = 2.0 * PI / self.boxv
                p.twopi_lx = 2.0 * PI / self.box[0]
                p.twopi_ly = 2.0 * PI / self.box[1]
                p.twopi_lz = 2.0 * PI / self.box[2]
                p.twopi_over_3v = 2.0 * PI / 3.0 / self.boxv
            self.fs, self.ts = self.o.frcshort(*self.x)
            v = v + self.deltsby2 * self.fs * rm
        self.fl, self.tl = self.o.frclong(*self.x)
        self.v = v + self.deltby2 * self.fl * rm
        self.op()


class SimanCpu:
    """The reference's simulated annealing, siman (src/siman.c:61-182; set-up
    init_siman, :8-59): cycles at T = t_initial, t_initial / mu_t, ... of
    velinit at T (RANMAR seeded by rmarin(ij, jk) in init_vars, whose own
    velinit the cycles follow) and n_tries vver_nh_3 steps from zeroed
    thermostats, ending after the first cycle run below t_min."""

    def __init__(self, system, run):
        from tests.hmc_cpu import HmcChain
        self.s, self.run = system, run
        # rmarin + init_vars' velinit (its draws are consumed, its velocities
        # replaced by the first cycle's)
        self.chain = HmcChain(system, run.ij, run.jk, run.nconstraint)
        self.x = self.chain.x
        eb = system.params["epsilon_base"]
        self.T, self.t_min = run.t_initial / eb, run.t_min / eb
        self.cycles = 0

    def anneal(self):
        r = self.run
        while True:
            self.chain.treq = self.T
            v = self.chain.velinit()
            nvt = NvtCpu(self.s, v, r.qq, r.qqs, r.nconstraint, x0=self.x)
            nvt.treq = self.T
            for _ in range(r.n_tries):
                nvt.step()
            self.x, self.v, self.nvt = nvt.x, nvt.v, nvt
            self.cycles += 1
            if self.T < self.t_min:
                return self
            self.T /= r.mu_t
