"""CPU restatement of one hybrid Monte Carlo chain (test infrastructure):
hmc() with canonical moves and the equilibration-phase delt adaptation
(src/hmc.c:23-61), md_move (src/dispmd.c:12-103),
vver (src/md.c:22-83), velinit (src/initvars.c:151-232) and gaussran
(src/random.c:113-138) over the oracle's frclong / frcshort and its RANMAR
stream, expression order as in the reference (numpy loops are vectorised
only where the reference's sum order does not matter for the checks)."""
from __future__ import annotations

import math

import numpy as np

from oracle.oracle import Oracle, Ranmar


class Gauss:
    """gaussran over one RANMAR stream (polar method, second deviate kept)."""

    def __init__(self, rng: Ranmar):
        self.rng = rng
        self.iset = False
        self.gset = 0.0
        self.used = 0  # uniform numbers consumed

    def uniform(self) -> float:
        self.used += 1
        return float(self.rng.next(1)[0])

    def __call__(self) -> float:
        if self.iset:
            self.iset = False
            return self.gset
        while True:
            v1 = 2.0 * self.uniform() - 1.0
            v2 = 2.0 * self.uniform() - 1.0
            rsq = v1 * v1 + v2 * v2
            if not (rsq >= 1.0 or rsq == 0.0):
                break
        fac = math.sqrt(-2.0 * math.log(rsq) / rsq)
        self.gset = v1 * fac
        self.iset = True
        return v2 * fac


class HmcChain:
    def __init__(self, system, ij, kl, nconstraint=0):
        self.s = system
        self.o = Oracle(system)
        self.g = Gauss(Ranmar(ij, kl))
        self.m = system.masses()
        self.treq = system.treq
        self.nfree = 3 * system.natom - nconstraint
        self.x = np.stack([system.xx, system.yy, system.zz]).astype(np.float64)
        self.v = self.velinit()
        self.first = True
        # counts[20] / counts[21] of the reference: canonical moves and
        # accepted ones since the last delt adjustment
        self.n_moves = 0
        self.n_accepted = 0

    def velinit(self):
        n = self.s.natom
        v = np.zeros((3, n))
        stdvtmp = math.sqrt(self.treq)
        for i in range(n):
            stdv = stdvtmp / math.sqrt(self.m[i])
            for d in range(3):
                v[d, i] = stdv * self.g()
        p = (v * self.m).sum(axis=1) / self.m.sum()
        v -= p[:, None]
        ukin = 0.5 * float((self.m * (v * v).sum(axis=0)).sum())
        tinst = 2.0 * ukin / self.nfree
        v *= math.sqrt(self.treq / tinst)
        self.ukin = 0.5 * float((self.m * (v * v).sum(axis=0)).sum())
        return v

    def forces(self):
        fl, tl = self.o.frclong(*self.x)
        fs, ts = self.o.frcshort(*self.x)
        return fl, tl, fs, ts

    def vver(self, fl, fs, delt, nstep_inner):
        r = 1.0 / self.m
        deltby2, delts = delt / 2.0, delt / nstep_inner
        deltsby2 = delts / 2.0
        self.v += deltby2 * fl * r
        ts = None
        for _ in range(nstep_inner):
            self.v += deltsby2 * fs * r
            self.x = self.x + delts * self.v
            fs, ts = self.o.frcshort(*self.x)
            self.v += deltsby2 * fs * r
        fl, tl = self.o.frclong(*self.x)
        self.v += deltby2 * fl * r
        self.ukin = 0.5 * float(((self.v * self.v).sum(axis=0) / r).sum())
        return fl, fs, tl, ts

    def move(self, delt, nstep_inner, nmd):
        """One canonical move; returns (accepted, dH)."""
        self.g.uniform()  # the move-type draw of hmc()
        if not self.first:
            self.v = self.velinit()
        self.first = False
        fl, tl, fs, ts = self.forces()
        x_old, ukin_old = self.x.copy(), self.ukin
        old = (tl["uinter"], ts["uintra"], tl["ushift"])
        for _ in range(nmd):
            fl, fs, tl, ts = self.vver(fl, fs, delt, nstep_inner)
        du = (tl["uinter"] - old[0]) + (ts["uintra"] - old[1]) + (self.ukin - ukin_old) + \
            (tl["ushift"] - old[2])
        dh = du / self.treq
        acc = dh <= 0.0 or self.g.uniform() < math.exp(-dh)
        if not acc:
            self.x = x_old
            self.ukin = ukin_old
        return acc, dh

    def run(self, nmove, delt, nstep_inner, nmd, adapt=None):
        """nmove moves of hmc(); adapt = (nstep_delt_adj_cycle, rreq_disp)
        while the chain is in its equilibration phase (bEquilibrium): every
        cycle of moves, delt *= 1 - rreq_disp + accepted / cycle
        (src/hmc.c:47-58).  Returns [(accepted, dH, delt after the move)]."""
        out = []
        for _ in range(nmove):
            acc, dh = self.move(delt, nstep_inner, nmd)
            self.n_moves += 1
            self.n_accepted += 1 if acc else 0
            if adapt is not None and self.n_moves == adapt[0]:
                rinst = self.n_accepted * 1.0 / adapt[0]
                delt = delt * (1.0 - adapt[1] + rinst)
                self.n_moves = self.n_accepted = 0
            out.append((acc, dh, delt))
        return out
