"""The CPU restatements of the reference's integrators (tests/md_cpu.py:
vver_nh_3 and vver) against the reference's own trajectories
(tests/golden/md_*: its compiled integrators run from velinit's
velocities, tests/golden/make_golden.py).  Both sides evaluate the same
FP64 expressions in the same order over bit-identical forces (the oracle's
frclong / frcshort are pinned to the reference's), so the states agree to
the last bit; the bound below leaves room for a different libm only."""
import os

import numpy as np
import pytest

from mspms2_amd import runio
from mspms2_amd.system import System
from tests.md_cpu import NptCpu, NveCpu, NvtCpu
from tests.util import GOLDEN

CASES = [("md_lj864_nvt", 1), ("md_lj864_nvt", 10), ("md_lj864_nvt", 100),
         ("md_water216_nvt", 1), ("md_water216_nvt", 10),
         ("md_lj864_npt", 1), ("md_lj864_npt", 10),
         ("md_water216_npt", 1), ("md_water216_npt", 5),
         ("md_water216_nve", 1), ("md_water216_nve", 10)]


def load(name, n):
    path = os.path.join(GOLDEN, name)
    rd = runio.RunDir.read(path)
    g = np.load(os.path.join(path, f"md{n}.npz"))
    return rd, System.from_rundir(rd), g


def rel(a, b):
    """||a - b|| / ||b||, or ||a|| where b is zero (no bonded forces)"""
    d = float(np.linalg.norm(np.asarray(a) - b))
    n = float(np.linalg.norm(b))
    return d / n if n > 0 else d


@pytest.mark.parametrize("name,n", CASES)
def test_restatement_matches_reference_trajectory(name, n):
    rd, s, g = load(name, n)
    if rd.run.what_ensemble == 1:
        c = NvtCpu(s, g["v0"], rd.run.qq, rd.run.qqs, rd.run.nconstraint)
    elif rd.run.what_ensemble == 2:
        c = NptCpu(s, g["v0"], rd.run.Qts, rd.run.Qbs, rd.run.nconstraint)
    else:
        c = NveCpu(s, g["v0"], rd.run.nconstraint)
    for _ in range(n):
        c.step()
    tol = 1e-12
    assert rel(c.x, g["x"]) <= tol
    assert rel(c.v, g["v"]) <= tol
    assert rel(c.fl, g["fl"]) <= tol
    assert rel(c.fs, g["fs"]) <= tol
    assert abs(c.ukin - g["md_ukin"]) <= tol * abs(g["md_ukin"])
    if rd.run.what_ensemble == 1:
        for k in ("ps", "pss", "ss", "sss", "unhts", "unhtss"):
            ref = float(g[f"md_{k}"])
            assert abs(getattr(c, k) - ref) <= tol * max(abs(ref), 1e-300), k
    if rd.run.what_ensemble == 2:
        assert np.allclose(c.box, g["box"], rtol=tol, atol=0)
        for k in ("vts", "vbs", "rts", "utsbs", "pinst", "tinst"):
            ref = float(g[f"md_{k}"])
            assert abs(getattr(c, k) - ref) <= tol * max(abs(ref), 1e-300), k


@pytest.mark.parametrize("name,nstep", [("md_lj864_nvt", 60), ("md_water216_nvt", 20)])
def test_nvt_extended_energy_is_conserved(name, nstep):
    """What the mathematics fixes: the Nose-Hoover dynamics conserves the
    extended energy H = ukin + uinter + uintra + unhts + unhtss (kinetic,
    potential and both thermostats' energies; src/nvtnh.c:157, :265) up to
    the integrator's O(dt^2) fluctuation, far below the kinetic energy's own
    swings while the thermostat pulls the temperature.  A sign or a term
    wrong in either thermostat breaks it; leaving out the inner thermostat's
    energy alone already gives a fluctuation of the kinetic energy's size."""
    rd, s, g = load(name, 1)
    c = NvtCpu(s, g["v0"], rd.run.qq, rd.run.qqs, rd.run.nconstraint)
    h, k = [], []
    for _ in range(nstep):
        c.step()
        up = c.tl["uinter"] + c.o.frcshort(*c.x)[1]["uintra"]
        h.append(c.ukin + up + c.unhts + c.unhtss)
        k.append(c.ukin)
    assert np.std(h) <= 0.1 * np.std(k)


def test_siman_restatement_matches_reference():
    """The annealing restatement (tests/md_cpu.py SimanCpu) against the
    reference's own siman on LJ-256 (tests/golden/siman_lj256: three cycles,
    150 K -> 100 K -> 66.7 K, 20 NVT steps each).  velinit's numpy sums are
    taken in another order than the reference's loop (1e-16 per velocity),
    so the bound is 1e-12."""
    path = os.path.join(GOLDEN, "siman_lj256")
    rd = runio.RunDir.read(path)
    s = System.from_rundir(rd)
    g = np.load(os.path.join(path, "siman.npz"))
    from tests.md_cpu import SimanCpu
    c = SimanCpu(s, rd.run).anneal()
    assert c.cycles == 3 and c.T == float(g["treq"])
    assert rel(c.x, g["x"]) <= 1e-12 and rel(c.v, g["v"]) <= 1e-12
