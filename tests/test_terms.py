"""Term level, on the CPU: the engine's term code (physics.h / terms.h, the
code the GPU kernels run, compiled for the host and exported as
mspms_term_pair / mspms_term_bonded) against the oracle's restatement of
ljfrc / ewald_real_frc / ewald_excl_frc / wolf_real_frc (src/ljfrc.c:21-71,
src/elecfrc.c:17-60) and bndfrc / aglfrc / dihfrc / impfrc
(src/bndfrc.c:17-78, src/aglfrc.c:17-190, src/dihfrc.c:17-284,
src/impfrc.c:17-191) on the same random inputs.  Both are FP64 in the
reference's expression order with host libm, so the results must be
bit-identical.  No GPU is used (the library is loaded, no HIP call is made).
"""
import ctypes as C

import numpy as np
import pytest

from mspms2_amd import abi, configs
from mspms2_amd import engine as E
from oracle import oracle as O
from tests.util import system_of

dp = C.POINTER(C.c_double)


def _libs():
    L = E.lib()
    L.mspms_term_pair.argtypes = [C.POINTER(abi.Params), C.c_int, C.c_double, C.c_double,
                                  C.c_double, dp, dp, dp]
    L.mspms_term_bonded.argtypes = [C.POINTER(abi.Species), C.c_int, C.c_int, dp, dp, dp, dp,
                                    dp, dp, dp, dp]
    R = O.lib()
    R.oracle_term_pair.argtypes = [C.POINTER(abi.Params), C.c_int, C.c_double, C.c_double,
                                   C.c_double, dp, dp, dp]
    R.oracle_term_bonded.argtypes = [C.POINTER(abi.Species), C.c_int, C.c_int, dp, dp, dp, dp,
                                     dp, dp, dp, dp]
    return L, R


def _params(system):
    p = system.c_params()
    E.derive_params(p)
    return p


PAIR_SYSTEMS = {
    "ewald": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3, kmax=5,
                                       ksqmax=27, seed=3),
    "wolf_switch": lambda: configs.mixture_box(elec="wolf", kappa=0.2, switch=True, seed=13),
    "lj_shift": lambda: configs.lj_fcc(),
}


@pytest.mark.parametrize("name", list(PAIR_SYSTEMS))
def test_pair_terms_bit_exact(name):
    L, R = _libs()
    s = system_of(PAIR_SYSTEMS[name]())
    p = _params(s)
    rng = np.random.default_rng(5)
    rc2 = max(p.rcutoffsq, p.rcutoffelecsq)
    out = [C.c_double() for _ in range(6)]
    n = 0
    for rsq in rng.uniform(0.3 * rc2 / 36, rc2, 400):
        a, b = rng.uniform(0.5, 1.5), rng.uniform(0.1, 3.0)
        qq = rng.uniform(-1.0, 1.0)
        for which, (x, y) in ((0, (a, b)), (1, (qq, 0.0)), (2, (qq, 0.0)), (3, (qq, 0.0)),
                               (4, (qq, 0.0))):
            for o in out:
                o.value = np.nan
            assert L.mspms_term_pair(C.byref(p), which, rsq, x, y, C.byref(out[0]),
                                     C.byref(out[1]), C.byref(out[2])) == 0
            R.oracle_term_pair(C.byref(p), which, rsq, x, y, C.byref(out[3]), C.byref(out[4]),
                               C.byref(out[5]))
            got = [o.value for o in out[:3]]
            ref = [o.value for o in out[3:]]
            assert got == ref, (name, which, rsq, got, ref)
            n += 1
    assert n == 2000


@pytest.mark.parametrize("name", ["bonded_zoo", "butane", "examples_water"])
def test_bonded_terms_bit_exact(name):
    """Every bond / angle / dihedral / improper of the species on strained
    copies of its template (sites displaced up to 0.1 A)."""
    import os
    from mspms2_amd import runio
    from mspms2_amd.system import System
    from tests.util import GOLDEN
    L, R = _libs()
    if name == "examples_water":
        s = System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, name)))
    else:
        s = system_of(configs.bonded_zoo() if name == "bonded_zoo" else configs.butane_box(64))
    sp, keep = s.c_species()
    rng = np.random.default_rng(11)
    checked = 0
    base = 0
    for k, q in enumerate(s.species):
        x0 = np.stack([s.xx[base:base + q.natom], s.yy[base:base + q.natom],
                       s.zz[base:base + q.natom]])
        base += q.natom * q.nmole
        nterm = {0: len(q.bnd_type), 1: len(q.agl_type), 2: len(q.dih_type), 3: len(q.imp_type)}
        for trial in range(20):
            x = x0 + rng.uniform(-0.1, 0.1, x0.shape) / s.params["sigma_base"]
            xx, yy, zz = (np.ascontiguousarray(v) for v in x)
            for kind, cnt in nterm.items():
                for t in range(cnt):
                    fa = np.zeros((3, q.natom))
                    fb = np.zeros((3, q.natom))
                    ua, va, ub, vb = (C.c_double() for _ in range(4))
                    ra = L.mspms_term_bonded(C.byref(sp[k]), kind, t, abi.dptr(xx), abi.dptr(yy),
                                             abi.dptr(zz), abi.dptr(fa[0]), abi.dptr(fa[1]),
                                             abi.dptr(fa[2]), C.byref(ua), C.byref(va))
                    rb = R.oracle_term_bonded(C.byref(sp[k]), kind, t, abi.dptr(xx),
                                              abi.dptr(yy), abi.dptr(zz), abi.dptr(fb[0]),
                                              abi.dptr(fb[1]), abi.dptr(fb[2]), C.byref(ub),
                                              C.byref(vb))
                    assert (ra == 0) == (rb == 0)
                    assert ua.value == ub.value and va.value == vb.value, (name, kind, t)
                    assert np.array_equal(fa, fb), (name, kind, t)
                    checked += 1
    assert checked > 0
    print(f"{name}: {checked} term evaluations bit-exact")


def test_term_errors():
    L, _ = _libs()
    s = system_of(configs.butane_box(64))
    p = _params(s)
    u = C.c_double()
    assert L.mspms_term_pair(C.byref(p), 7, 1.0, 1.0, 1.0, C.byref(u), C.byref(u),
                             C.byref(u)) == abi.MSPMS_EINVAL
    sp, keep = s.c_species()
    x = np.zeros(4)
    assert L.mspms_term_bonded(C.byref(sp[0]), 0, 99, abi.dptr(x), abi.dptr(x), abi.dptr(x),
                               abi.dptr(x), abi.dptr(x), abi.dptr(x), C.byref(u),
                               C.byref(u)) == abi.MSPMS_EINVAL
