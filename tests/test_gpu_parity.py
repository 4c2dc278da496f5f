"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on
identical seeded inputs.

Tolerances:
  FP64 mode   pair counts and pair-set hashes bit-exact; every energy within
              1e-10 relative (uelec / uinter / virial relative to the
              cancelling Ewald components); forces ||dF||/||F|| <= 1e-10.
  MIXED mode  forces ||dF||/||F|| <= 1e-5; energies within 1e-5 relative
              to the same scales; bonded / intramolecular / reciprocal terms
              stay FP64 and are held to 1e-10.
"""
import os

import numpy as np
import pytest

from mspms2_amd import configs, runio
from mspms2_amd.engine import Engine
from mspms2_amd.system import System
from oracle.oracle import Oracle
from tests.util import (GOLDEN, assert_long_close, assert_short_close, force_close, rel_rms,
                        system_of, tube_grid_system)

pytestmark = pytest.mark.gpu

CASES = {
    "lj864_jitter": lambda: configs.lj_fcc(jitter=0.05, seed=7),
    "lj864_lattice": lambda: configs.lj_fcc(),
    "water216_ewald": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                kmax=5, ksqmax=27, seed=3),
    "water216_wolf": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.2,
                                               elec="wolf", seed=4),
    "butane64": lambda: configs.butane_box(nmol=64, seed=5),
    "water512_ewald_k8": lambda: configs.water_box(n_side=8, L=8 * 3.104, rc=10.0, kappa=0.28,
                                                   kmax=8, ksqmax=64, seed=11),
    "mixture_ewald_ghost": lambda: configs.mixture_box(water4=True, seed=12),
    "mixture_wolf_switch": lambda: configs.mixture_box(elec="wolf", kappa=0.2, switch=True,
                                                       seed=13),
    "bonded_zoo": lambda: configs.bonded_zoo(),
    "tube300_hypergeo": lambda: configs.methane_tube(natom=300, length=60.0, box_xy=40.0,
                                                     tube_r=15.0, rc=10.0),
    "tube300_grid": "grid",
    "tasos_column": lambda: configs.tasos_column(),
    "examples_water": "golden",
    "mixture_water4_vacuum": "golden",
    "water216_coulomb": "golden",
    "tube_cell_myinterp": "golden",
    "tube_explicit": "golden",
    "tube_explicit_switch": "golden",
}

_cache = {}


def _case(name):
    if name not in _cache:
        if CASES[name] == "grid":
            s = tube_grid_system()
        elif CASES[name] == "golden":
            s = System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, name)))
        else:
            s = system_of(CASES[name]())
        o = Oracle(s)
        fl, tl = o.frclong()
        fs, ts = o.frcshort()
        _cache[name] = (s, fl, tl, fs, ts)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_frclong_fp64(name):
    s, fl, tl, _, _ = _case(name)
    print(f"case {name} fp64 natom={s.natom}", flush=True)
    eng = Engine(s, precision="fp64")
    f, t = eng.frclong()
    assert t["n_lj_pairs"] == tl["n_lj_pairs"]
    assert t["n_elec_pairs"] == tl["n_elec_pairs"]
    assert t["lj_pair_hash"] == tl["lj_pair_hash"]
    assert t["elec_pair_hash"] == tl["elec_pair_hash"]
    assert_long_close(t, tl, 1e-10, name)
    ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-10)
    assert ok, err
    eng.close()


@pytest.mark.parametrize("name", list(CASES))
def test_frclong_mixed(name):
    s, fl, tl, _, _ = _case(name)
    print(f"case {name} mixed natom={s.natom}", flush=True)
    eng = Engine(s, precision="mixed")
    f, t = eng.frclong()
    ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-5)
    print(f"  force rel rms {err:.3e}")
    assert ok, err
    assert_long_close(t, tl, 1e-5, name)
    # pair sets may differ only for pairs within FP32 rounding of the cutoff
    assert abs(t["n_lj_pairs"] - tl["n_lj_pairs"]) <= max(2, 1e-5 * tl["n_lj_pairs"])
    assert abs(t["n_elec_pairs"] - tl["n_elec_pairs"]) <= max(2, 1e-5 * tl["n_elec_pairs"])
    eng.close()


@pytest.mark.parametrize("name", ["water216_ewald", "butane64", "water512_ewald_k8"])
def test_frcshort(name):
    """Bonded terms on strained geometries (each site displaced up to
    0.05 A), so that every bond, angle and dihedral term is far from its
    minimum and the energies are not rounding noise."""
    s, _, _, _, _ = _case(name)
    rng = np.random.default_rng(17)
    d = 0.05 / s.params["sigma_base"]
    xx, yy, zz = (v + rng.uniform(-d, d, v.shape) for v in (s.xx, s.yy, s.zz))
    fs, ts = Oracle(s).frcshort(xx, yy, zz)
    eng = Engine(s, precision="mixed")
    f, t = eng.frcshort(xx, yy, zz)
    assert_short_close(t, ts, 1e-12, name)
    assert rel_rms(f, fs) <= 1e-12
    eng.close()


def test_repeat_is_deterministic():
    s, _, _, _, _ = _case("water216_ewald")
    eng = Engine(s, precision="mixed")
    f1, t1 = eng.frclong()
    f2, t2 = eng.frclong()
    # the pair part is atomic-free; the reciprocal force pass adds with FP64
    # atomics, so allow its last-bit reordering
    assert rel_rms(f2, f1) <= 1e-14
    for k in ("uvdw", "ureal", "ufourier", "uinter"):
        assert t1[k] == t2[k]
    eng.close()
