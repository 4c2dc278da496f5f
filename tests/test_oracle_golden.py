"""The oracle (FP64 C restatement) against the reference's own outputs.

tests/golden/*/ref.npz hold the dumps of the reference program built from
its sources (oracle/build_ref.sh) after its start-up path, frclong() and
frcshort().  Compiled the same way (gcc -std=c99 -O2, no contraction, the
same glibc libm), the restatement reproduces them bit for bit: energies,
virials, every force component.  The derived constants of the product's
mspms_derive_params (libmspms_gpu.so, host code) are checked bit for bit too.
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

from mspms2_amd import configs, runio
from mspms2_amd.system import System
from oracle import oracle as O
from tests.util import GOLDEN

SMALL = sorted(d for d in os.listdir(GOLDEN)
               if os.path.isfile(os.path.join(GOLDEN, d, "in.mspms"))
               and os.path.isfile(os.path.join(GOLDEN, d, "ref.npz")))
LONG_KEYS = ("uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uelec", "usflj",
             "uotherff", "uinter", "virial_inter")
SHORT_KEYS = ("ubond", "uangle", "udih", "uimp", "uintra", "virial_intra")
DERIVED = ("rcutoffsq", "rcutoffelecsq", "rcutonsq", "roff2_minus_ron2_cube", "shift1",
           "shift4", "kappa", "coulomb_prefactor", "twopi_lx", "twopi_ly", "twopi_lz",
           "bfactor_ewald", "vfactor_ewald", "twopi_over_3v", "wolfvcon1", "wolfvcon2",
           "wolfvcon3", "wolffcon1", "wolffcon2", "boxlx", "boxly", "boxlz", "f0")


def _golden(name):
    z = np.load(os.path.join(GOLDEN, name, "ref.npz"))
    sc = {k[2:]: float(z[k]) for k in z.files if k.startswith("s_")}
    return sc, z["fl"], z["fs"], z["idx"]


def _system(name):
    return System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, name)))


def test_fixture_set_present():
    assert len(SMALL) >= 10, SMALL


@pytest.mark.parametrize("name", SMALL)
def test_oracle_bit_exact(name):
    sc, fl, fs, idx = _golden(name)
    s = _system(name)
    o = O.Oracle(s)
    if s.params["fEwald_BC"] == 2 and s.params["iChargeType"] == 1:
        pass  # vacuum boundary: covered like the rest
    f1, t1 = o.frclong()
    f2, t2 = o.frcshort()
    for k in LONG_KEYS:
        assert t1[k] == sc[k], (k, t1[k], sc[k])
    for k in SHORT_KEYS:
        assert t2[k] == sc[k], (k, t2[k], sc[k])
    assert np.array_equal(f1[:, idx], fl)
    assert np.array_equal(f2[:, idx], fs)


@pytest.mark.parametrize("name", SMALL)
def test_derived_constants_bit_exact(name):
    """src/initvars.c:433-462 and :591-599, in the oracle and in the product
    (mspms_derive_params of libmspms_gpu.so, host code)."""
    sc, _, _, _ = _golden(name)
    s = _system(name)
    od = O.Oracle(s).derived()
    from mspms2_amd.engine import derive_params
    pd = derive_params(s.c_params())
    for k in DERIVED:
        if k in ("wolfvcon1", "wolfvcon2", "wolfvcon3", "wolffcon1", "wolffcon2") and \
                s.params["iChargeType"] != 2:
            continue
        if k in ("twopi_lx", "twopi_ly", "twopi_lz", "bfactor_ewald", "vfactor_ewald",
                 "twopi_over_3v") and s.params["iChargeType"] != 1:
            continue
        if k == "coulomb_prefactor" and s.params["iChargeType"] == 0:
            continue
        assert od[k] == sc[k], (k, od[k], sc[k])
        assert getattr(pd, k) == sc[k], (k, getattr(pd, k), sc[k])


def test_positions_reduced_like_the_reference():
    """src/readins.c:252-259: the run-directory reader reproduces the
    reference's reduced coordinates exactly (LJ-864: sigma_base 3.405)."""
    with tempfile.TemporaryDirectory() as tmp:
        if not O.reference_available():
            pytest.skip("reference build not present (only where the reference is)")
        d = O.run_reference(os.path.join(GOLDEN, "lj864_jitter"))
        s = _system("lj864_jitter")
        assert np.array_equal(s.xx, d["xx"]) and np.array_equal(s.zz, d["zz"])


@pytest.mark.slow
def test_water12288_regenerates_and_matches():
    """12,288-site Ewald water: the seeded generator reproduces
    the run files the fixture was made from (SHA-256), and the oracle
    matches the reference's energies and sampled forces bit for bit."""
    name = "water12288_ewald"
    sc, fl, fs, idx = _golden(name)
    with open(os.path.join(GOLDEN, name, "runfiles.sha256")) as f:
        want = f.read().strip()
    rd = configs.water_box(n_side=16, L=49.66, kappa=0.28, kmax=8, ksqmax=64, rc=10.0, seed=2)
    with tempfile.TemporaryDirectory() as tmp:
        rd.write(tmp)
        h = hashlib.sha256()
        for fn in ("in.mspms", "cfg.mspms", "xyz.mspms"):
            with open(os.path.join(tmp, fn), "rb") as f:
                h.update(f.read())
        assert h.hexdigest() == want
        s = System.from_rundir(runio.RunDir.read(tmp))
    f1, t1 = O.Oracle(s).frclong()
    for k in LONG_KEYS:
        assert t1[k] == sc[k], (k, t1[k], sc[k])
    assert np.array_equal(f1[:, idx], fl)


def test_reference_matches_fixtures_when_built():
    """Where the reference build exists, it still reproduces the committed
    fixtures (guards against stale fixtures)."""
    if not O.reference_available():
        pytest.skip("reference build not present")
    for name in ("water216_ewald", "bonded_zoo"):
        sc, fl, _, idx = _golden(name)
        d = O.run_reference(os.path.join(GOLDEN, name))
        assert d["uinter"] == sc["uinter"] and np.array_equal(d["fl"][:, idx], fl)
