"""The term-level drop-in symbols inside the reference's own program.

oracle/_ref/mspms_terms_harness is the reference's start-up path, frclong()
and frcshort() (src/frclong.c, src/frcshort.c and every caller, compiled
from their sources by oracle/build_ref.sh) with the per-term functions
replaced by the ones in mspms2_amd/csrc/dropin_terms.c: ljfrc,
ewald_real_frc, ewald_excl_frc, wolf_real_frc, xatom_coulomb_frc,
xmole_coulomb_frc, the bond / angle / dihedral / improper terms,
cal_sf_hypergeo, cal_sf_atom_explicit and get_values_from_grid, all over
libmspms_gpu.so's host term code.  Its dump must equal the golden dump of
the unmodified reference bit for bit: every energy, the virials and every
force component kept in tests/golden/*/ref.npz.

These runs are host code only (no GPU call); the test is skipped where the
harness was not built (no reference sources at build time).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.oracle import read_ref_dump
from tests.util import GOLDEN

HARNESS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "_ref", "mspms_terms_harness")

# one case per term family the symbols cover
CASES = ["water216_ewald", "water216_coulomb", "water216_wolf", "mixture_wolf_switch",
         "mixture_water4_vacuum", "tube300_hypergeo", "tube_explicit", "tube_explicit_switch",
         "tube_cell_myinterp", "bonded_zoo", "butane64", "examples_water"]
KEYS = ("uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uelec", "usflj", "uotherff",
        "uinter", "virial_inter", "ubond", "uangle", "udih", "uimp", "uintra", "virial_intra")

needs_harness = pytest.mark.skipif(not os.path.exists(HARNESS),
                                   reason="term harness not built (no reference sources)")


def _run(name):
    src = os.path.join(GOLDEN, name)
    with tempfile.TemporaryDirectory(prefix="mspms_terms_") as tmp:
        for fn in os.listdir(src):
            if fn.endswith(".npz"):
                continue
            with open(os.path.join(src, fn), "rb") as a, open(os.path.join(tmp, fn), "wb") as b:
                b.write(a.read())
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([HARNESS, tmp, out], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return read_ref_dump(out)


@needs_harness
def test_term_symbols_are_the_engines():
    """The harness resolves the term functions to dropin_terms.c, which
    calls into libmspms_gpu.so (mspms_term_*), not to the reference's."""
    r = subprocess.run(["nm", HARNESS], capture_output=True, text=True, check=True)
    undefined = {ln.split()[-1] for ln in r.stdout.splitlines() if ln.split()[0] == "U"}
    for sym in ("mspms_term_pair", "mspms_term_bonded", "mspms_term_sf", "mspms_term_myinterp"):
        assert sym in undefined, sym


@needs_harness
@pytest.mark.parametrize("name", [c for c in CASES if os.path.isdir(os.path.join(GOLDEN, c))])
def test_reference_program_with_engine_terms_bit_exact(name):
    z = np.load(os.path.join(GOLDEN, name, "ref.npz"))
    idx = z["idx"]
    d = _run(name)
    assert d["natom"] == int(z["natom"])
    for k in KEYS:
        assert d[k] == float(z[f"s_{k}"]), (name, k, d[k], float(z[f"s_{k}"]))
    assert np.array_equal(d["fl"][:, idx], z["fl"]), name
    assert np.array_equal(d["fs"][:, idx], z["fs"]), name
