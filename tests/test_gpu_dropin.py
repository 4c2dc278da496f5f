"""Drop-in integration on the GPU: the reference's own start-up path
(readins, init_vars: src/mspms2.c:39-49) linked with the engine's frclong /
frcshort (libmspms_dropin.a + libmspms_gpu.so) in place of src/frclong.c
and src/frcshort.c, on the golden run directories.  The dump must equal the
reference's (tests/golden/*/ref.npz) to 1e-10 in FP64 mode.

The harness binary is built from the reference sources by
oracle/build_ref.sh (here); it travels to the GPU box prebuilt.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests.util import GOLDEN, assert_long_close, assert_short_close, long_energy_floor

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(os.path.dirname(O.__file__), "_ref", "mspms_dropin_harness")

CASES = ["lj864_jitter", "water216_ewald", "water216_wolf", "butane64", "mixture_ewald",
         "mixture_wolf_switch", "bonded_zoo", "examples_water", "mixture_water4_vacuum",
         "water216_coulomb", "tube_cell_myinterp", "tube300_hypergeo", "tube_explicit",
         "tube_explicit_switch"]


def _golden(name):
    z = np.load(os.path.join(GOLDEN, name, "ref.npz"))
    sc = {k[2:]: float(z[k]) for k in z.files if k.startswith("s_")}
    return sc, z["fl"], z["fs"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dropin(name, precision):
    if not os.path.exists(HARNESS):
        pytest.fail(f"{HARNESS} missing: run __graft_entry__.build() where the reference is")
    print(f"dropin {name} {precision}", flush=True)
    sc, fl, fs = _golden(name)
    with tempfile.TemporaryDirectory(prefix="mspms_dropin_") as tmp:
        for fn in os.listdir(os.path.join(GOLDEN, name)):
            if fn.endswith(".mspms") or fn.endswith(".cubes"):
                shutil.copy(os.path.join(GOLDEN, name, fn), tmp)
        env = dict(os.environ, MSPMS_GPU_PRECISION=precision)
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([HARNESS, tmp, out], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        d = O.read_ref_dump(out)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert_long_close(d, sc, tol, name)
    assert_short_close(d, sc, 1e-10, name, floor=long_energy_floor(sc))
    fscale = max(np.linalg.norm(fl), (abs(sc["virial_inter"]) + abs(sc["uvdw"])) /
                 (d["natom"] * sc["rcutoff"]) * np.sqrt(fl.size))
    err = np.linalg.norm(d["fl"] - fl) / fscale
    print(f"  force rel {err:.2e}")
    assert err <= tol
    # frcshort forces: relative to the nonbonded force scale as well, since
    # rigid-template inputs put every bonded term at its minimum
    assert np.linalg.norm(d["fs"] - fs) <= 1e-10 * max(np.linalg.norm(fs), fscale)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dropin_tasos_grid_files(precision):
    """iSF_type 3 through the drop-in: the drop-in reads the TASOS section
    and its grid files itself (the Fortran library's copy is out of reach;
    the harness skips init_tasos_grid with --tasos-linked).  Against the
    oracle on the grid the Python reader loads: FP64 1e-10, mixed 1e-5."""
    from mspms2_amd import configs, runio
    from mspms2_amd.system import System
    if not os.path.exists(HARNESS):
        pytest.fail(f"{HARNESS} missing: run __graft_entry__.build() where the reference is")
    print(f"dropin tasos_column {precision}", flush=True)
    rd = configs.tasos_column()
    with tempfile.TemporaryDirectory(prefix="mspms_dropin_") as tmp:
        rd.write(tmp)
        s = System.from_rundir(runio.RunDir.read(tmp))
        fl, tl = O.Oracle(s).frclong()
        env = dict(os.environ, MSPMS_GPU_PRECISION=precision)
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([HARNESS, tmp, out, "--tasos-linked"], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        d = O.read_ref_dump(out)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert d["usflj"] == pytest.approx(tl["usflj"], rel=1e-10)
    assert_long_close(d, tl, tol, "tasos_column")
    err = np.linalg.norm(d["fl"] - fl) / np.linalg.norm(fl)
    print(f"  force rel {err:.2e}")
    assert err <= tol
