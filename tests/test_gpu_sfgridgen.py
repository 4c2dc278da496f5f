"""Solid-fluid grid generation on the GPU (mspms_sfgridgen) against the
reference's own generator, tools/sfgridgen.c: bit for bit against its
committed output (tests/golden/sfgridgen_ch3_1A) and against a live run of
it (oracle/_ref/sfgridgen, built from the reference's source) at 0.5 A."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.util import GOLDEN

pytestmark = pytest.mark.gpu

CASE = os.path.join(GOLDEN, "sfgridgen_ch3_1A")
REF_EXE = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "sfgridgen")


def test_device_tables_equal_reference_output():
    from mspms2_amd import sfgridgen as S
    g = S.read_input(os.path.join(CASE, "grid.in"))
    _, ref = S.read_output(os.path.join(CASE, "out.sfgridgen"))
    tab = S.generate(g)
    assert tab.shape == ref.shape
    assert np.array_equal(tab, ref)


def test_device_tables_equal_live_reference_run():
    from mspms2_amd import sfgridgen as S
    exe = os.path.abspath(REF_EXE)
    if not os.path.exists(exe):
        pytest.skip("reference generator not built (oracle/_ref/sfgridgen)")
    with tempfile.TemporaryDirectory() as tmp:
        src = open(os.path.join(CASE, "grid.in")).read().split("\n")
        src[1] = " ".join(["0.5"] + src[1].split()[1:])
        inp = os.path.join(tmp, "grid.in")
        open(inp, "w").write("\n".join(src))
        subprocess.run([exe, inp], cwd=tmp, check=True, capture_output=True, timeout=120)
        hdr, ref = S.read_output(os.path.join(tmp, "out.sfgridgen"))
        g = S.read_input(inp)
    tab = S.generate(g)
    print(f"nodes {hdr['ntot']}")
    assert np.array_equal(tab, ref)


def test_switch_with_rcuton_at_rcutoff_changes_nothing():
    """With the switching shell empty (rcuton = rcutoff) the switch branch
    gives the unswitched tables bit for bit."""
    from mspms2_amd import sfgridgen as S
    g = S.read_input(os.path.join(CASE, "grid.in"))
    a = S.generate(g)
    g.isLJswitch = 1
    b = S.generate(g)
    assert np.array_equal(a, b)
