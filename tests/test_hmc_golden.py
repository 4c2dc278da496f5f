"""The HMC restatement (tests/hmc_cpu.py) against the reference's own hmc()
(tests/golden/hmc_butane_*: oracle/ref_harness.c --hmc over the reference's
sources, src/hmc.c:23-70, src/dispmd.c:12-103).

Per move: the accept / reject decision identical, dH within 1e-9 (relative,
floor 1e-9: dH is a difference of energies of ~1e4 K divided by T, so
summation order moves it by ~1e-12), delt after the move bit for bit (the
adaptation is exact arithmetic on the accept counts); after the last move
the positions within 1e-9 of the box.  hmc_butane_eq runs 20 moves of the
equilibration phase with the delt adaptation every 5 moves,
hmc_butane_prod 12 production moves with delt fixed.
"""
import os

import numpy as np
import pytest

from mspms2_amd import runio
from mspms2_amd.system import System
from tests.hmc_cpu import HmcChain
from tests.util import GOLDEN

CASES = ["hmc_butane_eq", "hmc_butane_prod"]


def load(name):
    s = System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, name)))
    z = np.load(os.path.join(GOLDEN, name, "hmc.npz"))
    return s, {k: z[k] for k in z.files}


def adapt_of(run):
    """(cycle, rreq_disp) while the input starts in its equilibration phase
    (bEquilibrium = nstep_eq > 0, src/readins.c:66), else None."""
    return (run.nstep_delt_adj_cycle, run.rreq_disp) if run.nstep_eq > 0 else None


@pytest.mark.slow
@pytest.mark.parametrize("name", CASES)
def test_restatement_matches_reference_hmc(name):
    s, ref = load(name)
    r = s.run
    chain = HmcChain(s, r.ij, r.jk, r.nconstraint)
    assert np.abs(chain.v - ref["v0"]).max() <= 1e-13 * np.abs(ref["v0"]).max()
    nmove = len(ref["accept"])
    out = chain.run(nmove, s.delt, r.nstep_inner, r.nstep_md_per_hmc, adapt_of(r))
    acc = np.array([o[0] for o in out], dtype=np.int8)
    dh = np.array([o[1] for o in out])
    delt = np.array([o[2] for o in out])
    print(name, "".join(map(str, acc)), "".join(map(str, ref["accept"])))
    assert np.array_equal(acc, ref["accept"])
    assert np.all(np.abs(dh - ref["dH"]) <= 1e-9 * np.maximum(np.abs(ref["dH"]), 1.0))
    assert np.array_equal(delt, ref["delt"])
    L = s.params["boxlx"]
    assert np.abs(chain.x - ref["x"]).max() <= 1e-9 * L
    # the adaptation really ran in the equilibration case
    if adapt_of(r):
        assert len(set(ref["delt"].tolist())) > 1
    else:
        assert np.all(ref["delt"] == s.delt)
