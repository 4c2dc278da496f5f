"""The tube workload's restatements (tests/tube_ref.py) against the oracle,
on systems the oracle's all-pairs frclong finishes in seconds: the LJ pair
set (count and splitmix64 hash) bit for bit, uvdw / virial / forces to
1e-12; the wall's energy and forces bit for bit."""
import os

import numpy as np
import pytest

from mspms2_amd import configs, runio
from mspms2_amd.system import System, system_of
from oracle.oracle import Oracle
from tests.tube_ref import lj_pairs, wall
from tests.util import GOLDEN, tube_grid_system


@pytest.mark.parametrize("case", ["lj864_jitter", "tube2000"])
def test_lj_pairs_equal_oracle(case):
    if case == "lj864_jitter":
        s = System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, case)))
    else:
        s = system_of(configs.tube_loading(2000, length=200.0))
        s.params["iSF_type"] = 0
    f, t = Oracle(s).frclong()
    n, h, u, vir, F = lj_pairs(s)
    assert n == t["n_lj_pairs"] and h == t["lj_pair_hash"]
    assert u == pytest.approx(t["uvdw"], rel=1e-12)
    assert vir == pytest.approx(t["virial_inter"], rel=1e-12)
    assert np.linalg.norm(F - f) <= 1e-12 * np.linalg.norm(f)


def test_wall_equals_oracle():
    s = tube_grid_system()
    f, t = Oracle(s).frclong()
    s0 = System(species=s.species, params=dict(s.params, iSF_type=0), xx=s.xx, yy=s.yy,
                zz=s.zz, run=s.run)
    f0, _ = Oracle(s0).frclong()
    u, F = wall(s)
    assert u == t["usflj"]
    assert np.array_equal(F, f - f0) or np.linalg.norm(F - (f - f0)) <= 1e-14 * np.linalg.norm(F)
