"""Run-state files (mspms2_amd/rundata.py) against the reference's own:
sav.mspms, trj.mspms and ss.mspms written by its saveit / trajectory /
snapshot after 10 NVT steps of the water (tests/golden/md_water216_nvt/
ref10_*), and a restart from that checkpoint (starting option 2, its loadit
reading old.mspms) continued by its own integrator
(tests/golden/md_water216_nvt_restart)."""
import os

import numpy as np

from mspms2_amd import rundata, runio
from mspms2_amd.system import System
from tests.md_cpu import NvtCpu
from tests.util import GOLDEN

RUN = os.path.join(GOLDEN, "md_water216_nvt")
RESTART = os.path.join(GOLDEN, "md_water216_nvt_restart")


def _raw(path):
    with open(path, "rb") as f:
        return f.read()


def test_checkpoint_reads_the_reference_state_and_writes_its_bytes(tmp_path):
    g = np.load(os.path.join(RUN, "md10.npz"))
    path = os.path.join(RUN, "ref10_sav.mspms")
    ck = rundata.read_checkpoint(path, int(g["natom"]))
    assert np.array_equal(ck.x, g["x"]) and np.array_equal(ck.v, g["v"])
    for k in ("ps", "pss", "ss", "sss", "qq", "qqs", "gg", "ggs"):
        assert getattr(ck, k) == float(g[f"md_{k}"]), k
    assert np.array_equal(ck.box, g["box"])
    out = str(tmp_path / "sav.mspms")
    rundata.write_checkpoint(out, ck)
    assert _raw(out) == _raw(path)


def test_trajectory_frame_and_bytes(tmp_path):
    g = np.load(os.path.join(RUN, "md10.npz"))
    rd = runio.RunDir.read(RUN)
    s = System.from_rundir(rd)
    path = os.path.join(RUN, "ref10_trj.mspms")
    h, frames = rundata.read_trajectory(path)
    assert frames.shape == (1, 3, s.natom) and np.array_equal(frames[0], g["x"])
    mine = rundata.trajectory_header(rd, s)
    assert mine == h
    out = str(tmp_path / "trj.mspms")
    rundata.TrajectoryWriter(out, mine).append(g["x"])
    assert _raw(out) == _raw(path)


def test_snapshot_text(tmp_path):
    g = np.load(os.path.join(RUN, "md10.npz"))
    rd = runio.RunDir.read(RUN)
    path = os.path.join(RUN, "ref10_ss.mspms")
    istep, box, names, x = rundata.read_snapshot(path)
    assert names == rd.atom_names() and np.allclose(x, g["x"], rtol=0, atol=5e-7)
    out = str(tmp_path / "ss.mspms")
    rundata.write_snapshot(out, istep, g["box"], rd.atom_names(), g["x"])
    assert _raw(out) == _raw(path)


def test_restart_continues_the_reference_run():
    """The restart run directory loads old.mspms as the reference's loadit
    does, and the NVT restatement continued from it (thermostats included)
    reproduces the reference's own continuation bit for bit."""
    rd = runio.RunDir.read(RESTART)
    assert rd.checkpoint is not None
    s = System.from_rundir(rd)
    ck = rd.checkpoint
    assert np.array_equal(np.stack([s.xx, s.yy, s.zz]), ck.x) and np.array_equal(s.v0, ck.v)
    g = np.load(os.path.join(RESTART, "md5.npz"))
    assert np.array_equal(g["v0"], ck.v)
    c = NvtCpu(s, s.v0, ck.qq, ck.qqs, rd.run.nconstraint,
               state=dict(gg=ck.gg, ggs=ck.ggs, ss=ck.ss, ps=ck.ps, sss=ck.sss, pss=ck.pss))
    for _ in range(5):
        c.step()
    assert np.array_equal(c.x, g["x"]) and np.array_equal(c.v, g["v"])
    assert c.ps == float(g["md_ps"]) and c.pss == float(g["md_pss"])


def test_rundir_round_trip_keeps_the_checkpoint(tmp_path):
    rd = runio.RunDir.read(RESTART)
    rd.write(str(tmp_path))
    assert _raw(str(tmp_path / "old.mspms")) == _raw(os.path.join(RESTART, "old.mspms"))
    back = runio.RunDir.read(str(tmp_path))
    assert np.array_equal(back.checkpoint.x, rd.checkpoint.x)
