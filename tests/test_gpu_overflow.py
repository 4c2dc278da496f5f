"""Pair-list overflow inside a device RESPA run: every step saves its start
while no overflow is flagged, and mspms_sync puts the run's positions and
velocities back to the start of the first step that overflowed (here the
first step: the run's initial state, bit for bit), regrows the lists and
reports MSPMS_ENOMEM naming the step.  Forced by bounding the row capacity
(mspms_set_list_capacity) far below what 216 waters need; the repeat of the
steps after the regrowth equals the same steps on an engine that never
overflowed (FP64, to the last bits of the forces' summation order)."""
import numpy as np
import pytest

from mspms2_amd import configs
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine, EngineError
from mspms2_amd.system import system_of

pytestmark = pytest.mark.gpu


def _state(s, seed=5):
    rng = np.random.default_rng(seed)
    m = s.masses()
    v = rng.standard_normal((3, s.natom)) * np.sqrt(s.treq / m)
    return v - (v * m).sum(axis=1, keepdims=True) / m.sum()


def _buffers(eng, s, dev):
    import torch
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    pos, vel = t(np.stack([s.xx, s.yy, s.zz])), t(_state(s))
    fl, fs = t(eng.frclong()[0]), t(eng.frcshort()[0])
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    return pos, vel, fl, fs, lt, st, uk


@pytest.mark.parametrize("graph", [True, False])
def test_overflow_restores_run_start_and_repeat_matches(graph):
    import torch
    dev = torch.device("cuda:0")
    s = system_of(configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3, kmax=5,
                                    ksqmax=27, seed=3, nstep_inner=4))
    nin = s.run.nstep_inner
    # the reference run: 3 steps that never overflow
    ref = Engine(s, precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _buffers(ref, s, dev)
    ref.vver_run_device(pos, vel, fl, fs, s.delt, nin, 3, lt, st, uk, graph=graph)
    torch.cuda.synchronize()
    ref.sync()
    want_x, want_v = pos.cpu().numpy(), vel.cpu().numpy()
    ref.close()
    # the same run with rows bounded to 16 entries
    eng = Engine(s, precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _buffers(eng, s, dev)
    x0, v0 = pos.cpu().numpy(), vel.cpu().numpy()
    eng.set_list_capacity(16)
    eng.vver_run_device(pos, vel, fl, fs, s.delt, nin, 3, lt, st, uk, graph=graph)
    torch.cuda.synchronize()
    with pytest.raises(EngineError) as e:
        eng.sync()
    print(e.value, flush=True)
    assert "ENOMEM" in str(e.value) and "restored" in str(e.value) and "step 1" in str(e.value)
    assert np.array_equal(pos.cpu().numpy(), x0) and np.array_equal(vel.cpu().numpy(), v0)
    # forces of the restored state, then the steps again: the lists fit now
    eng.frclong_device(pos, fl, lt)
    eng.frcshort_device(pos, fs, st)
    eng.vver_run_device(pos, vel, fl, fs, s.delt, nin, 3, lt, st, uk, graph=graph)
    torch.cuda.synchronize()
    eng.sync()
    x, v = pos.cpu().numpy(), vel.cpu().numpy()
    assert np.abs(x - want_x).max() <= 1e-12 * np.abs(want_x).max()
    assert np.abs(v - want_v).max() <= 1e-12 * np.abs(want_v).max()
    eng.close()
