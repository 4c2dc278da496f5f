"""Batched replicas and graph-replayed RESPA steps.

A replica table packs independent boxes of uneven size (own box, cutoff,
k-vectors) into one engine; every kernel covers all replicas in one launch.
Per replica, the result must equal what the oracle gives for that box alone:
in FP64 mode pair counts and pair hashes exactly and every energy / force to
1e-10; in mixed mode forces to 1e-5.

The captured RESPA step (mspms_vver_run_device with a graph) is replayed
several times, with eager work and synchronisations in between, and each
replay is checked against the eager step from the same state.
"""
import numpy as np
import pytest

from mspms2_amd import configs
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
from oracle.oracle import Oracle
from tests.util import (assert_long_close, assert_short_close, force_close,
                        long_energy_floor, rel_rms, system_of)

pytestmark = pytest.mark.gpu

REPLICA_SETS = {
    # n-butane boxes of 64..210 molecules (rc = min(14 A, L/2) per box)
    "butane4": lambda: [configs.butane_box(nmol=n, seed=40 + k)
                        for k, n in enumerate((64, 150, 97, 210))],
    # Ewald water boxes of 125..343 molecules, one k-vector set per box
    "water3_ewald": lambda: [configs.water_box(n_side=ns, L=ns * 3.104, rc=7.5, kappa=0.3,
                                               kmax=5, ksqmax=27, seed=50 + ns)
                             for ns in (5, 7, 6)],
    "water3_wolf": lambda: [configs.water_box(n_side=ns, L=ns * 3.104, rc=7.5, kappa=0.2,
                                              elec="wolf", seed=60 + ns)
                            for ns in (6, 5, 7)],
}

_cache = {}


def _set(name):
    if name not in _cache:
        systems = [system_of(rd) for rd in REPLICA_SETS[name]()]
        refs = []
        for s in systems:
            o = Oracle(s)
            refs.append((o.frclong(), o.frcshort()))
        _cache[name] = (systems, refs)
    return _cache[name]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(REPLICA_SETS))
def test_replicas_match_oracle(name, precision):
    systems, refs = _set(name)
    print(f"case {name} {precision} natom={[s.natom for s in systems]}", flush=True)
    eng = Engine.replicas(systems, precision=precision)
    assert eng.natom == sum(s.natom for s in systems)
    f, terms = eng.frclong()
    fs, sterms = eng.frcshort()
    tol = 1e-10 if precision == "fp64" else 1e-5
    for r, (s, ((fl_ref, tl), (fs_ref, ts))) in enumerate(zip(systems, refs)):
        fr = eng.unpack(f)[r]
        if precision == "fp64":
            for k in ("n_lj_pairs", "n_elec_pairs", "lj_pair_hash", "elec_pair_hash"):
                assert terms[r][k] == tl[k], (r, k)
        ok, err = force_close(fr, fl_ref, tl, s.natom, s.params["rcutoff"], tol)
        assert ok, (r, err)
        assert_long_close(terms[r], tl, tol, f"{name}[{r}]")
        # rigid-template inputs: bonded terms at their minima, compare on the
        # nonbonded scales
        ok, err = force_close(eng.unpack(fs)[r], fs_ref, tl, s.natom, s.params["rcutoff"],
                              1e-12)
        assert ok, (r, err)
        assert_short_close(sterms[r], ts, 1e-12, f"{name}[{r}]", floor=long_energy_floor(tl))
    eng.close()


def _state(system, seed):
    rng = np.random.default_rng(seed)
    m = system.masses()
    v = rng.standard_normal((3, system.natom)) * np.sqrt(system.treq / m)
    return v - (v * m).sum(axis=1, keepdims=True) / m.sum()


def test_graph_replay_matches_eager():
    import torch
    dev = torch.device("cuda:0")
    s = system_of(configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3, kmax=5,
                                    ksqmax=27, seed=3, nstep_inner=4))
    eng = Engine(s, precision="mixed")
    n = s.natom
    pos0 = torch.tensor(np.stack([s.xx, s.yy, s.zz]), dtype=torch.float64, device=dev)
    vel0 = torch.tensor(_state(s, 5), dtype=torch.float64, device=dev)
    fl0 = torch.tensor(eng.frclong()[0], device=dev)
    fs0 = torch.tensor(eng.frcshort()[0], device=dev)

    def bufs():
        return dict(pos=pos0.clone(), vel=vel0.clone(), fl=fl0.clone(), fs=fs0.clone(),
                    lt=torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev),
                    st=torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev),
                    uk=torch.zeros(1, dtype=torch.float64, device=dev))

    nsteps = 3
    e = bufs()
    torch.cuda.synchronize()  # the engine runs on its own stream
    eng.vver_run_device(e["pos"], e["vel"], e["fl"], e["fs"], s.delt, 4, nsteps, e["lt"],
                        e["st"], e["uk"], graph=False)
    torch.cuda.synchronize()
    eng.sync()
    g = bufs()
    for rep in range(3):
        # reset the graph's buffers in place, eagerly, then replay
        for k, v0 in (("pos", pos0), ("vel", vel0), ("fl", fl0), ("fs", fs0)):
            g[k].copy_(v0)
        torch.cuda.synchronize()
        eng.vver_run_device(g["pos"], g["vel"], g["fl"], g["fs"], s.delt, 4, nsteps, g["lt"],
                            g["st"], g["uk"], graph=True)
        torch.cuda.synchronize()
        eng.sync()
        print(f"replay {rep}", flush=True)
        # last-bit differences come only from the FP64 atomics of the
        # reciprocal force pass
        for k in ("pos", "vel", "fl", "fs"):
            assert rel_rms(g[k].cpu().numpy(), e[k].cpu().numpy()) <= 1e-12, (rep, k)
        for k in ("lt", "st", "uk"):
            a, b = g[k].cpu().numpy(), e[k].cpu().numpy()
            assert np.allclose(a, b, rtol=1e-11, atol=0), (rep, k)
    eng.close()
