"""GPU parity at the sizes of the batched workloads.

n-butane replica table: 256 boxes of U{64..512} united-atom butanes
(277,824 sites) in one table.  frclong of the whole table, with every
replica of a sample (the smallest, the largest and four more) checked against
the oracle run on that box alone: FP64 pair counts and hashes equal, energies
and forces within 1e-10; mixed forces within 1e-5.  Then two HMC moves of
10 RESPA steps (1 fs, 5 inner steps) of every chain, FP64, with three
chains checked against the HMC restatement that is pinned to the
reference's own hmc(): accept decisions equal, dH within 1e-7 (relative,
floor 1e-3), positions within 1e-8.

Wolf water fitting set: 64 configurations of 1024 flexible waters (kappa
0.2 / A, rc 12 A) as one table.  Energies and forces of every configuration
against the oracle (1e-10), and the per-type parameter derivatives
dE/d(eps_O, sigma_O, q_O, q_H) of two configurations against central
differences of the oracle's energy (h = 1e-6 relative; bar 1e-6 relative to
max(|FD|, 1e-2 x energy scale / |p|), the FD rounding of cancelling terms).
"""
import copy

import numpy as np
import pytest

from mspms2_amd import configs
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
from mspms2_amd.system import system_of
from oracle.oracle import Oracle
from tests.hmc_cpu import HmcChain
from tests.util import assert_long_close, energy_scale, force_close

pytestmark = pytest.mark.gpu

_cache = {}


def _butanes():
    if "b" not in _cache:
        _cache["b"] = [system_of(rd) for rd in configs.butane_table()]
    return _cache["b"]


def _sample(systems):
    n = [s.natom for s in systems]
    return sorted({int(np.argmin(n)), int(np.argmax(n)), 0, 17, 100, len(systems) - 1})


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_butane_table_frclong_vs_oracle(prec):
    systems = _butanes()
    print(f"{len(systems)} boxes, {sum(s.natom for s in systems)} sites, {prec}", flush=True)
    eng = Engine.replicas(systems, precision=prec)
    f, terms = eng.frclong()
    parts = eng.unpack(f)
    for r in _sample(systems):
        s = systems[r]
        fl, tl = Oracle(s).frclong()
        t = terms[r]
        what = f"box {r} ({s.natom} sites) {prec}"
        if prec == "fp64":
            for key in ("n_lj_pairs", "lj_pair_hash"):
                assert t[key] == tl[key], (what, key)
            assert_long_close(t, tl, 1e-10, what)
            ok, err = force_close(parts[r], fl, tl, s.natom, s.params["rcutoff"], 1e-10)
        else:
            assert_long_close(t, tl, 1e-5, what)
            ok, err = force_close(parts[r], fl, tl, s.natom, s.params["rcutoff"], 1e-5)
        print(f"  {what}: force err {err:.2e}", flush=True)
        assert ok, (what, err)
    eng.close()


def test_butane_table_hmc_moves_vs_restatement():
    import torch
    dev = torch.device("cuda:0")
    systems = _butanes()
    nr = len(systems)
    eng = Engine.replicas(systems, precision="fp64")
    pos = torch.tensor(eng.packed_positions(), device=dev)
    vel = torch.zeros_like(pos)
    fl, fs = torch.zeros_like(pos), torch.zeros_like(pos)
    lt = torch.zeros(nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(nr * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(nr, dtype=torch.float64, device=dev)
    ij = [1000 + r for r in range(nr)]
    kl = [2000 + r for r in range(nr)]
    nc = [s.run.nconstraint for s in systems]
    eng.hmc_init(ij, kl, [s.treq for s in systems], nc, vel)
    inner, nmd, delt = systems[0].run.nstep_inner, 10, systems[0].delt
    nmove = 2
    acc_dev, dh_dev = [], []
    prev = np.zeros(nr, dtype=np.int64)
    for m in range(nmove):
        eng.hmc_run_device(pos, vel, fl, fs, delt, inner, nmd, 1, lt, st, uk, graph=True)
        torch.cuda.synchronize()
        acc, mov, dh = eng.hmc_stats()
        assert (mov == m + 1).all()
        acc_dev.append(acc - prev)
        dh_dev.append(dh.copy())
        prev = acc.copy()
    eng.sync()
    x = eng.unpack(pos.cpu().numpy())
    n = [s.natom for s in systems]
    for r in (int(np.argmin(n)), 100, int(np.argmax(n))):
        ch = HmcChain(systems[r], ij[r], kl[r], nc[r])
        for m in range(nmove):
            a, dh = ch.move(delt, inner, nmd)
            print(f"box {r} move {m}: cpu accept={a} dH={dh:.6e} | gpu accept="
                  f"{bool(acc_dev[m][r])} dH={dh_dev[m][r]:.6e}", flush=True)
            assert bool(acc_dev[m][r]) == a
            assert abs(dh_dev[m][r] - dh) <= 1e-7 * max(abs(dh), 1e-3)
        assert np.abs(x[r] - ch.x).max() <= 1e-8 * systems[r].params["boxlx"]
    eng.close()


def _wolf_set():
    if "w" not in _cache:
        _cache["w"] = [system_of(rd) for rd in configs.wolf_water_set()]
    return _cache["w"]


def test_wolf_set_forward_every_configuration():
    systems = _wolf_set()
    eng = Engine.replicas(systems, precision="fp64")
    f, terms, g = eng.param_grad()
    parts = eng.unpack(f)
    for r, s in enumerate(systems):
        fl, tl = Oracle(s).frclong()
        assert_long_close(terms[r], tl, 1e-10, f"configuration {r}")
        ok, err = force_close(parts[r], fl, tl, s.natom, s.params["rcutoff"], 1e-10)
        assert ok, (r, err)
    eng.close()


def _type_energy(system, p):
    """uinter with the water types (eps_O, sigma_O, q_O, q_H) set to p"""
    c = copy.deepcopy(system)
    w = c.species[0]
    w.epsilon[0], w.sigma[0], w.charge[0] = p[0], p[1], p[2]
    w.charge[1] = w.charge[2] = p[3]
    return Oracle(c).frclong()[1]["uinter"]


def test_wolf_set_type_derivatives_vs_central_differences():
    systems = _wolf_set()
    eng = Engine.replicas(systems, precision="fp64")
    _, terms, g = eng.param_grad()
    eng.close()
    for r in (0, len(systems) - 1):
        s = systems[r]
        w = s.species[0]
        p0 = np.array([w.epsilon[0], w.sigma[0], w.charge[0], w.charge[1]])
        # type derivative = sum of the derivatives of the sites of the type
        ga = np.array([g["epsilon"][r, 0], g["sigma"][r, 0], g["charge"][r, 0],
                       g["charge"][r, 1] + g["charge"][r, 2]])
        scale = energy_scale(terms[r]) + abs(terms[r]["uvdw"])
        for i in range(4):
            h = 1e-6 * abs(p0[i])
            pp, pm = p0.copy(), p0.copy()
            pp[i] += h
            pm[i] -= h
            gfd = (_type_energy(s, pp) - _type_energy(s, pm)) / (2 * h)
            floor = 1e-2 * scale / abs(p0[i])
            print(f"configuration {r} type parameter {i}: adjoint {ga[i]:.10e} FD {gfd:.10e}",
                  flush=True)
            assert abs(ga[i] - gfd) <= 1e-6 * max(abs(gfd), floor), (r, i, ga[i], gfd)
