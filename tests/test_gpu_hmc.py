"""Batched HMC on the device against the CPU restatement of the reference's
chain (tests/hmc_cpu.py: hmc / md_move / vver / velinit / gaussran over the
oracle's forces and RANMAR stream).

* The device RANMAR streams are the oracle's bit for bit (RANMAR values are
  multiples of 2^-24; the oracle's stream is pinned by the generator's
  known-answer test in test_closed_forms.py).
* Initial velocities agree to 1e-13 (gaussran's log / sqrt are the only
  non-exact operations).
* FP64 chains: per move the accept / reject decision is identical and dH
  agrees to 1e-7 relative (the trajectories differ only by summation order).
"""
import numpy as np
import pytest

from mspms2_amd import configs
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
from tests.hmc_cpu import HmcChain
from tests.util import system_of

pytestmark = pytest.mark.gpu


def _boxes():
    return [system_of(configs.butane_box(nmol=n, seed=70 + k)) for k, n in enumerate((64, 80))]


def _device_state(eng, systems, dev):
    import torch
    nr = len(systems)
    pos = torch.tensor(eng.packed_positions(), device=dev)
    vel = torch.zeros_like(pos)
    fl, fs = torch.zeros_like(pos), torch.zeros_like(pos)
    lt = torch.zeros(nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(nr * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(nr, dtype=torch.float64, device=dev)
    return pos, vel, fl, fs, lt, st, uk


def test_streams_and_initial_velocities():
    import torch
    dev = torch.device("cuda:0")
    systems = _boxes()
    eng = Engine.replicas(systems, precision="fp64")
    pos, vel, *_ = _device_state(eng, systems, dev)
    ij = [1000 + r for r in range(len(systems))]
    kl = [2000 + r for r in range(len(systems))]
    eng.hmc_init(ij, kl, [s.treq for s in systems], 0, vel)
    torch.cuda.synchronize()
    v = vel.cpu().numpy()
    for r, s in enumerate(systems):
        chain = HmcChain(s, ij[r], kl[r])
        vr = eng.unpack(v)[r]
        err = np.abs(vr - chain.v).max() / np.abs(chain.v).max()
        print(f"replica {r}: {chain.g.used} uniforms for velinit, velocity err {err:.2e}",
              flush=True)
        assert err <= 1e-13
        # the streams continue identically (raw uniforms, bit for bit)
        got = eng.hmc_draw(r, 257)
        want = chain.g.rng.next(257)
        assert np.array_equal(got, want)
    eng.close()


def test_fp64_chains_match_cpu():
    import torch
    dev = torch.device("cuda:0")
    systems = _boxes()
    nmove, nmd = 4, 4
    inner = systems[0].params.get("nstep_inner", 5) or 5
    delt = systems[0].delt
    eng = Engine.replicas(systems, precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _device_state(eng, systems, dev)
    ij = [1000 + r for r in range(len(systems))]
    kl = [2000 + r for r in range(len(systems))]
    eng.hmc_init(ij, kl, [s.treq for s in systems], 0, vel)
    chains = [HmcChain(s, ij[r], kl[r]) for r, s in enumerate(systems)]
    acc_dev = []
    dh_dev = []
    prev = np.zeros(len(systems), dtype=np.int64)
    for m in range(nmove):
        eng.hmc_run_device(pos, vel, fl, fs, delt, inner, nmd, 1, lt, st, uk, graph=True)
        torch.cuda.synchronize()
        acc, mov, dh = eng.hmc_stats()
        assert (mov == m + 1).all()
        acc_dev.append(acc - prev)
        dh_dev.append(dh.copy())
        prev = acc
    eng.sync()
    for r, ch in enumerate(chains):
        for m in range(nmove):
            a, dh = ch.move(delt, inner, nmd)
            print(f"replica {r} move {m}: cpu accept={a} dH={dh:.6e} | gpu accept="
                  f"{bool(acc_dev[m][r])} dH={dh_dev[m][r]:.6e}", flush=True)
            assert bool(acc_dev[m][r]) == a
            assert abs(dh_dev[m][r] - dh) <= 1e-7 * max(abs(dh), 1e-3)
    # the positions of each chain agree after the moves
    x = pos.cpu().numpy()
    for r, ch in enumerate(chains):
        err = np.abs(eng.unpack(x)[r] - ch.x).max()
        assert err <= 1e-8, err
    eng.close()


def test_mixed_chains_run_and_accept():
    """Mixed precision, graph replay: finite, and the acceptance rate of
    these short moves is neither 0 nor 1 for every chain."""
    import torch
    dev = torch.device("cuda:0")
    systems = [system_of(configs.butane_box(nmol=n, seed=90 + k))
               for k, n in enumerate((64, 128, 96, 200))]
    eng = Engine.replicas(systems, precision="mixed")
    pos, vel, fl, fs, lt, st, uk = _device_state(eng, systems, dev)
    eng.hmc_init([1000 + r for r in range(4)], [2000 + r for r in range(4)],
                 [s.treq for s in systems], 0, vel)
    eng.hmc_run_device(pos, vel, fl, fs, systems[0].delt, 5, 10, 20, lt, st, uk, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    acc, mov, dh = eng.hmc_stats()
    print("accepted", acc, "of", mov, flush=True)
    assert (mov == 20).all()
    assert bool(torch.isfinite(pos).all())
    assert acc.sum() > 0
    eng.close()


def _reference_chains():
    """The two reference HMC chains (tests/golden/hmc_butane_*, the
    reference's own hmc()) as one replica table."""
    from tests.test_hmc_golden import CASES, adapt_of, load
    out = [load(n) for n in CASES]
    return [s for s, _ in out], [ref for _, ref in out], [adapt_of(s.run) for s, _ in out]


@pytest.mark.parametrize("graph", [True, False])
def test_fp64_chains_match_reference_hmc(graph):
    """Both reference chains in one table, FP64: velinit's velocities to
    1e-13; per move the accept decision equal, dH within 1e-7 (relative,
    floor 1e-3) and each chain's delt (one adapting every 5 moves toward
    50 % acceptance, one fixed) bit for bit; positions after the last move
    within 1e-8 of the box."""
    import torch
    dev = torch.device("cuda:0")
    systems, refs, adapts = _reference_chains()
    eng = Engine.replicas(systems, precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _device_state(eng, systems, dev)
    runs = [s.run for s in systems]
    eng.hmc_init([r.ij for r in runs], [r.jk for r in runs], [s.treq for s in systems],
                 [r.nconstraint for r in runs], vel)
    torch.cuda.synchronize()
    v = vel.cpu().numpy()
    for k, ref in enumerate(refs):
        vr = eng.unpack(v)[k]
        assert np.abs(vr - ref["v0"]).max() <= 1e-13 * np.abs(ref["v0"]).max()
    eng.hmc_set_delt([s.delt for s in systems], [a[0] if a else 0 for a in adapts],
                     [a[1] if a else 0.0 for a in adapts])
    inner, nmd = runs[0].nstep_inner, runs[0].nstep_md_per_hmc
    nmove = min(len(ref["accept"]) for ref in refs)
    prev = np.zeros(len(systems), dtype=np.int64)
    for m in range(nmove):
        eng.hmc_run_device(pos, vel, fl, fs, systems[0].delt, inner, nmd, 1, lt, st, uk,
                           graph=graph)
        torch.cuda.synchronize()
        acc, mov, dh = eng.hmc_stats()
        delt = eng.hmc_get_delt()
        for k, ref in enumerate(refs):
            a = int(acc[k] - prev[k])
            print(f"chain {k} move {m}: ref accept={ref['accept'][m]} dH={ref['dH'][m]:.6e} "
                  f"delt={ref['delt'][m]:.6e} | gpu accept={a} dH={dh[k]:.6e} "
                  f"delt={delt[k]:.6e}", flush=True)
            assert a == int(ref["accept"][m])
            assert abs(dh[k] - ref["dH"][m]) <= 1e-7 * max(abs(ref["dH"][m]), 1e-3)
            assert delt[k] == ref["delt"][m]
        prev = acc.copy()
    eng.sync()
    if nmove == len(refs[1]["accept"]):
        x = eng.unpack(pos.cpu().numpy())[1]
        assert np.abs(x - refs[1]["x"]).max() <= 1e-8 * systems[1].params["boxlx"]
    eng.close()


def test_fp64_chain_reaches_reference_final_state():
    """The adapting chain alone for all of its reference moves: its final
    positions within 1e-8 of the box and its final delt bit for bit."""
    import torch
    dev = torch.device("cuda:0")
    systems, refs, adapts = _reference_chains()
    s, ref, ad = systems[0], refs[0], adapts[0]
    eng = Engine.replicas([s], precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _device_state(eng, [s], dev)
    r = s.run
    eng.hmc_init(r.ij, r.jk, s.treq, r.nconstraint, vel)
    eng.hmc_set_delt(s.delt, ad[0], ad[1])
    eng.hmc_run_device(pos, vel, fl, fs, s.delt, r.nstep_inner, r.nstep_md_per_hmc,
                       len(ref["accept"]), lt, st, uk, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    acc, mov, _ = eng.hmc_stats()
    assert int(mov[0]) == len(ref["accept"]) and int(acc[0]) == int(ref["accept"].sum())
    assert eng.hmc_get_delt()[0] == ref["delt"][-1]
    assert np.abs(pos.cpu().numpy() - ref["x"]).max() <= 1e-8 * s.params["boxlx"]
    eng.close()
