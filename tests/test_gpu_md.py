"""Device integrators against the reference's own trajectories
(tests/golden/md_*: its compiled integrators run from velinit's
velocities): Nose-Hoover NVT (mspms_nvt_run_device, vver_nh_3) and NVE
RESPA (mspms_vver_run_device, vver); NVT replica batches against the CPU
restatement (tests/md_cpu.py).

Tolerances: FP64 mode, ||dx|| / ||x|| and ||dv|| / ||v|| <= 1e-9 after 1 and
10 steps and <= 1e-7 after 100; thermostat momenta within 1e-8 of their
scale.  The device adds the sums over atoms in a tree order instead of left
to right (1e-16 relative per sum), and a sum that lands within that of the
Newton iteration's 1e-10 stopping test can take one pass more or less,
which moves the thermostat momentum by up to ~1e-10 of itself; the growth of
such differences over the trajectory sets the bounds.  Mixed mode: 1e-5
after 10 steps."""
import os

import numpy as np
import pytest

from mspms2_amd import configs, runio
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
from mspms2_amd.system import System
from tests.md_cpu import NvtCpu
from tests.util import GOLDEN, system_of

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = float(np.linalg.norm(np.asarray(a) - b))
    n = float(np.linalg.norm(b))
    return d / n if n > 0 else d


def _load(name, n):
    path = os.path.join(GOLDEN, name)
    rd = runio.RunDir.read(path)
    return rd, System.from_rundir(rd), np.load(os.path.join(path, f"md{n}.npz"))


def _run(eng, s, rd, v0, nsteps, graph, pos0=None):
    import torch
    dev = torch.device("cuda:0")
    x0 = np.stack([s.xx, s.yy, s.zz]) if pos0 is None else pos0
    f0, _ = eng.frclong(*x0)
    fs0, _ = eng.frcshort(*x0)
    pos = torch.tensor(x0, dtype=torch.float64, device=dev).contiguous()
    vel = torch.tensor(v0, dtype=torch.float64, device=dev).contiguous()
    fl = torch.tensor(f0, dtype=torch.float64, device=dev).contiguous()
    fs = torch.tensor(fs0, dtype=torch.float64, device=dev).contiguous()
    lt = torch.zeros(eng.nrep * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(eng.nrep * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(eng.nrep, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.nvt_run_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, nsteps, lt, st, uk,
                       graph=graph)
    torch.cuda.synchronize()
    eng.sync()
    return pos.cpu().numpy(), vel.cpu().numpy(), fl.cpu().numpy(), uk.cpu().numpy()


@pytest.mark.parametrize("name,n", [("md_lj864_nvt", 1), ("md_lj864_nvt", 10),
                                    ("md_lj864_nvt", 100), ("md_water216_nvt", 1),
                                    ("md_water216_nvt", 10)])
@pytest.mark.parametrize("graph", [False, True])
def test_fp64_matches_reference_trajectory(name, n, graph):
    rd, s, g = _load(name, n)
    print(f"case {name} steps={n} graph={graph} natom={s.natom}", flush=True)
    eng = Engine(s, precision="fp64")
    nfree = 3 * s.natom - rd.run.nconstraint
    eng.nvt_set(rd.run.qq, rd.run.qqs, s.treq, nfree)
    x, v, fl, uk = _run(eng, s, rd, g["v0"], n, graph)
    tol = 1e-7 if n >= 100 else 1e-9
    ex, ev = rel(x, g["x"]), rel(v, g["v"])
    print(f"  x {ex:.2e} v {ev:.2e} fl {rel(fl, g['fl']):.2e}")
    assert ex <= tol and ev <= tol
    assert rel(fl, g["fl"]) <= 100 * tol
    st = eng.nvt_get()[0]
    for k in ("ps", "pss"):
        ref = float(g[f"md_{k}"])
        scale = max(abs(ref), 1e-3)
        assert abs(st[k] - ref) <= 10 * tol * scale, (k, st[k], ref)
    assert abs(st["ukin"] - float(g["md_ukin"])) <= tol * abs(float(g["md_ukin"]))
    assert abs(uk[0] - st["ukin"]) <= 1e-15 * abs(st["ukin"])
    assert abs(st["tinst"] - float(g["md_tinst"])) <= tol * abs(float(g["md_tinst"]))
    eng.close()


def test_mixed_follows_reference_trajectory():
    rd, s, g = _load("md_water216_nvt", 10)
    eng = Engine(s, precision="mixed")
    eng.nvt_set(rd.run.qq, rd.run.qqs, s.treq, 3 * s.natom - rd.run.nconstraint)
    x, v, _, _ = _run(eng, s, rd, g["v0"], 10, True)
    print(f"mixed x {rel(x, g['x']):.2e} v {rel(v, g['v']):.2e}")
    assert rel(x, g["x"]) <= 1e-5 and rel(v, g["v"]) <= 1e-5
    eng.close()


def test_replica_batch_matches_per_replica_restatement():
    """Three LJ boxes of different sizes, seeds and temperatures in one
    table: every replica's chain equals the CPU restatement of its own."""
    boxes = [configs.lj_fcc(ncell=3, jitter=0.05, seed=11, temp=0.72),
             configs.lj_fcc(ncell=4, jitter=0.05, seed=12, temp=1.1),
             configs.lj_fcc(ncell=3, jitter=0.04, seed=13, temp=0.9)]
    systems = [system_of(b) for b in boxes]
    rng = np.random.default_rng(3)
    v0s = [rng.standard_normal((3, s.natom)) * np.sqrt(s.treq / s.masses()) for s in systems]
    nsteps = 8
    ref = []
    for b, s, v0 in zip(boxes, systems, v0s):
        c = NvtCpu(s, v0, b.run.qq, b.run.qqs, b.run.nconstraint)
        for _ in range(nsteps):
            c.step()
        ref.append(c)
    eng = Engine.replicas(systems, precision="fp64")
    eng.nvt_set([b.run.qq for b in boxes], [b.run.qqs for b in boxes],
                [s.treq for s in systems],
                [3 * s.natom - b.run.nconstraint for b, s in zip(boxes, systems)])
    s0 = systems[0]
    x, v, _, uk = _run(eng, s0, boxes[0], eng.pack(v0s), nsteps, True,
                       pos0=eng.packed_positions())
    states = eng.nvt_get()
    for r, c in enumerate(ref):
        xr, vr = eng.unpack(x)[r], eng.unpack(v)[r]
        print(f"replica {r}: x {rel(xr, c.x):.2e} v {rel(vr, c.v):.2e}")
        assert rel(xr, c.x) <= 1e-9 and rel(vr, c.v) <= 1e-9
        assert abs(states[r]["ps"] - c.ps) <= 1e-8 * max(abs(c.ps), 1e-3)
        assert abs(uk[r] - c.ukin) <= 1e-9 * c.ukin
    eng.close()


@pytest.mark.parametrize("n", [1, 10])
def test_nve_fp64_matches_reference_trajectory(n):
    """The device NVE RESPA step (mspms_vver_run_device) against the
    reference's own vver trajectory (tests/golden/md_water216_nve)."""
    import torch
    rd, s, g = _load("md_water216_nve", n)
    eng = Engine(s, precision="fp64")
    dev = torch.device("cuda:0")
    x0 = np.stack([s.xx, s.yy, s.zz])
    f0, _ = eng.frclong(*x0)
    fs0, _ = eng.frcshort(*x0)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    pos, vel, fl, fs = t(x0), t(g["v0"]), t(f0), t(fs0)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.vver_run_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, n, lt, st, uk, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    x, v = pos.cpu().numpy(), vel.cpu().numpy()
    print(f"nve {n}: x {rel(x, g['x']):.2e} v {rel(v, g['v']):.2e}")
    assert rel(x, g["x"]) <= 1e-9 and rel(v, g["v"]) <= 1e-9
    assert abs(uk.item() - float(g["md_ukin"])) <= 1e-9 * float(g["md_ukin"])
    eng.close()


@pytest.mark.parametrize("name,n", [("md_lj864_npt", 1), ("md_lj864_npt", 10),
                                    ("md_water216_npt", 1), ("md_water216_npt", 5)])
@pytest.mark.parametrize("graph", [False, True])
def test_npt_fp64_matches_reference_trajectory(name, n, graph):
    """The device NPT step (mspms_npt_run_device) against the reference's
    own npt_respa trajectory: positions, velocities, the box, thermostat and
    barostat velocities, the conserved quantity's extra part and the
    instantaneous pressure (same bounds as the NVT trajectories)."""
    import torch
    rd, s, g = _load(name, n)
    print(f"case {name} steps={n} graph={graph}", flush=True)
    eng = Engine(s, precision="fp64")
    nfree = 3 * s.natom - rd.run.nconstraint
    eng.npt_set(rd.run.Qts, rd.run.Qbs, s.treq, s.preq, nfree, lj_tail=bool(rd.run.isLJlrcOn))
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    pos, vel = t(np.stack([s.xx, s.yy, s.zz])), t(g["v0"])
    fl, fs = torch.zeros_like(pos), torch.zeros_like(pos)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    # forces and terms (virials) of the start, as after the reference's
    # start-up frclong / frcshort
    eng.frclong_device(pos, fl, lt)
    eng.frcshort_device(pos, fs, st)
    eng.sync()
    eng.npt_run_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, n, lt, st, uk, graph=graph)
    torch.cuda.synchronize()
    eng.sync()
    x, v = pos.cpu().numpy(), vel.cpu().numpy()
    state = eng.npt_get()[0]
    tol = 1e-9
    print(f"  x {rel(x, g['x']):.2e} v {rel(v, g['v']):.2e} box {state['boxl'][0]!r} "
          f"vs {float(g['box'][0])!r}")
    assert rel(x, g["x"]) <= tol and rel(v, g["v"]) <= tol
    assert np.allclose(state["boxl"], g["box"], rtol=1e-12, atol=0)
    for k in ("vts", "vbs", "rts", "utsbs", "pinst", "ukin", "tinst"):
        ref = float(g[f"md_{k}"])
        assert abs(state[k] - ref) <= 10 * tol * max(abs(ref), 1e-3), (k, state[k], ref)
    assert abs(uk.item() - state["ukin"]) <= 1e-15 * abs(state["ukin"])
    eng.close()


def test_nvt_restart_from_checkpoint_matches_reference():
    """A run directory with starting option 2: positions, velocities and both
    thermostats come from old.mspms (the reference's loadit); the device
    continues the run as the reference's own integrator does."""
    rd = runio.RunDir.read(os.path.join(GOLDEN, "md_water216_nvt_restart"))
    s = System.from_rundir(rd)
    ck = rd.checkpoint
    g = np.load(os.path.join(GOLDEN, "md_water216_nvt_restart", "md5.npz"))
    eng = Engine(s, precision="fp64")
    eng.nvt_set(ck.qq, ck.qqs, s.treq, 3 * s.natom - rd.run.nconstraint, ss=ck.ss, ps=ck.ps,
                sss=ck.sss, pss=ck.pss)
    x, v, _, _ = _run(eng, s, rd, s.v0, 5, True)
    print(f"restart x {rel(x, g['x']):.2e} v {rel(v, g['v']):.2e}")
    assert rel(x, g["x"]) <= 1e-9 and rel(v, g["v"]) <= 1e-9
    st = eng.nvt_get()[0]
    assert abs(st["ps"] - float(g["md_ps"])) <= 1e-8 * max(abs(float(g["md_ps"])), 1e-3)
    eng.close()


def _siman_buffers(eng, x0):
    import torch
    dev = torch.device("cuda:0")
    pos = torch.tensor(x0, dtype=torch.float64, device=dev).contiguous()
    vel = torch.zeros_like(pos)
    fl, fs = torch.zeros_like(pos), torch.zeros_like(pos)
    lt = torch.zeros(eng.nrep * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(eng.nrep * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(eng.nrep, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    return pos, vel, fl, fs, lt, st, uk


def test_siman_matches_reference():
    """Device simulated annealing (mspms_siman_run_device) against the
    reference's own siman (tests/golden/siman_lj256), FP64."""
    import torch
    path = os.path.join(GOLDEN, "siman_lj256")
    rd = runio.RunDir.read(path)
    s = System.from_rundir(rd)
    g = np.load(os.path.join(path, "siman.npz"))
    eb = s.params["epsilon_base"]
    eng = Engine(s, precision="fp64")
    pos, vel, fl, fs, lt, st, uk = _siman_buffers(eng, np.stack([s.xx, s.yy, s.zz]))
    nfree = 3 * s.natom - rd.run.nconstraint
    # rmarin(ij, jk) and init_vars' velinit at t_initial (init_siman sets treq)
    eng.hmc_init(rd.run.ij, rd.run.jk, rd.run.t_initial / eb, rd.run.nconstraint, vel)
    eng.nvt_set(rd.run.qq, rd.run.qqs, rd.run.t_initial / eb, nfree)
    cyc, tf = eng.siman_run_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, rd.run.n_tries,
                                   rd.run.t_initial / eb, rd.run.t_min / eb, rd.run.mu_t, lt, st,
                                   uk)
    torch.cuda.synchronize()
    eng.sync()
    x, v = pos.cpu().numpy(), vel.cpu().numpy()
    print(f"siman cycles {cyc} T {tf} x {rel(x, g['x']):.2e} v {rel(v, g['v']):.2e}")
    assert cyc[0] == 3 and tf[0] == float(g["treq"])
    assert rel(x, g["x"]) <= 1e-9 and rel(v, g["v"]) <= 1e-9
    eng.close()


def test_siman_replica_batch_matches_restatement():
    """Three LJ boxes with their own seeds and schedules (2, 3 and 4
    cycles) annealed in one table; each equals the restatement of its own."""
    import torch
    from tests.md_cpu import SimanCpu
    boxes, sched = [], [(150.0, 100.0, 1.6), (150.0, 80.0, 1.5), (200.0, 70.0, 1.5)]
    for k, (ti, tm, mu) in enumerate(sched):
        b = configs.lj_fcc(ncell=3, jitter=0.05, seed=30 + k)
        b.run.what_simulation, b.run.what_ensemble = 2, 1
        b.run.ij, b.run.jk = 1000 + k, 2000 + k
        b.run.n_tries = 6
        b.run.t_initial, b.run.t_min, b.run.mu_t = ti, tm, mu
        boxes.append(b)
    systems = [system_of(b) for b in boxes]
    ref = [SimanCpu(s, b.run).anneal() for s, b in zip(systems, boxes)]
    eng = Engine.replicas(systems, precision="fp64")
    eb = systems[0].params["epsilon_base"]
    pos, vel, fl, fs, lt, st, uk = _siman_buffers(eng, eng.packed_positions())
    eng.hmc_init([b.run.ij for b in boxes], [b.run.jk for b in boxes],
                 [b.run.t_initial / eb for b in boxes], [b.run.nconstraint for b in boxes], vel)
    eng.nvt_set([b.run.qq for b in boxes], [b.run.qqs for b in boxes],
                [b.run.t_initial / eb for b in boxes],
                [3 * s.natom - b.run.nconstraint for s, b in zip(systems, boxes)])
    cyc, tf = eng.siman_run_device(pos, vel, fl, fs, systems[0].delt, 1, 6,
                                   [b.run.t_initial / eb for b in boxes],
                                   [b.run.t_min / eb for b in boxes],
                                   [b.run.mu_t for b in boxes], lt, st, uk)
    torch.cuda.synchronize()
    eng.sync()
    x, v = eng.unpack(pos.cpu().numpy()), eng.unpack(vel.cpu().numpy())
    for r, c in enumerate(ref):
        print(f"replica {r}: cycles {cyc[r]} / {c.cycles} x {rel(x[r], c.x):.2e} "
              f"v {rel(v[r], c.v):.2e}")
        assert cyc[r] == c.cycles and tf[r] == c.T
        assert rel(x[r], c.x) <= 1e-9 and rel(v[r], c.v) <= 1e-9
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_without_histories_equals_single_steps(precision):
    """A device run of 6 steps reduces its terms only on the last step; its
    state and terms equal 6 single steps from the same start (the reciprocal
    force atomics leave last-bit differences only)."""
    import torch
    from mspms2_amd import abi
    rd, s, g = _load("md_water216_nve", 10)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    out = []
    for mode in ("run", "steps"):
        eng = Engine(s, precision=precision)
        x0 = np.stack([s.xx, s.yy, s.zz])
        f0, _ = eng.frclong(*x0)
        fs0, _ = eng.frcshort(*x0)
        pos, vel, fl, fs = t(x0), t(g["v0"]), t(f0), t(fs0)
        lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        uk = torch.zeros(1, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        if mode == "run":
            eng.vver_run_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, 6, lt, st, uk,
                                graph=True)
        else:
            for _ in range(6):
                eng.vver_step_device(pos, vel, fl, fs, s.delt, rd.run.nstep_inner, lt, st, uk)
        torch.cuda.synchronize()
        eng.sync()
        terms = abi.LongTerms.from_buffer_copy(lt.cpu().numpy().tobytes()).as_dict()
        out.append((pos.cpu().numpy(), vel.cpu().numpy(), terms, uk.item()))
        eng.close()
    (x1, v1, t1, k1), (x2, v2, t2, k2) = out
    assert rel(x1, x2) <= 1e-12 and rel(v1, v2) <= 1e-12
    assert abs(k1 - k2) <= 1e-12 * abs(k2)
    for k in ("uvdw", "ureal", "ufourier", "uinter", "virial_inter"):
        assert abs(t1[k] - t2[k]) <= 1e-10 * (abs(t2[k]) + abs(t2["ureal"])), k
    assert t1["n_lj_pairs"] == t2["n_lj_pairs"]


def _histories_vs_steps(s, nin, v0, precision, nsteps=6):
    """A device run of nsteps recording histories against nsteps single
    steps from the same start, each step's terms read after it."""
    import torch
    from mspms2_amd import abi
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    res = {}
    for mode in ("run", "steps"):
        eng = Engine(s, precision=precision)
        x0 = np.stack([s.xx, s.yy, s.zz])
        f0, _ = eng.frclong(*x0)
        fs0, _ = eng.frcshort(*x0)
        pos, vel, fl, fs = t(x0), t(v0), t(f0), t(fs0)
        lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        uk = torch.zeros(1, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        if mode == "run":
            lh = torch.zeros(nsteps * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
            sh = torch.zeros(nsteps * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
            kh = torch.zeros(nsteps, dtype=torch.float64, device=dev)
            eng.vver_run_device(pos, vel, fl, fs, s.delt, nin, nsteps, lt, st, uk, lhist=lh,
                                shist=sh, ukin_hist=kh, graph=True)
            torch.cuda.synchronize()
            eng.sync()
            L = (abi.LongTerms * nsteps).from_buffer_copy(lh.cpu().numpy().tobytes())
            S = (abi.ShortTerms * nsteps).from_buffer_copy(sh.cpu().numpy().tobytes())
            hist = [(L[k].as_dict(), S[k].as_dict(), float(kh[k].item())) for k in range(nsteps)]
        else:
            hist = []
            for _ in range(nsteps):
                eng.vver_step_device(pos, vel, fl, fs, s.delt, nin, lt, st, uk)
                torch.cuda.synchronize()
                hist.append((abi.LongTerms.from_buffer_copy(lt.cpu().numpy().tobytes()).as_dict(),
                             abi.ShortTerms.from_buffer_copy(st.cpu().numpy().tobytes()).as_dict(),
                             float(uk.item())))
            eng.sync()
        last = (abi.LongTerms.from_buffer_copy(lt.cpu().numpy().tobytes()).as_dict(),
                abi.ShortTerms.from_buffer_copy(st.cpu().numpy().tobytes()).as_dict(),
                float(uk.item()))
        res[mode] = (hist, last, pos.cpu().numpy())
        eng.close()
    (h1, last1, x1), (h2, last2, x2) = res["run"], res["steps"]
    assert rel(x1, x2) <= 1e-12
    for k in range(nsteps):
        (a, b, ka), (c, d, kc) = h1[k], h2[k]
        scale = abs(c["ureal"]) + abs(c["uvdw"]) + abs(c["uinter"])
        for key in ("uvdw", "ureal", "ufourier", "uvacuum", "uelec", "uinter", "virial_inter"):
            assert abs(a[key] - c[key]) <= 1e-10 * scale, (k, key, a[key], c[key])
        for key in ("n_lj_pairs", "n_elec_pairs"):
            assert a[key] == c[key], (k, key)
        for key in ("ubond", "uangle", "uintra", "virial_intra"):
            assert abs(b[key] - d[key]) <= 1e-10 * (abs(d[key]) + scale), (k, key)
        assert abs(ka - kc) <= 1e-12 * abs(kc)
    # the run's outputs are its last step's terms
    assert last1[0] == h1[-1][0] and last1[1] == h1[-1][1] and last1[2] == h1[-1][2]
    return h1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_histories_equal_single_steps(precision):
    """Every step's terms recorded by a device run (each step stores its
    workgroup sums; the terms are formed after the run) equal those of
    single steps from the same start, step by step."""
    rd, s, g = _load("md_water216_nve", 10)
    _histories_vs_steps(s, rd.run.nstep_inner, g["v0"], precision)


def test_run_histories_vacuum_boundary():
    """The vacuum-boundary energy of every step is recorded with it (the
    dipole term changes each step): the flexible water of md_water216_nve
    with the Ewald sum's vacuum boundary (src/frclong.c:1065-1126)."""
    rd, s, g = _load("md_water216_nve", 10)
    s.params["fEwald_BC"] = 2
    h = _histories_vs_steps(s, rd.run.nstep_inner, g["v0"], "fp64")
    vac = [h[k][0]["uvacuum"] for k in range(len(h))]
    assert len(set(vac)) > 1 and all(v != 0.0 for v in vac)
