"""Benchmark: RESPA MD of 12,288-site flexible water with Ewald on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Workload: 4096 flexible three-site waters (12,288 sites) on
a jittered 16^3 lattice in a 49.66 A box, Ewald alpha 0.28 / A, kmax 8,
KSQMAX 64 (2,108 k-vectors), rc 10 A; RESPA with the bonded terms in 8
inner steps of 0.25 fs and the nonbonded (frclong) in the 2 fs outer step.
A step is one device-resident RESPA velocity-Verlet outer step: 1 frclong
+ 8 frcshort + kicks/drifts, inputs resident in HBM, the launch shape picked
by the autotuner.  The timed region replays one captured step graph K times
(mspms_vver_run_device); the same K steps run eagerly before it with HIP
events around every kernel group on the launch stream, which give the
kernel durations of `roofline`.  Pair arithmetic FP32 with FP64
accumulation (FP64 reciprocal, bonded and integrator); the plain-FP64
engine is timed beside it.

Prints one JSON line (the benchmark contract) with `roofline` for the
dominant kernel and `cpu_baseline`: the reference's own frclong (oracle/_ref,
built from the reference sources) timed on the host, one thread.  Extra
legs: NVE energy drift (mixed and FP64) from an equilibrated state, batched
HMC (256 n-butane boxes, one table against one engine per box), the Wolf
water fitting set with the parameter adjoint, the LJ-864 NVT workload and
the grid generators.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DT_FS = 2.0
NSTEP_INNER = 8
PEAK_FP32_TFLOPS = 157.3   # MI355X published vector/matrix FP32
PEAK_FP64_TFLOPS = 78.6    # MI355X published FP64
PEAK_HBM_GBS = 8000.0


def progress(what):
    """A progress line on stderr (stdout carries only the JSON record)."""
    print(f"bench: {what}", file=sys.stderr, flush=True)


def make_workload():
    from mspms2_amd import configs
    from mspms2_amd.system import system_of
    rd = configs.water_box(n_side=16, L=49.66, kappa=0.28, kmax=8, ksqmax=64, rc=10.0,
                           seed=2, nstep_inner=NSTEP_INNER, delt_fs=DT_FS)
    return rd, system_of(rd)


def maxwell_velocities(system, seed=7):
    """Maxwell-Boltzmann velocities at treq, zero total momentum (the
    reference's velinit, src/initvars.c:151-232, with numpy's generator)."""
    rng = np.random.default_rng(seed)
    m = system.masses()
    v = rng.standard_normal((3, system.natom)) * np.sqrt(system.treq / m)
    p = (v * m).sum(axis=1, keepdims=True) / m.sum()
    return v - p


def run_gpu(system, precision, steps, warmup, profile=True, skin=0.0):
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    from mspms2_amd import abi
    import ctypes as C
    dev = torch.device("cuda:0")
    eng = Engine(system, precision=precision)
    n = system.natom
    pos = torch.tensor(np.stack([system.xx, system.yy, system.zz]), dtype=torch.float64,
                       device=dev).contiguous()
    vel = torch.tensor(maxwell_velocities(system), dtype=torch.float64, device=dev).contiguous()
    fl = torch.zeros((3, n), dtype=torch.float64, device=dev)
    fs = torch.zeros((3, n), dtype=torch.float64, device=dev)
    lterms = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    sterms = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    ukin = torch.zeros(1, dtype=torch.float64, device=dev)
    # every launch and both timing events on one explicit stream (default
    # priority): the engine's kernels, its per-kernel-group HIP events and
    # the step timer all see the same queue
    st = torch.cuda.Stream()
    stream = st.cuda_stream
    # forces of the starting positions (src/mspms2.c:50-51); the host call
    # also sizes the pair list
    f0, t0 = eng.frclong()
    fs0, _ = eng.frcshort()
    with torch.cuda.stream(st):
        fl.copy_(torch.tensor(f0))
        fs.copy_(torch.tensor(fs0))
    torch.cuda.synchronize()
    # launch shape (cell size x waves per i-cluster) picked by timing
    tune = eng.autotune(pos, repeats=5)
    # pair-list buffer (reduced units): lists reused until a site has moved
    # skin / 2 since the build
    eng.set_list_skin(skin)
    delt = system.delt
    # the autotuner times frclong calls that each build their list; the
    # RESPA step reuses lists and overlaps the reciprocal side stream, so
    # the shape is refined on graph-replayed steps (untimed, like warmup)
    best = None
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for cell in (6.0, 8.0, 12.0, 16.0):
        for wpc in (1, 2, 4, 8):
            eng.set_launch_shape(cell, wpc)
            eng.vver_run_device(pos, vel, fl, fs, delt, NSTEP_INNER, 8, lterms, sterms, ukin,
                                graph=True, stream=stream)
            torch.cuda.synchronize()
            ev0.record(st)
            eng.vver_run_device(pos, vel, fl, fs, delt, NSTEP_INNER, 60, lterms, sterms, ukin,
                                graph=True, stream=stream)
            ev1.record(st)
            torch.cuda.synchronize()
            eng.sync()
            t = ev0.elapsed_time(ev1) / 60
            if best is None or t < best[0]:
                best = (t, cell, wpc)
    eng.set_launch_shape(best[1], best[2])
    tune = dict(tune, cell_atoms=best[1], waves_per_cluster=best[2],
                refined_on="graph-replayed RESPA steps")
    for _ in range(warmup):
        eng.vver_step_device(pos, vel, fl, fs, delt, NSTEP_INNER, lterms, sterms, ukin, stream)
    torch.cuda.synchronize()
    eng.sync()
    # (1) eager steps with per-kernel-group HIP events on the launch stream:
    #     the kernel durations of the roofline
    if profile:
        eng.set_profiling(True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(st)
    for _ in range(steps):
        eng.vver_step_device(pos, vel, fl, fs, delt, NSTEP_INNER, lterms, sterms, ukin, stream)
    stop.record(st)
    torch.cuda.synchronize()
    eng.sync()
    ms_eager = start.elapsed_time(stop) / steps
    groups = eng.timings() if profile else {}
    eng.set_profiling(False)
    # (2) the same steps replayed from one captured RESPA-step graph: the
    #     timed region of the headline figure.  Every step reduces its
    #     energy terms, virial and kinetic energy and appends them to
    #     device histories (the per-step record the reference's main loop
    #     consumes through collect_aves, src/mspms2.c:101-102)
    lhist = torch.zeros(steps * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    shist = torch.zeros(steps * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    khist = torch.zeros(steps, dtype=torch.float64, device=dev)
    eng.vver_run_device(pos, vel, fl, fs, delt, NSTEP_INNER, 2, lterms, sterms, ukin, graph=True,
                        stream=stream)
    torch.cuda.synchronize()
    builds0 = eng.list_builds()
    t_wall = time.perf_counter()
    start.record(st)
    eng.vver_run_device(pos, vel, fl, fs, delt, NSTEP_INNER, steps, lterms, sterms, ukin,
                        lhist=lhist, shist=shist, ukin_hist=khist, graph=True, stream=stream)
    stop.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    eng.sync()
    ms = start.elapsed_time(stop) / steps
    builds = eng.list_builds() - builds0
    raw = lterms.cpu().numpy().view(np.uint8)
    lt = abi.LongTerms.from_buffer_copy(raw.tobytes()).as_dict()
    hl = (abi.LongTerms * steps).from_buffer_copy(lhist.cpu().numpy().tobytes())
    hist_ok = bool(np.isfinite([hl[k].uinter for k in range(steps)]).all()) and \
        hl[steps - 1].uinter == lt["uinter"]
    # the long forces and positions the last timed step left (launch shape
    # and list state of the timed region), for the accuracy check
    pos_after = pos.cpu().numpy().copy()
    f_after = fl.cpu().numpy().copy()
    # the same steps with only the last step's terms reduced (no consumer
    # of the others): an extra figure, not the headline
    start.record(st)
    eng.vver_run_device(pos, vel, fl, fs, delt, NSTEP_INNER, steps, lterms, sterms, ukin,
                        graph=True, stream=stream)
    stop.record(st)
    torch.cuda.synchronize()
    eng.sync()
    ms_last_only = start.elapsed_time(stop) / steps
    # the drop-in path (host buffers: positions up, forces down, every call)
    eng.set_profiling(False)
    hx, hy, hz = (np.ascontiguousarray(v) for v in pos.cpu().numpy())
    eng.frclong(hx, hy, hz)
    t_host = time.perf_counter()
    for _ in range(5):
        eng.frclong(hx, hy, hz)
    host_ms = 1e3 * (time.perf_counter() - t_host) / 5
    out = dict(ms_per_step=ms, wall_ms_per_step=1e3 * wall / steps, groups=groups, terms=lt,
               finite=bool(torch.isfinite(pos).all().item()) and hist_ok, f_initial=f0,
               t_initial=t0, tune=tune, host_frclong_ms=host_ms, eager_ms_per_step=ms_eager,
               list_builds_per_step=builds / steps, ms_per_step_last_terms_only=ms_last_only,
               pos_after=pos_after, f_after=f_after)
    eng.close()
    return out


def equilibrate(system, steps, skin=0.0, chunk=100):
    """Equilibration of the jittered lattice: `steps` device RESPA steps
    (mixed) in chunks of `chunk`, the velocities rescaled to the run's
    temperature after each chunk (the lattice start releases ~1e6 K of LJ
    overlap energy, too much for a Nose-Hoover start).  Returns the final
    (positions, velocities) as host arrays and the temperature before the
    last rescale."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    dev = torch.device("cuda:0")
    eng = Engine(system, precision="mixed")
    eng.set_list_skin(skin)
    nfree = 3 * system.natom - system.run.nconstraint
    pos = torch.tensor(np.stack([system.xx, system.yy, system.zz]), device=dev)
    vel = torch.tensor(maxwell_velocities(system), device=dev)
    fl = torch.tensor(eng.frclong()[0], device=dev)
    fs = torch.tensor(eng.frcshort()[0], device=dev)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    tinst = system.treq
    for done in range(0, steps, chunk):
        eng.vver_run_device(pos, vel, fl, fs, system.delt, NSTEP_INNER, min(chunk, steps - done),
                            lt, st, uk, graph=True)
        torch.cuda.synchronize()
        eng.sync()
        tinst = 2.0 * float(uk.item()) / nfree
        vel.mul_(np.sqrt(system.treq / tinst))
    torch.cuda.synchronize()
    out = pos.cpu().numpy(), vel.cpu().numpy(), tinst * system.params["epsilon_base"]
    eng.close()
    return out


def run_drift(system, precision, steps=1000, skin=0.0, start=None):
    """NVE energy conservation: `steps` RESPA outer steps replayed from the
    captured graph with per-step terms recorded on the device, from `start`
    = (positions, velocities) (an equilibrated state) or else the jittered
    lattice; returns the least-squares slope of H = ukin + uinter + uintra
    in K per ns per site (reduced energies are K here: epsilon_base = 1)
    and the rms fluctuation of H about the fit."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    from mspms2_amd import abi
    dev = torch.device("cuda:0")
    eng = Engine(system, precision=precision)
    eng.set_list_skin(skin)
    n = system.natom
    if start is None:
        start = (np.stack([system.xx, system.yy, system.zz]), maxwell_velocities(system))
    pos = torch.tensor(start[0], device=dev)
    vel = torch.tensor(start[1], device=dev)
    p0 = start[0]
    f0, _ = eng.frclong(p0[0], p0[1], p0[2])
    fs0, _ = eng.frcshort(p0[0], p0[1], p0[2])
    fl, fs = torch.tensor(f0, device=dev), torch.tensor(fs0, device=dev)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    lh = torch.zeros(steps * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    sh = torch.zeros(steps * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    kh = torch.zeros(steps, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    eng.vver_run_device(pos, vel, fl, fs, system.delt, NSTEP_INNER, steps, lt, st, uk, lhist=lh,
                        shist=sh, ukin_hist=kh, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    L = (abi.LongTerms * steps).from_buffer_copy(lh.cpu().numpy().tobytes())
    S = (abi.ShortTerms * steps).from_buffer_copy(sh.cpu().numpy().tobytes())
    H = np.array([L[k].uinter + S[k].uintra for k in range(steps)]) + kh.cpu().numpy()
    t_ns = DT_FS * 1e-6 * np.arange(1, steps + 1)
    slope, icpt = np.polyfit(t_ns, H, 1)
    eng.close()
    return {"steps": steps, "H_start_K": float(H[0]),
            "drift_K_per_ns_per_site": float(slope / n),
            "rms_about_fit_K_per_site": float(np.sqrt(np.mean((H - (slope * t_ns + icpt)) ** 2))
                                              / n),
            "finite": bool(np.isfinite(H).all())}


PROBE_COPY_MIB = 256


def run_probe(system, calls=5):
    """A short fixed workload for hardware-counter passes: the autotuned
    frclong of the benchmark system `calls` times (mixed, device buffers,
    every call with a list build)
    and a plain device-to-device copy of PROBE_COPY_MIB MiB, the reference
    transfer that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, Engine
    dev = torch.device("cuda:0")
    eng = Engine(system, precision="mixed")
    pos = torch.tensor(np.stack([system.xx, system.yy, system.zz]), device=dev)
    f = torch.zeros_like(pos)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    eng.frclong()
    torch.cuda.synchronize()
    tune = eng.autotune(pos, repeats=2)
    st = torch.cuda.Stream()
    # a 1e-7 nudge of every site before each call: each call rebuilds its
    # pair list, as a call after a move does
    nudge = torch.full_like(pos, 1e-7)
    torch.cuda.synchronize()
    for k in range(calls):
        pos.add_(nudge if k % 2 == 0 else -nudge)
        torch.cuda.synchronize()
        eng.frclong_device(pos, f, lt, stream=st.cuda_stream)
    torch.cuda.synchronize()
    eng.sync()
    n = PROBE_COPY_MIB * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    eng.close()
    print(json.dumps({"probe": "frclong x %d + copy %d MiB x 3" % (calls, PROBE_COPY_MIB),
                      "launch_shape": {k: tune[k] for k in ("cell_atoms", "waves_per_cluster")}}))


def run_trace(system, steps, skin, shape=None):
    """Only the timed region of the headline: `steps` mixed RESPA steps
    replayed from the captured graph (after the autotuner, or the launch
    shape given as (sites per cell, waves per i-cluster), and 10 eager
    steps), for kernel-trace runs; the last `steps` dispatches of each
    kernel are the graph's."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    dev = torch.device("cuda:0")
    eng = Engine(system, precision="mixed")
    pos = torch.tensor(np.stack([system.xx, system.yy, system.zz]), device=dev)
    vel = torch.tensor(maxwell_velocities(system), device=dev)
    fl = torch.tensor(eng.frclong()[0], device=dev)
    fs = torch.tensor(eng.frcshort()[0], device=dev)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    if shape is None:
        tune = eng.autotune(pos, repeats=2)
    else:
        eng.set_launch_shape(float(shape[0]), int(shape[1]))
        tune = {"cell_atoms": float(shape[0]), "waves_per_cluster": int(shape[1])}
    eng.set_list_skin(skin)
    for _ in range(10):
        eng.vver_step_device(pos, vel, fl, fs, system.delt, NSTEP_INNER, lt, st, uk)
    torch.cuda.synchronize()
    b0 = eng.list_builds()
    # per-step terms recorded, as in the timed region
    lh = torch.zeros(steps * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    sh = torch.zeros(steps * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    kh = torch.zeros(steps, dtype=torch.float64, device=dev)
    eng.vver_run_device(pos, vel, fl, fs, system.delt, NSTEP_INNER, steps, lt, st, uk, lhist=lh,
                        shist=sh, ukin_hist=kh, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    print(json.dumps({"trace": "%d graph-replayed mixed RESPA steps, per-step terms" % steps,
                      "list_builds": eng.list_builds() - b0,
                      "launch_shape": {k: tune[k] for k in ("cell_atoms", "waves_per_cluster")}}))
    eng.close()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_water12288.json")


def pmc_summary(fetch_csv, write_csv):
    """Per-kernel HBM bytes per launch from two rocprofv3 --pmc passes over
    `bench.py --probe` (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each
    scaled so that the reference copy of PROBE_COPY_MIB MiB reads and writes
    exactly its size.  Kernels are keyed by name (the launches of the chosen
    launch shape: the last `calls` dispatches of each name)."""
    import csv
    import re

    def load(path, counter):
        per = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = re.sub(r"\(anonymous namespace\)::", "", r["Kernel_Name"])
            name = re.sub(r"\(.*", "", name).replace("void ", "").strip()
            per.setdefault(name, []).append(float(r["Counter_Value"]) * 1024.0)
        return per

    fetch, write = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy = "__amd_rocclr_copyBuffer"
    size = PROBE_COPY_MIB * float(1 << 20)
    big = lambda v: [x for x in v if x > 0.5 * size]
    f_scale = size / np.mean(big(fetch[copy])[-3:])
    w_scale = size / np.mean(big(write[copy])[-3:])
    out = {"fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for name in sorted(set(fetch) & set(write)):
        if name == copy:
            continue
        f, w = fetch[name][-5:], write[name][-5:]
        out["kernels"][name] = {"fetch_bytes": float(np.mean(f)) * f_scale,
                                "write_bytes": float(np.mean(w)) * w_scale}
    return out


def traffic_of(kernels):
    """HBM bytes per launch summed over `kernels` (name prefixes) from the
    committed counter summary, or None where it is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    total, hit = 0.0, 0
    for pre in kernels:
        for name, v in t.items():
            if name.startswith(pre):
                total += v["fetch_bytes"] + v["write_bytes"]
                hit += 1
                break
    return total if hit == len(kernels) else None


def make_replicas(nrep=256, seed=2024):
    """Batched-replica workload: n-butane boxes of U{64..512} molecules at
    0.58 g/cm^3, rc = min(14 A, L/2), one seed per box."""
    from mspms2_amd import configs
    from mspms2_amd.system import system_of
    return [system_of(rd) for rd in configs.butane_table(nrep, seed)]


def _hmc_run(systems, moves, warmup, nmd, nstep_inner, dev):
    """HMC moves (nmd RESPA steps each) of one engine over `systems` (one
    replica table, one chain per replica), replayed from a captured move
    graph; returns (ms per move, i<j LJ pairs in range per force evaluation,
    accepted fraction, finite)."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    from mspms2_amd import abi
    eng = Engine.replicas(systems, precision="mixed")
    nr = len(systems)
    pos = torch.tensor(eng.packed_positions(), device=dev)
    vel = torch.zeros_like(pos)
    fl, fs = torch.zeros_like(pos), torch.zeros_like(pos)
    lt = torch.zeros(nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    stt = torch.zeros(nr * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(nr, dtype=torch.float64, device=dev)
    eng.hmc_init([1000 + r for r in range(nr)], [2000 + r for r in range(nr)],
                 [s.treq for s in systems], [s.run.nconstraint for s in systems], vel)
    st = torch.cuda.Stream()
    delt = systems[0].delt
    eng.hmc_run_device(pos, vel, fl, fs, delt, nstep_inner, nmd, warmup, lt, stt, uk,
                       graph=True, stream=st.cuda_stream)
    torch.cuda.synchronize()
    eng.sync()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(st)
    eng.hmc_run_device(pos, vel, fl, fs, delt, nstep_inner, nmd, moves, lt, stt, uk,
                       graph=True, stream=st.cuda_stream)
    b.record(st)
    torch.cuda.synchronize()
    eng.sync()
    raw = lt.cpu().numpy().view(np.uint8).tobytes()
    sz = LONG_TERMS_BYTES
    pairs = sum(abi.LongTerms.from_buffer_copy(raw[r * sz:(r + 1) * sz]).n_lj_pairs
                for r in range(nr))
    acc, mov, _ = eng.hmc_stats()
    finite = bool(torch.isfinite(pos).all().item())
    eng.close()
    return a.elapsed_time(b) / moves, pairs, float(acc.sum()) / max(1, int(mov.sum())), finite


def run_replica_batch(systems, moves=10, warmup=2, nmd=10, nstep_inner=5):
    """Batched HMC: every box's chain in one replica table, one captured move
    graph (velinit, forces, nmd RESPA steps of 1 fs with nstep_inner inner
    steps, Metropolis) against the same boxes run one after another, each
    its own engine and graph.  A move evaluates frclong nmd + 1 times."""
    import torch
    dev = torch.device("cuda:0")
    ms_b, pairs, acc, fin = _hmc_run(systems, moves, warmup, nmd, nstep_inner, dev)
    ms_seq, fin_s = 0.0, True
    for s in systems:
        ms1, _, _, f1 = _hmc_run([s], 2, 1, nmd, nstep_inner, dev)
        ms_seq += ms1
        fin_s = fin_s and f1
    evals = nmd + 1
    return {"workload": f"{len(systems)} n-butane boxes, 64-512 molecules each, "
                        f"{sum(s.natom for s in systems)} sites; HMC moves of {nmd} RESPA "
                        f"steps (1 fs / {nstep_inner}), one chain per box",
            "batched_ms_per_move": ms_b, "sequential_ms_per_move": ms_seq,
            "pairs_per_force_evaluation": int(pairs), "force_evaluations_per_move": evals,
            "batched_pair_interactions_per_s": evals * pairs / (ms_b * 1e-3),
            "sequential_pair_interactions_per_s": evals * pairs / (ms_seq * 1e-3),
            "batched_over_sequential": ms_seq / ms_b, "acceptance": acc,
            "finite": fin and fin_s}


def run_wolf_fit(nconf=64, calls=5, cpu=True):
    """The force-field fitting workload: nconf configurations of 1024 Wolf
    waters (kappa 0.2 / A, rc 12 A) in one replica table; one call gives
    every configuration's energy terms, forces and dE/d(epsilon, sigma,
    charge) per site in FP64 (mspms_energy_param_grad, host buffers: the
    positions go up and forces and derivatives come down in every call).
    CPU: the oracle's FP64 frclong of one configuration on one host core
    (a derivative by central differences would take two per parameter)."""
    from mspms2_amd import configs
    from mspms2_amd.engine import Engine
    from mspms2_amd.system import system_of
    systems = [system_of(rd) for rd in configs.wolf_water_set(nconf)]
    eng = Engine.replicas(systems, precision="fp64")
    x = eng.packed_positions()
    _, terms, _ = eng.param_grad(*x)
    t0 = time.perf_counter()
    for _ in range(calls):
        _, terms, g = eng.param_grad(*x)
    dt = (time.perf_counter() - t0) / calls
    eng.close()
    pairs = sum(t["n_elec_pairs"] for t in terms)
    out = {"workload": f"{nconf} x 1024 Wolf waters (kappa 0.2 / A, rc 12 A), energy + forces "
                       f"+ per-site dE/d(eps, sigma, q), FP64",
           "ms_per_call": 1e3 * dt, "configurations_per_s": nconf / dt,
           "pairs_per_configuration": pairs / nconf,
           "finite": bool(np.isfinite(g["charge"]).all())}
    if cpu:
        from oracle.oracle import Oracle
        o = Oracle(systems[0])
        o.frclong()
        t0 = time.perf_counter()
        o.frclong()
        tc = time.perf_counter() - t0
        out["cpu_port"] = {"seconds_per_configuration_forward": tc, "cores": 1,
                           "sample": "one frclong of configuration 0 (oracle restatement)"}
        out["speedup_vs_cpu_forward_only"] = tc * nconf / dt
    return out


TUBE_GRID_BYTES_PER_LOOKUP = 256 + 24 + 24   # 64 FP32 node values + position + force


def make_tube(nload=16):
    """The methane-in-tube workload: 16 loadings of 20,000-100,000 LJ
    methanes in a 100 x 100 x 4000 A box inside a 50 A smooth tube given as
    a 128 x 128 x 1024 FP32 tricubic Hermite grid (512 MiB) tabulated from
    the closed-form tube."""
    from mspms2_amd import configs, sfgrid
    from mspms2_amd.system import system_of
    h = 100.0 / 127.0
    grid = sfgrid.tube_grid([(3.73, 148.0)], R=50.0, x0=-50.0, y0=-50.0, nx=128, ny=128, hx=h,
                            hy=h, nz=1024, hz=4000.0 / 1023.0, sigma_s=3.37, eps_s=27.68)
    systems = [system_of(rd) for rd in configs.tube_loadings(nload)]
    for s in systems:
        s.grid = grid
    return systems


def run_tube(systems, steps=20, warmup=3, skin_a=1.0, cpu=True):
    """All loadings in one replica table, mixed, RESPA 2 fs outer / 4 inner
    steps with the wall's grid lookups in the inner loop (mspms_set_sf_split)
    and the pairs in the outer step; graph-replayed steps recording every
    step's terms.  The inner-loop kernel is the grid kernel: its roofline is
    HBM, 304 B per lookup (256 B of node values gathered, position, force)."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine, EngineError
    from mspms2_amd import abi
    dev = torch.device("cuda:0")
    nr = len(systems)
    ntot = sum(s.natom for s in systems)
    eng = Engine.replicas(systems, precision="mixed")
    eng.set_sf_split(True)
    eng.set_list_skin(skin_a / systems[0].params["sigma_base"])
    pos = torch.tensor(eng.packed_positions(), device=dev)
    vel = torch.tensor(np.concatenate([maxwell_velocities(s, seed=13) for s in systems], axis=1),
                       device=dev)
    fl = torch.tensor(eng.frclong()[0], device=dev)
    fs = torch.tensor(eng.frcshort()[0], device=dev)
    lt = torch.zeros(nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    stt = torch.zeros(nr * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(nr, dtype=torch.float64, device=dev)
    delt, nin = systems[0].delt, systems[0].run.nstep_inner
    st = torch.cuda.Stream()

    def phase(run):
        """Runs a phase of device steps; after a pair-list overflow
        (MSPMS_ENOMEM at the sync: the lists are widened and the positions
        and velocities are back at the start of the step that overflowed)
        the forces and terms are recomputed there and the phase runs again,
        as the library's overflow contract asks.  The loadings start from
        lattices and adsorb onto the wall within a few steps, so the longest
        pair-list row grows fast at first."""
        for attempt in range(4):
            run()
            torch.cuda.synchronize()
            try:
                eng.sync()
                return attempt
            except EngineError as e:
                if "ENOMEM" not in str(e) or attempt == 3:
                    raise
                progress(f"  pair-list overflow, lists widened; phase repeated ({e})")
                eng.frclong_device(pos, fl, lt)
                eng.frcshort_device(pos, fs, stt)
                torch.cuda.synchronize()
                eng.sync()

    phase(lambda: eng.vver_run_device(pos, vel, fl, fs, delt, nin, warmup, lt, stt, uk,
                                      graph=True, stream=st.cuda_stream))
    # launch shape (sites per cell x waves per i-cluster) by timing graph-
    # replayed steps (untimed, like the warmup); the loadings are far from
    # uniform (an adsorbed layer at the wall, a dilute core)
    best = None
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for cell in (3.0, 4.0, 6.0, 8.0):
        for wpc in (1, 2, 4):
            eng.set_launch_shape(cell, wpc)

            def trial():
                eng.vver_run_device(pos, vel, fl, fs, delt, nin, 4, lt, stt, uk, graph=True,
                                    stream=st.cuda_stream)
                e0.record(st)
                eng.vver_run_device(pos, vel, fl, fs, delt, nin, 8, lt, stt, uk, graph=True,
                                    stream=st.cuda_stream)
                e1.record(st)

            phase(trial)
            t = e0.elapsed_time(e1) / 8
            if best is None or t < best[0]:
                best = (t, cell, wpc)
    eng.set_launch_shape(best[1], best[2])
    phase(lambda: eng.vver_run_device(pos, vel, fl, fs, delt, nin, warmup, lt, stt, uk,
                                      graph=True, stream=st.cuda_stream))

    def eager():
        eng.set_profiling(True)
        for _ in range(steps):
            eng.vver_step_device(pos, vel, fl, fs, delt, nin, lt, stt, uk, st.cuda_stream)

    phase(eager)
    groups = eng.timings()
    eng.set_profiling(False)
    lh = torch.zeros(steps * nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed():
        a.record(st)
        eng.vver_run_device(pos, vel, fl, fs, delt, nin, steps, lt, stt, uk, lhist=lh,
                            graph=True, stream=st.cuda_stream)
        b.record(st)

    repeats = phase(timed)
    ms = a.elapsed_time(b) / steps
    raw = lt.cpu().numpy().tobytes()
    sz = LONG_TERMS_BYTES
    terms = [abi.LongTerms.from_buffer_copy(raw[r * sz:(r + 1) * sz]) for r in range(nr)]
    pairs = sum(t.n_lj_pairs for t in terms)
    finite = bool(torch.isfinite(pos).all().item()) and all(np.isfinite(t.uinter) for t in terms)
    eng.close()
    kms = groups.get("bonded", 0.0)
    gbytes = ntot * nin * TUBE_GRID_BYTES_PER_LOOKUP
    out = {"workload": f"{nr} methane loadings ({systems[0].natom}-{systems[-1].natom} sites, "
                       f"{ntot} in all) in a 50 A tube, 100 x 100 x 4000 A box, rc 12 A; "
                       f"128 x 128 x 1024 FP32 Hermite grid; RESPA {DT_FS:g} fs / {nin} with "
                       f"the grid in the inner loop",
           "ms_per_step": ms, "ns_per_day": DT_FS * 1e-6 * 86400e3 / ms,
           "launch_shape": {"cell_atoms": best[1], "waves_per_cluster": best[2]},
           "pairs_per_step": int(pairs), "pair_interactions_per_s": pairs / (ms * 1e-3),
           "grid_lookups_per_step": ntot * nin,
           "grid_kernel": {"name": "k_respa_inner_grid4 (the inner loop with the grid)",
                           "ms": kms, "bytes_per_launch": gbytes,
                           "achieved_GBps": gbytes / (kms * 1e-3) / 1e9 if kms > 0 else None,
                           "frac_of_hbm_peak": (gbytes / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS)
                           if kms > 0 else None},
           "groups_ms": groups, "finite": finite,
           "timed_region_repeats_after_overflow": repeats}
    if cpu:
        from oracle import oracle as O
        if O.reference_available():
            # the reference's frclong on the smallest loading with the
            # closed-form tube (iSF_type 1): the all-pairs loop plus
            # cal_sf_hypergeo per site, one core
            from mspms2_amd import configs
            rd = configs.tube_loading(systems[0].natom, seed=500)
            rd.run.iSF_type = 1
            with tempfile.TemporaryDirectory(prefix="mspms_bench_") as tmp:
                rd.write(tmp)
                d = O.run_reference(tmp, repeat=1, timeout=600)
            out["cpu_reference"] = {"seconds_per_frclong": d["frclong_seconds_per_call"],
                                    "cores": 1,
                                    "sample": f"one frclong of the {systems[0].natom}-site "
                                              f"loading, closed-form tube (iSF_type 1)"}
    return out


def run_lj_nvt(steps=1000, cpu=True):
    """The LJ-864 argon workload: FCC at rho* 0.8442, T* 0.722, rc* 2.5,
    Nose-Hoover NVT (vver_nh_3, qq = qqs = 100), dt* 0.005, `steps` steps
    graph-replayed on the device in each precision, against the
    reference's own integrator timed on one host core over the same steps
    (oracle/_ref, where it was built)."""
    import torch
    from mspms2_amd import configs
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
    from mspms2_amd.system import system_of
    rd = configs.lj_fcc(jitter=0.05, seed=7)
    s = system_of(rd)
    dev = torch.device("cuda:0")
    out = {"workload": "LJ-864 FCC argon, NVT vver_nh_3, rc* 2.5, dt* 0.005", "steps": steps}
    for prec in ("mixed", "fp64"):
        eng = Engine(s, precision=prec)
        eng.nvt_set(rd.run.qq, rd.run.qqs, s.treq, 3 * s.natom - rd.run.nconstraint)
        eng.set_list_skin(0.3)  # sigma units
        x0 = np.stack([s.xx, s.yy, s.zz])
        t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa
        pos, vel = t(x0), t(maxwell_velocities(s))
        fl, fs = t(eng.frclong()[0]), t(eng.frcshort()[0])
        lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        uk = torch.zeros(1, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        eng.nvt_run_device(pos, vel, fl, fs, s.delt, 1, 20, lt, st, uk, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.nvt_run_device(pos, vel, fl, fs, s.delt, 1, steps, lt, st, uk, graph=True)
        torch.cuda.synchronize()
        out[f"{prec}_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
        out[f"{prec}_tinst"] = eng.nvt_get()[0]["tinst"]
        eng.sync()
        eng.close()
    if cpu:
        from oracle import oracle as O
        if O.reference_available():
            with tempfile.TemporaryDirectory(prefix="mspms_bench_") as tmp:
                rd.write(tmp)
                d = O.run_reference(tmp, md_steps=steps, timeout=600)
            out["cpu_reference_ms_per_step"] = 1e3 * d["md_seconds_per_step"]
            out["cpu_cores"] = 1
            out["speedup_mixed_vs_cpu_reference"] = \
                out["cpu_reference_ms_per_step"] / out["mixed_ms_per_step"]
    return out


def run_sfgrid(h=0.1, cpu=True, cpu_h=None):
    """The reference's grid-generator workload: its own CH3 / (11,11)
    nanotube input (tools/CH3-SWNT11x11.in: 44 solid atoms, rcutoff 28 A,
    kept as tests/golden/sfgridgen_ch3_1A/grid.in) at its shipped 0.1 A
    interval.  GPU: the solid-fluid tables (bit-exact to the reference's
    generator) and the myinterp cube polynomials, each call timed whole
    (allocation, upload, kernels and the copy back to the host).  CPU: the
    reference's compiled generator (oracle/_ref/sfgridgen, built from
    tools/sfgridgen.c) on the same input at cpu_h (default: the same
    interval, about 13 s on one core at 0.1 A), per node."""
    cpu_h = h if cpu_h is None else cpu_h
    from mspms2_amd import myinterp as MI
    from mspms2_amd import sfgridgen as S
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                       "sfgridgen_ch3_1A", "grid.in")
    g = S.read_input(src)
    g.grid_itvl = h
    nx, ny, nz = S.dims(g)
    nn = nx * ny * nz
    S.generate(g)                      # first call: module load, pools
    t0 = time.perf_counter()
    S.generate(g)
    t_sfg = time.perf_counter() - t0
    t0 = time.perf_counter()
    nodes = MI.node_data(g)
    t_nodes = time.perf_counter() - t0
    t0 = time.perf_counter()
    MI.fit(nodes, (h, h, h))
    t_fit = time.perf_counter() - t0
    out = {"workload": f"sfgridgen CH3 / (11,11) SWNT input, {h} A interval, "
                       f"{nx}x{ny}x{nz} nodes",
           "nodes": nn, "sfgridgen_ms": 1e3 * t_sfg, "sfgridgen_nodes_per_s": nn / t_sfg,
           "myinterp_nodes_ms": 1e3 * t_nodes, "myinterp_fit_ms": 1e3 * t_fit,
           "cubes": (nx - 1) * (ny - 1) * (nz - 1)}
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "sfgridgen")
    if cpu and os.path.exists(exe):
        with tempfile.TemporaryDirectory(prefix="mspms_bench_") as tmp:
            lines = open(src).read().split("\n")
            lines[1] = " ".join([repr(cpu_h)] + lines[1].split()[1:])
            inp = os.path.join(tmp, "grid.in")
            with open(inp, "w") as f:
                f.write("\n".join(lines))
            g2 = S.read_input(inp)
            n2 = int(np.prod(S.dims(g2)))
            t0 = time.perf_counter()
            subprocess.run([exe, inp], cwd=tmp, check=True, capture_output=True, timeout=600)
            t_cpu = time.perf_counter() - t0
        out["cpu_reference"] = {"nodes_per_s": n2 / t_cpu, "cores": 1,
                                "sample": f"the same input at {cpu_h} A ({n2} nodes)"}
        out["speedup_vs_cpu_reference"] = out["sfgridgen_nodes_per_s"] / (n2 / t_cpu)
    return out


def run_reference_cpu(rd, repeat=2):
    """The reference's own frclong timed on one host core (oracle/_ref)."""
    from oracle import oracle as O
    with tempfile.TemporaryDirectory(prefix="mspms_bench_") as tmp:
        rd.write(tmp)
        if O.reference_available():
            d = O.run_reference(tmp, repeat=repeat, timeout=600)
            return dict(kind="reference", seconds_per_frclong=d["frclong_seconds_per_call"],
                        sample=f"{repeat} frclong calls of the 12,288-site system "
                               f"(oracle/_ref/mspms_ref_harness, reference sources)")
    # restatement fallback (same loop order, FP64)
    from mspms2_amd import runio
    from mspms2_amd.system import System
    s = System.from_rundir(rd)
    o = O.Oracle(s)
    t = time.perf_counter()
    for _ in range(repeat):
        o.frclong()
    return dict(kind="port", seconds_per_frclong=(time.perf_counter() - t) / repeat,
                sample=f"{repeat} frclong calls of the oracle restatement")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--fp64-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skin", type=float, default=5.0,
                    help="pair-list buffer in Angstrom (lists reused until a site moves half of it)")
    ap.add_argument("--skin-fp64", type=float, default=2.5,
                    help="pair-list buffer in Angstrom of the FP64 leg")
    ap.add_argument("--replicas", type=int, default=256,
                    help="boxes in the batched-replica leg (0: skip it)")
    ap.add_argument("--tube-loadings", type=int, default=16,
                    help="loadings of the methane-in-tube leg (0: skip it)")
    ap.add_argument("--tube-only", action="store_true",
                    help="run only the methane-in-tube leg and print its record")
    ap.add_argument("--wolf-configs", type=int, default=64,
                    help="configurations of the Wolf-water fitting leg (0: skip it)")
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="NVE steps of the energy-drift leg, per precision (0: skip it)")
    ap.add_argument("--equil-steps", type=int, default=2500,
                    help="velocity-rescaled steps that equilibrate the drift leg's start "
                         "(0: lattice start)")
    ap.add_argument("--lj-steps", type=int, default=1000,
                    help="steps of the LJ-864 NVT leg (0: skip it)")
    ap.add_argument("--sfgrid", type=float, default=0.1,
                    help="grid-generator leg at this interval (A); 0 skips it")
    ap.add_argument("--sfgrid-only", action="store_true",
                    help="run only the grid-generator leg and print its record")
    ap.add_argument("--probe", action="store_true",
                    help="only the short frclong + reference-copy workload for counter runs")
    ap.add_argument("--trace-steps", type=int, default=0,
                    help="only this many graph-replayed mixed steps, for kernel-trace runs")
    ap.add_argument("--shape", nargs=2, type=float, metavar=("CELL_ATOMS", "WPC"),
                    help="with --trace-steps: this launch shape instead of the autotuner's")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the per-kernel HBM traffic of two --pmc passes over --probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU per process; --gpus {args.gpus} is not "
              f"supported (replica batches shard across GPUs by replica)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary), indent=1))
        return
    if args.sfgrid_only:
        print(json.dumps(run_sfgrid(args.sfgrid, cpu=not args.no_cpu)))
        return
    if args.tube_only:
        print(json.dumps(run_tube(make_tube(args.tube_loadings), cpu=not args.no_cpu)))
        return
    rd, system = make_workload()
    if args.probe:
        run_probe(system)
        return
    if args.trace_steps > 0:
        run_trace(system, args.trace_steps, args.skin / system.params["sigma_base"], args.shape)
        return
    skin = args.skin / system.params["sigma_base"]
    progress("12,288-site water, mixed")
    mixed = run_gpu(system, "mixed", args.steps, args.warmup, skin=skin)
    progress("12,288-site water, FP64")
    # the exact kernel evaluates every entry of its lists (no pruning), so
    # it takes the shorter buffer
    fp64 = run_gpu(system, "fp64", args.fp64_steps, min(args.warmup, 5), profile=False,
                   skin=args.skin_fp64 / system.params["sigma_base"])

    ms = mixed["ms_per_step"]
    ns_day = DT_FS * 1e-6 * 86400e3 / ms
    t = mixed["terms"]
    n_el, n_lj = t["n_elec_pairs"], t["n_lj_pairs"]
    pairs_per_s = n_el / (ms * 1e-3)
    # force error of the timed mixed path (its launch shape and reused list,
    # the forces its last timed step computed) against the FP64 engine on
    # the same positions
    from mspms2_amd.engine import Engine
    pa, fm = mixed["pos_after"], mixed["f_after"]
    e64 = Engine(system, precision="fp64")
    ff, _ = e64.frclong(pa[0], pa[1], pa[2])
    e64.close()
    force_err = float(np.linalg.norm(fm - ff) / np.linalg.norm(ff))

    g = mixed["groups"]
    natom = system.natom
    nk = 2108
    # algorithmic work per launch (pair: 20 geometry + 24 LJ + 25 Ewald real
    # FLOP-eq per i<j pair in range; reciprocal: 29 FP64 FLOP-eq per (k, atom)
    # over the 1054 distinct +-k pairs: S(-k) = conj S(k) for real charges, so
    # the other half of the reference's 2108 vectors is no work)
    pair_flop = 20.0 * n_el + 24.0 * n_lj + 25.0 * n_el
    recip_flop = 29.0 * (nk // 2) * natom
    wpc = mixed["tune"]["waves_per_cluster"]
    cand = {
        "k_pair": (g.get("pair", 0.0), pair_flop, PEAK_FP32_TFLOPS,
                   ["k_pair_mx<1, %d, true, false>" % wpc]),
        "k_sk+k_sred+k_kforce": (g.get("recip_s", 0.0) + g.get("recip_f", 0.0), recip_flop,
                                 PEAK_FP64_TFLOPS, ["k_sk<", "k_sred", "k_kforce<"]),
    }
    def roof(name):
        kms, flop, peak, knames = cand[name]
        achieved = flop / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
        # HBM bytes per launch of the same kernels from the committed
        # counter summary (two rocprofv3 --pmc passes over `bench.py --probe`)
        traffic = traffic_of(knames)
        return {"bound": "valu", "kernel": name, "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                "traffic_unit": "bytes per launch", "flop_per_launch": flop,
                "hbm_GBps_at_traffic": (traffic / (kms * 1e-3) / 1e9) if traffic else None,
                "kernel_ms": kms}

    # the pair kernel is the step's largest kernel on the main stream (the
    # reciprocal group runs beside it on the side stream)
    roofline = dict(roof("k_pair"), groups_ms=g)
    roofline_recip = roof("k_sk+k_sred+k_kforce")

    cpu = None
    force_err_oracle = None
    if not args.no_cpu:
        progress("reference frclong on the host")
        c = run_reference_cpu(rd)
        cpu_step = c["seconds_per_frclong"]
        cpu = {"value": DT_FS * 1e-6 * 86400.0 / cpu_step, "unit": "ns/day", "cores": 1,
               "kind": c["kind"], "sample": c["sample"],
               "seconds_per_frclong": cpu_step}
        # the CPU leg's checker: the oracle's forces on the positions the
        # timed steps ended at, against the timed path's own forces
        from oracle.oracle import Oracle
        fo, _ = Oracle(system).frclong(pa[0], pa[1], pa[2])
        force_err_oracle = float(np.linalg.norm(fm - fo) / np.linalg.norm(fo))

    line = {
        "metric": "ns/day + pair-interactions/s, 12k-site water; batched/FP32-acc vs plain FP64",
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "wall_ms_per_step": mixed["wall_ms_per_step"],
        "eager_ms_per_step": mixed["eager_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs / f64 accumulate",
        "data": "synthetic (seeded jittered-lattice water, random orientations)",
        "config": {"workload": "water-12288 Ewald RESPA 2 fs / 8 x 0.25 fs", "natom": natom,
                   "kvectors": nk, "kvectors_distinct": nk // 2, "rc_A": 10.0, "kappa_per_A": 0.28, "box_A": 49.66,
                   "precision": "mixed",
                   "launch_shape": {k: mixed["tune"][k] for k in ("cell_atoms",
                                                                  "waves_per_cluster")},
                   "list_skin_A": args.skin, "fp64_list_skin_A": args.skin_fp64},
        "list_builds_per_step": mixed["list_builds_per_step"],
        "autotune_candidates_ms": mixed["tune"]["candidates"],
        "dropin_frclong_ms_pcie_inclusive": mixed["host_frclong_ms"],
        "pair_interactions_per_s": pairs_per_s,
        "pairs_in_cutoff": n_el,
        "fp64_ns_per_day": DT_FS * 1e-6 * 86400e3 / fp64["ms_per_step"],
        "fp64_ms_per_step": fp64["ms_per_step"],
        "mixed_over_fp64": fp64["ms_per_step"] / ms,
        "force_rms_rel_err_vs_fp64": force_err,
        "force_rms_rel_err_vs_oracle": force_err_oracle,
        "force_err_positions": "after the timed steps, timed launch shape and list state",
        "per_step_terms": "every timed step reduces and records its terms (device histories)",
        "ms_per_step_last_step_terms_only": mixed["ms_per_step_last_terms_only"],
        "ns_per_day_last_step_terms_only":
            DT_FS * 1e-6 * 86400e3 / mixed["ms_per_step_last_terms_only"],
        "finite": mixed["finite"] and fp64["finite"],
        "roofline": roofline,
        "roofline_recip_side_stream": roofline_recip,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_reference"] = ns_day / cpu["value"]
    if args.drift_steps > 0:
        side_leg(line, "energy_drift", "energy drift",
                 lambda: drift_leg(system, args.equil_steps, args.drift_steps, skin))
    if args.replicas > 0:
        side_leg(line, "replicas", "batched HMC replicas",
                 lambda: run_replica_batch(make_replicas(args.replicas)))
    if args.wolf_configs > 0:
        side_leg(line, "wolf_fit", "Wolf fitting set",
                 lambda: run_wolf_fit(args.wolf_configs, cpu=not args.no_cpu))
    if args.tube_loadings > 0:
        side_leg(line, "tube", "methane-in-tube loadings",
                 lambda: run_tube(make_tube(args.tube_loadings), cpu=not args.no_cpu))
    if args.lj_steps > 0:
        side_leg(line, "lj864_nvt", "LJ-864 NVT",
                 lambda: run_lj_nvt(args.lj_steps, cpu=not args.no_cpu))
    if args.sfgrid > 0:
        side_leg(line, "sfgrid", "grid generators",
                 lambda: run_sfgrid(args.sfgrid, cpu=not args.no_cpu))
    print(json.dumps(line))


def drift_leg(system, equil_steps, drift_steps, skin):
    """Both precisions from one equilibrated state (velocity-rescaled RESPA
    steps first), so that the drift shows the integrator and the precision,
    not relaxation of the lattice start."""
    start = None
    drift = {}
    if equil_steps > 0:
        p_eq, v_eq, t_eq = equilibrate(system, equil_steps, skin)
        start = (p_eq, v_eq)
        drift["start"] = (f"after {equil_steps} RESPA steps "
                          f"({DT_FS * equil_steps / 1000:g} ps) with velocities "
                          f"rescaled to {system.treq:g} K every 100 steps "
                          f"(T = {t_eq:.1f} K before the last rescale)")
    else:
        drift["start"] = "jittered lattice (not equilibrated)"
    for p in ("mixed", "fp64"):
        drift[p] = run_drift(system, p, drift_steps, skin, start)
    return drift


def side_leg(line, key, what, fn):
    """A measurement reported beside the headline: an error in it is
    recorded under its key (and on stderr) instead of ending the run, so the
    headline line is still printed."""
    progress(what)
    try:
        line[key] = fn()
    except Exception as e:  # noqa: BLE001 -- any failure of a side leg is reported
        line[key] = {"error": f"{type(e).__name__}: {e}"}
        progress(f"  {what} failed: {type(e).__name__}: {e}")


if __name__ == "__main__":
    main()
