"""GPU parity at the 12,288-site water workload and at every launch shape.

The workload: 4096 flexible three-site waters (12,288 sites) in a 49.66 A
box, Ewald alpha 0.28 / A, kmax 8, KSQMAX 64 (2,108 k-vectors), rc 10 A,
seeded as the committed fixture tests/golden/water12288_ewald (its run files
are regenerated here and checked by SHA-256).

The autotuner may pick any of 16 launch shapes (sites per binning cell 6, 8,
12, 16 x waves per i-cluster row 1, 2, 4, 8), and each shape is a different
binning and a different instantiation of the pair kernel, so every shape is
held to the bar:
  FP64   pair counts and pair-set hashes equal to the oracle's; energies
         within 1e-10 relative (uelec / uinter / virial relative to the
         cancelling Ewald components); forces ||dF|| / ||F|| <= 1e-10; the
         reference's own energies and 1,024 sampled forces (the fixture)
         within the same bar
  mixed  forces ||dF|| / ||F|| <= 1e-5, energies within 1e-5
Graph-replayed RESPA steps with the 2.5 A list buffer are compared with the
same steps rebuilding the list every step (FP64: positions 1e-12, forces
1e-9; mixed: forces 2e-6 between the two), and the mixed forces after the
steps with the oracle on the same positions (1e-5).
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

from mspms2_amd import configs, runio
from mspms2_amd.engine import Engine
from mspms2_amd.system import System
from oracle.oracle import Oracle
from tests.util import GOLDEN, assert_long_close, force_close, rel_rms, system_of

pytestmark = pytest.mark.gpu

SHAPES = [(c, w) for c in (6.0, 8.0, 12.0, 16.0) for w in (1, 2, 4, 8)]
SKIN_A = 2.5
BENCH_SKIN_A = 5.0
FIXTURE = "water12288_ewald"

_cache = {}


def _water12288():
    """(system, oracle forces, oracle terms, fixture scalars, fixture force
    sample, sample indices)."""
    if "w" in _cache:
        return _cache["w"]
    rd = configs.water_box(n_side=16, L=49.66, kappa=0.28, kmax=8, ksqmax=64, rc=10.0, seed=2)
    with open(os.path.join(GOLDEN, FIXTURE, "runfiles.sha256")) as f:
        want = f.read().strip()
    with tempfile.TemporaryDirectory() as tmp:
        rd.write(tmp)
        h = hashlib.sha256()
        for fn in ("in.mspms", "cfg.mspms", "xyz.mspms"):
            with open(os.path.join(tmp, fn), "rb") as f:
                h.update(f.read())
        assert h.hexdigest() == want, "run files differ from the fixture's"
        s = System.from_rundir(runio.RunDir.read(tmp))
    d = np.load(os.path.join(GOLDEN, FIXTURE, "ref.npz"))
    sc = {k[2:]: float(d[k]) for k in d.files if k.startswith("s_")}
    print(f"oracle frclong on {s.natom} sites", flush=True)
    fl, tl = Oracle(s).frclong()
    _cache["w"] = (s, fl, tl, sc, d["fl"], d["idx"])
    return _cache["w"]


def _water512():
    if "w512" not in _cache:
        s = system_of(configs.water_box(n_side=8, L=8 * 3.104, rc=10.0, kappa=0.28, kmax=8,
                                        ksqmax=64, seed=11))
        fl, tl = Oracle(s).frclong()
        _cache["w512"] = (s, fl, tl)
    return _cache["w512"]


def _check_fp64(s, f, t, fl, tl, what):
    for key in ("n_lj_pairs", "n_elec_pairs", "lj_pair_hash", "elec_pair_hash"):
        assert t[key] == tl[key], (what, key, t[key], tl[key])
    assert_long_close(t, tl, 1e-10, what)
    ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-10)
    assert ok, (what, err)


def _check_mixed(s, f, t, fl, tl, what):
    ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-5)
    print(f"  {what}: force rel rms {err:.3e}", flush=True)
    assert ok, (what, err)
    assert_long_close(t, tl, 1e-5, what)
    assert abs(t["n_elec_pairs"] - tl["n_elec_pairs"]) <= max(2, 1e-5 * tl["n_elec_pairs"])


def test_oracle_equals_fixture():
    """The live oracle on the regenerated run files equals the reference's
    own dump (energies bit for bit, sampled forces bit for bit)."""
    s, fl, tl, sc, fs_ref, idx = _water12288()
    for k in ("uvdw", "ureal", "uexcl", "ufourier", "uself", "uelec", "uinter",
              "virial_inter"):
        assert tl[k] == sc[k], k
    assert np.array_equal(fl[:, idx], fs_ref)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda x: f"{x[0]:g}x{x[1]}" if isinstance(x, tuple) else str(x))
def test_water12288_fp64_every_shape(shape):
    s, fl, tl, sc, fs_ref, idx = _water12288()
    what = f"water12288 fp64 shape {shape}"
    print(what, flush=True)
    eng = Engine(s, precision="fp64")
    eng.set_launch_shape(*shape)
    f, t = eng.frclong()
    eng.close()
    _check_fp64(s, f, t, fl, tl, what)
    # the reference's own numbers (the fixture) at the same bar
    assert_long_close(t, {**tl, **sc}, 1e-10, what + " vs fixture")
    assert np.linalg.norm(f[:, idx] - fs_ref) <= 1e-10 * np.linalg.norm(fs_ref)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda x: f"{x[0]:g}x{x[1]}" if isinstance(x, tuple) else str(x))
def test_water12288_mixed_every_shape(shape):
    s, fl, tl, sc, fs_ref, idx = _water12288()
    what = f"water12288 mixed shape {shape}"
    print(what, flush=True)
    eng = Engine(s, precision="mixed")
    eng.set_launch_shape(*shape)
    f, t = eng.frclong()
    # a second call on the same list-free path repeats the first to the
    # reordering of the FP64 force atomics
    f2, t2 = eng.frclong()
    eng.close()
    _check_mixed(s, f, t, fl, tl, what)
    assert rel_rms(f2, f) <= 1e-13
    assert t2["uinter"] == t["uinter"]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda x: f"{x[0]:g}x{x[1]}" if isinstance(x, tuple) else str(x))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_water512_every_shape(prec, shape):
    s, fl, tl = _water512()
    what = f"water512 {prec} shape {shape}"
    eng = Engine(s, precision=prec)
    eng.set_launch_shape(*shape)
    f, t = eng.frclong()
    eng.close()
    (_check_fp64 if prec == "fp64" else _check_mixed)(s, f, t, fl, tl, what)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda x: f"{x[0]:g}x{x[1]}" if isinstance(x, tuple) else str(x))
def test_list_reuse_every_shape(shape):
    """The buffered-list call sequence of the list-reuse tests (moves within
    the buffer, a rebuild, a box-length shift) on 512 waters at each shape,
    mixed, against the oracle."""
    from tests.test_gpu_list_reuse import _sequence
    s, skin, seq = _sequence("water512_ewald_k8")
    eng = Engine(s, precision="mixed")
    eng.set_launch_shape(*shape)
    eng.set_list_skin(skin)
    for k, (pos, nb, fl, tl) in enumerate(seq):
        f, t = eng.frclong(pos[0], pos[1], pos[2])
        what = f"shape {shape} call {k + 1}"
        assert eng.list_builds() == nb, what
        ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-5)
        assert ok, (what, err)
        assert_long_close(t, tl, 5e-5, what)
    eng.close()


def _respa_run(s, prec, shape, skin_a, nsteps, seed=5):
    """nsteps graph-replayed RESPA outer steps (8 inner steps) from the
    fixture's positions; returns (positions, long forces, long terms
    buffer, list builds)."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES
    from tests.test_gpu_replicas import _state
    dev = torch.device("cuda:0")
    eng = Engine(s, precision=prec)
    eng.set_launch_shape(*shape)
    eng.set_list_skin(skin_a / s.params["sigma_base"])
    pos = torch.tensor(np.stack([s.xx, s.yy, s.zz]), dtype=torch.float64, device=dev)
    vel = torch.tensor(_state(s, seed), dtype=torch.float64, device=dev)
    fl = torch.tensor(eng.frclong()[0], device=dev)
    fs = torch.tensor(eng.frcshort()[0], device=dev)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    b0 = eng.list_builds()
    # the fixture's run input: 2 fs outer step, 8 inner steps
    eng.vver_run_device(pos, vel, fl, fs, s.delt, s.run.nstep_inner, nsteps, lt, st, uk,
                        graph=True)
    torch.cuda.synchronize()
    eng.sync()
    out = (pos.cpu().numpy(), fl.cpu().numpy(), lt.cpu().numpy(), eng.list_builds() - b0)
    eng.close()
    return out


@pytest.mark.parametrize("skin", [SKIN_A, BENCH_SKIN_A], ids=lambda x: f"skin{x:g}")
def test_graph_skin_fp64_matches_rebuild_every_step(skin):
    """FP64, launch shape (6 sites per cell, 2 waves): 12 steps with a
    2.5 A or 5 A buffer (the bench's) against 12 steps rebuilding every
    step."""
    s = _water12288()[0]
    pf, ff, lf, nf = _respa_run(s, "fp64", (6.0, 2), 0.0, 12)
    pb, fb, lb, nb = _respa_run(s, "fp64", (6.0, 2), skin, 12)
    print(f"builds: fresh {nf}, buffered {nb}", flush=True)
    assert nb < nf
    assert rel_rms(pb, pf) <= 1e-12
    assert rel_rms(fb, ff) <= 1e-9
    assert np.allclose(lb, lf, rtol=1e-9, atol=0, equal_nan=True)


@pytest.mark.parametrize("shape,skin", [((6.0, 2), SKIN_A), ((8.0, 4), SKIN_A),
                                        ((12.0, 8), SKIN_A), ((16.0, 1), SKIN_A),
                                        ((12.0, 4), BENCH_SKIN_A)],
                         ids=lambda x: f"{x[0]:g}x{x[1]}" if isinstance(x, tuple) else f"skin{x:g}")
def test_graph_skin_mixed_matches_oracle(shape, skin):
    """Mixed, 2.5 A buffer (and the bench's 5 A at its usual shape),
    graph-replayed: after 12 steps the device's long forces and terms equal
    the oracle's on the same positions (1e-5), and equal a run rebuilding
    every step to FP32 rounding."""
    from mspms2_amd import abi
    s = _water12288()[0]
    pb, fb, lb, nb = _respa_run(s, "mixed", shape, skin, 12)
    pf, ff, lf, nf = _respa_run(s, "mixed", shape, 0.0, 12)
    print(f"shape {shape}: builds fresh {nf}, buffered {nb}", flush=True)
    assert nb < nf
    x0 = np.stack([s.xx, s.yy, s.zz])
    assert rel_rms(pb - x0, pf - x0) <= 1e-5
    fo, to = Oracle(s).frclong(pb[0], pb[1], pb[2])
    t = abi.LongTerms.from_buffer_copy(lb.view(np.uint8).tobytes()).as_dict()
    _check_mixed(s, fb, t, fo, to, f"shape {shape} after 12 buffered steps")
