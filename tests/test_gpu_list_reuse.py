"""Buffered pair lists (mspms_set_list_skin).

A list built with the cutoff + skin is reused while every site stays within
skin / 2 of its position at the build.  Every call must give what a fresh
evaluation gives: the oracle's pair counts and pair-set hashes bit for bit in
FP64 mode, energies and forces within 1e-10 (a reused row keeps the build's
order, so only the summation order changes); forces within 1e-5 in mixed
mode, energies within 5e-5 there: the random moves below push sites into
close contacts (repulsive pair terms of 10^3-10^4 K in the butane box) whose
FP32 rounding, amplified six times by r^-12, shifts uvdw by ~1e-5 of its
total on a fresh list as much as on a reused one.

Each case runs one sequence of frclong calls:
  1    the starting positions (first build)
  2-4  random moves of every site, at most 0.9 * skin / 2 from the positions
       of call 1 (reuse; 2-10 % of the sites cross a box face on the way)
  5    the positions of call 1 with one listed (not fully ghost) site moved
       by 0.6 * skin (rebuild)
  6    one molecule of call 5 shifted by a whole box length along x: the
       same configuration with other unwrapped coordinates (rebuild)
  7    the positions of call 6 again (reuse, zero displacement)
and checks the build count after each call against the rule applied on the
host.
"""
import numpy as np
import pytest

from mspms2_amd import abi, configs
from mspms2_amd.engine import Engine
from oracle.oracle import Oracle
from tests.util import assert_long_close, force_close, rel_rms, system_of

pytestmark = pytest.mark.gpu

SKIN_A = 1.0  # skin in Angstrom

CASES = {
    "water512_ewald_k8": lambda: configs.water_box(n_side=8, L=8 * 3.104, rc=10.0, kappa=0.28,
                                                   kmax=8, ksqmax=64, seed=11),
    "mixture_ewald_ghost": lambda: configs.mixture_box(water4=True, seed=12),
    "butane64": lambda: configs.butane_box(nmol=64, seed=5),
    "water216_wolf": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.2,
                                               elec="wolf", seed=4),
}

_cache = {}


def _sequence(name):
    """[(xx, yy, zz, builds expected after the call, oracle (f, t))]"""
    if name in _cache:
        return _cache[name]
    s = system_of(CASES[name]())
    skin = SKIN_A / s.params["sigma_base"]
    rng = np.random.default_rng(23)
    p0 = np.stack([s.xx, s.yy, s.zz])
    # translate the whole system (energies and forces are unchanged) so that
    # site 0 sits 0.1 skin inside the lower face of each axis: the sites of
    # its lattice planes then cross the faces under the moves below
    L = np.array([s.params["boxlx"], s.params["boxly"], s.params["boxlz"]])[:, None]
    p0 = p0 + (0.1 * skin - np.mod(p0[:, :1], L))
    seq = [(p0, 1)]
    for _ in range(3):
        # uniform direction, radius up to 0.9 * skin / 2
        d = rng.normal(size=p0.shape)
        d /= np.linalg.norm(d, axis=0)
        d *= 0.45 * skin * rng.uniform(0.0, 1.0, size=p0.shape[1]) ** (1 / 3)
        seq.append((p0 + d, 1))
    # per-site full-ghost flags (such sites are never in the list)
    full = np.concatenate([np.tile(sp.ghost_type == abi.GHOST_FULL, sp.nmole)
                           for sp in s.species])
    p = p0.copy()
    p[0, np.flatnonzero(~full)[7]] += 0.6 * skin
    seq.append((p, 2))
    q = p.copy()
    q[0, :s.species[0].natom] += s.params["boxlx"]  # the first molecule's sites
    seq.append((q, 3))
    seq.append((q.copy(), 3))
    # the rule itself: rebuild when a listed site is farther than skin / 2
    # from its position at the last build
    built, nb_rule = None, 0
    for pos, nb in seq:
        if built is None or np.sqrt(((pos - built)[:, ~full] ** 2).sum(axis=0)).max() > 0.5 * skin:
            built, nb_rule = pos, nb_rule + 1
        assert nb_rule == nb
    o = Oracle(s)
    out = []
    for pos, nb in seq:
        f, t = o.frclong(pos[0], pos[1], pos[2])
        out.append((pos, nb, f, t))
    _cache[name] = (s, skin, out)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_reuse_fp64_matches_oracle(name):
    s, skin, seq = _sequence(name)
    print(f"case {name} fp64 natom={s.natom} skin={skin:.4f}", flush=True)
    eng = Engine(s, precision="fp64")
    eng.set_list_skin(skin)
    for k, (pos, nb, fl, tl) in enumerate(seq):
        f, t = eng.frclong(pos[0], pos[1], pos[2])
        what = f"{name} call {k + 1}"
        assert eng.list_builds() == nb, what
        for key in ("n_lj_pairs", "n_elec_pairs", "lj_pair_hash", "elec_pair_hash"):
            assert t[key] == tl[key], (what, key)
        assert_long_close(t, tl, 1e-10, what)
        ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-10)
        assert ok, (what, err)
    eng.close()


@pytest.mark.parametrize("name", list(CASES))
def test_reuse_mixed_matches_oracle(name):
    s, skin, seq = _sequence(name)
    print(f"case {name} mixed natom={s.natom}", flush=True)
    eng = Engine(s, precision="mixed")
    eng.set_list_skin(skin)
    for k, (pos, nb, fl, tl) in enumerate(seq):
        f, t = eng.frclong(pos[0], pos[1], pos[2])
        what = f"{name} call {k + 1}"
        assert eng.list_builds() == nb, what
        ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-5)
        assert ok, (what, err)
        assert_long_close(t, tl, 5e-5, what)
        assert abs(t["n_elec_pairs"] - tl["n_elec_pairs"]) <= max(2, 1e-5 * tl["n_elec_pairs"])
    eng.close()


def test_reused_list_equals_fresh_list_mixed():
    """The mixed path on a reused list against the same positions on a
    fresh list: only the FP32 offsets differ (build-time wrapped position
    minus centre plus the displacement, against the current wrapped
    position minus the new centre), so the forces agree to FP32 rounding."""
    s, skin, seq = _sequence("water512_ewald_k8")
    a = Engine(s, precision="mixed")
    a.set_list_skin(skin)
    b = Engine(s, precision="mixed")
    for pos, _, _, _ in seq[:4]:
        fa, ta = a.frclong(pos[0], pos[1], pos[2])
        fb, tb = b.frclong(pos[0], pos[1], pos[2])
        assert rel_rms(fa, fb) <= 2e-6
    assert a.list_builds() == 1 and b.list_builds() == 4
    a.close()
    b.close()


def test_graph_run_with_skin_matches_rebuild_every_step():
    """Device RESPA steps replayed from a graph with a buffered list against
    the same steps rebuilding every step (skin 0), FP64."""
    import torch
    from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES
    from tests.test_gpu_replicas import _state
    dev = torch.device("cuda:0")
    s = system_of(configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3, kmax=5,
                                    ksqmax=27, seed=3, nstep_inner=4))
    skin = SKIN_A / s.params["sigma_base"]
    out = {}
    nsteps = 12
    for label, sk in (("fresh", 0.0), ("buffered", skin)):
        eng = Engine(s, precision="fp64")
        eng.set_list_skin(sk)
        pos = torch.tensor(np.stack([s.xx, s.yy, s.zz]), dtype=torch.float64, device=dev)
        vel = torch.tensor(_state(s, 5), dtype=torch.float64, device=dev)
        fl = torch.tensor(eng.frclong()[0], device=dev)
        fs = torch.tensor(eng.frcshort()[0], device=dev)
        lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
        uk = torch.zeros(1, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        eng.vver_run_device(pos, vel, fl, fs, s.delt, 4, nsteps, lt, st, uk, graph=True)
        torch.cuda.synchronize()
        eng.sync()
        out[label] = (pos.cpu().numpy(), fl.cpu().numpy(), lt.cpu().numpy(), eng.list_builds())
        print(f"{label}: builds {out[label][3]}", flush=True)
        eng.close()
    pf, ff, lf, nf = out["fresh"]
    pb, fb, lb, nb = out["buffered"]
    assert nb < nf
    assert rel_rms(pb, pf) <= 1e-12
    assert rel_rms(fb, ff) <= 1e-9
    # the pair counts and hashes in the terms, read as doubles, must be the
    # same bits (a hash can read as NaN)
    assert np.allclose(lb, lf, rtol=1e-9, atol=0, equal_nan=True)
