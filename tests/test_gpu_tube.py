"""GPU parity on the methane-in-tube workload at its real size.

16 loadings of 20,000 to 100,000 LJ methanes (3.73 A, 148 K, rc 12 A) in a
100 x 100 x 4000 A periodic box, inside a smooth tube of radius 50 A given
as a tricubic Hermite grid of 128 x 128 x 1024 nodes (FP32, 8 values per
node: 512 MiB) tabulated from the closed-form tube (sigma_sf 3.55 A,
eps_sf 64 K by Lorentz-Berthelot with the carbon wall's 3.37 A, 27.68 K).

At this size the oracle's all-pairs frclong is out of reach, so the pair
part is pinned by an exact restatement in numpy: candidate pairs from a
periodic k-d tree, then the reference's own arithmetic on them (minimum
image d - L rint(d / L), r^2 summed x, y, z, the cutoff test r^2 < rc^2),
so the (i < j) pair set and its splitmix64 hash are the reference's; LJ
energies, virial and forces summed in numpy.  The wall is pinned by the
oracle's tricubic Hermite lookup, site by site.
  FP64   pair count and hash equal; uvdw, virial, usflj within 1e-10
         relative; forces ||dF|| / ||F|| <= 1e-10; the pair forces sum to
         zero (1e-12 of their norm)
  mixed  forces within 1e-5
The RESPA split (mspms_set_sf_split) moves the wall's forces from frclong to
frcshort: each side is checked on its own, and one device RESPA step with
the wall in the inner loop (FP64) equals its numpy restatement (positions
and velocities within 1e-10 relative).  All 16 loadings run as one replica
table.
"""
import numpy as np
import pytest

from mspms2_amd import abi, configs, sfgrid
from mspms2_amd.engine import LONG_TERMS_BYTES, SHORT_TERMS_BYTES, Engine
from mspms2_amd.system import system_of
from tests.tube_ref import lj_pairs, wall

pytestmark = pytest.mark.gpu

_cache = {}


def tube_grid():
    if "grid" not in _cache:
        h = 100.0 / 127.0
        _cache["grid"] = sfgrid.tube_grid([(3.73, 148.0)], R=50.0, x0=-50.0, y0=-50.0, nx=128,
                                          ny=128, hx=h, hy=h, nz=1024, hz=4000.0 / 1023.0,
                                          sigma_s=3.37, eps_s=27.68)
    return _cache["grid"]


def loading(k):
    key = ("loading", k)
    if key not in _cache:
        rds = configs.tube_loadings()
        s = system_of(rds[k])
        s.grid = tube_grid()
        _cache[key] = s
    return _cache[key]


def ref_parts(k):
    key = ("ref", k)
    if key not in _cache:
        s = loading(k)
        print(f"loading {k}: {s.natom} sites, numpy pairs + oracle wall", flush=True)
        _cache[key] = (lj_pairs(s), wall(s))
    return _cache[key]


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("k", [0, 15])
def test_loading_fp64_vs_reference_arithmetic(k):
    s = loading(k)
    (npair, h, uvdw, vir, Fp), (usf, Fw) = ref_parts(k)
    assert np.abs(Fp.sum(axis=1)).max() <= 1e-12 * np.linalg.norm(Fp)
    eng = Engine(s, precision="fp64")
    f, t = eng.frclong()
    print(f"  {npair} pairs; engine {t['n_lj_pairs']}", flush=True)
    assert t["n_lj_pairs"] == npair
    assert t["lj_pair_hash"] == h
    assert t["uvdw"] == pytest.approx(uvdw, rel=1e-10)
    assert t["virial_inter"] == pytest.approx(vir, rel=1e-10)
    assert t["usflj"] == pytest.approx(usf, rel=1e-10)
    assert t["uinter"] == pytest.approx(uvdw + usf, rel=1e-10)
    assert _rel(f, Fp + Fw) <= 1e-10
    # the split: the wall leaves frclong for frcshort
    eng.set_sf_split(True)
    f1, t1 = eng.frclong()
    fs, ts = eng.frcshort()
    assert t1["usflj"] == 0.0 and _rel(f1, Fp) <= 1e-10
    assert ts["usflj"] == pytest.approx(usf, rel=1e-10) and _rel(fs, Fw) <= 1e-10
    assert ts["uintra"] == 0.0
    eng.close()


@pytest.mark.parametrize("k", [0, 15])
def test_loading_mixed_vs_reference_arithmetic(k):
    s = loading(k)
    (npair, h, uvdw, vir, Fp), (usf, Fw) = ref_parts(k)
    eng = Engine(s, precision="mixed")
    f, t = eng.frclong()
    eng.close()
    err = _rel(f, Fp + Fw)
    print(f"  mixed force err {err:.2e}", flush=True)
    assert err <= 1e-5
    assert t["uvdw"] == pytest.approx(uvdw, rel=1e-5)
    assert t["usflj"] == pytest.approx(usf, rel=1e-10)
    assert abs(t["n_lj_pairs"] - npair) <= max(2, 1e-5 * npair)


def test_respa_step_with_wall_in_inner_loop():
    """One device RESPA step (FP64, 4 inner steps) of the 20,000-site loading
    with the wall split into the inner loop, against its restatement: the
    long half kick with the pair forces, then per inner step a half kick
    with the wall's forces, the drift, the wall's forces at the new
    positions and a half kick; the pair forces at the end and the closing
    half kick (src/md.c:22-83 with frcshort = the wall)."""
    import torch
    dev = torch.device("cuda:0")
    s = loading(0)
    (npair, h, uvdw, vir, Fp), (usf, Fw) = ref_parts(0)
    n = s.natom
    rng = np.random.default_rng(3)
    m = s.masses()
    v0 = rng.standard_normal((3, n)) * np.sqrt(s.treq / m)
    x0 = np.stack([s.xx, s.yy, s.zz])
    nin = s.run.nstep_inner
    delt = s.delt
    eng = Engine(s, precision="fp64")
    eng.set_sf_split(True)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    pos, vel, fl, fs = t(x0), t(v0), t(Fp), t(Fw)
    lt = torch.zeros(LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(1, dtype=torch.float64, device=dev)
    eng.vver_step_device(pos, vel, fl, fs, delt, nin, lt, st, uk)
    torch.cuda.synchronize()
    eng.sync()
    xd, vd = pos.cpu().numpy(), vel.cpu().numpy()
    terms = abi.LongTerms.from_buffer_copy(lt.cpu().numpy().tobytes()).as_dict()
    eng.close()
    # restatement
    r = 1.0 / m
    deltby2, delts = delt / 2.0, delt / nin
    deltsby2 = delts / 2.0
    x, v, fsr = x0.copy(), v0 + deltby2 * Fp * r, Fw
    for _ in range(nin):
        v = v + deltsby2 * fsr * r
        x = x + delts * v
        u_w, fsr = wall(s, x)
        v = v + deltsby2 * fsr * r
    _, _, uv, _, flr = lj_pairs(s, x)
    v = v + deltby2 * flr * r
    assert _rel(xd - x0, x - x0) <= 1e-10
    assert _rel(vd, v) <= 1e-10
    assert terms["usflj"] == pytest.approx(u_w, rel=1e-10)
    assert terms["uinter"] == pytest.approx(uv + u_w, rel=1e-10)


def test_sixteen_loadings_as_one_table():
    """All 16 loadings in one replica table: per-replica pair counts and
    energies of the first and last loading equal the numpy restatement
    (FP64), and mixed RESPA steps with the wall in the inner loop run
    finite."""
    import torch
    dev = torch.device("cuda:0")
    systems = [loading(k) for k in range(16)]
    eng = Engine.replicas(systems, precision="fp64")
    f, terms = eng.frclong()
    parts = eng.unpack(f)
    for k in (0, 15):
        (npair, h, uvdw, vir, Fp), (usf, Fw) = ref_parts(k)
        assert terms[k]["n_lj_pairs"] == npair and terms[k]["lj_pair_hash"] == h
        assert terms[k]["usflj"] == pytest.approx(usf, rel=1e-10)
        assert _rel(parts[k], Fp + Fw) <= 1e-10
    eng.close()
    eng = Engine.replicas(systems, precision="mixed")
    eng.set_sf_split(True)
    eng.set_list_skin(1.0)
    nr = len(systems)
    pos = torch.tensor(eng.packed_positions(), device=dev)
    vel = torch.zeros_like(pos)
    fl = torch.tensor(eng.frclong()[0], device=dev)
    fs = torch.tensor(eng.frcshort()[0], device=dev)
    lt = torch.zeros(nr * LONG_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    st = torch.zeros(nr * SHORT_TERMS_BYTES // 8, dtype=torch.float64, device=dev)
    uk = torch.zeros(nr, dtype=torch.float64, device=dev)
    eng.vver_run_device(pos, vel, fl, fs, systems[0].delt, systems[0].run.nstep_inner, 5, lt,
                        st, uk, graph=True)
    torch.cuda.synchronize()
    eng.sync()
    assert bool(torch.isfinite(pos).all())
    raw = lt.cpu().numpy().tobytes()
    sz = LONG_TERMS_BYTES
    for k in range(nr):
        tk = abi.LongTerms.from_buffer_copy(raw[k * sz:(k + 1) * sz])
        assert np.isfinite(tk.uinter) and tk.usflj < 0.0
    eng.close()
