"""Parameter adjoint: dE/d(epsilon, sigma, charge) of uinter from the HIP
kernels (mspms_energy_param_grad) against central finite differences of the
oracle's energy (h = 1e-6 relative), per template site, plus the energy and
forces of the same call against the oracle (1e-10).  The reference has no
parameter derivatives; what pins this function is the finite difference of
the reference energy (oracle, bit-exact with the reference on the golden
cases).

Tolerance: |g - g_fd| <= 1e-6 * max(|g_fd|, floor), where floor = 1e-2 *
(sum of |energy components|) / |p| covers the FD rounding of energies that
cancel (Ewald real / reciprocal / self / exclusion).
"""
import copy

import numpy as np
import pytest

from mspms2_amd import abi, configs
from mspms2_amd.engine import Engine
from oracle.oracle import Oracle
from tests.util import assert_long_close, energy_scale, force_close, system_of

pytestmark = pytest.mark.gpu

CASES = {
    # the model of the differentiable Wolf water workload (alpha 0.2 / A,
    # rc 12 A) at 512 molecules
    "water512_wolf_rc12": lambda: configs.water_box(n_side=8, L=8 * 3.104, rc=12.0, kappa=0.2,
                                                    elec="wolf", seed=21),
    "water216_ewald": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                kmax=5, ksqmax=27, seed=3),
    "mixture_wolf_switch": lambda: configs.mixture_box(elec="wolf", kappa=0.2, switch=True,
                                                       seed=13),
    "bonded_zoo": lambda: configs.bonded_zoo(),
    "mixture_water4_vacuum": lambda: configs.mixture_box(water4=True, bc=abi.EWALD_BC_VACUUM,
                                                         seed=12),
}


def _energy(system, sp, site, name, value):
    s = copy.deepcopy(system)
    getattr(s.species[sp], name)[site] = value
    return Oracle(s).frclong()[1]["uinter"]


def _fd(system, sp, site, name):
    p = float(getattr(system.species[sp], name)[site])
    h = 1e-6 * abs(p) if p != 0.0 else 1e-6
    return (_energy(system, sp, site, name, p + h) - _energy(system, sp, site, name, p - h)) / (
        2 * h), p


@pytest.mark.parametrize("name", list(CASES))
def test_param_grad_vs_finite_differences(name):
    s = system_of(CASES[name]())
    print(f"case {name} natom={s.natom}", flush=True)
    fl, tl = Oracle(s).frclong()
    eng = Engine(s, precision="mixed")  # the adjoint is FP64 in every mode
    f, t, g = eng.param_grad()
    assert_long_close(t, tl, 1e-10, name)
    ok, err = force_close(f, fl, tl, s.natom, s.params["rcutoff"], 1e-10)
    assert ok, err
    scale = energy_scale(tl) + abs(tl["uvdw"])
    gsite = 0
    checked = 0
    for k, sp in enumerate(s.species):
        for a in range(sp.natom):
            for pname, key in (("epsilon", "epsilon"), ("sigma", "sigma"), ("charge", "charge")):
                val = float(getattr(sp, pname)[a])
                if pname != "charge" and val == 0.0:
                    continue   # no derivative of sqrt(eps_i eps_j) / a zero sigma
                gfd, p = _fd(s, k, a, pname)
                ga = g[key][0, gsite + a]
                floor = 1e-2 * scale / max(abs(p), 1e-12)
                assert abs(ga - gfd) <= 1e-6 * max(abs(gfd), floor), \
                    f"{name} species {k} site {a} d/d{pname}: {ga!r} vs FD {gfd!r}"
                checked += 1
        gsite += sp.natom
    print(f"  {checked} site derivatives checked", flush=True)
    eng.close()


def test_autograd_harness_water_types():
    """torch.autograd over per-type parameters (eps_O, sigma_O, q_O, q_H)
    of two Wolf-water replicas, against FD of the oracle energy sum."""
    import torch
    from mspms2_amd.autograd import NonbondedEnergy, site_params
    systems = [system_of(configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.2,
                                           elec="wolf", seed=30 + r)) for r in range(2)]
    eng = Engine.replicas(systems, precision="mixed")
    eps0, sig0, q0 = site_params(systems[0])
    # types: O (site 0), H (sites 1, 2)
    ptype = torch.tensor([eps0[0], sig0[0], q0[0], q0[1]], dtype=torch.float64,
                         requires_grad=True)
    eps = torch.cat([ptype[0:1], eps0[1:2], eps0[2:3]])
    sig = torch.cat([ptype[1:2], sig0[1:2], sig0[2:3]])
    q = torch.cat([ptype[2:3], ptype[3:4], ptype[3:4]])
    pos = torch.tensor(eng.packed_positions(), requires_grad=True)
    e = NonbondedEnergy.apply(pos, eps, sig, q, eng)
    e.sum().backward()
    gtype = ptype.grad.numpy()
    # dE/dpos = -F of each replica
    for r, s in enumerate(systems):
        fl, tl = Oracle(s).frclong()
        assert e[r].item() == pytest.approx(tl["uinter"], rel=1e-10, abs=1e-10 * energy_scale(tl))
        gp = pos.grad[:, eng.rep_atom0[r]:eng.rep_atom0[r + 1]].numpy()
        ok, err = force_close(-gp, fl, tl, s.natom, s.params["rcutoff"], 1e-10)
        assert ok, err

    def total(p):
        tot = 0.0
        for s in systems:
            c = copy.deepcopy(s)
            w = c.species[0]
            w.epsilon[0], w.sigma[0], w.charge[0] = p[0], p[1], p[2]
            w.charge[1] = w.charge[2] = p[3]
            tot += Oracle(c).frclong()[1]["uinter"]
        return tot

    p0 = ptype.detach().numpy().copy()
    scale = sum(energy_scale(Oracle(s).frclong()[1]) for s in systems)
    for i in range(4):
        h = 1e-6 * abs(p0[i])
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        gfd = (total(pp) - total(pm)) / (2 * h)
        floor = 1e-2 * scale / abs(p0[i])
        assert abs(gtype[i] - gfd) <= 1e-6 * max(abs(gfd), floor), (i, gtype[i], gfd)
    eng.close()
