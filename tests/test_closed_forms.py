"""Pins the oracle against mathematics, independently of the reference.

  - LJ: brute-force numpy pair sum (energy, virial, forces) on LJ-864
  - Ewald: rock-salt NaCl reproduces the Madelung constant 1.7475645946...
  - momentum: sum of all frclong forces vanishes
  - forces = -gradient of the energies by central differences, for every
    term kind (LJ, Ewald real + reciprocal + exclusion, Wolf, harmonic /
    Morse / FENE bonds, harmonic / Toukan-Rahman angles, OPLS dihedrals,
    the hypergeometric tube)
  - tricubic Hermite grid: exact on any tricubic polynomial
  - RANMAR known-answer test (src/random.c:28-34)
  - InitCheckUniques flags on a molecule with shared ending pairs
"""
import ctypes as C
import os

import numpy as np
import pytest

from mspms2_amd import abi, configs, runio
from mspms2_amd.runio import RunDir, RunInput, Site, SpeciesSpec
from mspms2_amd.system import System
from oracle import oracle as O
from tests.util import GOLDEN, system_of

MADELUNG_NACL = 1.7475645946331821906


def _golden_system(name):
    return System.from_rundir(runio.RunDir.read(os.path.join(GOLDEN, name)))


def test_lj_bruteforce_lattice_and_jitter():
    for rd in (configs.lj_fcc(), configs.lj_fcc(jitter=0.05, seed=7)):
        s = system_of(rd)
        f, t = O.Oracle(s).frclong()
        x = np.stack([s.xx, s.yy, s.zz], 1)
        L = s.params["boxlx"]
        rc2 = s.params["rcutoff"] ** 2
        d = x[:, None, :] - x[None, :, :]
        d -= L * np.rint(d / L)
        r2 = (d * d).sum(-1)
        np.fill_diagonal(r2, np.inf)
        m = r2 < rc2
        ir6 = np.where(m, 1.0 / np.where(m, r2, 1.0) ** 3, 0.0)
        u = 0.5 * np.sum(4.0 * (ir6 * ir6 - ir6))
        w = 0.5 * np.sum(np.where(m, 24.0 * (2 * ir6 * ir6 - ir6), 0.0))
        fij = np.where(m, 24.0 * (2 * ir6 * ir6 - ir6) / np.where(m, r2, 1.0), 0.0)
        fb = np.einsum("ij,ijk->ik", fij, d).T
        assert t["uvdw"] == pytest.approx(u, rel=1e-12)
        assert t["virial_inter"] == pytest.approx(w, rel=1e-12)
        assert np.linalg.norm(f - fb) <= 1e-10 * max(np.linalg.norm(fb), 1e-300) + 1e-9
        assert t["n_lj_pairs"] == int(m.sum() // 2)
    # FCC lattice energy per atom at rho* = 0.8442, rc = 2.5 (measured value
    # quoted for the reference build)
    s = system_of(configs.lj_fcc())
    _, t = O.Oracle(s).frclong()
    assert t["uvdw"] / 864 == pytest.approx(-6.7734, abs=5e-5)


def _nacl_rundir(ncell=4, a=5.64, kappa=0.39, kmax=13):
    """Rock salt of single-site ions (no LJ) in a cube of ncell^3 cells."""
    basis_na = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    basis_cl = basis_na + np.array([0.5, 0, 0])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell)
                      for k in range(ncell)], dtype=float)
    na = ((cells[:, None] + basis_na[None]) * a).reshape(-1, 3)
    cl = ((cells[:, None] + basis_cl[None]) * a).reshape(-1, 3)
    L = ncell * a
    mk = lambda nm, q: SpeciesSpec(name=nm, nmole=len(na), sites=[
        Site(type=0, name=nm, xx=0, yy=0, zz=0, aw=0.023, epsilon=0.0, sigma=1.0, charge=q)])
    run = RunInput(boxlx=L, boxly=L, boxlz=L, rcuton=L / 2 * 0.975, rcutoff=L / 2 * 0.975,
                   rcutoffelec=L / 2 * 0.975, iChargeType=abi.ELEC_EWALD, kappa=kappa,
                   kmaxx=kmax, kmaxy=kmax, kmaxz=kmax, ksqmax=kmax * kmax, fEwald_BC=1)
    return RunDir(run=run, sysname="nacl", species=[mk("Na", 1.0), mk("Cl", -1.0)],
                  xyz=np.vstack([na, cl])), a


def test_ewald_madelung_constant():
    rd, a = _nacl_rundir()
    s = system_of(rd)
    f, t = O.Oracle(s).frclong()
    npair = s.natom // 2
    expect = -MADELUNG_NACL * npair * 167100.80840 / (a / 2)
    assert t["uelec"] == pytest.approx(expect, rel=2e-7)
    # every ion of the perfect lattice is at a force minimum
    assert np.abs(f).max() <= 1e-6 * 167100.8 / (a / 2) ** 2


@pytest.mark.parametrize("name", ["water216_ewald", "mixture_ewald", "water216_wolf",
                                  "mixture_water4_vacuum"])
def test_momentum_conservation(name):
    s = _golden_system(name)
    f, _ = O.Oracle(s).frclong()
    scale = np.abs(f).sum()
    assert np.abs(f.sum(axis=1)).max() <= 1e-11 * scale


def _fd(fun, s, atoms, h):
    """central differences of fun(xx, yy, zz) along x/y/z of `atoms`"""
    out = []
    for i in atoms:
        g = []
        for arr in ("xx", "yy", "zz"):
            v = [s.xx.copy(), s.yy.copy(), s.zz.copy()]
            k = "xx yy zz".split().index(arr)
            v[k][i] += h
            ep = fun(*v)
            v[k][i] -= 2 * h
            em = fun(*v)
            g.append(-(ep - em) / (2 * h))
        out.append(g)
    return np.array(out).T


@pytest.mark.parametrize("name,which", [
    ("water216_ewald", "long"), ("water216_wolf", "long"), ("mixture_ewald", "long"),
    ("mixture_wolf_switch", "long"), ("tube300_hypergeo", "long"), ("bonded_zoo", "long"),
    ("tube_explicit", "long"), ("tube_explicit_switch", "long"),
    ("butane64", "short"), ("examples_water", "short"), ("bonded_zoo_nofene", "short"),
])
def test_forces_are_minus_gradient(name, which):
    if name == "bonded_zoo_nofene":
        # every bonded type (OPLS and CHARMM dihedrals, CHARMM improper,
        # Morse / harmonic bonds) except FENE, whose energy has the opposite
        # sign of its force in the reference (test_fene_sign_defect)
        rd = configs.bonded_zoo()
        sp = rd.species[0]
        sp.bonds = [(b[0], b[1], abi.BOND_HARMONIC, 9.0e5, 1.5, 0.0) if b[2] == abi.BOND_FENE
                    else b for b in sp.bonds]
        s = system_of(rd)
    else:
        s = _golden_system(name)
    o = O.Oracle(s)
    rng = np.random.default_rng(3)
    atoms = rng.choice(s.natom, size=6, replace=False)
    h = 1e-6 * max(1.0, 1.0 / s.params["sigma_base"])
    if which == "long":
        f, _ = o.frclong()
        fd = _fd(lambda x, y, z: o.frclong(x, y, z)[1]["uinter"], s, atoms, h)
    else:
        f, _ = o.frcshort()
        fd = _fd(lambda x, y, z: o.frcshort(x, y, z)[1]["uintra"], s, atoms, h)
    an = f[:, atoms]
    scale = max(np.abs(an).max(), 1e-300)
    assert np.abs(an - fd).max() <= 2e-5 * scale, (an, fd)


@pytest.mark.parametrize("name", ["tube_explicit", "tube_explicit_switch"])
def test_explicit_wall_bruteforce(name):
    """cal_sf_atom_explicit (src/sffrc.c:1213-1314): usflj and the wall's
    forces equal a numpy sum of 4 eps_ij [(s/r)^12 - (s/r)^6] over (non-ghost
    fluid site, wall atom) pairs within the cutoff, minimum image, Lorentz-
    Berthelot mixing, times the switch (rc^2 - r^2)^2 (rc^2 + 2 r^2 - 3 ron^2)
    / (rc^2 - ron^2)^3 beyond ron when it is on; the forces are its exact
    gradient.  The wall's share of the oracle's forces is the difference to
    the same system without the wall (iSF_type 0)."""
    s = _golden_system(name)
    p = s.params
    f, t = O.Oracle(s).frclong()
    p0 = dict(p, iSF_type=0)
    s0 = System(species=s.species, params=p0, xx=s.xx, yy=s.yy, zz=s.zz, run=s.run)
    f0, t0 = O.Oracle(s0).frclong()
    x = np.stack([s.xx, s.yy, s.zz], 1)
    w = np.stack([p["solid_xx"], p["solid_yy"], p["solid_zz"]], 1)
    L = np.array([p["boxlx"], p["boxly"], p["boxlz"]])
    sig = np.concatenate([np.tile(sp.sigma, sp.nmole) for sp in s.species])
    eps = np.concatenate([np.tile(sp.epsilon, sp.nmole) for sp in s.species])
    real = np.concatenate([np.tile(sp.ghost_type == abi.GHOST_NONE, sp.nmole)
                           for sp in s.species])
    d = x[:, None, :] - w[None, :, :]
    d -= L * np.rint(d / L)
    r2 = (d * d).sum(-1)
    rc2 = p["rcutoff"] ** 2
    m = (r2 < rc2) & real[:, None]
    sij = 0.5 * (sig + p["solid_sigma"])[:, None]
    eij = np.sqrt(eps * p["solid_epsilon"])[:, None]
    q = np.where(m, (sij * sij) / np.where(m, r2, 1.0), 0.0)
    q6 = q ** 3
    u = 4.0 * eij * (q6 * q6 - q6)                   # pair energies
    du = -4.0 * eij * (12 * q6 * q6 - 6 * q6) / np.where(m, r2, 1.0)  # dU/d(r^2) * 2
    if p["isLJswitchOn"]:
        ron2 = p["rcuton"] ** 2
        den = (rc2 - ron2) ** 3
        sw = np.where(r2 < ron2, 1.0, (rc2 - r2) ** 2 * (rc2 + 2 * r2 - 3 * ron2) / den)
        dsw = np.where(r2 < ron2, 0.0, 6.0 * (rc2 - r2) * (ron2 - r2) / den)  # dS/d(r^2)
        du = du * sw + 2.0 * u * dsw
        u = u * sw
    u = np.where(m, u, 0.0)
    fw = -np.einsum("ij,ijk->ik", np.where(m, du, 0.0), d).T
    assert t["usflj"] == pytest.approx(u.sum(), rel=1e-12)
    assert np.linalg.norm((f - f0) - fw) <= 1e-11 * np.linalg.norm(fw)
    # ghost sites within the cutoff of wall atoms exist, so the skip matters
    assert t0["usflj"] == 0.0 and ((r2 < rc2) & ~real[:, None]).any()


def test_fene_sign_defect():
    """src/bndfrc.c:55-59: u = Kb rmax^2 log(1 - r^2/rmax^2) but
    fij = -2 Kb rmax^2 / (rmax^2 - r^2), the force of -u.  The restatement
    keeps the reference's behaviour: force = +gradient of the FENE energy."""
    rd = configs.bonded_zoo(nmole=4)
    sp = rd.species[0]
    sp.bonds = [b for b in sp.bonds if b[2] == abi.BOND_FENE]
    sp.angles, sp.dihedrals, sp.impropers = [], [], []
    s = system_of(rd)
    o = O.Oracle(s)
    f, t = o.frcshort()
    assert t["ubond"] < 0.0
    fd = _fd(lambda x, y, z: o.frcshort(x, y, z)[1]["uintra"], s, range(6), 1e-6)
    an = f[:, :6]
    assert np.abs(an + fd).max() <= 1e-6 * np.abs(an).max()


def test_hermite_exact_on_tricubic_polynomials():
    """The tricubic Hermite interpolant (mylibtasos/cubicinterpolation.f90)
    reproduces any polynomial of degree <= 3 in each variable, with its
    gradient, from node values and derivatives."""
    rng = np.random.default_rng(5)
    Cc = rng.integers(-3, 4, size=(4, 4, 4)).astype(float)
    n, hgrid = 6, 0.5
    g = np.arange(n) * hgrid
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")

    def p(x, y, z, dx=0, dy=0, dz=0):
        tot = 0.0
        for a in range(4):
            for b in range(4):
                for c in range(4):
                    ca = 1.0
                    for k in range(dx):
                        ca *= (a - k)
                    cb = 1.0
                    for k in range(dy):
                        cb *= (b - k)
                    cc = 1.0
                    for k in range(dz):
                        cc *= (c - k)
                    if ca == 0 or cb == 0 or cc == 0:
                        continue
                    tot = tot + Cc[a, b, c] * ca * cb * cc * x ** (a - dx) * y ** (b - dy) * \
                        z ** (c - dz)
        return tot
    fields = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1),
              (1, 1, 1)]
    nodes = np.stack([p(X, Y, Z, *d) for d in fields])          # [8][x][y][z]
    nodes = np.ascontiguousarray(nodes.transpose(0, 3, 2, 1)).astype(np.float32)  # [8][z][y][x]
    assert np.array_equal(nodes.astype(np.float64),
                          np.stack([p(X, Y, Z, *d) for d in fields]).transpose(0, 3, 2, 1))
    grid = abi.SFGrid(ntype=1, nx=n, ny=n, nz=n, x0=0.0, y0=0.0, z0=0.0, hx=hgrid, hy=hgrid,
                      hz=hgrid, nodes=abi.fptr(nodes))
    L = O.lib()
    minpe = L.oracle_grid_min_pe(C.byref(grid), 0)
    for _ in range(40):
        x, y = rng.uniform(0, (n - 1) * hgrid, 2)
        z = rng.uniform(0, (n - 2) * hgrid)
        pe = C.c_double()
        f3 = (C.c_double * 3)()
        rc = L.oracle_hermite_site(C.byref(grid), 0, minpe, x, y, z, C.byref(pe), f3)
        assert rc == 0
        val = p(x, y, z)
        if val >= 1.5 * minpe:
            assert pe.value == pytest.approx(val, rel=1e-12, abs=1e-9)
        for k, d in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
            assert -f3[k] == pytest.approx(p(x, y, z, *d), rel=1e-11, abs=1e-8)
    # outside the x/y range: EDOM (mylibtasos/forces.f90:59-79 flags it)
    pe = C.c_double()
    f3 = (C.c_double * 3)()
    assert L.oracle_hermite_site(C.byref(grid), 0, minpe, -0.1, 0.2, 0.2, C.byref(pe),
                                 f3) == abi.MSPMS_EDOM


def test_ranmar_known_answer():
    """src/random.c:28-34: seeds 1802 / 9373, skip 20000, the next six
    numbers times 4096^2."""
    r = O.Ranmar(1802, 9373)
    r.next(20000)
    got = r.next(6) * 4096.0 * 4096.0
    assert list(got) == [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0, 10633180.0]


def test_check_uniques_quirks():
    """InitCheckUniques (src/initvars.c:13-146) on the zoo molecule: the
    dihedral (3,2,1,0) repeats (0,1,2,3)'s ending pair and is dropped; the
    dihedral (0,5,1,2) shares ending pair (0,2) with angle 0, which makes
    the reference clear isDih_unique[0] (it indexes by the angle, :66).
    The product's rule and the oracle's agree."""
    s = system_of(configs.bonded_zoo())
    sp, keep = s.c_species()
    au, du = O.check_uniques(sp[0])
    assert list(du) == [0, 1, 1, 1, 0, 1]
    assert list(au) == [1, 1, 1, 1, 1]
    from mspms2_amd.engine import lib
    au2 = np.zeros(5, dtype=np.uint8)
    du2 = np.zeros(6, dtype=np.uint8)
    assert lib().mspms_check_uniques(C.byref(sp[0]), abi.u8ptr(au2), abi.u8ptr(du2)) == 0
    assert list(au2) == list(au) and list(du2) == list(du)


def test_grid_reproduces_tube_closed_form():
    """The Hermite grid tabulated from the hypergeometric tube (0.25 A
    spacing; the interpolation error falls as h^4) reproduces the
    closed-form energy and forces of cal_sf_hypergeo (src/sffrc.c:1120-1209)
    inside the pore."""
    from tests.util import tube_grid_system
    sg = tube_grid_system(hx=0.25)
    sh = _golden_system("tube300_hypergeo")
    assert np.array_equal(sg.xx, sh.xx)
    fg, tg = O.Oracle(sg).frclong()
    fh, th = O.Oracle(sh).frclong()
    assert tg["usflj"] == pytest.approx(th["usflj"], rel=1e-4)
    # pair forces are identical; the solid-fluid part differs by the
    # interpolation error only
    assert np.linalg.norm(fg - fh) <= 1e-3 * np.linalg.norm(fh)
