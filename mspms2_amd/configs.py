"""Seeded synthetic workloads, written as mspms2 run directories.

Each function returns a `RunDir` (in/cfg/xyz.mspms content) that the
reference's own start-up path can read unchanged.  Positions come from
numpy's PCG64 generator with a fixed seed, so the same call gives the same
bytes on every machine with this numpy.

  lj_fcc          LJ-864 argon on an FCC lattice, rho*=0.8442, rc=2.5 sigma
  water_box       flexible 3-site water (SPC/Fw-like charges/LJ, harmonic
                  O-H and H-O-H), Ewald or Wolf
  wolf_water_set  seeded configurations of 1024 waters with Wolf sums (the
                  force-field fitting workload)
  butane_table    256 n-butane boxes of 64-512 molecules (replica table)
  tube_loading    LJ methane filling a 50 A smooth tube, 100 x 100 x 4000 A box
  tube_loadings   its 16 loadings from 20,000 to 100,000 sites
  butane_box      united-atom TraPPE n-butane
  methane_tube    LJ methane inside a smooth cylindrical pore
  explicit_tube   LJ methane and a ghost-site diatomic around an explicit
                  carbon nanotube wall (iSF_type 2, ATOMEXPLICIT)
  reference_water the reference's Examples/Water inputs (TR water, Morse)
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import abi
from .runio import RunDir, RunInput, Site, SpeciesSpec

AVOGADRO = 6.0221415e23
RGAS = 8.314472
BOLTZMAN_CONSTANT = 1.3806505e-23


def _time_base(mass_base, eps_base, sigma_base):
    return math.sqrt(mass_base * 1.0e-27 / (eps_base * BOLTZMAN_CONSTANT)) * sigma_base * 1.0e5


def _random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
                  np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3),
                  np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _min_image_dist2(a, b, L):
    d = a - b
    d -= L * np.rint(d / L)
    return np.sum(d * d, axis=-1)


# --------------------------------------------------------------------------- #
def lj_fcc(ncell: int = 6, rho: float = 0.8442, temp: float = 0.722,
           rc: float = 2.5, jitter: float = 0.0, seed: int = 1,
           ensemble: int = 1, nstep: int = 1000) -> RunDir:
    """LJ fluid: N=4*ncell^3 atoms on an FCC lattice (reduced units given
    via argon bases sigma 3.405 A, epsilon 119.8 K, mass 1 reduced)."""
    sb, eb = 3.405, 119.8
    aw = 0.039948
    mb = aw / AVOGADRO / 1.0e-27
    n = 4 * ncell ** 3
    L = (n / rho) ** (1.0 / 3.0)
    a = L / ncell
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell)
                      for k in range(ncell)], dtype=np.float64)
    pos = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    if jitter:
        rng = np.random.default_rng(seed)
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape) * a
    tb = _time_base(mb, eb, sb)
    run = RunInput(title="LJ-864 FCC argon", ij=1802, jk=9373, sigma_base=sb,
                   epsilon_base=eb, mass_base=mb, treq=temp * eb, preq=1.0,
                   boxlx=L * sb, boxly=L * sb, boxlz=L * sb, rcuton=rc * sb,
                   rcutoff=rc * sb, rcutoffelec=rc * sb, isLJlrcOn=1, nstep=nstep,
                   nstep_ave=100, nstep_print=100, delt=0.005 * tb, nstep_inner=1,
                   f0=0.0, what_simulation=0, what_ensemble=ensemble, nconstraint=3,
                   iChargeType=abi.ELEC_NONE, qq=100.0, qqs=100.0)
    sp = SpeciesSpec(name="argon", nmole=n, sites=[
        Site(type=0, name="Ar", xx=0.0, yy=0.0, zz=0.0, aw=aw, epsilon=eb, sigma=sb,
             charge=0.0, ghost_type=0, interp_type=1)])
    return RunDir(run=run, sysname="lj-fcc", species=[sp], xyz=pos * sb)


# --------------------------------------------------------------------------- #
WATER_Q_O, WATER_Q_H = -0.8476, 0.4238
WATER_SIGMA_O, WATER_EPS_O = 3.166, 0.650e3 / RGAS        # Angstrom, K
KCAL_TO_K = 4184.0 / RGAS
WATER_KB = 1059.162 * KCAL_TO_K        # K/A^2, U = Kb/2 (r-r0)^2
WATER_KTHETA = 75.90 * KCAL_TO_K       # K/rad^2, U = Kt/2 (th-th0)^2
WATER_R0 = 1.0
WATER_THETA0 = 109.47 * math.pi / 180.0


def water_species(nmole: int) -> SpeciesSpec:
    """Flexible three-site water with harmonic O-H bonds and H-O-H angle;
    H sites carry charge only (GHOST_LJ, as the reference's Water example)."""
    h = WATER_THETA0 / 2
    tmpl = np.array([[0.0, 0.0, 0.0],
                     [WATER_R0 * math.sin(h), 0.0, WATER_R0 * math.cos(h)],
                     [-WATER_R0 * math.sin(h), 0.0, WATER_R0 * math.cos(h)]])
    sites = [
        Site(type=0, name="O", xx=tmpl[0, 0], yy=tmpl[0, 1], zz=tmpl[0, 2], aw=0.0159994,
             epsilon=WATER_EPS_O, sigma=WATER_SIGMA_O, charge=WATER_Q_O, ghost_type=0,
             interp_type=1),
        Site(type=1, name="H", xx=tmpl[1, 0], yy=tmpl[1, 1], zz=tmpl[1, 2], aw=0.001008,
             epsilon=0.0, sigma=0.0, charge=WATER_Q_H, ghost_type=1, interp_type=2),
        Site(type=1, name="H", xx=tmpl[2, 0], yy=tmpl[2, 1], zz=tmpl[2, 2], aw=0.001008,
             epsilon=0.0, sigma=0.0, charge=WATER_Q_H, ghost_type=1, interp_type=2),
    ]
    return SpeciesSpec(
        name="water", nmole=nmole, sites=sites,
        bonds=[(0, 1, abi.BOND_HARMONIC, WATER_KB, WATER_R0, 0.0),
               (0, 2, abi.BOND_HARMONIC, WATER_KB, WATER_R0, 0.0)],
        angles=[(1, 0, 2, abi.ANGLE_HARMONIC, WATER_KTHETA, WATER_THETA0, 0.0, 0.0, 0.0)])


def _spin_tilt_rotation(rng, max_tilt):
    """Rotation by a uniform angle about z after a tilt of at most max_tilt
    radians about a random horizontal axis."""
    phi, psi = 2 * np.pi * rng.random(), 2 * np.pi * rng.random()
    th = max_tilt * rng.random()
    a = np.array([np.cos(psi), np.sin(psi), 0.0])
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    T = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * (K @ K)
    Rz = np.array([[np.cos(phi), -np.sin(phi), 0], [np.sin(phi), np.cos(phi), 0], [0, 0, 1]])
    return Rz @ T


def _place_rigid(rng, centers, tmpl, L, min_dist, max_tries=400, recenter=False,
                 jitter=None, rotation=None):
    """Random orientation per molecule (uniform, or `rotation(rng)`),
    rejecting contacts closer than min_dist to molecules already placed
    (cell-list check).  Every 20 failed tries the centre is redrawn:
    uniformly in the box (recenter) or within +-jitter of its lattice site
    (jitter); running out of tries is an error."""
    nmol, nsite = centers.shape[0], tmpl.shape[0]
    out = np.empty((nmol, nsite, 3))
    ncell = max(1, int(L // max(min_dist, 1e-9)))
    h = L / ncell
    grid = {}
    md2 = min_dist * min_dist
    for m in range(nmol):
        c0 = centers[m].copy()
        if jitter:
            c0 = centers[m] + rng.uniform(-jitter, jitter, 3)
        ok = False
        for attempt in range(max_tries):
            if attempt and attempt % 20 == 0:
                if recenter:
                    c0 = rng.uniform(0, L, 3)
                elif jitter:
                    c0 = centers[m] + rng.uniform(-jitter, jitter, 3)
            R = _random_rotations(rng, 1)[0] if rotation is None else rotation(rng)
            xyz = c0 + tmpl @ R.T
            ok = True
            cells = set()
            for s in range(nsite):
                w = xyz[s] - L * np.floor(xyz[s] / L)
                c = tuple((w // h).astype(int) % ncell)
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dz in (-1, 0, 1):
                            cc = ((c[0] + dx) % ncell, (c[1] + dy) % ncell,
                                  (c[2] + dz) % ncell)
                            for p in grid.get(cc, ()):
                                if _min_image_dist2(xyz[s], p, L) < md2:
                                    ok = False
                                    break
                            if not ok:
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
                cells.add(c)
            if ok:
                break
        if not ok:
            raise RuntimeError(f"could not place molecule {m} without contacts < {min_dist} A")
        out[m] = xyz
        for s in range(nsite):
            w = xyz[s] - L * np.floor(xyz[s] / L)
            c = tuple((w // h).astype(int) % ncell)
            grid.setdefault(c, []).append(xyz[s].copy())
    return out


def water_box(n_side: int = 16, L: float = 49.66, seed: int = 2,
              elec: str = "ewald", kappa: float = 0.28, kmax: int = 8,
              ksqmax: int = 64, rc: float = 10.0, jitter: float = 0.3,
              nstep_inner: int = 8, delt_fs: float = 2.0, min_dist: float = 1.6,
              ensemble: int = 0, nmol: Optional[int] = None) -> RunDir:
    """Flexible three-site water on a jittered cubic lattice with uniform
    random orientations (n_side=16: 12,288 sites; Ewald or Wolf).  With nmol,
    nmol lattice sites of the n_side^3 lattice are taken at random."""
    rng = np.random.default_rng(seed)
    a = L / n_side
    idx = np.array([[i, j, k] for i in range(n_side) for j in range(n_side)
                    for k in range(n_side)], dtype=np.float64)
    if nmol is not None:
        if nmol > len(idx):
            raise ValueError("nmol exceeds the lattice")
        idx = idx[np.sort(rng.permutation(len(idx))[:nmol])]
    nmol = len(idx)
    centers = (idx + 0.5) * a
    sp = water_species(nmol)
    tmpl = np.array([[s.xx, s.yy, s.zz] for s in sp.sites])
    tmpl = tmpl - (tmpl * np.array([15.9994, 1.008, 1.008])[:, None]).sum(0) / 18.0154
    xyz = _place_rigid(rng, centers, tmpl, L, min_dist, jitter=jitter).reshape(-1, 3)
    ct = {"ewald": abi.ELEC_EWALD, "wolf": abi.ELEC_WOLF, "none": abi.ELEC_NONE,
          "coulomb": abi.ELEC_SIMPLE_COULOMB}[elec]
    run = RunInput(title=f"water {nmol} {elec}", ij=1802, jk=9373, sigma_base=1.0,
                   epsilon_base=1.0, mass_base=1.0, treq=298.0, preq=1.0, boxlx=L,
                   boxly=L, boxlz=L, rcuton=rc, rcutoff=rc, rcutoffelec=rc,
                   isLJlrcOn=0, delt=delt_fs, nstep_inner=nstep_inner, f0=0.0,
                   what_simulation=0, what_ensemble=ensemble, nconstraint=3,
                   iChargeType=ct, kappa=kappa, kmaxx=kmax, kmaxy=kmax, kmaxz=kmax,
                   ksqmax=ksqmax, fEwald_BC=abi.EWALD_BC_TINFOIL, fEwald_Dim=3,
                   qq=200.0, qqs=200.0)
    return RunDir(run=run, sysname="water", species=[sp], xyz=xyz)


def water4_species(nmole: int) -> SpeciesSpec:
    """Four-site water variant: the negative charge sits on a GHOST_FULL site
    M 0.15 A along the bisector (skipped by every pair loop of the
    reference, kept in the Ewald sums), O carries LJ only."""
    sp = water_species(nmole)
    o, h1, h2 = (np.array([s.xx, s.yy, s.zz]) for s in sp.sites)
    bis = (h1 + h2) / 2 - o
    mpos = o + 0.15 * bis / np.linalg.norm(bis)
    sp.sites[0].charge = 0.0
    sp.sites.append(Site(type=2, name="M", xx=mpos[0], yy=mpos[1], zz=mpos[2], aw=0.0,
                         epsilon=0.0, sigma=0.0, charge=WATER_Q_O, ghost_type=2, interp_type=0))
    return sp


def co_species(nmole: int) -> SpeciesSpec:
    """Charged diatomic (CO-like), Morse bond; a second species for the
    cross-species loops of src/frclong.c:74-80."""
    sites = [Site(type=0, name="C", xx=0.0, yy=0.0, zz=0.0, aw=0.012011, epsilon=35.0,
                  sigma=3.55, charge=0.2, ghost_type=0, interp_type=1),
             Site(type=1, name="O", xx=0.0, yy=0.0, zz=1.128, aw=0.0159994, epsilon=61.6,
                  sigma=2.95, charge=-0.2, ghost_type=0, interp_type=2)]
    return SpeciesSpec(name="co", nmole=nmole, sites=sites,
                       bonds=[(0, 1, abi.BOND_MORSE, 1.3e5, 1.128, 2.3)])


def mixture_box(n_water: int = 150, n_co: int = 60, L: float = 22.0, seed: int = 8,
                elec: str = "ewald", kappa: float = 0.3, kmax: int = 5, ksqmax: int = 27,
                rc: float = 9.0, water4: bool = False, bc: int = abi.EWALD_BC_TINFOIL,
                switch: bool = False) -> RunDir:
    """Two species (water, CO) at random positions and orientations."""
    rng = np.random.default_rng(seed)
    wsp = water4_species(n_water) if water4 else water_species(n_water)
    csp = co_species(n_co)
    out = []
    for sp, n in ((wsp, n_water), (csp, n_co)):
        tmpl = np.array([[s.xx, s.yy, s.zz] for s in sp.sites])
        tmpl = tmpl - tmpl.mean(0)
        cen = rng.uniform(0, L, size=(n, 3))
        out.append((sp, cen, tmpl))
    xyz = []
    placed = np.zeros((0, 3))
    for sp, cen, tmpl in out:
        for c in cen:
            for _ in range(500):
                R = _random_rotations(rng, 1)[0]
                cand = c + tmpl @ R.T
                if placed.shape[0] == 0:
                    ok = True
                else:
                    d2 = np.min(_min_image_dist2(cand[:, None, :], placed[None, :, :], L))
                    ok = d2 > 2.0 ** 2
                if ok:
                    break
                c = rng.uniform(0, L, 3)
            xyz.append(cand)
            placed = np.vstack([placed, cand])
    ct = {"ewald": abi.ELEC_EWALD, "wolf": abi.ELEC_WOLF, "none": abi.ELEC_NONE,
          "coulomb": abi.ELEC_SIMPLE_COULOMB}[elec]
    run = RunInput(title="water/CO mixture", sigma_base=1.0, epsilon_base=1.0, mass_base=1.0,
                   treq=300.0, boxlx=L, boxly=L, boxlz=L,
                   rcuton=(0.8 * rc if switch else rc), rcutoff=rc, rcutoffelec=rc,
                   isLJswitchOn=1 if switch else 0, iChargeType=ct, kappa=kappa, kmaxx=kmax,
                   kmaxy=kmax, kmaxz=kmax, ksqmax=ksqmax, fEwald_BC=bc, delt=1.0,
                   nstep_inner=4)
    return RunDir(run=run, sysname="mixture", species=[wsp, csp], xyz=np.vstack(xyz))


def bonded_zoo(nmole: int = 40, L: float = 30.0, seed: int = 9, f0: float = 0.5) -> RunDir:
    """A 6-site synthetic molecule with every bonded type the reference has:
    harmonic / Morse / FENE bonds, harmonic angles, OPLS and CHARMM
    dihedrals (one pair sharing ending atoms, for InitCheckUniques), a
    CHARMM improper; charged sites, 1-4 pairs with f0 scaling, Ewald."""
    rng = np.random.default_rng(seed)
    tmpl = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [2.0, 1.4, 0.0], [3.5, 1.5, 0.3],
                     [1.9, -0.9, 1.1], [0.2, 1.2, -0.8]])
    sites = []
    for k in range(6):
        sites.append(Site(type=k, name=f"X{k}", xx=tmpl[k, 0], yy=tmpl[k, 1], zz=tmpl[k, 2],
                          aw=0.014, epsilon=40.0 + 10 * k, sigma=3.0 + 0.1 * k,
                          charge=[0.3, -0.2, 0.1, -0.25, 0.15, -0.1][k], ghost_type=0,
                          interp_type=1))
    sp = SpeciesSpec(
        name="zoo", nmole=nmole, sites=sites,
        bonds=[(0, 1, abi.BOND_HARMONIC, 9.0e5, 1.5, 0.0), (1, 2, abi.BOND_MORSE, 4.0e4, 1.45, 2.0),
               (2, 3, abi.BOND_FENE, 1.5e4, 2.2, 0.0), (1, 4, abi.BOND_HARMONIC, 8.0e5, 1.45, 0.0),
               (0, 5, abi.BOND_HARMONIC, 8.5e5, 1.5, 0.0)],
        angles=[(0, 1, 2, abi.ANGLE_HARMONIC, 6.0e4, 1.95, 0, 0, 0),
                (1, 2, 3, abi.ANGLE_HARMONIC, 6.2e4, 1.9, 0, 0, 0),
                (0, 1, 4, abi.ANGLE_HARMONIC, 5.0e4, 1.9, 0, 0, 0),
                (2, 1, 4, abi.ANGLE_HARMONIC, 5.5e4, 1.95, 0, 0, 0),
                (1, 0, 5, abi.ANGLE_HARMONIC, 5.8e4, 1.92, 0, 0, 0)],
        dihedrals=[(0, 1, 2, 3, abi.DIH_OPLS_COSIN, 0.0, 710.06, -136.38, 1582.64),
                   (4, 1, 2, 3, abi.DIH_CHARMM, 300.0, 3.0, 0.0, 0.0),
                   (5, 0, 1, 2, abi.DIH_OPLS_COSIN, 10.0, 500.0, 100.0, 800.0),
                   (5, 0, 1, 4, abi.DIH_CHARMM, 200.0, 2.0, 180.0, 0.0),
                   (3, 2, 1, 0, abi.DIH_OPLS_COSIN, 0.0, 100.0, 50.0, 20.0),
                   # ending pair (0,2) = angle 0's: InitCheckUniques then marks
                   # isDih_unique by the angle index (src/initvars.c:66)
                   (0, 5, 1, 2, abi.DIH_OPLS_COSIN, 0.0, 300.0, 0.0, 0.0)],
        impropers=[(1, 0, 2, 4, abi.IMP_CHARMM, 4.0e3, 0.0)])
    cen = rng.uniform(0, L, size=(nmole, 3))
    xyz = _place_rigid(rng, cen, tmpl - tmpl.mean(0), L, 2.6, recenter=True).reshape(-1, 3)
    xyz = xyz + rng.uniform(-0.05, 0.05, size=xyz.shape)
    run = RunInput(title="bonded zoo", sigma_base=1.0, epsilon_base=1.0, mass_base=1.0,
                   treq=300.0, boxlx=L, boxly=L, boxlz=L, rcuton=11.0, rcutoff=11.0,
                   rcutoffelec=11.0, iChargeType=abi.ELEC_EWALD, kappa=0.3, kmaxx=5, kmaxy=5,
                   kmaxz=5, ksqmax=27, f0=f0, delt=1.0, nstep_inner=4)
    return RunDir(run=run, sysname="zoo", species=[sp], xyz=xyz)


# --------------------------------------------------------------------------- #
BUTANE_DENSITY = 0.58 * AVOGADRO / 58.12 / 1.0e24   # molecules / A^3
BUTANE_TMPL = np.array([[0.0, 0.0, 0.0], [-0.5133, -0.7260, 1.2574],
                        [-2.0533, -0.7235, 1.2588], [-2.5667, -1.4500, 2.5160]])


def butane_species(nmole: int) -> SpeciesSpec:
    """TraPPE-UA n-butane: CH3 (3.75 A, 98 K), CH2 (3.95 A, 46 K), bonds
    1.54 A (stiff harmonic), angle 114 deg, TraPPE cosine dihedral in the
    reference's c1..c4 convention (root cfg.mspms:11-18)."""
    s = []
    for k, (nm, sg, ep, aw) in enumerate([("CH3", 3.75, 98.0, 0.015035),
                                          ("CH2", 3.95, 46.0, 0.014027),
                                          ("CH2", 3.95, 46.0, 0.014027),
                                          ("CH3", 3.75, 98.0, 0.015035)]):
        s.append(Site(type=0 if nm == "CH3" else 1, name=nm, xx=BUTANE_TMPL[k, 0],
                      yy=BUTANE_TMPL[k, 1], zz=BUTANE_TMPL[k, 2], aw=aw, epsilon=ep,
                      sigma=sg, charge=0.0, ghost_type=0, interp_type=1 if nm == "CH3" else 2))
    kb, r0 = 905800.0, 1.54
    kt, t0 = 62500.0, 114.0 * math.pi / 180.0
    return SpeciesSpec(
        name="n-butane", nmole=nmole, sites=s,
        bonds=[(0, 1, abi.BOND_HARMONIC, kb, r0, 0.0), (1, 2, abi.BOND_HARMONIC, kb, r0, 0.0),
               (2, 3, abi.BOND_HARMONIC, kb, r0, 0.0)],
        angles=[(0, 1, 2, abi.ANGLE_HARMONIC, kt, t0, 0.0, 0.0, 0.0),
                (1, 2, 3, abi.ANGLE_HARMONIC, kt, t0, 0.0, 0.0, 0.0)],
        dihedrals=[(0, 1, 2, 3, abi.DIH_OPLS_COSIN, 0.0, 710.06, -136.38, 1582.64)])


def butane_box(nmol: int = 128, seed: int = 3, rc: float = 14.0,
               min_dist: float = 2.8, delt_fs: float = 1.0,
               nstep_inner: int = 5) -> RunDir:
    """One box of united-atom n-butane at 0.58 g/cm^3, rc = min(rc, L/2)."""
    rng = np.random.default_rng(seed)
    L = (nmol / BUTANE_DENSITY) ** (1.0 / 3.0)
    rcut = min(rc, 0.5 * L * (1 - 1e-12))
    # chains roughly along z on a lattice of columns (>= 6.9 A apart along
    # z, about 4.4 A apart across): a random isotropic packing of rigid
    # chains jams at this density
    nz = max(1, int(L / 6.9))
    nxy = int(math.ceil(math.sqrt(nmol / nz)))
    sites = np.array([[i, j, k] for k in range(nz) for i in range(nxy)
                      for j in range(nxy)], dtype=np.float64)
    pick = rng.permutation(len(sites))[:nmol]
    centers = (sites[pick] + 0.5) * np.array([L / nxy, L / nxy, L / nz])
    tmpl = BUTANE_TMPL - BUTANE_TMPL.mean(0)
    # principal (longest) axis of the chain along z
    w, V = np.linalg.eigh(tmpl.T @ tmpl)
    tmpl = tmpl @ V[:, ::-1][:, [1, 2, 0]]
    xyz = _place_rigid(rng, centers, tmpl, L, min_dist, jitter=0.2,
                       rotation=lambda g: _spin_tilt_rotation(g, 0.15)).reshape(-1, 3)
    run = RunInput(title=f"n-butane {nmol}", ij=1000 + seed % 30000, jk=2000 + seed % 28000,
                   sigma_base=1.0, epsilon_base=1.0, mass_base=1.0, treq=300.0, preq=1.0,
                   boxlx=L, boxly=L, boxlz=L, rcuton=rcut, rcutoff=rcut, rcutoffelec=rcut,
                   isLJlrcOn=0, delt=delt_fs, nstep_inner=nstep_inner, f0=0.0,
                   what_simulation=1, what_ensemble=0, nconstraint=3,
                   iChargeType=abi.ELEC_NONE, nstep_md_per_hmc=10)
    return RunDir(run=run, sysname="n-butane", species=[butane_species(nmol)], xyz=xyz)


# --------------------------------------------------------------------------- #
def methane_tube(natom: int = 2000, length: float = 400.0, box_xy: float = 100.0,
                 tube_r: float = 50.0, seed: int = 5, rc: float = 12.0,
                 sf: int = abi.SF_HYPERGEO, r_fill: Optional[float] = None,
                 min_dist: float = 3.0) -> RunDir:
    """LJ methane (3.73 A, 148 K) in a smooth carbon tube of radius
    tube_r centred on the z axis, axial box `length` (Angstrom)."""
    rng = np.random.default_rng(seed)
    r_fill = tube_r - 4.0 if r_fill is None else r_fill
    pts = []
    cell = min_dist
    grid = {}
    while len(pts) < natom:
        m = 4 * (natom - len(pts))
        r = r_fill * np.sqrt(rng.random(m))
        th = 2 * np.pi * rng.random(m)
        z = length * rng.random(m)
        cand = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
        for c in cand:
            key = (int(c[0] // cell), int(c[1] // cell), int((c[2] % length) // cell))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        kk = (key[0] + dx, key[1] + dy,
                              (key[2] + dz) % max(1, int(length // cell)))
                        for p in grid.get(kk, ()):
                            d = c - p
                            d[2] -= length * np.rint(d[2] / length)
                            if d @ d < min_dist * min_dist:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pts.append(c)
                grid.setdefault(key, []).append(c)
                if len(pts) == natom:
                    break
    xyz = np.array(pts)
    run = RunInput(title=f"methane tube {natom}", sigma_base=1.0, epsilon_base=1.0,
                   mass_base=1.0, treq=300.0, boxlx=box_xy, boxly=box_xy, boxlz=length,
                   rcuton=rc, rcutoff=rc, rcutoffelec=rc, isLJlrcOn=0, delt=2.0,
                   nstep_inner=4, iChargeType=abi.ELEC_NONE, iSF_type=sf,
                   tube_sigma=3.37, tube_epsilon=27.68, tubes=[(0.0, 0.0, tube_r)])
    sp = SpeciesSpec(name="methane", nmole=natom, sites=[
        Site(type=0, name="C", xx=0.0, yy=0.0, zz=0.0, aw=0.016043, epsilon=148.0,
             sigma=3.73, charge=0.0, ghost_type=0, interp_type=1)])
    return RunDir(run=run, sysname="methane-tube", species=[sp], xyz=xyz)


# --------------------------------------------------------------------------- #
def tube_cell_fluid(uclx: float, ucly: float, uclz: float, ncell_z: int = 12,
                    n_ch3: int = 16, n_dimer: int = 4, r_fill: float = 3.5,
                    rc: float = 7.0, seed: int = 21, min_dist: float = 3.3,
                    files=("ch3.cubes", "ch2.cubes")) -> RunDir:
    """United-atom CH3 sites (3.75 A, 98 K, table type 1) and CH2-CH2 dimers
    (3.95 A, 46 K, bond 1.54 A, table type 2) inside a nanotube whose unit
    cell (uclx, ucly, uclz, Angstrom) the myinterp cube files describe; the
    box is one cell across and ncell_z cells long, the fluid within r_fill of
    the axis (uclx/2, ucly/2).  The MYINTERP section names `files`; the
    caller supplies their contents (RunDir.cube_files)."""
    rng = np.random.default_rng(seed)
    Lz = ncell_z * uclz
    cx, cy = uclx / 2.0, ucly / 2.0
    pts = []

    def fits(p):
        for q in pts:
            d = p - q
            d[2] -= Lz * np.rint(d[2] / Lz)
            if d @ d < min_dist * min_dist:
                return False
        return True

    def point():
        r = r_fill * math.sqrt(rng.random())
        th = 2 * math.pi * rng.random()
        return np.array([cx + r * math.cos(th), cy + r * math.sin(th), Lz * rng.random()])

    ch3 = []
    while len(ch3) < n_ch3:
        p = point()
        if fits(p):
            pts.append(p)
            ch3.append(p)
    dim = []
    bond = 1.54
    while len(dim) < n_dimer:
        p = point()
        u = rng.normal(size=3)
        u[:2] *= 0.3                      # roughly along the axis
        u /= np.linalg.norm(u)
        q = p + bond * u
        if math.hypot(q[0] - cx, q[1] - cy) > r_fill or not fits(p) or not fits(q):
            continue
        pts += [p, q]
        dim.append((p, q))
    xyz = np.array(ch3 + [v for pq in dim for v in pq])
    xyz[:, 2] %= Lz
    sp1 = SpeciesSpec(name="CH3", nmole=n_ch3, sites=[
        Site(type=0, name="CH3", xx=0.0, yy=0.0, zz=0.0, aw=0.015035, epsilon=98.0,
             sigma=3.75, charge=0.0, ghost_type=0, interp_type=1)])
    sp2 = SpeciesSpec(name="ethylene-ua", nmole=n_dimer, sites=[
        Site(type=1, name="CH2", xx=0.0, yy=0.0, zz=0.0, aw=0.014027, epsilon=46.0,
             sigma=3.95, charge=0.0, ghost_type=0, interp_type=2),
        Site(type=1, name="CH2", xx=bond, yy=0.0, zz=0.0, aw=0.014027, epsilon=46.0,
             sigma=3.95, charge=0.0, ghost_type=0, interp_type=2)],
        bonds=[(0, 1, abi.BOND_HARMONIC, 905800.0, bond, 0.0)])
    rcut = min(rc, 0.5 * min(uclx, ucly) * (1 - 1e-12))
    run = RunInput(title="fluid in a nanotube cell (myinterp)", sigma_base=1.0,
                   epsilon_base=1.0, mass_base=1.0, treq=300.0, boxlx=uclx, boxly=ucly,
                   boxlz=Lz, rcuton=rcut, rcutoff=rcut, rcutoffelec=rcut, isLJlrcOn=0,
                   delt=1.0, nstep_inner=4, iChargeType=abi.ELEC_NONE,
                   iSF_type=abi.SF_MY_INTERP, mi_ucl=(uclx, ucly, uclz),
                   mi_center=(cx, cy, 0.0), mi_files=list(files))
    return RunDir(run=run, sysname="tube-fluid", species=[sp1, sp2], xyz=xyz)


# --------------------------------------------------------------------------- #
def armchair_tube(n: int, nper: int, bond: float = 1.42) -> np.ndarray:
    """Carbon atoms of an (n, n) armchair nanotube on the z axis, nper axial
    periods of a = sqrt(3) * bond: the graphene sheet (lattice vectors
    a1, a2 = a (sqrt(3)/2, +-1/2), two-atom basis) cut to the chiral vector
    C = n (a1 + a2) and the translation T = a1 - a2, then rolled.  Returns
    (4 n nper, 3) Angstrom."""
    a = math.sqrt(3.0) * bond
    cx, ty = n * a * math.sqrt(3.0), nper * a
    pts = set()
    m = 2 * (n + nper) + 4
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            X = 0.5 * math.sqrt(3.0) * a * (i + j)
            Y = 0.5 * a * (i - j)
            for dx in (0.0, a / math.sqrt(3.0)):
                u, v = (X + dx) % cx, Y % ty
                pts.add((round(u, 9) % round(cx, 9), round(v, 9) % round(ty, 9)))
    uv = np.array(sorted(pts))
    R = cx / (2.0 * math.pi)
    th = 2.0 * math.pi * uv[:, 0] / cx
    out = np.stack([R * np.cos(th), R * np.sin(th), uv[:, 1]], axis=1)
    assert out.shape[0] == 4 * n * nper, out.shape
    return out


def ghost_diatomic_species(nmole: int) -> SpeciesSpec:
    """A CH3 bead bonded to a ghost (GHOST_LJ) site: the ghost check of the
    solid-fluid loops (src/frclong.c:1154-1199) skips its second site."""
    sites = [Site(type=0, name="CH3", xx=0.0, yy=0.0, zz=0.0, aw=0.015035, epsilon=98.0,
                  sigma=3.75, charge=0.0, ghost_type=abi.GHOST_NONE, interp_type=1),
             Site(type=1, name="X", xx=0.0, yy=0.0, zz=1.0, aw=0.001, epsilon=0.0,
                  sigma=0.0, charge=0.0, ghost_type=abi.GHOST_LJ, interp_type=1)]
    return SpeciesSpec(name="ch3x", nmole=nmole, sites=sites,
                       bonds=[(0, 1, abi.BOND_HARMONIC, 2.0e5, 1.0, 0.0)])


def explicit_tube(n: int = 10, nper: int = 9, box_xy: float = 30.0, n_in: int = 5,
                  n_out: int = 60, n_dimer: int = 20, rc: float = 10.0, switch: bool = False,
                  seed: int = 31, min_dist: float = 3.0) -> RunDir:
    """Methane (3.73 A, 148 K) and CH3-ghost dimers around an explicit
    (n, n) carbon nanotube (iSF_type 2): the wall's atoms are the
    ATOMEXPLICIT section (sigma 3.4 A, epsilon 28 K), the box is periodic
    with the tube's period along z.  n_in methanes sit on the axis inside
    the tube, the rest at random outside it."""
    rng = np.random.default_rng(seed)
    wall = armchair_tube(n, nper)
    R = float(np.hypot(wall[0, 0], wall[0, 1]))
    lz = nper * math.sqrt(3.0) * 1.42
    L = np.array([box_xy, box_xy, lz])

    def far(c, pts, d):
        if len(pts) == 0:
            return True
        dd = np.asarray(pts) - c
        dd -= L * np.rint(dd / L)
        return bool(np.min((dd * dd).sum(axis=1)) > d * d)

    fluid = [np.array([0.0, 0.0, (k + 0.5) * lz / n_in]) for k in range(n_in)]
    while len(fluid) < n_in + n_out:
        c = rng.uniform(-0.5, 0.5, 3) * L
        if np.hypot(c[0], c[1]) > R + min_dist and far(c, wall, min_dist) and \
                far(c, fluid, min_dist):
            fluid.append(c)
    dimers = []
    while len(dimers) < n_dimer:
        c = rng.uniform(-0.5, 0.5, 3) * L
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c2 = c + d
        if np.hypot(c[0], c[1]) > R + min_dist and far(c, wall, min_dist) and \
                far(c, fluid, min_dist) and far(c, [x for p in dimers for x in p], min_dist):
            dimers.append((c, c2))
    xyz = np.vstack([np.array(fluid)] + [np.array(p) for p in dimers])
    run = RunInput(title=f"explicit ({n},{n}) tube", sigma_base=1.0, epsilon_base=1.0,
                   mass_base=1.0, treq=300.0, boxlx=box_xy, boxly=box_xy, boxlz=lz,
                   rcuton=(0.8 * rc if switch else rc), rcutoff=rc, rcutoffelec=rc,
                   isLJswitchOn=1 if switch else 0, isLJlrcOn=0, delt=2.0, nstep_inner=4,
                   iChargeType=abi.ELEC_NONE, iSF_type=abi.SF_ATOM_EXPLICIT,
                   solid_type=0, solid_sigma=3.4, solid_epsilon=28.0, solid_charge=0.0,
                   solid_atoms=[("C", *map(float, w)) for w in wall])
    ch4 = SpeciesSpec(name="methane", nmole=n_in + n_out, sites=[
        Site(type=0, name="C", xx=0.0, yy=0.0, zz=0.0, aw=0.016043, epsilon=148.0,
             sigma=3.73, charge=0.0, ghost_type=0, interp_type=1)])
    return RunDir(run=run, sysname="explicit-tube", species=[ch4, ghost_diatomic_species(n_dimer)],
                  xyz=xyz)


# --------------------------------------------------------------------------- #
def wolf_water_set(nconf: int = 64, nmol: int = 1024, L: float = 31.284, kappa: float = 0.2,
                   rc: float = 12.0, seed0: int = 400) -> list:
    """nconf seeded configurations of nmol flexible waters (water_box's
    model) at the same density as the 12,288-site box, Wolf damping kappa
    (1/A) with cutoff rc (A): random occupations of the 11^3 lattice."""
    n_side = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    return [water_box(n_side=n_side, L=L, seed=seed0 + k, elec="wolf", kappa=kappa, rc=rc,
                      nmol=nmol) for k in range(nconf)]


def butane_table(nrep: int = 256, seed: int = 2024) -> list:
    """The batched-replica workload: nrep n-butane boxes of U{64..512}
    molecules (one seeded draw), box r seeded 1000 + r, at 0.58 g/cm^3 with
    rc = min(14 A, L/2)."""
    rng = np.random.default_rng(seed)
    n = rng.integers(64, 513, size=nrep)
    return [butane_box(nmol=int(k), seed=1000 + r) for r, k in enumerate(n)]


# --------------------------------------------------------------------------- #
def tube_loading(natom: int, seed: int = 0, R: float = 50.0, box_xy: float = 100.0,
                 length: float = 4000.0, rc: float = 12.0, wall_gap: float = 4.0,
                 min_dist: float = 3.0, delt_fs: float = 2.0, nstep_inner: int = 4) -> RunDir:
    """natom LJ methanes (3.73 A, 148 K) inside a smooth tube of radius R on
    the z axis (iSF_type 3: the tube as a tricubic Hermite grid), periodic
    box box_xy^2 x length (A), cutoff rc.  Sites are a random choice of the
    points of a cubic lattice inside r < R - wall_gap, spaced so that about
    1.2 natom points fit, each jittered by up to (spacing - min_dist) / 2 per
    axis; vectorized, so that 100,000 sites take a fraction of a second."""
    rng = np.random.default_rng(seed)
    rf = R - wall_gap
    a = (math.pi * rf * rf * length / (1.2 * natom)) ** (1.0 / 3.0)
    while True:
        if a < min_dist:
            raise ValueError("the tube cannot hold that many sites")
        nxy = int(2 * rf / a) + 1
        nz = int(length / a)
        g = np.arange(nxy) * a - (nxy - 1) * a / 2.0
        X, Y = np.meshgrid(g, g, indexing="ij")
        keep = X * X + Y * Y < (rf - 0.5 * (a - min_dist)) ** 2
        if keep.sum() * nz >= natom:
            break
        a *= 0.98
    xy = np.stack([X[keep], Y[keep]], axis=1)
    zs = (np.arange(nz) + 0.5) * (length / nz)
    pts = np.concatenate([np.column_stack([xy, np.full(len(xy), z)]) for z in zs])
    pick = np.sort(rng.permutation(len(pts))[:natom])
    j = 0.5 * (min(a, length / nz) - min_dist)
    xyz = pts[pick] + rng.uniform(-j, j, size=(natom, 3))
    run = RunInput(title=f"methane tube loading {natom}", sigma_base=1.0, epsilon_base=1.0,
                   mass_base=1.0, treq=300.0, boxlx=box_xy, boxly=box_xy, boxlz=length,
                   rcuton=rc, rcutoff=rc, rcutoffelec=rc, isLJlrcOn=0, delt=delt_fs,
                   nstep_inner=nstep_inner, iChargeType=abi.ELEC_NONE, iSF_type=abi.SF_TASOS,
                   tube_sigma=3.37, tube_epsilon=27.68, tubes=[(0.0, 0.0, R)])
    sp = SpeciesSpec(name="methane", nmole=natom, sites=[
        Site(type=0, name="C", xx=0.0, yy=0.0, zz=0.0, aw=0.016043, epsilon=148.0,
             sigma=3.73, charge=0.0, ghost_type=0, interp_type=1)])
    return RunDir(run=run, sysname="methane-tube", species=[sp], xyz=xyz)


def tube_loadings(n: int = 16, nmin: int = 20000, nmax: int = 100000, seed0: int = 500):
    """n loadings N_i = nmin + i (nmax - nmin) / (n - 1), i = 0..n-1."""
    return [tube_loading(int(round(nmin + i * (nmax - nmin) / (n - 1))), seed=seed0 + i)
            for i in range(n)]


def tasos_column(nsite: int = 8, auc: float = 4.0, cuc: float = 1.2, length: float = 48.0,
                 box_xy: float = 30.0, rc: float = 10.0, n_dimer: int = 3, amp: float = 20.0,
                 seed: int = 41) -> RunDir:
    """Methanes (3.73 A, 148 K) and CH3-ghost dimers in a column along z,
    with the solid-fluid field given as the Fortran library's Hermite grid
    files (iSF_type 3, the TASOS section): a square unit cell of side auc
    and height cuc, sampled by InitPotentialGrid's 0.2 A grid
    (mylibtasos/initialize.f90:96-101), whose field is the smooth tube of
    radius 6 A around the column's axis plus amp * cos(2 pi z / cuc), so
    that the lookup's z wrap and z derivatives take part.  The sites stay
    within 1.2 A of the axis, inside the grid's x-y range."""
    from . import sfgrid
    rng = np.random.default_rng(seed)
    (ngx, ngy, ngz), (hx, hy, hz) = sfgrid.tasos_dims(auc, auc, cuc)
    nx, ny, nz = ngx + 2, ngy + 2, ngz + 2
    centre = 0.5 * auc
    g = sfgrid.tube_grid([(3.73, 148.0)], R=6.0, x0=0.0, y0=0.0, nx=nx, ny=ny, hx=hx, hy=hy,
                         nz=nz, hz=hz, sigma_s=3.37, eps_s=27.68, center=(centre, centre))
    tab = np.asarray(g["nodes"][0], dtype=np.float64)
    zs = hz * np.arange(nz)
    w = 2.0 * np.pi / cuc
    tab[0] += amp * np.cos(w * zs)[:, None, None]
    tab[3] += -amp * w * np.sin(w * zs)[:, None, None]
    ntot = nsite + n_dimer
    zk = (np.arange(ntot) + 0.5) * (length / ntot)
    xy = centre + rng.uniform(-1.2, 1.2, size=(ntot, 2))
    pts = np.column_stack([xy, zk])
    order = rng.permutation(ntot)
    meth = pts[order[:nsite]]
    dim = pts[order[nsite:]]
    ghost = dim + np.array([0.0, 0.0, 1.0])
    xyz = np.concatenate([meth, np.stack([dim, ghost], axis=1).reshape(-1, 3)])
    run = RunInput(title="column on a Hermite grid (TASOS files)", sigma_base=1.0,
                   epsilon_base=1.0, mass_base=1.0, treq=148.0, boxlx=box_xy, boxly=box_xy,
                   boxlz=length, rcuton=rc, rcutoff=rc, rcutoffelec=rc, isLJlrcOn=0, delt=2.0,
                   nstep_inner=4, iChargeType=abi.ELEC_NONE, iSF_type=abi.SF_TASOS,
                   tasos_uc=(auc, auc, cuc), tasos_radius=6.0,
                   tasos_ncell=(1, 1, int(round(length / cuc))), tasos_ngrid=1,
                   tasos_files=["ch4.grid"])
    sp = SpeciesSpec(name="methane", nmole=nsite, sites=[
        Site(type=0, name="C", xx=0.0, yy=0.0, zz=0.0, aw=0.016043, epsilon=148.0,
             sigma=3.73, charge=0.0, ghost_type=0, interp_type=1)])
    return RunDir(run=run, sysname="tasos-column", species=[sp, ghost_diatomic_species(n_dimer)],
                  xyz=xyz, grid_files={"ch4.grid": tab})
