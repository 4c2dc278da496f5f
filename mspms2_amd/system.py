"""Reduced-unit system description, built the way the reference builds it.

`System.from_rundir` applies the unit reduction of src/readins.c:33-260 and
the section readers of src/initvars.c:416-462 / src/sffrc.c:35-69 with the
same floating-point expressions, so every reduced parameter and coordinate is
bit-identical to the reference's globals.  `System.c_structs()` packs it into
the C ABI structs of include/mspms_gpu.h (mspms2_amd.abi).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi
from .runio import RunDir

AVOGADRO = 6.0221415e23          # src/defs.h:89
BOLTZMAN_CONSTANT = 1.3806505e-23  # src/defs.h:90
COULOMB_PREFACTOR = 167100.80840   # src/defs.h:94
C3_PI_SQ_THETA = 11.3105666436484  # src/defs.h:99


@dataclass
class ReducedSpecies:
    """SAMPLE_MOLECULE after src/readins.c:118-231 (reduced units)."""
    name: str
    nmole: int
    natom: int
    ghost_type: np.ndarray
    interp_type: np.ndarray
    aw: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray
    charge: np.ndarray
    mw: float
    bnd_idx: np.ndarray
    bnd_type: np.ndarray
    bnd_para: np.ndarray
    agl_idx: np.ndarray
    agl_type: np.ndarray
    agl_para: np.ndarray
    dih_idx: np.ndarray
    dih_type: np.ndarray
    dih_para: np.ndarray
    imp_idx: np.ndarray
    imp_type: np.ndarray
    imp_para: np.ndarray
    nnbp: int = 0
    agl_unique: Optional[np.ndarray] = None
    dih_unique: Optional[np.ndarray] = None


def _arr(x, dt, shape):
    a = np.asarray(x, dtype=dt)
    if a.size == 0:
        return np.zeros(shape, dtype=dt)
    return np.ascontiguousarray(a.reshape(shape))


def reduce_species(sp, sb: float, eb: float, mb: float) -> ReducedSpecies:
    """src/readins.c:118-231, expression for expression."""
    n = sp.natom
    aw = np.empty(n)
    sig = np.empty(n)
    eps = np.empty(n)
    mw = 0.0
    for k, s in enumerate(sp.sites):
        sig[k] = s.sigma / sb
        eps[k] = s.epsilon / eb
        aw[k] = (s.aw / AVOGADRO) / (mb * 1.0e-27)
        mw += aw[k]
    bpara = []
    for (i, j, t, kb, req, alpha) in sp.bonds:
        kb = kb / eb if t == abi.BOND_MORSE else kb * sb * sb / eb
        bpara.append((kb, req / sb, alpha / sb))
    apara = []
    for a in sp.angles:
        t = a[3]
        p1, p2, p3, p4, p5 = a[4:9]
        if t == abi.ANGLE_TR_WATER:
            p1 = p1 * sb * sb / eb
            p2 = p2 * sb * sb / eb
            p3 = p3 * sb * sb / eb
            p4 = p4 / sb
            p5 = p5 / sb
        else:
            p1 = p1 / eb
        apara.append((p1, p2, p3, p4, p5))
    dpara = []
    for d in sp.dihedrals:
        t = d[4]
        c = list(d[5:9])
        if t == abi.DIH_OPLS_COSIN:
            c = [v / eb for v in c]
        elif t == abi.DIH_CHARMM:
            c[0] = c[0] / eb
        dpara.append(tuple(c))
    ipara = []
    for m in sp.impropers:
        ko, om = m[5], m[6]
        if m[4] == abi.IMP_CHARMM:
            ko = ko / eb
        ipara.append((ko, om))
    nb, na, nd, ni = len(sp.bonds), len(sp.angles), len(sp.dihedrals), len(sp.impropers)
    return ReducedSpecies(
        name=sp.name, nmole=sp.nmole, natom=n,
        ghost_type=np.array([s.ghost_type for s in sp.sites], dtype=np.int32),
        interp_type=np.array([s.interp_type for s in sp.sites], dtype=np.int32),
        aw=aw, sigma=sig, epsilon=eps,
        charge=np.array([s.charge for s in sp.sites], dtype=np.float64), mw=mw,
        bnd_idx=_arr([b[:2] for b in sp.bonds], np.int32, (nb, 2)),
        bnd_type=_arr([b[2] for b in sp.bonds], np.int32, (nb,)),
        bnd_para=_arr(bpara, np.float64, (nb, 3)),
        agl_idx=_arr([a[:3] for a in sp.angles], np.int32, (na, 3)),
        agl_type=_arr([a[3] for a in sp.angles], np.int32, (na,)),
        agl_para=_arr(apara, np.float64, (na, 5)),
        dih_idx=_arr([d[:4] for d in sp.dihedrals], np.int32, (nd, 4)),
        dih_type=_arr([d[4] for d in sp.dihedrals], np.int32, (nd,)),
        dih_para=_arr(dpara, np.float64, (nd, 4)),
        imp_idx=_arr([m[:4] for m in sp.impropers], np.int32, (ni, 4)),
        imp_type=_arr([m[4] for m in sp.impropers], np.int32, (ni,)),
        imp_para=_arr(ipara, np.float64, (ni, 2)),
        nnbp=len(sp.nbp))


@dataclass
class System:
    """One mspms2 system in the reference's reduced units."""
    species: List[ReducedSpecies]
    params: dict                     # mspms_params input fields
    xx: np.ndarray
    yy: np.ndarray
    zz: np.ndarray
    run: Optional[object] = None     # the RunInput it came from
    time_base: float = 1.0
    delt: float = 0.0
    treq: float = 0.0
    preq: float = 0.0                # reduced pressure (src/readins.c:38-47)
    v0: Optional[np.ndarray] = None  # velocities of a restart (old.mspms), reduced
    checkpoint: Optional[object] = None
    grid: Optional[dict] = None      # SF_TASOS grid (mspms_sf_grid fields)
    cubes: Optional[dict] = None     # SF_MY_INTERP tables (myinterp.engine_tables)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def natom(self) -> int:
        return sum(s.natom * s.nmole for s in self.species)

    @property
    def nsite(self) -> int:
        return sum(s.natom for s in self.species)

    @staticmethod
    def from_rundir(rd: RunDir) -> "System":
        r = rd.run
        sb, eb, mb = r.sigma_base, r.epsilon_base, r.mass_base
        # src/readins.c:37-59
        time_base = np.sqrt(mb * 1.0e-27 / (eb * BOLTZMAN_CONSTANT)) * sb * 1.0e5
        species = [reduce_species(sp, sb, eb, mb) for sp in rd.species]
        xyz = rd.xyz
        xx = np.ascontiguousarray(xyz[:, 0] / sb)
        yy = np.ascontiguousarray(xyz[:, 1] / sb)
        zz = np.ascontiguousarray(xyz[:, 2] / sb)
        # the ELECTROSTATIC section is read only for Ewald (kappa, KMAX*,
        # KSQMAX, boundary) and Wolf (kappa) (src/initvars.c:421-459)
        ewald = r.iChargeType == abi.ELEC_EWALD
        p = dict(
            sigma_base=sb, epsilon_base=eb,
            boxlx=r.boxlx / sb, boxly=r.boxly / sb, boxlz=r.boxlz / sb,
            rcuton=r.rcuton / sb, rcutoff=r.rcutoff / sb, rcutoffelec=r.rcutoffelec / sb,
            isLJswitchOn=r.isLJswitchOn, f0=r.f0, iChargeType=r.iChargeType,
            kappa=r.kappa * sb if r.iChargeType in (abi.ELEC_EWALD, abi.ELEC_WOLF) else 0.0,
            kmaxx=r.kmaxx if ewald else 0, kmaxy=r.kmaxy if ewald else 0,
            kmaxz=r.kmaxz if ewald else 0, ksqmax=r.ksqmax if ewald else 0,
            fEwald_BC=r.fEwald_BC if ewald else 0, iSF_type=r.iSF_type,
            iExternal_FF_type=r.iExternal_FF_type)
        if r.iSF_type == abi.SF_HYPERGEO:
            p.update(ntube=len(r.tubes),
                     tube_x=np.array([t[0] / sb for t in r.tubes]),
                     tube_y=np.array([t[1] / sb for t in r.tubes]),
                     tube_r=np.array([t[2] / sb for t in r.tubes]),
                     const_3pisq_theta=C3_PI_SQ_THETA * sb * sb,
                     solid_sigma=r.tube_sigma / sb, solid_epsilon=r.tube_epsilon / eb)
        if r.iSF_type == abi.SF_ATOM_EXPLICIT:
            # init_sf_atom_explicit (src/sffrc.c:101-128): reduced by division
            if not r.solid_atoms:
                raise ValueError("iSF_type 2 needs the ATOMEXPLICIT section")
            sa = np.array([a[1:] for a in r.solid_atoms], dtype=np.float64)
            p.update(solid_natom=len(r.solid_atoms),
                     solid_xx=sa[:, 0] / sb, solid_yy=sa[:, 1] / sb, solid_zz=sa[:, 2] / sb,
                     solid_sigma=r.solid_sigma / sb, solid_epsilon=r.solid_epsilon / eb)
        ck = rd.checkpoint
        if ck is not None:
            # loadit (src/loadsave.c:60-128): positions, velocities and box of
            # the checkpoint, in reduced units.  The reference keeps the
            # Ewald constants init_charge derived from the input file's box
            # (src/initvars.c:660-681 run before loadit); they are derived
            # from the checkpoint's box here, which is the same box for
            # every run that does not change it (NVE, NVT, HMC).
            xx = np.ascontiguousarray(ck.x[0])
            yy = np.ascontiguousarray(ck.x[1])
            zz = np.ascontiguousarray(ck.x[2])
            p.update(boxlx=ck.box[0], boxly=ck.box[1], boxlz=ck.box[2])
        cubes = None
        if r.iSF_type == abi.SF_MY_INTERP:
            from .myinterp import engine_tables
            if not rd.cube_files:
                raise ValueError("iSF_type 4 needs the MYINTERP section and its cube files")
            cubes = engine_tables(r.mi_ucl, r.mi_center,
                                  [rd.cube_files[n] for n in r.mi_files])
        grid = None
        if r.iSF_type == abi.SF_TASOS and rd.grid_files:
            from .sfgrid import tasos_grid
            grid = tasos_grid([rd.grid_files[n] for n in r.tasos_files[:r.tasos_ngrid]],
                              *r.tasos_uc)
        pressure_base = BOLTZMAN_CONSTANT * eb * 1.0e25 / (sb * sb * sb)
        s = System(species=species, params=p, xx=xx, yy=yy, zz=zz, run=r,
                   time_base=time_base, delt=r.delt / time_base, treq=r.treq / eb,
                   preq=r.preq / pressure_base,
                   v0=None if ck is None else np.array(ck.v), checkpoint=ck, cubes=cubes,
                   grid=grid)
        return s

    # ------------------------------------------------------------------ #
    def site_arrays(self):
        """Per-atom (species, site, molecule) index arrays in atom order."""
        sp_id, site_id, mol_id = [], [], []
        mol = 0
        for k, s in enumerate(self.species):
            sp_id.append(np.full(s.natom * s.nmole, k, dtype=np.int32))
            site_id.append(np.tile(np.arange(s.natom, dtype=np.int32), s.nmole))
            mol_id.append(np.repeat(np.arange(mol, mol + s.nmole, dtype=np.int32), s.natom))
            mol += s.nmole
        return (np.concatenate(sp_id), np.concatenate(site_id), np.concatenate(mol_id))

    def charges(self) -> np.ndarray:
        return np.concatenate([np.tile(s.charge, s.nmole) for s in self.species])

    def masses(self) -> np.ndarray:
        return np.concatenate([np.tile(s.aw, s.nmole) for s in self.species])

    def c_species(self):
        arr = (abi.Species * len(self.species))()
        keep = []
        for k, s in enumerate(self.species):
            c = arr[k]
            c.natom, c.nmole = s.natom, s.nmole
            c.ghost_type = abi.iptr(s.ghost_type)
            c.interp_type = abi.iptr(s.interp_type)
            c.aw, c.sigma = abi.dptr(s.aw), abi.dptr(s.sigma)
            c.epsilon, c.charge = abi.dptr(s.epsilon), abi.dptr(s.charge)
            c.nbond = len(s.bnd_type)
            c.bnd_idx, c.bnd_type, c.bnd_para = (abi.iptr(s.bnd_idx), abi.iptr(s.bnd_type),
                                                 abi.dptr(s.bnd_para))
            c.nangle = len(s.agl_type)
            c.agl_idx, c.agl_type, c.agl_para = (abi.iptr(s.agl_idx), abi.iptr(s.agl_type),
                                                 abi.dptr(s.agl_para))
            c.agl_unique = abi.u8ptr(s.agl_unique)
            c.ndih = len(s.dih_type)
            c.dih_idx, c.dih_type, c.dih_para = (abi.iptr(s.dih_idx), abi.iptr(s.dih_type),
                                                 abi.dptr(s.dih_para))
            c.dih_unique = abi.u8ptr(s.dih_unique)
            c.nimp = len(s.imp_type)
            c.imp_idx, c.imp_type, c.imp_para = (abi.iptr(s.imp_idx), abi.iptr(s.imp_type),
                                                 abi.dptr(s.imp_para))
            c.nnbp = s.nnbp
            keep.append(s)
        return arr, keep

    def c_params(self, overrides: Optional[dict] = None) -> abi.Params:
        p = abi.Params()
        d = dict(self.params)
        if overrides:
            d.update(overrides)
        keep = []
        for k, v in d.items():
            if k in ("tube_x", "tube_y", "tube_r", "solid_xx", "solid_yy", "solid_zz"):
                a = np.ascontiguousarray(v, dtype=np.float64)
                keep.append(a)
                setattr(p, k, abi.dptr(a))
            else:
                setattr(p, k, v)
        if self.grid is not None:
            g = abi.SFGrid()
            for k in ("ntype", "nx", "ny", "nz", "x0", "y0", "z0", "hx", "hy", "hz"):
                setattr(g, k, self.grid[k])
            g.zperiod = float(self.grid.get("zperiod", 0.0))
            nodes = np.ascontiguousarray(self.grid["nodes"], dtype=np.float32)
            keep.append(nodes)
            g.nodes = abi.fptr(nodes)
            keep.append(g)
            p.grid = C.pointer(g)
        if self.cubes is not None:
            from .myinterp import c_cubes
            q, k2 = c_cubes(self.cubes)
            keep += k2 + [q]
            p.cubes = C.pointer(q)
        p._keep = keep
        return p


def system_of(rundir_obj: RunDir) -> "System":
    """RunDir -> System through a write / read round trip of the run files,
    so that a generated workload is read by the same text path as an
    existing run directory (src/readins.c:21-263)."""
    import tempfile

    from . import runio
    with tempfile.TemporaryDirectory(prefix="mspms_rd_") as tmp:
        rundir_obj.write(tmp)
        return System.from_rundir(runio.RunDir.read(tmp))
