"""mspms2 run-directory formats: in.mspms, cfg.mspms, xyz.mspms.

Reader and writer for the three text files the reference reads at start-up
(src/readins.c:21-263 and the keyword sections of src/initvars.c:248-472,
src/sffrc.c:16-78).  Values stay in the files' units (Angstrom, K, kg/mol,
fs); mspms2_amd.system applies the reference's unit reduction.

Numbers are written with 17 significant digits so that a write/read round
trip (and the reference's own sscanf("%lf")) reproduces every double exactly.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


def _g(x) -> str:
    return repr(float(x)) if np.isfinite(x) else str(x)


@dataclass
class Site:
    """One atom record of cfg.mspms (src/readins.c:118-126)."""
    type: int
    name: str
    xx: float
    yy: float
    zz: float
    aw: float          # kg/mol
    epsilon: float     # K
    sigma: float       # Angstrom
    charge: float      # e
    ghost_type: int = 0
    interp_type: int = 0
    ee: Optional[float] = None
    ff: Optional[float] = None
    gg: Optional[float] = None


@dataclass
class SpeciesSpec:
    """One species block of cfg.mspms (src/readins.c:101-243)."""
    name: str
    nmole: int
    sites: List[Site]
    bonds: List[Tuple] = field(default_factory=list)      # (i, j, type, Kb, Req, alpha)
    angles: List[Tuple] = field(default_factory=list)     # (i, j, k, type, p1..p5)
    dihedrals: List[Tuple] = field(default_factory=list)  # (i, j, k, l, type, c1..c4)
    impropers: List[Tuple] = field(default_factory=list)  # (i, j, k, l, type, komega, omega0)
    nbp: List[Tuple] = field(default_factory=list)        # (i, j)

    @property
    def natom(self) -> int:
        return len(self.sites)


@dataclass
class RunInput:
    """The fixed-line part of in.mspms (src/readins.c:29-82) plus the keyword
    sections used by the force path."""
    title: str = "mspms2_amd run"
    ij: int = 1802
    jk: int = 9373
    sigma_base: float = 1.0
    epsilon_base: float = 1.0
    mass_base: float = 1.0
    treq: float = 298.0
    preq: float = 1.0
    boxlx: float = 10.0
    boxly: float = 10.0
    boxlz: float = 10.0
    rcuton: float = 5.0
    rcutoff: float = 5.0
    rcutoffelec: float = 5.0
    isLJlrcOn: int = 0
    isLJswitchOn: int = 0
    iStart_option: int = 1
    nstep: int = 10
    nstep_eq: int = 0
    nstep_ave: int = 10
    nstep_print: int = 1
    nstep_save: int = 1000000
    nstep_ss: int = 1000000
    nstep_trj: int = 0
    delt: float = 1.0
    nstep_inner: int = 1
    f0: float = 0.0
    what_simulation: int = 0
    what_ensemble: int = 0
    nconstraint: int = 3
    iInterMolePotType: int = 0
    iChargeType: int = 0
    iSF_type: int = 0
    iExternal_FF_type: int = 0
    # ELECTROSTATIC section (src/initvars.c:416-456)
    kappa: float = 0.0
    kmaxx: int = 0
    kmaxy: int = 0
    kmaxz: int = 0
    ksqmax: int = 0
    fEwald_BC: int = 1
    fEwald_Dim: int = 3
    # MDNVT (src/initvars.c:310-347) / MDNPT
    qq: float = 100.0
    qqs: float = 100.0
    Qts: float = 1000.0
    Qbs: float = 1000.0
    # HYPERGEO (src/sffrc.c:35-69)
    tube_sigma: float = 3.4
    tube_epsilon: float = 28.0
    tubes: List[Tuple[float, float, float]] = field(default_factory=list)
    # ATOMEXPLICIT (init_sf_atom_explicit, src/sffrc.c:81-147): solid type,
    # uniform solid sigma (A), epsilon (K) and charge, then one
    # (name, x, y, z) line per solid atom (A); written only when atoms are
    # given
    solid_type: int = 0
    solid_sigma: float = 3.4
    solid_epsilon: float = 28.0
    solid_charge: float = 0.0
    solid_atoms: List[Tuple[str, float, float, float]] = field(default_factory=list)
    # HMC (src/initvars.c:248-290)
    nstep_md_per_hmc: int = 10
    pdisp: float = 1.0
    pvolm: float = 0.0
    pmake: float = 0.0
    pkill: float = 0.0
    rreq_disp: float = 0.5
    rreq_volm: float = 0.5
    delv: float = 100.0
    nstep_delt_adj_cycle: int = 100
    nstep_delv_adj_cycle: int = 100
    # SIMAN (init_siman, src/siman.c:8-59): NVT steps per temperature, print
    # interval; initial and final temperature (K), cooling divisor
    n_tries: int = 20
    iters_fixed_t: int = 10
    t_initial: float = 300.0
    t_min: float = 100.0
    mu_t: float = 1.5
    # MYINTERP (init_my_interp, src/sffrc.c:262-297): unit cell and centre
    # (Angstrom), one cube-polynomial file per table type; written only
    # when files are named
    mi_ucl: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    mi_center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    mi_files: List[str] = field(default_factory=list)
    # TASOS (init_tasos_grid, src/sffrc.c:149-221): unit cell a b c and the
    # tube radius (Angstrom), cell counts, the number of grids read, then
    # one grid file per unique atom; written only when files are named
    tasos_uc: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    tasos_radius: float = 0.0
    tasos_ncell: Tuple[int, int, int] = (1, 1, 1)
    tasos_ngrid: int = 0
    tasos_files: List[str] = field(default_factory=list)


_FIXED = [  # (attribute, format) in src/readins.c:31-82 order
    ("ij", "d"), ("jk", "d"), ("sigma_base", "g"), ("epsilon_base", "g"),
    ("mass_base", "g"), ("treq", "g"), ("preq", "g"), ("boxlx", "g"), ("boxly", "g"),
    ("boxlz", "g"), ("rcuton", "g"), ("rcutoff", "g"), ("rcutoffelec", "g"),
    ("isLJlrcOn", "d"), ("isLJswitchOn", "d"), ("iStart_option", "d"), ("nstep", "d"),
    ("nstep_eq", "d"), ("nstep_ave", "d"), ("nstep_print", "d"), ("nstep_save", "d"),
    ("nstep_ss", "d"), ("nstep_trj", "d"), ("delt", "g"), ("nstep_inner", "d"),
    ("f0", "g"), ("what_simulation", "d"), ("what_ensemble", "d"), ("nconstraint", "d"),
    ("iInterMolePotType", "d"), ("iChargeType", "d"), ("iSF_type", "d"),
    ("iExternal_FF_type", "d"),
]


def _fmt(v, kind):
    return str(int(v)) if kind == "d" else _g(v)


def write_in(path: str, r: RunInput) -> None:
    lines = [r.title.strip().replace("\n", " ") or "mspms2_amd run"]
    for name, kind in _FIXED:
        lines.append(f"{_fmt(getattr(r, name), kind):<24s} {name}")
    lines.append("")
    lines.append("   Electrostatic")
    if r.iChargeType == 2:
        lines.append(f"{_g(r.kappa):<24s} kappa (1/A)")
    else:
        lines += [f"{_g(r.kappa):<24s} kappa (1/A)", f"{r.kmaxx:<24d} kmaxx",
                  f"{r.kmaxy:<24d} kmaxy", f"{r.kmaxz:<24d} kmaxz",
                  f"{r.ksqmax:<24d} ksqmax", f"{r.fEwald_BC:<24d} boundary",
                  f"{r.fEwald_Dim:<24d} dimension"]
    lines.append("")
    lines.append("   MDNVT")
    lines.append(f"{_g(r.qq)} {_g(r.qqs)}      qq qqs")
    lines.append("")
    lines.append("   MDNPT")
    lines.append(f"{_g(r.Qts)} {_g(r.Qbs)}      Qts Qbs")
    lines.append("")
    lines.append("   HyperGeo")
    lines.append(f"{len(r.tubes):<24d} ntube")
    lines.append(f"{_g(r.tube_sigma):<24s} tube sigma (A)")
    lines.append(f"{_g(r.tube_epsilon):<24s} tube epsilon (K)")
    for (tx, ty, tr) in r.tubes:
        lines.append(f"{_g(tx)} {_g(ty)} {_g(tr)}      x y radius (A)")
    if r.solid_atoms:
        lines.append("")
        lines.append("   AtomExplicit")
        lines.append(f"{len(r.solid_atoms):<24d} solid_natom")
        lines.append(f"{r.solid_type} {_g(r.solid_sigma)} {_g(r.solid_epsilon)} "
                     f"{_g(r.solid_charge)}      type sigma (A) epsilon (K) charge")
        for (nm, sx, sy, sz) in r.solid_atoms:
            lines.append(f"{nm} {repr(float(sx))} {repr(float(sy))} {repr(float(sz))}")
    lines.append("")
    lines.append("  HMC")
    for name, kind in [("nstep_md_per_hmc", "d"), ("pdisp", "g"), ("pvolm", "g"),
                       ("pmake", "g"), ("pkill", "g"), ("rreq_disp", "g"),
                       ("rreq_volm", "g"), ("delv", "g"), ("nstep_delt_adj_cycle", "d"),
                       ("nstep_delv_adj_cycle", "d")]:
        lines.append(f"{_fmt(getattr(r, name), kind):<24s} {name}")
    lines.append("")
    lines.append("   SIMAN")
    lines.append(f"{r.n_tries} {r.iters_fixed_t}      n_tries iters_fixed_t")
    lines.append(f"{_g(r.t_initial)} {_g(r.t_min)} {_g(r.mu_t)}      t_initial t_min mu_t (K)")
    if r.mi_files:
        lines.append("")
        lines.append("   MYINTERP")
        lines.append(" ".join(repr(float(v)) for v in r.mi_ucl) + "      uclx ucly uclz (A)")
        lines.append(" ".join(repr(float(v)) for v in r.mi_center) +
                     "      xcenter ycenter zcenter (A)")
        lines.append(f"{len(r.mi_files):<24d} nunique_atom")
        lines += list(r.mi_files)
    if r.tasos_files:
        lines.append("")
        lines.append("   TASOS")
        lines.append(" ".join(repr(float(v)) for v in r.tasos_uc) +
                     f" {repr(float(r.tasos_radius))}      auc buc cuc radius (A)")
        lines.append(" ".join(str(int(v)) for v in r.tasos_ncell) + "      na nb nc")
        lines.append(f"{r.tasos_ngrid:<24d} nspecies_yang")
        lines.append(f"{len(r.tasos_files):<24d} tnuatoms")
        lines += list(r.tasos_files)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _section(lines, keyword):
    for i, ln in enumerate(lines):
        tok = ln.split()
        if tok and tok[0].upper() == keyword:
            return i + 1
    return None


def read_in(path: str) -> RunInput:
    with open(path) as f:
        lines = f.read().split("\n")
    r = RunInput()
    r.title = lines[0]
    for k, (name, kind) in enumerate(_FIXED):
        tok = lines[1 + k].split()[0]
        setattr(r, name, int(tok) if kind == "d" else float(tok))
    i = _section(lines, "ELECTROSTATIC")
    if i is not None:
        r.kappa = float(lines[i].split()[0])
        if r.iChargeType != 2:
            r.kmaxx, r.kmaxy, r.kmaxz, r.ksqmax, r.fEwald_BC, r.fEwald_Dim = (
                int(lines[i + k].split()[0]) for k in range(1, 7))
    i = _section(lines, "MDNVT")
    if i is not None:
        a = lines[i].split()
        r.qq, r.qqs = float(a[0]), float(a[1])
    i = _section(lines, "MDNPT")
    if i is not None:
        a = lines[i].split()
        r.Qts, r.Qbs = float(a[0]), float(a[1])
    i = _section(lines, "HYPERGEO")
    if i is not None:
        nt = int(lines[i].split()[0])
        r.tube_sigma = float(lines[i + 1].split()[0])
        r.tube_epsilon = float(lines[i + 2].split()[0])
        r.tubes = [tuple(float(v) for v in lines[i + 3 + t].split()[:3]) for t in range(nt)]
    i = _section(lines, "ATOMEXPLICIT")
    if i is not None:
        # src/sffrc.c:101-128: the count; "type sigma epsilon charge" on
        # one line (read twice there: %d, then %d %lf %lf %lf); then one
        # "name x y z" per atom
        n = int(lines[i].split()[0])
        a = lines[i + 1].split()
        r.solid_type = int(a[0])
        r.solid_sigma, r.solid_epsilon, r.solid_charge = float(a[1]), float(a[2]), float(a[3])
        toks = " ".join(lines[i + 2:]).split()
        r.solid_atoms = [(toks[4 * k], float(toks[4 * k + 1]), float(toks[4 * k + 2]),
                          float(toks[4 * k + 3])) for k in range(n)]
    i = _section(lines, "HMC")
    if i is not None:
        vals = [lines[i + k].split()[0] for k in range(10)]
        r.nstep_md_per_hmc = int(vals[0])
        r.pdisp, r.pvolm, r.pmake, r.pkill, r.rreq_disp, r.rreq_volm, r.delv = map(float, vals[1:8])
        r.nstep_delt_adj_cycle, r.nstep_delv_adj_cycle = int(vals[8]), int(vals[9])
    i = _section(lines, "SIMAN")
    if i is not None:
        a, b = lines[i].split(), lines[i + 1].split()
        r.n_tries, r.iters_fixed_t = int(a[0]), int(a[1])
        r.t_initial, r.t_min, r.mu_t = float(b[0]), float(b[1]), float(b[2])
    i = _section(lines, "MYINTERP")
    if i is not None:
        r.mi_ucl = tuple(float(v) for v in lines[i].split()[:3])
        r.mi_center = tuple(float(v) for v in lines[i + 1].split()[:3])
        n = int(lines[i + 2].split()[0])
        r.mi_files = [lines[i + 3 + k].split()[0] for k in range(n)]
    i = _section(lines, "TASOS")
    if i is not None:
        a = lines[i].split()
        r.tasos_uc = (float(a[0]), float(a[1]), float(a[2]))
        r.tasos_radius = float(a[3])
        r.tasos_ncell = tuple(int(v) for v in lines[i + 1].split()[:3])
        r.tasos_ngrid = int(lines[i + 2].split()[0])
        n = int(lines[i + 3].split()[0])
        r.tasos_files = [lines[i + 4 + k].split()[0] for k in range(n)]
    return r


def write_cfg(path: str, sysname: str, species: List[SpeciesSpec]) -> None:
    out = [sysname.split()[0] if sysname.split() else "system", f"{len(species)} species"]
    for sp in species:
        out.append(sp.name.split()[0])
        out.append(f"{sp.nmole} molecules")
        out.append(f"{sp.natom} atoms per molecule")
        for s in sp.sites:
            ee = s.xx if s.ee is None else s.ee
            ff = s.yy if s.ff is None else s.ff
            gg = s.zz if s.gg is None else s.gg
            out.append(" ".join([str(int(s.type)), s.name] + [_g(v) for v in (
                s.xx, s.yy, s.zz, ee, ff, gg, s.aw, s.epsilon, s.sigma, s.charge)] +
                [str(int(s.ghost_type)), str(int(s.interp_type))]))
        out.append(f"{len(sp.bonds)} bonds per molecule")
        for b in sp.bonds:
            out.append(" ".join([str(int(v)) for v in b[:3]] + [_g(v) for v in b[3:6]]))
        out.append(f"{len(sp.angles)} angles per molecule")
        for a in sp.angles:
            out.append(" ".join([str(int(v)) for v in a[:4]] + [_g(v) for v in a[4:9]]))
        out.append(f"{len(sp.dihedrals)} dihedrals per molecule")
        for d in sp.dihedrals:
            out.append(" ".join([str(int(v)) for v in d[:5]] + [_g(v) for v in d[5:9]]))
        out.append(f"{len(sp.impropers)} impropers per molecule")
        for m in sp.impropers:
            out.append(" ".join([str(int(v)) for v in m[:5]] + [_g(v) for v in m[5:7]]))
        out.append(f"{len(sp.nbp)} nonbonded per molecule")
        for n in sp.nbp:
            out.append(f"{int(n[0])} {int(n[1])}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_cfg(path: str) -> Tuple[str, List[SpeciesSpec]]:
    with open(path) as f:
        lines = [ln for ln in f.read().split("\n")]
    pos = 0

    def nxt():
        nonlocal pos
        while pos < len(lines) and not lines[pos].strip():
            pos += 1
        ln = lines[pos]
        pos += 1
        return ln.split()

    sysname = nxt()[0]
    nspecie = int(nxt()[0])
    species = []
    for _ in range(nspecie):
        name = nxt()[0]
        nmole = int(nxt()[0])
        natom = int(nxt()[0])
        sites = []
        for _a in range(natom):
            t = nxt()
            sites.append(Site(type=int(t[0]), name=t[1], xx=float(t[2]), yy=float(t[3]),
                              zz=float(t[4]), ee=float(t[5]), ff=float(t[6]), gg=float(t[7]),
                              aw=float(t[8]), epsilon=float(t[9]), sigma=float(t[10]),
                              charge=float(t[11]), ghost_type=int(t[12]),
                              interp_type=int(t[13])))
        sp = SpeciesSpec(name=name, nmole=nmole, sites=sites)
        for _b in range(int(nxt()[0])):
            t = nxt()
            sp.bonds.append((int(t[0]), int(t[1]), int(t[2]), float(t[3]), float(t[4]),
                             float(t[5])))
        for _b in range(int(nxt()[0])):
            t = nxt()
            sp.angles.append(tuple([int(v) for v in t[:4]] + [float(v) for v in t[4:9]]))
        for _b in range(int(nxt()[0])):
            t = nxt()
            sp.dihedrals.append(tuple([int(v) for v in t[:5]] + [float(v) for v in t[5:9]]))
        for _b in range(int(nxt()[0])):
            t = nxt()
            sp.impropers.append(tuple([int(v) for v in t[:5]] + [float(v) for v in t[5:7]]))
        for _b in range(int(nxt()[0])):
            t = nxt()
            sp.nbp.append((int(t[0]), int(t[1])))
        species.append(sp)
    return sysname, species


def write_xyz(path: str, names: List[str], xyz: np.ndarray, comment: str = "") -> None:
    xyz = np.asarray(xyz, dtype=np.float64)
    out = [str(len(names)), comment.replace("\n", " ")]
    for n, (x, y, z) in zip(names, xyz):
        out.append(f"{n} {_g(x)} {_g(y)} {_g(z)}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_xyz(path: str, natom: int) -> np.ndarray:
    with open(path) as f:
        lines = f.read().split("\n")
    xyz = np.empty((natom, 3), dtype=np.float64)
    k = 0
    for ln in lines[2:]:
        t = ln.split()
        if len(t) < 4:
            continue
        xyz[k] = (float(t[1]), float(t[2]), float(t[3]))
        k += 1
        if k == natom:
            break
    if k != natom:
        raise ValueError(f"{path}: {k} coordinates for {natom} atoms")
    return xyz


@dataclass
class RunDir:
    run: RunInput
    sysname: str
    species: List[SpeciesSpec]
    xyz: np.ndarray            # (natom, 3) Angstrom, reference atom order
    # old.mspms, read where the starting option is not NEW (init_vars ->
    # loadit, src/initvars.c:678-681): it replaces the positions, the
    # velocities, the thermostat / barostat variables and the box
    checkpoint: Optional[object] = None
    # myinterp cube files (file name of the MYINTERP section -> CubeFile)
    cube_files: Optional[dict] = None
    # Hermite grid files of the TASOS section (file name -> [8][nz][ny][nx]
    # node array, sfgrid.read_tasos_table), the first nspecies_yang of them
    grid_files: Optional[dict] = None

    @property
    def natom(self) -> int:
        return sum(sp.natom * sp.nmole for sp in self.species)

    def atom_names(self) -> List[str]:
        out = []
        for sp in self.species:
            for _ in range(sp.nmole):
                out += [s.name for s in sp.sites]
        return out

    def write(self, path: str) -> str:
        os.makedirs(path, exist_ok=True)
        write_in(os.path.join(path, "in.mspms"), self.run)
        write_cfg(os.path.join(path, "cfg.mspms"), self.sysname, self.species)
        write_xyz(os.path.join(path, "xyz.mspms"), self.atom_names(), self.xyz)
        if self.checkpoint is not None:
            from .rundata import LOAD, write_checkpoint
            write_checkpoint(os.path.join(path, LOAD), self.checkpoint)
        if self.cube_files:
            from .myinterp import write_cubes
            for name, cf in self.cube_files.items():
                write_cubes(os.path.join(path, name), cf)
        if self.grid_files:
            from .sfgrid import write_tasos_grid
            for name, tab in self.grid_files.items():
                write_tasos_grid(os.path.join(path, name), tab, *self.run.tasos_uc)
        return path

    @staticmethod
    def read(path: str) -> "RunDir":
        run = read_in(os.path.join(path, "in.mspms"))
        sysname, species = read_cfg(os.path.join(path, "cfg.mspms"))
        natom = sum(sp.natom * sp.nmole for sp in species)
        xyz = read_xyz(os.path.join(path, "xyz.mspms"), natom)
        ck = None
        old = os.path.join(path, "old.mspms")
        if run.iStart_option != 1 and os.path.exists(old):
            from .rundata import read_checkpoint
            ck = read_checkpoint(old, natom)
        cubes = None
        if run.iSF_type == 4 and run.mi_files:
            from .myinterp import read_cubes
            cubes = {n: read_cubes(os.path.join(path, n)) for n in run.mi_files}
        grids = None
        if run.iSF_type == 3 and run.tasos_files:
            # Read_Grids reads nspecies_yang grids (mylibtasos/grid.f90:69)
            from .sfgrid import read_tasos_table
            grids = {n: read_tasos_table(os.path.join(path, n), *run.tasos_uc)
                     for n in run.tasos_files[:run.tasos_ngrid]}
        return RunDir(run=run, sysname=sysname, species=species, xyz=xyz, checkpoint=ck,
                      cube_files=cubes, grid_files=grids)
