"""Writes the golden fixtures of tests/golden from the reference build.

    python tests/golden/make_golden.py      (needs oracle/_ref, i.e. the
                                             reference checkout; run here)

For each case: the run directory (in/cfg/xyz.mspms, the reference's own
input format) and ref.npz with the reference's outputs after its start-up
path + frclong() + frcshort() (oracle/ref_harness.c): every energy /
virial global, the derived constants, the long and short forces.  The
12,288-site water case keeps only the seed, the SHA-256 of its run files,
all scalars and the forces of every 12th atom.

examples_water holds the input files of the reference's Examples/Water
(flexible Toukan-Rahman water with Morse bonds), unchanged.

MD cases (md_*): the run directory and md<N>.npz with the state after N
steps of the reference's own integrator for the input's ensemble (vver,
vver_nh_3), run from velinit's velocities (stored as v0): positions,
velocities, both forces, the energies of the last evaluation and the
thermostat scalars.
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mspms2_amd import abi, configs  # noqa: E402
from oracle import oracle as O  # noqa: E402

REF = os.environ.get("MSPMS_REF_DIR", "/root/reference")

CASES = {
    "lj864_lattice": lambda: configs.lj_fcc(),
    "lj864_jitter": lambda: configs.lj_fcc(jitter=0.05, seed=7),
    "water216_ewald": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                kmax=5, ksqmax=27, seed=3),
    "water216_wolf": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.2,
                                               elec="wolf", seed=4),
    "water216_coulomb": lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0,
                                                  elec="coulomb", seed=4),
    "butane64": lambda: configs.butane_box(nmol=64, seed=5),
    "tube300_hypergeo": lambda: configs.methane_tube(natom=300, length=60.0, box_xy=40.0,
                                                     tube_r=15.0, rc=10.0),
    "mixture_ewald": lambda: configs.mixture_box(),
    "mixture_water4_vacuum": lambda: configs.mixture_box(water4=True, bc=abi.EWALD_BC_VACUUM,
                                                         seed=12),
    "mixture_wolf_switch": lambda: configs.mixture_box(elec="wolf", kappa=0.2, switch=True,
                                                       seed=13),
    "bonded_zoo": lambda: configs.bonded_zoo(),
    # explicit (10,10) nanotube wall (iSF_type 2), plain LJ and switched LJ
    "tube_explicit": lambda: configs.explicit_tube(),
    "tube_explicit_switch": lambda: configs.explicit_tube(switch=True, seed=32),
}

BIG = {
    "water12288_ewald": lambda: configs.water_box(n_side=16, L=49.66, kappa=0.28, kmax=8,
                                                  ksqmax=64, rc=10.0, seed=2),
}


MD_CASES = {
    # the LJ-864 argon workload: FCC, T* 0.722, NVT qq = qqs = 100, dt* 0.005
    "md_lj864_nvt": (lambda: configs.lj_fcc(jitter=0.05, seed=7), (1, 10, 100)),
    # flexible water, Ewald, RESPA with 4 inner steps, NVT qq = qqs = 200
    "md_water216_nvt": (lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                  kmax=5, ksqmax=27, seed=3, nstep_inner=4,
                                                  ensemble=1), (1, 10)),
    # NPT (npt_respa): LJ with the tail correction, and the water
    "md_lj864_npt": (lambda: configs.lj_fcc(jitter=0.05, seed=7, ensemble=2), (1, 10)),
    "md_water216_npt": (lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                  kmax=5, ksqmax=27, seed=3, nstep_inner=4,
                                                  ensemble=2), (1, 5)),
    # the same water, NVE (vver)
    "md_water216_nve": (lambda: configs.water_box(n_side=6, L=6 * 3.104, rc=9.0, kappa=0.3,
                                                  kmax=5, ksqmax=27, seed=3, nstep_inner=4,
                                                  ensemble=0), (1, 10)),
}


# the reference's own run-state files of a run (saveit / trajectory /
# snapshot), and a restart from its checkpoint (starting option 2, loadit)
FILE_CASE = ("md_water216_nvt", 10)
RESTART = ("md_water216_nvt_restart", 5)


def save_md(path, d, nstep):
    sc = {k: np.float64(v) for k, v in d.items() if isinstance(v, float)}
    md = {f"md_{k}": np.float64(v) for k, v in d["md"].items()}
    np.savez_compressed(os.path.join(path, f"md{nstep}.npz"),
                        x=np.stack([d["xx"], d["yy"], d["zz"]]), v=d["v"], v0=d["v0"],
                        box=np.array([d["boxlx"], d["boxly"], d["boxlz"]]),
                        fl=d["fl"], fs=d["fs"], natom=d["natom"],
                        **{f"s_{k}": v for k, v in sc.items()}, **md)


def _sha(path):
    h = hashlib.sha256()
    for fn in ("in.mspms", "cfg.mspms", "xyz.mspms"):
        with open(os.path.join(path, fn), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def save(path, d, keep_every=1):
    sc = {k: np.float64(v) for k, v in d.items() if isinstance(v, float)}
    idx = np.arange(0, d["natom"], keep_every)
    np.savez_compressed(os.path.join(path, "ref.npz"), fl=d["fl"][:, idx], fs=d["fs"][:, idx],
                        idx=idx, natom=d["natom"], **{f"s_{k}": v for k, v in sc.items()})


def siman_lj256():
    """LJ-256 annealed by the reference's siman: 150 K -> 100 K -> 66.7 K
    (t_min 80 K, mu_t 1.5), 20 NVT steps per temperature."""
    rd = configs.lj_fcc(ncell=4, jitter=0.05, seed=9)
    rd.run.what_simulation, rd.run.what_ensemble = 2, 1
    rd.run.n_tries, rd.run.iters_fixed_t = 20, 10
    rd.run.t_initial, rd.run.t_min, rd.run.mu_t = 150.0, 80.0, 1.5
    return rd


def make_siman_case():
    path = os.path.join(HERE, "siman_lj256")
    os.makedirs(path, exist_ok=True)
    siman_lj256().write(path)
    d = O.run_reference(path, siman=True)
    np.savez_compressed(os.path.join(path, "siman.npz"), x=np.stack([d["xx"], d["yy"], d["zz"]]),
                        v=d["v"], treq=d["md"]["treq"], natom=d["natom"])
    print("siman_lj256", d["md"]["treq"])


def make_sfgridgen_case():
    """The reference's grid generator (oracle/_ref/sfgridgen, built from
    tools/sfgridgen.c) on its own CH3 / (11,11) nanotube input
    (tools/CH3-SWNT11x11.in) at a 1 A interval (16 x 16 x 4 nodes)."""
    path = os.path.join(HERE, "sfgridgen_ch3_1A")
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(REF, "tools", "CH3-SWNT11x11.in")) as f:
        lines = f.read().split("\n")
    t = lines[1].split()
    lines[1] = " ".join(["1.0"] + t[1:])
    with open(os.path.join(path, "grid.in"), "w") as f:
        f.write("\n".join(lines))
    exe = os.path.join(ROOT, "oracle", "_ref", "sfgridgen")
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([exe, os.path.join(path, "grid.in")], cwd=tmp, check=True,
                       capture_output=True)
        shutil.copy(os.path.join(tmp, "out.sfgridgen"), os.path.join(path, "out.sfgridgen"))
    print("sfgridgen_ch3_1A", os.path.getsize(os.path.join(path, "out.sfgridgen")))


HMC_CASES = {
    # 48 butanes, 3 fs, 20 moves of the equilibration phase with the delt
    # adaptation every 5 moves toward 50 % acceptance (src/hmc.c:47-61)
    "hmc_butane_eq": (48, 5, 20, 20, 20),
    # 40 butanes, 3 fs, 12 production moves (nstep_eq 0: delt fixed)
    "hmc_butane_prod": (40, 6, 0, 12, 12),
}


def hmc_case(name):
    nmol, seed, neq, nstep, _ = HMC_CASES[name]
    rd = configs.butane_box(nmol=nmol, seed=seed, delt_fs=3.0, nstep_inner=5)
    r = rd.run
    r.nstep_eq, r.nstep = neq, nstep
    r.nstep_delt_adj_cycle = 5
    r.rreq_disp = 0.5
    return rd


def make_hmc_cases():
    """The reference's own hmc() on two small n-butane boxes: per move the
    accept flag, dH, delt after the move, uinter, uintra and ukin, and the
    final positions and velocities."""
    for name, (_, _, _, _, nmove) in HMC_CASES.items():
        path = os.path.join(HERE, name)
        os.makedirs(path, exist_ok=True)
        hmc_case(name).write(path)
        d = O.run_reference(path, hmc=nmove)
        h = d["hmc"]
        np.savez_compressed(os.path.join(path, "hmc.npz"),
                            accept=np.array([m["accept"] for m in h], dtype=np.int8),
                            **{k: np.array([m[k] for m in h])
                               for k in ("dH", "delt", "uinter", "uintra", "ukin")},
                            x=np.stack([d["xx"], d["yy"], d["zz"]]), v=d["v"], v0=d["v0"],
                            natom=d["natom"])
        print(name, d["natom"], "".join(str(m["accept"]) for m in h), h[-1]["delt"])


def make_myinterp_case():
    """CH3 sites and CH2 dimers in the (11,11) nanotube cell of
    sfgridgen_ch3_1A (the reference's tools/CH3-SWNT11x11.in), iSF_type 4:
    two cube files at a 2 A interval written by the CPU restatement of the
    generator (oracle/myinterp_gen.py), evaluated by the reference's own
    init_my_interp / get_values_from_grid."""
    from mspms2_amd import myinterp as MI
    from mspms2_amd import sfgridgen as S
    from oracle import myinterp_gen as MG
    g = S.read_input(os.path.join(HERE, "sfgridgen_ch3_1A", "grid.in"))
    g.grid_itvl = 2.0
    files = {}
    for fname, (fs, fe) in (("ch3.cubes", (3.75, 98.0)), ("ch2.cubes", (3.95, 46.0))):
        g.f_sigma, g.f_epsilon = fs, fe
        nodes = MG.lj_nodes(g)
        files[fname] = MI.CubeFile(ucl=(g.uclx, g.ucly, g.uclz),
                                   center=(g.uclx / 2.0, g.ucly / 2.0, 0.0),
                                   ngrid=nodes.shape[2:], h=(2.0, 2.0, 2.0),
                                   coef=MG.fit(nodes, (2.0, 2.0, 2.0)))
    rd = configs.tube_cell_fluid(g.uclx, g.ucly, g.uclz, files=tuple(files))
    rd.cube_files = files
    path = os.path.join(HERE, "tube_cell_myinterp")
    os.makedirs(path, exist_ok=True)
    rd.write(path)
    d = O.run_reference(path)
    save(path, d)
    print("tube_cell_myinterp", d["natom"], d["usflj"], d["uinter"])


def make_file_cases():
    name, n = FILE_CASE
    src = os.path.join(HERE, name)
    with tempfile.TemporaryDirectory() as out:
        O.run_reference(src, md_steps=n, files_out=out)
        for fn in ("sav.mspms", "trj.mspms", "ss.mspms"):
            shutil.copy(os.path.join(out, fn), os.path.join(src, f"ref{n}_{fn}"))
    rname, rn = RESTART
    dst = os.path.join(HERE, rname)
    os.makedirs(dst, exist_ok=True)
    for fn in ("cfg.mspms", "xyz.mspms"):
        shutil.copy(os.path.join(src, fn), os.path.join(dst, fn))
    with open(os.path.join(src, "in.mspms")) as f:
        lines = f.read().split("\n")
    # line 17 of in.mspms: the starting option (src/readins.c:62)
    assert lines[16].split()[1] == "iStart_option"
    lines[16] = f"{2:<24d} iStart_option"
    with open(os.path.join(dst, "in.mspms"), "w") as f:
        f.write("\n".join(lines))
    shutil.copy(os.path.join(src, f"ref{n}_sav.mspms"), os.path.join(dst, "old.mspms"))
    d = O.run_reference(dst, md_steps=rn)
    save_md(dst, d, rn)
    print(rname, rn, d["md"]["ukin"])


def main(only=()):
    """Every fixture, or with names given only those single-evaluation
    cases (CASES), and "hmc" for the HMC chains."""
    if not O.reference_available():
        raise SystemExit("oracle/_ref is not built (needs the reference checkout)")
    if only:
        if "hmc" in only:
            make_hmc_cases()
            only = [n for n in only if n != "hmc"]
        for name in only:
            path = os.path.join(HERE, name)
            os.makedirs(path, exist_ok=True)
            CASES[name]().write(path)
            d = O.run_reference(path)
            save(path, d)
            print(name, d["natom"], d["uinter"], d["usflj"], d["uintra"])
        return
    for name, fn in CASES.items():
        path = os.path.join(HERE, name)
        os.makedirs(path, exist_ok=True)
        fn().write(path)
        d = O.run_reference(path)
        save(path, d)
        print(name, d["natom"], d["uinter"], d["uintra"])
    ew = os.path.join(HERE, "examples_water")
    os.makedirs(ew, exist_ok=True)
    for fnm in ("in.mspms", "cfg.mspms", "xyz.mspms"):
        shutil.copy(os.path.join(REF, "Examples", "Water", fnm), os.path.join(ew, fnm))
    d = O.run_reference(ew)
    save(ew, d)
    print("examples_water", d["natom"], d["uinter"], d["uintra"])
    for name, (fn, steps) in MD_CASES.items():
        path = os.path.join(HERE, name)
        os.makedirs(path, exist_ok=True)
        fn().write(path)
        for n in steps:
            d = O.run_reference(path, md_steps=n)
            save_md(path, d, n)
            print(name, n, d["md"]["ukin"], d["md"]["ps"])
    make_file_cases()
    make_siman_case()
    make_hmc_cases()
    make_sfgridgen_case()
    make_myinterp_case()
    for name, fn in BIG.items():
        path = os.path.join(HERE, name)
        os.makedirs(path, exist_ok=True)
        with tempfile.TemporaryDirectory() as tmp:
            fn().write(tmp)
            d = O.run_reference(tmp)
            with open(os.path.join(path, "runfiles.sha256"), "w") as f:
                f.write(_sha(tmp) + "\n")
        save(path, d, keep_every=12)
        print(name, d["natom"], d["uinter"], d["uintra"])


if __name__ == "__main__":
    main(sys.argv[1:])
