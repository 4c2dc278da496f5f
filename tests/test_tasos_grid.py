"""The Fortran grid library's input side (iSF_type 3): grid geometry, grid
files, the TASOS section, the z period and the PE sign flip.

- InitPotentialGrid without symmetry (mylibtasos/initialize.f90:96-101):
  ng = ANINT(a / 0.2), rounding halves away from zero (Python's round()
  would take 2.5 to 2).
- Read_Grids (mylibtasos/grid.f90:54-129): one list-directed record of 8
  values per node, x fastest; Fortran D exponents and comma separators read.
- init_tasos_grid's section (src/sffrc.c:173-200) reads back as written.
- The lookup's z period: PBCType 2 wraps z by the cell's c = ngz h
  (pbc.f90:82-86), one cell less than the ngz + 2 nodes span.  A grid of
  ngz + 2 layers with zperiod = c gives what the same grid cut to ngz + 1
  layers gives with the default period (nz - 1) h, everywhere.
- The PE sign flip (forces.f90:166-171): below 1.5 times the minimum node F
  the energy is -F.  Fields that are positive everywhere reach it.
- The reference's own frclong with the engine's call_tasos_forces, which
  reads the TASOS section's files itself (oracle/_ref/mspms_terms_harness
  --tasos-linked), equals the oracle on the grid the Python reader loads.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from mspms2_amd import abi, configs, runio, sfgrid
from mspms2_amd.system import System
from oracle import oracle as O

HARNESS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "_ref", "mspms_terms_harness")


def test_grid_dims_round_half_away_from_zero():
    (ngx, ngy, ngz), (hx, hy, hz) = sfgrid.tasos_dims(0.5, 4.0, 1.1)
    assert (ngx, ngy, ngz) == (3, 20, 6)        # 2.5 -> 3, 20, 5.5 -> 6
    assert hx == 0.5 / 3 and hy == 4.0 / 20 and hz == 1.1 / 6


def test_grid_file_round_trip_and_fortran_notation():
    rng = np.random.default_rng(3)
    auc, buc, cuc = 0.6, 0.4, 0.6
    (ngx, ngy, ngz), _ = sfgrid.tasos_dims(auc, buc, cuc)
    tab = rng.standard_normal((8, ngz + 2, ngy + 2, ngx + 2)) * 1e3
    with tempfile.TemporaryDirectory() as tmp:
        fn = os.path.join(tmp, "g.grid")
        sfgrid.write_tasos_grid(fn, tab, auc, buc, cuc)
        back = sfgrid.read_tasos_table(fn, auc, buc, cuc)
        assert np.array_equal(back, tab)
        # the same values as a Fortran program writes them: D exponents,
        # commas, records split over lines
        fn2 = os.path.join(tmp, "f.grid")
        with open(fn2, "w") as f:
            for rec in tab.reshape(8, -1).T:
                v = [f"{x:.16E}".replace("E", "D") for x in rec]
                f.write(", ".join(v[:5]) + "\n  " + " ".join(v[5:]) + "\n")
        assert np.array_equal(sfgrid.read_tasos_table(fn2, auc, buc, cuc), tab)
        short = os.path.join(tmp, "short.grid")
        np.savetxt(short, tab.reshape(8, -1).T[:-1])
        with pytest.raises(ValueError):
            sfgrid.read_tasos_table(short, auc, buc, cuc)


def test_tasos_section_round_trip():
    rd = configs.tasos_column()
    with tempfile.TemporaryDirectory() as tmp:
        rd.write(tmp)
        r = runio.read_in(os.path.join(tmp, "in.mspms"))
        assert r.iSF_type == abi.SF_TASOS
        assert r.tasos_uc == rd.run.tasos_uc and r.tasos_radius == rd.run.tasos_radius
        assert r.tasos_ncell == rd.run.tasos_ncell and r.tasos_ngrid == 1
        assert r.tasos_files == ["ch4.grid"]
        rd2 = runio.RunDir.read(tmp)
        assert np.array_equal(rd2.grid_files["ch4.grid"], rd.grid_files["ch4.grid"])
        s = System.from_rundir(rd2)
    g = s.grid
    assert (g["nx"], g["ny"], g["nz"]) == (22, 22, 8)
    assert g["zperiod"] == 1.2 and g["x0"] == g["y0"] == g["z0"] == 0.0


def _c_grid(g):
    nodes = np.ascontiguousarray(g["nodes"], dtype=np.float32)
    cg = abi.SFGrid(ntype=g["ntype"], nx=g["nx"], ny=g["ny"], nz=g["nz"], x0=g["x0"],
                    y0=g["y0"], z0=g["z0"], hx=g["hx"], hy=g["hy"], hz=g["hz"],
                    nodes=abi.fptr(nodes), zperiod=g.get("zperiod", 0.0))
    return cg, nodes


def _lookup(cg, minpe, x, y, z):
    pe = C.c_double()
    f3 = (C.c_double * 3)()
    rc = O.lib().oracle_hermite_site(C.byref(cg), 0, minpe, x, y, z, C.byref(pe), f3)
    return rc, pe.value, list(f3)


def test_z_period_of_the_fortran_layout():
    rd = configs.tasos_column()
    g = sfgrid.tasos_grid([rd.grid_files["ch4.grid"]], *rd.run.tasos_uc)
    cut = dict(g, nz=g["nz"] - 1, zperiod=0.0, nodes=g["nodes"][:, :, :-1])
    cg, k1 = _c_grid(g)
    cc, k2 = _c_grid(cut)
    minpe = O.lib().oracle_grid_min_pe(C.byref(cg), 0)
    rng = np.random.default_rng(7)
    for _ in range(200):
        x, y = rng.uniform(0.0, 4.19, 2)
        z = rng.uniform(-30.0, 30.0)
        a = _lookup(cg, minpe, x, y, z)
        b = _lookup(cc, minpe, x, y, z)
        assert a[0] == 0 and a == b, (x, y, z, a, b)
    # the period is c, not (nz - 1) h: one period up is the same point (to
    # the rounding of the wrap)
    a = _lookup(cg, minpe, 2.0, 2.0, 0.37)
    b = _lookup(cg, minpe, 2.0, 2.0, 0.37 + 1.2)
    assert abs(a[1] - b[1]) <= 1e-9 * abs(a[1])


def test_pe_sign_flip_below_one_and_a_half_minimum():
    """forces.f90:166-171 on a field positive everywhere: at the node of
    the minimum, F = minPE < 1.5 minPE, so PE = -F; where F >= 1.5 minPE,
    PE = F.  Forces are -grad F either way."""
    rd = configs.tasos_column()
    tab = rd.grid_files["ch4.grid"].copy()
    tab[0] += -tab[0].min() + 50.0          # F >= 50 K everywhere
    g = sfgrid.tasos_grid([tab], *rd.run.tasos_uc)
    cg, keep = _c_grid(g)
    minpe = O.lib().oracle_grid_min_pe(C.byref(cg), 0)
    assert minpe == pytest.approx(50.0, abs=1e-4)
    # nodes a lookup can sit on: the last x, y and z layers start no cell
    inner = g["nodes"][0, 0, :-2, :-1, :-1]
    k = np.unravel_index(np.argmin(inner), inner.shape)
    fmin = float(inner[k])
    assert fmin < 1.5 * minpe
    z, y, x = (k[0] * g["hz"], k[1] * g["hy"], k[2] * g["hx"])
    rc, pe, f = _lookup(cg, minpe, x, y, z)
    assert rc == 0 and pe == pytest.approx(-fmin, rel=1e-12)
    kmax = np.unravel_index(np.argmax(inner), inner.shape)
    z, y, x = (kmax[0] * g["hz"], kmax[1] * g["hy"], kmax[2] * g["hx"])
    rc, pe, f = _lookup(cg, minpe, x, y, z)
    assert rc == 0 and pe == pytest.approx(float(inner[kmax]), rel=1e-12) and pe > 1.5 * minpe


@pytest.mark.skipif(not os.path.exists(HARNESS),
                    reason="term harness not built (no reference sources)")
def test_reference_frclong_with_linked_grid_reader_equals_oracle():
    rd = configs.tasos_column()
    with tempfile.TemporaryDirectory() as tmp:
        rd.write(tmp)
        s = System.from_rundir(runio.RunDir.read(tmp))
        f, t = O.Oracle(s).frclong()
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([HARNESS, tmp, out, "--tasos-linked"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        d = O.read_ref_dump(out)
    for k in ("usflj", "uvdw", "uinter", "virial_inter"):
        assert d[k] == t[k], (k, d[k], t[k])
    assert t["usflj"] < 0.0
    assert np.array_equal(d["fl"], f)
