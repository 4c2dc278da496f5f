"""The myinterp solid-fluid model on the CPU: cube files, the run-input
section, and the CPU restatement of the generator (oracle/myinterp_gen.py).

The reference evaluates the cube polynomials (src/sffrc.c:951-1089) but
ships no generator.  The evaluator is pinned bit for bit against the
reference by tests/test_oracle_golden.py (case tube_cell_myinterp) and,
on the GPU, by the parity and drop-in tests.  The generator is pinned here
by mathematics: node data against the reference's own grid generator and
finite differences, exact reproduction of the 32-monomial space, the null
space of the corner conditions, and fourth-order convergence to the direct
LJ sum."""
import math
import os

import numpy as np
import pytest

from mspms2_amd import myinterp as MI
from mspms2_amd import runio
from mspms2_amd import sfgridgen as S
from mspms2_amd.system import System
from oracle import myinterp_gen as MG
from oracle.oracle import Oracle
from tests.util import GOLDEN

CASE = os.path.join(GOLDEN, "tube_cell_myinterp")
GRID = os.path.join(GOLDEN, "sfgridgen_ch3_1A")


def _space_poly(rng, shape_origin_h):
    """node data (4, 7, nx, ny, nz) of four random polynomials of the space"""
    (nx, ny, nz), org, h = shape_origin_h
    X, Y, Z = np.meshgrid(org[0] + h[0] * np.arange(nx), org[1] + h[1] * np.arange(ny),
                          org[2] + h[2] * np.arange(nz), indexing="ij")
    co = rng.normal(size=(4, 32))
    nodes = np.zeros((4, 7, nx, ny, nz))
    for f in range(4):
        for d, dv in enumerate(MG.DERIVS):
            nodes[f, d] = sum(co[f, k] * MG._dmono(e, dv, (X, Y, Z))
                              for k, e in enumerate(MG.EXPONENTS))
    return co, nodes


def test_cube_file_round_trip(tmp_path):
    for fn in ("ch3.cubes", "ch2.cubes"):
        src = os.path.join(CASE, fn)
        cf = MI.read_cubes(src)
        assert cf.ngrid == (9, 9, 3) and cf.h == (2.0, 2.0, 2.0)
        assert cf.coef.shape == (128, 8 * 8 * 2)
        out = str(tmp_path / fn)
        MI.write_cubes(out, cf)
        assert open(out, "rb").read() == open(src, "rb").read()


def test_run_input_section_round_trip(tmp_path):
    rd = runio.RunDir.read(CASE)
    r = rd.run
    assert r.iSF_type == 4 and r.mi_files == ["ch3.cubes", "ch2.cubes"]
    assert r.mi_ucl == (14.916, 14.916, 2.4595)
    rd.write(str(tmp_path))
    for fn in ("in.mspms", "ch3.cubes", "ch2.cubes"):
        assert open(tmp_path / fn, "rb").read() == open(os.path.join(CASE, fn), "rb").read()


def test_engine_tables_follow_init_my_interp():
    """xmin = xcenter - uclx/2, ymin likewise, zmax = uclz; the lattice of
    the files; a file with another unit cell is refused."""
    rd = runio.RunDir.read(CASE)
    s = System.from_rundir(rd)
    q = s.cubes
    assert (q["ntype"], q["ncx"], q["ncy"], q["ncz"]) == (2, 8, 8, 2)
    assert q["xmin"] == 7.458 - 14.916 / 2.0 and q["zmax"] == 2.4595
    bad = MI.CubeFile(ucl=(14.9, 14.916, 2.4595), center=(0, 0, 0), ngrid=(9, 9, 3),
                      h=(2.0, 2.0, 2.0), coef=np.zeros((128, 128)))
    with pytest.raises(ValueError, match="unit cell"):
        MI.engine_tables(rd.run.mi_ucl, rd.run.mi_center, [bad])


def test_restated_monomial_order_is_the_reference_evaluators():
    """The usflj of the reference-exact evaluator (the C oracle, bit-exact
    to the reference on this case) equals the numpy restatement's sum over
    sites, which uses oracle/myinterp_gen.EXPONENTS: the monomial order of
    the generator is the one the reference evaluates."""
    s = System.from_rundir(runio.RunDir.read(CASE))
    _, t = Oracle(s).frclong()
    q = s.cubes
    total = 0.0
    site_types = []
    for sp in s.species:
        site_types += list(sp.interp_type) * sp.nmole
    for k in range(s.natom):
        x, y, z = s.xx[k], s.yy[k], s.zz[k]
        z = z - q["zmax"] * np.rint(z / q["zmax"])
        z = z + q["zmax"] if z < 0 else z
        ix, iy, iz = int((x - q["xmin"]) / q["hx"]), int((y - q["ymin"]) / q["hy"]), int(z / q["hz"])
        c = (ix * q["ncy"] + iy) * q["ncz"] + iz
        coef = q["coef"][site_types[k] - 1][:, c:c + 1]
        total += MG.evaluate(coef, np.array([x]), np.array([y]), np.array([z]))[0, 0]
    assert abs(total * 0.120272219 - t["usflj"]) <= 1e-9 * abs(t["usflj"])


def test_null_space_of_the_corner_conditions():
    """u v (1-u)(1-v) and its analogues vanish with their gradient at every
    corner: value and gradient data cannot fix them (rank 29)."""
    M = np.array([[MG._dmono(e, d, c) for e in MG.EXPONENTS] for c in MG.CORNERS
                  for d in MG.DERIVS[:4]])
    assert np.linalg.matrix_rank(M) == 29
    for a, b in ((0, 1), (0, 2), (1, 2)):
        N = np.zeros(32)
        for pa, pb, w in ((1, 1, 1.0), (2, 1, -1.0), (1, 2, -1.0), (2, 2, 1.0)):
            e = [0, 0, 0]
            e[a], e[b] = pa, pb
            N[MG.EXPONENTS.index(tuple(e))] += w
        assert np.abs(M @ N).max() == 0.0


def test_fit_reproduces_the_polynomial_space():
    rng = np.random.default_rng(5)
    spec = ((5, 4, 3), (0.3, -0.4, 0.2), (0.7, 0.9, 1.1))
    co, nodes = _space_poly(rng, spec)
    out = MG.fit(nodes, spec[2], spec[1])
    want = np.repeat(co.reshape(128, 1), out.shape[1], axis=1)
    assert np.abs(out - want).max() <= 1e-9


def test_node_data_equals_the_reference_generator():
    """E, grad E and the force Jacobian (= -Hessian) at the nodes equal the
    reference generator's tables (tools/sfgridgen.c output, 1 A)."""
    g = S.read_input(os.path.join(GRID, "grid.in"))
    _, tab = S.read_output(os.path.join(GRID, "out.sfgridgen"))
    nd = MG.lj_nodes(g)

    def close(a, b):
        return np.abs(a - b).max() <= 1e-13 * np.abs(b).max()

    assert close(nd[0, 0], tab[0])
    for d in range(3):
        assert close(nd[0, 1 + d], tab[1 + d])          # dE/dx_d
        assert close(nd[1 + d, 0], -tab[1 + d])         # F_d
        for e in range(3):
            assert close(nd[1 + d, 1 + e], tab[4 + 3 * d + e])  # dF_d/dx_e
    # mixed second derivatives of E: the Jacobian's off-diagonal, negated
    assert close(nd[0, 4], -tab[5]) and close(nd[0, 5], -tab[6]) and close(nd[0, 6], -tab[9])


def test_third_derivatives_are_differences_of_the_hessian():
    """Moving every solid atom by -delta along z moves the nodes by +delta:
    central differences of first and second derivatives give the mixed
    second derivatives (xz, yz) of E and F."""
    g = S.read_input(os.path.join(GRID, "grid.in"))
    g.grid_itvl = 2.0
    base = MG.lj_nodes(g)
    dl = 1e-4
    shifted = []
    for sgn in (1, -1):
        h = S.read_input(os.path.join(GRID, "grid.in"))
        h.grid_itvl = 2.0
        h.s_xyz = h.s_xyz.copy()
        h.s_xyz[2] -= sgn * dl
        shifted.append(MG.lj_nodes(h))
    fd = (shifted[0] - shifted[1]) / (2 * dl)          # d/dz of every field
    # d/dz of (dFx/dx) = d2Fx/dxdz ; d/dz of (dFy/dy) = d2Fy/dydz
    for f, src_d, mix_d in ((1, 1, 5), (2, 2, 6), (0, 1, 5)):
        a, b = base[f, mix_d], fd[f, src_d]
        scale = np.abs(a).max()
        assert np.abs(a - b).max() <= 1e-5 * scale, (f, np.abs(a - b).max(), scale)


def test_fourth_order_convergence_to_the_direct_sum():
    """Energy of the fitted polynomial against the direct LJ sum at random
    points near the tube axis: the error falls by >= 12x per
    halving of the interval (fourth order gives 16x)."""
    g = S.read_input(os.path.join(GRID, "grid.in"))
    rng = np.random.default_rng(3)
    pts = []
    while len(pts) < 40:
        r, th = 3.0 * math.sqrt(rng.random()), 2 * math.pi * rng.random()
        pts.append((g.uclx / 2 + r * math.cos(th), g.ucly / 2 + r * math.sin(th),
                    g.uclz * rng.random()))
    sig = (g.f_sigma + g.s_sigma) / 2
    eps = math.sqrt(g.f_epsilon * g.s_epsilon) * MG.RGAS

    def direct(p):
        e = 0.0
        for pp in range(-12, 13):
            d = np.stack([p[0] - g.s_xyz[0], p[1] - g.s_xyz[1], p[2] - g.s_xyz[2] + pp * g.uclz])
            q = (d * d).sum(0)
            s6 = (sig * sig / q[q < g.rcutoff ** 2]) ** 3
            e += (4 * eps * (s6 * s6 - s6)).sum()
        return e

    ref = np.array([direct(p) for p in pts])
    errs = []
    for h in (0.5, 0.25):
        g.grid_itvl = h
        nodes = MG.lj_nodes(g)
        coef = MG.fit(nodes, (h, h, h))
        _, ny, nz = nodes.shape[2:]
        err = 0.0
        for p, e in zip(pts, ref):
            c = (int(p[0] / h) * (ny - 1) + int(p[1] / h)) * (nz - 1) + int(p[2] / h)
            v = MG.evaluate(coef[:, c:c + 1], *(np.array([t]) for t in p))[0, 0]
            err = max(err, abs(v - e))
        errs.append(err)
    assert errs[1] * 12 <= errs[0], errs
    assert errs[1] <= 1e-4 * np.abs(ref).max(), errs
