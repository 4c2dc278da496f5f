"""The myinterp generator on the GPU (mspms_myinterp_nodes /
mspms_myinterp_fit) against its CPU restatement (oracle/myinterp_gen.py),
and the engine's myinterp evaluation out of range.  The evaluation itself
is pinned against the reference by the parity and drop-in tests (case
tube_cell_myinterp).

Tolerances: node data 1e-12 relative (same sums, different order); fitted
polynomials compared through their values at random points of each cube,
1e-9 relative to the value scale (the absolute-coordinate coefficients
cancel by orders of magnitude there); polynomial-space reproduction 1e-9."""
import os

import numpy as np
import pytest

from mspms2_amd import myinterp as MI
from mspms2_amd import runio
from mspms2_amd import sfgridgen as S
from mspms2_amd.engine import Engine, EngineError
from mspms2_amd.system import System
from oracle import myinterp_gen as MG
from tests.util import GOLDEN

pytestmark = pytest.mark.gpu

GRID = os.path.join(GOLDEN, "sfgridgen_ch3_1A")
CASE = os.path.join(GOLDEN, "tube_cell_myinterp")


def _grid(h):
    g = S.read_input(os.path.join(GRID, "grid.in"))
    g.grid_itvl = h
    return g


def test_node_data_equals_restatement():
    g = _grid(1.0)
    got = MI.node_data(g)
    want = MG.lj_nodes(g)
    for f in range(4):
        for d in range(7):
            a, b = got[f, d], want[f, d]
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), (f, d)


def test_fit_equals_restatement():
    g = _grid(1.0)
    nodes = MG.lj_nodes(g)
    got = MI.fit(nodes, (1.0, 1.0, 1.0))
    want = MG.fit(nodes, (1.0, 1.0, 1.0))
    assert got.shape == want.shape
    rng = np.random.default_rng(2)
    _, nx, ny, nz = nodes.shape[1:]
    ncz, ncy = nz - 1, ny - 1
    for c in rng.integers(0, got.shape[1], 60):
        ix, iy, iz = c // (ncy * ncz), (c // ncz) % ncy, c % ncz
        p = [np.array([v + rng.random()]) for v in (ix, iy, iz)]
        a = MG.evaluate(got[:, c:c + 1], *p)[:, 0]
        b = MG.evaluate(want[:, c:c + 1], *p)[:, 0]
        scale = np.abs(nodes[:, 0]).max(axis=(1, 2, 3))
        assert np.all(np.abs(a - b) <= 1e-9 * scale), (c, a, b)


def test_fit_reproduces_the_polynomial_space():
    rng = np.random.default_rng(5)
    nx, ny, nz = 5, 4, 3
    org, h = (0.3, -0.4, 0.2), (0.7, 0.9, 1.1)
    X, Y, Z = np.meshgrid(org[0] + h[0] * np.arange(nx), org[1] + h[1] * np.arange(ny),
                          org[2] + h[2] * np.arange(nz), indexing="ij")
    co = rng.normal(size=(4, 32))
    nodes = np.zeros((4, 7, nx, ny, nz))
    for f in range(4):
        for d, dv in enumerate(MG.DERIVS):
            nodes[f, d] = sum(co[f, k] * MG._dmono(e, dv, (X, Y, Z))
                              for k, e in enumerate(MG.EXPONENTS))
    out = MI.fit(nodes, h, org)
    assert np.abs(out - np.repeat(co.reshape(128, 1), out.shape[1], axis=1)).max() <= 1e-9


def test_generated_file_drives_the_engine(tmp_path):
    """generate() -> cube files -> run directory -> engine: the engine's
    usflj and forces equal the C oracle's on the same files (FP64)."""
    from oracle.oracle import Oracle
    g = _grid(0.5)
    files = {}
    for fn, (fs, fe) in (("ch3.cubes", (3.75, 98.0)), ("ch2.cubes", (3.95, 46.0))):
        g.f_sigma, g.f_epsilon = fs, fe
        files[fn] = MI.generate(g)
    rd = runio.RunDir.read(CASE)
    rd.cube_files = files
    rd.write(str(tmp_path))
    s = System.from_rundir(runio.RunDir.read(str(tmp_path)))
    fo, to = Oracle(s).frclong()
    eng = Engine(s, precision="fp64")
    f, t = eng.frclong()
    eng.close()
    print(f"usflj engine {t['usflj']:.12e} oracle {to['usflj']:.12e}")
    assert abs(t["usflj"] - to["usflj"]) <= 1e-10 * abs(to["usflj"])
    assert np.linalg.norm(f - fo) <= 1e-10 * np.linalg.norm(fo)


def test_site_outside_the_tables_is_refused():
    s = System.from_rundir(runio.RunDir.read(CASE))
    s.xx = s.xx.copy()
    s.xx[0] = -5.0 / s.params["sigma_base"]   # cube index -3 along x
    eng = Engine(s, precision="fp64")
    with pytest.raises(EngineError, match="solid-fluid grid"):
        eng.frclong()
    eng.close()
