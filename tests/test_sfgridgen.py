"""Grid-generator formats and mathematics on the CPU (mspms2_amd/sfgridgen.py).

tests/golden/sfgridgen_ch3_1A holds the reference generator's own output
(oracle/_ref/sfgridgen, built from tools/sfgridgen.c) for its CH3 / (11,11)
nanotube input at a 1 A interval.  Checked here: the input and output
readers and the writer (byte-identical), and what the mathematics fixes
about the tables -- the energy equals a direct LJ sum over the solid atoms
and their z images, the gradient equals central differences of that sum,
and the nine second-derivative tables, the Jacobian of the force -grad E,
equal minus the central-difference Hessian of the sum and are symmetric."""
import math
import os

import numpy as np
import pytest

from tests.util import GOLDEN

CASE = os.path.join(GOLDEN, "sfgridgen_ch3_1A")
RGAS = 8.314472


def _lj(g, x, y, z):
    """direct sum: 4 eps (s^12 - s^6) over atoms and z images within rc"""
    sig = 0.5 * (g.f_sigma + g.s_sigma)
    eps = math.sqrt(g.f_epsilon * g.s_epsilon) * RGAS
    nuc = 1
    while nuc * g.uclz < g.rcutoff:
        nuc += 1
    u = 0.0
    for p in range(-nuc, nuc + 1):
        d = np.stack([x - g.s_xyz[0], y - g.s_xyz[1], z - g.s_xyz[2] + p * g.uclz])
        r2 = (d * d).sum(axis=0)
        m = r2 < g.rcutoff ** 2
        s6 = (sig * sig / r2[m]) ** 3
        u += float((4.0 * eps * (s6 * s6 - s6)).sum())
    return u


def _load():
    from mspms2_amd import sfgridgen as S
    g = S.read_input(os.path.join(CASE, "grid.in"))
    hdr, tab = S.read_output(os.path.join(CASE, "out.sfgridgen"))
    return S, g, hdr, tab


def test_reader_and_header():
    S, g, hdr, tab = _load()
    assert g.s_natom == 44 and g.grid_itvl == 1.0 and not g.isLJswitch
    assert hdr["n"] == (16, 16, 4) and hdr["ntot"] == 1024 and hdr["h"] == (1.0, 1.0, 1.0)
    assert hdr["ucl"] == (g.uclx, g.ucly, g.uclz)
    assert tab.shape == (13, 16, 16, 4) and np.isfinite(tab).all()


def test_dims_through_the_library():
    """mspms_sfgridgen_dims (host arithmetic only, no device call)"""
    S, g, hdr, tab = _load()
    assert S.dims(g) == hdr["n"]
    g.grid_itvl = 0.1
    assert S.dims(g) == (151, 151, 26)


def test_writer_reproduces_the_file(tmp_path):
    S, g, hdr, tab = _load()
    out = str(tmp_path / "out.sfgridgen")
    S.write_output(out, g, tab)
    assert open(out, "rb").read() == open(os.path.join(CASE, "out.sfgridgen"), "rb").read()


@pytest.mark.parametrize("node", [(3, 4, 1), (8, 8, 2), (12, 2, 3), (6, 13, 0)])
def test_tables_are_the_lj_sum_and_its_derivatives(node):
    S, g, hdr, tab = _load()
    i, j, k = node
    x, y, z = i * g.grid_itvl, j * g.grid_itvl, k * g.grid_itvl
    e = _lj(g, x, y, z)
    assert abs(tab[0, i, j, k] - e) <= 1e-10 * abs(e)
    h = 1e-4
    grad = [(_lj(g, x + h, y, z) - _lj(g, x - h, y, z)) / (2 * h),
            (_lj(g, x, y + h, z) - _lj(g, x, y - h, z)) / (2 * h),
            (_lj(g, x, y, z + h) - _lj(g, x, y, z - h)) / (2 * h)]
    scale = max(abs(v) for v in grad) + 1e-3 * abs(e)
    for d in range(3):
        assert abs(tab[1 + d, i, j, k] - grad[d]) <= 1e-5 * scale, d
    # force Jacobian dF/dr = -Hessian of E: symmetric, and equal to minus the
    # central-difference second derivatives of the direct sum
    J = tab[4:13, i, j, k].reshape(3, 3)
    assert np.allclose(J, J.T, rtol=1e-12, atol=1e-12 * np.abs(J).max())
    p0 = np.array([x, y, z])

    def E(*dd):
        q = p0.copy()
        for ax, sgn in dd:
            q[ax] += sgn * h
        return _lj(g, *q)

    hess = np.zeros((3, 3))
    for a in range(3):
        hess[a, a] = (E((a, 1)) - 2 * e + E((a, -1))) / (h * h)
        for b in range(a + 1, 3):
            hess[a, b] = hess[b, a] = (E((a, 1), (b, 1)) - E((a, 1), (b, -1))
                                       - E((a, -1), (b, 1)) + E((a, -1), (b, -1))) / (4 * h * h)
    assert np.allclose(J, -hess, rtol=0, atol=1e-4 * np.abs(J).max()), (J, -hess)
