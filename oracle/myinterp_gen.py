"""CPU restatement of the myinterp cube-polynomial generator (numpy).

TEST INFRASTRUCTURE ONLY (the checker of mspms_myinterp_nodes /
mspms_myinterp_fit, and the writer of the golden cube files).

The reference reads and evaluates the 32-coefficient cube polynomials
(src/sffrc.c:228-1089) but ships no generator, so there is no reference
algorithm to follow here.  Its pins are mathematical:
  * node values, gradients and Hessians equal the reference's own grid
    generator tools/sfgridgen.c (same sum, different program);
  * third derivatives equal central differences of the Hessian;
  * a polynomial of the 32-monomial space, given through its node data, is
    reproduced coefficient for coefficient;
  * the 32 x 56 map is built with numpy's pseudo-inverse, independently of
    the engine's Gauss-Jordan construction.
"""
from __future__ import annotations

import math

import numpy as np

RGAS = 8.314472   # tools/sfgridgen.c:27

# exponents of the 32 monomials of Amatrix_ele_assign_value (src/sffrc.c:771-802)
EXPONENTS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1),
             (0, 2, 0), (0, 1, 1), (0, 0, 2), (3, 0, 0), (2, 1, 0), (1, 2, 0), (2, 0, 1),
             (1, 0, 2), (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3), (1, 1, 1), (3, 1, 0),
             (2, 2, 0), (1, 3, 0), (3, 0, 1), (2, 0, 2), (1, 0, 3), (0, 3, 1), (0, 2, 2),
             (0, 1, 3), (2, 1, 1), (1, 2, 1), (1, 1, 2)]
DERIVS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
CORNERS = [(c & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)]


def _dmono(e, d, p):
    """derivative d of monomial with exponents e at point p (scalars or arrays)"""
    r = 1.0
    for a in range(3):
        if d[a] > e[a]:
            return 0.0 * p[0]
        r = r * math.perm(e[a], d[a]) * p[a] ** (e[a] - d[a])
    return r


def lj_nodes(g):
    """(4, 7, nx, ny, nz): E, Fx, Fy, Fz (F = -grad E) at the grid generator's
    nodes with (value, d/dx, d/dy, d/dz, d2/dxdy, d2/dxdz, d2/dydz)."""
    nx = int(math.floor(g.uclx / g.grid_itvl)) + 2
    ny = int(math.floor(g.ucly / g.grid_itvl)) + 2
    nz = int(math.floor(g.uclz / g.grid_itvl)) + 2
    X, Y, Z = np.meshgrid(g.grid_itvl * np.arange(nx), g.grid_itvl * np.arange(ny),
                          g.grid_itvl * np.arange(nz), indexing="ij")
    sig = (g.f_sigma + g.s_sigma) / 2.0
    eps = math.sqrt(g.f_epsilon * g.s_epsilon) * RGAS
    nuc = 1
    while nuc * g.uclz < g.rcutoff:
        nuc += 1
    D = {}
    keys = ["", "x", "y", "z", "xx", "yy", "zz", "xy", "xz", "yz",
            "xxy", "xxz", "xyy", "xyz", "xzz", "yyz", "yzz"]
    for k in keys:
        D[k] = np.zeros_like(X)
    for p in range(-nuc, nuc + 1):
        for a in range(g.s_natom):
            d = {"x": X - g.s_xyz[0, a], "y": Y - g.s_xyz[1, a],
                 "z": Z - g.s_xyz[2, a] + p * g.uclz}
            q = d["x"] ** 2 + d["y"] ** 2 + d["z"] ** 2
            m = q < g.rcutoff ** 2
            qs = np.where(m, q, 1.0)
            s6 = (sig * sig / qs) ** 3
            s12 = s6 * s6
            # U(q) and its q-derivatives, q = r^2
            u = [4 * eps * (s12 - s6), 4 * eps * (-6 * s12 + 3 * s6) / qs,
                 4 * eps * (42 * s12 - 12 * s6) / qs ** 2,
                 4 * eps * (-336 * s12 + 60 * s6) / qs ** 3]
            for k in keys:
                # d^n U / dx_a.. = sum over pairings (Faa di Bruno for U(|d|^2))
                idx = list(k)
                n = len(idx)
                if n == 0:
                    t = u[0]
                elif n == 1:
                    t = 2 * u[1] * d[idx[0]]
                elif n == 2:
                    t = 4 * u[2] * d[idx[0]] * d[idx[1]] + (2 * u[1] if idx[0] == idx[1] else 0)
                else:
                    a_, b_, c_ = idx
                    t = 8 * u[3] * d[a_] * d[b_] * d[c_]
                    t = t + 4 * u[2] * ((d[c_] if a_ == b_ else 0) + (d[b_] if a_ == c_ else 0)
                                        + (d[a_] if b_ == c_ else 0))
                D[k] = D[k] + np.where(m, t, 0.0)
    E = [D[""], D["x"], D["y"], D["z"], D["xy"], D["xz"], D["yz"]]
    Fx = [-D["x"], -D["xx"], -D["xy"], -D["xz"], -D["xxy"], -D["xxz"], -D["xyz"]]
    Fy = [-D["y"], -D["xy"], -D["yy"], -D["yz"], -D["xyy"], -D["xyz"], -D["yyz"]]
    Fz = [-D["z"], -D["xz"], -D["yz"], -D["zz"], -D["xyz"], -D["xzz"], -D["yzz"]]
    return np.array([E, Fx, Fy, Fz])


def fit_map():
    """(32, 56): cube-local data (corner c, derivative d) -> local coefficients"""
    M = np.array([[_dmono(e, d, c) for e in EXPONENTS] for c in CORNERS for d in DERIVS[:4]])
    A = np.linalg.pinv(M)                       # minimum-norm least squares, (32, 32)
    G = np.zeros((32, 56))
    for c in range(8):
        for d in range(4):
            G[:, c * 7 + d] = A[:, c * 4 + d]
    # the null space of M: u v (1-u)(1-v) and its xz, yz analogues
    pairs = [(0, 1), (0, 2), (1, 2)]
    for n, (a, b) in enumerate(pairs):
        N = np.zeros(32)
        for (pa, pb, w) in [(1, 1, 1.0), (2, 1, -1.0), (1, 2, -1.0), (2, 2, 1.0)]:
            e = [0, 0, 0]
            e[a], e[b] = pa, pb
            N[EXPONENTS.index(tuple(e))] += w
        dmix = DERIVS[4 + n]
        row = np.zeros(56)
        for c, cp in enumerate(CORNERS):
            sgn = (1 - 2 * cp[a]) * (1 - 2 * cp[b])
            row[c * 7 + 4 + n] += sgn / 8.0
            dv = np.array([_dmono(e, dmix, cp) for e in EXPONENTS])
            row -= sgn / 8.0 * (dv @ G)
        G = G + np.outer(N, row)
    return G


def fit(nodes, h, origin=(0.0, 0.0, 0.0)):
    """(128, ncube) absolute-coordinate coefficients (file order) from node
    data (4, 7, nx, ny, nz)."""
    _, _, nx, ny, nz = nodes.shape
    hx, hy, hz = h
    G = fit_map()
    ix, iy, iz = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1),
                             indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    scale = [1.0, hx, hy, hz, hx * hy, hx * hz, hy * hz]
    xa = origin[0] + hx * ix
    ya = origin[1] + hy * iy
    za = origin[2] + hz * iz
    out = np.zeros((128, ix.size))
    for f in range(4):
        data = np.zeros((ix.size, 56))
        for c, (cx, cy, cz) in enumerate(CORNERS):
            for d in range(7):
                data[:, c * 7 + d] = nodes[f, d, ix + cx, iy + cy, iz + cz] * scale[d]
        loc = data @ G.T                       # (ncube, 32) in u, v, w
        for k, (a, b, c) in enumerate(EXPONENTS):
            w = loc[:, k] / (hx ** a * hy ** b * hz ** c)
            for i in range(a + 1):
                for j in range(b + 1):
                    for l in range(c + 1):
                        t = (math.comb(a, i) * math.comb(b, j) * math.comb(c, l)
                             * (-xa) ** (a - i) * (-ya) ** (b - j) * (-za) ** (c - l))
                        out[f * 32 + EXPONENTS.index((i, j, l)), :] += w * t
    return out


def evaluate(coef, x, y, z):
    """the four polynomials (128 rows, one cube's column each) at points"""
    v = np.array([x ** a * y ** b * z ** c for (a, b, c) in EXPONENTS])
    return np.array([(coef[32 * f:32 * f + 32] * v).sum(axis=0) for f in range(4)])
