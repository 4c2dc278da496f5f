"""Restatements for the methane-in-tube workload at sizes beyond the
oracle's all-pairs loop (test infrastructure).

lj_pairs: the LJ pair part of frclong (src/frclong.c:105-147,
src/ljfrc.c:21-71) for one single-site species in numpy: candidate pairs from
a periodic k-d tree, then the reference's arithmetic on them (minimum image
d - L rint(d / L), r^2 summed x, y, z, the test r^2 < rc^2), so the (i < j)
pair set and its splitmix64 hash are the reference's.
wall: the oracle's tricubic Hermite lookup (mylibtasos/forces.f90), site by
site.  Both are pinned to the oracle in tests/test_tube_ref.py.
"""
import ctypes as C

import numpy as np

from mspms2_amd import abi
from oracle import oracle as O


def splitmix64(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def lj_pairs(s, x=None):
    """The reference's LJ pair part of frclong (src/frclong.c:105-147,
    src/ljfrc.c:21-71) for one single-site species: (count, hash, uvdw,
    virial, forces)."""
    from scipy.spatial import cKDTree
    p = s.params
    X = np.stack([s.xx, s.yy, s.zz]) if x is None else x
    L = np.array([p["boxlx"], p["boxly"], p["boxlz"]])
    rc = p["rcutoff"]
    w = np.mod(X.T, L)
    w = np.where(w >= L, 0.0, w)
    i, j = cKDTree(w, boxsize=L).query_pairs(rc * (1 + 1e-9) + 1e-9, output_type="ndarray").T
    swap = i > j
    i, j = np.where(swap, j, i), np.where(swap, i, j)
    d = X[:, i] - X[:, j]
    d = d - L[:, None] * np.rint(d / L[:, None])
    rsq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    m = rsq < rc * rc
    i, j, d, rsq = i[m], j[m], d[:, m], rsq[m]
    sp = s.species[0]
    sig, eps = sp.sigma[0], sp.epsilon[0]
    sij, eij = 0.5 * (sig + sig), np.sqrt(eps * eps)
    r2 = sij * sij / rsq
    r6 = r2 * r2 * r2
    r12 = r6 * r6
    u = 4.0 * eij * (r12 - r6)
    f = 24.0 * eij * (r12 + r12 - r6) / rsq
    F = np.zeros_like(X)
    for k in range(3):
        np.add.at(F[k], i, f * d[k])
        np.add.at(F[k], j, -f * d[k])
    key = (i.astype(np.uint64) << np.uint64(32)) | j.astype(np.uint64)
    h = int(np.sum(splitmix64(key), dtype=np.uint64))
    return len(i), h, float(u.sum()), float((f * rsq).sum()), F


def _c_grid(g):
    G = abi.SFGrid()
    for k in ("ntype", "nx", "ny", "nz", "x0", "y0", "z0", "hx", "hy", "hz"):
        setattr(G, k, g[k])
    nodes = np.ascontiguousarray(g["nodes"], dtype=np.float32)
    G.nodes = abi.fptr(nodes)
    return G, nodes


def wall(s, x=None):
    """The oracle's tricubic Hermite lookup (mylibtasos/forces.f90) for every
    site: (usflj, forces) in reduced units."""
    X = np.stack([s.xx, s.yy, s.zz]) if x is None else x
    G, keep = _c_grid(s.grid)
    L = O.lib()
    minpe = L.oracle_grid_min_pe(C.byref(G), 0)
    sb, eb = s.params["sigma_base"], s.params["epsilon_base"]
    pe = C.c_double()
    f3 = (C.c_double * 3)()
    F = np.zeros_like(X)
    u = 0.0
    for a in range(X.shape[1]):
        rc = L.oracle_hermite_site(C.byref(G), 0, minpe, X[0, a] * sb, X[1, a] * sb,
                                   X[2, a] * sb, C.byref(pe), f3)
        assert rc == 0, a
        u += pe.value / eb
        F[:, a] = np.array(f3[:]) * (sb / eb)
    return u, F
