"""Solid-fluid potential grids for the tricubic Hermite lookup (SF_TASOS).

`tube_grid` tabulates the smooth-cylinder (hypergeometric) LJ potential of
cal_sf_hypergeo (src/sffrc.c:1092-1209) on a node lattice, with the eight
fields the Hermite interpolant needs per node (mylibtasos/grid.f90:81-99
reads the same eight per node): F, dF/dx, dF/dy, dF/dz, d2F/dxdy,
d2F/dxdz, d2F/dydz, d3F/dxdydz, in K and Angstrom.  The tube is
z-invariant, so every z derivative is zero.  F and dF/dr come from the
closed form (energy and its force expression); d2F/dr2 is a central
difference of the closed-form dF/dr.

Layout: nodes[type][field][iz][iy][ix] float32 (mspms_sf_grid).
"""
from __future__ import annotations

import numpy as np

C3_PI_SQ_THETA = 11.3105666436484  # src/defs.h:99
CLAMP = 1.0e30                      # keeps nodes on the wall finite in FP32


def _hypergeo(a, b, c, z, rr):
    """the 11-term series of src/sffrc.c:1092-1116 (vectorized)"""
    zn = np.ones_like(z)
    an1 = np.ones_like(z)
    sv = np.ones_like(z)
    sf = np.zeros_like(z)
    for i in range(11):
        an = (a + i) * (b + i) / ((c + i) * (1 + i))
        an1 = an1 * an
        zn = zn * z
        sv = sv + an1 * zn
        sf = sf - 2.0 * (i + 1.0) * an1 * zn / rr
    return sv, sf


def tube_u_du(r, R, sigma_f, eps_f, sigma_s, eps_s, c3pisq=C3_PI_SQ_THETA):
    """Energy u(r) and du/dr of a fluid site at distance r from the axis of a
    smooth tube of radius R (cal_sf_hypergeo, Lorentz-Berthelot mixing)."""
    r = np.asarray(r, dtype=np.float64)
    sij = 0.5 * (sigma_f + sigma_s)
    eij = np.sqrt(eps_f * eps_s)
    rsq = r * r
    Rsq = R * R
    u = np.empty_like(r)
    f = np.empty_like(r)
    ins = r < R
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ri, rsi = r[ins], rsq[ins]
        ljc = c3pisq * eij * sij * sij
        k = sij * R / (Rsq - rsi)
        z = rsi / Rsq
        hr, hfr = _hypergeo(-4.5, -4.5, 1.0, z, ri)
        k4 = k ** 4
        r1 = 0.65625 * k ** 10
        r2 = -hfr + hr * (20.0 * ri / (Rsq - rsi))
        ha, hfa = _hypergeo(-1.5, -1.5, 1.0, z, ri)
        a2 = -hfa + ha * (8.0 * ri / (Rsq - rsi))
        u[ins] = ljc * (r1 * hr - k4 * ha)
        f[ins] = -ljc * (r1 * r2 - k4 * a2)
        out = ~ins
        ro, rso = r[out], rsq[out]
        ljc = c3pisq * eij * sij * sij * R
        k = sij * ro / (rso - Rsq)
        z = Rsq / rso
        hr, hfr = _hypergeo(-4.5, -4.5, 1.0, z, ro)
        k4 = k ** 4
        r1 = 0.65625 * k ** 10
        r2 = hfr / ro + hr * (-20.0 / (rso - Rsq) + 9.0 / rso)
        ha, hfa = _hypergeo(-1.5, -1.5, 1.0, z, ro)
        a2 = hfa / ro + ha * (-8.0 / (rso - Rsq) + 3.0 / rso)
        u[out] = ljc * (r1 * hr - k4 * ha) / ro
        f[out] = -ljc * (r1 * r2 - k4 * a2)
    # the reference's radial force is -du/dr
    return u, -f


def tube_grid(site_params, R, x0, y0, nx, ny, hx, hy, nz, hz, sigma_s, eps_s,
              center=(0.0, 0.0), c3pisq=C3_PI_SQ_THETA, dr=1e-4):
    """Node table for each (sigma_f, eps_f) in site_params (one grid type
    each, interp_type = index + 1).  Returns the mspms_sf_grid fields."""
    xs = x0 + hx * np.arange(nx)
    ys = y0 + hy * np.arange(ny)
    X, Y = np.meshgrid(xs - center[0], ys - center[1], indexing="xy")   # [ny][nx]
    r = np.sqrt(X * X + Y * Y)
    r = np.where(r == 0.0, 1e-12, r)
    nodes = np.zeros((len(site_params), 8, nz, ny, nx), dtype=np.float32)
    for t, (sf, ef) in enumerate(site_params):
        u, du = tube_u_du(r, R, sf, ef, sigma_s, eps_s, c3pisq)
        _, dup = tube_u_du(r + dr, R, sf, ef, sigma_s, eps_s, c3pisq)
        _, dum = tube_u_du(r - dr, R, sf, ef, sigma_s, eps_s, c3pisq)
        d2u = (dup - dum) / (2 * dr)
        fx = du * X / r
        fy = du * Y / r
        fxy = d2u * X * Y / (r * r) - du * X * Y / (r ** 3)
        planes = [u, fx, fy, np.zeros_like(u), fxy, np.zeros_like(u), np.zeros_like(u),
                  np.zeros_like(u)]
        for k, pl in enumerate(planes):
            pl = np.nan_to_num(pl, nan=CLAMP, posinf=CLAMP, neginf=-CLAMP)
            nodes[t, k, :, :, :] = np.clip(pl, -CLAMP, CLAMP)[None, :, :]
    return dict(ntype=len(site_params), nx=nx, ny=ny, nz=nz, x0=x0, y0=y0, z0=0.0, hx=hx,
                hy=hy, hz=hz, nodes=nodes)


# ---- the Fortran library's grid files ------------------------------------
# init_tasos_grid (src/sffrc.c:149-221) reads the TASOS section of in.mspms:
# the unit cell a, b, c and the tube radius (Angstrom), the cell counts na,
# nb, nc, the number of grids, the number of unique atoms and one grid file
# name per unique atom.  InitPotentialGrid (mylibtasos/initialize.f90:41-155,
# USESYMMETRY off, its compiled default: global.f90:235) sizes the grid
# from the unit cell and grid_spacing = 0.2 A (global.f90:214):
# ng = ANINT(a / 0.2), h = a / ng, ng + 2 nodes per axis from the origin;
# Read_Grids (grid.f90:54-129) reads, per grid, one list-directed record of
# 8 values (F, dF/dx, dF/dy, dF/dz, d2F/dxdy, d2F/dxdz, d2F/dydz,
# d3F/dxdydz) per node, x fastest, then y, then z.  Positions wrap in z by c
# only (PBCType 2, pbc.f90:82-86).

TASOS_GRID_SPACING = 0.2


def _anint(v):
    """Fortran ANINT: nearest integer, halves away from zero."""
    return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def tasos_dims(auc, buc, cuc, spacing=TASOS_GRID_SPACING):
    """(ngx, ngy, ngz), (hx, hy, hz) of InitPotentialGrid without symmetry
    (mylibtasos/initialize.f90:96-101)."""
    ng = [_anint(a / spacing) for a in (auc, buc, cuc)]
    h = [a / n for a, n in zip((auc, buc, cuc), ng)]
    return tuple(ng), tuple(h)


def write_tasos_grid(path, nodes, auc, buc, cuc):
    """One grid file as Read_Grids reads it: nodes is [8][nz][ny][nx] with
    (nx, ny, nz) = (ngx + 2, ngy + 2, ngz + 2) of the unit cell; one record
    of 8 values per node, x fastest.  Values are written with 17
    significant digits, so FP64 grids read back exactly."""
    (ngx, ngy, ngz), _ = tasos_dims(auc, buc, cuc)
    nodes = np.asarray(nodes, dtype=np.float64)
    if nodes.shape != (8, ngz + 2, ngy + 2, ngx + 2):
        raise ValueError(f"grid of shape {nodes.shape}; the unit cell needs "
                         f"(8, {ngz + 2}, {ngy + 2}, {ngx + 2})")
    flat = nodes.reshape(8, -1).T
    np.savetxt(path, flat, fmt="%.17g")


def _read_records(path, count):
    with open(path) as f:
        text = f.read()
    # list-directed input: blanks, commas or newlines separate values;
    # Fortran writes D exponents
    vals = np.array(text.replace(",", " ").replace("D", "E").replace("d", "e").split(),
                    dtype=np.float64)
    if vals.size < 8 * count:
        raise ValueError(f"{path}: {vals.size // 8} node records, {count} needed")
    return vals[:8 * count].reshape(count, 8)


def read_tasos_table(path, auc, buc, cuc):
    """One grid file (Read_Grids' records for one grid) as a
    [8][ngz+2][ngy+2][ngx+2] FP64 array."""
    (ngx, ngy, ngz), _ = tasos_dims(auc, buc, cuc)
    nx, ny, nz = ngx + 2, ngy + 2, ngz + 2
    return np.ascontiguousarray(_read_records(path, nx * ny * nz).T.reshape(8, nz, ny, nx))


def tasos_grid(tables, auc, buc, cuc):
    """The mspms_sf_grid fields of the Fortran library's grids: one table
    type per grid in the TASOS section's order (the site interp_type selects
    one, 1-based), FP32 node values, nodes from the origin, z period c."""
    (ngx, ngy, ngz), (hx, hy, hz) = tasos_dims(auc, buc, cuc)
    nx, ny, nz = ngx + 2, ngy + 2, ngz + 2
    nodes = np.zeros((len(tables), 8, nz, ny, nx), dtype=np.float32)
    for t, tab in enumerate(tables):
        if np.shape(tab) != (8, nz, ny, nx):
            raise ValueError(f"grid {t} of shape {np.shape(tab)}, the unit cell needs "
                             f"(8, {nz}, {ny}, {nx})")
        nodes[t] = np.asarray(tab, dtype=np.float32)
    return dict(ntype=len(tables), nx=nx, ny=ny, nz=nz, x0=0.0, y0=0.0, z0=0.0, hx=hx, hy=hy,
                hz=hz, zperiod=float(cuc), nodes=nodes)


def read_tasos_grids(files, auc, buc, cuc):
    """The grids of the named files (Read_Grids) as the mspms_sf_grid
    fields."""
    return tasos_grid([read_tasos_table(f, auc, buc, cuc) for f in files], auc, buc, cuc)
