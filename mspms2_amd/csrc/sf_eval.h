// sf_eval.h -- the solid-fluid term of one site (src/frclong.c:1145-1209),
// shared by the frclong pass (k_sf.hip) and the RESPA inner loop when the
// wall is split into the short-range forces (k_integrate.hip).
//   SF_HYPERGEO       smooth-cylinder LJ tube, hypergeometric series
//                     (cal_sf_hypergeo, src/sffrc.c:1092-1209), FP64
//   SF_ATOM_EXPLICIT  LJ against explicit wall atoms
//                     (cal_sf_atom_explicit, src/sffrc.c:1213-1314), FP64
//   SF_TASOS          tricubic Hermite interpolation of an FP32 node grid
//                     (call_tasos_forces, src/sffrc.c:1322-1337 ->
//                     CForce_atom, mylibtasos/forces.f90:6-173): a gather of
//                     8 corners x 8 values = 256 B per site from the
//                     node-major device table, FP64 arithmetic
//   SF_MY_INTERP      32-monomial polynomials per cube (get_values_from_grid,
//                     src/sffrc.c:951-1089): one contiguous KiB per site
//                     (cube-major tables), FP64, the reference's order
#pragma once
#include "engine.h"

namespace sfeval {

__host__ __device__ inline double mic(double d, double L) {
#pragma clang fp contract(off)
    return d - L * rint(d / L);
}

// hypergeo (src/sffrc.c:1092-1116)
__host__ __device__ inline void hypergeo(double a, double b, double c, double z, double rr, double* sumV,
                                double* sumF) {
#pragma clang fp contract(off)
    double zn = 1.0, an1 = 1.0;
    double sv = 1.0, sf = 0.0;
    for (int i = 0; i <= 10; i++) {
        double an = (a + i) * (b + i) / ((c + i) * (1 + i));
        an1 = an1 * an;
        zn *= z;
        sv = sv + an1 * zn;
        sf = sf - 2.0 * (i + 1.0) * an1 * zn / rr;
    }
    *sumV = sv;
    *sumF = sf;
}

// 1-D cubic Hermite basis on t in [0,1], spacing h: value weights of the two
// nodes, slope weights (times h), and their t-derivatives divided by h
// (mylibtasos/cubicauxiliary.f90:1-60: g, -f, h w(t), -h w(1-t)).
__host__ __device__ inline void hermite_1d(double t, double h, double b[4], double db[4]) {
    double t2 = t * t, t3 = t2 * t;
    b[0] = 2.0 * t3 - 3.0 * t2 + 1.0;
    b[1] = 3.0 * t2 - 2.0 * t3;
    b[2] = h * (t3 - 2.0 * t2 + t);
    b[3] = h * (t3 - t2);
    db[0] = (6.0 * t2 - 6.0 * t) / h;
    db[1] = (6.0 * t - 6.0 * t2) / h;
    db[2] = 3.0 * t2 - 4.0 * t + 1.0;
    db[3] = 3.0 * t2 - 2.0 * t;
}

// The solid-fluid energy of a GHOST_NONE site at (xi, yi, zi) (reduced
// units); its force is added to (fxi, fyi, fzi) as the reference adds it to
// fxl.  A site outside a grid's tables sets *edom and adds nothing.  Host
// callers pass tables in host memory.
// The 32 monomials of Amatrix_ele_assign_value (src/sffrc.c:750-803) at
// absolute (X, Y, Z) and the four cube polynomials (E, Fx, Fy, Fz), in the
// reference's order; q(f, m) is coefficient m of polynomial f.
template <class Coef>
__host__ __device__ inline void mi_poly(const Coef& q, double X, double Y, double Z,
                                        double r[4]) {
#pragma clang fp contract(off)
    const double x2 = X * X, x3 = x2 * X, y2 = Y * Y, y3 = y2 * Y, z2 = Z * Z, z3 = z2 * Z;
    const double v[32] = {1.0,        X,          Y,          Z,      x2,     X * Y,
                          X * Z,      y2,         Y * Z,      z2,     x3,     x2 * Y,
                          X * y2,     x2 * Z,     X * z2,     y3,     y2 * Z, Y * z2,
                          z3,         X * Y * Z,  x3 * Y,     x2 * y2, X * y3, x3 * Z,
                          x2 * z2,    X * z3,     y3 * Z,     y2 * z2, Y * z3, x2 * Y * Z,
                          X * y2 * Z, X * Y * z2};
#pragma unroll
    for (int f = 0; f < 4; f++) {
        double a = q(f, 0) * v[0];
#pragma unroll
        for (int m = 1; m < 32; m++) a = a + q(f, m) * v[m];
        r[f] = a;
    }
}

// get_values_from_grid's cube location (src/sffrc.c:968-987) for reduced
// (x, y, z): the absolute coordinates (z folded into [0, zmax)) and the cube
// index, or -1 outside the tables (the reference reads past them)
__host__ __device__ inline int mi_cube(double xi, double yi, double zi, double sigma_base,
                                       double x0, double y0, double zmax, double hx, double hy,
                                       double hz, int nx, int ny, int nz, double* X, double* Y,
                                       double* Z) {
#pragma clang fp contract(off)
    const double fxx = xi * sigma_base, fyy = yi * sigma_base;
    double fzz = zi * sigma_base;
    const double tx = (fxx - x0) / hx, tyy = (fyy - y0) / hy;
    fzz = fzz - zmax * rint(fzz / zmax);  // periodic in z
    if (fzz < 0) fzz = zmax + fzz;
    const double tz = (fzz - 0.0) / hz;   // zmin = 0
    *X = fxx;
    *Y = fyy;
    *Z = fzz;
    if (!(tx > -1.0 && tx < nx && tyy > -1.0 && tyy < ny && tz > -1.0 && tz < nz)) return -1;
    const int iAx = (int)tx, iAy = (int)tyy, iAz = (int)tz;
    return iAx * ny * nz + iAy * nz + iAz;
}

__host__ __device__ inline double sf_site(const SFDev& sf, const RepDev& R, const SiteDev& S,
                                          double xi, double yi, double zi, double& fxi,
                                          double& fyi, double& fzi, int* edom,
                                          double* fij_last = nullptr) {
#pragma clang fp contract(off)
    double u = 0.0;
    if (sf.type == MSPMS_SF_HYPERGEO) {
        double sij = 0.5 * (S.sigma + sf.solid_sigma);
        double eij = sqrt(S.epsilon * sf.solid_epsilon);
        double hfx = fxi, hfy = fyi;
        for (int t = 0; t < sf.ntube; t++) {
            double rx = mic(xi - sf.tube[3 * t], R.L[0]);
            double ry = mic(yi - sf.tube[3 * t + 1], R.L[1]);
            double rsq = rx * rx + ry * ry, r = sqrt(rsq);
            double TR = sf.tube[3 * t + 2], TRsq = TR * TR;
            double ljc, k, zf, hr, hfr, ha, hfa, k2, k4, k10, r1, r2, a1, a2, fij;
            if (r < TR) {
                ljc = sf.c3pisq * eij * sij * sij;
                k = sij * TR / (TRsq - rsq);
                zf = rsq / TRsq;
                hypergeo(-4.5, -4.5, 1.0, zf, r, &hr, &hfr);
                k2 = k * k;
                k4 = k2 * k2;
                k10 = k4 * k4 * k2;
                r1 = 0.65625 * k10;
                r2 = -hfr + hr * (20.0 * r / (TRsq - rsq));
                hypergeo(-1.5, -1.5, 1.0, zf, r, &ha, &hfa);
                a1 = -k4;
                a2 = -hfa + ha * (8.0 * r / (TRsq - rsq));
                u = ljc * (r1 * hr + a1 * ha);
                fij = -ljc * (r1 * r2 + a1 * a2);
            } else {
                ljc = sf.c3pisq * eij * sij * sij * TR;
                k = sij * r / (rsq - TRsq);
                zf = TRsq / rsq;
                hypergeo(-4.5, -4.5, 1.0, zf, r, &hr, &hfr);
                k2 = k * k;
                k4 = k2 * k2;
                k10 = k4 * k4 * k2;
                r1 = 0.65625 * k10;
                r2 = hfr / r + hr * (-20.0 / (rsq - TRsq) + 9.0 / rsq);
                hypergeo(-1.5, -1.5, 1.0, zf, r, &ha, &hfa);
                a1 = -k4;
                a2 = hfa / r + ha * (-8.0 / (rsq - TRsq) + 3.0 / rsq);
                u = ljc * (r1 * hr + a1 * ha) / r;
                fij = -ljc * (r1 * r2 + a1 * a2);
            }
            hfx += fij * rx / r;
            hfy += fij * ry / r;
            if (fij_last) *fij_last = fij;
        }
        fxi = hfx;
        fyi = hfy;
    } else if (sf.type == MSPMS_SF_ATOM_EXPLICIT) {
        double sij = 0.5 * (S.sigma + sf.solid_sigma);
        double eij = sqrt(S.epsilon * sf.solid_epsilon);
        double efx = fxi, efy = fyi, efz = fzi, us = 0.0;
        for (int j = 0; j < sf.nsolid; j++) {
            double rx = mic(xi - sf.solid[3 * j], R.L[0]);
            double ry = mic(yi - sf.solid[3 * j + 1], R.L[1]);
            double rz = mic(zi - sf.solid[3 * j + 2], R.L[2]);
            double rsq = rx * rx + ry * ry + rz * rz;
            if (!(rsq < R.rc2_lj)) continue;
            double sw = 0.0;
            if (R.switch_on) {
                if (rsq < R.rcon2)
                    sw = 1.0;
                else
                    sw = (R.rc2_lj - rsq) * (R.rc2_lj - rsq) * (R.rc2_lj + 2.0 * rsq - 3.0 * R.rcon2) /
                         R.roff2_ron2_cube;
            }
            double rr = sij * sij / rsq, r6 = rr * rr * rr, r12 = r6 * r6, d = r12 - r6;
            double ut = eij * d, fij;
            us += R.switch_on ? ut * sw : ut;
            if (R.switch_on) {
                if (rsq < R.rcon2)
                    fij = 24.0 * eij * (d + r12) / rsq;
                else
                    fij = 24.0 * eij * (d + r12) / rsq * sw -
                          4.0 * ut * 12.0 * (R.rc2_lj - rsq) * (R.rcon2 - rsq) / R.roff2_ron2_cube;
            } else {
                fij = 24.0 * eij * (d + r12) / rsq;
            }
            efx += fij * rx;
            efy += fij * ry;
            efz += fij * rz;
            if (fij_last) *fij_last = fij;
        }
        fxi = efx;
        fyi = efy;
        fzi = efz;
        u = us * 4.0;
    } else if (sf.type == MSPMS_SF_TASOS) {
        // call_tasos_forces: Angstrom in, K and K/A out, reduced back
        const int ty = S.interp - 1;
        if (ty < 0 || ty >= sf.ntype) {
            *edom = 1;
            return 0.0;
        }
        double xp = xi * sf.sigma_base - sf.x0, yp = yi * sf.sigma_base - sf.y0,
               zp = zi * sf.sigma_base - sf.z0;
        const double pz = sf.zperiod;
        zp = zp - floor(zp / pz) * pz;  // mylibtasos/pbc.f90:82-86
        int ix = (int)floor(xp / sf.hx), iy = (int)floor(yp / sf.hy), iz = (int)floor(zp / sf.hz);
        if (!(ix >= 0 && iy >= 0 && iz >= 0 && ix <= sf.nx - 2 && iy <= sf.ny - 2 &&
              iz <= sf.nz - 2)) {
            *edom = 1;
            return 0.0;
        }
        double bx[4], by[4], bz[4], dbx[4], dby[4], dbz[4];
        hermite_1d((xp - ix * sf.hx) / sf.hx, sf.hx, bx, dbx);
        hermite_1d((yp - iy * sf.hy) / sf.hy, sf.hy, by, dby);
        hermite_1d((zp - iz * sf.hz) / sf.hz, sf.hz, bz, dbz);
        const size_t nn = (size_t)sf.nx * sf.ny * sf.nz;
        // node-major device layout: a node's eight values (32 B) in
        // derivative-bit order o = (x, y, z) bits, so a corner is two float4
        const float4* base = reinterpret_cast<const float4*>(sf.nodes) + (size_t)ty * 2 * nn;
        double fval = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
            const size_t node = ((size_t)(iz + cz) * sf.ny + (iy + cy)) * sf.nx + (ix + cx);
            const float4 lo = base[2 * node], hi = base[2 * node + 1];
            const float vals[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int o = 0; o < 8; o++) {
                const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
                const double v = (double)vals[o];
                const double wx = bx[cx + 2 * ox], wy = by[cy + 2 * oy], wz = bz[cz + 2 * oz];
                fval += v * wx * wy * wz;
                gx += v * dbx[cx + 2 * ox] * wy * wz;
                gy += v * wx * dby[cy + 2 * oy] * wz;
                gz += v * wx * wy * dbz[cz + 2 * oz];
            }
        }
        double pe = (fval < 1.5 * sf.min_pe[ty]) ? -fval : fval;  // forces.f90:166-171
        const double fconv = sf.sigma_base / sf.epsilon_base;
        u = pe / sf.epsilon_base;
        fxi += -gx * fconv;
        fyi += -gy * fconv;
        fzi += -gz * fconv;
    } else if (sf.type == MSPMS_SF_MY_INTERP) {
        // get_values_from_grid (src/sffrc.c:951-1089) and its caller
        // (src/frclong.c:1193-1207): the same expressions in the same order
        const int ty = S.interp - 1;  // init_my_interp decrements interp_type
        if (ty < 0 || ty >= sf.ntype) {
            *edom = 1;
            return 0.0;
        }
        double X, Y, Z;
        const int cubeidx = mi_cube(xi, yi, zi, sf.sigma_base, sf.x0, sf.y0, sf.zmax, sf.hx,
                                    sf.hy, sf.hz, sf.nx, sf.ny, sf.nz, &X, &Y, &Z);
        if (cubeidx < 0) {
            *edom = 1;
            return 0.0;
        }
        const double* q =
            sf.cubes + ((size_t)ty * ((size_t)sf.nx * sf.ny * sf.nz) + (size_t)cubeidx) * 128;
        double r[4];
        mi_poly([q](int f, int m) { return q[f * 32 + m]; }, X, Y, Z, r);
        const double R_RGAS = 0.120272219;  // src/defs.h:96
        u = r[0] * R_RGAS / sf.epsilon_base;
        fxi += r[1] * R_RGAS * sf.sigma_base / sf.epsilon_base;
        fyi += r[2] * R_RGAS * sf.sigma_base / sf.epsilon_base;
        fzi += r[3] * R_RGAS * sf.sigma_base / sf.epsilon_base;
    }
    return u;
}

}  // namespace sfeval
