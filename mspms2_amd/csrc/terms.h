// terms.h -- per-molecule term tables and the bonded terms (FP64, device).
//
// bonded_term() restates bndfrc (src/bndfrc.c:17-78), aglfrc
// (src/aglfrc.c:17-190), dihfrc (src/dihfrc.c:17-284) and impfrc
// (src/impfrc.c:17-191) with the same expressions, contraction disabled.
#pragma once
#include <hip/hip_runtime.h>
#include "physics.h"
#include "../../include/mspms_gpu.h"

enum {
    TERM_PAIR12 = 0,  // frclong intramolecular: bond ending pair
    TERM_PAIR13,      //   unique angle ending pair
    TERM_PAIR14,      //   unique dihedral ending pair
    TERM_BOND,        // frcshort
    TERM_ANGLE,
    TERM_DIH,
    TERM_IMP
};

struct TermDev {
    int kind, type, idx[4];
    double para[5];
};

struct SpeciesDev {
    int natom, site0;      // site0: global site id of the template's site 0
    int term0, nintra;     // frclong intramolecular pair terms
    int bterm0, nbterm;    // frcshort terms, in the reference's order
    int pad[2];
    double mw;             // molecular mass, sum of aw in site order
};

#define ST_SINGULAR 8  // improper geometry the reference exit(1)s on

__host__ __device__ inline void charmm_geom(int i1, int i2, int i3, int i4, const double* xx,
                                   const double* yy, const double* zz, double g[18]) {
#pragma clang fp contract(off)
    double rxjk = xx[i3] - xx[i2], ryjk = yy[i3] - yy[i2], rzjk = zz[i3] - zz[i2];
    double rxij = xx[i1] - xx[i2], ryij = yy[i1] - yy[i2], rzij = zz[i1] - zz[i2];
    double rxjl = xx[i4] - xx[i2], ryjl = yy[i4] - yy[i2], rzjl = zz[i4] - zz[i2];
    double aa1 = ryjk * rzij - rzjk * ryij, bb1 = rzjk * rxij - rxjk * rzij,
           cc1 = rxjk * ryij - rxij * ryjk;
    double aa2 = ryjl * rzjk - rzjl * ryjk, bb2 = rzjl * rxjk - rxjl * rzjk,
           cc2 = rxjl * ryjk - rxjk * ryjl;
    g[0] = aa1; g[1] = bb1; g[2] = cc1; g[3] = aa2; g[4] = bb2; g[5] = cc2;
    g[6] = sqrt(aa1 * aa1 + bb1 * bb1 + cc1 * cc1);
    g[7] = sqrt(aa2 * aa2 + bb2 * bb2 + cc2 * cc2);
    g[8] = (aa1 * aa2 + bb1 * bb2 + cc1 * cc2);
    g[9] = rxjk; g[10] = ryjk; g[11] = rzjk;
    g[12] = rxij; g[13] = ryij; g[14] = rzij;
    g[15] = rxjl; g[16] = ryjl; g[17] = rzjl;
}

// the four-atom gradient shared by src/dihfrc.c:423-468 and
// src/impfrc.c:140-181
__host__ __device__ inline void charmm_forces(int i1, int i2, int i3, int i4, const double* xx,
                                     const double* yy, const double* zz, double* fxs,
                                     double* fys, double* fzs, double temp, const double g[18]) {
#pragma clang fp contract(off)
    const double aa1 = g[0], bb1 = g[1], cc1 = g[2], aa2 = g[3], bb2 = g[4], cc2 = g[5];
    const double t1 = g[6], t2 = g[7], t3 = g[8];
    const double rxjk = g[9], ryjk = g[10], rzjk = g[11], rxij = g[12], ryij = g[13],
                 rzij = g[14], rxjl = g[15], ryjl = g[16], rzjl = g[17];
    double fij = temp / (t2 * t1 * t1);
    fxs[i1] = fxs[i1] + fij * ((bb2 * rzjk - cc2 * ryjk) * t1 - t3 / t1 * (bb1 * rzjk - cc1 * ryjk));
    fys[i1] = fys[i1] + fij * ((-aa2 * rzjk + cc2 * rxjk) * t1 - t3 / t1 * (-aa1 * rzjk + cc1 * rxjk));
    fzs[i1] = fzs[i1] + fij * ((aa2 * ryjk - bb2 * rxjk) * t1 - t3 / t1 * (aa1 * ryjk - bb1 * rxjk));
    fij = temp / (t1 * t1 * t2 * t2);
    double rxik = xx[i1] - xx[i3], ryik = yy[i1] - yy[i3], rzik = zz[i1] - zz[i3];
    double rxkl = xx[i4] - xx[i3], rykl = yy[i4] - yy[i3], rzkl = zz[i4] - zz[i3];
    fxs[i2] = fxs[i2] + fij * ((bb2 * rzik - bb1 * rzkl - cc2 * ryik + cc1 * rykl) * t1 * t2 -
                               t3 * (t2 / t1 * (bb1 * rzik - cc1 * ryik) + t1 / t2 * (-bb2 * rzkl + cc2 * rykl)));
    fys[i2] = fys[i2] + fij * ((-aa2 * rzik + aa1 * rzkl + cc2 * rxik - cc1 * rxkl) * t1 * t2 -
                               t3 * (t2 / t1 * (-aa1 * rzik + cc1 * rxik) + t1 / t2 * (aa2 * rzkl - cc2 * rxkl)));
    fzs[i2] = fzs[i2] + fij * ((aa2 * ryik - aa1 * rykl - bb2 * rxik + bb1 * rxkl) * t1 * t2 -
                               t3 * (t2 / t1 * (aa1 * ryik - bb1 * rxik) + t1 / t2 * (-aa2 * rykl + bb2 * rxkl)));
    fxs[i3] = fxs[i3] + fij * ((-bb2 * rzij + bb1 * rzjl + cc2 * ryij - cc1 * ryjl) * t1 * t2 -
                               t3 * (t2 / t1 * (-bb1 * rzij + cc1 * ryij) + t1 / t2 * (bb2 * rzjl - cc2 * ryjl)));
    fys[i3] = fys[i3] + fij * ((aa2 * rzij - aa1 * rzjl - cc2 * rxij + cc1 * rxjl) * t1 * t2 -
                               t3 * (t2 / t1 * (aa1 * rzij - cc1 * rxij) + t1 / t2 * (-aa2 * rzjl + cc2 * rxjl)));
    fzs[i3] = fzs[i3] + fij * ((-aa2 * ryij + aa1 * ryjl + bb2 * rxij - bb1 * rxjl) * t1 * t2 -
                               t3 * (t2 / t1 * (-aa1 * ryij + bb1 * rxij) + t1 / t2 * (aa2 * ryjl - bb2 * rxjl)));
    fij = temp / (t1 * t2 * t2);
    fxs[i4] = fxs[i4] + fij * ((-bb1 * rzjk + cc1 * ryjk) * t2 - t3 / t2 * (-bb2 * rzjk + cc2 * ryjk));
    fys[i4] = fys[i4] + fij * ((aa1 * rzjk - cc1 * rxjk) * t2 - t3 / t2 * (aa2 * rzjk - cc2 * rxjk));
    fzs[i4] = fzs[i4] + fij * ((-aa1 * ryjk + bb1 * rxjk) * t2 - t3 / t2 * (-aa2 * ryjk + bb2 * rxjk));
}

// Returns the energy class (0 bond, 1 angle, 2 dihedral, 3 improper).
// Atom k of the term sits at index base + idx[k] * stride of the arrays
// (stride 1: the global SoA arrays; a per-molecule LDS slice otherwise).
// *singular is set for an improper at the geometry impfrc exits on
// (src/impfrc.c:96-134); the term then adds nothing.  Host and device.
__host__ __device__ inline int bonded_term(const TermDev& T, int base, const double* xx,
                                           const double* yy, const double* zz, double* fxs,
                                           double* fys, double* fzs, double* u, double* vir,
                                           int* singular, int stride = 1) {
#pragma clang fp contract(off)
    const double* pr = T.para;
    if (T.kind == TERM_BOND) {
        int i1 = base + T.idx[0] * stride, i2 = base + T.idx[1] * stride;
        double rx = xx[i1] - xx[i2], ry = yy[i1] - yy[i2], rz = zz[i1] - zz[i2];
        double rsq = rx * rx + ry * ry + rz * rz, r = sqrt(rsq), fij = 0.0;
        if (T.type == MSPMS_BOND_HARMONIC) {
            double dr = r - pr[1];
            *u = 0.5 * pr[0] * dr * dr;
            fij = -pr[0] * dr / r;
        } else if (T.type == MSPMS_BOND_MORSE) {
            double dr = r - pr[1], De = pr[0], e = exp(-pr[2] * dr), om = 1.0 - e;
            *u = De * om * om;
            fij = -2.0 * De * pr[2] * om * e / r;
        } else {  // FENE
            double rmax2 = pr[1] * pr[1];
            *u = pr[0] * rmax2 * log(1.0 - rsq / rmax2);
            fij = -2.0 * pr[0] * rmax2 / (rmax2 - rsq);
        }
        *vir = fij * rsq;
        fxs[i1] += fij * rx;
        fys[i1] += fij * ry;
        fzs[i1] += fij * rz;
        fxs[i2] -= fij * rx;
        fys[i2] -= fij * ry;
        fzs[i2] -= fij * rz;
        return 0;
    }
    if (T.kind == TERM_ANGLE) {
        int ia = base + T.idx[0] * stride, ib = base + T.idx[1] * stride,
            ic = base + T.idx[2] * stride;
        if (T.type == MSPMS_ANGLE_HARMONIC) {
            double xab = xx[ia] - xx[ib], yab = yy[ia] - yy[ib], zab = zz[ia] - zz[ib];
            double xbc = xx[ic] - xx[ib], ybc = yy[ic] - yy[ib], zbc = zz[ic] - zz[ib];
            double rab = sqrt(xab * xab + yab * yab + zab * zab), rrab = 1.0 / rab;
            xab = xab * rrab;
            yab = yab * rrab;
            zab = zab * rrab;
            double rbc = sqrt(xbc * xbc + ybc * ybc + zbc * zbc), rrbc = 1.0 / rbc;
            xbc = xbc * rrbc;
            ybc = ybc * rrbc;
            zbc = zbc * rrbc;
            double cost = xab * xbc + yab * ybc + zab * zbc;
            double theta = acos(cost);
            double sint = fmax(1.0e-8, sqrt(1.0 - cost * cost));
            double dth = theta - pr[1];
            *u = 0.5 * pr[0] * dth * dth;
            double gam = pr[0] * dth / sint;
            double fxa = gam * (xbc - xab * cost) * rrab, fya = gam * (ybc - yab * cost) * rrab,
                   fza = gam * (zbc - zab * cost) * rrab;
            double fxc = gam * (xab - xbc * cost) * rrbc, fyc = gam * (yab - ybc * cost) * rrbc,
                   fzc = gam * (zab - zbc * cost) * rrbc;
            *vir = 0.0;
            fxs[ia] += fxa;
            fys[ia] += fya;
            fzs[ia] += fza;
            fxs[ib] = fxs[ib] - fxa - fxc;
            fys[ib] = fys[ib] - fya - fyc;
            fzs[ib] = fzs[ib] - fza - fzc;
            fxs[ic] += fxc;
            fys[ic] += fyc;
            fzs[ic] += fzc;
        } else {  // ANGLE_TR_WATER
            double v13x = xx[ia] - xx[ic], v13y = yy[ia] - yy[ic], v13z = zz[ia] - zz[ic];
            double v12x = xx[ia] - xx[ib], v12y = yy[ia] - yy[ib], v12z = zz[ia] - zz[ib];
            double v23x = xx[ib] - xx[ic], v23y = yy[ib] - yy[ic], v23z = zz[ib] - zz[ic];
            double r13 = sqrt(v13x * v13x + v13y * v13y + v13z * v13z);
            double r12 = sqrt(v12x * v12x + v12y * v12y + v12z * v12z);
            double r23 = sqrt(v23x * v23x + v23y * v23y + v23z * v23z);
            double d3 = r13 - pr[3], d1 = r12 - pr[4], d2 = r23 - pr[4];
            double kt = pr[0], krt = pr[1], krr = pr[2];
            *u = 0.5 * kt * d3 * d3 + krt * d3 * (d1 + d2) + krr * d1 * d2;
            double t1 = -kt * d3 / r13, t21 = -krt * d3 / r12, t22 = -krt * (d1 + d2) / r13,
                   t3 = -krr * d2 / r12;
            double fxa = (t1 * v13x + t21 * v12x + t22 * v13x + t3 * v12x);
            double fya = (t1 * v13y + t21 * v12y + t22 * v13y + t3 * v12y);
            double fza = (t1 * v13z + t21 * v12z + t22 * v13z + t3 * v12z);
            fxs[ia] += fxa;
            fys[ia] += fya;
            fzs[ia] += fza;
            t21 = -krt * d3 / r23;
            t22 = krt * d3 / r12;
            double t31 = -krr * d1 / r23, t32 = krr * d2 / r12;
            double fxb = (t21 * v23x + t22 * v12x + t31 * v23x + t32 * v12x);
            double fyb = (t21 * v23y + t22 * v12y + t31 * v23y + t32 * v12y);
            double fzb = (t21 * v23z + t22 * v12z + t31 * v23z + t32 * v12z);
            fxs[ib] += fxb;
            fys[ib] += fyb;
            fzs[ib] += fzb;
            t1 = kt * d3 / r13;
            t21 = krt * d3 / r23;
            t22 = krt * (d1 + d2) / r13;
            t3 = krr * d1 / r23;
            double fxc = (t1 * v13x + t21 * v23x + t22 * v13x + t3 * v23x);
            double fyc = (t1 * v13y + t21 * v23y + t22 * v13y + t3 * v23y);
            double fzc = (t1 * v13z + t21 * v23z + t22 * v13z + t3 * v23z);
            fxs[ic] += fxc;
            fys[ic] += fyc;
            fzs[ic] += fzc;
            double v = fxa * v12x + fya * v12y + fza * v12z;
            *vir = v - fxc * v23x - fyc * v23y - fzc * v23z;
        }
        return 1;
    }
    if (T.kind == TERM_DIH) {
        int ia = base + T.idx[0] * stride, ib = base + T.idx[1] * stride,
            ic = base + T.idx[2] * stride, id = base + T.idx[3] * stride;
        if (T.type == MSPMS_DIH_OPLS_COSIN) {
            double xab = xx[ia] - xx[ib], yab = yy[ia] - yy[ib], zab = zz[ia] - zz[ib];
            double xbc = xx[ib] - xx[ic], ybc = yy[ib] - yy[ic], zbc = zz[ib] - zz[ic];
            double xcd = xx[ic] - xx[id], ycd = yy[ic] - yy[id], zcd = zz[ic] - zz[id];
            double rrbc = 1.0 / sqrt(xbc * xbc + ybc * ybc + zbc * zbc);
            double pbx = yab * zbc - zab * ybc, pby = zab * xbc - xab * zbc, pbz = xab * ybc - yab * xbc;
            double pb2 = pbx * pbx + pby * pby + pbz * pbz, rpb1 = 1.0 / sqrt(pb2), rpb2 = rpb1 * rpb1;
            double pcx = ybc * zcd - zbc * ycd, pcy = zbc * xcd - xbc * zcd, pcz = xbc * ycd - ybc * xcd;
            double pc2 = pcx * pcx + pcy * pcy + pcz * pcz, rpc1 = 1.0 / sqrt(pc2), rpc2 = rpc1 * rpc1;
            double pbpc = pbx * pcx + pby * pcy + pbz * pcz;
            double cost = pbpc * rpb1 * rpc1;
            double sint = (xbc * (pcy * pbz - pcz * pby) + ybc * (pbx * pcz - pbz * pcx) +
                           zbc * (pcx * pby - pcy * pbx)) * (rpb1 * rpc1 * rrbc);
            double theta = atan2(sint, cost);
            sint = copysign(fmax(1.0e-8, fabs(sint)), sint);
            double rsint = 1.0 / sint;
            *u = pr[0] + 0.5 * (pr[1] * (1.0 + cos(theta)) + pr[2] * (1.0 - cos(2.0 * theta)) +
                                pr[3] * (1.0 + cos(3.0 * theta)));
            *vir = 0.0;
            double dopls = -0.5 * (pr[1] * sin(theta) - 2.0 * pr[2] * sin(2.0 * theta) +
                                   3.0 * pr[3] * sin(3.0 * theta));
            double gam = dopls * rpb1 * rpc1 * rsint;
            double fax = gam * ((-pcy * zbc + pcz * ybc) - pbpc * rpb2 * (-pby * zbc + pbz * ybc));
            double fay = gam * ((pcx * zbc - pcz * xbc) - pbpc * rpb2 * (pbx * zbc - pbz * xbc));
            double faz = gam * ((-pcx * ybc + pcy * xbc) - pbpc * rpb2 * (-pbx * ybc + pby * xbc));
            double fcx = gam * ((-pcy * zab + pcz * yab) - pbpc * rpb2 * (-pby * zab + pbz * yab));
            double fcy = gam * ((pcx * zab - pcz * xab) - pbpc * rpb2 * (pbx * zab - pbz * xab));
            double fcz = gam * ((-pcx * yab + pcy * xab) - pbpc * rpb2 * (-pbx * yab + pby * xab));
            double fbx = gam * ((-pby * zcd + pbz * ycd) - pbpc * rpc2 * (-pcy * zcd + pcz * ycd));
            double fby = gam * ((pbx * zcd - pbz * xcd) - pbpc * rpc2 * (pcx * zcd - pcz * xcd));
            double fbz = gam * ((-pbx * ycd + pby * xcd) - pbpc * rpc2 * (-pcx * ycd + pcy * xcd));
            double fdx = gam * ((-pby * zbc + pbz * ybc) - pbpc * rpc2 * (-pcy * zbc + pcz * ybc));
            double fdy = gam * ((pbx * zbc - pbz * xbc) - pbpc * rpc2 * (pcx * zbc - pcz * xbc));
            double fdz = gam * ((-pbx * ybc + pby * xbc) - pbpc * rpc2 * (-pcx * ybc + pcy * xbc));
            fxs[ia] += fax;
            fys[ia] += fay;
            fzs[ia] += faz;
            fxs[ib] += (-fax - fcx + fbx);
            fys[ib] += (-fay - fcy + fby);
            fzs[ib] += (-faz - fcz + fbz);
            fxs[ic] += fcx - fbx - fdx;
            fys[ic] += fcy - fby - fdy;
            fzs[ic] += fcz - fbz - fdz;
            fxs[id] += fdx;
            fys[id] += fdy;
            fzs[id] += fdz;
        } else {  // DIH_CHARMM
            double g[18];
            charmm_geom(ia, ib, ic, id, xx, yy, zz, g);
            double kphi = pr[0], nper = pr[1], d0 = pr[2];
            double phi = g[8] / (g[6] * g[7]);
            double alpha = phi;
            phi = 180.0 - (acos(phi) * 180.0 / MSPMS_PI);
            *u = kphi * (1 + cos((nper * phi - d0) / 180 * MSPMS_PI));
            *vir = 0.0;
            double temp;
            if (fabs(sin(phi * MSPMS_PI / 180)) <= 1e-10)
                temp = -nper * nper * kphi * cos((nper * phi - d0) * MSPMS_PI / 180) / alpha;
            else
                temp = nper * kphi * sin((nper * phi - d0) * MSPMS_PI / 180) / sin(phi * MSPMS_PI / 180);
            charmm_forces(ia, ib, ic, id, xx, yy, zz, fxs, fys, fzs, temp, g);
        }
        return 2;
    }
    // TERM_IMP, IMP_CHARMM
    {
        int ia = base + T.idx[0] * stride, ib = base + T.idx[1] * stride,
            ic = base + T.idx[2] * stride, id = base + T.idx[3] * stride;
        double kom = pr[0], om0 = pr[1];
        double g[18];
        charmm_geom(ia, ib, ic, id, xx, yy, zz, g);
        double phi = g[8] / (g[6] * g[7]);
        double alpha = phi;
        phi = 180 - (acos(phi) * 180 / MSPMS_PI);
        double uu = kom * (phi - om0) * (phi - om0);
        uu *= (MSPMS_PI / 180) * (MSPMS_PI / 180);
        *u = uu;
        *vir = 0.0;
        double ang = acos(alpha) * 180 / MSPMS_PI;
        bool sing = ((fabs(om0) <= 1e-5) && (fabs(ang) <= 1e-20)) ||
                        ((fabs(om0 - 180) <= 1e-5) && (fabs(ang - 180) <= 1e-20)) ||
                        ((fabs(om0) <= 1e-5) && (fabs(ang - 180) <= 1e-20)) ||
                        ((fabs(om0 - 180) <= 1e-5) && (fabs(ang) <= 1e-20));
        if (sing) {
            *singular = 1;
            return 3;
        }
        double temp = -2.0 * kom * (phi - om0) * MSPMS_PI / 180 / sin(phi * MSPMS_PI / 180);
        charmm_forces(ia, ib, ic, id, xx, yy, zz, fxs, fys, fzs, temp, g);
        return 3;
    }
}
