/*
 * mspms_oracle.c -- CPU FP64 restatement of mspms2's frclong / frcshort.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * Written from the reference's behaviour: each routine below names the
 * reference lines it follows.  The arithmetic expressions and the loop order
 * are kept identical to the reference's so that the FP64 results match it
 * bit for bit when both are compiled with gcc -std=c99 -O2 (no contraction).
 * Deliberate differences from the reference, all documented:
 *   - src/frclong.c:86 fix: j-molecule atom offset (NaN otherwise).
 *   - state is passed by argument instead of global variables.
 *   - undefined-behaviour inputs are rejected (oracle_validate).
 */
#include "mspms_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_REF 3.141592653589793 /* src/defs.h:100 */
#define TOL_REF 1.0e-8           /* src/defs.h:94 TOLERANCE */

/* ------------------------------------------------------------------------- */
/* pair statistics: order-independent hash of (i<j) pairs                     */
static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static uint64_t pair_key(int i, int j)
{
    int a = i < j ? i : j, b = i < j ? j : i;
    return splitmix64(((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
}

/* ------------------------------------------------------------------------- */
/* term kernels                                                               */

/* ljfrc (src/ljfrc.c:21-71) */
static void lj_term(const mspms_params* p, double rijsq, double sigmaij,
                    double epsilonij, double* uij, double* fij, double* ushift)
{
    double sw = 0.0, r_rijsq, r6, r12, s2, s6;
    if (p->isLJswitchOn) {
        if (rijsq < p->rcutonsq)
            sw = 1.0;
        else
            sw = (p->rcutoffsq - rijsq) * (p->rcutoffsq - rijsq) *
                 (p->rcutoffsq + 2.0 * rijsq - 3.0 * p->rcutonsq) /
                 p->roff2_minus_ron2_cube;
    }
    r_rijsq = 1.0 / rijsq;
    s2 = sigmaij * sigmaij;
    r6 = s2 * r_rijsq;
    r6 = r6 * r6 * r6;
    r12 = r6 * r6;
    *uij = 4.0 * epsilonij * (r12 - r6);
    if (p->isLJswitchOn) {
        if (rijsq < p->rcutonsq) {
            *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq;
        } else {
            *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq * sw -
                   *uij * 12.0 * (p->rcutoffsq - rijsq) * (p->rcutonsq - rijsq) /
                       p->roff2_minus_ron2_cube;
            *uij *= sw;
        }
        *ushift = 0.0;
    } else {
        *fij = 24.0 * epsilonij * (2 * r12 - r6) * r_rijsq;
        s6 = s2 * s2 * s2;
        *ushift = epsilonij * s6 * (s6 * p->shift1 - 1.0);
    }
}

/* ewald_real_frc (src/elecfrc.c:17-33) */
static void ewald_real(const mspms_params* p, double rijsq, double qq,
                       double* uij, double* fij)
{
    double rij = sqrt(rijsq);
    double ut = p->coulomb_prefactor * qq / rij;
    double t1, t2;
    *uij = ut * erfc(p->kappa * rij);
    t1 = p->kappa * rij;
    t2 = ut * erfc(t1);
    *fij = (t2 + ut * 2.0 * t1 * exp(-t1 * t1) / sqrt(PI_REF)) / rijsq;
}

/* ewald_excl_frc (src/elecfrc.c:35-46) */
static void ewald_excl(const mspms_params* p, double rijsq, double qq,
                       double* uij, double* fij)
{
    double rij = sqrt(rijsq);
    *uij = p->coulomb_prefactor * qq / rij;
    *fij = -*uij / rijsq;
}

/* wolf_real_frc (src/elecfrc.c:48-60) */
static void wolf_real(const mspms_params* p, double rijsq, double qq,
                      double* uij, double* fij)
{
    double rij = sqrt(rijsq);
    *uij = p->coulomb_prefactor * qq *
           (erfc(p->kappa * rij) / rij + p->wolfvcon1 +
            p->wolfvcon2 * (rij - p->rcutoffelec));
    *fij = p->coulomb_prefactor * (qq / rij) *
           (erfc(p->kappa * rij) / rijsq +
            p->wolffcon1 * exp(-(p->kappa * rij) * (p->kappa * rij)) / rij +
            p->wolffcon2);
}

/* xatom_coulomb_frc (src/elecfrc.c:159-170) */
static void xatom_coulomb(const mspms_params* p, double rijsq, double qq,
                          double* uij, double* fij)
{
    double rij = sqrt(rijsq);
    *uij = p->coulomb_prefactor * qq / rij;
    *fij = *uij / rijsq;
}

/* ------------------------------------------------------------------------- */
/* accumulator state shared by the frclong sections                           */
typedef struct {
    const mspms_params* p;
    const double *xx, *yy, *zz;
    double *fxl, *fyl, *fzl;
    double uvdw, ushift, ureal, uexcl, ucoul, virial;
    int64_t nlj, nel;
    uint64_t hlj, hel;
} lr_state;

static double mic(double d, double L) { return d - L * rint(d / L); }

static void add_pair_force(lr_state* s, int ii, int jj, double fij, double rx,
                           double ry, double rz)
{
    double fx = fij * rx, fy = fij * ry, fz = fij * rz;
    s->fxl[ii] += fx;
    s->fyl[ii] += fy;
    s->fzl[ii] += fz;
    s->fxl[jj] -= fx;
    s->fyl[jj] -= fy;
    s->fzl[jj] -= fz;
}

/* cal_com_and_efg_one (src/misc.c:72-103): COM and inner coordinates */
static void com_inner(const mspms_species* sp, const double* xx,
                      const double* yy, const double* zz, int iabs, double mw,
                      double com[3], double* ex, double* fy, double* gz)
{
    int a;
    com[0] = com[1] = com[2] = 0.0;
    for (a = 0; a < sp->natom; a++) {
        com[0] += xx[iabs + a] * sp->aw[a];
        com[1] += yy[iabs + a] * sp->aw[a];
        com[2] += zz[iabs + a] * sp->aw[a];
    }
    com[0] /= mw;
    com[1] /= mw;
    com[2] /= mw;
    for (a = 0; a < sp->natom; a++) {
        ex[a] = xx[iabs + a] - com[0];
        fy[a] = yy[iabs + a] - com[1];
        gz[a] = zz[iabs + a] - com[2];
    }
}

static double species_mw(const mspms_species* sp)
{
    /* src/readins.c:115,138: mw accumulated site by site */
    double mw = 0.0;
    int a;
    for (a = 0; a < sp->natom; a++) mw += sp->aw[a];
    return mw;
}

/* xmole_coulomb_frc (src/elecfrc.c:67-154) */
static void xmole_coulomb(lr_state* s, const mspms_species* spi, int iabs,
                          const mspms_species* spj, int jabs)
{
    const mspms_params* p = s->p;
    double ci[3], cj[3], exi[256], fyi[256], gzi[256], exj[256], fyj[256],
        gzj[256];
    double rxcm, rycm, rzcm, rcmsq;
    int mm, nn;
    com_inner(spi, s->xx, s->yy, s->zz, iabs, species_mw(spi), ci, exi, fyi, gzi);
    com_inner(spj, s->xx, s->yy, s->zz, jabs, species_mw(spj), cj, exj, fyj, gzj);
    rxcm = mic(ci[0] - cj[0], p->boxlx);
    rycm = mic(ci[1] - cj[1], p->boxly);
    rzcm = mic(ci[2] - cj[2], p->boxlz);
    rcmsq = rxcm * rxcm + rycm * rycm + rzcm * rzcm;
    if (!(rcmsq < p->rcutoffelecsq)) return;
    for (mm = 0; mm < spi->natom; mm++) {
        if (spi->ghost_type[mm] == MSPMS_GHOST_FULL) continue;
        for (nn = 0; nn < spj->natom; nn++) {
            double rx, ry, rz, rsq, rij, uij, fij;
            if (spj->ghost_type[nn] == MSPMS_GHOST_FULL) continue;
            rx = exi[mm] + rxcm - exj[nn];
            ry = fyi[mm] + rycm - fyj[nn];
            rz = gzi[mm] + rzcm - gzj[nn];
            rsq = rx * rx + ry * ry + rz * rz;
            rij = sqrt(rsq);
            uij = p->coulomb_prefactor * spi->charge[mm] * spj->charge[nn] / rij;
            s->ucoul += uij;
            s->virial += uij;
            fij = uij / rsq;
            add_pair_force(s, iabs + mm, jabs + nn, fij, rx, ry, rz);
        }
    }
}

/* Inter-molecular loop, src/frclong.c:52-220 with the :86 fix. */
static void inter_molecular(lr_state* s, int nspecie, const mspms_species* sp)
{
    const mspms_params* p = s->p;
    const int ct = p->iChargeType;
    int isum_atom = 0, iS, jS;
    for (iS = 0; iS < nspecie; iS++) {
        const mspms_species* pi = sp + iS;
        int jsum_atom = isum_atom;
        for (jS = iS; jS < nspecie; jS++) {
            const mspms_species* pj = sp + jS;
            int same = (iS == jS);
            int nimole = same ? pi->nmole - 1 : pi->nmole;
            int iabs = isum_atom, im;
            for (im = 0; im < nimole; im++) {
                /* fixed :86: first j atom is the next molecule (same species)
                 * or the first molecule of species jS */
                int jabs = same ? iabs + pi->natom : jsum_atom;
                int jm0 = same ? im + 1 : 0, jm;
                for (jm = jm0; jm < pj->nmole; jm++) {
                    int mm, nn;
                    for (mm = 0; mm < pi->natom; mm++) {
                        int ii;
                        double xi, yi, zi, fxi, fyi, fzi;
                        if (pi->ghost_type[mm] == MSPMS_GHOST_FULL) continue;
                        ii = iabs + mm;
                        xi = s->xx[ii];
                        yi = s->yy[ii];
                        zi = s->zz[ii];
                        fxi = s->fxl[ii];
                        fyi = s->fyl[ii];
                        fzi = s->fzl[ii];
                        for (nn = 0; nn < pj->natom; nn++) {
                            int jj;
                            double rx, ry, rz, rsq, uij, fij, ush, fx, fy, fz;
                            if (pj->ghost_type[nn] == MSPMS_GHOST_FULL) continue;
                            jj = jabs + nn;
                            rx = xi - s->xx[jj];
                            ry = yi - s->yy[jj];
                            rz = zi - s->zz[jj];
                            rx = mic(rx, p->boxlx);
                            ry = mic(ry, p->boxly);
                            rz = mic(rz, p->boxlz);
                            rsq = rx * rx + ry * ry + rz * rz;
                            if (pi->ghost_type[mm] != MSPMS_GHOST_LJ &&
                                pj->ghost_type[nn] != MSPMS_GHOST_LJ &&
                                rsq < p->rcutoffsq) {
                                double sij = 0.5 * (pi->sigma[mm] + pj->sigma[nn]);
                                double eij = sqrt(pi->epsilon[mm] * pj->epsilon[nn]);
                                lj_term(p, rsq, sij, eij, &uij, &fij, &ush);
                                s->uvdw += uij;
                                s->ushift += ush;
                                s->virial += fij * rsq;
                                fx = fij * rx;
                                fy = fij * ry;
                                fz = fij * rz;
                                fxi += fx;
                                fyi += fy;
                                fzi += fz;
                                s->fxl[jj] -= fx;
                                s->fyl[jj] -= fy;
                                s->fzl[jj] -= fz;
                                s->nlj++;
                                s->hlj += pair_key(ii, jj);
                            }
                            if (ct == MSPMS_ELEC_EWALD || ct == MSPMS_ELEC_WOLF) {
                                double qq = pi->charge[mm] * pj->charge[nn];
                                if (rsq < p->rcutoffelecsq) {
                                    if (ct == MSPMS_ELEC_EWALD) {
                                        ewald_real(p, rsq, qq, &uij, &fij);
                                        s->ureal += uij;
                                    } else {
                                        wolf_real(p, rsq, qq, &uij, &fij);
                                        s->ureal += uij;
                                        s->virial += fij * rsq;
                                    }
                                    fx = fij * rx;
                                    fy = fij * ry;
                                    fz = fij * rz;
                                    fxi += fx;
                                    fyi += fy;
                                    fzi += fz;
                                    s->fxl[jj] -= fx;
                                    s->fyl[jj] -= fy;
                                    s->fzl[jj] -= fz;
                                    s->nel++;
                                    s->hel += pair_key(ii, jj);
                                }
                            }
                        }
                        s->fxl[ii] = fxi;
                        s->fyl[ii] = fyi;
                        s->fzl[ii] = fzi;
                    }
                    if (ct == MSPMS_ELEC_SIMPLE_COULOMB)
                        xmole_coulomb(s, pi, iabs, pj, jabs);
                    jabs += pj->natom;
                }
                iabs += pi->natom;
            }
            jsum_atom += pj->natom * pj->nmole;
        }
        isum_atom += pi->natom * pi->nmole;
    }
}

/* The real-space electrostatic part shared by every intramolecular pair
 * class: exclusion on the raw separation (bonds, angles only) then the MIC
 * real-space term (src/frclong.c:316-408, :499-592, :708-786). */
static void intra_elec(lr_state* s, int ii, int jj, double qq, double rx,
                       double ry, double rz, int with_excl)
{
    const mspms_params* p = s->p;
    const int ct = p->iChargeType;
    double rsq, uij, fij;
    if (ct == MSPMS_ELEC_EWALD || ct == MSPMS_ELEC_WOLF) {
        if (with_excl) {
            rsq = rx * rx + ry * ry + rz * rz;
            if (rsq < p->rcutoffelecsq) {
                ewald_excl(p, rsq, qq, &uij, &fij);
                s->uexcl += uij;
                if (ct == MSPMS_ELEC_WOLF) s->virial += fij * rsq;
                add_pair_force(s, ii, jj, fij, rx, ry, rz);
            }
        }
        rx = mic(rx, p->boxlx);
        ry = mic(ry, p->boxly);
        rz = mic(rz, p->boxlz);
        rsq = rx * rx + ry * ry + rz * rz;
        if (rsq < p->rcutoffelecsq) {
            if (ct == MSPMS_ELEC_EWALD) {
                ewald_real(p, rsq, qq, &uij, &fij);
                s->ureal += uij;
            } else {
                wolf_real(p, rsq, qq, &uij, &fij);
                s->ureal += uij;
                s->virial += fij * rsq;
            }
            add_pair_force(s, ii, jj, fij, rx, ry, rz);
        }
    } else if (ct == MSPMS_ELEC_SIMPLE_COULOMB && !with_excl) {
        /* only the 1-4 (and nbp) sections have a simple-Coulomb branch */
        rx = mic(rx, p->boxlx);
        ry = mic(ry, p->boxly);
        rz = mic(rz, p->boxlz);
        rsq = rx * rx + ry * ry + rz * rz;
        if (rsq < p->rcutoffelecsq) {
            xatom_coulomb(p, rsq, qq, &uij, &fij);
            s->ucoul += uij;
            s->virial += fij * rsq;
            add_pair_force(s, ii, jj, fij, rx, ry, rz);
        }
    }
}

/* LJ of an intramolecular pair on a given separation (no shift energy). */
static void intra_lj(lr_state* s, const mspms_species* sp, int mm, int nn,
                     int ii, int jj, double rx, double ry, double rz,
                     double escale)
{
    double rsq = rx * rx + ry * ry + rz * rz, uij, fij, ush;
    if (!(rsq < s->p->rcutoffsq)) return;
    lj_term(s->p, rsq, 0.5 * (sp->sigma[mm] + sp->sigma[nn]),
            escale * sqrt(sp->epsilon[mm] * sp->epsilon[nn]), &uij, &fij, &ush);
    s->uvdw += uij;
    s->virial += fij * rsq;
    add_pair_force(s, ii, jj, fij, rx, ry, rz);
}

static int is_long(const mspms_params* p, double rx, double ry, double rz)
{
    return fabs(rx) > p->boxlx * 0.5 || fabs(ry) > p->boxly * 0.5 ||
           fabs(rz) > p->boxlz * 0.5;
}

/* Bond (1-2) and unique-angle (1-3) ending pairs, src/frclong.c:234-409 and
 * :413-593.  Returns the isSameMole value the reference leaves behind, or -1
 * if the pair was skipped before setting it (never happens: the skip comes
 * after isSameMole=false). */
static int intra_12_13(lr_state* s, const mspms_species* sp, int base, int mm,
                       int nn)
{
    const mspms_params* p = s->p;
    int ii = base + mm, jj = base + nn;
    double rx = s->xx[ii] - s->xx[jj];
    double ry = s->yy[ii] - s->yy[jj];
    double rz = s->zz[ii] - s->zz[jj];
    int same = 1;
    if (is_long(p, rx, ry, rz)) {
        same = 0;
        if (sp->ghost_type[mm] == MSPMS_GHOST_FULL ||
            sp->ghost_type[nn] == MSPMS_GHOST_FULL)
            return 0;
    }
    if (!same) {
        /* src/frclong.c:277-312: LJ on the MIC separation; the raw separation
         * is then restored through `int` variables (:223), i.e. truncated. */
        double mx = mic(rx, p->boxlx), my = mic(ry, p->boxly), mz = mic(rz, p->boxlz);
        double rsq = mx * mx + my * my + mz * mz;
        if (sp->ghost_type[mm] != MSPMS_GHOST_LJ &&
            sp->ghost_type[nn] != MSPMS_GHOST_LJ && rsq < p->rcutoffsq)
            intra_lj(s, sp, mm, nn, ii, jj, mx, my, mz, 1.0);
        rx = (double)(int)rx;
        ry = (double)(int)ry;
        rz = (double)(int)rz;
    }
    intra_elec(s, ii, jj, sp->charge[mm] * sp->charge[nn], rx, ry, rz, 1);
    return same;
}

/* Unique dihedral 1-4 ending pair, src/frclong.c:597-787.  `same` is the
 * isSameMole value carried over from the bond / angle sections (:622,:632
 * compare instead of assign). */
static void intra_14(lr_state* s, const mspms_species* sp, int base, int mm,
                     int nn, int same, double f0)
{
    const mspms_params* p = s->p;
    int ii = base + mm, jj = base + nn;
    double rx = s->xx[ii] - s->xx[jj];
    double ry = s->yy[ii] - s->yy[jj];
    double rz = s->zz[ii] - s->zz[jj];
    if (is_long(p, rx, ry, rz)) {
        if (sp->ghost_type[mm] == MSPMS_GHOST_FULL ||
            sp->ghost_type[nn] == MSPMS_GHOST_FULL)
            return;
    }
    if (!same) {
        double mx = mic(rx, p->boxlx), my = mic(ry, p->boxly), mz = mic(rz, p->boxlz);
        double rsq = mx * mx + my * my + mz * mz;
        if (sp->ghost_type[mm] != MSPMS_GHOST_LJ &&
            sp->ghost_type[nn] != MSPMS_GHOST_LJ && rsq < p->rcutoffsq)
            intra_lj(s, sp, mm, nn, ii, jj, mx, my, mz, 1.0);
        rx = (double)(int)rx;
        ry = (double)(int)ry;
        rz = (double)(int)rz;
    } else {
        intra_lj(s, sp, mm, nn, ii, jj, rx, ry, rz, f0);
    }
    intra_elec(s, ii, jj, sp->charge[mm] * sp->charge[nn], rx, ry, rz, 0);
}

/* get_specie_and_relative_atom_id (src/misc.c:50-66) */
static void site_of(int nspecie, const mspms_species* sp, int abs_id,
                    int* specie, int* site)
{
    int s, a0 = 0;
    for (s = 0; s < nspecie; s++) {
        int n = sp[s].natom * sp[s].nmole;
        if (abs_id < a0 + n) {
            *specie = s;
            *site = (abs_id - a0) % sp[s].natom;
            return;
        }
        a0 += n;
    }
    *specie = 0;
    *site = 0;
}

/* hypergeo (src/sffrc.c:1092-1116) */
static void hypergeo(double a, double b, double c, double z, double rr,
                     double* sumV, double* sumF)
{
    double zn = 1.0, an1 = 1.0, an;
    int i;
    *sumV = 1.0;
    *sumF = 0.0;
    for (i = 0; i <= 10; i++) {
        an = (a + i) * (b + i) / ((c + i) * (1 + i));
        an1 = an1 * an;
        zn *= z;
        *sumV = *sumV + an1 * zn;
        *sumF = *sumF - 2.0 * (i + 1.0) * an1 * zn / rr;
    }
}

/* cal_sf_hypergeo (src/sffrc.c:1120-1209) */
static void sf_hypergeo(const mspms_params* p, const mspms_species* sp,
                        int site, double xi, double yi, double* fx, double* fy,
                        double* u)
{
    double sij = 0.5 * (sp->sigma[site] + p->solid_sigma);
    double eij = sqrt(sp->epsilon[site] * p->solid_epsilon);
    double fxi = *fx, fyi = *fy, uh = 0.0;
    int t;
    for (t = 0; t < p->ntube; t++) {
        double rx = mic(xi - p->tube_x[t], p->boxlx);
        double ry = mic(yi - p->tube_y[t], p->boxly);
        double rsq = rx * rx + ry * ry, r = sqrt(rsq);
        double R = p->tube_r[t], Rsq = R * R;
        double ljc, k, z, hr, hfr, ha, hfa, k2, k4, k10, r1, r2, a1, a2, fij;
        if (r < R) {
            ljc = p->const_3pisq_theta * eij * sij * sij;
            k = sij * R / (Rsq - rsq);
            z = rsq / Rsq;
            hypergeo(-4.5, -4.5, 1.0, z, r, &hr, &hfr);
            k2 = k * k;
            k4 = k2 * k2;
            k10 = k4 * k4 * k2;
            r1 = 0.65625 * k10;
            r2 = -hfr + hr * (20.0 * r / (Rsq - rsq));
            hypergeo(-1.5, -1.5, 1.0, z, r, &ha, &hfa);
            a1 = -k4;
            a2 = -hfa + ha * (8.0 * r / (Rsq - rsq));
            uh = ljc * (r1 * hr + a1 * ha);
            fij = -ljc * (r1 * r2 + a1 * a2);
        } else {
            ljc = p->const_3pisq_theta * eij * sij * sij * R;
            k = sij * r / (rsq - Rsq);
            z = Rsq / rsq;
            hypergeo(-4.5, -4.5, 1.0, z, r, &hr, &hfr);
            k2 = k * k;
            k4 = k2 * k2;
            k10 = k4 * k4 * k2;
            r1 = 0.65625 * k10;
            r2 = hfr / r + hr * (-20.0 / (rsq - Rsq) + 9.0 / rsq);
            hypergeo(-1.5, -1.5, 1.0, z, r, &ha, &hfa);
            a1 = -k4;
            a2 = hfa / r + ha * (-8.0 / (rsq - Rsq) + 3.0 / rsq);
            uh = ljc * (r1 * hr + a1 * ha) / r;
            fij = -ljc * (r1 * r2 + a1 * a2);
        }
        fxi += fij * rx / r;
        fyi += fij * ry / r;
    }
    *u = uh; /* overwritten per tube, src/sffrc.c:1201 */
    *fx = fxi;
    *fy = fyi;
}

/* cal_sf_atom_explicit (src/sffrc.c:1213-1314) */
static void sf_atom_explicit(const mspms_params* p, const mspms_species* sp,
                             int site, double xi, double yi, double zi,
                             double* fx, double* fy, double* fz, double* u)
{
    double sij = 0.5 * (sp->sigma[site] + p->solid_sigma);
    double eij = sqrt(sp->epsilon[site] * p->solid_epsilon);
    double usum = 0.0, fxi = *fx, fyi = *fy, fzi = *fz;
    int j;
    for (j = 0; j < p->solid_natom; j++) {
        double rx = mic(xi - p->solid_xx[j], p->boxlx);
        double ry = mic(yi - p->solid_yy[j], p->boxly);
        double rz = mic(zi - p->solid_zz[j], p->boxlz);
        double rsq = rx * rx + ry * ry + rz * rz;
        double sw = 0.0, rr, r6, r12, d, ut, fij;
        if (!(rsq < p->rcutoffsq)) continue;
        if (p->isLJswitchOn) {
            if (rsq < p->rcutonsq)
                sw = 1.0;
            else
                sw = (p->rcutoffsq - rsq) * (p->rcutoffsq - rsq) *
                     (p->rcutoffsq + 2.0 * rsq - 3.0 * p->rcutonsq) /
                     p->roff2_minus_ron2_cube;
        }
        rr = sij * sij / rsq;
        r6 = rr * rr * rr;
        r12 = r6 * r6;
        d = r12 - r6;
        ut = eij * d;
        usum += p->isLJswitchOn ? ut * sw : ut;
        if (p->isLJswitchOn) {
            if (rsq < p->rcutonsq)
                fij = 24.0 * eij * (d + r12) / rsq;
            else
                fij = 24.0 * eij * (d + r12) / rsq * sw -
                      4.0 * ut * 12.0 * (p->rcutoffsq - rsq) * (p->rcutonsq - rsq) /
                          p->roff2_minus_ron2_cube;
        } else {
            fij = 24.0 * eij * (d + r12) / rsq;
        }
        fxi += fij * rx;
        fyi += fij * ry;
        fzi += fij * rz;
    }
    *fx = fxi;
    *fy = fyi;
    *fz = fzi;
    *u = usum * 4.0;
}

/* ------------------------------------------------------------------------- */
/* tricubic Hermite grid (mylibtasos)                                         */

/* 1-D cubic Hermite basis on t in [0,1] with node spacing h:
 * value weights at the two nodes and derivative weights (times h), plus
 * their t-derivatives divided by h.  mylibtasos/cubicauxiliary.f90:1-60
 * writes the same four polynomials as g, -f, h*w(t), -h*w(1-t). */
static void hermite_1d(double t, double h, double b[4], double db[4])
{
    double t2 = t * t, t3 = t2 * t;
    b[0] = 2.0 * t3 - 3.0 * t2 + 1.0;          /* value at node 0 */
    b[1] = 3.0 * t2 - 2.0 * t3;                /* value at node 1 */
    b[2] = h * (t3 - 2.0 * t2 + t);            /* slope at node 0 */
    b[3] = h * (t3 - t2);                      /* slope at node 1 */
    db[0] = (6.0 * t2 - 6.0 * t) / h;
    db[1] = (6.0 * t - 6.0 * t2) / h;
    db[2] = 3.0 * t2 - 4.0 * t + 1.0;
    db[3] = 3.0 * t2 - 2.0 * t;
}

static double grid_node(const mspms_sf_grid* g, int type, int field, int ix,
                        int iy, int iz)
{
    size_t n = (size_t)g->nx * g->ny * g->nz;
    size_t node = ((size_t)iz * g->ny + iy) * g->nx + ix;
    return (double)g->nodes[((size_t)type * 8 + field) * n + node];
}

/* CForce_atom + CHermInt3D (mylibtasos/forces.f90:6-173,
 * cubicinterpolation.f90:9-116): interpolant of F over the cell holding
 * (x,y,z) from the 8 corner nodes' F, first, mixed second and mixed third
 * derivatives; Force = -grad; PE sign flipped when below 1.5*min node F
 * (forces.f90:166-171). */
double oracle_grid_min_pe(const mspms_sf_grid* g, int itype)
{
    size_t k, n = (size_t)g->nx * g->ny * g->nz;
    double minpe = 10e10; /* grid.f90:61, :102 */
    for (k = 0; k < n; k++) {
        double v = (double)g->nodes[((size_t)itype * 8) * n + k];
        if (v < minpe) minpe = v;
    }
    return minpe;
}

int oracle_hermite_site(const mspms_sf_grid* g, int itype, double min_pe,
                        double x, double y, double z, double* pe,
                        double* force3)
{
    double xp = x - g->x0, yp = y - g->y0, zp = z - g->z0;
    double pz = g->zperiod > 0.0 ? g->zperiod : (g->nz - 1) * g->hz;
    double tx, ty, tz, bx[4], by[4], bz[4], dbx[4], dby[4], dbz[4];
    double fval = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    int ix, iy, iz, cx, cy, cz;
    zp = zp - floor(zp / pz) * pz; /* pbc.f90:82-86, nanotube case */
    ix = (int)floor(xp / g->hx);
    iy = (int)floor(yp / g->hy);
    iz = (int)floor(zp / g->hz);
    if (ix < 0 || iy < 0 || iz < 0 || ix > g->nx - 2 || iy > g->ny - 2 ||
        iz > g->nz - 2)
        return MSPMS_EDOM;
    tx = (xp - ix * g->hx) / g->hx;
    ty = (yp - iy * g->hy) / g->hy;
    tz = (zp - iz * g->hz) / g->hz;
    hermite_1d(tx, g->hx, bx, dbx);
    hermite_1d(ty, g->hy, by, dby);
    hermite_1d(tz, g->hz, bz, dbz);
    /* field index for derivative orders (dx,dy,dz) in {0,1}^3 */
    for (cz = 0; cz < 2; cz++)
        for (cy = 0; cy < 2; cy++)
            for (cx = 0; cx < 2; cx++) {
                int o;
                for (o = 0; o < 8; o++) {
                    int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
                    static const int field_of[8] = {0, 1, 2, 4, 3, 5, 6, 7};
                    double v = grid_node(g, itype, field_of[o], ix + cx, iy + cy,
                                         iz + cz);
                    double wx = bx[cx + 2 * ox], wy = by[cy + 2 * oy],
                           wz = bz[cz + 2 * oz];
                    fval += v * wx * wy * wz;
                    gx += v * dbx[cx + 2 * ox] * wy * wz;
                    gy += v * wx * dby[cy + 2 * oy] * wz;
                    gz += v * wx * wy * dbz[cz + 2 * oz];
                }
            }
    force3[0] = -gx;
    force3[1] = -gy;
    force3[2] = -gz;
    *pe = (fval < 1.5 * min_pe) ? -fval : fval;
    return MSPMS_OK;
}

/* get_values_from_grid (src/sffrc.c:951-1089): the position in Angstrom
 * (reduced times sigma_base), the cube of its lower corner, the 32 monomials
 * of Amatrix_ele_assign_value (:750-803) and four left-to-right dot
 * products; J/mol and J/mol/Angstrom out.  Returns EDOM for a cube index
 * outside the tables (the reference reads past them). */
static int oracle_myinterp_site(const mspms_sf_cubes* q, int type, double fxx, double fyy,
                                double fzz, double* usf, double fsf[3])
{
    double tx, ty, tz, x2, x3, y2, y3, z2, z3, v[32];
    const double* const* tb = q->tables + 128 * type;
    int iAx, iAy, iAz, cube, f, m;
    tx = (fxx - q->xmin) / q->hx;
    ty = (fyy - q->ymin) / q->hy;
    fzz = fzz - q->zmax * rint(fzz / q->zmax);
    if (fzz < 0) fzz = q->zmax + fzz;
    tz = (fzz - 0.0) / q->hz;
    if (!(tx > -1.0 && tx < q->ncx && ty > -1.0 && ty < q->ncy && tz > -1.0 && tz < q->ncz))
        return MSPMS_EDOM;
    iAx = (int)tx;
    iAy = (int)ty;
    iAz = (int)tz;
    cube = iAx * q->ncy * q->ncz + iAy * q->ncz + iAz;
    x2 = fxx * fxx; x3 = x2 * fxx;
    y2 = fyy * fyy; y3 = y2 * fyy;
    z2 = fzz * fzz; z3 = z2 * fzz;
    v[0] = 1.0; v[1] = fxx; v[2] = fyy; v[3] = fzz; v[4] = x2; v[5] = fxx * fyy;
    v[6] = fxx * fzz; v[7] = y2; v[8] = fyy * fzz; v[9] = z2; v[10] = x3; v[11] = x2 * fyy;
    v[12] = fxx * y2; v[13] = x2 * fzz; v[14] = fxx * z2; v[15] = y3; v[16] = y2 * fzz;
    v[17] = fyy * z2; v[18] = z3; v[19] = fxx * fyy * fzz; v[20] = x3 * fyy; v[21] = x2 * y2;
    v[22] = fxx * y3; v[23] = x3 * fzz; v[24] = x2 * z2; v[25] = fxx * z3; v[26] = y3 * fzz;
    v[27] = y2 * z2; v[28] = fyy * z3; v[29] = x2 * fyy * fzz; v[30] = fxx * y2 * fzz;
    v[31] = fxx * fyy * z2;
    for (f = 0; f < 4; f++) {
        double a = tb[32 * f][cube] * v[0];
        for (m = 1; m < 32; m++) a = a + tb[32 * f + m][cube] * v[m];
        if (f == 0)
            *usf = a;
        else
            fsf[f - 1] = a;
    }
    return MSPMS_OK;
}

/* ------------------------------------------------------------------------- */
int oracle_validate(int nspecie, const mspms_species* sp, const mspms_params* p)
{
    int s, k;
    if (nspecie < 1) return MSPMS_EINVAL;
    if (p->iExternal_FF_type != 0) return MSPMS_EINVAL;
    if (p->iChargeType < 0 || p->iChargeType > 3) return MSPMS_EINVAL;
    if (p->iSF_type < 0 || p->iSF_type > 4) return MSPMS_EINVAL;
    if (p->iSF_type == MSPMS_SF_TASOS && !p->grid) return MSPMS_EINVAL;
    if (p->iSF_type == MSPMS_SF_MY_INTERP &&
        (!p->cubes || !p->cubes->tables || p->cubes->ntype < 1))
        return MSPMS_EINVAL;
    /* cal_sf_hypergeo leaves *uij unset when ntube == 0 */
    if (p->iSF_type == MSPMS_SF_HYPERGEO && p->ntube < 1) return MSPMS_EINVAL;
    for (s = 0; s < nspecie; s++) {
        const mspms_species* q = sp + s;
        if (q->natom < 1 || q->natom > 256 || q->nmole < 0) return MSPMS_EINVAL;
        if (q->nnbp != 0) return MSPMS_EINVAL;
        for (k = 0; k < q->nbond; k++)
            if (q->bnd_type[k] < 1 || q->bnd_type[k] > 3) return MSPMS_EINVAL;
        for (k = 0; k < q->nangle; k++)
            if (q->agl_type[k] < 1 || q->agl_type[k] > 2) return MSPMS_EINVAL;
        for (k = 0; k < q->ndih; k++) {
            if (q->dih_type[k] < 1 || q->dih_type[k] > 2) return MSPMS_EINVAL;
            if (q->dih_type[k] == MSPMS_DIH_CHARMM &&
                !(fabs(q->dih_para[4 * k]) >= TOL_REF))
                return MSPMS_EINVAL;
        }
        for (k = 0; k < q->nimp; k++) {
            if (q->imp_type[k] != MSPMS_IMP_CHARMM) return MSPMS_EINVAL;
            if (!(fabs(q->imp_para[2 * k]) >= 1e-5)) return MSPMS_EINVAL;
        }
        /* the 1-4 section reads isSameMole left by the bond/angle sections */
        if (q->ndih > 0 && q->nbond == 0 && q->nangle == 0) return MSPMS_EINVAL;
    }
    return MSPMS_OK;
}

static void uniques_or_given(const mspms_species* q, uint8_t* au, uint8_t* du)
{
    if (q->agl_unique && q->dih_unique) {
        memcpy(au, q->agl_unique, (size_t)q->nangle);
        memcpy(du, q->dih_unique, (size_t)q->ndih);
    } else {
        oracle_check_uniques(q, au, du);
        if (q->agl_unique) memcpy(au, q->agl_unique, (size_t)q->nangle);
        if (q->dih_unique) memcpy(du, q->dih_unique, (size_t)q->ndih);
    }
}

/* frclong, src/frclong.c:20-1233 */
int oracle_frclong(int nspecie, const mspms_species* sp, const mspms_params* p,
                   const double* xx, const double* yy, const double* zz,
                   double* fxl, double* fyl, double* fzl, mspms_long_terms* t)
{
    lr_state s;
    int natom = 0, sI, ii, rc;
    int carried_same = 1;
    rc = oracle_validate(nspecie, sp, p);
    if (rc) return rc;
    for (sI = 0; sI < nspecie; sI++) natom += sp[sI].natom * sp[sI].nmole;
    memset(&s, 0, sizeof(s));
    s.p = p;
    s.xx = xx;
    s.yy = yy;
    s.zz = zz;
    s.fxl = fxl;
    s.fyl = fyl;
    s.fzl = fzl;
    memset(t, 0, sizeof(*t));
    for (ii = 0; ii < natom; ii++) fxl[ii] = fyl[ii] = fzl[ii] = 0.0;

    inter_molecular(&s, nspecie, sp);

    /* intramolecular corrections, src/frclong.c:222-982 */
    {
        int base = 0;
        for (sI = 0; sI < nspecie; sI++) {
            const mspms_species* q = sp + sI;
            uint8_t au[256], du[256];
            int m, k;
            uniques_or_given(q, au, du);
            for (m = 0; m < q->nmole; m++) {
                for (k = 0; k < q->nbond; k++)
                    carried_same = intra_12_13(&s, q, base, q->bnd_idx[2 * k],
                                               q->bnd_idx[2 * k + 1]);
                for (k = 0; k < q->nangle; k++) {
                    if (!au[k]) continue;
                    carried_same = intra_12_13(&s, q, base, q->agl_idx[3 * k],
                                               q->agl_idx[3 * k + 2]);
                }
                for (k = 0; k < q->ndih; k++) {
                    if (!du[k]) continue;
                    intra_14(&s, q, base, q->dih_idx[4 * k], q->dih_idx[4 * k + 3],
                             carried_same, p->f0);
                }
                base += q->natom;
            }
        }
    }
    s.ushift *= p->shift4; /* :985 */

    t->uvdw = s.uvdw;
    t->ushift = s.ushift;
    t->ureal = s.ureal;
    t->uexcl = s.uexcl;
    t->uelec = 0.0;
    t->ufourier = 0.0;
    t->uself = 0.0;
    t->uvacuum = 0.0;

    if (p->iChargeType == MSPMS_ELEC_EWALD) {
        /* reciprocal sum, src/frclong.c:988-1048 */
        int kx, ky, kz;
        double uf = 0.0, us = 0.0;
        double* q = (double*)malloc(sizeof(double) * (size_t)(natom > 0 ? natom : 1));
        for (ii = 0; ii < natom; ii++) {
            int spc, site;
            site_of(nspecie, sp, ii, &spc, &site);
            q[ii] = sp[spc].charge[site];
        }
        for (kx = -p->kmaxx; kx <= p->kmaxx; kx++)
            for (ky = -p->kmaxy; ky <= p->kmaxy; ky++)
                for (kz = -p->kmaxz; kz <= p->kmaxz; kz++) {
                    int ksq = kx * kx + ky * ky + kz * kz;
                    double rkx, rky, rkz, rksq, kvec, sr = 0.0, si = 0.0;
                    if (!(ksq <= p->ksqmax && ksq != 0)) continue;
                    rkx = p->twopi_lx * kx;
                    rky = p->twopi_ly * ky;
                    rkz = p->twopi_lz * kz;
                    rksq = rkx * rkx + rky * rky + rkz * rkz;
                    kvec = exp(-p->bfactor_ewald * rksq) / rksq;
                    for (ii = 0; ii < natom; ii++) {
                        double tt = rkx * xx[ii] + rky * yy[ii] + rkz * zz[ii];
                        sr = sr + q[ii] * cos(tt);
                        si = si + q[ii] * sin(tt);
                    }
                    uf += kvec * (sr * sr + si * si);
                    for (ii = 0; ii < natom; ii++) {
                        double tt = rkx * xx[ii] + rky * yy[ii] + rkz * zz[ii];
                        double fij = 2.0 * q[ii] * (sr * sin(tt) - si * cos(tt)) *
                                     p->vfactor_ewald * kvec * p->coulomb_prefactor;
                        fxl[ii] += fij * rkx;
                        fyl[ii] += fij * rky;
                        fzl[ii] += fij * rkz;
                    }
                }
        uf *= p->vfactor_ewald * p->coulomb_prefactor;
        for (ii = 0; ii < natom; ii++) us += q[ii] * q[ii];
        us *= p->coulomb_prefactor * sqrt(p->kappa * p->kappa / PI_REF);
        t->ufourier = uf;
        t->uself = us;
        t->uelec = s.ureal + uf - us - s.uexcl;
        if (p->fEwald_BC == MSPMS_EWALD_BC_VACUUM) {
            /* src/frclong.c:1065-1126 */
            double qrx = 0.0, qry = 0.0, qrz = 0.0, uv;
            int base = 0, m, a;
            for (sI = 0; sI < nspecie; sI++) {
                const mspms_species* qs = sp + sI;
                double mw = species_mw(qs);
                for (m = 0; m < qs->nmole; m++) {
                    double com[3], ex[256], fy[256], gz[256];
                    com_inner(qs, xx, yy, zz, base, mw, com, ex, fy, gz);
                    com[0] -= p->boxlx * rint(com[0] / p->boxlx);
                    com[1] -= p->boxly * rint(com[1] / p->boxly);
                    com[2] -= p->boxlz * rint(com[2] / p->boxlz);
                    for (a = 0; a < qs->natom; a++) {
                        ex[a] = ex[a] + com[0];
                        fy[a] = fy[a] + com[1];
                        gz[a] = gz[a] + com[2];
                    }
                    /* :1095-1109 sums q * reconstructed position in atom
                     * order, which is this molecule-by-molecule order */
                    for (a = 0; a < qs->natom; a++) {
                        double qa = qs->charge[a];
                        qrx = qrx + qa * ex[a];
                        qry = qry + qa * fy[a];
                        qrz = qrz + qa * gz[a];
                    }
                    base += qs->natom;
                }
            }
            uv = qrx * qrx + qry * qry + qrz * qrz;
            uv *= p->twopi_over_3v * p->coulomb_prefactor;
            for (ii = 0; ii < natom; ii++) {
                double fij = -2.0 * p->twopi_over_3v * q[ii] * p->coulomb_prefactor;
                fxl[ii] += fij * qrx;
                fyl[ii] += fij * qry;
                fzl[ii] += fij * qrz;
            }
            t->uvacuum = uv;
            t->uelec += uv;
        }
        s.virial += t->uelec;
        free(q);
    } else if (p->iChargeType == MSPMS_ELEC_WOLF) {
        double us = 0.0;
        for (ii = 0; ii < natom; ii++) {
            int spc, site;
            double qi;
            site_of(nspecie, sp, ii, &spc, &site);
            qi = sp[spc].charge[site];
            us += qi * qi;
        }
        us *= p->coulomb_prefactor * p->wolfvcon3;
        t->uself = us;
        t->uelec = s.ureal - us - s.uexcl;
    } else if (p->iChargeType == MSPMS_ELEC_SIMPLE_COULOMB) {
        t->uelec = s.ucoul;
    }

    /* solid-fluid, src/frclong.c:1145-1209 */
    t->usflj = 0.0;
    if (p->iSF_type != MSPMS_SF_NONE) {
        double minpe[64];
        if (p->iSF_type == MSPMS_SF_TASOS) {
            int ty;
            if (p->grid->ntype > 64) return MSPMS_EINVAL;
            for (ty = 0; ty < p->grid->ntype; ty++)
                minpe[ty] = oracle_grid_min_pe(p->grid, ty);
        }
        for (ii = 0; ii < natom; ii++) {
            int spc, site;
            double uij = 0.0;
            site_of(nspecie, sp, ii, &spc, &site);
            if (sp[spc].ghost_type[site] != MSPMS_GHOST_NONE) continue;
            if (p->iSF_type == MSPMS_SF_HYPERGEO) {
                sf_hypergeo(p, sp + spc, site, xx[ii], yy[ii], &fxl[ii], &fyl[ii], &uij);
                t->usflj += uij;
            } else if (p->iSF_type == MSPMS_SF_ATOM_EXPLICIT) {
                sf_atom_explicit(p, sp + spc, site, xx[ii], yy[ii], zz[ii],
                                 &fxl[ii], &fyl[ii], &fzl[ii], &uij);
                t->usflj += uij;
            } else if (p->iSF_type == MSPMS_SF_TASOS) {
                /* call_tasos_forces, src/sffrc.c:1322-1337 */
                double f3[3], pe;
                int ty = sp[spc].interp_type[site] - 1;
                int r = (ty < 0 || ty >= p->grid->ntype)
                            ? MSPMS_EINVAL
                            : oracle_hermite_site(p->grid, ty, minpe[ty],
                                            xx[ii] * p->sigma_base,
                                            yy[ii] * p->sigma_base,
                                            zz[ii] * p->sigma_base, &pe, f3);
                if (r) return r;
                pe /= p->epsilon_base;
                f3[0] *= (p->sigma_base / p->epsilon_base);
                f3[1] *= (p->sigma_base / p->epsilon_base);
                f3[2] *= (p->sigma_base / p->epsilon_base);
                t->usflj += pe;
                fxl[ii] += f3[0];
                fyl[ii] += f3[1];
                fzl[ii] += f3[2];
            } else if (p->iSF_type == MSPMS_SF_MY_INTERP) {
                /* src/frclong.c:1193-1207; R_RGAS src/defs.h:96 */
                const double R_RGAS = 0.120272219;
                double f3[3] = {0.0, 0.0, 0.0}, ue = 0.0;
                int ty = sp[spc].interp_type[site] - 1;
                int r = (ty < 0 || ty >= p->cubes->ntype)
                            ? MSPMS_EDOM
                            : oracle_myinterp_site(p->cubes, ty, xx[ii] * p->sigma_base,
                                                   yy[ii] * p->sigma_base,
                                                   zz[ii] * p->sigma_base, &ue, f3);
                if (r) return r;
                t->usflj += ue * R_RGAS / p->epsilon_base;
                fxl[ii] += f3[0] * R_RGAS * p->sigma_base / p->epsilon_base;
                fyl[ii] += f3[1] * R_RGAS * p->sigma_base / p->epsilon_base;
                fzl[ii] += f3[2] * R_RGAS * p->sigma_base / p->epsilon_base;
            }
        }
    }
    t->uotherff = 0.0;
    t->uinter = t->uvdw + t->uelec + t->usflj + t->uotherff; /* :1223 */
    t->virial_inter = s.virial;
    t->n_lj_pairs = s.nlj;
    t->n_elec_pairs = s.nel;
    t->lj_pair_hash = s.hlj;
    t->elec_pair_hash = s.hel;
    return MSPMS_OK;
}

/* ------------------------------------------------------------------------- */
/* bonded terms                                                               */

/* bndfrc (src/bndfrc.c:17-78) */
static void bond_term(const mspms_species* q, int k, int base, const double* xx,
                      const double* yy, const double* zz, double* fxs,
                      double* fys, double* fzs, double* u, double* vir)
{
    int i1 = base + q->bnd_idx[2 * k], i2 = base + q->bnd_idx[2 * k + 1];
    const double* pr = q->bnd_para + 3 * k; /* Kb, Req, alpha */
    double rx = xx[i1] - xx[i2], ry = yy[i1] - yy[i2], rz = zz[i1] - zz[i2];
    double rsq = rx * rx + ry * ry + rz * rz, r = sqrt(rsq), fij = 0.0;
    switch (q->bnd_type[k]) {
    case MSPMS_BOND_HARMONIC: {
        double dr = r - pr[1];
        *u = 0.5 * pr[0] * dr * dr;
        fij = -pr[0] * dr / r;
        break;
    }
    case MSPMS_BOND_MORSE: {
        double dr = r - pr[1], De = pr[0], e = exp(-pr[2] * dr), om = 1.0 - e;
        *u = De * om * om;
        fij = -2.0 * De * pr[2] * om * e / r;
        break;
    }
    case MSPMS_BOND_FENE: {
        double rmax2 = pr[1] * pr[1];
        *u = pr[0] * rmax2 * log(1.0 - rsq / rmax2);
        fij = -2.0 * pr[0] * rmax2 / (rmax2 - rsq);
        break;
    }
    }
    *vir = fij * rsq;
    fxs[i1] += fij * rx;
    fys[i1] += fij * ry;
    fzs[i1] += fij * rz;
    fxs[i2] -= fij * rx;
    fys[i2] -= fij * ry;
    fzs[i2] -= fij * rz;
}

/* aglfrc (src/aglfrc.c:17-190) */
static void angle_term(const mspms_species* q, int k, int base, const double* xx,
                       const double* yy, const double* zz, double* fxs,
                       double* fys, double* fzs, double* u, double* vir)
{
    int ia = base + q->agl_idx[3 * k], ib = base + q->agl_idx[3 * k + 1],
        ic = base + q->agl_idx[3 * k + 2];
    const double* pr = q->agl_para + 5 * k;
    if (q->agl_type[k] == MSPMS_ANGLE_HARMONIC) {
        double xab = xx[ia] - xx[ib], yab = yy[ia] - yy[ib], zab = zz[ia] - zz[ib];
        double xbc = xx[ic] - xx[ib], ybc = yy[ic] - yy[ib], zbc = zz[ic] - zz[ib];
        double rab = sqrt(xab * xab + yab * yab + zab * zab), rrab = 1.0 / rab;
        double rbc, rrbc, cost, theta, sint, dth, gam;
        double fxa, fya, fza, fxc, fyc, fzc;
        xab = xab * rrab;
        yab = yab * rrab;
        zab = zab * rrab;
        rbc = sqrt(xbc * xbc + ybc * ybc + zbc * zbc);
        rrbc = 1.0 / rbc;
        xbc = xbc * rrbc;
        ybc = ybc * rrbc;
        zbc = zbc * rrbc;
        cost = xab * xbc + yab * ybc + zab * zbc;
        theta = acos(cost);
        sint = fmax(1.0e-8, sqrt(1.0 - cost * cost));
        dth = theta - pr[1];
        *u = 0.5 * pr[0] * dth * dth;
        gam = pr[0] * dth / sint;
        fxa = gam * (xbc - xab * cost) * rrab;
        fya = gam * (ybc - yab * cost) * rrab;
        fza = gam * (zbc - zab * cost) * rrab;
        fxc = gam * (xab - xbc * cost) * rrbc;
        fyc = gam * (yab - ybc * cost) * rrbc;
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
    } else { /* ANGLE_TR_WATER, src/aglfrc.c:107-185 */
        double v13x = xx[ia] - xx[ic], v13y = yy[ia] - yy[ic], v13z = zz[ia] - zz[ic];
        double v12x = xx[ia] - xx[ib], v12y = yy[ia] - yy[ib], v12z = zz[ia] - zz[ib];
        double v23x = xx[ib] - xx[ic], v23y = yy[ib] - yy[ic], v23z = zz[ib] - zz[ic];
        double r13 = sqrt(v13x * v13x + v13y * v13y + v13z * v13z);
        double r12 = sqrt(v12x * v12x + v12y * v12y + v12z * v12z);
        double r23 = sqrt(v23x * v23x + v23y * v23y + v23z * v23z);
        double d3 = r13 - pr[3], d1 = r12 - pr[4], d2 = r23 - pr[4];
        double kt = pr[0], krt = pr[1], krr = pr[2];
        double t1, t21, t22, t3, t31, t32;
        double fxa, fya, fza, fxb, fyb, fzb, fxc, fyc, fzc;
        *u = 0.5 * kt * d3 * d3 + krt * d3 * (d1 + d2) + krr * d1 * d2;
        t1 = -kt * d3 / r13;
        t21 = -krt * d3 / r12;
        t22 = -krt * (d1 + d2) / r13;
        t3 = -krr * d2 / r12;
        fxa = (t1 * v13x + t21 * v12x + t22 * v13x + t3 * v12x);
        fya = (t1 * v13y + t21 * v12y + t22 * v13y + t3 * v12y);
        fza = (t1 * v13z + t21 * v12z + t22 * v13z + t3 * v12z);
        fxs[ia] += fxa;
        fys[ia] += fya;
        fzs[ia] += fza;
        t21 = -krt * d3 / r23;
        t22 = krt * d3 / r12;
        t31 = -krr * d1 / r23;
        t32 = krr * d2 / r12;
        fxb = (t21 * v23x + t22 * v12x + t31 * v23x + t32 * v12x);
        fyb = (t21 * v23y + t22 * v12y + t31 * v23y + t32 * v12y);
        fzb = (t21 * v23z + t22 * v12z + t31 * v23z + t32 * v12z);
        fxs[ib] += fxb;
        fys[ib] += fyb;
        fzs[ib] += fzb;
        t1 = kt * d3 / r13;
        t21 = krt * d3 / r23;
        t22 = krt * (d1 + d2) / r13;
        t3 = krr * d1 / r23;
        fxc = (t1 * v13x + t21 * v23x + t22 * v13x + t3 * v23x);
        fyc = (t1 * v13y + t21 * v23y + t22 * v13y + t3 * v23y);
        fzc = (t1 * v13z + t21 * v23z + t22 * v13z + t3 * v23z);
        fxs[ic] += fxc;
        fys[ic] += fyc;
        fzs[ic] += fzc;
        *vir = fxa * v12x + fya * v12y + fza * v12z;
        *vir = *vir - fxc * v23x - fyc * v23y - fzc * v23z;
    }
}

/* CHARMM dihedral / improper force core shared by src/dihfrc.c:366-469 and
 * src/impfrc.c:55-181: the same four-atom gradient with a prefactor `temp`. */
static void charmm_forces(int i1, int i2, int i3, int i4, const double* xx,
                          const double* yy, const double* zz, double* fxs,
                          double* fys, double* fzs, double temp, double aa1,
                          double bb1, double cc1, double aa2, double bb2,
                          double cc2, double t1, double t2, double t3,
                          double rxjk, double ryjk, double rzjk, double rxij,
                          double ryij, double rzij, double rxjl, double ryjl,
                          double rzjl)
{
    double fij, rxik, ryik, rzik, rxkl, rykl, rzkl;
    fij = temp / (t2 * t1 * t1);
    fxs[i1] = fxs[i1] + fij * ((bb2 * rzjk - cc2 * ryjk) * t1 - t3 / t1 * (bb1 * rzjk - cc1 * ryjk));
    fys[i1] = fys[i1] + fij * ((-aa2 * rzjk + cc2 * rxjk) * t1 - t3 / t1 * (-aa1 * rzjk + cc1 * rxjk));
    fzs[i1] = fzs[i1] + fij * ((aa2 * ryjk - bb2 * rxjk) * t1 - t3 / t1 * (aa1 * ryjk - bb1 * rxjk));
    fij = temp / (t1 * t1 * t2 * t2);
    rxik = xx[i1] - xx[i3];
    ryik = yy[i1] - yy[i3];
    rzik = zz[i1] - zz[i3];
    rxkl = xx[i4] - xx[i3];
    rykl = yy[i4] - yy[i3];
    rzkl = zz[i4] - zz[i3];
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

/* plane normals of CHARMM dihedrals/impropers (src/dihfrc.c:377-402) */
static void charmm_geom(int i1, int i2, int i3, int i4, const double* xx,
                        const double* yy, const double* zz, double g[18])
{
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

/* dihfrc (src/dihfrc.c:17-284) */
static void dih_term(const mspms_species* q, int k, int base, const double* xx,
                     const double* yy, const double* zz, double* fxs,
                     double* fys, double* fzs, double* u, double* vir)
{
    int ia = base + q->dih_idx[4 * k], ib = base + q->dih_idx[4 * k + 1],
        ic = base + q->dih_idx[4 * k + 2], id = base + q->dih_idx[4 * k + 3];
    const double* c = q->dih_para + 4 * k;
    if (q->dih_type[k] == MSPMS_DIH_OPLS_COSIN) {
        double xab = xx[ia] - xx[ib], yab = yy[ia] - yy[ib], zab = zz[ia] - zz[ib];
        double xbc = xx[ib] - xx[ic], ybc = yy[ib] - yy[ic], zbc = zz[ib] - zz[ic];
        double xcd = xx[ic] - xx[id], ycd = yy[ic] - yy[id], zcd = zz[ic] - zz[id];
        double rrbc = 1.0 / sqrt(xbc * xbc + ybc * ybc + zbc * zbc);
        double pbx = yab * zbc - zab * ybc, pby = zab * xbc - xab * zbc,
               pbz = xab * ybc - yab * xbc;
        double pb2 = pbx * pbx + pby * pby + pbz * pbz, rpb1 = 1.0 / sqrt(pb2),
               rpb2 = rpb1 * rpb1;
        double pcx = ybc * zcd - zbc * ycd, pcy = zbc * xcd - xbc * zcd,
               pcz = xbc * ycd - ybc * xcd;
        double pc2 = pcx * pcx + pcy * pcy + pcz * pcz, rpc1 = 1.0 / sqrt(pc2),
               rpc2 = rpc1 * rpc1;
        double pbpc = pbx * pcx + pby * pcy + pbz * pcz;
        double cost = pbpc * rpb1 * rpc1;
        double sint = (xbc * (pcy * pbz - pcz * pby) + ybc * (pbx * pcz - pbz * pcx) +
                       zbc * (pcx * pby - pcy * pbx)) *
                      (rpb1 * rpc1 * rrbc);
        double theta = atan2(sint, cost), rsint, vopls, dopls, gam;
        double fax, fay, faz, fcx, fcy, fcz, fbx, fby, fbz, fdx, fdy, fdz;
        sint = copysign(fmax(1.0e-8, fabs(sint)), sint);
        rsint = 1.0 / sint;
        vopls = c[0] + 0.5 * (c[1] * (1.0 + cos(theta)) + c[2] * (1.0 - cos(2.0 * theta)) +
                              c[3] * (1.0 + cos(3.0 * theta)));
        *u = vopls;
        *vir = 0.0;
        dopls = -0.5 * (c[1] * sin(theta) - 2.0 * c[2] * sin(2.0 * theta) +
                        3.0 * c[3] * sin(3.0 * theta));
        gam = dopls * rpb1 * rpc1 * rsint;
        fax = gam * ((-pcy * zbc + pcz * ybc) - pbpc * rpb2 * (-pby * zbc + pbz * ybc));
        fay = gam * ((pcx * zbc - pcz * xbc) - pbpc * rpb2 * (pbx * zbc - pbz * xbc));
        faz = gam * ((-pcx * ybc + pcy * xbc) - pbpc * rpb2 * (-pbx * ybc + pby * xbc));
        fcx = gam * ((-pcy * zab + pcz * yab) - pbpc * rpb2 * (-pby * zab + pbz * yab));
        fcy = gam * ((pcx * zab - pcz * xab) - pbpc * rpb2 * (pbx * zab - pbz * xab));
        fcz = gam * ((-pcx * yab + pcy * xab) - pbpc * rpb2 * (-pbx * yab + pby * xab));
        fbx = gam * ((-pby * zcd + pbz * ycd) - pbpc * rpc2 * (-pcy * zcd + pcz * ycd));
        fby = gam * ((pbx * zcd - pbz * xcd) - pbpc * rpc2 * (pcx * zcd - pcz * xcd));
        fbz = gam * ((-pbx * ycd + pby * xcd) - pbpc * rpc2 * (-pcx * ycd + pcy * xcd));
        fdx = gam * ((-pby * zbc + pbz * ybc) - pbpc * rpc2 * (-pcy * zbc + pcz * ybc));
        fdy = gam * ((pbx * zbc - pbz * xbc) - pbpc * rpc2 * (pcx * zbc - pcz * xbc));
        fdz = gam * ((-pbx * ybc + pby * xbc) - pbpc * rpc2 * (-pcx * ybc + pcy * xbc));
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
    } else { /* DIH_CHARMM, src/dihfrc.c:358-470 */
        double kphi = c[0], nper = c[1], d0 = c[2], g[18], phi, alpha, temp;
        charmm_geom(ia, ib, ic, id, xx, yy, zz, g);
        phi = g[8] / (g[6] * g[7]);
        alpha = phi;
        phi = 180.0 - (acos(phi) * 180.0 / PI_REF);
        *u = kphi * (1 + cos((nper * phi - d0) / 180 * PI_REF));
        *vir = 0.0;
        if (fabs(sin(phi * PI_REF / 180)) <= 1e-10)
            temp = -nper * nper * kphi * cos((nper * phi - d0) * PI_REF / 180) / alpha;
        else
            temp = nper * kphi * sin((nper * phi - d0) * PI_REF / 180) / sin(phi * PI_REF / 180);
        charmm_forces(ia, ib, ic, id, xx, yy, zz, fxs, fys, fzs, temp, g[0], g[1],
                      g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], g[11],
                      g[12], g[13], g[14], g[15], g[16], g[17]);
    }
}

/* impfrc (src/impfrc.c:17-191).  The reference exit(1)s on its two
 * singular geometries; those are reported here as MSPMS_EINVAL. */
static int imp_term(const mspms_species* q, int k, int base, const double* xx,
                    const double* yy, const double* zz, double* fxs,
                    double* fys, double* fzs, double* u, double* vir)
{
    int ia = base + q->imp_idx[4 * k], ib = base + q->imp_idx[4 * k + 1],
        ic = base + q->imp_idx[4 * k + 2], id = base + q->imp_idx[4 * k + 3];
    double kom = q->imp_para[2 * k], om0 = q->imp_para[2 * k + 1];
    double g[18], phi, alpha, temp, ang;
    charmm_geom(ia, ib, ic, id, xx, yy, zz, g);
    phi = g[8] / (g[6] * g[7]);
    alpha = phi;
    phi = 180 - (acos(phi) * 180 / PI_REF);
    *u = kom * (phi - om0) * (phi - om0);
    *u *= (PI_REF / 180) * (PI_REF / 180);
    *vir = 0.0;
    ang = acos(alpha) * 180 / PI_REF;
    if ((fabs(om0) <= 1e-5) && (fabs(ang) <= 1e-20)) return MSPMS_EINVAL;
    if ((fabs(om0 - 180) <= 1e-5) && (fabs(ang - 180) <= 1e-20)) return MSPMS_EINVAL;
    if ((fabs(om0) <= 1e-5) && (fabs(ang - 180) <= 1e-20)) return MSPMS_EINVAL;
    if ((fabs(om0 - 180) <= 1e-5) && (fabs(ang) <= 1e-20)) return MSPMS_EINVAL;
    temp = -2.0 * kom * (phi - om0) * PI_REF / 180 / sin(phi * PI_REF / 180);
    charmm_forces(ia, ib, ic, id, xx, yy, zz, fxs, fys, fzs, temp, g[0], g[1], g[2],
                  g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], g[11], g[12],
                  g[13], g[14], g[15], g[16], g[17]);
    return MSPMS_OK;
}

/* frcshort, src/frcshort.c:20-97 */
int oracle_frcshort(int nspecie, const mspms_species* sp, const mspms_params* p,
                    const double* xx, const double* yy, const double* zz,
                    double* fxs, double* fys, double* fzs, mspms_short_terms* t)
{
    double ub = 0.0, ua = 0.0, ud = 0.0, ui = 0.0, vir = 0.0;
    int natom = 0, s, m, k, base = 0, rc;
    (void)p;
    rc = oracle_validate(nspecie, sp, p);
    if (rc) return rc;
    for (s = 0; s < nspecie; s++) natom += sp[s].natom * sp[s].nmole;
    for (k = 0; k < natom; k++) fxs[k] = fys[k] = fzs[k] = 0.0;
    for (s = 0; s < nspecie; s++) {
        const mspms_species* q = sp + s;
        for (m = 0; m < q->nmole; m++) {
            double u = 0.0, v = 0.0;
            for (k = 0; k < q->nbond; k++) {
                bond_term(q, k, base, xx, yy, zz, fxs, fys, fzs, &u, &v);
                ub += u;
                vir += v;
            }
            for (k = 0; k < q->nangle; k++) {
                angle_term(q, k, base, xx, yy, zz, fxs, fys, fzs, &u, &v);
                ua += u;
                vir += v;
            }
            for (k = 0; k < q->ndih; k++) {
                dih_term(q, k, base, xx, yy, zz, fxs, fys, fzs, &u, &v);
                ud += u;
                vir += v;
            }
            for (k = 0; k < q->nimp; k++) {
                rc = imp_term(q, k, base, xx, yy, zz, fxs, fys, fzs, &u, &v);
                if (rc) return rc;
                ui += u;
                vir += v;
            }
            base += q->natom;
        }
    }
    t->ubond = ub;
    t->uangle = ua;
    t->udih = ud;
    t->uimp = ui;
    t->uintra = ub + ua + ud + ui;
    t->virial_intra = vir;
    return MSPMS_OK;
}

/* ------------------------------------------------------------------------- */
void oracle_check_uniques(const mspms_species* q, uint8_t* au, uint8_t* du)
{
    /* src/initvars.c:13-146.  Writes beyond the current counts land in the
     * reference's unused template slots; they are dropped here. */
    int i, j;
    uint8_t dih[256], agl[256];
    for (i = 0; i < 256; i++) dih[i] = agl[i] = 1;
    for (i = 0; i < q->ndih - 1; i++) {
        if (!dih[i]) continue;
        for (j = i + 1; j < q->ndih; j++) {
            if (!dih[j]) continue;
            if ((q->dih_idx[4 * i] == q->dih_idx[4 * j] &&
                 q->dih_idx[4 * i + 3] == q->dih_idx[4 * j + 3]) ||
                (q->dih_idx[4 * i] == q->dih_idx[4 * j + 3] &&
                 q->dih_idx[4 * i + 3] == q->dih_idx[4 * j]))
                dih[j] = 0;
        }
    }
    for (i = 0; i < q->ndih; i++) {
        if (!dih[i]) continue;
        for (j = 0; j < q->nangle; j++)
            if ((q->dih_idx[4 * i] == q->agl_idx[3 * j] &&
                 q->dih_idx[4 * i + 3] == q->agl_idx[3 * j + 2]) ||
                (q->dih_idx[4 * i] == q->agl_idx[3 * j + 2] &&
                 q->dih_idx[4 * i + 3] == q->agl_idx[3 * j]))
                dih[j] = 0; /* :66 marks by the angle index */
    }
    for (i = 0; i < q->ndih; i++) {
        if (!dih[i]) continue;
        for (j = 0; j < q->nbond; j++)
            if ((q->dih_idx[4 * i] == q->bnd_idx[2 * j] &&
                 q->dih_idx[4 * i + 3] == q->bnd_idx[2 * j + 1]) ||
                (q->dih_idx[4 * i] == q->bnd_idx[2 * j + 1] &&
                 q->dih_idx[4 * i + 3] == q->bnd_idx[2 * j]))
                dih[j] = 0; /* :86 marks by the bond index */
    }
    for (i = 0; i < q->nangle - 1; i++) {
        if (!agl[i]) continue;
        for (j = i + 1; j < q->nangle; j++) {
            if (!agl[j]) continue;
            if ((q->agl_idx[3 * i] == q->agl_idx[3 * j] &&
                 q->agl_idx[3 * i + 2] == q->agl_idx[3 * j + 2]) ||
                (q->agl_idx[3 * i] == q->agl_idx[3 * j + 2] &&
                 q->agl_idx[3 * i + 2] == q->agl_idx[3 * j]))
                agl[j] = 0;
        }
    }
    for (i = 0; i < q->nangle; i++) {
        if (!agl[i]) continue;
        for (j = 0; j < q->nbond; j++)
            if ((q->agl_idx[3 * i] == q->bnd_idx[2 * j] &&
                 q->agl_idx[3 * i + 2] == q->bnd_idx[2 * j + 1]) ||
                (q->agl_idx[3 * i] == q->bnd_idx[2 * j + 1] &&
                 q->agl_idx[3 * i + 2] == q->bnd_idx[2 * j]))
                agl[j] = 0; /* :134 marks by the bond index */
    }
    for (i = 0; i < q->nangle; i++) au[i] = agl[i];
    for (i = 0; i < q->ndih; i++) du[i] = dih[i];
}

/* src/initvars.c:591-599 (cutoffs, shift) and :433-462 (Ewald, Wolf, C) */
int oracle_derive_params(mspms_params* p)
{
    double fmin;
    p->rcutoffsq = p->rcutoff * p->rcutoff;
    p->rcutoffelecsq = p->rcutoffelec * p->rcutoffelec;
    p->rcutonsq = p->rcuton * p->rcuton;
    p->roff2_minus_ron2_cube = (p->rcutoffsq - p->rcutonsq) *
                               (p->rcutoffsq - p->rcutonsq) *
                               (p->rcutoffsq - p->rcutonsq);
    p->shift1 = pow((1 / p->rcutoff), 6.0);
    p->shift4 = 4.0 * p->shift1;
    if (p->iChargeType == MSPMS_ELEC_EWALD) {
        p->bfactor_ewald = 1.0 / (4.0 * p->kappa * p->kappa);
        p->vfactor_ewald = 2.0 * PI_REF / (p->boxlx * p->boxly * p->boxlz);
        p->twopi_lx = 2.0 * PI_REF / p->boxlx;
        p->twopi_ly = 2.0 * PI_REF / p->boxly;
        p->twopi_lz = 2.0 * PI_REF / p->boxlz;
        p->twopi_over_3v = 2.0 * PI_REF / 3.0 / p->boxlx / p->boxly / p->boxlz;
    } else if (p->iChargeType == MSPMS_ELEC_WOLF) {
        double k = p->kappa, rc = p->rcutoffelec;
        p->wolfvcon1 = -erfc(k * rc) / rc;
        p->wolfvcon2 = erfc(k * rc) / p->rcutoffelecsq +
                       2.0 * k * exp(-(k * rc) * (k * rc)) / (sqrt(PI_REF) * rc);
        p->wolfvcon3 = k / sqrt(PI_REF) + erfc(k * rc) / (2.0 * rc);
        p->wolffcon1 = 2.0 * k / sqrt(PI_REF);
        p->wolffcon2 = -p->wolfvcon2;
    }
    if (p->iChargeType != MSPMS_ELEC_NONE)
        p->coulomb_prefactor = 167100.80840 / p->epsilon_base / p->sigma_base;
    fmin = p->boxlx;
    fmin = fmin < p->boxly ? fmin : p->boxly;
    fmin = fmin < p->boxlz ? fmin : p->boxlz;
    fmin /= 2.0;
    if (p->rcuton > fmin || p->rcutoff > fmin || p->rcutoffelec > fmin)
        return MSPMS_ERANGE;
    return MSPMS_OK;
}

/* ------------------------------------------------------------------------- */
/* RANMAR (src/random.c:14-109) */
void oracle_rmarin(oracle_ranmar* r, int ij, int kl)
{
    int i, j, k, l, ii, jj, m;
    double s, t;
    i = (ij / 177) % 177 + 2;
    j = ij % 177 + 2;
    k = (kl / 169) % 178 + 1;
    l = kl % 169;
    for (ii = 1; ii <= 97; ii++) {
        s = 0.0;
        t = 0.5;
        for (jj = 1; jj <= 24; jj++) {
            m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t *= 0.5;
        }
        r->u[ii] = s;
    }
    r->c = 362436.0 / 16777216.0;
    r->cd = 7654321.0 / 16777216.0;
    r->cm = 16777213.0 / 16777216.0;
    r->i97 = 97;
    r->j97 = 33;
}

void oracle_ranmar_next(oracle_ranmar* r, double* rvec, int len)
{
    int v;
    for (v = 0; v < len; v++) {
        double uni = r->u[r->i97] - r->u[r->j97];
        if (uni < 0.0) uni += 1.0;
        r->u[r->i97] = uni;
        if (--r->i97 == 0) r->i97 = 97;
        if (--r->j97 == 0) r->j97 = 97;
        r->c -= r->cd;
        if (r->c < 0.0) r->c += r->cm;
        uni -= r->c;
        if (uni < 0.0) uni += 1.0;
        rvec[v] = uni;
    }
}

/* ------------------------------------------------------------------------- */
/* term-level entry points: the reference's pair terms and bonded terms with  */
/* the globals passed explicitly (checker for the engine's mspms_term_*)      */

void oracle_term_pair(const mspms_params* p, int which, double rijsq, double a,
                      double b, double* u, double* f, double* ushift)
{
    *ushift = 0.0;
    switch (which) {
    case 0: lj_term(p, rijsq, a, b, u, f, ushift); break; /* ljfrc */
    case 1: ewald_real(p, rijsq, a, u, f); break;         /* ewald_real_frc */
    case 2: ewald_excl(p, rijsq, a, u, f); break;         /* ewald_excl_frc */
    case 4:                                               /* xatom_coulomb_frc */
        *u = p->coulomb_prefactor * a / sqrt(rijsq);      /* src/elecfrc.c:164-167 */
        *f = *u / rijsq;
        break;
    default: wolf_real(p, rijsq, a, u, f); break;         /* wolf_real_frc */
    }
}

int oracle_term_bonded(const mspms_species* q, int kind, int k, const double* xx,
                       const double* yy, const double* zz, double* fxs,
                       double* fys, double* fzs, double* u, double* vir)
{
    switch (kind) {
    case 0: bond_term(q, k, 0, xx, yy, zz, fxs, fys, fzs, u, vir); return 0;
    case 1: angle_term(q, k, 0, xx, yy, zz, fxs, fys, fzs, u, vir); return 0;
    case 2: dih_term(q, k, 0, xx, yy, zz, fxs, fys, fzs, u, vir); return 0;
    default: return imp_term(q, k, 0, xx, yy, zz, fxs, fys, fzs, u, vir);
    }
}
