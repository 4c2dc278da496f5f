/*
 * dropin_terms.c -- the reference's term-level functions with their own
 * signatures (src/funcs.h), evaluated by the engine's term code
 * (mspms_term_pair / mspms_term_bonded, the code the GPU kernels run) with
 * the program's globals.  A separate archive member of libmspms_dropin.a:
 * the linker takes it only for a program that leaves these symbols
 * undefined, so a program that keeps the reference's own ljfrc.o,
 * bndfrc.o, ... links unchanged.
 *
 *   ljfrc              src/ljfrc.c:21       ewald_real_frc  src/elecfrc.c:17
 *   ewald_excl_frc     src/elecfrc.c:35     wolf_real_frc   src/elecfrc.c:48
 *   xmole_coulomb_frc  src/elecfrc.c:67     xatom_coulomb_frc src/elecfrc.c:159
 *   bndfrc             src/bndfrc.c:17      aglfrc          src/aglfrc.c:17
 *   dihfrc             src/dihfrc.c:17      impfrc          src/impfrc.c:17
 *   cal_sf_hypergeo    src/sffrc.c:1118     cal_sf_atom_explicit src/sffrc.c:1213
 *   get_values_from_grid src/sffrc.c:951    call_tasos_forces src/sffrc.c:1322
 *
 * Errors print a message and exit(1), as the reference does.
 * call_tasos_forces always stops: its Hermite grid lives in the Fortran
 * library's module variables, which C cannot read (the context API takes
 * the grid as mspms_sf_grid instead).
 */
#include <stdio.h>
#include <stdlib.h>

#include <math.h>
#include <stdbool.h>

#include "../../include/mspms_dropin.h"
#include "../../include/mspms_gpu.h"

#include "ref_layout.h"

#define MAXSP 3 /* NSPECIE_MAX, src/molevars.h:15 */

extern ref_sample_molecule sample_mole[];
extern double xx[], yy[], zz[];
extern double fxs[], fys[], fzs[], fxl[], fyl[], fzl[];
extern double ex[], fy[], gz[];             /* inner coordinates (src/molevars.h:80) */
extern double mole_xx[], mole_yy[], mole_zz[];  /* centres of mass (:87) */
extern double boxlx, boxly, boxlz, rcutoffelecsq, coulomb_prefactor;
extern double uelec, virial_inter;          /* ucoulomb is uelec (src/vars.h:185) */
void mspms_dropin_read_params(mspms_params* p);
void mspms_dropin_species(mspms_species* sp);

static void check(const char* what, int rc)
{
    if (!rc) return;
    fprintf(stderr, "mspms gpu %s failed (%d): %s\n", what, rc, mspms_last_error());
    exit(1);
}

static int pair(int which, double rijsq, double a, double b, double* uij, double* fij,
                double* ushift)
{
    mspms_params p;
    double sh = 0.0;
    mspms_dropin_read_params(&p);
    check("pair term", mspms_term_pair(&p, which, rijsq, a, b, uij, fij, ushift ? ushift : &sh));
    return 0;
}

int ljfrc(double rijsq, double sigmaij, double epsilonij, double* uij, double* fij,
          double* uijshift)
{
    return pair(MSPMS_TERM_LJ, rijsq, sigmaij, epsilonij, uij, fij, uijshift);
}

int ewald_real_frc(double rijsq, double chargeij, double* uij, double* fij)
{
    return pair(MSPMS_TERM_EWALD_REAL, rijsq, chargeij, 0.0, uij, fij, NULL);
}

int ewald_excl_frc(double rijsq, double chargeij, double* uij, double* fij)
{
    return pair(MSPMS_TERM_EWALD_EXCL, rijsq, chargeij, 0.0, uij, fij, NULL);
}

int wolf_real_frc(double rijsq, double chargeij, double* uij, double* fij)
{
    return pair(MSPMS_TERM_WOLF_REAL, rijsq, chargeij, 0.0, uij, fij, NULL);
}

static int bonded(int kind, int iSpecie, int k, int iabs, double* uij, double* virial)
{
    mspms_species sp[MAXSP];
    mspms_dropin_species(sp);
    check("bonded term", mspms_term_bonded(sp + iSpecie, kind, k, xx + iabs, yy + iabs,
                                           zz + iabs, fxs + iabs, fys + iabs, fzs + iabs, uij,
                                           virial));
    return 0;
}

int bndfrc(int iSpecie, int iBond, int iabs, double* uij, double* virial)
{
    return bonded(MSPMS_TERM_BOND, iSpecie, iBond, iabs, uij, virial);
}

int aglfrc(int iSpecie, int iAngle, int iabs, double* uij, double* virial)
{
    return bonded(MSPMS_TERM_ANGLE, iSpecie, iAngle, iabs, uij, virial);
}

int dihfrc(int iSpecie, int iDih, int iabs, double* uij, double* virial)
{
    return bonded(MSPMS_TERM_DIH, iSpecie, iDih, iabs, uij, virial);
}

int impfrc(int iSpecie, int iImp, int iabs, double* uij, double* virial)
{
    return bonded(MSPMS_TERM_IMP, iSpecie, iImp, iabs, uij, virial);
}

int xatom_coulomb_frc(double rijsq, double chargeij, double* uij, double* fij)
{
    return pair(MSPMS_TERM_COULOMB, rijsq, chargeij, 0.0, uij, fij, NULL);
}

/* centre of mass of one molecule into mole_xx[iMole] and its sites' inner
 * coordinates into ex/fy/gz (the state cal_com_and_efg_one leaves,
 * src/misc.c:72-103) */
static void com_and_inner(int iSpecie, int iMole, int iabs, int natom_mol)
{
    const ref_sample_molecule* m = sample_mole + iSpecie;
    int k;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (k = 0; k < natom_mol; k++) {
        cx += xx[iabs + k] * m->aw[k];
        cy += yy[iabs + k] * m->aw[k];
        cz += zz[iabs + k] * m->aw[k];
    }
    mole_xx[iMole] = cx / m->mw;
    mole_yy[iMole] = cy / m->mw;
    mole_zz[iMole] = cz / m->mw;
    for (k = 0; k < natom_mol; k++) {
        ex[iabs + k] = xx[iabs + k] - mole_xx[iMole];
        fy[iabs + k] = yy[iabs + k] - mole_yy[iMole];
        gz[iabs + k] = zz[iabs + k] - mole_zz[iMole];
    }
}

static double mic(double d, double L) { return d - L * rint(d / L); }

/* simple Coulomb between two whole molecules, cut off on the distance of
 * their centres of mass: ucoulomb and virial_inter gain the pair energies,
 * fxl.. the forces; site separations are built from the minimum-image
 * centre separation and the inner coordinates */
int xmole_coulomb_frc(int iSpecie, int iMole, int iabs, int iAtom, int jSpecie, int jMole,
                      int jabs, int jAtom)
{
    const ref_sample_molecule* mi = sample_mole + iSpecie;
    const ref_sample_molecule* mj = sample_mole + jSpecie;
    mspms_params p;
    double rx, ry, rz;
    int a, b;
    com_and_inner(iSpecie, iMole, iabs, iAtom);
    com_and_inner(jSpecie, jMole, jabs, jAtom);
    rx = mic(mole_xx[iMole] - mole_xx[jMole], boxlx);
    ry = mic(mole_yy[iMole] - mole_yy[jMole], boxly);
    rz = mic(mole_zz[iMole] - mole_zz[jMole], boxlz);
    if (!(rx * rx + ry * ry + rz * rz < rcutoffelecsq)) return 0;
    mspms_dropin_read_params(&p);
    for (a = 0; a < iAtom; a++) {
        const int ii = iabs + a;
        if (mi->ghost_type[a] == MSPMS_GHOST_FULL) continue;
        for (b = 0; b < jAtom; b++) {
            const int jj = jabs + b;
            double sx, sy, sz, rsq, u, f, sh;
            if (mj->ghost_type[b] == MSPMS_GHOST_FULL) continue;
            sx = ex[ii] + rx - ex[jj];
            sy = fy[ii] + ry - fy[jj];
            sz = gz[ii] + rz - gz[jj];
            rsq = sx * sx + sy * sy + sz * sz;
            /* u = (C q_i) q_j / r, the product in the reference's order */
            p.coulomb_prefactor = coulomb_prefactor * mi->charge[a];
            check("xmole_coulomb_frc",
                  mspms_term_pair(&p, MSPMS_TERM_COULOMB, rsq, mj->charge[b], 0.0, &u, &f, &sh));
            uelec += u;
            virial_inter += u;
            fxl[ii] += f * sx;
            fyl[ii] += f * sy;
            fzl[ii] += f * sz;
            fxl[jj] -= f * sx;
            fyl[jj] -= f * sy;
            fzl[jj] -= f * sz;
        }
    }
    return 0;
}

static int sf_term(int ii, int iSpecie, int iAtom, double* uij, double* fij)
{
    mspms_params p;
    mspms_species sp[MAXSP];
    double fx = fxl[ii], fy_ = fyl[ii], fz_ = fzl[ii];
    mspms_dropin_read_params(&p);
    mspms_dropin_species(sp);
    check("solid-fluid term", mspms_term_sf(&p, sp + iSpecie, iAtom, xx[ii], yy[ii], zz[ii], &fx,
                                            &fy_, &fz_, uij, fij));
    fxl[ii] = fx;
    fyl[ii] = fy_;
    fzl[ii] = fz_;
    return 0;
}

int cal_sf_hypergeo(int ii, int iSpecie, int iAtom, double* uij, double* fij)
{
    return sf_term(ii, iSpecie, iAtom, uij, fij);
}

int cal_sf_atom_explicit(int ii, int iSpecie, int iAtom, double* uij, double* fij)
{
    return sf_term(ii, iSpecie, iAtom, uij, fij);
}

int get_values_from_grid(double fxx, double fyy, double fzz, int type, double* usf, double* fsf)
{
    mspms_params p;
    mspms_dropin_read_params(&p);
    check("get_values_from_grid", mspms_term_myinterp(&p, fxx, fyy, fzz, type, usf, fsf));
    return 0;
}

/* the grid the drop-in reads from the TASOS section's files (the Fortran
 * library's copy is out of reach); reduced position in, reduced energy and
 * force out, as the reference's wrapper returns them */
int call_tasos_forces(int itype, double fxx, double fyy, double fzz, double* usflj_tasos,
                      double* tasos_force)
{
    mspms_params p;
    mspms_dropin_read_params(&p);
    check("call_tasos_forces",
          mspms_term_hermite(&p, itype, fxx, fyy, fzz, usflj_tasos, tasos_force));
    return 0;
}
