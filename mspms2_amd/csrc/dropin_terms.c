/*
 * dropin_terms.c -- the reference's term-level functions with their own
 * signatures (src/funcs.h), evaluated by the engine's term code
 * (mspms_term_pair / mspms_term_bonded, the code the GPU kernels run) with
 * the program's globals.  A separate archive member of libmspms_dropin.a:
 * the linker takes it only for a program that leaves these symbols
 * undefined, so a program that keeps the reference's own ljfrc.o,
 * bndfrc.o, ... links unchanged.
 *
 *   ljfrc          src/ljfrc.c:21        ewald_real_frc  src/elecfrc.c:17
 *   ewald_excl_frc src/elecfrc.c:35      wolf_real_frc   src/elecfrc.c:48
 *   bndfrc         src/bndfrc.c:17       aglfrc          src/aglfrc.c:17
 *   dihfrc         src/dihfrc.c:17       impfrc          src/impfrc.c:17
 *
 * Errors print a message and exit(1), as the reference does.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/mspms_dropin.h"
#include "../../include/mspms_gpu.h"

#define MAXSP 3 /* NSPECIE_MAX, src/molevars.h:15 */

extern double xx[], yy[], zz[];
extern double fxs[], fys[], fzs[];
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
