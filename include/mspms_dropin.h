/*
 * mspms_dropin.h -- the drop-in symbols of libmspms_dropin.a.
 *
 * Link libmspms_dropin.a + libmspms_gpu.so into the mspms2 program in place
 * of src/frclong.c and src/frcshort.c.  The two functions keep the
 * reference's prototypes (src/funcs.h:18 `int frclong();`, :26
 * `int frcshort();`) and its global-state contract: they read the
 * positions, topology (sample_mole, nspecie, nmole_per_specie) and run
 * constants from the program's globals (src/molevars.h, src/vars.h) and
 * overwrite fxl/fyl/fzl (frclong) or fxs/fys/fzs (frcshort) and the energy
 * / virial globals, as the reference does.
 *
 * Per call: positions go host->device (FP64), the GPU evaluates, forces and
 * energies come back.  The device context is created on the first call and
 * rebuilt when the topology changes; run constants are re-read every call
 * (NPT / volume moves change the box).
 *
 * Environment: MSPMS_GPU_PRECISION=fp64|mixed (default mixed),
 *              MSPMS_GPU_DEVICE=<index> (default 0).
 * Errors: the reference's fail-stop behaviour -- a message on stderr and
 * exit(1) -- replaces its unconditional return 0.
 */
#ifndef MSPMS_DROPIN_H_
#define MSPMS_DROPIN_H_

#ifdef __cplusplus
extern "C" {
#endif

int frclong(void);   /* replaces src/frclong.c:20 */
int frcshort(void);  /* replaces src/frcshort.c:20 */

/* Term-level functions with the reference's signatures (src/funcs.h:19-30),
 * evaluated by the engine's term code on the host with the program's
 * globals (archive member dropin_terms.o: linked only where the program
 * leaves them undefined, so the reference's own term objects still win). */
int ljfrc(double rijsq, double sigmaij, double epsilonij, double* uij, double* fij,
          double* uijshift);                                        /* src/ljfrc.c:21 */
int ewald_real_frc(double rijsq, double chargeij, double* uij, double* fij); /* src/elecfrc.c:17 */
int ewald_excl_frc(double rijsq, double chargeij, double* uij, double* fij); /* src/elecfrc.c:35 */
int wolf_real_frc(double rijsq, double chargeij, double* uij, double* fij);  /* src/elecfrc.c:48 */
int bndfrc(int iSpecie, int iBond, int iabs, double* uij, double* virial_ij);  /* src/bndfrc.c:17 */
int aglfrc(int iSpecie, int iAngle, int iabs, double* uij, double* virial_ij); /* src/aglfrc.c:17 */
int dihfrc(int iSpecie, int iDih, int iabs, double* uij, double* virial_ij);   /* src/dihfrc.c:17 */
int impfrc(int iSpecie, int iImp, int iabs, double* uij, double* virial_ij);   /* src/impfrc.c:17 */
int xatom_coulomb_frc(double rijsq, double chargeij, double* uij, double* fij); /* src/elecfrc.c:159 */
/* molecule-molecule Coulomb, cut off on the centre-of-mass distance; adds to
 * uelec, virial_inter and fxl/fyl/fzl as the reference does */
int xmole_coulomb_frc(int iSpecie, int iMole, int iabs, int iAtom, int jSpecie, int jMole,
                      int jabs, int jAtom);                         /* src/elecfrc.c:67 */
/* solid-fluid wall of one site: uij, and the site's force added to fxl.. */
int cal_sf_hypergeo(int ii, int iSpecie, int iAtom, double* uij, double* fij); /* src/sffrc.c:1120 */
int cal_sf_atom_explicit(int ii, int iSpecie, int iAtom, double* uij,
                         double* fij);                              /* src/sffrc.c:1213 */
/* cubic-interpolation table lookup (the MYINTERP tables read at start-up) */
int get_values_from_grid(double fxx, double fyy, double fzz, int type, double* usf,
                         double* fsf);                              /* src/sffrc.c:951 */
/* The Hermite grid of the Fortran library is not reachable from here: this
 * symbol prints a message and exits; give the grid to the context API
 * (mspms_sf_grid in mspms_gpu.h) instead. */
int call_tasos_forces(int itype, double fxx, double fyy, double fzz, double* usflj_tasos,
                      double* tasos_force);                         /* src/sffrc.c:1322 */

/* Releases the device context (optional; the OS reclaims it at exit). */
void mspms_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
