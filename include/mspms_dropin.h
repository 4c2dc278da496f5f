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

/* Releases the device context (optional; the OS reclaims it at exit). */
void mspms_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
