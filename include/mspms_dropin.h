/*
 * mspms_dropin.h -- the drop-in symbols of libmspms_dropin.a.
 *
 * Link libmspms_dropin.a + libmspms_gpu.so into the mspms2 program in place
 * of src/frclong.c and src/frcshort.c.  The two functions keep the
 * reference's prototypes (src/funcs.h:225 `int frclong();`, :233
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

/* Releases the device context (optional; the OS reclaims it at exit). */
void mspms_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
