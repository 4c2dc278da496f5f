/*
 * mspms_oracle.h -- CPU FP64 restatement of mspms2's force routines.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the GPU engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (mspms2_amd, libmspms_gpu.so) never calls it.
 *
 * The restatement follows the reference's loop order and arithmetic
 * expressions so that, built with gcc -std=c99 -O2 (no FP contraction, as the
 * reference is built), it reproduces the reference's FP64 results bit for
 * bit.  It carries the one-line fix of src/frclong.c:86 (the j-molecule atom
 * offset), without which the reference yields NaN for >2 molecules of a
 * species.  Every function cites the reference lines it follows.
 */
#ifndef MSPMS_ORACLE_H_
#define MSPMS_ORACLE_H_

#include "../include/mspms_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* frclong (src/frclong.c:20-1233).  xx/yy/zz: [natom] reduced, unwrapped.
 * Returns MSPMS_EINVAL for configurations the reference handles with
 * undefined behaviour (see oracle_validate). */
int oracle_frclong(int nspecie, const mspms_species* sp, const mspms_params* p,
                   const double* xx, const double* yy, const double* zz,
                   double* fxl, double* fyl, double* fzl, mspms_long_terms* t);

/* frcshort (src/frcshort.c:20-97). */
int oracle_frcshort(int nspecie, const mspms_species* sp, const mspms_params* p,
                    const double* xx, const double* yy, const double* zz,
                    double* fxs, double* fys, double* fzs, mspms_short_terms* t);

/* Rejects what the reference leaves undefined: BOND/ANGLE/DIH/IMP_NONE
 * (uninitialised energy, src/bndfrc.c:40-41), nnbp > 0 (src/frclong.c:793
 * indexes nbp_idx with iDih), unknown types (exit(1) in the reference). */
int oracle_validate(int nspecie, const mspms_species* sp, const mspms_params* p);

/* InitCheckUniques (src/initvars.c:13-146), index quirks included. */
void oracle_check_uniques(const mspms_species* sp, uint8_t* agl_unique,
                          uint8_t* dih_unique);

/* Derived constants (src/initvars.c:433-462, :591-599). */
int oracle_derive_params(mspms_params* p);

/* Tricubic Hermite solid-fluid lookup of one site: restatement of
 * CForce_atom (mylibtasos/forces.f90:6-173) and CHermInt3D
 * (mylibtasos/cubicinterpolation.f90:9-116) on the grid layout of
 * mspms_sf_grid.  Positions in Angstrom; PE in K, force in K/A.
 * Returns MSPMS_EDOM when x or y is outside the grid. */
int oracle_hermite_site(const mspms_sf_grid* g, int itype, double min_pe,
                        double x, double y, double z, double* pe,
                        double* force3);
/* minPE of Read_Grids (mylibtasos/grid.f90:61,:102): min node F of a type. */
double oracle_grid_min_pe(const mspms_sf_grid* g, int itype);

/* One RANMAR stream (src/random.c:14-109), for the known-answer test. */
typedef struct oracle_ranmar { double u[98]; double c, cd, cm; int i97, j97; } oracle_ranmar;
void oracle_rmarin(oracle_ranmar* r, int ij, int kl);
void oracle_ranmar_next(oracle_ranmar* r, double* rvec, int len);

/* Term level: which = 0 ljfrc (a sigma_ij, b epsilon_ij), 1 ewald_real_frc,
 * 2 ewald_excl_frc, 3 wolf_real_frc (a = q_i q_j); kind = 0 bndfrc, 1 aglfrc,
 * 2 dihfrc, 3 impfrc on the molecule coordinates xx.. (site indices of the
 * template), forces added into fxs.. */
void oracle_term_pair(const mspms_params* p, int which, double rijsq, double a,
                      double b, double* u, double* f, double* ushift);
int oracle_term_bonded(const mspms_species* q, int kind, int k, const double* xx,
                       const double* yy, const double* zz, double* fxs,
                       double* fys, double* fzs, double* u, double* vir);

#ifdef __cplusplus
}
#endif
#endif
