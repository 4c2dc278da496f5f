/*
 * ref_harness.c -- driver around the reference's own force routines.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked against objects compiled from the
 * reference's sources (see build_ref.sh); never part of the product.
 *
 * In a run directory holding in.mspms / cfg.mspms / xyz.mspms it runs the
 * reference's start-up path (src/mspms2.c:39-51: opening, readins,
 * fnValidateInput, init_vars, fnValidateInit), then frclong() and
 * frcshort(), and writes every output global to a binary dump:
 *
 *   "MSPMSREF" | int32 version=1 | int32 natom | int32 nscalar |
 *   double scalar[nscalar] | double xx,yy,zz,fxl,fyl,fzl,fxs,fys,fzs [natom]
 *
 * With --md-steps N (version 2) it then runs N steps of the integrator the
 * input selects, as the reference's main loop does (src/mspms2.c:64-77,
 * :96-98: vver, vver_nh_3 or npt_respa), and the dump holds the state after
 * them, followed by
 *   int32 nmd | double md[nmd] (MD_SCALARS below) |
 *   double vx0,vy0,vz0 [natom] (velinit's velocities, before step 1) |
 *   double vx,vy,vz [natom] (after step N)
 *
 * With --write-files it finally writes the reference's own trajectory
 * frame, snapshot and checkpoint of that state into RUNDIR (trajectory(),
 * snapshot(), saveit(): trj.mspms, ss.mspms, sav.mspms).
 *
 * With --hmc N it runs N moves of the reference's hybrid Monte Carlo
 * driver hmc() (src/hmc.c:23-70: velinit + frclong + frcshort on every move
 * but the first, md_move over nstep_md_per_hmc vver steps, Metropolis on
 * dH, the delt adaptation while bEquilibrium) as the main loop calls it
 * (src/mspms2.c:99-100, istep from nstep_start, within nstep_end), prints
 * one line per move
 *   hmc_move K ACCEPT dH delt uinter uintra ukin
 * and dumps the state after the last move (version 2).
 *
 * With --siman it runs the reference's simulated annealing driver siman()
 * (src/siman.c:61-182: cycles of velinit + vver_nh_3 steps, cooling
 * between them) instead, and dumps the state it ends in (version 2, no
 * integrator timing).
 *
 * Usage: mspms_ref_harness RUNDIR OUTFILE [--repeat K] [--md-steps N] [--write-files]
 *                          [--siman] [--hmc N]
 *   --repeat K times K further frclong() calls and prints seconds/call;
 *   --md-steps N also prints the integrator's seconds per step.
 *
 * Configurations that would reach code absent from this build (the
 * Fortran grid library, iSF_type 3; the GSL metal-cluster force field,
 * iExternal_FF_type != 0) are refused before init_vars.  iSF_type 4
 * (myinterp, src/sffrc.c:228-1089) is plain C and runs.
 * --tasos-linked: iSF_type 3 where the harness is linked with the engine's
 * frclong or call_tasos_forces, which read the TASOS section's grid files
 * themselves: init_vars runs with iSF_type 0, so that init_tasos_grid (which
 * calls into the absent library) is skipped, and iSF_type 3 is restored
 * before frclong.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "mspms2.h"

static void put(FILE* f, const double* v, int n) { fwrite(v, sizeof(double), (size_t)n, f); }

static double v0[3][NATOM_MAX];

int main(int argc, char** argv)
{
    int repeat = 0, nmd = 0, files = 0, sa = 0, nhmc = 0, tasos = 0, k;
    char outpath[4096];
    FILE* fo;
    if (argc < 3) {
        fprintf(stderr, "usage: %s RUNDIR OUTFILE [--repeat K]\n", argv[0]);
        return 2;
    }
    for (k = 3; k < argc; k++)
        if (!strcmp(argv[k], "--repeat") && k + 1 < argc)
            repeat = atoi(argv[++k]);
        else if (!strcmp(argv[k], "--md-steps") && k + 1 < argc)
            nmd = atoi(argv[++k]);
        else if (!strcmp(argv[k], "--write-files"))
            files = 1;
        else if (!strcmp(argv[k], "--siman"))
            sa = 1;
        else if (!strcmp(argv[k], "--hmc") && k + 1 < argc)
            nhmc = atoi(argv[++k]);
        else if (!strcmp(argv[k], "--tasos-linked"))
            tasos = 1;
    if (argv[2][0] == '/') {
        snprintf(outpath, sizeof outpath, "%s", argv[2]);
    } else {
        char cwd[2048];
        if (!getcwd(cwd, sizeof cwd)) return 2;
        snprintf(outpath, sizeof outpath, "%s/%s", cwd, argv[2]);
    }
    if (chdir(argv[1]) != 0) {
        perror("chdir");
        return 2;
    }
    opening();
    readins();
    if ((iSF_type == SF_NANOTUBE_TASOS && !tasos) || iExternal_FF_type != FF_NONE) {
        fprintf(stderr, "ref_harness: iSF_type=%d / iExternal_FF_type=%d need "
                        "libraries absent from this build\n",
                iSF_type, iExternal_FF_type);
        return 3;
    }
    fnValidateInput();
    if (tasos && iSF_type == SF_NANOTUBE_TASOS) {
        iSF_type = SF_NONE;
        init_vars();
        iSF_type = SF_NANOTUBE_TASOS;
    } else {
        init_vars();
    }
    fnValidateInit();
    frclong();
    frcshort();
    if (repeat > 0) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (k = 0; k < repeat; k++) frclong();
        clock_gettime(CLOCK_MONOTONIC, &t1);
        printf("frclong_seconds_per_call %.9f\n",
               ((t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec)) / repeat);
    }
    for (k = 0; k < natom; k++) {
        v0[0][k] = vx[k];
        v0[1][k] = vy[k];
        v0[2][k] = vz[k];
    }
    if (sa) {
        siman();
        nmd = 1;  /* version-2 dump of the final state */
    } else if (nhmc > 0) {
        if (what_simulation != HYBRID_MONTE_CARLO || nstep_start + nhmc - 1 > nstep_end) {
            fprintf(stderr, "ref_harness: --hmc needs an HMC input with at least %d steps "
                            "in its current phase\n", nhmc);
            return 3;
        }
        for (k = 0; k < nhmc; k++) {
            istep = nstep_start + k;
            hmc();
            printf("hmc_move %d %d %.17g %.17g %.17g %.17g %.17g\n", k, isAccept ? 1 : 0, dH,
                   delt, uinter, uintra, ukin);
        }
        nmd = 1;  /* version-2 dump of the final state */
    } else if (nmd > 0) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (k = 0; k < nmd; k++) {
            if (what_ensemble == NVT)
                vver_nh_3();
            else if (what_ensemble == NPT)
                npt_respa();
            else
                vver();
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        printf("md_seconds_per_step %.9f\n",
               ((t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec)) / nmd);
    }
    if (files) {
        trajectory();
        fclose(fptrj);
        snapshot();
        saveit();
    }
    fo = fopen(outpath, "wb");
    if (!fo) {
        perror("fopen");
        return 2;
    }
    {
        const double sc[] = {
            uvdw, ushift, ureal, uexcl, ufourier, uself, uvacuum, uelec,
            usflj, uotherff, uinter, virial_inter,
            ubond, uangle, udih, uimp, uintra, virial_intra,
            boxlx, boxly, boxlz, rcutoff, rcutoffsq, rcutoffelec, rcutoffelecsq,
            rcuton, rcutonsq, roff2_minus_ron2_cube, shift1, shift4,
            kappa, coulomb_prefactor, TWOPI_LX, TWOPI_LY, TWOPI_LZ,
            Bfactor_ewald, Vfactor_ewald, twopi_over_3v,
            wolfvcon1, wolfvcon2, wolfvcon3, wolffcon1, wolffcon2,
            f0, sigma_base, epsilon_base, uljlrc, pljlrc, time_base, delt,
            delts, const_3pisq_theta};
        int nsc = (int)(sizeof sc / sizeof sc[0]), ver = nmd > 0 ? 2 : 1;
        fwrite("MSPMSREF", 1, 8, fo);
        fwrite(&ver, sizeof ver, 1, fo);
        fwrite(&natom, sizeof natom, 1, fo);
        fwrite(&nsc, sizeof nsc, 1, fo);
        put(fo, sc, nsc);
    }
    put(fo, xx, natom);
    put(fo, yy, natom);
    put(fo, zz, natom);
    put(fo, fxl, natom);
    put(fo, fyl, natom);
    put(fo, fzl, natom);
    put(fo, fxs, natom);
    put(fo, fys, natom);
    put(fo, fzs, natom);
    if (nmd > 0) {
        /* MD_SCALARS */
        const double md[] = {
            (double)nmd, (double)what_ensemble, (double)nfree, (double)nstep_inner, treq,
            preq, delt, delts, qq, qqs, gg, ggs, ss, ps, sss, pss, unhts, unhtss, ukin, tinst,
            Qts, Qbs, vts, vbs, rts, utsbs, boxv, pinst};
        int n = (int)(sizeof md / sizeof md[0]);
        fwrite(&n, sizeof n, 1, fo);
        put(fo, md, n);
        put(fo, v0[0], natom);
        put(fo, v0[1], natom);
        put(fo, v0[2], natom);
        put(fo, vx, natom);
        put(fo, vy, natom);
        put(fo, vz, natom);
    }
    fclose(fo);
    return 0;
}
