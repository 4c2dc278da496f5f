/*
 * dropin.c -- frclong() / frcshort() over the reference program's globals,
 * evaluated by the GPU engine (libmspms_gpu.so).  See mspms_dropin.h.  The
 * globals are read into the C ABI's tables by dropin_globals.c.
 */
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mspms_dropin.h"
#include "../../include/mspms_gpu.h"

/* ---- the reference's globals these two functions write (src/vars.h) ------ */
extern int natom, nspecie;
extern int nmole_per_specie[];
extern double xx[], yy[], zz[];
extern double fxl[], fyl[], fzl[], fxs[], fys[], fzs[];
extern int iChargeType, fEwald_BC;
extern double uvdw, ushift, ureal, uexcl, ufourier, uself, uvacuum, uelec, usflj, uotherff;
extern double uinter, virial_inter;
extern double ubond, uangle, udih, uimp, uintra, virial_intra;

void mspms_dropin_read_params(mspms_params* p);
void mspms_dropin_species(mspms_species* sp);

/* ---- cached device context ------------------------------------------------ */
#define MAXSP 3 /* NSPECIE_MAX, src/molevars.h:15 */
static mspms_ctx* g_ctx;
static int g_nspecie = -1, g_natom = -1, g_nmole[MAXSP];
static mspms_params g_params;

static void die(const char* what, int rc)
{
    fprintf(stderr, "mspms gpu %s failed (%d): %s\n", what, rc, mspms_last_error());
    exit(1);
}

static void ensure_context(void)
{
    int s, rc, same = (g_ctx != NULL && g_nspecie == nspecie && g_natom == natom);
    mspms_params p;
    mspms_species sp[MAXSP];
    if (nspecie < 1 || nspecie > MAXSP) die("setup (nspecie)", MSPMS_EINVAL);
    for (s = 0; same && s < nspecie; s++) same = g_nmole[s] == nmole_per_specie[s];
    mspms_dropin_read_params(&p);
    if (!g_ctx) {
        const char* prec = getenv("MSPMS_GPU_PRECISION");
        const char* dev = getenv("MSPMS_GPU_DEVICE");
        int precision = (prec && !strcmp(prec, "fp64")) ? MSPMS_PREC_FP64 : MSPMS_PREC_MIXED;
        const char* skin = getenv("MSPMS_GPU_SKIN");  /* Angstrom */
        rc = mspms_create(&g_ctx, dev ? atoi(dev) : 0, precision);
        if (rc) die("mspms_create", rc);
        if (skin) {
            /* positions are in units of sigma_base */
            rc = mspms_set_list_skin(g_ctx, atof(skin) / p.sigma_base);
            if (rc) die("mspms_set_list_skin", rc);
        }
    }
    if (!same) {
        mspms_dropin_species(sp);
        rc = mspms_set_system(g_ctx, nspecie, sp, &p);
        if (rc) die("mspms_set_system", rc);
        g_nspecie = nspecie;
        g_natom = natom;
        for (s = 0; s < nspecie; s++) g_nmole[s] = nmole_per_specie[s];
        g_params = p;
    } else if (memcmp(&p, &g_params, sizeof p) != 0) {
        rc = mspms_set_params(g_ctx, &p);
        if (rc) die("mspms_set_params", rc);
        g_params = p;
    }
}

int frclong(void)
{
    mspms_long_terms t;
    int rc;
    ensure_context();
    rc = mspms_frclong_host(g_ctx, xx, yy, zz, fxl, fyl, fzl, &t);
    if (rc) die("frclong", rc);
    uvdw = t.uvdw;
    ushift = t.ushift;
    ureal = t.ureal;
    uexcl = t.uexcl;
    uelec = t.uelec;
    if (iChargeType == MSPMS_ELEC_EWALD) {
        /* the reference leaves these untouched for other charge types */
        ufourier = t.ufourier;
        uself = t.uself;
        if (fEwald_BC == MSPMS_EWALD_BC_VACUUM) uvacuum = t.uvacuum;
    } else if (iChargeType == MSPMS_ELEC_WOLF) {
        uself = t.uself;
    }
    usflj = t.usflj;
    uotherff = t.uotherff;
    uinter = t.uinter;
    virial_inter = t.virial_inter;
    return 0;
}

int frcshort(void)
{
    mspms_short_terms t;
    int rc;
    ensure_context();
    rc = mspms_frcshort_host(g_ctx, xx, yy, zz, fxs, fys, fzs, &t);
    if (rc) die("frcshort", rc);
    ubond = t.ubond;
    uangle = t.uangle;
    udih = t.udih;
    uimp = t.uimp;
    uintra = t.uintra;
    virial_intra = t.virial_intra;
    return 0;
}

void mspms_dropin_release(void)
{
    if (g_ctx) mspms_destroy(g_ctx);
    g_ctx = NULL;
    g_nspecie = g_natom = -1;
}

