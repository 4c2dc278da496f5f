/*
 * dropin_globals.c -- the reference program's globals read into the C ABI's
 * parameter block and species tables, shared by the drop-in frclong() /
 * frcshort() (dropin.c) and the term-level functions (dropin_terms.c).
 *
 * The reference defines its state as global variables in headers
 * (src/molevars.h:75-104, src/vars.h:30-212) and builds with -fcommon; this
 * file declares the ones the force routines use as `extern`, with the
 * reference's exact types.  SAMPLE_MOLECULE's layout (src/molevars.h:40-73,
 * limits :18-23) is restated in ref_sample_molecule; a test compares its
 * field offsets with the reference's compiled layout.
 */
#include <ctype.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mspms_dropin.h"
#include "../../include/mspms_gpu.h"

#include "ref_layout.h"

/* ---- the reference's globals (src/molevars.h, src/vars.h) ---------------- */
extern ref_sample_molecule sample_mole[];
extern int natom, nspecie;
extern int nmole_per_specie[];
extern double xx[], yy[], zz[];
extern double fxl[], fyl[], fzl[], fxs[], fys[], fzs[];
extern double sigma_base, epsilon_base;
extern double boxlx, boxly, boxlz;
extern double rcuton, rcutoff, rcutoffelec, rcutoffsq, rcutonsq, rcutoffelecsq;
extern double roff2_minus_ron2_cube, shift1, shift4, f0;
extern int isLJswitchOn, iChargeType, iSF_type, iExternal_FF_type;
extern double kappa;
extern int KMAXX, KMAXY, KMAXZ, KSQMAX, fEwald_BC;
extern double TWOPI_LX, TWOPI_LY, TWOPI_LZ, Bfactor_ewald, Vfactor_ewald, twopi_over_3v;
extern double wolfvcon1, wolfvcon2, wolfvcon3, wolffcon1, wolffcon2;
extern double coulomb_prefactor;
extern double uvdw, ushift, ureal, uexcl, ufourier, uself, uvacuum, uelec, usflj, uotherff;
extern double uinter, virial_inter;
extern double ubond, uangle, udih, uimp, uintra, virial_intra;
extern int ntube;
extern double *hgntc_xx, *hgntc_yy, *hgnt_radius;
extern double const_3pisq_theta;
extern double *solid_sigma, *solid_epsilon;
extern int solid_natom;
extern double *solid_xx, *solid_yy, *solid_zz;
/* myinterp cube tables (src/vars.h:319-472), NUNIQUE_ATOM_MAX = 5 */
#define MI_32(X)                                                                        \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) \
    X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29)  \
    X(30) X(31)
#define MI_DECL(k) extern double *ene##k[], *fxa##k[], *fya##k[], *fza##k[];
MI_32(MI_DECL)
extern double xmin, ymin, zmax, grid_itvl_x, grid_itvl_y, grid_itvl_z;
extern int ncube_x, ncube_y, ncube_z;

/* ---- parameter copies the C ABI takes ------------------------------------- */
#define MAXSP 3 /* NSPECIE_MAX, src/molevars.h:15 */
/* interleaved parameter copies the C ABI takes */
static double g_bnd[MAXSP][3 * REF_SAMPLE_NTERM_MAX], g_agl[MAXSP][5 * REF_SAMPLE_NTERM_MAX],
    g_dih[MAXSP][4 * REF_SAMPLE_NTERM_MAX], g_imp[MAXSP][2 * REF_SAMPLE_NTERM_MAX];
static unsigned char g_au[MAXSP][REF_SAMPLE_NTERM_MAX], g_du[MAXSP][REF_SAMPLE_NTERM_MAX];
static const double* g_mi_tables[5 * 128];
static mspms_sf_cubes g_cubes;
/* interp_type as the cfg file gives it: init_my_interp decrements the
 * reference's copy in place (src/sffrc.c:249-257) */
static int g_interp[MAXSP][REF_SAMPLE_NATOM_MAX];

/* ---- the Fortran library's Hermite grids (iSF_type 3) ---------------------
 * init_tasos_grid (src/sffrc.c:149-221) hands the TASOS section of in.mspms
 * to mylibtasos, which sizes the grid (InitPotentialGrid,
 * mylibtasos/initialize.f90:96-101: USESYMMETRY off, its compiled default,
 * global.f90:235; grid_spacing 0.2 A, global.f90:214) and reads one file of
 * list-directed records per grid (Read_Grids, grid.f90:54-129).  That state
 * lives in Fortran module variables C cannot reach, so the drop-in reads the
 * same section and files itself, once, from the run directory. */
static mspms_sf_grid g_grid;
static float* g_grid_nodes;

static int fortran_anint(double v) { return (int)(floor(fabs(v) + 0.5) * (v >= 0 ? 1 : -1)); }

static void grid_fail(const char* what, const char* name)
{
    fprintf(stderr, "mspms gpu: TASOS grid: %s%s%s\n", what, name ? " " : "", name ? name : "");
    exit(1);
}

/* the whole file, with Fortran D exponents and comma separators made
 * readable for strtod */
static char* slurp(const char* path)
{
    FILE* f = fopen(path, "rb");
    long n;
    char* buf;
    size_t k;
    if (!f) grid_fail("cannot open", path);
    if (fseek(f, 0, SEEK_END) != 0 || (n = ftell(f)) < 0 || fseek(f, 0, SEEK_SET) != 0)
        grid_fail("cannot size", path);
    buf = (char*)malloc((size_t)n + 1);
    if (!buf || fread(buf, 1, (size_t)n, f) != (size_t)n) grid_fail("cannot read", path);
    fclose(f);
    buf[n] = 0;
    for (k = 0; k < (size_t)n; k++) {
        if (buf[k] == 'D' || buf[k] == 'd') buf[k] = 'E';
        else if (buf[k] == ',') buf[k] = ' ';
    }
    return buf;
}

static void read_tasos_grids(void)
{
    char line[4096], key[128], names[12][200];
    double auc, buc, cuc, radius, h[3];
    int na, nb, nc, nspecies, tnuatoms, ng[3], t, k, found = 0;
    size_t nn;
    FILE* f = fopen("in.mspms", "r"); /* INPUT, src/defs.h:4 */
    if (!f) grid_fail("cannot open", "in.mspms");
    while (fgets(line, sizeof line, f)) {
        if (sscanf(line, "%127s", key) != 1) continue;
        for (k = 0; key[k]; k++) key[k] = (char)toupper((unsigned char)key[k]);
        if (!strcmp(key, "TASOS")) {
            found = 1;
            break;
        }
    }
    if (!found) grid_fail("no TASOS section in", "in.mspms");
    if (!fgets(line, sizeof line, f) ||
        sscanf(line, "%lf %lf %lf %lf", &auc, &buc, &cuc, &radius) != 4 ||
        !fgets(line, sizeof line, f) || sscanf(line, "%d %d %d", &na, &nb, &nc) != 3 ||
        !fgets(line, sizeof line, f) || sscanf(line, "%d", &nspecies) != 1 ||
        !fgets(line, sizeof line, f) || sscanf(line, "%d", &tnuatoms) != 1 || tnuatoms < 1 ||
        tnuatoms > 12 || nspecies < 1 || nspecies > tnuatoms)
        grid_fail("malformed TASOS section in", "in.mspms");
    for (t = 0; t < tnuatoms; t++)
        if (!fgets(line, sizeof line, f) || sscanf(line, "%199s", names[t]) != 1)
            grid_fail("missing grid file name in", "in.mspms");
    fclose(f);
    (void)radius;
    (void)na;
    (void)nb;
    (void)nc;
    ng[0] = fortran_anint(auc / 0.2);
    ng[1] = fortran_anint(buc / 0.2);
    ng[2] = fortran_anint(cuc / 0.2);
    if (ng[0] < 1 || ng[1] < 1 || ng[2] < 1) grid_fail("unit cell smaller than one grid step", 0);
    h[0] = auc / ng[0];
    h[1] = buc / ng[1];
    h[2] = cuc / ng[2];
    nn = (size_t)(ng[0] + 2) * (ng[1] + 2) * (ng[2] + 2);
    g_grid_nodes = (float*)malloc(sizeof(float) * 8 * nn * (size_t)nspecies);
    if (!g_grid_nodes) grid_fail("out of memory", 0);
    for (t = 0; t < nspecies; t++) {
        char* text = slurp(names[t]);
        char* s = text;
        size_t node;
        /* one record of 8 values per node, x fastest (grid.f90:76-99) */
        for (node = 0; node < nn; node++)
            for (k = 0; k < 8; k++) {
                char* e;
                double v = strtod(s, &e);
                if (e == s) grid_fail("too few node records in", names[t]);
                s = e;
                g_grid_nodes[((size_t)t * 8 + k) * nn + node] = (float)v;
            }
        free(text);
    }
    g_grid.ntype = nspecies;
    g_grid.nx = ng[0] + 2;
    g_grid.ny = ng[1] + 2;
    g_grid.nz = ng[2] + 2;
    g_grid.x0 = g_grid.y0 = g_grid.z0 = 0.0;
    g_grid.hx = h[0];
    g_grid.hy = h[1];
    g_grid.hz = h[2];
    g_grid.nodes = g_grid_nodes;
    g_grid.zperiod = cuc; /* PBCType 2 wraps z by the cell's c (pbc.f90:82-86) */
}

void mspms_dropin_read_params(mspms_params* p)
{
    memset(p, 0, sizeof *p);
    p->sigma_base = sigma_base;
    p->epsilon_base = epsilon_base;
    p->boxlx = boxlx;
    p->boxly = boxly;
    p->boxlz = boxlz;
    p->rcuton = rcuton;
    p->rcutoff = rcutoff;
    p->rcutoffelec = rcutoffelec;
    p->isLJswitchOn = isLJswitchOn;
    p->f0 = f0;
    p->iChargeType = iChargeType;
    p->kappa = kappa;
    p->kmaxx = KMAXX;
    p->kmaxy = KMAXY;
    p->kmaxz = KMAXZ;
    p->ksqmax = KSQMAX;
    p->fEwald_BC = fEwald_BC;
    p->iSF_type = iSF_type;
    p->iExternal_FF_type = iExternal_FF_type;
    p->rcutoffsq = rcutoffsq;
    p->rcutonsq = rcutonsq;
    p->rcutoffelecsq = rcutoffelecsq;
    p->roff2_minus_ron2_cube = roff2_minus_ron2_cube;
    p->shift1 = shift1;
    p->shift4 = shift4;
    p->coulomb_prefactor = coulomb_prefactor;
    p->twopi_lx = TWOPI_LX;
    p->twopi_ly = TWOPI_LY;
    p->twopi_lz = TWOPI_LZ;
    p->bfactor_ewald = Bfactor_ewald;
    p->vfactor_ewald = Vfactor_ewald;
    p->twopi_over_3v = twopi_over_3v;
    p->wolfvcon1 = wolfvcon1;
    p->wolfvcon2 = wolfvcon2;
    p->wolfvcon3 = wolfvcon3;
    p->wolffcon1 = wolffcon1;
    p->wolffcon2 = wolffcon2;
    if (iSF_type == MSPMS_SF_HYPERGEO) {
        p->ntube = ntube;
        p->tube_x = hgntc_xx;
        p->tube_y = hgntc_yy;
        p->tube_r = hgnt_radius;
        p->const_3pisq_theta = const_3pisq_theta;
        p->solid_sigma = solid_sigma ? *solid_sigma : 0.0;
        p->solid_epsilon = solid_epsilon ? *solid_epsilon : 0.0;
    } else if (iSF_type == MSPMS_SF_ATOM_EXPLICIT) {
        p->solid_natom = solid_natom;
        p->solid_xx = solid_xx;
        p->solid_yy = solid_yy;
        p->solid_zz = solid_zz;
        p->solid_sigma = solid_sigma ? *solid_sigma : 0.0;
        p->solid_epsilon = solid_epsilon ? *solid_epsilon : 0.0;
    } else if (iSF_type == MSPMS_SF_TASOS) {
        if (!g_grid_nodes) read_tasos_grids();
        p->grid = &g_grid;
    } else if (iSF_type == MSPMS_SF_MY_INTERP) {
        /* init_my_interp filled ene0..fza31[t] for its nunique_atom files */
        int t, n = 0;
        while (n < 5 && ene0[n]) n++;
        for (t = 0; t < n; t++) {
#define MI_FILL(k)                                \
    g_mi_tables[t * 128 + (k)] = ene##k[t];       \
    g_mi_tables[t * 128 + 32 + (k)] = fxa##k[t];  \
    g_mi_tables[t * 128 + 64 + (k)] = fya##k[t];  \
    g_mi_tables[t * 128 + 96 + (k)] = fza##k[t];
            MI_32(MI_FILL)
#undef MI_FILL
        }
        g_cubes.ntype = n;
        g_cubes.ncx = ncube_x;
        g_cubes.ncy = ncube_y;
        g_cubes.ncz = ncube_z;
        g_cubes.xmin = xmin;
        g_cubes.ymin = ymin;
        g_cubes.zmax = zmax;
        g_cubes.hx = grid_itvl_x;
        g_cubes.hy = grid_itvl_y;
        g_cubes.hz = grid_itvl_z;
        g_cubes.tables = g_mi_tables;
        p->cubes = &g_cubes;
    }
}

void mspms_dropin_species(mspms_species* out)
{
    int s, k;
    for (s = 0; s < nspecie; s++) {
        const ref_sample_molecule* m = sample_mole + s;
        mspms_species* q = out + s;
        memset(q, 0, sizeof *q);
        q->natom = m->natom;
        q->nmole = nmole_per_specie[s];
        q->ghost_type = m->ghost_type;
        if (iSF_type == MSPMS_SF_MY_INTERP) {
            for (k = 0; k < m->natom; k++) g_interp[s][k] = m->interp_type[k] + 1;
            q->interp_type = g_interp[s];
        } else {
            q->interp_type = m->interp_type;
        }
        q->aw = m->aw;
        q->sigma = m->sigma;
        q->epsilon = m->epsilon;
        q->charge = m->charge;
        q->nbond = m->nbond;
        q->bnd_idx = &m->bnd_idx[0][0];
        q->bnd_type = m->bnd_type;
        for (k = 0; k < m->nbond; k++) {
            g_bnd[s][3 * k] = m->bnd_para_0[k];
            g_bnd[s][3 * k + 1] = m->bnd_para_1[k];
            g_bnd[s][3 * k + 2] = m->bnd_para_2[k];
        }
        q->bnd_para = g_bnd[s];
        q->nangle = m->nangle;
        q->agl_idx = &m->agl_idx[0][0];
        q->agl_type = m->agl_type;
        for (k = 0; k < m->nangle; k++) {
            g_agl[s][5 * k] = m->agl_para_1[k];
            g_agl[s][5 * k + 1] = m->agl_para_2[k];
            g_agl[s][5 * k + 2] = m->agl_para_3[k];
            g_agl[s][5 * k + 3] = m->agl_para_4[k];
            g_agl[s][5 * k + 4] = m->agl_para_5[k];
            g_au[s][k] = m->isAngle_unique[k] ? 1 : 0;
        }
        q->agl_para = g_agl[s];
        q->agl_unique = g_au[s];
        q->ndih = m->ndih;
        q->dih_idx = &m->dih_idx[0][0];
        q->dih_type = m->dih_type;
        for (k = 0; k < m->ndih; k++) {
            g_dih[s][4 * k] = m->dih_para_1[k];
            g_dih[s][4 * k + 1] = m->dih_para_2[k];
            g_dih[s][4 * k + 2] = m->dih_para_3[k];
            g_dih[s][4 * k + 3] = m->dih_para_4[k];
            g_du[s][k] = m->isDih_unique[k] ? 1 : 0;
        }
        q->dih_para = g_dih[s];
        q->dih_unique = g_du[s];
        q->nimp = m->nimp;
        q->imp_idx = &m->imp_idx[0][0];
        q->imp_type = m->imp_type;
        for (k = 0; k < m->nimp; k++) {
            g_imp[s][2 * k] = m->imp_para_0[k];
            g_imp[s][2 * k + 1] = m->imp_para_1[k];
        }
        q->imp_para = g_imp[s];
        q->nnbp = m->nnbp;
    }
}
