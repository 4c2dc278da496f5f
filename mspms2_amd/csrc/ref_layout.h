/*
 * ref_layout.h -- SAMPLE_MOLECULE of the reference (src/molevars.h:40-73,
 * limits :18-23, string lengths src/defs.h:97-98), restated field for field
 * so that the drop-in can read the reference program's sample_mole[]
 * global.  tests/test_dropin_layout.py compares every offset with the
 * reference's compiled layout.
 */
#ifndef MSPMS_REF_LAYOUT_H_
#define MSPMS_REF_LAYOUT_H_
#include <stdbool.h>

#define REF_SAMPLE_NATOM_MAX 200 /* src/molevars.h:18-23 */
#define REF_SAMPLE_NTERM_MAX 100
#define REF_SHORT_STRING 5       /* src/defs.h:97-98 */
#define REF_LONG_STRING 200

typedef struct {
    char atom_name[REF_SAMPLE_NATOM_MAX][REF_SHORT_STRING];
    int type[REF_SAMPLE_NATOM_MAX], ghost_type[REF_SAMPLE_NATOM_MAX],
        interp_type[REF_SAMPLE_NATOM_MAX];
    double aw[REF_SAMPLE_NATOM_MAX], epsilon[REF_SAMPLE_NATOM_MAX],
        sigma[REF_SAMPLE_NATOM_MAX], charge[REF_SAMPLE_NATOM_MAX];
    double xx[REF_SAMPLE_NATOM_MAX], yy[REF_SAMPLE_NATOM_MAX], zz[REF_SAMPLE_NATOM_MAX];
    double ee[REF_SAMPLE_NATOM_MAX], ff[REF_SAMPLE_NATOM_MAX], gg[REF_SAMPLE_NATOM_MAX];
    char mole_name[REF_LONG_STRING];
    double mw;
    int natom, nbond, nangle, ndih, nimp, nnbp;
    int bnd_idx[REF_SAMPLE_NTERM_MAX][2];
    int bnd_type[REF_SAMPLE_NTERM_MAX];
    double bnd_para_0[REF_SAMPLE_NTERM_MAX], bnd_para_1[REF_SAMPLE_NTERM_MAX],
        bnd_para_2[REF_SAMPLE_NTERM_MAX];
    int agl_idx[REF_SAMPLE_NTERM_MAX][3];
    int agl_type[REF_SAMPLE_NTERM_MAX];
    double agl_para_1[REF_SAMPLE_NTERM_MAX], agl_para_2[REF_SAMPLE_NTERM_MAX],
        agl_para_3[REF_SAMPLE_NTERM_MAX], agl_para_4[REF_SAMPLE_NTERM_MAX],
        agl_para_5[REF_SAMPLE_NTERM_MAX];
    bool isAngle_unique[REF_SAMPLE_NTERM_MAX];
    int dih_idx[REF_SAMPLE_NTERM_MAX][4];
    int dih_type[REF_SAMPLE_NTERM_MAX];
    double dih_para_1[REF_SAMPLE_NTERM_MAX], dih_para_2[REF_SAMPLE_NTERM_MAX],
        dih_para_3[REF_SAMPLE_NTERM_MAX], dih_para_4[REF_SAMPLE_NTERM_MAX];
    bool isDih_unique[REF_SAMPLE_NTERM_MAX];
    int imp_idx[REF_SAMPLE_NTERM_MAX][4];
    int imp_type[REF_SAMPLE_NTERM_MAX];
    double imp_para_0[REF_SAMPLE_NTERM_MAX], imp_para_1[REF_SAMPLE_NTERM_MAX];
    double nbp_idx[REF_SAMPLE_NTERM_MAX][2];
} ref_sample_molecule;

#endif
