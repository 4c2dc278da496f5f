/*
 * mspms_gpu.h -- C ABI of the MI355X (gfx950) force engine for mspms2.
 *
 * The engine evaluates mspms2's two force routines on one GPU:
 *   frclong  : inter-molecular LJ + Ewald (real, reciprocal, self, exclusion,
 *              vacuum) or Wolf summation, plus the solid-fluid terms
 *              (reference: src/frclong.c:20-1233)
 *   frcshort : bonds, angles, dihedrals, impropers
 *              (reference: src/frcshort.c:20-97)
 *
 * Two ways in:
 *   1. Drop-in symbols `int frclong(void)` / `int frcshort(void)` with the
 *      reference's global-state contract (declared in mspms_dropin.h, built
 *      into libmspms_dropin.a); they replace src/frclong.c and src/frcshort.c
 *      at link time.
 *   2. The context API below: plain pointers and sizes, no global state,
 *      several independent contexts per process, batched replicas.
 *
 * Units are the reference's reduced units (lengths / sigma_base, energies /
 * epsilon_base; src/readins.c:33-59).  All host arrays are FP64 SoA, exactly
 * the reference's xx/yy/zz and fxl/fyl/fzl layout (src/molevars.h:79-83).
 *
 * Every function returns 0 on success or a negative errno-style code
 * (MSPMS_E*); the reference instead prints and calls exit(1)
 * (src/bndfrc.c:60-62, src/aglfrc.c:186-188, src/dihfrc.c:278-280).
 * mspms_last_error() returns a message for the last failure on the thread.
 */
#ifndef MSPMS_GPU_H_
#define MSPMS_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define MSPMS_OK       0
#define MSPMS_EINVAL (-22) /* unknown term type, bad argument, unsupported mode */
#define MSPMS_EDOM   (-33) /* solid-fluid grid lookup out of range            */
#define MSPMS_ERANGE (-34) /* cutoff larger than half the box                 */
#define MSPMS_ENOMEM (-12) /* device allocation failed                        */
#define MSPMS_EIO    (-5)  /* HIP runtime failure                             */
#define MSPMS_ENOTFINITE (-74) /* non-finite position or force                 */

/* ---- enumerations: values identical to src/defs.h:24-75 ------------------- */
#define MSPMS_GHOST_NONE 0
#define MSPMS_GHOST_LJ   1
#define MSPMS_GHOST_FULL 2

#define MSPMS_BOND_NONE     0
#define MSPMS_BOND_HARMONIC 1
#define MSPMS_BOND_MORSE    2
#define MSPMS_BOND_FENE     3

#define MSPMS_ANGLE_NONE     0
#define MSPMS_ANGLE_HARMONIC 1
#define MSPMS_ANGLE_TR_WATER 2

#define MSPMS_DIH_NONE       0
#define MSPMS_DIH_OPLS_COSIN 1
#define MSPMS_DIH_CHARMM     2

#define MSPMS_IMP_NONE   0
#define MSPMS_IMP_CHARMM 1

#define MSPMS_SF_NONE              0
#define MSPMS_SF_HYPERGEO          1
#define MSPMS_SF_ATOM_EXPLICIT     2
#define MSPMS_SF_TASOS             3 /* tricubic Hermite grid (mylibtasos) */
#define MSPMS_SF_MY_INTERP         4 /* 32-coefficient cube polynomials   */

#define MSPMS_ELEC_NONE           0
#define MSPMS_ELEC_EWALD          1
#define MSPMS_ELEC_WOLF           2
#define MSPMS_ELEC_SIMPLE_COULOMB 3

#define MSPMS_EWALD_BC_TINFOIL 1
#define MSPMS_EWALD_BC_VACUUM  2

/* precision modes */
#define MSPMS_PREC_FP64  0 /* reference operation order, FP64 everywhere      */
#define MSPMS_PREC_MIXED 1 /* FP32 pair arithmetic, FP64 accumulation,
                              FP64 Ewald reciprocal and bonded terms          */

/* ---- molecule template (one per species) ------------------------------------
 * Mirrors SAMPLE_MOLECULE (src/molevars.h:40-73) as flat arrays.  Parameters
 * are already reduced as src/readins.c:118-231 leaves them.  Index arrays are
 * row-major: bnd_idx[2*b+k], agl_idx[3*a+k], dih_idx[4*d+k], imp_idx[4*m+k].
 * agl_unique / dih_unique are the isAngle_unique / isDih_unique flags that
 * InitCheckUniques (src/initvars.c:13-146) computes; pass NULL to let the
 * engine compute them with that routine's rules. */
typedef struct mspms_species {
    int natom;                 /* sites per molecule                      */
    int nmole;                 /* molecules of this species               */
    const int*    ghost_type;  /* [natom]                                 */
    const int*    interp_type; /* [natom] grid table index (SF_TASOS)     */
    const double* aw;          /* [natom] reduced site mass               */
    const double* sigma;       /* [natom]                                 */
    const double* epsilon;     /* [natom]                                 */
    const double* charge;      /* [natom] (e)                             */

    int nbond;
    const int*    bnd_idx;     /* [2*nbond]                               */
    const int*    bnd_type;    /* [nbond]                                 */
    const double* bnd_para;    /* [3*nbond]: Kb, Req, alpha               */

    int nangle;
    const int*    agl_idx;     /* [3*nangle]                              */
    const int*    agl_type;    /* [nangle]                                */
    const double* agl_para;    /* [5*nangle]: para_1..para_5              */
    const uint8_t* agl_unique; /* [nangle] or NULL                        */

    int ndih;
    const int*    dih_idx;     /* [4*ndih]                                */
    const int*    dih_type;    /* [ndih]                                  */
    const double* dih_para;    /* [4*ndih]: c1..c4                        */
    const uint8_t* dih_unique; /* [ndih] or NULL                          */

    int nimp;
    const int*    imp_idx;     /* [4*nimp]                                */
    const int*    imp_type;    /* [nimp]                                  */
    const double* imp_para;    /* [2*nimp]: komega, omega0                */

    int nnbp;                  /* must be 0: src/frclong.c:793 reads
                                  nbp_idx with the dihedral counter       */
} mspms_species;

/* ---- solid-fluid tables -------------------------------------------------- */
/* Tricubic Hermite grid, the data of mylibtasos Read_Grids
 * (mylibtasos/grid.f90:54-129, global.f90:214-227) in SoA node order
 * node = (iz*ny + iy)*nx + ix.  8 fields per node and table type:
 * F, dF/dx, dF/dy, dF/dz, d2F/dxdy, d2F/dxdz, d2F/dydz, d3F/dxdydz, in
 * Kelvin and Angstrom (the Fortran's units; the engine converts as
 * src/sffrc.c:1322-1337 does).  Grid spans [x0, x0+(nx-1)*hx] etc.; z is
 * periodic with period zperiod (0: (nz-1)*hz; at most (nz-1)*hz), x and y
 * are not (out of range = EDOM).  A grid read from the Fortran library's
 * files has nx = ngx+2, ny = ngy+2, nz = ngz+2 nodes from the origin and
 * zperiod = the unit cell's c = ngz*hz (mylibtasos/initialize.f90:96-101,
 * pbc.f90:82-86); the minimum node F, which sets the PE sign flip of
 * forces.f90:166-171, is taken over all nodes as grid.f90:102 does. */
typedef struct mspms_sf_grid {
    int ntype;                 /* number of interp types                  */
    int nx, ny, nz;
    double x0, y0, z0;         /* Angstrom                                */
    double hx, hy, hz;         /* Angstrom                                */
    const float* nodes;        /* [ntype][8][nz][ny][nx] FP32             */
    double zperiod;            /* Angstrom; 0: (nz-1)*hz                  */
} mspms_sf_grid;

/* Cube polynomials of the "myinterp" model, the data init_my_interp reads
 * (src/sffrc.c:228-748, globals src/vars.h:319-472).  Per cube and table
 * type, four polynomials (energy in J/mol, then the force components in
 * J/mol/Angstrom) of 32 coefficients each over the monomials of
 * Amatrix_ele_assign_value (src/sffrc.c:750-797) in ABSOLUTE Angstrom
 * coordinates.  tables[t*128 + f*32 + m] points at the ncube values of
 * coefficient m of polynomial f (the reference's ene0..31, fxa0..31,
 * fya0..31, fza0..31 of type t), cube = (ix*ncy + iy)*ncz + iz.  A site at
 * x (Angstrom) uses cube ix = (int)((x - xmin)/hx), likewise y; z is
 * wrapped into [0, zmax) first (src/sffrc.c:969-983).  Table type =
 * interp_type - 1 (init_my_interp decrements it, src/sffrc.c:249-257). */
typedef struct mspms_sf_cubes {
    int ntype;                 /* grid files (unique atoms), 1..5         */
    int ncx, ncy, ncz;         /* cubes per axis = nodes - 1              */
    double xmin, ymin, zmax;   /* Angstrom: xcenter - uclx/2, ycenter -
                                  ucly/2, uclz                            */
    double hx, hy, hz;         /* grid interval (Angstrom)                */
    const double* const* tables;  /* [ntype*128] host arrays of ncube     */
} mspms_sf_cubes;

/* ---- run parameters ---------------------------------------------------------
 * Field names follow src/vars.h:30-212.  Fill the "input" fields, then call
 * mspms_derive_params() to compute the derived constants with the exact
 * expressions of src/initvars.c:397-472 and :591-599, or fill them yourself
 * (the drop-in copies the reference's globals verbatim). */
typedef struct mspms_params {
    /* inputs */
    double sigma_base, epsilon_base;       /* src/readins.c:33-35          */
    double boxlx, boxly, boxlz;            /* reduced                      */
    double rcuton, rcutoff, rcutoffelec;   /* reduced                      */
    int    isLJswitchOn;
    double f0;                             /* 1-4 LJ scale                 */
    int    iChargeType;                    /* MSPMS_ELEC_*                 */
    double kappa;                          /* reduced (A^-1 * sigma_base)  */
    int    kmaxx, kmaxy, kmaxz, ksqmax;
    int    fEwald_BC;                      /* MSPMS_EWALD_BC_*             */
    int    iSF_type;                       /* MSPMS_SF_*                   */
    int    iExternal_FF_type;              /* must be 0                    */

    /* derived */
    double rcutoffsq, rcutonsq, rcutoffelecsq, roff2_minus_ron2_cube;
    double shift1, shift4;
    double coulomb_prefactor;
    double twopi_lx, twopi_ly, twopi_lz;
    double bfactor_ewald, vfactor_ewald, twopi_over_3v;
    double wolfvcon1, wolfvcon2, wolfvcon3, wolffcon1, wolffcon2;

    /* SF_HYPERGEO (src/sffrc.c:16-78): reduced tube data */
    int    ntube;
    const double* tube_x;                  /* [ntube] hgntc_xx             */
    const double* tube_y;                  /* [ntube] hgntc_yy             */
    const double* tube_r;                  /* [ntube] hgnt_radius          */
    double const_3pisq_theta;
    double solid_sigma, solid_epsilon;     /* reduced                      */

    /* SF_ATOM_EXPLICIT (src/sffrc.c:81-160): reduced solid atoms */
    int    solid_natom;
    const double* solid_xx;
    const double* solid_yy;
    const double* solid_zz;

    /* SF_TASOS */
    const mspms_sf_grid* grid;

    /* SF_MY_INTERP */
    const mspms_sf_cubes* cubes;
} mspms_params;

/* Outputs of frclong: the globals src/frclong.c writes (src/vars.h:173-210),
 * plus pair statistics.  n_lj_pairs / n_elec_pairs count inter-molecular
 * (i<j) site pairs inside rcutoff / rcutoffelec (src/frclong.c:123-125,
 * :157); lj_pair_hash / elec_pair_hash are order-independent sums of
 * splitmix64((uint64)i<<32 | j) over those pairs, i < j numbered within the
 * replica (0 = its first atom). */
typedef struct mspms_long_terms {
    double uvdw, ushift, ureal, uexcl, ufourier, uself, uvacuum, uelec;
    double usflj, uotherff, uinter, virial_inter;
    int64_t  n_lj_pairs, n_elec_pairs;
    uint64_t lj_pair_hash, elec_pair_hash;
} mspms_long_terms;

/* Outputs of frcshort (src/frcshort.c:26-87).  usflj: the solid-fluid
 * energy when mspms_set_sf_split has moved the wall into frcshort and
 * frcshort ran on its own (0 otherwise; a RESPA step reports the wall's
 * energy in its long terms, as src/frclong.c:1223 does). */
typedef struct mspms_short_terms {
    double ubond, uangle, udih, uimp, uintra, virial_intra;
    double usflj;
} mspms_short_terms;

/* Parameter gradients of the nonbonded energy uinter (LJ + Ewald or Wolf,
 * with the intramolecular corrections, self and reciprocal terms) with
 * respect to the per-site template values epsilon, sigma, charge (the
 * reduced values of mspms_species), one row per replica:
 * [replica][global site id], global site id = sum of natom of earlier
 * species + site.  A parameter shared by several sites (an atom type) has
 * the sum of their entries as its derivative.  With epsilon_i = 0 < epsilon_j
 * for an interacting pair, dE/d(epsilon_i) is +-inf (the mixing rule
 * sqrt(eps_i eps_j) has no derivative there).  No reference counterpart:
 * the reference has no derivatives with respect to parameters. */
typedef struct mspms_param_grad {
    double* dE_depsilon;   /* [nsite] */
    double* dE_dsigma;     /* [nsite] */
    double* dE_dcharge;    /* [nsite] */
} mspms_param_grad;

/* ---- helpers ------------------------------------------------------------ */

/* Computes the derived fields of *p from its inputs with the expressions of
 * src/initvars.c:433-462 (Ewald, Wolf, Coulomb prefactor) and :591-599
 * (cutoff squares, switch factor, shift).  Validates cutoff <= L/2
 * (src/validation.c:613-626): returns MSPMS_ERANGE otherwise. */
int mspms_derive_params(mspms_params* p);

/* Fills unique[] with the isAngle_unique / isDih_unique rules of
 * InitCheckUniques (src/initvars.c:13-146), including that routine's
 * index quirks (src/initvars.c:66,86,134 mark entries by the wrong loop
 * index).  agl_unique: [nangle], dih_unique: [ndih]. */
int mspms_check_uniques(const mspms_species* sp, uint8_t* agl_unique,
                        uint8_t* dih_unique);

const char* mspms_last_error(void);

/* ---- context API -------------------------------------------------------- */
typedef struct mspms_ctx mspms_ctx;

/* Creates a context on HIP device `device` with MSPMS_PREC_* `precision`. */
int mspms_create(mspms_ctx** out, int device, int precision);
int mspms_destroy(mspms_ctx* ctx);

/* Topology + parameters of one system (copied; caller keeps ownership). */
int mspms_set_system(mspms_ctx* ctx, int nspecie, const mspms_species* sp,
                     const mspms_params* p);

/* Replaces only the parameters (e.g. a new box after a volume move,
 * src/nptnh.c:156-200 / src/volchg.c); the topology stays. */
int mspms_set_params(mspms_ctx* ctx, const mspms_params* p);

/* Batched replicas: `nrep` independent boxes sharing the species templates
 * set by mspms_set_system.  nmole[r*nspecie + s] molecules of species s in
 * replica r; params[r] gives each replica's box, cutoffs and Ewald/Wolf
 * constants.  Positions and forces of all replicas are packed replica after
 * replica, each replica's atoms in the reference order.  One call of
 * mspms_frclong_* / mspms_frcshort_* then evaluates every replica in a
 * single launch per kernel; terms[] receives one entry per replica. */
int mspms_set_replicas(mspms_ctx* ctx, int nrep, const int* nmole,
                       const mspms_params* params);

int mspms_natom(const mspms_ctx* ctx);      /* total sites, all replicas */
int mspms_nreplica(const mspms_ctx* ctx);

/* Host-buffer evaluation: uploads x/y/z, evaluates, downloads forces and
 * terms (one mspms_*_terms per replica).  Forces are overwritten, as the
 * reference overwrites fxl / fxs (src/frclong.c:39-42, src/frcshort.c:33-36).
 */
int mspms_frclong_host(mspms_ctx* ctx, const double* xx, const double* yy,
                       const double* zz, double* fxl, double* fyl, double* fzl,
                       mspms_long_terms* terms);
int mspms_frcshort_host(mspms_ctx* ctx, const double* xx, const double* yy,
                        const double* zz, double* fxs, double* fys,
                        double* fzs, mspms_short_terms* terms);

/* Device-buffer evaluation on `stream` (a hipStream_t, NULL = the context's
 * stream).  All pointers are device pointers; terms has one entry per
 * replica.  Asynchronous: synchronize the stream before reading. */
int mspms_frclong_device(mspms_ctx* ctx, const double* d_xx,
                         const double* d_yy, const double* d_zz, double* d_fxl,
                         double* d_fyl, double* d_fzl,
                         mspms_long_terms* d_terms, void* stream);
int mspms_frcshort_device(mspms_ctx* ctx, const double* d_xx,
                          const double* d_yy, const double* d_zz,
                          double* d_fxs, double* d_fys, double* d_fzs,
                          mspms_short_terms* d_terms, void* stream);

/* Term-level entry points, evaluated on the host by the term code of the GPU
 * kernels (FP64, the reference's expression order), with the run globals
 * the reference's term functions read passed as `p`.
 * mspms_term_pair replaces, with which =
 *   MSPMS_TERM_LJ          ljfrc(rijsq, sigmaij=a, epsilonij=b, &u, &f, &ushift)
 *                          (src/ljfrc.c:21-71)
 *   MSPMS_TERM_EWALD_REAL  ewald_real_frc(rijsq, chargeij=a, &u, &f)
 *                          (src/elecfrc.c:17-33)
 *   MSPMS_TERM_EWALD_EXCL  ewald_excl_frc(rijsq, chargeij=a, &u, &f)
 *                          (src/elecfrc.c:35-46)
 *   MSPMS_TERM_WOLF_REAL   wolf_real_frc(rijsq, chargeij=a, &u, &f)
 *                          (src/elecfrc.c:48-60)
 *   MSPMS_TERM_COULOMB     xatom_coulomb_frc(rijsq, chargeij=a, &u, &f)
 *                          (src/elecfrc.c:159-170)
 * mspms_term_bonded replaces bndfrc / aglfrc / dihfrc / impfrc(iSpecie,
 * index, iabs, &u, &virial) (src/bndfrc.c:17, src/aglfrc.c:17,
 * src/dihfrc.c:17, src/impfrc.c:17) for the molecule whose sites are at
 * xx[0..natom) (the reference's xx + iabs), adding its forces into
 * fxs[0..natom) like the reference does into fxs + iabs.  A singular
 * improper returns MSPMS_EINVAL where impfrc exits. */
enum {
    MSPMS_TERM_LJ = 0, MSPMS_TERM_EWALD_REAL = 1, MSPMS_TERM_EWALD_EXCL = 2,
    MSPMS_TERM_WOLF_REAL = 3, MSPMS_TERM_COULOMB = 4
};
enum { MSPMS_TERM_BOND = 0, MSPMS_TERM_ANGLE = 1, MSPMS_TERM_DIH = 2, MSPMS_TERM_IMP = 3 };
int mspms_term_pair(const mspms_params* p, int which, double rijsq, double a,
                    double b, double* uij, double* fij, double* ushift);
int mspms_term_bonded(const mspms_species* sp, int kind, int index,
                      const double* xx, const double* yy, const double* zz,
                      double* fxs, double* fys, double* fzs, double* u,
                      double* vir);
/* The solid-fluid term of one site, evaluated on the host by the code the
 * kernels run: cal_sf_hypergeo (iSF_type 1, src/sffrc.c:1118-1209) and
 * cal_sf_atom_explicit (iSF_type 2, src/sffrc.c:1213-1314) for site `site`
 * of *sp at reduced (x, y, z); the site's force is added to (*fx, *fy, *fz)
 * as the reference adds it to fxl, the energy goes to *u and the scalar
 * force of the last tube / wall atom in range to *fij (left unchanged when
 * none is).  Other iSF_type values: MSPMS_EINVAL. */
int mspms_term_sf(const mspms_params* p, const mspms_species* sp, int site, double x,
                  double y, double z, double* fx, double* fy, double* fz, double* u,
                  double* fij);
/* get_values_from_grid (src/sffrc.c:951-1089): the cube polynomials of table
 * `type` (the reference's decremented interp type, 0-based) of p->cubes at
 * reduced (x, y, z); *usf in J/mol and fsf[3] in J/mol/Angstrom, the
 * reference's raw units.  A cube outside the tables: MSPMS_EDOM. */
/* call_tasos_forces (src/sffrc.c:1322-1337) over CForce_atom
 * (mylibtasos/forces.f90:6-173): the tricubic Hermite grid p->grid of table
 * `type` (the site's interp_type, 1-based) at reduced (x, y, z); *u and
 * force3[3] in reduced units, as call_tasos_forces returns them (PE with the
 * sign flip below 1.5 times the minimum node F).  A position outside the
 * grid or a type without a table: MSPMS_EDOM.  The node table is converted
 * once per grid and kept. */
int mspms_term_hermite(const mspms_params* p, int type, double x, double y, double z,
                       double* u, double* force3);
int mspms_term_myinterp(const mspms_params* p, double x, double y, double z, int type,
                        double* usf, double* fsf);

/* Adjoint of the nonbonded energy with respect to the per-site force-field
 * parameters, in FP64 whatever the context precision.  Host buffers; grad
 * arrays hold nreplica * nsite entries.  Also returns the FP64 frclong
 * forces (= -dE/dr) and terms like mspms_frclong_host.  Solid-fluid terms
 * are not differentiated (MSPMS_EINVAL when a replica has one). */
int mspms_energy_param_grad(mspms_ctx* ctx, const double* xx, const double* yy,
                            const double* zz, double* fxl, double* fyl,
                            double* fzl, mspms_long_terms* terms,
                            mspms_param_grad* grad);

/* One RESPA velocity-Verlet outer step in the NVE ensemble, device
 * resident: replaces vver (src/md.c:22-83, the integrator of
 * src/mspms2.c:99-132 and md_move src/dispmd.c:42-46) with every buffer on
 * the device.  pos/vel/fl/fs are (3, N) FP64 SoA device arrays; fl and fs
 * must hold the long / short forces of the current positions (one
 * mspms_frclong_device + mspms_frcshort_device call before the first step,
 * as src/mspms2.c:50-51 does) and are updated to those of the new
 * positions.  lterms / sterms / ukin receive one value per replica: the
 * frclong terms, the frcshort terms of the last inner substep and the
 * kinetic energy (src/md.c:76-78). */
int mspms_vver_step_device(mspms_ctx* ctx, double* pos, double* vel, double* fl,
                           double* fs, double delt, int nstep_inner,
                           mspms_long_terms* lterms, mspms_short_terms* sterms,
                           double* ukin, void* stream);

/* nsteps calls of mspms_vver_step_device with the same arguments.  With
 * use_graph != 0 one outer step (all kernels of every replica) is captured
 * once as a HIP graph on the context stream and replayed on `stream`; the
 * graph is kept for the same argument set until the system is rebuilt.
 * lhist / shist / ukin_hist (device, optional) receive every step's terms
 * and kinetic energies, nsteps * nreplica entries, step-major: each step's
 * closing reduction pass appends them itself (no copy per step), which is
 * the per-step record the reference's collect_aves() accumulates
 * (src/mspms2.c:101-102, src/averages.c:19-95).  Without any history only
 * the last step's terms are reduced.  No counterpart in the reference: its
 * outer loop (src/mspms2.c:99-132) calls the integrator once per step from
 * the host. */
int mspms_vver_run_device(mspms_ctx* ctx, double* pos, double* vel, double* fl,
                          double* fs, double delt, int nstep_inner, int nsteps,
                          mspms_long_terms* lterms, mspms_short_terms* sterms,
                          double* ukin, mspms_long_terms* lhist,
                          mspms_short_terms* shist, double* ukin_hist,
                          int use_graph, void* stream);

/* Batched hybrid Monte Carlo, one independent chain per replica: hmc()
 * (src/hmc.c:23-70) with canonical (MD) moves, md_move (src/dispmd.c:12-103)
 * and velinit (src/initvars.c:151-232), all on the device.
 * mspms_hmc_init seeds each replica's RANMAR stream (rmarin(ij[r], kl[r]),
 * src/random.c:14-57; 0 <= ij <= 31328, 0 <= kl <= 30081) and draws its
 * initial velocities into d_vel ((3, N) device array); treq[r] is the
 * reduced temperature and nconstraint[r] the constrained degrees of freedom
 * (nfree = 3 N_r - nconstraint[r], src/initvars.c:588).
 * mspms_hmc_run_device runs nmove moves of every chain: the move-type draw,
 * fresh velocities (all but the first move), forces of the current
 * positions, nmd RESPA steps of delt (nstep_inner inner steps), then
 * Metropolis on dH = [d uinter + d uintra + d ukin + d ushift] / treq;
 * a rejected chain gets its positions back.  use_graph replays one
 * captured move.  mspms_hmc_stats gives per replica the accepted and
 * attempted moves and the last dH; mspms_hmc_draw takes n numbers from a
 * replica's stream (the generator's known-answer test). */
int mspms_hmc_init(mspms_ctx* ctx, const int* ij, const int* kl, const double* treq,
                   const int* nconstraint, double* d_vel);
/* Per-chain outer time steps and hmc()'s equilibration-phase delt
 * adaptation (src/hmc.c:47-58).  delt[r] is replica r's reduced outer step;
 * adj_cycle[r] = nstep_delt_adj_cycle (0, or adj_cycle null: delt fixed) and
 * rreq_disp[r] the target acceptance: after every adj_cycle moves of a chain
 * its delt becomes delt * (1 - rreq_disp + accepted / adj_cycle), the
 * reference's update while bEquilibrium holds.  From this call on
 * mspms_hmc_run_device steps each chain with its own delt (its delt argument
 * is ignored); mspms_hmc_get_delt reads the current steps back.  A later
 * mspms_set_system / mspms_set_replicas ends it.  No counterpart in the
 * reference, which keeps one global delt (src/vars.h). */
int mspms_hmc_set_delt(mspms_ctx* ctx, const double* delt, const int* adj_cycle,
                       const double* rreq_disp);
int mspms_hmc_get_delt(mspms_ctx* ctx, double* delt);
int mspms_hmc_run_device(mspms_ctx* ctx, double* pos, double* vel, double* fl,
                         double* fs, double delt, int nstep_inner, int nmd, int nmove,
                         mspms_long_terms* lterms, mspms_short_terms* sterms,
                         double* ukin, int use_graph, void* stream);
int mspms_hmc_stats(mspms_ctx* ctx, long long* accepted, long long* moves,
                    double* last_dH);
int mspms_hmc_draw(mspms_ctx* ctx, int replica, int n, double* out);

/* Nose-Hoover NVT molecular dynamics on device buffers: the reference's
 * vver_nh_3 (src/nvtnh.c:11-221, selected by src/mspms2.c:66-69), one
 * thermostat on the outer step (ss, ps, mass qq) and one on the RESPA
 * inner steps (sss, pss, mass qqs), degrees of freedom gg = ggs = nfree
 * (init_nvt, src/initvars.c:298-347: ss = ps = sss = pss = 0 at the
 * start).  mspms_nvt_set uploads one state per replica (all fields but
 * the outputs unhts, unhtss, ukin and tinst); mspms_nvt_run_device runs
 * nsteps steps of every replica from pos / vel with fl / fs holding the
 * long and short forces of pos (as after the reference's start-up
 * frclong / frcshort), leaving lterms / sterms of the last step and ukin
 * [nrep]; replayed from one captured graph when use_graph is set.
 * mspms_nvt_get downloads the states (thermostat variables, energies
 * unhts / unhtss, ukin, tinst = 2 ukin / nfree). */
typedef struct mspms_nvt_state {
    double qq, qqs;            /* thermostat masses (outer, inner)          */
    double gg, ggs;            /* thermostat degrees of freedom (nfree)     */
    double treq;               /* reduced temperature                       */
    double ss, ps, sss, pss;   /* thermostat positions and momenta          */
    double unhts, unhtss;      /* thermostat energies (outputs)             */
    double ukin, tinst;        /* kinetic energy, temperature (outputs)     */
    int nfree, pad;
} mspms_nvt_state;
int mspms_nvt_set(mspms_ctx* ctx, const mspms_nvt_state* states);
int mspms_nvt_get(mspms_ctx* ctx, mspms_nvt_state* states);
int mspms_nvt_run_device(mspms_ctx* ctx, double* pos, double* vel, double* fl, double* fs,
                         double delt, int nstep_inner, int nsteps, mspms_long_terms* lterms,
                         mspms_short_terms* sterms, double* ukin, int use_graph, void* stream);

/* Solid-fluid grid generation on the device: the reference's offline
 * generator tools/sfgridgen.c:71-352.  At every node (x0 + h i, y0 + h j,
 * h k) of a grid over one unit cell of a z-periodic solid (nx = floor(uclx /
 * h) + 2 nodes, likewise y and z; x0 = y0 = 0 as the generator centres the
 * cell) it sums the fluid-solid LJ potential (Lorentz-Berthelot mixing,
 * epsilon converted to J/mol, optional switch between rcuton and rcutoff)
 * over the solid atoms and their z images within the cutoff, in the
 * generator's order and FP64 expressions, so that the tables equal its
 * out.sfgridgen bit for bit.  out: 13 fields of nx*ny*nz nodes in the
 * generator's order (energy, dE/dx, dE/dy, dE/dz, then the Jacobian of the
 * force -grad E, dFx/dx, dFx/dy, dFx/dz, dFy/dx .. dFz/dz), node = (i ny + j)
 * nz + k.  With the switch on, a pair in the switching shell adds the Jacobian
 * terms of the
 * node's previous pair (the generator's empty else branch,
 * tools/sfgridgen.c:254-256); they start at zero here. */
typedef struct mspms_sfgridgen_spec {
    double f_sigma, f_epsilon;        /* fluid site (Angstrom, K)          */
    double grid_itvl, rcuton, rcutoff;  /* Angstrom                        */
    int isLJswitch, s_natom;          /* solid atoms: 1..1000               */
    double uclx, ucly, uclz;          /* unit cell (Angstrom)               */
    double s_sigma, s_epsilon;        /* solid site (Angstrom, K)           */
    const double *sx, *sy, *sz;       /* [s_natom] solid atoms (Angstrom)   */
} mspms_sfgridgen_spec;
int mspms_sfgridgen_dims(const mspms_sfgridgen_spec* spec, int* nx, int* ny, int* nz);
int mspms_sfgridgen(int device, const mspms_sfgridgen_spec* spec, double* out);
/* mspms_sfgridgen without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture preparation, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * solid atoms are staged in a fixed per-GPU device array of 3 x 1000 FP64
 * (24,000 bytes, held until the process ends); more than 1000 solid atoms is
 * refused with MSPMS_EINVAL.  The 13 x nx*ny*nz tables (mspms_sfgridgen_dims)
 * are written into the caller's device buffer d_out on `stream` (null: the
 * null stream); the call returns when they are written.  Calls that overlap
 * on one GPU share the staging array and race. */
int mspms_sfgridgen_noalloc(int device, const mspms_sfgridgen_spec* spec, double* d_out,
                            void* stream);

/* Generator of the myinterp cube polynomials (the reference reads them,
 * src/sffrc.c:228-748, but ships no program that writes them).
 *
 * mspms_myinterp_nodes: at the nodes of mspms_sfgridgen (same spec, same
 * order node = (i ny + j) nz + k, switch off only) the fluid-solid LJ sum
 * and its derivatives up to third order, as node_data[q*nnode + node] with
 * q = f*7 + d for f in (E, Fx, Fy, Fz) (J/mol and J/mol/Angstrom, F = -grad
 * E) and d in (value, d/dx, d/dy, d/dz, d2/dxdy, d2/dxdz, d2/dydz).
 *
 * mspms_myinterp_fit: per cube of the node lattice (nx-1 by ny-1 by nz-1
 * cubes, lower corner (x0 + hx i, y0 + hy j, z0 + hz k)) and per
 * polynomial f, the 32 coefficients that fit the value and gradient at the
 * eight corners (least squares: the 32 conditions have rank 29); the three
 * functions that vanish with their gradient at every corner, u v (1-u)(1-v)
 * and its xz and yz analogues in cube-local coordinates, are added with the
 * weights that fit the mixed second derivatives at the corners.  A
 * polynomial of the 32-monomial space is reproduced exactly.  Output as the
 * file stores it: coef[(f*32 + m)*ncube + cube], absolute coordinates. */
int mspms_myinterp_nodes(int device, const mspms_sfgridgen_spec* spec, double* node_data);
int mspms_myinterp_fit(int device, int nx, int ny, int nz, double x0, double y0, double z0,
                       double hx, double hy, double hz, const double* node_data, double* coef);

/* Simulated annealing of every replica on device buffers: the reference's
 * siman (src/siman.c:61-182; init_siman, :8-59).  Cycle k of replica r runs
 * at T = t_initial[r] / mu_t[r]^k (reduced): velinit at T from the
 * replica's RANMAR stream, frclong / frcshort of the current positions,
 * then n_tries vver_nh_3 steps from zeroed thermostats (qq, qqs, nfree of
 * mspms_nvt_set).  A replica stops after its first cycle run below t_min[r]
 * (later cycles of the others leave it as it is), at most max_cycles
 * cycles.  Needs mspms_hmc_init (streams seeded as rmarin(ij, kl) and the
 * start-up velinit done, as in init_vars) and mspms_nvt_set.  Outputs:
 * cycles[r] run and t_final[r], the temperature of the last cycle; pos,
 * vel, fl, fs, lterms, sterms, ukin as after mspms_nvt_run_device. */
int mspms_siman_run_device(mspms_ctx* ctx, double* pos, double* vel, double* fl, double* fs,
                           double delt, int nstep_inner, int n_tries, const double* t_initial,
                           const double* t_min, const double* mu_t, int max_cycles,
                           int* cycles, double* t_final, mspms_long_terms* lterms,
                           mspms_short_terms* sterms, double* ukin, void* stream);

/* NPT molecular dynamics on device buffers: the reference's npt_respa with
 * the Nose-Hoover thermostat / barostat operator npt_nh_operator
 * (src/nptnh.c:10-234, selected by src/mspms2.c:70-73).  The box scales by
 * exp(delts vbs) after every RESPA inner step, and with it the replica's
 * box-dependent tables (cell grid, Ewald weights, 2 pi / L, 2 pi / V) and,
 * with lj_tail (the reference's isLJlrcOn), the LJ tail pressure
 * pljlrc = sum pljlrc_term[m][n] n_m n_n / V^2 (src/misc.c:131-148).
 * mspms_npt_set resets every replica to the box of its parameters and
 * uploads one state per replica (Qts, Qbs, treq, preq, vts, vbs, rts,
 * nfree, lj_tail; init_npt_respa, src/initvars.c:351-396, starts from
 * vts = sqrt((nfree+1)/Qts), vbs = sqrt((nfree+1)/Qbs), rts = 0).
 * mspms_npt_run_device runs nsteps steps of every replica; on entry fl /
 * fs and lterms / sterms hold the forces and terms of pos (the barostat
 * reads their virials), on exit those of the new positions, with ukin
 * [nrep].  The pair lists are rebuilt every step (the box moves).
 * mspms_npt_get downloads the states, current boxes included (outputs
 * utsbs, the conserved quantity's thermostat / barostat part, ukin, tinst
 * and pinst).  After a run, mspms_set_params with the new box keeps later
 * rebuilds of the tables (launch shape, skin, parameters) consistent. */
typedef struct mspms_npt_state {
    double Qts, Qbs;             /* thermostat and barostat masses            */
    double treq, preq;           /* reduced temperature and pressure          */
    double vts, vbs, rts;        /* thermostat and barostat velocities,
                                    thermostat position                        */
    double pljlrc;               /* LJ tail pressure of the box (output)      */
    double utsbs, ukin, tinst, pinst;  /* outputs                             */
    double boxl[3], boxv;        /* current box (output)                      */
    int nfree, lj_tail;
} mspms_npt_state;
int mspms_npt_set(mspms_ctx* ctx, const mspms_npt_state* states);
int mspms_npt_get(mspms_ctx* ctx, mspms_npt_state* states);
int mspms_npt_run_device(mspms_ctx* ctx, double* pos, double* vel, double* fl, double* fs,
                         double delt, int nstep_inner, int nsteps, mspms_long_terms* lterms,
                         mspms_short_terms* sterms, double* ukin, int use_graph, void* stream);

/* Synchronises the device and reports the status accumulated by the
 * device-buffer evaluations since the previous mspms_sync (list overflow,
 * non-finite input, grid out of range, singular improper), which the
 * host-buffer calls check by themselves; the status is cleared.  On a
 * pair-list overflow the lists are regrown before MSPMS_ENOMEM is returned.
 * If the overflow happened in NVE RESPA steps (mspms_vver_step_device /
 * mspms_vver_run_device, and the steps of HMC moves), the positions and
 * velocities of that run are put back to the start of the first step that
 * overflowed, i.e. the end of the last good step (every step saves its start
 * while no overflow is flagged; the message names the step); the caller
 * recomputes the forces and terms (mspms_frclong_device /
 * mspms_frcshort_device) and repeats the steps from there.  For the other
 * device calls (NVT, NPT, annealing, plain evaluations) nothing is put back:
 * keep a copy of the state from the previous sync to repeat them.  HMC chain
 * statistics and random streams are not put back. */
int mspms_sync(mspms_ctx* ctx);

/* The context's own HIP stream (hipStream_t). */
void* mspms_stream(const mspms_ctx* ctx);

/* Launch-shape autotuner.  Times `repeats` frclong evaluations of the
 * device positions d_xx.. for each candidate shape -- sites per binning
 * cell (6, 8, 12, 16; this sets the cluster geometry and so the pair-list
 * length) x waves sharing an i-cluster's pair-list row in the pair kernel
 * (1, 2, 4, 8) -- and keeps the fastest for the context (until the
 * next mspms_autotune / mspms_set_launch_shape).  Results are the same for
 * every shape up to FP64 summation order.  No counterpart in the
 * reference (a CPU code with nothing to tune). */
#define MSPMS_TUNE_MAX 16
typedef struct mspms_tune_result {
    double cell_atoms;               /* chosen sites per binning cell   */
    int    waves_per_cluster;        /* chosen waves per i-cluster      */
    int    ncand;                    /* candidates timed                */
    double ms;                       /* mean frclong time of the choice */
    double cand_cell_atoms[MSPMS_TUNE_MAX];
    int    cand_wpc[MSPMS_TUNE_MAX];
    double cand_ms[MSPMS_TUNE_MAX];
} mspms_tune_result;
int mspms_autotune(mspms_ctx* ctx, const double* d_xx, const double* d_yy,
                   const double* d_zz, int repeats, mspms_tune_result* out);
/* Sets the launch shape by hand (cell_atoms in 2..64; waves 1, 2, 4, 8). */
int mspms_set_launch_shape(mspms_ctx* ctx, double cell_atoms, int waves_per_cluster);

/* Pair-list buffer.  The reference tests every molecule pair on every call
 * (src/frclong.c:52-125); this engine finds candidate pairs through a
 * cluster-pair list built with radius max(rcutoff, rcutoffelec) + skin and
 * reuses it while every site stays within skin / 2 of its position at the
 * list's build, checked on the device at each frclong (no host round
 * trip; a reuse only refreshes the sites' positions).  Which pairs
 * interact is decided per pair with the cutoff test either way, so the
 * pair sets and results do not depend on the skin; a larger skin trades
 * longer pair-list rows for fewer builds.  skin in 0..8 (length units of
 * the positions, default 0: a list is reused only for unchanged
 * positions).  mspms_list_builds reports the number of builds since the
 * last change of system, parameters, launch shape or skin. */
/* RESPA split of the solid-fluid wall (on != 0): the wall's forces move
 * from frclong to frcshort, so a device RESPA step evaluates them in every
 * inner step with the bonded terms (the fast forces) and frclong keeps the
 * pairs and the electrostatics.  A RESPA step still reports the wall's
 * energy in the long terms (usflj, in uinter); a standalone frcshort reports
 * it in mspms_short_terms.usflj and a standalone frclong leaves it out.
 * Supported by the NVE RESPA step (mspms_vver_*) and HMC; the NVT / NPT
 * integrators refuse it.  No counterpart in the reference, which evaluates
 * the wall in frclong (src/frclong.c:1145-1209). */
int mspms_set_sf_split(mspms_ctx* ctx, int on);
/* Upper bound on the pair-list capacity per cluster row (entries; 0: none,
 * the default).  Rows that need more overflow (MSPMS_ENOMEM, see mspms_sync,
 * which lifts the bound).  No counterpart in the reference. */
int mspms_set_list_capacity(mspms_ctx* ctx, int max_entries);
int mspms_set_list_skin(mspms_ctx* ctx, double skin);
int mspms_list_builds(mspms_ctx* ctx, long long* builds);

/* Kernel-group timing with HIP events on the evaluation stream: switch on,
 * evaluate, then read the mean milliseconds per call of each group
 * (0 clustering + pair list, 1 nonbonded pairs, 2 intramolecular,
 * 3 Ewald S(k) + reduction, 4 Ewald forces, 5 finalize, 6 bonded / RESPA
 * inner loop, 7 solid-fluid).  Device-buffer calls are accumulated at
 * mspms_sync(); steps replayed from a graph are not timed. */
int mspms_set_profiling(mspms_ctx* ctx, int on);
int mspms_last_timings(const mspms_ctx* ctx, double* ms, int n);

#ifdef __cplusplus
}
#endif
#endif /* MSPMS_GPU_H_ */
