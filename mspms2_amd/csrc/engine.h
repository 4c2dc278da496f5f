// engine.h -- internal data layout shared by the host orchestration
// (engine.cpp) and the gfx950 kernels (k_*.hip).  Not part of the C ABI.
//
// HBM layout (all replicas packed back to back):
//   per atom      xx,yy,zz / fxl.. FP64 SoA (caller's arrays), atom_site,
//                 atom_mol (int), scratch for binning
//   per site      SiteDev table (a few hundred entries at most)
//   per molecule  MolDev {atom0, species, rep, natom, the species' site and term ranges}
//   per replica   RepDev: box, cutoffs, electrostatics constants, cell grid
//   per cluster   8 consecutive slots of the spatially sorted arrays:
//                 slot_orig (atom id), slot_info (mol, site),
//                 slot_x64 (FP64 unwrapped position, exact mode),
//                 slot_off (FP32 offset from the cluster centre + charge,
//                 mixed mode); cl_center (FP64), cl_ext (FP32 half extent)
//   pair list     per i-cluster a fixed-capacity row of packed j-cluster
//                 entries (j << 7 | ambiguous << 6 | image shift bits)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "terms.h"

#define MSPMS_CL 8            // atoms per cluster
#define MSPMS_MAX_SITES 1024  // site table entries kept in LDS
#define MSPMS_REDUCE_SLOTS 8  // per-block partial sums per kernel

// energy partial-sum slots (FP64) of the frclong reduction
enum {
    ES_UVDW = 0, ES_USHIFT, ES_UREAL, ES_UEXCL, ES_VIR, ES_UFOURIER, ES_USF,
    ES_NSLOT
};
// integer partial-sum slots (pair statistics)
enum { IS_NLJ = 0, IS_NEL, IS_HLJ, IS_HEL, IS_NSLOT };

struct SiteDev {
    double sigma, epsilon, charge, aw;
    double sqrt_eps;          // sqrt(epsilon), LB mixing
    float sigma_f, sqrt_eps_f, charge_f, half_sigma_f;
    int ghost, species, site, interp;
};

struct MolDev {
    int atom0, species, rep, natom;
    // the species' template fields, copied per molecule so that the
    // per-molecule kernels reach a molecule's terms and sites in one
    // dependent load less (no species lookup in between)
    int site0, term0, nintra, bterm0, nbterm;
    int pad[3];
};

struct RepDev {
    double L[3], invL[3];
    double rc_lj, rc2_lj, rc_el, rc2_el, rc_list;
    double rcon2, roff2_ron2_cube, shift1, shift4, f0;
    double C;                 // coulomb_prefactor
    double kappa;
    double wvc1, wvc2, wvc3, wfc1, wfc2;
    double vfactor, bfactor, twopi_l[3];
    double self_sum_q2;       // sum of q^2 over the replica's atoms
    int charge_type, switch_on, sf_type, pad0;
    int atom0, natom, mol0, nmol;
    // cell grid for clustering: columns along z, cells (cx,cy,cz)
    int ncx, ncy, ncz, cell0;  // cell0: first global cell of the replica
    int col0, ncol, pad1, pad2;
    double h[3], inv_h[3];
    // Ewald reciprocal work tables (per replica ranges)
    int krow0, nkrow, k0, nk;
    int kmax[3], ksqmax;
    double twopi_over_3v;     // vacuum boundary term, src/initvars.c:441
    int ewald_bc, pad3;
};

// Ewald k-vectors grouped in rows of fixed (kx, ky), kz contiguous
struct KRow {
    int kx, ky, kz_lo, kz_hi;  // kz range (inclusive); (0,0,0) excluded
    int k0;                   // index of (kx,ky,kz_lo) in the k list
    int rep;
    int nk;                   // number of k in the row
    int pad;
};

// the three groups of scratch used by one frclong evaluation
struct HistDev;

struct ClusterBufs {
    int* atom_cell;      // [natom]
    int* atom_slot;      // [natom]
    double* atom_wz;     // [natom] wrapped z
    int* cell_count;     // [ncell]  zero between calls
    int* cell_start;     // [ncell+1]
    int* cell_members;   // [natom]
    int* col_cell0;      // [ncol+1] first global cell of each column (static)
    int* col_rep;        // [ncol]   replica of each column (static)
    int* col_cl0;        // [ncol+1] first cluster of each column
    int* cl_col;         // [maxcl]
    int* ncl_total;      // [1]
    int* slot_orig;      // [8*maxcl]
    int2* slot_info;     // [8*maxcl] (mol, site)
    double4* slot_x64;   // [8*maxcl] FP64 unwrapped (x,y,z,-), current call
    double4* slot_w64;   // [8*maxcl] FP64 wrapped at the last build (x,y,z,charge)
    double4* atom_x0;    // [natom] unwrapped position at the last build
    int* lstate;         // [LS_N] list state: LS_VALID, LS_NEED (rebuild
                         //   wanted by this call), LS_BUILDS (builds so far),
                         //   LS_LAYOUT
    float4* slot_off;    // [8*maxcl] FP32 offset from centre + charge
    float4* slot_par;    // [8*maxcl] FP32 sigma, sqrt(eps) (0: no LJ),
                         //   molecule id and flags as int bits (SLOT_*)
    double4* cl_center;  // [maxcl]
    float4* cl_ext;      // [maxcl]
    float4* cl_box;      // [2 maxcl] the cluster's current box relative to its
                         //   build-time centre: (mid, -) then (half extent, -);
                         //   written by the build and refreshed by
                         //   k_list_check on every reuse
    float4* plist;       // [maxcl*cap] entries: FP32 centre separation
                         //   Ci - (Cj + image shift) (Ci - Cj for an
                         //   ambiguous image) and the entry bits in .w
    int* plist_n;        // [maxcl]
    int* atom_gslot;     // [natom] slot of each atom at the last build (-1: not
                         //   binned, a full ghost)
    double* fslot;       // [3 * 8*maxcl] pair forces of the half-list kernel in
                         //   slot order (x, y, z planes): its atomics hit one
                         //   64-B line per j-cluster and component; zeroed by
                         //   k_list_check, folded into the atoms' forces
                         //   after the pair kernel (SlotForces)
    HistDev* hist;       // the run's history descriptor (k_list_check ticks it)
    int* bin_ctr;        // [1] k_bin_scan arrival count (zero between builds)
    int maxcl, cap, ncell, ncol;
    int half;            // layout this call wants: 1 = each cluster pair in one
                         //   row (mixed pair kernel, j forces by atomics),
                         //   0 = both rows (exact and adjoint kernels)
};
// LS_LAYOUT: the `half` of the list last built (a list of the other layout
// is rebuilt, decided on the device so that captured graphs stay valid)
enum { LS_VALID = 0, LS_NEED = 1, LS_BUILDS = 2, LS_LAYOUT = 3, LS_N = 4 };

// frclong / RESPA-step finalize: nb (<= FIN_BLOCKS) workgroups of FIN_T
// threads per replica, enough for one partial of each kind per thread in the
// largest replica; their sums in part / partu, per-replica arrival counters
// (zero between calls)
#define FIN_BLOCKS 64
#define FIN_T 256
#define FIN_ND 17
// Per-step term histories of a device run, appended by the finalize pass
// of every RESPA step (one slot per step and replica); `on` is 0 outside a
// run that records them.  Lives in device memory at a fixed address, so a
// captured step graph appends to whatever histories the current run set.
// Histories of a device run: every step's finalize stores its workgroup
// sums in the step's slot of part / partu (the step's k_list_check advances
// *tick first), and k_hist_terms forms all steps' terms after the run, in
// the order the step's own last workgroup would.
struct HistDev {
    mspms_long_terms* lh;   // [cap][nrep] or null
    mspms_short_terms* sh;  // [cap][nrep] or null
    double* kh;             // [cap][nrep] or null
    int nrep, cap, on;
    double* part;               // [cap][nrep][nb][FIN_ND]
    unsigned long long* partu;  // [cap][nrep][nb][4]
    int nb;
    int* tick;                  // steps begun in the run
    double* uvac;               // [cap][nrep] each step's vacuum energy, or null
};

struct FinBufs {
    double* part;              // [nrep][nb][FIN_ND]
    unsigned long long* partu; // [nrep][nb][4]
    int* count;                // [nrep]
    int nb;
    const int* on;             // null, or a device flag: 0 = this pass only
                               //   clears the rebuild request (a step whose
                               //   terms nobody reads)
    const HistDev* hist;       // null, or the run's history descriptor
    int* hstep;                // [nrep] next history slot of each replica
};

struct Status {
    int flags;           // bit 0: non-finite position, bit 1: list overflow,
                         // bit 2: grid out of range
    int max_list;        // largest pair-list row seen
    int pad[2];
};

// slot_par.w flags
#define SLOT_VALID 1   // a site (not padding)
#define SLOT_LJ 2      // takes part in LJ (not GHOST_LJ)

#define ST_NONFINITE 1
#define ST_OVERFLOW 2
#define ST_GRID_EDOM 4

// ---- kernel launchers (k_*.hip) ----------------------------------------
struct LaunchCtx {
    hipStream_t stream;
    int natom, nmol, nsite, nrep;
    int max_mol_atoms;        // largest molecule (atoms) of any species
    int max_mol_terms;        // most bonded terms of one molecule of any species
    int pair_wpc;             // waves per i-cluster of the mixed pair kernel
    const SiteDev* sites;
    const MolDev* mols;
    const int* atom_mol;
    const int* atom_site;
    const RepDev* reps;
    Status* status;
    double* snap;             // null, or (6, natom): positions and velocities
                              //   at the start of the last RESPA step that
                              //   began without a pair-list overflow
    int* snap_count;          // steps that wrote the snapshot
    const double* rep_delt;   // null, or each replica's own outer time step
                              //   (HMC with delt adaptation): integrator
                              //   kernels then derive deltby2, delts and
                              //   deltsby2 from it as the host does
};

struct RecipBufs {
    KRow* rows;          // reference order, per replica ranges
    KRow* rows_s;        // same rows sorted by length (S pass)
    double* kvec;        // [nk_total] exp(-B k^2)/k^2 (host libm)
    int* k_rep;          // [nk_total]
    int4* work_s;        // (rep, chunk, first atom, atoms)
    int nwork_s;
    int4* work_f;        // (rep, first atom, first row, end row)
    int nwork_f;
    int* rep_chunk0;     // [nrep+1]
    double2* s_part;     // [nchunk][nk_stride]
    int nk_stride;
    double2* ak;         // [nk_total] force coefficients
    double* uk;          // [nk_total] kvec |S|^2
    int nk_total, ntab_max, nkz_max;
    int nrow;            // rows of all replicas
    int nkrow_max;       // most rows of one replica
    double2* eatom;      // [3][natom] exp(i 2 pi x_d / L_d) per atom and axis,
                         //   written by k_sk, read by k_kforce
};

// solid-fluid model (shared by all replicas), passed by value to k_sf
struct SFDev {
    int type, ntube, nsolid, ntype;
    const double* tube;      // [3*ntube] x, y, radius (reduced)
    const double* solid;     // [3*nsolid] x, y, z (reduced)
    double solid_sigma, solid_epsilon, c3pisq;
    int nx, ny, nz, pad;     // Hermite grid (Angstrom, K)
    double x0, y0, z0, hx, hy, hz;
    double zperiod;          // Hermite grid's z period (Angstrom)
    double sigma_base, epsilon_base;
    const float* nodes;      // [ntype][nz][ny][nx][8]: a node's F, Fx, Fy, Fxy,
                             //   Fz, Fxz, Fyz, Fxyz (derivative-bit order)
    const double* min_pe;    // [ntype] min node F (mylibtasos/grid.f90:102)
    // SF_MY_INTERP: nx, ny, nz = cubes per axis, x0, y0 = xmin, ymin
    double zmax;
    const double* cubes;     // [ntype][cube][4][32]
};

// RANMAR stream of one replica (src/random.c:14-109) and the cached second
// deviate of gaussran (src/random.c:113-138)
struct RanmarDev {
    double u[98];
    double c, cd, cm;
    int i97, j97;
    int iset, pad;
    double gset;
};

// batched HMC chain state of one replica
struct HmcRep {
    double treq, system_mass;   // reduced temperature, sum of the site masses
    int nfree, first;           // 3N - nconstraint; first move pending
    double ukin;                // kinetic energy of the current velocities
    double uinter_old, uintra_old, ukin_old, ushift_old;
    double last_dH;
    long long accepted, moves;
    // delt adaptation of the equilibration phase (src/hmc.c:47-58):
    // every adj_cycle moves delt *= 1 - rreq + accepted / adj_cycle
    // (adj_cycle 0: off); counts since the last adjustment
    double rreq;
    int adj_cycle, adj_moves, adj_accepted, pad;
};

// velinit of every replica (active: only where active[r] != 0; null: all)
void launch_hmc_init(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const double* aw_atom,
                     double* vel, const int* active = nullptr);
void launch_hmc_begin(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const double* aw_atom,
                      const double* pos, double* pos_old, double* vel);
void launch_hmc_mark(const LaunchCtx& c, HmcRep* hmc, const void* lterms, const void* sterms);
void launch_hmc_end(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const void* lterms,
                    const void* sterms, const double* ukin, double* pos, const double* pos_old,
                    double* rep_delt = nullptr);
void launch_hmc_draw(const LaunchCtx& c, RanmarDev* rng, int rep, int n, double* out);

// NPT state of one replica (k_npt.hip): the thermostat and barostat of
// npt_nh_operator, the LJ tail-pressure coefficients and the outputs
#define NPT_MAX_LRC 16
struct NptDev {
    double Qts, Qbs, treq, preq;
    double vts, vbs, rts;
    double pljlrc, boxv;
    double utsbs, ukin, tinst, pinst;
    double lrc[NPT_MAX_LRC];    // pljlrc_term[m][n] n_m n_n, (m, n) order
    int nlrc, nfree;
};
void launch_npt_op(const LaunchCtx& c, NptDev* st, bool end, double* vel, const double* fl,
                   const double* fs, const void* lterms, const void* sterms, double deltby2,
                   double deltsby2, double dt_outer4, double dt_outer8, double* ukin);
void launch_npt_inner(const LaunchCtx& c, RepDev* reps, NptDev* st, bool first, double* pos,
                      double* vel, const double* fs, double delts, double deltsby2);
void launch_kvec(const LaunchCtx& c, const RecipBufs& rb, int nrow);

// Nose-Hoover NVT (k_nvt.hip), one workgroup per replica
struct mspms_nvt_state;
void launch_nvt_drift(const LaunchCtx& c, mspms_nvt_state* st, bool first, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz, const double* fl,
                      const double* fs, double deltby2, double delts, double deltsby2,
                      double delts_sqby2, const int* active);
void launch_nvt_solve(const LaunchCtx& c, mspms_nvt_state* st, bool outer, double* carry,
                      double* vx, double* vy, double* vz, const double* f, double* vn,
                      double dt, double dtby2, double dt_sqby2, double* ukin,
                      const int* active);

// [type][128][ncube] (the file's order) -> [type][cube][128]
void launch_cube_major(hipStream_t s, const double* src, double* dst, int ntype, int ncube);
void launch_sf(const LaunchCtx& c, const SFDev& sf, const double* x,
               const double* y, const double* z, double* fx, double* fy,
               double* fz, double* usf);
// the Hermite grid's C-ABI layout [type][field][node] into the device's
// node-major [type][node][8] records (derivative-bit order)
void launch_grid_node_major(hipStream_t s, const float* src, float* dst, size_t nn, int ntype);

// Buffered (Verlet) cluster list: the list is built with rc_list = cutoff +
// skin and reused while every atom stays within skin/2 of its position at
// the build (force: rebuild regardless); reused calls only refresh the slot
// positions.  A rebuild request made here is cleared by the finalize kernel.
void launch_cluster_build(const LaunchCtx& c, ClusterBufs& b, const double* x,
                          const double* y, const double* z, double* fx,
                          double* fy, double* fz, double half_skin2, int force);
// gq/ge/gs: per-atom derivatives of the energy with respect to the atom's
// own charge, epsilon and sigma (parameter adjoint, FP64); null: none
void launch_pair(const LaunchCtx& c, const ClusterBufs& b, int precision,
                 int charge_type, double* fx, double* fy, double* fz,
                 double* epart, unsigned long long* ipart, int want_hash,
                 double* gq = nullptr, double* ge = nullptr,
                 double* gs = nullptr);
// The closing half kick of a RESPA step (src/md.c:75-80), fused into the
// last kernel that adds long forces when nothing adds after it: per atom
// v += deltby2 f / m and kin = m v.v (vx null: no kick).
struct Kick {
    double *vx, *vy, *vz, *kin;
    double deltby2;
    // null, or per-replica time steps: deltby2 = rep_delt[rep] / 2.0 with the
    // atom's replica from atom_mol / mols
    const double* rep_delt;
    const int* atom_mol;
    const MolDev* mols;
    const double* rmass;  // [natom] 1 / site mass
};
// A RESPA step's per-atom closing pass run by its finalize: fx += frec (the
// side stream's reciprocal and intramolecular forces) and the half kick,
// before the partials are reduced (also on gated-off steps)
// The half-list pair forces in slot order, added to the atoms' forces by the
// first per-atom pass after the pair kernel (fx[a] += f[gslot[a]] per plane).
struct SlotForces {
    const double* f;     // null: nothing to fold
    const int* gslot;
    size_t nslot;        // plane stride
};
struct FinAtoms {
    double *fx, *fy, *fz;
    const double* frec;  // (3, natom)
    Kick K;
    int natom;
    SlotForces S;
};
void launch_fold_slots(const LaunchCtx& c, const SlotForces& S, double* fx, double* fy,
                       double* fz);
void launch_intra(const LaunchCtx& c, const SpeciesDev* species,
                  const TermDev* terms, const double* x, const double* y,
                  const double* z, double* fx, double* fy, double* fz,
                  double* epart, double* gq = nullptr, double* ge = nullptr,
                  double* gs = nullptr, const double* frec = nullptr,
                  const Kick* kick = nullptr, bool zero_first = false);
// (zero_first: each molecule's thread sets its atoms' fx..fz to zero before
// adding, so fx..fz need no clearing pass)
// fx += (slot forces, when given) + frec ((3, N) reciprocal-space forces),
// and a capturable zero fill
void launch_add_frec(const LaunchCtx& c, double* fx, double* fy, double* fz,
                     const double* frec, const struct Kick* kick = nullptr,
                     const SlotForces* slots = nullptr);
void launch_zero(const LaunchCtx& c, double* p, size_t n);
void launch_bonded(const LaunchCtx& c, const SpeciesDev* species,
                   const TermDev* terms, const double* x, const double* y,
                   const double* z, double* fx, double* fy, double* fz,
                   double* epart);
int recip_nkz_supported(int kmaxz);
void launch_recip_s(const LaunchCtx& c, const RecipBufs& rb, const double* x,
                    const double* y, const double* z);
void launch_recip_f(const LaunchCtx& c, const RecipBufs& rb, const double* x,
                    const double* y, const double* z, double* fx, double* fy,
                    double* fz, double* gq = nullptr);
// per (replica, site): sum of the per-atom derivatives over the replica's
// atoms of that site, plus the self-energy term -2 q C coef (Ewald / Wolf)
// (work items: rep, global site, first atom of the site, molecule count;
// outputs [nrep][nsite])
void launch_grad_sites(const LaunchCtx& c, const SpeciesDev* species,
                       const int4* work, int nwork, const double* gq,
                       const double* ge, const double* gs, double* out_q,
                       double* out_e, double* out_s);
void launch_finalize_long(const LaunchCtx& c, const int* col_cl0,
                          const double* ep_pair,
                          const unsigned long long* ip_pair, int have_pair,
                          const double* ep_intra, int have_intra,
                          const double* uk, int have_recip,
                          const double* ep_sf, int have_sf, const double* uvac,
                          const double* ep_coul, void* terms, int* lstate,
                          const double* ep_short, void* sterms, const double* kin,
                          double* ukin, const FinBufs& fb,
                          const struct FinAtoms* atoms = nullptr);
// simple Coulomb between molecules (iChargeType 3): COMs into com[nmol],
// per-molecule energy partials into ep[nmol]; molecules of <= 16 sites
int coulomb_mol_supported(int max_mol_atoms);
void launch_coulomb_mol(const LaunchCtx& c, const SpeciesDev* species, const double* x,
                        const double* y, const double* z, double4* com, double* fx,
                        double* fy, double* fz, double* ep);
// Ewald vacuum boundary (src/frclong.c:1065-1126): per replica the dipole
// of the COM-wrapped molecules, uvac[rep], and its force; gq (adjoint,
// optional) receives d uvacuum / d q_i
void launch_vacuum(const LaunchCtx& c, const SpeciesDev* species, const double* x,
                   const double* y, const double* z, double* fx, double* fy, double* fz,
                   double* uvac, double* gq);
// sets a device run's history descriptor and clears the per-replica slots
void launch_hist_terms(const LaunchCtx& c, const HistDev& h, int nsteps, void* lterms,
                       void* sterms, double* ukin);
void launch_set_hist(const LaunchCtx& c, HistDev* d, const HistDev& v, int* hstep);
void launch_finalize_short(const LaunchCtx& c, const double* epart, void* terms,
                           const double* ep_sf = nullptr);
void launch_respa_inner(const LaunchCtx& c, const SpeciesDev* species,
                        const TermDev* terms, double* x, double* y, double* z,
                        double* vx, double* vy, double* vz, const double* fxl,
                        const double* fyl, const double* fzl, double* fxs,
                        double* fys, double* fzs, double deltby2, double delts,
                        double deltsby2, int nstep_inner, double* epart,
                        const SFDev* sf = nullptr, double* usf = nullptr);
// (sf: the solid-fluid wall added to the short-range forces in every inner
// step, the last step's per-site energies in usf)
// closing half kick; kin[i] = v.v / r_mass per atom (the step's finalize
// kernel reduces it into ukin)
void launch_respa_final(const LaunchCtx& c, double* vx, double* vy, double* vz,
                        const double* fxl, const double* fyl,
                        const double* fzl, double deltby2, double* kin);
