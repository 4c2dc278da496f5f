// engine.cpp -- host side of the MI355X force engine: the C ABI of
// include/mspms_gpu.h.  Builds the device tables (sites, molecules,
// replicas, bonded terms, Ewald k rows and work lists), owns the device
// buffers and enqueues the kernels of k_*.hip on one HIP stream.
//
// Reference boundary: src/funcs.h:225,233 (int frclong(), int frcshort())
// and the globals they read and write (src/molevars.h, src/vars.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mspms_gpu.h"
#include "engine.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPC(expr)                                                                \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(MSPMS_EIO, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

const double PI_REF = 3.141592653589793;  // src/defs.h:100

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int ensure(size_t count) {
        if (count <= n && p) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
            return fail(MSPMS_ENOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
        n = count;
        return 0;
    }
    int upload(const std::vector<T>& v, hipStream_t s) {
        int rc = ensure(v.size());
        if (rc) return rc;
        if (!v.empty()) HIPC(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct SpeciesHost {
    int natom, nmole;
    std::vector<int> ghost, interp;
    std::vector<double> aw, sigma, epsilon, charge;
    std::vector<int> bnd_idx, bnd_type;
    std::vector<double> bnd_para;
    std::vector<int> agl_idx, agl_type;
    std::vector<double> agl_para;
    std::vector<uint8_t> agl_unique;
    std::vector<int> dih_idx, dih_type;
    std::vector<double> dih_para;
    std::vector<uint8_t> dih_unique;
    std::vector<int> imp_idx, imp_type;
    std::vector<double> imp_para;
    int nbond() const { return (int)bnd_type.size(); }
    int nangle() const { return (int)agl_type.size(); }
    int ndih() const { return (int)dih_type.size(); }
    int nimp() const { return (int)imp_type.size(); }
};

template <class T>
std::vector<T> copy_arr(const T* p, size_t n) {
    return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T());
}

}  // namespace

// the error path of the other translation units (mspms_last_error)
__attribute__((visibility("hidden"))) int mspms_fail(int code, const char* what) {
    return fail(code, "%s", what);
}

// -------------------------------------------------------------------------
struct mspms_ctx {
    int device = 0, precision = MSPMS_PREC_MIXED;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t side = nullptr;          // concurrent branch of frclong
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    std::vector<SpeciesHost> sp;
    int nrep = 0;
    std::vector<int> nmole;              // [nrep*nspecie]
    std::vector<mspms_params> params;    // [nrep]
    std::vector<std::vector<double>> tube_keep;
    int natom = 0, nmol = 0, nsite = 0;
    int charge_union = MSPMS_ELEC_NONE;  // charge type shared by the replicas
    bool any_intra = false, any_bonded = false, any_recip = false, any_sf = false;
    int max_mol_terms = 0;  // most bonded terms of one molecule (any species)
    bool sf_split = false;  // the solid-fluid wall in frcshort / the inner loop
    bool any_vacuum = false, any_coul = false;
    int want_hash = -1;                  // -1: FP64 only
    // launch shape, set by mspms_autotune: sites per binning cell and waves
    // per i-cluster of the mixed pair kernel
    double cell_atoms = 8.0;
    int pair_wpc = 4;
    // pair-list buffer (mspms_set_list_skin): lists are built with cutoff +
    // skin and reused while every atom stays within skin / 2 of its
    // position at the build; force_rebuild: rebuild on every call
    double skin = 0.0;
    bool force_rebuild = false;

    // static device tables
    DevBuf<SiteDev> d_sites;
    DevBuf<SpeciesDev> d_species;
    DevBuf<TermDev> d_terms;
    DevBuf<MolDev> d_mols;
    DevBuf<int> d_atom_mol, d_atom_site;
    DevBuf<RepDev> d_reps;
    std::vector<RepDev> reps;

    // clustering
    ClusterBufs cb{};
    DevBuf<int> d_atom_cell, d_atom_slot, d_cell_count, d_cell_start, d_cell_members,
        d_col_cell0, d_col_rep, d_col_cl0, d_cl_col, d_ncl_total, d_slot_orig, d_plist_n,
        d_bin_ctr;
    DevBuf<double> d_atom_wz;
    DevBuf<int2> d_slot_info;
    DevBuf<double4> d_slot_x64, d_slot_w64, d_cl_center, d_atom_x0;
    DevBuf<int> d_lstate;
    DevBuf<float4> d_slot_off, d_slot_par, d_cl_ext, d_cl_box;
    DevBuf<float4> d_plist;
    DevBuf<int> d_atom_gslot;
    DevBuf<double2> d_eatom;
    DevBuf<double> d_fslot;
    int cap_hint = 0;
    int cap_max = 1;  // no row holds more entries than a replica has clusters
    int cap_limit = 0;  // caller's bound on the row capacity (0: none)

    // reciprocal
    RecipBufs rb{};
    DevBuf<KRow> d_rows, d_rows_s;
    DevBuf<double> d_kvec, d_uk;
    DevBuf<int> d_k_rep, d_rep_chunk0;
    DevBuf<int4> d_work_s, d_work_f;
    DevBuf<double2> d_s_part, d_ak;

    // partial sums and outputs
    DevBuf<double> d_ep_pair, d_ep_intra, d_ep_short, d_ep_sf, d_uvac, d_ep_coul;
    DevBuf<double4> d_mcom;
    FinBufs fin{};
    DevBuf<double> d_fin_part;
    DevBuf<unsigned long long> d_fin_partu;
    DevBuf<int> d_fin_count;
    DevBuf<int> d_fin_on;    // gate of the RESPA step's finalize (1: run)
    DevBuf<double> d_snap;   // (6, natom) start of the last step begun
                             //   without a pair-list overflow (RESPA runs)
    DevBuf<int> d_snap_count;  // steps that wrote it since the last sync
    double* snap_pos = nullptr;  // the buffers of the run it belongs to
    double* snap_vel = nullptr;
    DevBuf<HistDev> d_hist;  // per-step term histories of a device run
    DevBuf<double> d_hpart;  // their workgroup sums, [steps][nrep][nb][FIN_ND]
    DevBuf<unsigned long long> d_hpartu;
    DevBuf<double> d_huvac;  // each step's vacuum energies, [steps][nrep]
    DevBuf<int> d_hstep;     // [nrep] next history slot

    // solid-fluid tables
    SFDev sf{};
    DevBuf<double> d_sf_tube, d_sf_solid, d_sf_minpe;
    DevBuf<float> d_sf_nodes;
    size_t sf_nodes_uploaded = 0;
    const float* sf_nodes_src = nullptr;
    DevBuf<double> d_sf_cubes;                  // [ntype][cube][128]
    std::vector<const double*> sf_cubes_src;    // the tables last uploaded
    DevBuf<unsigned long long> d_ip_pair;
    DevBuf<mspms_long_terms> d_long;
    DevBuf<mspms_short_terms> d_short;
    DevBuf<Status> d_status;
    Status* h_status = nullptr;  // pinned

    // host-API staging and integrator scratch
    DevBuf<double> d_pos, d_fl, d_fs, d_kin;
    DevBuf<double> d_frec;               // (3, N) reciprocal forces (side stream)
    // parameter adjoint: per-atom derivatives, per-(replica, site) sums
    DevBuf<double> d_grad, d_grad_out;
    DevBuf<int4> d_grad_work;

    // profiling: a pool of events, one (start, stop) pair per recorded
    // kernel group, read back at the next synchronisation
    bool profiling = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<int> pending;          // group of each recorded pair
    std::vector<double> prof_ms;       // summed milliseconds per group
    std::vector<long> prof_n;          // recorded launches per group
    bool capturing = false;            // no event records inside a capture

    // one RESPA outer step captured as a graph (mspms_vver_run_device),
    // valid for the argument set in gkey until the next build
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_kind = -1;                 // GRAPH_*: what the cached graph runs

    // batched HMC chains (mspms_hmc_init)
    DevBuf<RanmarDev> d_rng;
    DevBuf<HmcRep> d_hmc;
    DevBuf<double> d_aw_atom, d_pos_old;
    DevBuf<double> d_rmass;  // [natom] 1 / site mass (the kicks' factor)
    bool hmc_ready = false;
    DevBuf<double> d_hmc_delt;          // [nrep] each chain's outer step
    bool hmc_rep_delt = false;          // chains step with d_hmc_delt
    const double* cur_rep_delt = nullptr;  // integrator kernels' per-replica
                                           // steps while an HMC move is enqueued

    // Nose-Hoover NVT (mspms_nvt_set): states, the last solve's sum m v^2,
    // the Newton iterate (3, N)
    DevBuf<mspms_nvt_state> d_nvt;
    DevBuf<double> d_nvt_carry, d_nvt_vn;
    DevBuf<int> d_active;                // replicas the NVT kernels advance
    bool nvt_ready = false;
    // NPT (mspms_npt_set)
    DevBuf<NptDev> d_npt;
    bool npt_ready = false;
    struct GraphKey {
        void *pos, *vel, *fl, *fs, *lt, *st, *uk;
        double delt;
        int nstep_inner;
        int per = 1;  // steps per graph
        bool operator==(const GraphKey& o) const {
            return pos == o.pos && vel == o.vel && fl == o.fl && fs == o.fs && lt == o.lt &&
                   st == o.st && uk == o.uk && delt == o.delt && nstep_inner == o.nstep_inner &&
                   per == o.per;
        }
    } gkey{};
};

namespace {

void drop_graph(mspms_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

enum { GRAPH_VVER = 0, GRAPH_HMC = 1, GRAPH_NVT = 2, GRAPH_NPT = 3 };

// Makes the cached graph the capture of enqueue(stream) for (kind, key),
// capturing it again when either differs.
template <typename F>
int ensure_graph(mspms_ctx* c, hipStream_t s, int kind, const mspms_ctx::GraphKey& key,
                 F&& enqueue) {
    if (c->graph_exec && key == c->gkey && c->graph_kind == kind) return 0;
    drop_graph(c);
    // the capture stream must not have work in flight from the caller
    HIPC(hipStreamSynchronize(s));
    HIPC(hipStreamSynchronize(c->stream));
    c->capturing = true;
    HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(c->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->capturing = false;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(MSPMS_EIO, "graph capture: %s", hipGetErrorString(e));
    c->graph = g;
    e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(c);
        return fail(MSPMS_EIO, "graph instantiate: %s", hipGetErrorString(e));
    }
    c->gkey = key;
    c->graph_kind = kind;
    return 0;
}

LaunchCtx launch_ctx(mspms_ctx* c, hipStream_t s) {
    LaunchCtx L;
    L.stream = s;
    L.natom = c->natom;
    L.nmol = c->nmol;
    L.nsite = c->nsite;
    L.nrep = c->nrep;
    L.pair_wpc = c->pair_wpc;
    L.max_mol_atoms = 0;
    for (const SpeciesHost& q : c->sp) L.max_mol_atoms = std::max(L.max_mol_atoms, q.natom);
    L.max_mol_terms = c->max_mol_terms;
    L.rep_delt = c->cur_rep_delt;
    L.snap = nullptr;
    L.snap_count = nullptr;
    L.sites = c->d_sites.p;
    L.mols = c->d_mols.p;
    L.atom_mol = c->d_atom_mol.p;
    L.atom_site = c->d_atom_site.p;
    L.reps = c->d_reps.p;
    L.status = c->d_status.p;
    return L;
}

// InitCheckUniques (src/initvars.c:13-146) including its index quirks:
// the dihedral-vs-angle and dihedral-vs-bond tests mark isDih_unique by the
// angle / bond index (:66, :86), the angle-vs-bond test marks isAngle_unique
// by the bond index (:134).  Marks past the current counts are dropped.
void check_uniques(const int* agl_idx, int nangle, const int* dih_idx, int ndih,
                   const int* bnd_idx, int nbond, uint8_t* au, uint8_t* du) {
    std::vector<uint8_t> dih(std::max({ndih, nangle, nbond, 1}), 1),
        agl(std::max({nangle, nbond, 1}), 1);
    for (int i = 0; i < ndih - 1; i++) {
        if (!dih[i]) continue;
        for (int j = i + 1; j < ndih; j++) {
            if (!dih[j]) continue;
            if ((dih_idx[4 * i] == dih_idx[4 * j] && dih_idx[4 * i + 3] == dih_idx[4 * j + 3]) ||
                (dih_idx[4 * i] == dih_idx[4 * j + 3] && dih_idx[4 * i + 3] == dih_idx[4 * j]))
                dih[j] = 0;
        }
    }
    for (int i = 0; i < ndih; i++) {
        if (!dih[i]) continue;
        for (int j = 0; j < nangle; j++)
            if ((dih_idx[4 * i] == agl_idx[3 * j] && dih_idx[4 * i + 3] == agl_idx[3 * j + 2]) ||
                (dih_idx[4 * i] == agl_idx[3 * j + 2] && dih_idx[4 * i + 3] == agl_idx[3 * j]))
                dih[j] = 0;
    }
    for (int i = 0; i < ndih; i++) {
        if (!dih[i]) continue;
        for (int j = 0; j < nbond; j++)
            if ((dih_idx[4 * i] == bnd_idx[2 * j] && dih_idx[4 * i + 3] == bnd_idx[2 * j + 1]) ||
                (dih_idx[4 * i] == bnd_idx[2 * j + 1] && dih_idx[4 * i + 3] == bnd_idx[2 * j]))
                dih[j] = 0;
    }
    for (int i = 0; i < nangle - 1; i++) {
        if (!agl[i]) continue;
        for (int j = i + 1; j < nangle; j++) {
            if (!agl[j]) continue;
            if ((agl_idx[3 * i] == agl_idx[3 * j] && agl_idx[3 * i + 2] == agl_idx[3 * j + 2]) ||
                (agl_idx[3 * i] == agl_idx[3 * j + 2] && agl_idx[3 * i + 2] == agl_idx[3 * j]))
                agl[j] = 0;
        }
    }
    for (int i = 0; i < nangle; i++) {
        if (!agl[i]) continue;
        for (int j = 0; j < nbond; j++)
            if ((agl_idx[3 * i] == bnd_idx[2 * j] && agl_idx[3 * i + 2] == bnd_idx[2 * j + 1]) ||
                (agl_idx[3 * i] == bnd_idx[2 * j + 1] && agl_idx[3 * i + 2] == bnd_idx[2 * j]))
                agl[j] = 0;
    }
    for (int i = 0; i < nangle; i++) au[i] = agl[i];
    for (int i = 0; i < ndih; i++) du[i] = dih[i];
}

int validate_species(const mspms_species& q) {
    if (q.natom < 1 || q.natom > 256) return fail(MSPMS_EINVAL, "natom %d out of 1..256", q.natom);
    if (q.nmole < 0) return fail(MSPMS_EINVAL, "negative nmole");
    if (q.nnbp != 0)
        return fail(MSPMS_EINVAL,
                    "non-bonded pair lists are not supported: src/frclong.c:793 indexes "
                    "nbp_idx with the dihedral counter");
    for (int k = 0; k < q.nbond; k++) {
        int t = q.bnd_type[k];
        if (t < MSPMS_BOND_HARMONIC || t > MSPMS_BOND_FENE)
            return fail(MSPMS_EINVAL, "bond type %d unsupported (reference: unknown or "
                                      "uninitialised energy)", t);
    }
    for (int k = 0; k < q.nangle; k++) {
        int t = q.agl_type[k];
        if (t < MSPMS_ANGLE_HARMONIC || t > MSPMS_ANGLE_TR_WATER)
            return fail(MSPMS_EINVAL, "angle type %d unsupported", t);
    }
    for (int k = 0; k < q.ndih; k++) {
        int t = q.dih_type[k];
        if (t < MSPMS_DIH_OPLS_COSIN || t > MSPMS_DIH_CHARMM)
            return fail(MSPMS_EINVAL, "dihedral type %d unsupported", t);
        if (t == MSPMS_DIH_CHARMM && !(std::fabs(q.dih_para[4 * k]) >= 1.0e-8))
            return fail(MSPMS_EINVAL, "CHARMM dihedral with |kphi| < TOLERANCE leaves the "
                                      "reference's energy uninitialised");
    }
    for (int k = 0; k < q.nimp; k++) {
        if (q.imp_type[k] != MSPMS_IMP_CHARMM)
            return fail(MSPMS_EINVAL, "improper type %d unsupported", q.imp_type[k]);
        if (!(std::fabs(q.imp_para[2 * k]) >= 1e-5))
            return fail(MSPMS_EINVAL, "CHARMM improper with |komega| < 1e-5 leaves the "
                                      "reference's energy uninitialised");
    }
    if (q.ndih > 0 && q.nbond == 0 && q.nangle == 0)
        return fail(MSPMS_EINVAL, "dihedrals without bonds or angles: the 1-4 section "
                                  "would read an unset isSameMole");
    auto chk_idx = [&](const int* idx, int n, int w) {
        for (int k = 0; k < n * w; k++)
            if (idx[k] < 0 || idx[k] >= q.natom) return false;
        return true;
    };
    if (!chk_idx(q.bnd_idx, q.nbond, 2) || !chk_idx(q.agl_idx, q.nangle, 3) ||
        !chk_idx(q.dih_idx, q.ndih, 4) || !chk_idx(q.imp_idx, q.nimp, 4))
        return fail(MSPMS_EINVAL, "term index outside the molecule");
    return 0;
}

int validate_params(const mspms_params& p) {
    if (!(p.boxlx > 0 && p.boxly > 0 && p.boxlz > 0))
        return fail(MSPMS_EINVAL, "invalid box %g %g %g", p.boxlx, p.boxly, p.boxlz);
    double fmin = std::min(p.boxlx, std::min(p.boxly, p.boxlz)) / 2.0;
    if (p.rcuton > fmin || p.rcutoff > fmin || p.rcutoffelec > fmin)
        return fail(MSPMS_ERANGE, "cutoff larger than half the box (src/validation.c:617)");
    if (p.iExternal_FF_type != 0)
        return fail(MSPMS_EINVAL, "external force fields (metal cluster) are not supported");
    if (p.iChargeType < 0 || p.iChargeType > 3) return fail(MSPMS_EINVAL, "bad iChargeType");
    if (p.iChargeType == MSPMS_ELEC_EWALD) {
        if (p.kmaxx < 0 || p.kmaxy < 0 || p.kmaxz < 0 || p.kmaxx > 16 || p.kmaxy > 16 ||
            p.kmaxz > 16)
            return fail(MSPMS_EINVAL, "KMAX must be within 0..16");
    }
    if (p.iSF_type < MSPMS_SF_NONE || p.iSF_type > MSPMS_SF_MY_INTERP)
        return fail(MSPMS_EINVAL, "solid-fluid type %d unsupported", p.iSF_type);
    if (p.iSF_type == MSPMS_SF_HYPERGEO && (p.ntube < 1 || !p.tube_x || !p.tube_y || !p.tube_r))
        return fail(MSPMS_EINVAL, "hypergeometric tube needs ntube >= 1 and tube arrays");
    if (p.iSF_type == MSPMS_SF_ATOM_EXPLICIT && p.solid_natom > 0 &&
        (!p.solid_xx || !p.solid_yy || !p.solid_zz))
        return fail(MSPMS_EINVAL, "explicit solid needs solid_xx/yy/zz");
    if (p.iSF_type == MSPMS_SF_TASOS &&
        (!p.grid || !p.grid->nodes || p.grid->ntype < 1 || p.grid->nx < 2 || p.grid->ny < 2 ||
         p.grid->nz < 2 || !(p.grid->hx > 0) || !(p.grid->hy > 0) || !(p.grid->hz > 0) ||
         !(p.grid->zperiod >= 0.0) || p.grid->zperiod > (p.grid->nz - 1) * p.grid->hz))
        return fail(MSPMS_EINVAL, "SF grid missing or malformed");
    if (p.iSF_type == MSPMS_SF_MY_INTERP) {
        const mspms_sf_cubes* q = p.cubes;
        // NUNIQUE_ATOM_MAX = 5 grid files (src/defs.h:78)
        if (!q || !q->tables || q->ntype < 1 || q->ntype > 5 || q->ncx < 1 || q->ncy < 1 ||
            q->ncz < 1 || !(q->hx > 0) || !(q->hy > 0) || !(q->hz > 0) || !(q->zmax > 0) ||
            (double)q->ncx * q->ncy * q->ncz > 2147483647.0)
            return fail(MSPMS_EINVAL, "myinterp cube tables missing or malformed");
        for (int k = 0; k < 128 * q->ntype; k++)
            if (!q->tables[k]) return fail(MSPMS_EINVAL, "myinterp table %d is null", k);
    }
    return 0;
}

}  // namespace

// -------------------------------------------------------------------------
extern "C" const char* mspms_last_error(void) { return g_err.c_str(); }

extern "C" int mspms_check_uniques(const mspms_species* sp, uint8_t* au, uint8_t* du) {
    if (!sp) return fail(MSPMS_EINVAL, "null species");
    check_uniques(sp->agl_idx, sp->nangle, sp->dih_idx, sp->ndih, sp->bnd_idx, sp->nbond, au, du);
    return 0;
}

// src/initvars.c:591-599 and :433-462
extern "C" int mspms_derive_params(mspms_params* p) {
    if (!p) return fail(MSPMS_EINVAL, "null params");
    volatile double rc = p->rcutoff;  // keep the reference's evaluation order
    p->rcutoffsq = p->rcutoff * p->rcutoff;
    p->rcutoffelecsq = p->rcutoffelec * p->rcutoffelec;
    p->rcutonsq = p->rcuton * p->rcuton;
    p->roff2_minus_ron2_cube = (p->rcutoffsq - p->rcutonsq) * (p->rcutoffsq - p->rcutonsq) *
                               (p->rcutoffsq - p->rcutonsq);
    p->shift1 = std::pow((1 / rc), 6.0);
    p->shift4 = 4.0 * p->shift1;
    if (p->iChargeType == MSPMS_ELEC_EWALD) {
        p->bfactor_ewald = 1.0 / (4.0 * p->kappa * p->kappa);
        p->vfactor_ewald = 2.0 * PI_REF / (p->boxlx * p->boxly * p->boxlz);
        p->twopi_lx = 2.0 * PI_REF / p->boxlx;
        p->twopi_ly = 2.0 * PI_REF / p->boxly;
        p->twopi_lz = 2.0 * PI_REF / p->boxlz;
        p->twopi_over_3v = 2.0 * PI_REF / 3.0 / p->boxlx / p->boxly / p->boxlz;
    } else if (p->iChargeType == MSPMS_ELEC_WOLF) {
        double k = p->kappa, rce = p->rcutoffelec;
        p->wolfvcon1 = -std::erfc(k * rce) / rce;
        p->wolfvcon2 = std::erfc(k * rce) / p->rcutoffelecsq +
                       2.0 * k * std::exp(-(k * rce) * (k * rce)) / (std::sqrt(PI_REF) * rce);
        p->wolfvcon3 = k / std::sqrt(PI_REF) + std::erfc(k * rce) / (2.0 * rce);
        p->wolffcon1 = 2.0 * k / std::sqrt(PI_REF);
        p->wolffcon2 = -p->wolfvcon2;
    }
    if (p->iChargeType != MSPMS_ELEC_NONE)
        p->coulomb_prefactor = 167100.80840 / p->epsilon_base / p->sigma_base;
    double fmin = std::min(p->boxlx, std::min(p->boxly, p->boxlz)) / 2.0;
    if (p->rcuton > fmin || p->rcutoff > fmin || p->rcutoffelec > fmin)
        return fail(MSPMS_ERANGE, "cutoff larger than half the box");
    return 0;
}

extern "C" int mspms_create(mspms_ctx** out, int device, int precision) {
    if (!out) return fail(MSPMS_EINVAL, "null out");
    if (precision != MSPMS_PREC_FP64 && precision != MSPMS_PREC_MIXED)
        return fail(MSPMS_EINVAL, "unknown precision %d", precision);
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MSPMS_EINVAL, "no HIP device %d", device);
    HIPC(hipSetDevice(device));
    mspms_ctx* c = new mspms_ctx();
    c->device = device;
    c->precision = precision;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MSPMS_EIO, "hipStreamCreate failed");
    }
    c->own_stream = true;
    // side stream (default priority) for the reciprocal S(k) pass, which
    // overlaps the clustering and the pair kernel; joined by events
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        mspms_destroy(c);
        return fail(MSPMS_EIO, "hipStreamCreate / hipEventCreate failed");
    }
    if (c->d_status.ensure(1) ||
        hipHostMalloc((void**)&c->h_status, sizeof(Status), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(MSPMS_ENOMEM, "status allocation failed");
    }
    memset(c->h_status, 0, sizeof(Status));
    (void)hipMemcpy(c->d_status.p, c->h_status, sizeof(Status), hipMemcpyHostToDevice);
    *out = c;
    return 0;
}

extern "C" int mspms_destroy(mspms_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    c->d_sites.release();
    c->d_species.release();
    c->d_terms.release();
    c->d_mols.release();
    c->d_atom_mol.release();
    c->d_atom_site.release();
    c->d_reps.release();
    c->d_atom_cell.release();
    c->d_atom_slot.release();
    c->d_cell_count.release();
    c->d_cell_start.release();
    c->d_cell_members.release();
    c->d_col_cell0.release();
    c->d_col_rep.release();
    c->d_col_cl0.release();
    c->d_cl_col.release();
    c->d_ncl_total.release();
    c->d_slot_orig.release();
    c->d_plist_n.release();
    c->d_bin_ctr.release();
    c->d_atom_wz.release();
    c->d_slot_info.release();
    c->d_slot_x64.release();
    c->d_slot_w64.release();
    c->d_atom_x0.release();
    c->d_lstate.release();
    c->d_cl_center.release();
    c->d_slot_off.release();
    c->d_slot_par.release();
    c->d_cl_ext.release();
    c->d_cl_box.release();
    c->d_plist.release();
    c->d_atom_gslot.release();
    c->d_eatom.release();
    c->d_fslot.release();
    c->d_rows.release();
    c->d_rows_s.release();
    c->d_kvec.release();
    c->d_uk.release();
    c->d_k_rep.release();
    c->d_rep_chunk0.release();
    c->d_work_s.release();
    c->d_work_f.release();
    c->d_s_part.release();
    c->d_ak.release();
    c->d_ep_pair.release();
    c->d_ep_intra.release();
    c->d_ep_short.release();
    c->d_ep_sf.release();
    c->d_uvac.release();
    c->d_ep_coul.release();
    c->d_mcom.release();
    c->d_fin_part.release();
    c->d_fin_partu.release();
    c->d_fin_count.release();
    c->d_hpart.release();
    c->d_hpartu.release();
    c->d_huvac.release();
    c->d_ip_pair.release();
    c->d_long.release();
    c->d_short.release();
    c->d_status.release();
    c->d_pos.release();
    c->d_fl.release();
    c->d_fs.release();
    c->d_kin.release();
    c->d_frec.release();
    c->d_grad.release();
    c->d_grad_out.release();
    c->d_grad_work.release();
    c->d_rng.release();
    c->d_hmc.release();
    c->d_aw_atom.release();
    c->d_rmass.release();
    c->d_pos_old.release();
    c->d_hmc_delt.release();
    c->d_nvt.release();
    c->d_npt.release();
    c->d_nvt_carry.release();
    c->d_nvt_vn.release();
    c->d_active.release();
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

namespace {

// Builds every device table from sp / nmole / params.
// each atom's 1 / mass, as the kicks divide (the same division, once)
int upload_rmass(mspms_ctx* c, const std::vector<SiteDev>& sites,
                 const std::vector<int>& atom_site, hipStream_t s) {
    std::vector<double> rm(atom_site.size());
    for (size_t a = 0; a < atom_site.size(); a++) rm[a] = 1.0 / sites[atom_site[a]].aw;
    return c->d_rmass.upload(rm, s);
}

int build(mspms_ctx* c) {
    HIPC(hipSetDevice(c->device));
    drop_graph(c);  // buffers may move
    const int ns = (int)c->sp.size();
    const hipStream_t s = c->stream;
    // ---- sites, species, terms ----
    std::vector<SiteDev> sites;
    std::vector<SpeciesDev> spd;
    std::vector<TermDev> terms;
    c->any_intra = c->any_bonded = false;
    c->max_mol_terms = 0;
    for (int k = 0; k < ns; k++) {
        const SpeciesHost& q = c->sp[k];
        SpeciesDev d{};
        d.natom = q.natom;
        d.site0 = (int)sites.size();
        d.mw = 0.0;
        for (int a = 0; a < q.natom; a++) d.mw += q.aw[a];  // src/readins.c:115,138
        for (int a = 0; a < q.natom; a++) {
            SiteDev t{};
            t.sigma = q.sigma[a];
            t.epsilon = q.epsilon[a];
            t.charge = q.charge[a];
            t.aw = q.aw[a];
            t.sqrt_eps = std::sqrt(q.epsilon[a]);
            t.sigma_f = (float)q.sigma[a];
            t.sqrt_eps_f = (float)std::sqrt(q.epsilon[a]);
            t.charge_f = (float)q.charge[a];
            t.half_sigma_f = (float)(0.5 * q.sigma[a]);
            t.ghost = q.ghost[a];
            t.species = k;
            t.site = a;
            t.interp = q.interp[a];
            sites.push_back(t);
        }
        // frclong intramolecular pairs: bonds, unique angles, unique dihedrals
        d.term0 = (int)terms.size();
        for (int b = 0; b < q.nbond(); b++) {
            TermDev t{};
            t.kind = TERM_PAIR12;
            t.idx[0] = q.bnd_idx[2 * b];
            t.idx[1] = q.bnd_idx[2 * b + 1];
            terms.push_back(t);
        }
        for (int a = 0; a < q.nangle(); a++) {
            if (!q.agl_unique[a]) continue;
            TermDev t{};
            t.kind = TERM_PAIR13;
            t.idx[0] = q.agl_idx[3 * a];
            t.idx[1] = q.agl_idx[3 * a + 2];
            terms.push_back(t);
        }
        for (int h = 0; h < q.ndih(); h++) {
            if (!q.dih_unique[h]) continue;
            TermDev t{};
            t.kind = TERM_PAIR14;
            t.idx[0] = q.dih_idx[4 * h];
            t.idx[1] = q.dih_idx[4 * h + 3];
            terms.push_back(t);
        }
        d.nintra = (int)terms.size() - d.term0;
        // frcshort terms in src/frcshort.c:47-79 order
        d.bterm0 = (int)terms.size();
        for (int b = 0; b < q.nbond(); b++) {
            TermDev t{};
            t.kind = TERM_BOND;
            t.type = q.bnd_type[b];
            t.idx[0] = q.bnd_idx[2 * b];
            t.idx[1] = q.bnd_idx[2 * b + 1];
            for (int p = 0; p < 3; p++) t.para[p] = q.bnd_para[3 * b + p];
            terms.push_back(t);
        }
        for (int a = 0; a < q.nangle(); a++) {
            TermDev t{};
            t.kind = TERM_ANGLE;
            t.type = q.agl_type[a];
            for (int p = 0; p < 3; p++) t.idx[p] = q.agl_idx[3 * a + p];
            for (int p = 0; p < 5; p++) t.para[p] = q.agl_para[5 * a + p];
            terms.push_back(t);
        }
        for (int h = 0; h < q.ndih(); h++) {
            TermDev t{};
            t.kind = TERM_DIH;
            t.type = q.dih_type[h];
            for (int p = 0; p < 4; p++) t.idx[p] = q.dih_idx[4 * h + p];
            for (int p = 0; p < 4; p++) t.para[p] = q.dih_para[4 * h + p];
            terms.push_back(t);
        }
        for (int m = 0; m < q.nimp(); m++) {
            TermDev t{};
            t.kind = TERM_IMP;
            t.type = q.imp_type[m];
            for (int p = 0; p < 4; p++) t.idx[p] = q.imp_idx[4 * m + p];
            for (int p = 0; p < 2; p++) t.para[p] = q.imp_para[2 * m + p];
            terms.push_back(t);
        }
        d.nbterm = (int)terms.size() - d.bterm0;
        c->max_mol_terms = std::max(c->max_mol_terms, d.nbterm);
        if (d.nintra > 0) c->any_intra = true;
        if (d.nbterm > 0) c->any_bonded = true;
        spd.push_back(d);
    }
    c->nsite = (int)sites.size();

    // ---- atoms, molecules, replicas ----
    std::vector<MolDev> mols;
    std::vector<int> atom_mol, atom_site;
    std::vector<RepDev> reps(c->nrep);
    std::vector<int> col_cell0(1, 0), col_rep;
    int cell0 = 0, col0 = 0;
    c->charge_union = c->params[0].iChargeType;
    c->any_recip = false;
    c->any_vacuum = false;
    c->any_coul = c->charge_union == MSPMS_ELEC_SIMPLE_COULOMB;
    if (c->any_coul) {
        int mx = 0;
        for (const SpeciesHost& q : c->sp) mx = std::max(mx, q.natom);
        if (!coulomb_mol_supported(mx))
            return fail(MSPMS_EINVAL, "simple Coulomb supports molecules of at most 16 sites");
    }
    c->any_sf = false;
    std::vector<KRow> rows, rows_s;
    std::vector<double> kvec;
    std::vector<int> k_rep, rep_chunk0(1, 0);
    std::vector<int4> work_s, work_f;
    int ntab_max = 1, nkz_max = 1, nk_stride = 1, nkrow_max = 1;
    size_t max_rep_cl = 1;
    double max_est = 0.0;
    for (int r = 0; r < c->nrep; r++) {
        const mspms_params& P = c->params[r];
        if (P.iChargeType != c->charge_union)
            return fail(MSPMS_EINVAL, "all replicas must share one electrostatics type");
        RepDev& R = reps[r];
        memset(&R, 0, sizeof R);
        R.atom0 = (int)atom_mol.size();
        R.mol0 = (int)mols.size();
        double q2 = 0.0;
        int nreal = 0;
        for (int k = 0; k < ns; k++) {
            const SpeciesHost& q = c->sp[k];
            int nm = c->nmole[(size_t)r * ns + k];
            for (int m = 0; m < nm; m++) {
                MolDev M;
                M.atom0 = (int)atom_mol.size();
                M.species = k;
                M.rep = r;
                M.natom = q.natom;
                M.site0 = spd[k].site0;
                M.term0 = spd[k].term0;
                M.nintra = spd[k].nintra;
                M.bterm0 = spd[k].bterm0;
                M.nbterm = spd[k].nbterm;
                M.pad[0] = M.pad[1] = M.pad[2] = 0;
                for (int a = 0; a < q.natom; a++) {
                    atom_mol.push_back((int)mols.size());
                    atom_site.push_back(spd[k].site0 + a);
                    q2 += q.charge[a] * q.charge[a];  // src/frclong.c:1051-1057 order
                    if (q.ghost[a] != MSPMS_GHOST_FULL) nreal++;
                }
                mols.push_back(M);
            }
        }
        R.natom = (int)atom_mol.size() - R.atom0;
        R.nmol = (int)mols.size() - R.mol0;
        R.L[0] = P.boxlx;
        R.L[1] = P.boxly;
        R.L[2] = P.boxlz;
        for (int d = 0; d < 3; d++) R.invL[d] = 1.0 / R.L[d];
        R.rc_lj = P.rcutoff;
        R.rc2_lj = P.rcutoffsq;
        R.rc_el = P.rcutoffelec;
        R.rc2_el = P.rcutoffelecsq;
        double rmax = P.iChargeType == MSPMS_ELEC_NONE ? P.rcutoff
                                                       : std::max(P.rcutoff, P.rcutoffelec);
        R.rc_list = rmax * (1.0 + 1e-9) + 1e-9 + c->skin;
        R.rcon2 = P.rcutonsq;
        R.roff2_ron2_cube = P.roff2_minus_ron2_cube;
        R.shift1 = P.shift1;
        R.shift4 = P.shift4;
        R.f0 = P.f0;
        R.C = P.coulomb_prefactor;
        R.kappa = P.kappa;
        R.wvc1 = P.wolfvcon1;
        R.wvc2 = P.wolfvcon2;
        R.wvc3 = P.wolfvcon3;
        R.wfc1 = P.wolffcon1;
        R.wfc2 = P.wolffcon2;
        R.vfactor = P.vfactor_ewald;
        R.bfactor = P.bfactor_ewald;
        R.twopi_l[0] = P.twopi_lx;
        R.twopi_l[1] = P.twopi_ly;
        R.twopi_l[2] = P.twopi_lz;
        R.self_sum_q2 = q2;
        R.charge_type = P.iChargeType;
        R.switch_on = P.isLJswitchOn;
        R.sf_type = P.iSF_type;
        R.twopi_over_3v = P.twopi_over_3v;
        R.ewald_bc = P.iChargeType == MSPMS_ELEC_EWALD ? P.fEwald_BC : 0;
        if (R.ewald_bc == MSPMS_EWALD_BC_VACUUM) c->any_vacuum = true;
        // cell grid: cubes holding ~cell_atoms real sites; columns along z
        double vol = P.boxlx * P.boxly * P.boxlz;
        double dens = std::max(nreal, 1) / vol;
        double side = std::cbrt(c->cell_atoms / dens);
        int nc[3];
        for (int d = 0; d < 3; d++) {
            nc[d] = std::max(1, (int)std::floor(R.L[d] / side));
            nc[d] = std::min(nc[d], 1024);
            R.h[d] = R.L[d] / nc[d];
            R.inv_h[d] = 1.0 / R.h[d];
        }
        R.ncx = nc[0];
        R.ncy = nc[1];
        R.ncz = nc[2];
        R.cell0 = cell0;
        R.col0 = col0;
        R.ncol = nc[0] * nc[1];
        for (int g = 0; g < R.ncol; g++) {
            col_cell0.push_back(col_cell0.back() + nc[2]);
            col_rep.push_back(r);
        }
        cell0 += R.ncol * nc[2];
        col0 += R.ncol;
        max_rep_cl = std::max(max_rep_cl, (size_t)(nreal / MSPMS_CL + R.ncol + 1));
        double reach = 2.0 * R.rc_list + 3.0 * side;
        max_est = std::max(max_est, 1.5 * reach * reach * reach * dens / MSPMS_CL + 64.0);
        // Ewald reciprocal rows in the reference's k order (src/frclong.c:996-1003)
        R.kmax[0] = P.kmaxx;
        R.kmax[1] = P.kmaxy;
        R.kmax[2] = P.kmaxz;
        R.ksqmax = P.ksqmax;
        R.krow0 = (int)rows.size();
        R.k0 = (int)kvec.size();
        if (P.iChargeType == MSPMS_ELEC_EWALD) {
            c->any_recip = true;
            // The reference sums over k and -k (src/frclong.c:988-1048).  For
            // real charges S(-k) = conj S(k), so -k gives the same energy
            // and force terms as k: only the half space kx > 0, or kx = 0
            // and ky > 0, or kx = ky = 0 and kz > 0 is enumerated, each k
            // with twice its weight (exact in floating point).
            auto half = [](int kx, int ky, int kz) {
                return kx > 0 || (kx == 0 && (ky > 0 || (ky == 0 && kz > 0)));
            };
            for (int kx = 0; kx <= P.kmaxx; kx++)
                for (int ky = -P.kmaxy; ky <= P.kmaxy; ky++) {
                    KRow row{};
                    row.kx = kx;
                    row.ky = ky;
                    row.rep = r;
                    int n = 0;
                    for (int kz = -P.kmaxz; kz <= P.kmaxz; kz++) {
                        int ksq = kx * kx + ky * ky + kz * kz;
                        bool ok = ksq <= P.ksqmax && ksq != 0 && half(kx, ky, kz);
                        if (ok) {
                            if (n == 0) {
                                row.kz_lo = kz;
                                row.k0 = (int)kvec.size();
                            }
                            double rkx = P.twopi_lx * kx, rky = P.twopi_ly * ky,
                                   rkz = P.twopi_lz * kz;
                            volatile double rksq = rkx * rkx + rky * rky + rkz * rkz;
                            kvec.push_back(2.0 * (std::exp(-P.bfactor_ewald * rksq) / rksq));
                            k_rep.push_back(r);
                            row.kz_hi = kz;
                            n++;
                        }
                        // close the row at a gap (kz = 0 of the (0,0) row)
                        bool next_ok = false;
                        if (kz < P.kmaxz) {
                            int k2 = kx * kx + ky * ky + (kz + 1) * (kz + 1);
                            next_ok = k2 <= P.ksqmax && k2 != 0 && half(kx, ky, kz + 1);
                        }
                        if (n > 0 && !next_ok) {
                            row.nk = n;
                            rows.push_back(row);
                            n = 0;
                        }
                    }
                }
            R.nkrow = (int)rows.size() - R.krow0;
            R.nk = (int)kvec.size() - R.k0;
            nk_stride = std::max(nk_stride, R.nk);
            ntab_max = std::max(ntab_max, P.kmaxx + P.kmaxy + std::max(P.kmaxz, 1) + 3);
            nkz_max = std::max(nkz_max, 2 * P.kmaxz + 1);
            nkrow_max = std::max(nkrow_max, R.nkrow);
            // S pass: rows sorted by length; one workgroup per SK_CHUNK-atom
            // chunk (the kernel's tables hold up to 64).  Smaller chunks give
            // more waves but take more of the GPU from the pair kernel that
            // runs beside this pass: 64 measured fastest for the RESPA step
            std::vector<KRow> rs(rows.begin() + R.krow0, rows.end());
            std::stable_sort(rs.begin(), rs.end(),
                             [](const KRow& a, const KRow& b) { return a.nk > b.nk; });
            rows_s.insert(rows_s.end(), rs.begin(), rs.end());
            const int SK_CHUNK = 64;
            int nchunk = (R.natom + SK_CHUNK - 1) / SK_CHUNK;
            for (int ch = 0; ch < nchunk; ch++) {
                int a0 = R.atom0 + SK_CHUNK * ch;
                work_s.push_back(make_int4(r, rep_chunk0.back() + ch, a0,
                                           std::min(SK_CHUNK, R.atom0 + R.natom - a0)));
            }
            rep_chunk0.push_back(rep_chunk0.back() + nchunk);
        } else {
            rep_chunk0.push_back(rep_chunk0.back());
        }
        if (P.iSF_type != MSPMS_SF_NONE) c->any_sf = true;
    }
    if (c->any_recip) {
        // force pass: (atom block of 64) x (row group), ~4096 waves in total
        const long KF_WAVES = 4096;
        long blocks = 0;
        for (int r = 0; r < c->nrep; r++) blocks += (reps[r].natom + 63) / 64;
        for (int r = 0; r < c->nrep; r++) {
            const RepDev& R = reps[r];
            int nb = (R.natom + 63) / 64;
            int groups = (int)std::max(
                1L, std::min((long)R.nkrow, (KF_WAVES + blocks - 1) / std::max(blocks, 1L)));
            // split rows into `groups` ranges of about equal k count
            std::vector<int> cut(1, R.krow0);
            double per = (double)R.nk / groups, acc = 0.0;
            for (int i = R.krow0; i < R.krow0 + R.nkrow; i++) {
                acc += rows[i].nk;
                if (acc >= per * (double)cut.size() && (int)cut.size() < groups && i + 1 < R.krow0 + R.nkrow)
                    cut.push_back(i + 1);
            }
            cut.push_back(R.krow0 + R.nkrow);
            for (int b = 0; b < nb; b++)
                for (size_t g = 0; g + 1 < cut.size(); g++)
                    if (cut[g + 1] > cut[g])
                        work_f.push_back(make_int4(r, R.atom0 + 64 * b, cut[g], cut[g + 1]));
        }
    }
    c->natom = (int)atom_mol.size();
    c->nmol = (int)mols.size();
    c->reps = reps;

    int rc = 0;
    if ((rc = c->d_sites.upload(sites, s)) || (rc = c->d_species.upload(spd, s)) ||
        (rc = c->d_terms.upload(terms, s)) || (rc = c->d_mols.upload(mols, s)) ||
        (rc = c->d_atom_mol.upload(atom_mol, s)) || (rc = c->d_atom_site.upload(atom_site, s)) ||
        (rc = upload_rmass(c, sites, atom_site, s)) ||
        (rc = c->d_reps.upload(reps, s)))
        return rc;

    // ---- clustering buffers ----
    ClusterBufs& b = c->cb;
    int ncell = cell0, ncol = col0;
    size_t maxcl = (size_t)(c->natom / MSPMS_CL + ncol + 1);
    c->cap_max = (int)max_rep_cl;
    int cap = (int)std::min<double>((double)max_rep_cl, std::max(max_est, (double)c->cap_hint));
    if (c->cap_limit > 0) cap = std::min(cap, c->cap_limit);
    cap = std::max(cap, 16);
    if ((rc = c->d_atom_cell.ensure(c->natom)) || (rc = c->d_atom_slot.ensure(c->natom)) ||
        (rc = c->d_atom_wz.ensure(c->natom)) || (rc = c->d_cell_count.ensure(ncell)) ||
        (rc = c->d_cell_start.ensure(ncell + 1)) || (rc = c->d_cell_members.ensure(c->natom)) ||
        (rc = c->d_col_cell0.upload(col_cell0, s)) || (rc = c->d_col_rep.upload(col_rep, s)) ||
        (rc = c->d_col_cl0.ensure(ncol + 1)) || (rc = c->d_cl_col.ensure(maxcl)) ||
        (rc = c->d_ncl_total.ensure(1)) || (rc = c->d_slot_orig.ensure(maxcl * MSPMS_CL)) ||
        (rc = c->d_slot_info.ensure(maxcl * MSPMS_CL)) ||
        (rc = c->d_slot_x64.ensure(maxcl * MSPMS_CL)) ||
        (rc = c->d_slot_w64.ensure(maxcl * MSPMS_CL)) || (rc = c->d_atom_x0.ensure(c->natom)) ||
        (rc = c->d_lstate.ensure(LS_N)) ||
        (rc = c->d_slot_off.ensure(maxcl * MSPMS_CL)) ||
        (rc = c->d_slot_par.ensure(maxcl * MSPMS_CL)) || (rc = c->d_cl_center.ensure(maxcl)) ||
        (rc = c->d_cl_ext.ensure(maxcl)) || (rc = c->d_cl_box.ensure(2 * maxcl)) || (rc = c->d_plist.ensure(maxcl * (size_t)cap)) ||
        (rc = c->d_plist_n.ensure(maxcl)) || (rc = c->d_atom_gslot.ensure(c->natom)) ||
        (rc = c->d_fslot.ensure(3 * maxcl * MSPMS_CL)) || (rc = c->d_bin_ctr.ensure(1)))
        return rc;
    HIPC(hipMemsetAsync(c->d_bin_ctr.p, 0, sizeof(int), s));
    HIPC(hipMemsetAsync(c->d_cell_count.p, 0, sizeof(int) * (size_t)ncell, s));
    HIPC(hipMemsetAsync(c->d_lstate.p, 0, sizeof(int) * LS_N, s));  // no list yet
    b.atom_cell = c->d_atom_cell.p;
    b.atom_slot = c->d_atom_slot.p;
    b.atom_wz = c->d_atom_wz.p;
    b.cell_count = c->d_cell_count.p;
    b.cell_start = c->d_cell_start.p;
    b.cell_members = c->d_cell_members.p;
    b.col_cell0 = c->d_col_cell0.p;
    b.col_rep = c->d_col_rep.p;
    b.col_cl0 = c->d_col_cl0.p;
    b.cl_col = c->d_cl_col.p;
    b.ncl_total = c->d_ncl_total.p;
    b.slot_orig = c->d_slot_orig.p;
    b.slot_info = c->d_slot_info.p;
    b.slot_x64 = c->d_slot_x64.p;
    b.slot_w64 = c->d_slot_w64.p;
    b.atom_x0 = c->d_atom_x0.p;
    b.lstate = c->d_lstate.p;
    b.slot_off = c->d_slot_off.p;
    b.slot_par = c->d_slot_par.p;
    b.cl_center = c->d_cl_center.p;
    b.cl_ext = c->d_cl_ext.p;
    b.cl_box = c->d_cl_box.p;
    b.plist = c->d_plist.p;
    b.plist_n = c->d_plist_n.p;
    b.atom_gslot = c->d_atom_gslot.p;
    b.fslot = c->d_fslot.p;
    b.hist = nullptr;  // set with the finalize buffers below
    b.bin_ctr = c->d_bin_ctr.p;
    b.maxcl = (int)maxcl;
    b.cap = cap;
    b.ncell = ncell;
    b.ncol = ncol;

    // ---- reciprocal buffers ----
    RecipBufs& R = c->rb;
    memset(&R, 0, sizeof R);
    if (c->any_recip) {
        int nchunk = rep_chunk0.back();
        if ((rc = c->d_rows.upload(rows, s)) || (rc = c->d_rows_s.upload(rows_s, s)) ||
            (rc = c->d_kvec.upload(kvec, s)) || (rc = c->d_k_rep.upload(k_rep, s)) ||
            (rc = c->d_work_s.upload(work_s, s)) || (rc = c->d_work_f.upload(work_f, s)) ||
            (rc = c->d_rep_chunk0.upload(rep_chunk0, s)) ||
            (rc = c->d_s_part.ensure((size_t)nchunk * nk_stride)) ||
            (rc = c->d_ak.ensure(kvec.size())) || (rc = c->d_uk.ensure(kvec.size())) ||
            (rc = c->d_eatom.ensure(3 * (size_t)c->natom)))
            return rc;
        R.rows = c->d_rows.p;
        R.nrow = (int)rows.size();
        R.rows_s = c->d_rows_s.p;
        R.kvec = c->d_kvec.p;
        R.k_rep = c->d_k_rep.p;
        R.work_s = c->d_work_s.p;
        R.nwork_s = (int)work_s.size();
        R.work_f = c->d_work_f.p;
        R.nwork_f = (int)work_f.size();
        R.rep_chunk0 = c->d_rep_chunk0.p;
        R.s_part = c->d_s_part.p;
        R.nk_stride = nk_stride;
        R.ak = c->d_ak.p;
        R.uk = c->d_uk.p;
        R.nk_total = (int)kvec.size();
        R.ntab_max = ntab_max;
        R.nkz_max = nkz_max;
        R.nkrow_max = nkrow_max;
        R.eatom = c->d_eatom.p;
    }
    // ---- solid-fluid tables (replica 0's model, shared) ----
    memset(&c->sf, 0, sizeof c->sf);
    if (c->any_sf) {
        const mspms_params& P = c->params[0];
        for (int r = 1; r < c->nrep; r++)
            if (c->params[r].iSF_type != P.iSF_type)
                return fail(MSPMS_EINVAL, "replicas must share one solid-fluid model");
        SFDev& F = c->sf;
        F.type = P.iSF_type;
        F.solid_sigma = P.solid_sigma;
        F.solid_epsilon = P.solid_epsilon;
        F.c3pisq = P.const_3pisq_theta;
        F.sigma_base = P.sigma_base;
        F.epsilon_base = P.epsilon_base;
        if (F.type == MSPMS_SF_HYPERGEO) {
            std::vector<double> t;
            for (int k = 0; k < P.ntube; k++) {
                t.push_back(P.tube_x[k]);
                t.push_back(P.tube_y[k]);
                t.push_back(P.tube_r[k]);
            }
            if ((rc = c->d_sf_tube.upload(t, s))) return rc;
            F.ntube = P.ntube;
            F.tube = c->d_sf_tube.p;
        } else if (F.type == MSPMS_SF_ATOM_EXPLICIT) {
            std::vector<double> t;
            for (int k = 0; k < P.solid_natom; k++) {
                t.push_back(P.solid_xx[k]);
                t.push_back(P.solid_yy[k]);
                t.push_back(P.solid_zz[k]);
            }
            if ((rc = c->d_sf_solid.upload(t, s))) return rc;
            F.nsolid = P.solid_natom;
            F.solid = c->d_sf_solid.p;
        } else if (F.type == MSPMS_SF_MY_INTERP) {
            const mspms_sf_cubes& Q = *P.cubes;
            const size_t ncube = (size_t)Q.ncx * Q.ncy * Q.ncz, nt = (size_t)Q.ntype;
            std::vector<const double*> src(Q.tables, Q.tables + 128 * nt);
            // the tables are large (up to GiB): upload once per source
            if (c->sf_cubes_src != src || c->d_sf_cubes.n < nt * ncube * 128) {
                DevBuf<double> stage;
                if ((rc = c->d_sf_cubes.ensure(nt * ncube * 128)) ||
                    (rc = stage.ensure(nt * ncube * 128)))
                    return rc;
                hipError_t e = hipSuccess;
                for (size_t k = 0; k < 128 * nt && e == hipSuccess; k++)
                    e = hipMemcpyAsync(stage.p + k * ncube, Q.tables[k], ncube * sizeof(double),
                                       hipMemcpyHostToDevice, s);
                if (e == hipSuccess) {
                    launch_cube_major(s, stage.p, c->d_sf_cubes.p, (int)nt, (int)ncube);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                stage.release();
                if (e != hipSuccess)
                    return fail(MSPMS_EIO, "myinterp table upload: %s", hipGetErrorString(e));
                c->sf_cubes_src = src;
            }
            F.ntype = Q.ntype;
            F.nx = Q.ncx;
            F.ny = Q.ncy;
            F.nz = Q.ncz;
            F.x0 = Q.xmin;
            F.y0 = Q.ymin;
            F.z0 = 0.0;
            F.zmax = Q.zmax;
            F.hx = Q.hx;
            F.hy = Q.hy;
            F.hz = Q.hz;
            F.cubes = c->d_sf_cubes.p;
        } else {
            const mspms_sf_grid& G = *P.grid;
            size_t nn = (size_t)G.nx * G.ny * G.nz, tot = nn * 8 * (size_t)G.ntype;
            // the node table is large (up to GiB): uploaded once per source,
            // through a staging buffer, and turned into node-major records
            if (c->sf_nodes_src != G.nodes || c->sf_nodes_uploaded != tot) {
                DevBuf<float> stage;
                struct Release {
                    DevBuf<float>& b;
                    ~Release() { b.release(); }
                } release_stage{stage};
                if ((rc = c->d_sf_nodes.ensure(tot)) || (rc = stage.ensure(tot))) return rc;
                HIPC(hipMemcpyAsync(stage.p, G.nodes, tot * sizeof(float), hipMemcpyHostToDevice,
                                    s));
                launch_grid_node_major(s, stage.p, c->d_sf_nodes.p, nn, G.ntype);
                HIPC(hipGetLastError());
                HIPC(hipStreamSynchronize(s));
                c->sf_nodes_src = G.nodes;
                c->sf_nodes_uploaded = tot;
            }
            std::vector<double> mn(G.ntype, 10e10);  // grid.f90:61
            for (int t = 0; t < G.ntype; t++)
                for (size_t k = 0; k < nn; k++)
                    mn[t] = std::min(mn[t], (double)G.nodes[(size_t)t * 8 * nn + k]);
            if ((rc = c->d_sf_minpe.upload(mn, s))) return rc;
            F.ntype = G.ntype;
            F.nx = G.nx;
            F.ny = G.ny;
            F.nz = G.nz;
            F.x0 = G.x0;
            F.y0 = G.y0;
            F.z0 = G.z0;
            F.hx = G.hx;
            F.hy = G.hy;
            F.hz = G.hz;
            F.zperiod = G.zperiod > 0.0 ? G.zperiod : (G.nz - 1) * G.hz;
            F.nodes = c->d_sf_nodes.p;
            F.min_pe = c->d_sf_minpe.p;
        }
    }

    // ---- partials and outputs ----
    if ((rc = c->d_ep_pair.ensure(maxcl * 4)) || (rc = c->d_uvac.ensure(c->nrep)) ||
        (rc = c->d_ep_coul.ensure(c->nmol)) || (rc = c->d_mcom.ensure(c->nmol)) || (rc = c->d_ip_pair.ensure(maxcl * 4)) ||
        (rc = c->d_ep_intra.ensure((size_t)c->nmol * 4)) ||
        (rc = c->d_ep_short.ensure((size_t)c->nmol * 5)) ||
        (rc = c->d_ep_sf.ensure((size_t)c->natom)) || (rc = c->d_long.ensure(c->nrep)) ||
        (rc = c->d_short.ensure(c->nrep)) || (rc = c->d_pos.ensure((size_t)3 * c->natom)) ||
        (rc = c->d_frec.ensure((size_t)3 * c->natom)) ||
        (rc = c->d_fl.ensure((size_t)3 * c->natom)) || (rc = c->d_fs.ensure((size_t)3 * c->natom)) ||
        (rc = c->d_kin.ensure((size_t)c->natom)) ||
        (rc = c->d_fin_part.ensure((size_t)c->nrep * FIN_BLOCKS * FIN_ND)) ||
        (rc = c->d_fin_partu.ensure((size_t)c->nrep * FIN_BLOCKS * 4)) ||
        (rc = c->d_fin_count.ensure((size_t)c->nrep)) || (rc = c->d_fin_on.ensure(1)) ||
        (rc = c->d_hist.ensure(1)) || (rc = c->d_hstep.ensure((size_t)c->nrep)) ||
        (rc = c->d_snap.ensure((size_t)6 * c->natom)) || (rc = c->d_snap_count.ensure(1)))
        return rc;
    HIPC(hipMemsetAsync(c->d_snap_count.p, 0, sizeof(int), s));
    c->snap_pos = c->snap_vel = nullptr;
    HIPC(hipMemsetAsync(c->d_fin_count.p, 0, sizeof(int) * (size_t)c->nrep, s));
    HIPC(hipMemsetD32Async((hipDeviceptr_t)c->d_fin_on.p, 1, 1, s));
    HIPC(hipMemsetAsync(c->d_hist.p, 0, sizeof(HistDev), s));
    HIPC(hipMemsetAsync(c->d_hstep.p, 0, sizeof(int) * (size_t)c->nrep, s));
    c->fin.part = c->d_fin_part.p;
    c->fin.partu = c->d_fin_partu.p;
    c->fin.count = c->d_fin_count.p;
    c->fin.on = nullptr;
    c->fin.hist = c->d_hist.p;
    c->cb.hist = c->d_hist.p;
    c->fin.hstep = c->d_hstep.p;
    {
        // one partial of each kind per thread in the largest replica
        size_t most = 1;
        for (const RepDev& R : c->reps)
            most = std::max({most, (size_t)R.natom, (size_t)R.nmol, (size_t)R.nk,
                             (size_t)(R.natom / MSPMS_CL + R.ncol + 1)});
        c->fin.nb = (int)std::min<size_t>(FIN_BLOCKS, (most + FIN_T - 1) / FIN_T);
    }
    HIPC(hipStreamSynchronize(s));
    return 0;
}

int set_species(mspms_ctx* c, int nspecie, const mspms_species* sp) {
    if (nspecie < 1 || !sp) return fail(MSPMS_EINVAL, "need at least one species");
    std::vector<SpeciesHost> out;
    for (int k = 0; k < nspecie; k++) {
        const mspms_species& q = sp[k];
        int rc = validate_species(q);
        if (rc) return rc;
        SpeciesHost h;
        h.natom = q.natom;
        h.nmole = q.nmole;
        h.ghost = copy_arr(q.ghost_type, q.natom);
        h.interp = copy_arr(q.interp_type, q.natom);
        h.aw = copy_arr(q.aw, q.natom);
        h.sigma = copy_arr(q.sigma, q.natom);
        h.epsilon = copy_arr(q.epsilon, q.natom);
        h.charge = copy_arr(q.charge, q.natom);
        h.bnd_idx = copy_arr(q.bnd_idx, 2 * (size_t)q.nbond);
        h.bnd_type = copy_arr(q.bnd_type, q.nbond);
        h.bnd_para = copy_arr(q.bnd_para, 3 * (size_t)q.nbond);
        h.agl_idx = copy_arr(q.agl_idx, 3 * (size_t)q.nangle);
        h.agl_type = copy_arr(q.agl_type, q.nangle);
        h.agl_para = copy_arr(q.agl_para, 5 * (size_t)q.nangle);
        h.dih_idx = copy_arr(q.dih_idx, 4 * (size_t)q.ndih);
        h.dih_type = copy_arr(q.dih_type, q.ndih);
        h.dih_para = copy_arr(q.dih_para, 4 * (size_t)q.ndih);
        h.imp_idx = copy_arr(q.imp_idx, 4 * (size_t)q.nimp);
        h.imp_type = copy_arr(q.imp_type, q.nimp);
        h.imp_para = copy_arr(q.imp_para, 2 * (size_t)q.nimp);
        h.agl_unique.assign(std::max(q.nangle, 1), 1);
        h.dih_unique.assign(std::max(q.ndih, 1), 1);
        check_uniques(h.agl_idx.data(), q.nangle, h.dih_idx.data(), q.ndih, h.bnd_idx.data(),
                      q.nbond, h.agl_unique.data(), h.dih_unique.data());
        if (q.agl_unique) h.agl_unique.assign(q.agl_unique, q.agl_unique + q.nangle);
        if (q.dih_unique) h.dih_unique.assign(q.dih_unique, q.dih_unique + q.ndih);
        h.agl_unique.resize(std::max(q.nangle, 1), 1);
        h.dih_unique.resize(std::max(q.ndih, 1), 1);
        out.push_back(std::move(h));
    }
    c->sp = std::move(out);
    return 0;
}

}  // namespace

extern "C" int mspms_set_system(mspms_ctx* c, int nspecie, const mspms_species* sp,
                                const mspms_params* p) {
    if (!c || !p) return fail(MSPMS_EINVAL, "null argument");
    int rc = validate_params(*p);
    if (rc) return rc;
    if ((rc = set_species(c, nspecie, sp))) return rc;
    c->nrep = 1;
    c->nmole.resize(nspecie);
    for (int k = 0; k < nspecie; k++) c->nmole[k] = sp[k].nmole;
    c->params.assign(1, *p);
    c->hmc_ready = false;  // the chains belong to the previous layout
    c->hmc_rep_delt = false;
    c->nvt_ready = false;
    c->npt_ready = false;
    return build(c);
}

extern "C" int mspms_set_params(mspms_ctx* c, const mspms_params* p) {
    if (!c || !p) return fail(MSPMS_EINVAL, "null argument");
    if (c->sp.empty()) return fail(MSPMS_EINVAL, "mspms_set_system first");
    int rc = validate_params(*p);
    if (rc) return rc;
    for (int r = 0; r < c->nrep; r++) c->params[r] = *p;
    return build(c);
}

extern "C" int mspms_set_replicas(mspms_ctx* c, int nrep, const int* nmole,
                                  const mspms_params* params) {
    if (!c || nrep < 1 || !nmole || !params) return fail(MSPMS_EINVAL, "bad arguments");
    if (c->sp.empty()) return fail(MSPMS_EINVAL, "mspms_set_system first");
    const int ns = (int)c->sp.size();
    for (int r = 0; r < nrep; r++) {
        int rc = validate_params(params[r]);
        if (rc) return rc;
        for (int k = 0; k < ns; k++)
            if (nmole[(size_t)r * ns + k] < 0) return fail(MSPMS_EINVAL, "negative nmole");
    }
    c->nrep = nrep;
    c->nmole.assign(nmole, nmole + (size_t)nrep * ns);
    c->params.assign(params, params + nrep);
    c->hmc_ready = false;
    c->hmc_rep_delt = false;
    c->nvt_ready = false;
    c->npt_ready = false;
    return build(c);
}

extern "C" int mspms_natom(const mspms_ctx* c) { return c ? c->natom : 0; }
extern "C" int mspms_nreplica(const mspms_ctx* c) { return c ? c->nrep : 0; }

// ---- evaluation -----------------------------------------------------------
namespace {

enum {
    PROF_CLUSTER = 0, PROF_PAIR, PROF_INTRA, PROF_RECIP_S, PROF_RECIP_F, PROF_FINAL, PROF_BONDED,
    PROF_SF,
    PROF_N
};

struct ProfScope {
    mspms_ctx* c;
    hipStream_t s;
    int slot;
    size_t idx = 0;
    bool on = false;
    ProfScope(mspms_ctx* c_, hipStream_t s_, int slot_) : c(c_), s(s_), slot(slot_) {
        if (!c->profiling || c->capturing) return;
        if (c->ev_used + 2 > c->ev.size()) {
            size_t old = c->ev.size(), grow = std::max<size_t>(64, old);
            c->ev.resize(old + grow);
            for (size_t k = old; k < c->ev.size(); k++)
                if (hipEventCreate(&c->ev[k]) != hipSuccess) {
                    c->ev.resize(k);
                    return;
                }
            if (c->ev_used + 2 > c->ev.size()) return;
        }
        idx = c->ev_used;
        c->ev_used += 2;
        on = true;
        (void)hipEventRecord(c->ev[idx], s);
    }
    ~ProfScope() {
        if (!on) return;
        (void)hipEventRecord(c->ev[idx + 1], s);
        c->pending.push_back(slot);
    }
};

__global__ void k_reset_status(Status* st) {
    st->flags = 0;
    st->max_list = 0;
}

// The frclong terms from the partial sums; with sterms, also the frcshort
// terms of the inner steps (ep_short) and the kinetic energy (kin).
// gated: the pass runs only while the context's finalize flag is 1 (the
// device vver run clears it for the steps whose terms nobody reads)
void launch_finalize(mspms_ctx* c, const LaunchCtx& L, mspms_long_terms* terms,
                     mspms_short_terms* sterms = nullptr, double* ukin = nullptr,
                     bool gated = false, const FinAtoms* atoms = nullptr) {
    FinBufs fb = c->fin;
    fb.on = gated ? c->d_fin_on.p : nullptr;
    // with the wall split into the inner loop, a RESPA step (sterms) still
    // reports its energy, which the last inner step left in ep_sf; a
    // standalone frclong leaves it out
    const bool have_sf = c->any_sf && (!c->sf_split || sterms);
    launch_finalize_long(L, c->cb.col_cl0, c->d_ep_pair.p, c->d_ip_pair.p, 1, c->d_ep_intra.p,
                         c->any_intra ? 1 : 0, c->rb.uk, c->any_recip ? 1 : 0, c->d_ep_sf.p,
                         have_sf ? 1 : 0, c->any_vacuum ? c->d_uvac.p : nullptr,
                         c->any_coul ? c->d_ep_coul.p : nullptr, terms, c->cb.lstate,
                         sterms ? c->d_ep_short.p : nullptr, sterms, sterms ? c->d_kin.p : nullptr,
                         sterms ? ukin : nullptr, fb, atoms);
}

// grad (optional): per-atom parameter derivatives gq | ge | gs, 3 * natom;
// with it the pair path is the FP64 one whatever the context precision
// finalize = false leaves the partial sums for the caller's finalize pass
// (enqueue_vver reduces them together with the step's other partials).
// kick (optional): the RESPA step's closing half kick, applied by the last
// kernel that adds long forces when no kernel adds after it; *kicked tells
// whether it was (otherwise the caller runs it).
// frec_deferred (optional): with the reciprocal forces on the side stream,
// leave their addition (and the kick) to the caller's finalize pass and
// report it
// the slot-ordered forces of the half-list (mixed) pair kernel, or none
SlotForces slot_forces(const mspms_ctx* c) {
    if (!c->cb.half) return SlotForces{};
    return SlotForces{c->cb.fslot, c->cb.atom_gslot, (size_t)c->cb.maxcl * MSPMS_CL};
}

int enqueue_long(mspms_ctx* c, const double* x, const double* y, const double* z, double* fx,
                 double* fy, double* fz, mspms_long_terms* terms, hipStream_t s,
                 double* grad = nullptr, bool finalize = true, const Kick* kick = nullptr,
                 bool* kicked = nullptr, bool* frec_deferred = nullptr) {
    const size_t n = (size_t)c->natom;
    double* gq = grad;
    double* ge = grad ? grad + n : nullptr;
    double* gs = grad ? grad + 2 * n : nullptr;
    // the status word is sticky: the host API clears it before each call,
    // device-API runs accumulate it until mspms_sync reads it
    LaunchCtx L = launch_ctx(c, s);
    const int hash = c->want_hash < 0 ? (c->precision == MSPMS_PREC_FP64) : c->want_hash;
    // fork: the Ewald S(k) pass only reads positions, so it runs on the side
    // stream beside the clustering and the pair kernel; the force pass,
    // which adds into the forces the pair kernel writes, waits for it
    // Without the adjoint the force pass runs there too, into its own (3, N)
    // buffer that joins the long forces in k_intra (or k_add_frec); the
    // adjoint keeps it on the main stream (its charge derivative adds into
    // the buffer the pair kernel writes).
    const bool recip_side = c->any_recip && !grad;
    double* frec = recip_side ? c->d_frec.p : nullptr;
    if (c->any_recip) {
        HIPC(hipEventRecord(c->ev_fork, s));
        HIPC(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        LaunchCtx Ls = launch_ctx(c, c->side);
        if (recip_side) {
            // the intramolecular corrections need positions only: they run
            // first on this stream, into the same (3, N) buffer as the
            // reciprocal forces, beside the list build and the pair kernel
            if (c->any_intra) {
                // every atom belongs to one molecule thread, which clears its
                // atoms' entries first: no separate zero fill
                ProfScope p(c, c->side, PROF_INTRA);
                launch_intra(Ls, c->d_species.p, c->d_terms.p, x, y, z, frec, frec + n,
                             frec + 2 * n, c->d_ep_intra.p, nullptr, nullptr, nullptr, nullptr,
                             nullptr, true);
            } else {
                launch_zero(Ls, frec, 3 * n);
            }
        }
        {
            ProfScope p(c, c->side, PROF_RECIP_S);
            launch_recip_s(Ls, c->rb, x, y, z);
        }
        if (recip_side) {
            ProfScope p(c, c->side, PROF_RECIP_F);
            launch_recip_f(Ls, c->rb, x, y, z, frec, frec + n, frec + 2 * n, nullptr);
        }
        HIPC(hipEventRecord(c->ev_join, c->side));
    }
    {
        ProfScope p(c, s, PROF_CLUSTER);
        // the mixed pair kernel takes each cluster pair once; the exact and
        // adjoint kernels want both rows (a list of the other layout is
        // rebuilt on the device)
        c->cb.half = (!grad && c->precision == MSPMS_PREC_MIXED) ? 1 : 0;
        launch_cluster_build(L, c->cb, x, y, z, fx, fy, fz, 0.25 * c->skin * c->skin,
                             c->force_rebuild ? 1 : 0);
    }
    {
        ProfScope p(c, s, PROF_PAIR);
        launch_pair(L, c->cb, grad ? MSPMS_PREC_FP64 : c->precision, c->charge_union, fx, fy,
                    fz, c->d_ep_pair.p, c->d_ip_pair.p, grad ? 1 : hash, gq, ge, gs);
    }
    // the half-list kernel's slot-ordered forces join the atoms' forces in
    // the first per-atom pass after it
    const SlotForces slots = slot_forces(c);
    if (c->any_recip) HIPC(hipStreamWaitEvent(s, c->ev_join, 0));  // join
    // nothing adds long forces after k_intra / k_add_frec
    const bool sf_long = c->any_sf && !c->sf_split;
    const bool fuse = kick && !grad && !c->any_vacuum && !c->any_coul && !sf_long &&
                      (c->any_intra || recip_side);
    if (kicked) *kicked = fuse;
    const bool defer = recip_side && frec_deferred;
    if (frec_deferred) *frec_deferred = defer;
    if (defer) {
        // the caller's finalize folds the slot forces, adds frec and kicks
        if (kicked) *kicked = false;
    } else if (recip_side) {
        // reciprocal forces and intramolecular corrections join the long
        // forces, one thread per atom
        ProfScope p(c, s, PROF_INTRA);
        launch_add_frec(L, fx, fy, fz, frec, fuse ? kick : nullptr, &slots);
    } else {
        if (slots.f) launch_fold_slots(L, slots, fx, fy, fz);
        if (c->any_intra) {
            ProfScope p(c, s, PROF_INTRA);
            launch_intra(L, c->d_species.p, c->d_terms.p, x, y, z, fx, fy, fz, c->d_ep_intra.p,
                         gq, ge, gs, nullptr, fuse ? kick : nullptr);
        }
    }
    if (c->any_recip) {
        if (!recip_side) {
            ProfScope p(c, s, PROF_RECIP_F);
            launch_recip_f(L, c->rb, x, y, z, fx, fy, fz, gq);
        }
        if (c->any_vacuum)
            launch_vacuum(L, c->d_species.p, x, y, z, fx, fy, fz, c->d_uvac.p, gq);
    }
    if (c->any_coul) {
        ProfScope p(c, s, PROF_INTRA);
        launch_coulomb_mol(L, c->d_species.p, x, y, z, c->d_mcom.p, fx, fy, fz, c->d_ep_coul.p);
    }
    if (sf_long) {
        ProfScope p(c, s, PROF_SF);
        launch_sf(L, c->sf, x, y, z, fx, fy, fz, c->d_ep_sf.p);
    }
    if (finalize) {
        ProfScope p(c, s, PROF_FINAL);
        launch_finalize(c, L, terms);
    }
    HIPC(hipGetLastError());
    return 0;
}

int enqueue_short(mspms_ctx* c, const double* x, const double* y, const double* z, double* fx,
                  double* fy, double* fz, mspms_short_terms* terms, hipStream_t s) {
    LaunchCtx L = launch_ctx(c, s);
    ProfScope p(c, s, PROF_BONDED);
    launch_bonded(L, c->d_species.p, c->d_terms.p, x, y, z, fx, fy, fz, c->d_ep_short.p);
    const bool sf_short = c->any_sf && c->sf_split;
    if (sf_short) launch_sf(L, c->sf, x, y, z, fx, fy, fz, c->d_ep_sf.p);
    launch_finalize_short(L, c->d_ep_short.p, terms, sf_short ? c->d_ep_sf.p : nullptr);
    HIPC(hipGetLastError());
    return 0;
}

// One RESPA velocity-Verlet outer step, NVE (src/md.c:22-83) on device
// (3, N) SoA buffers; kernels only, so that it can be captured.
int enqueue_vver(mspms_ctx* c, double* pos, double* vel, double* fl, double* fs, double delt,
                 int nstep_inner, mspms_long_terms* lterms, mspms_short_terms* sterms,
                 double* ukin, hipStream_t s) {
    const size_t n = (size_t)c->natom;
    // src/initvars.c:605-607
    const double deltby2 = delt / 2.0, delts = delt / nstep_inner, deltsby2 = delts / 2.0;
    LaunchCtx L = launch_ctx(c, s);
    // each step saves its start while no overflow is flagged (mspms_sync
    // puts it back after one)
    L.snap = c->d_snap.p;
    L.snap_count = c->d_snap_count.p;
    c->snap_pos = pos;
    c->snap_vel = vel;
    {
        ProfScope p(c, s, PROF_BONDED);
        const bool sf_inner = c->any_sf && c->sf_split;
        launch_respa_inner(L, c->d_species.p, c->d_terms.p, pos, pos + n, pos + 2 * n, vel,
                           vel + n, vel + 2 * n, fl, fl + n, fl + 2 * n, fs, fs + n, fs + 2 * n,
                           deltby2, delts, deltsby2, nstep_inner, c->d_ep_short.p,
                           sf_inner ? &c->sf : nullptr, c->d_ep_sf.p);
    }
    const Kick kick{vel,        vel + n,    vel + 2 * n, c->d_kin.p, deltby2,
                    L.rep_delt, L.atom_mol, L.mols,      c->d_rmass.p};
    bool kicked = false, deferred = false;
    int rc = enqueue_long(c, pos, pos + n, pos + 2 * n, fl, fl + n, fl + 2 * n, lterms, s,
                          nullptr, false, &kick, &kicked, &deferred);
    if (rc) return rc;
    if (!kicked && !deferred)
        launch_respa_final(L, vel, vel + n, vel + 2 * n, fl, fl + n, fl + 2 * n, deltby2,
                           c->d_kin.p);
    {
        // (the side stream's forces and the closing kick,) then frclong,
        // frcshort (last inner step) and kinetic terms in one pass
        ProfScope p(c, s, PROF_FINAL);
        const FinAtoms A{fl, fl + n, fl + 2 * n, c->d_frec.p, kick, c->natom, slot_forces(c)};
        launch_finalize(c, L, lterms, sterms, ukin, true, deferred ? &A : nullptr);
    }
    HIPC(hipGetLastError());
    return 0;
}

int read_status(mspms_ctx* c, hipStream_t s) {
    HIPC(hipMemcpyAsync(c->h_status, c->d_status.p, sizeof(Status), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return 0;
}

// Headroom of the pair-list rows: when the longest row seen reaches 3/4 of
// the capacity, the rows are widened to twice that row, so that later
// builds of a moving system stay within it (in inhomogeneous systems, such
// as fluid adsorbing on a wall, the longest row can grow by a third within
// a few steps).
// The rows are copied into the wider buffer, so a valid list stays valid and
// the build count goes on; only the captured graph (which holds the old
// buffer) is dropped.  Called after a status read, outside any device run.
int keep_list_headroom(mspms_ctx* c) {
    ClusterBufs& b = c->cb;
    const int seen = c->h_status->max_list;
    if (c->cap_limit > 0 || seen * 4 < b.cap * 3) return 0;
    const int cap = std::min(c->cap_max, 2 * seen + 16);
    c->cap_hint = std::max(c->cap_hint, cap);
    if (cap <= b.cap) return 0;
    DevBuf<float4> wide;
    int rc = wide.ensure((size_t)b.maxcl * cap);
    if (rc) return rc;
    HIPC(hipMemcpy2DAsync(wide.p, (size_t)cap * sizeof(float4), b.plist,
                          (size_t)b.cap * sizeof(float4), (size_t)b.cap * sizeof(float4),
                          (size_t)b.maxcl, hipMemcpyDeviceToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    std::swap(c->d_plist.p, wide.p);
    std::swap(c->d_plist.n, wide.n);
    wide.release();
    drop_graph(c);
    b.plist = c->d_plist.p;
    b.cap = cap;
    return 0;
}

void accumulate_profile(mspms_ctx* c) {
    if (!c->profiling || c->pending.empty()) return;
    (void)hipEventSynchronize(c->ev[2 * c->pending.size() - 1]);
    for (size_t p = 0; p < c->pending.size(); p++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[2 * p], c->ev[2 * p + 1]) == hipSuccess) {
            c->prof_ms[c->pending[p]] += ms;
            c->prof_n[c->pending[p]]++;
        }
    }
    c->pending.clear();
    c->ev_used = 0;
}

}  // namespace

extern "C" int mspms_frclong_device(mspms_ctx* c, const double* x, const double* y,
                                    const double* z, double* fx, double* fy, double* fz,
                                    mspms_long_terms* terms, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return enqueue_long(c, x, y, z, fx, fy, fz, terms, s);
}

extern "C" int mspms_frcshort_device(mspms_ctx* c, const double* x, const double* y,
                                     const double* z, double* fx, double* fy, double* fz,
                                     mspms_short_terms* terms, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return enqueue_short(c, x, y, z, fx, fy, fz, terms, s);
}

extern "C" int mspms_frclong_host(mspms_ctx* c, const double* xx, const double* yy,
                                  const double* zz, double* fxl, double* fyl, double* fzl,
                                  mspms_long_terms* terms) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!xx || !yy || !zz || !fxl || !fyl || !fzl || !terms)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    const size_t n = (size_t)c->natom;
    double* dx = c->d_pos.p;
    double *dfx = c->d_fl.p, *dfy = dfx + n, *dfz = dfy + n;
    HIPC(hipMemcpyAsync(dx, xx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + n, yy, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + 2 * n, zz, n * sizeof(double), hipMemcpyHostToDevice, s));
    for (int attempt = 0; attempt < 4; attempt++) {
        hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, c->d_status.p);
        int rc = enqueue_long(c, dx, dx + n, dx + 2 * n, dfx, dfy, dfz, c->d_long.p, s);
        if (rc) return rc;
        if ((rc = read_status(c, s))) return rc;
        if (c->h_status->flags & ST_OVERFLOW) {
            // pair-list row too short: grow to the largest row seen and redo
            c->cap_hint = std::max(c->cap_hint * 2, c->h_status->max_list + 64);
            if ((rc = build(c))) return rc;
            continue;
        }
        accumulate_profile(c);
        break;
    }
    HIPC(hipMemcpyAsync(fxl, dfx, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fyl, dfy, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fzl, dfz, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(terms, c->d_long.p, sizeof(mspms_long_terms) * c->nrep,
                        hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (c->h_status->flags & ST_OVERFLOW)
        return fail(MSPMS_ENOMEM, "pair list overflow after growing");
    if (c->h_status->flags & ST_NONFINITE) return fail(MSPMS_ENOTFINITE, "non-finite position");
    if (c->h_status->flags & ST_GRID_EDOM)
        return fail(MSPMS_EDOM, "site outside the solid-fluid grid or without a grid type");
    return keep_list_headroom(c);
}

extern "C" int mspms_frcshort_host(mspms_ctx* c, const double* xx, const double* yy,
                                   const double* zz, double* fxs, double* fys, double* fzs,
                                   mspms_short_terms* terms) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!xx || !yy || !zz || !fxs || !fys || !fzs || !terms)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    const size_t n = (size_t)c->natom;
    double* dx = c->d_pos.p;
    double *dfx = c->d_fs.p, *dfy = dfx + n, *dfz = dfy + n;
    HIPC(hipMemcpyAsync(dx, xx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + n, yy, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + 2 * n, zz, n * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, c->d_status.p);
    int rc = enqueue_short(c, dx, dx + n, dx + 2 * n, dfx, dfy, dfz, c->d_short.p, s);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(fxs, dfx, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fys, dfy, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fzs, dfz, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(terms, c->d_short.p, sizeof(mspms_short_terms) * c->nrep,
                        hipMemcpyDeviceToHost, s));
    if ((rc = read_status(c, s))) return rc;
    if (c->h_status->flags & ST_SINGULAR)
        return fail(MSPMS_EINVAL, "improper at a singular geometry (src/impfrc.c:96-134)");
    return 0;
}

// Parameter adjoint of uinter (FP64): the frclong evaluation of
// mspms_frclong_host plus, per replica and global site id, dE/d(epsilon),
// dE/d(sigma), dE/d(charge).  The pair, intramolecular and reciprocal
// kernels write each atom's derivative with respect to its own parameters
// (every inter-molecular pair is visited from both atoms, so each visit
// differentiates only the visiting atom's side); k_grad_sites sums them
// over the atoms that share a template site.
extern "C" int mspms_energy_param_grad(mspms_ctx* c, const double* xx, const double* yy,
                                       const double* zz, double* fxl, double* fyl, double* fzl,
                                       mspms_long_terms* terms, mspms_param_grad* grad) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!xx || !yy || !zz || !fxl || !fyl || !fzl || !terms || !grad || !grad->dE_depsilon ||
        !grad->dE_dsigma || !grad->dE_dcharge)
        return fail(MSPMS_EINVAL, "null argument");
    if (c->any_sf)
        return fail(MSPMS_EINVAL, "parameter gradients of the solid-fluid term are not provided");
    if (c->any_coul)
        return fail(MSPMS_EINVAL, "parameter gradients of simple Coulomb are not provided");
    HIPC(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    const size_t n = (size_t)c->natom;
    const int ns = (int)c->sp.size();
    // work items (rep, global site, first atom of the site, molecule count)
    std::vector<int4> work;
    {
        int atom = 0;
        for (int r = 0; r < c->nrep; r++) {
            int site0 = 0;
            for (int k = 0; k < ns; k++) {
                const int nm = c->nmole[(size_t)r * ns + k], na = c->sp[k].natom;
                for (int a = 0; a < na; a++) work.push_back(make_int4(r, site0 + a, atom + a, nm));
                atom += nm * na;
                site0 += na;
            }
        }
    }
    const size_t nout = (size_t)c->nrep * c->nsite;
    int rc = 0;
    if ((rc = c->d_grad.ensure(3 * n)) || (rc = c->d_grad_out.ensure(3 * nout)) ||
        (rc = c->d_grad_work.upload(work, s)))
        return rc;
    double* dx = c->d_pos.p;
    double *dfx = c->d_fl.p, *dfy = dfx + n, *dfz = dfy + n;
    HIPC(hipMemcpyAsync(dx, xx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + n, yy, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dx + 2 * n, zz, n * sizeof(double), hipMemcpyHostToDevice, s));
    for (int attempt = 0; attempt < 4; attempt++) {
        hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, c->d_status.p);
        // the pair kernel writes the entries of binned sites only; GHOST_FULL
        // sites (never binned) start from zero for the kernels that add
        HIPC(hipMemsetAsync(c->d_grad.p, 0, 3 * n * sizeof(double), s));
        rc = enqueue_long(c, dx, dx + n, dx + 2 * n, dfx, dfy, dfz, c->d_long.p, s, c->d_grad.p);
        if (rc) return rc;
        if ((rc = read_status(c, s))) return rc;
        if (c->h_status->flags & ST_OVERFLOW) {
            c->cap_hint = std::max(c->cap_hint * 2, c->h_status->max_list + 64);
            if ((rc = build(c))) return rc;
            // build() may move every buffer, the uploaded positions included
            if ((rc = c->d_grad.ensure(3 * n)) || (rc = c->d_grad_out.ensure(3 * nout)) ||
                (rc = c->d_grad_work.upload(work, s)))
                return rc;
            dx = c->d_pos.p;
            dfx = c->d_fl.p;
            dfy = dfx + n;
            dfz = dfy + n;
            HIPC(hipMemcpyAsync(dx, xx, n * sizeof(double), hipMemcpyHostToDevice, s));
            HIPC(hipMemcpyAsync(dx + n, yy, n * sizeof(double), hipMemcpyHostToDevice, s));
            HIPC(hipMemcpyAsync(dx + 2 * n, zz, n * sizeof(double), hipMemcpyHostToDevice, s));
            continue;
        }
        break;
    }
    double* go = c->d_grad_out.p;
    {
        LaunchCtx L = launch_ctx(c, s);
        const double* g = c->d_grad.p;
        launch_grad_sites(L, c->d_species.p, c->d_grad_work.p, (int)work.size(), g, g + n,
                          g + 2 * n, go, go + nout, go + 2 * nout);
        HIPC(hipGetLastError());
    }
    HIPC(hipMemcpyAsync(fxl, dfx, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fyl, dfy, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(fzl, dfz, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(terms, c->d_long.p, sizeof(mspms_long_terms) * c->nrep,
                        hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(grad->dE_dcharge, go, nout * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(grad->dE_depsilon, go + nout, nout * sizeof(double),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(grad->dE_dsigma, go + 2 * nout, nout * sizeof(double),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (c->h_status->flags & ST_OVERFLOW)
        return fail(MSPMS_ENOMEM, "pair list overflow after growing");
    if (c->h_status->flags & ST_NONFINITE) return fail(MSPMS_ENOTFINITE, "non-finite position");
    return 0;
}

// One RESPA velocity-Verlet outer step, NVE (src/md.c:22-83), all buffers
// on the device, (3, N) SoA: pos, vel, long forces fl (in: forces of the
// current positions, out: of the new ones), short forces fs (same).
extern "C" int mspms_vver_step_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                      double* fs, double delt, int nstep_inner,
                                      mspms_long_terms* lterms, mspms_short_terms* sterms,
                                      double* ukin, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (nstep_inner < 1) return fail(MSPMS_EINVAL, "nstep_inner must be >= 1");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return enqueue_vver(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, s);
}

// Steps per captured graph of a device run: four steps in one graph drop
// three of the four graph boundaries per four steps (0.6 us per step at the
// 12,288-site water step of ~97 us, A/B on one box).
constexpr int VVER_GRAPH_STEPS = 4;

// nsteps RESPA outer steps; with use_graph VVER_GRAPH_STEPS steps (one step
// for shorter runs) are captured once as a HIP graph (on the context stream)
// and replayed on `stream`, the last nsteps % VVER_GRAPH_STEPS steps
// enqueued directly.  Optional histories receive each step's terms (nsteps
// * nrep entries) and kinetic energies (nsteps * nrep doubles).
extern "C" int mspms_vver_run_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                     double* fs, double delt, int nstep_inner, int nsteps,
                                     mspms_long_terms* lterms, mspms_short_terms* sterms,
                                     double* ukin, mspms_long_terms* lhist,
                                     mspms_short_terms* shist, double* ukin_hist, int use_graph,
                                     void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (nstep_inner < 1 || nsteps < 0) return fail(MSPMS_EINVAL, "bad step counts");
    if (!pos || !vel || !fl || !fs || !lterms || !sterms || !ukin)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const int nr = c->nrep;
    const int per = nsteps >= VVER_GRAPH_STEPS ? VVER_GRAPH_STEPS : 1;
    if (use_graph) {
        mspms_ctx::GraphKey k{pos, vel, fl, fs, lterms, sterms, ukin, delt, nstep_inner, per};
        int rc = ensure_graph(c, s, GRAPH_VVER, k, [&](hipStream_t cs) {
            for (int q = 0; q < per; q++) {
                int rq = enqueue_vver(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms,
                                      ukin, cs);
                if (rq) return rq;
            }
            return 0;
        });
        if (rc) return rc;
    }
    // without histories only the last step's terms are read: the steps
    // before it skip the finalize pass (its gate is set between launches,
    // outside any capture)
    const bool hist = lhist || shist || ukin_hist;
    const bool gate = !hist && nsteps > 1;
    // the gate is open and the histories are off again when this call
    // returns, whatever the path
    struct Reopen {
        mspms_ctx* c;
        hipStream_t s;
        bool armed, hist_on;
        ~Reopen() {
            if (armed) (void)hipMemsetD32Async((hipDeviceptr_t)c->d_fin_on.p, 1, 1, s);
            if (hist_on) {
                LaunchCtx L = launch_ctx(c, s);
                launch_set_hist(L, c->d_hist.p, HistDev{}, c->d_hstep.p);
            }
        }
    } reopen{c, s, false, false};
    if (gate) {
        HIPC(hipMemsetD32Async((hipDeviceptr_t)c->d_fin_on.p, 0, 1, s));
        reopen.armed = true;
    }
    // the overflow snapshot counts this run's steps
    HIPC(hipMemsetAsync(c->d_snap_count.p, 0, sizeof(int), s));
    HistDev hd{};
    if (hist && nsteps > 0) {
        // every step's finalize stores its workgroup sums in slot t; the
        // terms of all steps are formed after the last one
        const size_t nsum = (size_t)nsteps * nr * c->fin.nb;
        int rc = 0;
        if ((rc = c->d_hpart.ensure(nsum * FIN_ND)) || (rc = c->d_hpartu.ensure(nsum * 4)) ||
            (rc = c->d_huvac.ensure((size_t)nsteps * nr)))
            return rc;
        LaunchCtx L = launch_ctx(c, s);
        hd = HistDev{lhist, shist, ukin_hist, nr, nsteps, 1, c->d_hpart.p, c->d_hpartu.p,
                     c->fin.nb, c->d_hstep.p, c->any_vacuum ? c->d_huvac.p : nullptr};
        launch_set_hist(L, c->d_hist.p, hd, c->d_hstep.p);
        HIPC(hipGetLastError());
        reopen.hist_on = true;
    }
    for (int t = 0; t < nsteps;) {
        // a graph of `per` steps while that many remain, then single steps
        const bool grp = use_graph && nsteps - t >= per;
        const int adv = grp ? per : 1;
        if (gate && t + adv >= nsteps) {
            // the launch that holds the last step reduces its terms (the
            // other steps of a graph reduce theirs too: the last one wins)
            HIPC(hipMemsetD32Async((hipDeviceptr_t)c->d_fin_on.p, 1, 1, s));
            reopen.armed = false;
        }
        t += adv;
        if (grp) {
            HIPC(hipGraphLaunch(c->graph_exec, s));
        } else {
            int rc = enqueue_vver(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, s);
            if (rc) return rc;
        }
    }
    if (hd.on) {
        LaunchCtx L = launch_ctx(c, s);
        launch_hist_terms(L, hd, nsteps, lterms, sterms, ukin);
        HIPC(hipGetLastError());
    }
    return 0;
}

// ---- batched HMC ------------------------------------------------------------
namespace {

// rmarin (src/random.c:14-57): the 97 lags and the arithmetic sequence of
// Marsaglia and Zaman's generator from the two seeds (James' formulation)
void ranmar_seed(RanmarDev& r, int ij, int kl) {
    int i = (ij / 177) % 177 + 2, j = ij % 177 + 2, k = (kl / 169) % 178 + 1, l = kl % 169;
    r.u[0] = 0.0;
    for (int lag = 1; lag <= 97; lag++) {
        double s = 0.0, t = 0.5;
        for (int bit = 0; bit < 24; bit++) {
            const int m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t *= 0.5;
        }
        r.u[lag] = s;
    }
    r.c = 362436.0 / 16777216.0;
    r.cd = 7654321.0 / 16777216.0;
    r.cm = 16777213.0 / 16777216.0;
    r.i97 = 97;
    r.j97 = 33;
    r.iset = 0;
    r.pad = 0;
    r.gset = 0.0;
}

// one HMC move of every replica (kernels only: capturable)
int enqueue_hmc_move(mspms_ctx* c, double* pos, double* vel, double* fl, double* fs, double delt,
                     int nstep_inner, int nmd, mspms_long_terms* lterms,
                     mspms_short_terms* sterms, double* ukin, hipStream_t s) {
    const size_t n = (size_t)c->natom;
    LaunchCtx L = launch_ctx(c, s);
    launch_hmc_begin(L, c->d_hmc.p, c->d_rng.p, c->d_aw_atom.p, pos, c->d_pos_old.p, vel);
    int rc = enqueue_long(c, pos, pos + n, pos + 2 * n, fl, fl + n, fl + 2 * n, lterms, s);
    if (rc) return rc;
    rc = enqueue_short(c, pos, pos + n, pos + 2 * n, fs, fs + n, fs + 2 * n, sterms, s);
    if (rc) return rc;
    launch_hmc_mark(L, c->d_hmc.p, lterms, sterms);
    // with per-chain steps the integrator kernels read each replica's delt
    double* rep_delt = c->hmc_rep_delt ? c->d_hmc_delt.p : nullptr;
    struct Unset {
        mspms_ctx* c;
        ~Unset() { c->cur_rep_delt = nullptr; }
    } unset{c};
    c->cur_rep_delt = rep_delt;
    for (int t = 0; t < nmd; t++)
        if ((rc = enqueue_vver(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, s)))
            return rc;
    launch_hmc_end(L, c->d_hmc.p, c->d_rng.p, lterms, sterms, ukin, pos, c->d_pos_old.p,
                   rep_delt);
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" int mspms_hmc_init(mspms_ctx* c, const int* ij, const int* kl, const double* treq,
                              const int* nconstraint, double* d_vel) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!ij || !kl || !treq || !nconstraint || !d_vel) return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    const int nr = c->nrep, ns = (int)c->sp.size();
    std::vector<RanmarDev> rng(nr);
    std::vector<HmcRep> hmc(nr);
    std::vector<double> aw_atom;
    aw_atom.reserve(c->natom);
    for (int r = 0; r < nr; r++) {
        if (ij[r] < 0 || ij[r] > 31328 || kl[r] < 0 || kl[r] > 30081)
            return fail(MSPMS_EINVAL, "RANMAR seeds out of range (0..31328, 0..30081)");
        if (!(treq[r] > 0.0)) return fail(MSPMS_EINVAL, "treq must be positive");
        ranmar_seed(rng[r], ij[r], kl[r]);
        HmcRep h{};
        h.treq = treq[r];
        double mass = 0.0;
        int natom_r = 0;
        for (int k = 0; k < ns; k++) {
            const SpeciesHost& q = c->sp[k];
            const int nm = c->nmole[(size_t)r * ns + k];
            double mw = 0.0;
            for (int a = 0; a < q.natom; a++) mw += q.aw[a];
            mass += mw * nm;  // src/initvars.c:566-570
            natom_r += nm * q.natom;
            for (int m = 0; m < nm; m++)
                for (int a = 0; a < q.natom; a++) aw_atom.push_back(q.aw[a]);
        }
        h.system_mass = mass;
        h.nfree = 3 * natom_r - nconstraint[r];  // src/initvars.c:588
        if (h.nfree <= 0) return fail(MSPMS_EINVAL, "no degrees of freedom");
        hmc[r] = h;
    }
    int rc = 0;
    const hipStream_t s = c->stream;
    if ((rc = c->d_rng.upload(rng, s)) || (rc = c->d_hmc.upload(hmc, s)) ||
        (rc = c->d_aw_atom.upload(aw_atom, s)) ||
        (rc = c->d_pos_old.ensure((size_t)3 * c->natom)))
        return rc;
    drop_graph(c);
    LaunchCtx L = launch_ctx(c, s);
    launch_hmc_init(L, c->d_hmc.p, c->d_rng.p, c->d_aw_atom.p, d_vel);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s));
    c->hmc_ready = true;
    c->hmc_rep_delt = false;  // fresh chains step with the run's delt
    return 0;
}

extern "C" int mspms_hmc_set_delt(mspms_ctx* c, const double* delt, const int* adj_cycle,
                                  const double* rreq_disp) {
    if (!c || !c->hmc_ready) return fail(MSPMS_EINVAL, "mspms_hmc_init first");
    if (!delt) return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    const int nr = c->nrep;
    std::vector<HmcRep> h(nr);
    const hipStream_t s = c->stream;
    HIPC(hipMemcpyAsync(h.data(), c->d_hmc.p, sizeof(HmcRep) * nr, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    std::vector<double> d(delt, delt + nr);
    for (int r = 0; r < nr; r++) {
        if (!(d[r] > 0.0)) return fail(MSPMS_EINVAL, "replica %d: delt must be positive", r);
        const int cyc = adj_cycle ? adj_cycle[r] : 0;
        if (cyc < 0) return fail(MSPMS_EINVAL, "replica %d: negative adjustment cycle", r);
        h[r].adj_cycle = cyc;
        h[r].rreq = rreq_disp ? rreq_disp[r] : 0.0;
        h[r].adj_moves = 0;
        h[r].adj_accepted = 0;
    }
    int rc;
    if ((rc = c->d_hmc.upload(h, s)) || (rc = c->d_hmc_delt.upload(d, s))) return rc;
    HIPC(hipStreamSynchronize(s));
    drop_graph(c);
    c->hmc_rep_delt = true;
    return 0;
}

extern "C" int mspms_hmc_get_delt(mspms_ctx* c, double* delt) {
    if (!c || !c->hmc_ready || !c->hmc_rep_delt)
        return fail(MSPMS_EINVAL, "mspms_hmc_set_delt first");
    if (!delt) return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    HIPC(hipMemcpyAsync(delt, c->d_hmc_delt.p, sizeof(double) * c->nrep, hipMemcpyDeviceToHost,
                        c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mspms_hmc_run_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                    double* fs, double delt, int nstep_inner, int nmd, int nmove,
                                    mspms_long_terms* lterms, mspms_short_terms* sterms,
                                    double* ukin, int use_graph, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!c->hmc_ready) return fail(MSPMS_EINVAL, "mspms_hmc_init first");
    if (nstep_inner < 1 || nmd < 1 || nmove < 0) return fail(MSPMS_EINVAL, "bad step counts");
    if (!pos || !vel || !fl || !fs || !lterms || !sterms || !ukin)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (use_graph) {
        mspms_ctx::GraphKey k{pos, vel, fl, fs, lterms, sterms, ukin, delt,
                              nstep_inner * 65536 + nmd};
        int rc = ensure_graph(c, s, GRAPH_HMC, k, [&](hipStream_t cs) {
            return enqueue_hmc_move(c, pos, vel, fl, fs, delt, nstep_inner, nmd, lterms, sterms,
                                    ukin, cs);
        });
        if (rc) return rc;
        for (int m = 0; m < nmove; m++) HIPC(hipGraphLaunch(c->graph_exec, s));
        return 0;
    }
    for (int m = 0; m < nmove; m++) {
        int rc = enqueue_hmc_move(c, pos, vel, fl, fs, delt, nstep_inner, nmd, lterms, sterms,
                                  ukin, s);
        if (rc) return rc;
    }
    return 0;
}

// ---- Nose-Hoover NVT ----------------------------------------------------------
namespace {

// One vver_nh_3 step (src/nvtnh.c:11-221) of every replica on device (3, N)
// buffers; kernels only, so that it can be captured.
int enqueue_nvt(mspms_ctx* c, double* pos, double* vel, double* fl, double* fs, double delt,
                int nstep_inner, mspms_long_terms* lterms, mspms_short_terms* sterms,
                double* ukin, hipStream_t s) {
    const size_t n = (size_t)c->natom;
    // src/initvars.c:604-607, :329-330
    const double deltby2 = delt / 2.0, delts = delt / nstep_inner, deltsby2 = delts / 2.0;
    const double delt_sqby2 = delt * delt / 2.0, delts_sqby2 = delts * delts / 2.0;
    LaunchCtx L = launch_ctx(c, s);
    {
        ProfScope p(c, s, PROF_BONDED);
        for (int ll = 0; ll < nstep_inner; ll++) {
            launch_nvt_drift(L, c->d_nvt.p, ll == 0, pos, pos + n, pos + 2 * n, vel, vel + n,
                             vel + 2 * n, fl, fs, deltby2, delts, deltsby2, delts_sqby2,
                             c->d_active.p);
            launch_bonded(L, c->d_species.p, c->d_terms.p, pos, pos + n, pos + 2 * n, fs, fs + n,
                          fs + 2 * n, c->d_ep_short.p);
            launch_nvt_solve(L, c->d_nvt.p, false, c->d_nvt_carry.p, vel, vel + n, vel + 2 * n,
                             fs, c->d_nvt_vn.p, delts, deltsby2, delts_sqby2, nullptr,
                             c->d_active.p);
        }
    }
    int rc = enqueue_long(c, pos, pos + n, pos + 2 * n, fl, fl + n, fl + 2 * n, lterms, s,
                          nullptr, false);
    if (rc) return rc;
    launch_nvt_solve(L, c->d_nvt.p, true, c->d_nvt_carry.p, vel, vel + n, vel + 2 * n, fl,
                     c->d_nvt_vn.p, delt, deltby2, delt_sqby2, ukin, c->d_active.p);
    {
        ProfScope p(c, s, PROF_FINAL);
        launch_finalize(c, L, lterms, sterms, nullptr);
    }
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" int mspms_nvt_set(mspms_ctx* c, const mspms_nvt_state* st) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!st) return fail(MSPMS_EINVAL, "null argument");
    for (int r = 0; r < c->nrep; r++)
        if (!(st[r].qq > 0.0) || !(st[r].qqs > 0.0) || !(st[r].treq > 0.0) || st[r].nfree < 1)
            return fail(MSPMS_EINVAL, "replica %d: qq, qqs, treq and nfree must be positive", r);
    HIPC(hipSetDevice(c->device));
    int rc;
    if ((rc = c->d_nvt.ensure(c->nrep)) || (rc = c->d_nvt_carry.ensure(c->nrep)) ||
        (rc = c->d_nvt_vn.ensure((size_t)3 * c->natom)) ||
        (rc = c->d_active.upload(std::vector<int>(c->nrep, 1), c->stream)))
        return rc;
    HIPC(hipMemcpyAsync(c->d_nvt.p, st, sizeof(mspms_nvt_state) * c->nrep,
                        hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    c->nvt_ready = true;
    return 0;
}

extern "C" int mspms_nvt_get(mspms_ctx* c, mspms_nvt_state* st) {
    if (!c || !c->nvt_ready) return fail(MSPMS_EINVAL, "mspms_nvt_set first");
    if (!st) return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(st, c->d_nvt.p, sizeof(mspms_nvt_state) * c->nrep, hipMemcpyDeviceToHost));
    return 0;
}

// Simulated annealing (siman, src/siman.c:61-182) of every replica: cycles
// of velinit at T (each replica's RANMAR stream), the forces of the current
// positions and n_tries NVT steps from zeroed thermostats; T /= mu_t after
// each cycle, a replica stopping after its first cycle run below t_min (its
// later cycles leave it as it is).  Host-driven: one graph-replayed NVT run
// per cycle; needs mspms_hmc_init (streams, masses) and mspms_nvt_set.
extern "C" int mspms_siman_run_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                      double* fs, double delt, int nstep_inner, int n_tries,
                                      const double* t_initial, const double* t_min,
                                      const double* mu_t, int max_cycles, int* cycles,
                                      double* t_final, mspms_long_terms* lterms,
                                      mspms_short_terms* sterms, double* ukin, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (c->any_sf && c->sf_split)
        return fail(MSPMS_EINVAL, "the solid-fluid RESPA split is for the NVE step and HMC");
    if (!c->hmc_ready) return fail(MSPMS_EINVAL, "mspms_hmc_init first (velocity streams)");
    if (!c->nvt_ready) return fail(MSPMS_EINVAL, "mspms_nvt_set first");
    if (!pos || !vel || !fl || !fs || !t_initial || !t_min || !mu_t || !lterms || !sterms ||
        !ukin)
        return fail(MSPMS_EINVAL, "null argument");
    if (nstep_inner < 1 || n_tries < 0 || max_cycles < 0) return fail(MSPMS_EINVAL, "bad counts");
    const int nr = c->nrep;
    for (int r = 0; r < nr; r++)
        if (!(t_initial[r] > 0.0) || !(mu_t[r] > 1.0))
            return fail(MSPMS_EINVAL, "replica %d: t_initial must be positive, mu_t above 1", r);
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const size_t n = (size_t)c->natom;
    std::vector<double> T(t_initial, t_initial + nr);
    std::vector<int> act(nr, 1), cyc(nr, 0);
    std::vector<HmcRep> hmc(nr);
    std::vector<mspms_nvt_state> nvt(nr);
    int rc = 0;
    for (int k = 0; k < max_cycles && !rc; k++) {
        bool any = false;
        for (int r = 0; r < nr; r++) any = any || act[r];
        if (!any) break;
        HIPC(hipStreamSynchronize(s));
        HIPC(hipMemcpy(hmc.data(), c->d_hmc.p, sizeof(HmcRep) * nr, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(nvt.data(), c->d_nvt.p, sizeof(mspms_nvt_state) * nr,
                       hipMemcpyDeviceToHost));
        for (int r = 0; r < nr; r++) {
            if (!act[r]) continue;
            // src/siman.c:98 and the reset of :169-177 (init_nvt's at the start)
            hmc[r].treq = T[r];
            mspms_nvt_state& v = nvt[r];
            v.treq = T[r];
            v.ss = v.ps = v.sss = v.pss = v.unhts = v.unhtss = 0.0;
            v.gg = v.ggs = (double)v.nfree;
        }
        HIPC(hipMemcpy(c->d_hmc.p, hmc.data(), sizeof(HmcRep) * nr, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(c->d_nvt.p, nvt.data(), sizeof(mspms_nvt_state) * nr,
                       hipMemcpyHostToDevice));
        HIPC(hipMemcpy(c->d_active.p, act.data(), sizeof(int) * nr, hipMemcpyHostToDevice));
        LaunchCtx L = launch_ctx(c, s);
        launch_hmc_init(L, c->d_hmc.p, c->d_rng.p, c->d_aw_atom.p, vel, c->d_active.p);
        if ((rc = enqueue_long(c, pos, pos + n, pos + 2 * n, fl, fl + n, fl + 2 * n, lterms, s)))
            break;
        if ((rc = enqueue_short(c, pos, pos + n, pos + 2 * n, fs, fs + n, fs + 2 * n, sterms, s)))
            break;
        rc = mspms_nvt_run_device(c, pos, vel, fl, fs, delt, nstep_inner, n_tries, lterms,
                                  sterms, ukin, 1, s);
        for (int r = 0; r < nr; r++) {
            if (!act[r]) continue;
            cyc[r]++;
            if (T[r] < t_min[r])
                act[r] = 0;  // src/siman.c:141-144
            else
                T[r] /= mu_t[r];
        }
    }
    HIPC(hipStreamSynchronize(s));
    std::vector<int> ones(nr, 1);
    HIPC(hipMemcpy(c->d_active.p, ones.data(), sizeof(int) * nr, hipMemcpyHostToDevice));
    if (rc) return rc;
    for (int r = 0; r < nr; r++) {
        if (cycles) cycles[r] = cyc[r];
        if (t_final) t_final[r] = T[r];  // the last cycle's (the next one's if unfinished)
    }
    return 0;
}

extern "C" int mspms_nvt_run_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                    double* fs, double delt, int nstep_inner, int nsteps,
                                    mspms_long_terms* lterms, mspms_short_terms* sterms,
                                    double* ukin, int use_graph, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (c->any_sf && c->sf_split)
        return fail(MSPMS_EINVAL, "the solid-fluid RESPA split is for the NVE step and HMC");
    if (!c->nvt_ready) return fail(MSPMS_EINVAL, "mspms_nvt_set first");
    if (nstep_inner < 1 || nsteps < 0) return fail(MSPMS_EINVAL, "bad step counts");
    if (!pos || !vel || !fl || !fs || !lterms || !sterms || !ukin)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (use_graph) {
        mspms_ctx::GraphKey k{pos, vel, fl, fs, lterms, sterms, ukin, delt, nstep_inner};
        int rc = ensure_graph(c, s, GRAPH_NVT, k, [&](hipStream_t cs) {
            return enqueue_nvt(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, cs);
        });
        if (rc) return rc;
        for (int t = 0; t < nsteps; t++) HIPC(hipGraphLaunch(c->graph_exec, s));
        return 0;
    }
    for (int t = 0; t < nsteps; t++) {
        int rc = enqueue_nvt(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, s);
        if (rc) return rc;
    }
    return 0;
}

// ---- NPT -----------------------------------------------------------------------
namespace {

// One npt_respa step (src/nptnh.c:133-234) of every replica on device (3, N)
// buffers; lterms / sterms hold the terms of the current positions on
// entry (their virials drive the barostat) and of the new ones on exit.
int enqueue_npt(mspms_ctx* c, double* pos, double* vel, double* fl, double* fs, double delt,
                int nstep_inner, mspms_long_terms* lterms, mspms_short_terms* sterms,
                double* ukin, hipStream_t s) {
    const size_t n = (size_t)c->natom;
    // src/initvars.c:604-610
    const double deltby2 = delt / 2.0, delts = delt / nstep_inner, deltsby2 = delts / 2.0;
    const double dt_outer4 = delt / 4.0, dt_outer8 = delt / 8.0;
    LaunchCtx L = launch_ctx(c, s);
    NptDev* st = c->d_npt.p;
    {
        ProfScope p(c, s, PROF_BONDED);
        launch_npt_op(L, st, false, vel, fl, fs, lterms, sterms, deltby2, deltsby2, dt_outer4,
                      dt_outer8, nullptr);
        for (int ll = 0; ll < nstep_inner; ll++) {
            launch_npt_inner(L, c->d_reps.p, st, ll == 0, pos, vel, fs, delts, deltsby2);
            launch_bonded(L, c->d_species.p, c->d_terms.p, pos, pos + n, pos + 2 * n, fs, fs + n,
                          fs + 2 * n, c->d_ep_short.p);
        }
    }
    // the box changed: new Ewald weights, and the pair lists are rebuilt
    if (c->any_recip) launch_kvec(L, c->rb, c->rb.nrow);
    const bool force = c->force_rebuild;
    c->force_rebuild = true;
    int rc = enqueue_long(c, pos, pos + n, pos + 2 * n, fl, fl + n, fl + 2 * n, lterms, s,
                          nullptr, false);
    c->force_rebuild = force;
    if (rc) return rc;
    {
        ProfScope p(c, s, PROF_FINAL);
        launch_finalize(c, L, lterms, sterms, nullptr);
    }
    launch_npt_op(L, st, true, vel, fl, fs, lterms, sterms, deltby2, deltsby2, dt_outer4,
                  dt_outer8, ukin);
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" int mspms_npt_set(mspms_ctx* c, const mspms_npt_state* st) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!st) return fail(MSPMS_EINVAL, "null argument");
    const int ns = (int)c->sp.size();
    if (ns * ns > NPT_MAX_LRC) return fail(MSPMS_EINVAL, "NPT supports at most 4 species");
    for (int r = 0; r < c->nrep; r++)
        if (!(st[r].Qts > 0.0) || !(st[r].Qbs > 0.0) || !(st[r].treq > 0.0) || st[r].nfree < 1)
            return fail(MSPMS_EINVAL, "replica %d: Qts, Qbs, treq and nfree must be positive", r);
    // every replica back to the box of its parameters (an earlier NPT run
    // moved the device copy)
    int rc = build(c);
    if (rc) return rc;
    std::vector<NptDev> h(c->nrep);
    for (int r = 0; r < c->nrep; r++) {
        NptDev& D = h[r];
        D = NptDev{};
        D.Qts = st[r].Qts;
        D.Qbs = st[r].Qbs;
        D.treq = st[r].treq;
        D.preq = st[r].preq;
        D.vts = st[r].vts;
        D.vbs = st[r].vbs;
        D.rts = st[r].rts;
        D.nfree = st[r].nfree;
        const mspms_params& P = c->params[r];
        D.boxv = P.boxlx * P.boxly * P.boxlz;
        if (st[r].lj_tail) {
            // InitLJlrcCommonTerms (src/initvars.c:474-521) and the
            // molecule-count products of calculate_ljlrc (src/misc.c:131-148)
            const double pi = 3.141592653589793;
            const double t1 = (32.0 / 9.0) * pi * std::pow(P.rcutoff, -9.0);
            const double t2 = -(16.0 / 3.0) * pi * std::pow(P.rcutoff, -3.0);
            for (int m = 0; m < ns; m++)
                for (int nn = 0; nn < ns; nn++) {
                    double term = 0.0;
                    const SpeciesHost &A = c->sp[m], &B = c->sp[nn];
                    for (int i = 0; i < A.natom; i++)
                        for (int j = 0; j < B.natom; j++) {
                            const double sij = 0.5 * (A.sigma[i] + B.sigma[j]);
                            const double eij = std::sqrt(A.epsilon[i] * B.epsilon[j]);
                            const double temp2 = eij * std::pow(sij, 3.0);
                            const double temp3 = std::pow(sij, 9.0) * t1 + std::pow(sij, 3.0) * t2;
                            term += temp3 * temp2;
                        }
                    D.lrc[D.nlrc++] = term * c->nmole[(size_t)r * ns + m] *
                                      c->nmole[(size_t)r * ns + nn];
                }
            double p = 0.0;
            for (int k = 0; k < D.nlrc; k++) p += D.lrc[k] / D.boxv / D.boxv;
            D.pljlrc = p;
        }
    }
    HIPC(hipSetDevice(c->device));
    if ((rc = c->d_npt.upload(h, c->stream))) return rc;
    HIPC(hipStreamSynchronize(c->stream));
    c->npt_ready = true;
    return 0;
}

extern "C" int mspms_npt_get(mspms_ctx* c, mspms_npt_state* st) {
    if (!c || !c->npt_ready) return fail(MSPMS_EINVAL, "mspms_npt_set first");
    if (!st) return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    HIPC(hipDeviceSynchronize());
    std::vector<NptDev> h(c->nrep);
    std::vector<RepDev> reps(c->nrep);
    HIPC(hipMemcpy(h.data(), c->d_npt.p, sizeof(NptDev) * c->nrep, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(reps.data(), c->d_reps.p, sizeof(RepDev) * c->nrep, hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; r++) {
        const NptDev& D = h[r];
        mspms_npt_state& o = st[r];
        o.Qts = D.Qts;
        o.Qbs = D.Qbs;
        o.treq = D.treq;
        o.preq = D.preq;
        o.vts = D.vts;
        o.vbs = D.vbs;
        o.rts = D.rts;
        o.pljlrc = D.pljlrc;
        o.utsbs = D.utsbs;
        o.ukin = D.ukin;
        o.tinst = D.tinst;
        o.pinst = D.pinst;
        for (int d = 0; d < 3; d++) o.boxl[d] = reps[r].L[d];
        o.boxv = D.boxv;
        o.nfree = D.nfree;
        o.lj_tail = D.nlrc > 0;
    }
    return 0;
}

extern "C" int mspms_npt_run_device(mspms_ctx* c, double* pos, double* vel, double* fl,
                                    double* fs, double delt, int nstep_inner, int nsteps,
                                    mspms_long_terms* lterms, mspms_short_terms* sterms,
                                    double* ukin, int use_graph, void* stream) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (c->any_sf && c->sf_split)
        return fail(MSPMS_EINVAL, "the solid-fluid RESPA split is for the NVE step and HMC");
    if (!c->npt_ready) return fail(MSPMS_EINVAL, "mspms_npt_set first");
    if (nstep_inner < 1 || nsteps < 0) return fail(MSPMS_EINVAL, "bad step counts");
    if (!pos || !vel || !fl || !fs || !lterms || !sterms || !ukin)
        return fail(MSPMS_EINVAL, "null argument");
    HIPC(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (use_graph) {
        mspms_ctx::GraphKey k{pos, vel, fl, fs, lterms, sterms, ukin, delt, nstep_inner};
        int rc = ensure_graph(c, s, GRAPH_NPT, k, [&](hipStream_t cs) {
            return enqueue_npt(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, cs);
        });
        if (rc) return rc;
        for (int t = 0; t < nsteps; t++) HIPC(hipGraphLaunch(c->graph_exec, s));
        return 0;
    }
    for (int t = 0; t < nsteps; t++) {
        int rc = enqueue_npt(c, pos, vel, fl, fs, delt, nstep_inner, lterms, sterms, ukin, s);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int mspms_hmc_stats(mspms_ctx* c, long long* accepted, long long* moves,
                               double* last_dH) {
    if (!c || !c->hmc_ready) return fail(MSPMS_EINVAL, "mspms_hmc_init first");
    HIPC(hipSetDevice(c->device));
    HIPC(hipDeviceSynchronize());
    std::vector<HmcRep> h(c->nrep);
    HIPC(hipMemcpy(h.data(), c->d_hmc.p, sizeof(HmcRep) * c->nrep, hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; r++) {
        if (accepted) accepted[r] = h[r].accepted;
        if (moves) moves[r] = h[r].moves;
        if (last_dH) last_dH[r] = h[r].last_dH;
    }
    return 0;
}

extern "C" int mspms_hmc_draw(mspms_ctx* c, int replica, int n, double* out) {
    if (!c || !c->hmc_ready) return fail(MSPMS_EINVAL, "mspms_hmc_init first");
    if (replica < 0 || replica >= c->nrep || n < 0 || !out)
        return fail(MSPMS_EINVAL, "bad arguments");
    if (n == 0) return 0;
    HIPC(hipSetDevice(c->device));
    DevBuf<double> tmp;
    int rc = tmp.ensure((size_t)n);
    if (rc) return rc;
    LaunchCtx L = launch_ctx(c, c->stream);
    launch_hmc_draw(L, c->d_rng.p, replica, n, tmp.p);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, tmp.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    tmp.release();
    return 0;
}

// ---- autotuning -------------------------------------------------------------
// Times frclong on the given device positions for each launch shape
// (sites per binning cell x waves per i-cluster of the mixed pair kernel),
// keeps the fastest.  Every shape gives the same result up to summation
// order; only the time differs.
extern "C" int mspms_autotune(mspms_ctx* c, const double* x, const double* y, const double* z,
                              int repeats, mspms_tune_result* out) {
    if (!c || c->sp.empty()) return fail(MSPMS_EINVAL, "context has no system");
    if (!x || !y || !z || repeats < 1) return fail(MSPMS_EINVAL, "bad arguments");
    HIPC(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    // the caller's settings come back on every exit path; the shape only
    // changes when a tuned shape has been built
    struct Restore {
        mspms_ctx* c;
        bool prof, rebuild;
        double cell;
        int wpc;
        bool keep_shape = false;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Restore() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            c->profiling = prof;
            c->force_rebuild = rebuild;
            if (!keep_shape && (c->cell_atoms != cell || c->pair_wpc != wpc)) {
                c->cell_atoms = cell;
                c->pair_wpc = wpc;
                (void)build(c);
            }
        }
    } g{c, c->profiling, c->force_rebuild, c->cell_atoms, c->pair_wpc};
    c->profiling = false;
    // every timed call builds its list, as a call after a move would
    c->force_rebuild = true;
    HIPC(hipEventCreate(&g.e0));
    HIPC(hipEventCreate(&g.e1));
    hipEvent_t e0 = g.e0, e1 = g.e1;
    const double cells[] = {6.0, 8.0, 12.0, 16.0};
    const int wpcs[] = {1, 2, 4, 8};
    const int nw = 4;
    mspms_tune_result res{};
    double best = 1e300;
    int rc = 0;
    for (double cell : cells) {
        c->cell_atoms = cell;
        if ((rc = build(c))) break;
        for (int iw = 0; iw < nw && !rc; iw++) {
            c->pair_wpc = wpcs[iw];
            const size_t n = (size_t)c->natom;
            double *fx = c->d_fl.p, *fy = fx + n, *fz = fy + n;
            // one untimed call, regrowing the lists until they fit
            for (int attempt = 0; attempt < 4; attempt++) {
                hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, c->d_status.p);
                if ((rc = enqueue_long(c, x, y, z, fx, fy, fz, c->d_long.p, s))) break;
                if ((rc = read_status(c, s))) break;
                if (!(c->h_status->flags & ST_OVERFLOW)) break;
                c->cap_hint = std::max(c->cap_hint * 2, c->h_status->max_list + 64);
                if ((rc = build(c))) break;
                fx = c->d_fl.p;
                fy = fx + n;
                fz = fy + n;
            }
            if (rc) break;
            HIPC(hipEventRecord(e0, s));
            for (int r = 0; r < repeats && !rc; r++)
                rc = enqueue_long(c, x, y, z, fx, fy, fz, c->d_long.p, s);
            HIPC(hipEventRecord(e1, s));
            HIPC(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, e0, e1));
            const double t = ms / repeats;
            if (res.ncand < MSPMS_TUNE_MAX) {
                res.cand_cell_atoms[res.ncand] = cell;
                res.cand_wpc[res.ncand] = c->pair_wpc;
                res.cand_ms[res.ncand] = t;
                res.ncand++;
            }
            if (t < best) {
                best = t;
                res.cell_atoms = cell;
                res.waves_per_cluster = c->pair_wpc;
                res.ms = t;
            }
        }
        if (rc) break;
    }
    if (rc) return rc;
    c->cell_atoms = res.cell_atoms;
    c->pair_wpc = res.waves_per_cluster;
    if ((rc = build(c))) return rc;
    g.keep_shape = true;
    if (out) *out = res;
    return 0;
}

extern "C" int mspms_set_launch_shape(mspms_ctx* c, double cell_atoms, int waves_per_cluster) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    if (!(cell_atoms >= 2.0 && cell_atoms <= 64.0))
        return fail(MSPMS_EINVAL, "cell_atoms must be within 2..64");
    if (waves_per_cluster != 1 && waves_per_cluster != 2 && waves_per_cluster != 4 &&
        waves_per_cluster != 8)
        return fail(MSPMS_EINVAL, "waves_per_cluster must be 1, 2, 4 or 8");
    c->cell_atoms = cell_atoms;
    c->pair_wpc = waves_per_cluster;
    return c->sp.empty() ? 0 : build(c);
}

extern "C" int mspms_set_list_capacity(mspms_ctx* c, int max_entries) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    if (max_entries < 0) return fail(MSPMS_EINVAL, "negative capacity");
    c->cap_limit = max_entries;
    return c->sp.empty() ? 0 : build(c);
}

extern "C" int mspms_set_sf_split(mspms_ctx* c, int on) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    if (c->sf_split != (on != 0)) drop_graph(c);  // captured steps change
    c->sf_split = on != 0;
    return 0;
}

extern "C" int mspms_set_list_skin(mspms_ctx* c, double skin) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    if (!(skin >= 0.0 && skin <= 8.0)) return fail(MSPMS_EINVAL, "skin must be within 0..8");
    c->skin = skin;
    return c->sp.empty() ? 0 : build(c);
}

extern "C" int mspms_list_builds(mspms_ctx* c, long long* builds) {
    if (!c || !builds) return fail(MSPMS_EINVAL, "null argument");
    *builds = 0;
    if (c->sp.empty() || !c->cb.lstate) return 0;
    HIPC(hipSetDevice(c->device));
    int st[LS_N];
    HIPC(hipMemcpyAsync(st, c->cb.lstate, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    *builds = (long long)(unsigned)st[LS_BUILDS];
    return 0;
}

// ---- profiling ------------------------------------------------------------
extern "C" int mspms_set_profiling(mspms_ctx* c, int on) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    HIPC(hipSetDevice(c->device));
    HIPC(hipDeviceSynchronize());
    c->profiling = on != 0;
    c->pending.clear();
    c->ev_used = 0;
    c->prof_ms.assign(PROF_N, 0.0);
    c->prof_n.assign(PROF_N, 0);
    return 0;
}

// ms[k] = mean milliseconds per call of kernel group k since profiling was
// switched on: 0 clustering + pair list, 1 nonbonded pairs, 2 intramolecular,
// 3 reciprocal S(k) + reduction, 4 reciprocal forces, 5 finalize,
// 6 bonded / RESPA inner loop.  Returns the number of groups written.
extern "C" int mspms_last_timings(const mspms_ctx* c, double* ms, int n) {
    if (!c || !ms) return 0;
    int k = 0;
    for (; k < n && k < PROF_N; k++)
        ms[k] = (k < (int)c->prof_n.size() && c->prof_n[k]) ? c->prof_ms[k] / c->prof_n[k] : 0.0;
    return k;
}

// Synchronises the context stream, accumulates profiling and reports the
// status of the last device-API evaluation.
extern "C" int mspms_sync(mspms_ctx* c) {
    if (!c) return fail(MSPMS_EINVAL, "null ctx");
    HIPC(hipSetDevice(c->device));
    // device-buffer calls may have run on a caller's stream
    HIPC(hipDeviceSynchronize());
    int rc = read_status(c, c->stream);
    if (rc) return rc;
    accumulate_profile(c);
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, c->stream, c->d_status.p);
    HIPC(hipStreamSynchronize(c->stream));
    int steps_ok = 0;
    HIPC(hipMemcpy(&steps_ok, c->d_snap_count.p, sizeof(int), hipMemcpyDeviceToHost));
    HIPC(hipMemsetAsync(c->d_snap_count.p, 0, sizeof(int), c->stream));
    if (c->h_status->flags & ST_OVERFLOW) {
        // the evaluations since the last sync dropped pairs.  A RESPA run
        // gets back the positions and velocities it had at the start of the
        // first step that overflowed (the end of the last good step); then
        // the lists are sized for the largest row seen, so that a repeat of
        // the steps fits
        bool restored = false;
        if (c->snap_pos && c->snap_vel && steps_ok > 0) {
            const size_t n3 = (size_t)3 * c->natom;
            HIPC(hipMemcpy(c->snap_pos, c->d_snap.p, n3 * sizeof(double),
                           hipMemcpyDeviceToDevice));
            HIPC(hipMemcpy(c->snap_vel, c->d_snap.p + n3, n3 * sizeof(double),
                           hipMemcpyDeviceToDevice));
            restored = true;
        }
        c->snap_pos = c->snap_vel = nullptr;
        c->cap_hint = std::max(c->cap_hint * 2, c->h_status->max_list + 64);
        c->cap_limit = 0;
        drop_graph(c);
        if ((rc = build(c))) return rc;
        if (restored)
            return fail(MSPMS_ENOMEM,
                        "pair list overflow: positions and velocities restored to the start of "
                        "the RESPA step that overflowed (step %d since the last sync or run "
                        "start); forces and terms are stale; lists regrown",
                        steps_ok);
        return fail(MSPMS_ENOMEM, "pair list overflow (lists regrown; repeat the calls)");
    }
    if (c->h_status->flags & ST_NONFINITE) return fail(MSPMS_ENOTFINITE, "non-finite position");
    if (c->h_status->flags & ST_GRID_EDOM)
        return fail(MSPMS_EDOM, "site outside the solid-fluid grid or without a grid type");
    if (c->h_status->flags & ST_SINGULAR) return fail(MSPMS_EINVAL, "singular improper");
    return keep_list_headroom(c);
}

extern "C" void* mspms_stream(const mspms_ctx* c) { return c ? (void*)c->stream : nullptr; }
