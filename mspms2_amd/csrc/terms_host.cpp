// terms_host.cpp -- term-level entry points of the C ABI, evaluated on the
// host by the same term code the gfx950 kernels run (physics.h, terms.h),
// with the reference's run globals passed as an mspms_params block:
//   mspms_term_pair    ljfrc (src/ljfrc.c:21-71), ewald_real_frc,
//                      ewald_excl_frc, wolf_real_frc (src/elecfrc.c:17-60)
//   mspms_term_bonded  bndfrc / aglfrc / dihfrc / impfrc (src/bndfrc.c:17-78,
//                      src/aglfrc.c:17-190, src/dihfrc.c:17-284,
//                      src/impfrc.c:17-191) on one molecule's coordinates
//   mspms_term_sf      cal_sf_hypergeo / cal_sf_atom_explicit for one site
//                      (src/sffrc.c:1118-1314)
//   mspms_term_myinterp  get_values_from_grid (src/sffrc.c:951-1089)
//   mspms_term_hermite   call_tasos_forces (src/sffrc.c:1322-1337) over the
//                      tricubic Hermite grid (mylibtasos/forces.f90:6-173)
// They exist for term-level parity tests and for callers that evaluate
// single terms; the force evaluation itself runs on the GPU only.
#include <algorithm>
#include <mutex>
#include <vector>

#include "engine.h"
#include "physics.h"
#include "sf_eval.h"
#include "terms.h"
#include "../../include/mspms_gpu.h"

int mspms_fail(int code, const char* what);

extern "C" int mspms_term_pair(const mspms_params* p, int which, double rijsq, double a,
                               double b, double* uij, double* fij, double* ushift) {
    if (!p || !uij || !fij || !ushift) return mspms_fail(MSPMS_EINVAL, "null argument");
    *ushift = 0.0;
    switch (which) {
    case MSPMS_TERM_LJ:
        lj64(rijsq, a, b, p->isLJswitchOn, p->rcutoffsq, p->rcutonsq, p->roff2_minus_ron2_cube,
             p->shift1, uij, fij, ushift);
        return 0;
    case MSPMS_TERM_EWALD_REAL:
        ewald_real64(rijsq, a, p->coulomb_prefactor, p->kappa, uij, fij);
        return 0;
    case MSPMS_TERM_EWALD_EXCL:
        ewald_excl64(rijsq, a, p->coulomb_prefactor, uij, fij);
        return 0;
    case MSPMS_TERM_WOLF_REAL:
        wolf_real64(rijsq, a, p->coulomb_prefactor, p->kappa, p->rcutoffelec, p->wolfvcon1,
                    p->wolfvcon2, p->wolffcon1, p->wolffcon2, uij, fij);
        return 0;
    case MSPMS_TERM_COULOMB:
        coulomb64(rijsq, a, p->coulomb_prefactor, uij, fij);
        return 0;
    default:
        return mspms_fail(MSPMS_EINVAL, "unknown pair term");
    }
}

extern "C" int mspms_term_bonded(const mspms_species* q, int kind, int k, const double* xx,
                                 const double* yy, const double* zz, double* fxs, double* fys,
                                 double* fzs, double* u, double* vir) {
    if (!q || !xx || !yy || !zz || !fxs || !fys || !fzs || !u || !vir)
        return mspms_fail(MSPMS_EINVAL, "null argument");
    TermDev T{};
    switch (kind) {
    case MSPMS_TERM_BOND:
        if (k < 0 || k >= q->nbond) return mspms_fail(MSPMS_EINVAL, "bond index");
        T.kind = TERM_BOND;
        T.type = q->bnd_type[k];
        for (int i = 0; i < 2; i++) T.idx[i] = q->bnd_idx[2 * k + i];
        for (int i = 0; i < 3; i++) T.para[i] = q->bnd_para[3 * k + i];
        break;
    case MSPMS_TERM_ANGLE:
        if (k < 0 || k >= q->nangle) return mspms_fail(MSPMS_EINVAL, "angle index");
        T.kind = TERM_ANGLE;
        T.type = q->agl_type[k];
        for (int i = 0; i < 3; i++) T.idx[i] = q->agl_idx[3 * k + i];
        for (int i = 0; i < 5; i++) T.para[i] = q->agl_para[5 * k + i];
        break;
    case MSPMS_TERM_DIH:
        if (k < 0 || k >= q->ndih) return mspms_fail(MSPMS_EINVAL, "dihedral index");
        T.kind = TERM_DIH;
        T.type = q->dih_type[k];
        for (int i = 0; i < 4; i++) T.idx[i] = q->dih_idx[4 * k + i];
        for (int i = 0; i < 4; i++) T.para[i] = q->dih_para[4 * k + i];
        break;
    case MSPMS_TERM_IMP:
        if (k < 0 || k >= q->nimp) return mspms_fail(MSPMS_EINVAL, "improper index");
        T.kind = TERM_IMP;
        T.type = q->imp_type[k];
        for (int i = 0; i < 4; i++) T.idx[i] = q->imp_idx[4 * k + i];
        for (int i = 0; i < 2; i++) T.para[i] = q->imp_para[2 * k + i];
        break;
    default:
        return mspms_fail(MSPMS_EINVAL, "unknown bonded term");
    }
    int singular = 0;
    *u = 0.0;
    *vir = 0.0;
    bonded_term(T, 0, xx, yy, zz, fxs, fys, fzs, u, vir, &singular);
    if (singular)
        return mspms_fail(MSPMS_EINVAL, "improper at a singular geometry (src/impfrc.c:96-134)");
    return 0;
}

extern "C" int mspms_term_sf(const mspms_params* p, const mspms_species* sp, int site, double x,
                             double y, double z, double* fx, double* fy, double* fz, double* u,
                             double* fij) {
    if (!p || !sp || !fx || !fy || !fz || !u || !fij)
        return mspms_fail(MSPMS_EINVAL, "null argument");
    if (site < 0 || site >= sp->natom) return mspms_fail(MSPMS_EINVAL, "site index");
    SFDev F{};
    F.type = p->iSF_type;
    std::vector<double> tab;
    if (p->iSF_type == MSPMS_SF_HYPERGEO) {
        F.ntube = p->ntube;
        for (int t = 0; t < p->ntube; t++) {
            tab.push_back(p->tube_x[t]);
            tab.push_back(p->tube_y[t]);
            tab.push_back(p->tube_r[t]);
        }
        F.tube = tab.data();
        F.c3pisq = p->const_3pisq_theta;
    } else if (p->iSF_type == MSPMS_SF_ATOM_EXPLICIT) {
        F.nsolid = p->solid_natom;
        for (int k = 0; k < p->solid_natom; k++) {
            tab.push_back(p->solid_xx[k]);
            tab.push_back(p->solid_yy[k]);
            tab.push_back(p->solid_zz[k]);
        }
        F.solid = tab.data();
    } else {
        return mspms_fail(MSPMS_EINVAL, "mspms_term_sf: iSF_type 1 or 2");
    }
    F.solid_sigma = p->solid_sigma;
    F.solid_epsilon = p->solid_epsilon;
    RepDev R{};
    R.L[0] = p->boxlx;
    R.L[1] = p->boxly;
    R.L[2] = p->boxlz;
    R.rc2_lj = p->rcutoffsq;
    R.rcon2 = p->rcutonsq;
    R.roff2_ron2_cube = p->roff2_minus_ron2_cube;
    R.switch_on = p->isLJswitchOn;
    SiteDev S{};
    S.sigma = sp->sigma[site];
    S.epsilon = sp->epsilon[site];
    S.ghost = sp->ghost_type[site];
    int edom = 0;
    double gx = *fx, gy = *fy, gz = *fz;
    *u = sfeval::sf_site(F, R, S, x, y, z, gx, gy, gz, &edom, fij);
    *fx = gx;
    *fy = gy;
    *fz = gz;
    return 0;
}

extern "C" int mspms_term_myinterp(const mspms_params* p, double x, double y, double z, int type,
                                   double* usf, double* fsf) {
    if (!p || !p->cubes || !usf || !fsf) return mspms_fail(MSPMS_EINVAL, "null argument");
    const mspms_sf_cubes& Q = *p->cubes;
    if (type < 0 || type >= Q.ntype) return mspms_fail(MSPMS_EINVAL, "myinterp table type");
    double X, Y, Z;
    const int cube = sfeval::mi_cube(x, y, z, p->sigma_base, Q.xmin, Q.ymin, Q.zmax, Q.hx, Q.hy,
                                     Q.hz, Q.ncx, Q.ncy, Q.ncz, &X, &Y, &Z);
    if (cube < 0) return mspms_fail(MSPMS_EDOM, "position outside the myinterp tables");
    // the file's order: table (type * 128 + polynomial * 32 + coefficient)[cube]
    const double* const* T = Q.tables + (size_t)type * 128;
    double r[4];
    sfeval::mi_poly([T, cube](int f, int m) { return T[f * 32 + m][cube]; }, X, Y, Z, r);
    *usf = r[0];
    fsf[0] = r[1];
    fsf[1] = r[2];
    fsf[2] = r[3];
    return 0;
}

// The Hermite grid of p->grid as the device holds it (node-major 32-B
// records, minimum node F per type), converted once per grid and kept for
// later calls on the same node table.
namespace {
struct HostGrid {
    const float* src = nullptr;
    size_t count = 0;
    std::vector<float> nodes;
    std::vector<double> min_pe;
};
std::mutex g_host_grid_mu;
HostGrid g_host_grid;
}  // namespace

extern "C" int mspms_term_hermite(const mspms_params* p, int type, double x, double y, double z,
                                  double* u, double* force3) {
    if (!p || !u || !force3) return mspms_fail(MSPMS_EINVAL, "null argument");
    const mspms_sf_grid* G = p->grid;
    if (!G || !G->nodes || G->ntype < 1 || G->nx < 2 || G->ny < 2 || G->nz < 2 ||
        !(G->hx > 0) || !(G->hy > 0) || !(G->hz > 0) || !(G->zperiod >= 0.0) ||
        G->zperiod > (G->nz - 1) * G->hz)
        return mspms_fail(MSPMS_EINVAL, "mspms_term_hermite: grid missing or malformed");
    const size_t nn = (size_t)G->nx * G->ny * G->nz, tot = nn * 8 * (size_t)G->ntype;
    std::lock_guard<std::mutex> lock(g_host_grid_mu);
    HostGrid& H = g_host_grid;
    if (H.src != G->nodes || H.count != tot) {
        static const int field_of[8] = {0, 1, 2, 4, 3, 5, 6, 7};
        H.nodes.assign(tot, 0.f);
        H.min_pe.assign(G->ntype, 10e10);  // grid.f90:61
        for (int t = 0; t < G->ntype; t++)
            for (size_t k = 0; k < nn; k++) {
                for (int o = 0; o < 8; o++)
                    H.nodes[((size_t)t * nn + k) * 8 + o] =
                        G->nodes[((size_t)t * 8 + field_of[o]) * nn + k];
                H.min_pe[t] = std::min(H.min_pe[t], (double)G->nodes[(size_t)t * 8 * nn + k]);
            }
        H.src = G->nodes;
        H.count = tot;
    }
    SFDev F{};
    F.type = MSPMS_SF_TASOS;
    F.ntype = G->ntype;
    F.nx = G->nx;
    F.ny = G->ny;
    F.nz = G->nz;
    F.x0 = G->x0;
    F.y0 = G->y0;
    F.z0 = G->z0;
    F.hx = G->hx;
    F.hy = G->hy;
    F.hz = G->hz;
    F.zperiod = G->zperiod > 0.0 ? G->zperiod : (G->nz - 1) * G->hz;
    F.sigma_base = p->sigma_base;
    F.epsilon_base = p->epsilon_base;
    F.nodes = H.nodes.data();
    F.min_pe = H.min_pe.data();
    RepDev R{};
    SiteDev S{};
    S.ghost = MSPMS_GHOST_NONE;
    S.interp = type;
    int edom = 0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const double pe = sfeval::sf_site(F, R, S, x, y, z, fx, fy, fz, &edom);
    if (edom) return mspms_fail(MSPMS_EDOM, "site outside the Hermite grid or without a grid type");
    *u = pe;
    force3[0] = fx;
    force3[1] = fy;
    force3[2] = fz;
    return 0;
}
