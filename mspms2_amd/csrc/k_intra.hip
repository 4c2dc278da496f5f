// k_intra.hip -- per-molecule kernels (gfx950), FP64 in every precision
// mode:
//   k_intra          intramolecular corrections of frclong: bonded 1-2 and
//                    unique 1-3 ending pairs (Ewald exclusion + real space,
//                    src/frclong.c:234-593), unique 1-4 pairs (f0-scaled LJ +
//                    real space, :597-787)
//   k_bonded         frcshort: bonds, angles, dihedrals, impropers
//                    (src/frcshort.c:20-97 and the term files)
//   k_finalize       per-replica deterministic reduction of every partial
//   k_finalize_short   sum and the reference's energy bookkeeping (one
//                    RESPA step: long, short and kinetic in one pass)
//                    (src/frclong.c:985, :1048-1061, :1128, :1142, :1223)
// One thread owns one molecule and therefore its atoms: forces are
// read-modify-written without atomics, in the reference's term order.
// These terms are O(N_bonded) and carry the largest cancellations of the
// electrostatics (exclusion vs real space), so they stay FP64.
#include "engine.h"
#include "physics.h"
#include "terms.h"
#include "wave.h"
#include "../../include/mspms_gpu.h"

namespace {

struct IntraAcc {
    double uvdw, ureal, uexcl, vir;
};

__device__ inline double mic(double d, double L) {
#pragma clang fp contract(off)
    return d - L * rint(d / L);
}

__device__ inline void add_pair(double* fx, double* fy, double* fz, int ii, int jj, double f,
                                double rx, double ry, double rz) {
#pragma clang fp contract(off)
    double ax = f * rx, ay = f * ry, az = f * rz;
    fx[ii] += ax;
    fy[ii] += ay;
    fz[ii] += az;
    fx[jj] -= ax;
    fy[jj] -= ay;
    fz[jj] -= az;
}

// per-atom parameter-derivative outputs of the adjoint (null: not wanted)
struct GradOut {
    double *gq, *ge, *gs;
};

template <bool GRAD>
__device__ inline void intra_lj(const RepDev& R, const SiteDev& a, const SiteDev& b, int ii,
                                int jj, double rx, double ry, double rz, double escale,
                                IntraAcc& acc, double* fx, double* fy, double* fz,
                                const GradOut& G) {
#pragma clang fp contract(off)
    double rsq = rx * rx + ry * ry + rz * rz;
    if (!(rsq < R.rc2_lj)) return;
    double u, f, ush;
    const double e0 = sqrt(a.epsilon * b.epsilon), sij = 0.5 * (a.sigma + b.sigma);
    lj64(rsq, sij, escale * e0, R.switch_on, R.rc2_lj, R.rcon2, R.roff2_ron2_cube, R.shift1, &u,
         &f, &ush);
    acc.uvdw += u;
    acc.vir += f * rsq;
    add_pair(fx, fy, fz, ii, jj, f, rx, ry, rz);
    if (GRAD) {
        double dds, dde;
        lj64_dparam(rsq, sij, escale * e0, R.switch_on, R.rc2_lj, R.rcon2, R.roff2_ron2_cube,
                    &dds, &dde);
        G.gs[ii] += 0.5 * dds;
        G.gs[jj] += 0.5 * dds;
        G.ge[ii] += dde * escale * deps_mix(e0, b.epsilon);
        G.ge[jj] += dde * escale * deps_mix(e0, a.epsilon);
    }
}

// electrostatics of an intramolecular pair (src/frclong.c:316-408 with
// with_excl, :708-786 without)
template <bool GRAD>
__device__ inline void intra_elec(const RepDev& R, double qq, double qa, double qb, int ii,
                                  int jj, double rx, double ry, double rz, bool with_excl,
                                  IntraAcc& acc, double* fx, double* fy, double* fz,
                                  const GradOut& G) {
#pragma clang fp contract(off)
    const int ct = R.charge_type;
    if (ct != MSPMS_ELEC_EWALD && ct != MSPMS_ELEC_WOLF) return;
    double rsq, u, f;
    if (with_excl) {
        rsq = rx * rx + ry * ry + rz * rz;
        if (rsq < R.rc2_el) {
            ewald_excl64(rsq, qq, R.C, &u, &f);
            acc.uexcl += u;
            if (ct == MSPMS_ELEC_WOLF) acc.vir += f * rsq;
            add_pair(fx, fy, fz, ii, jj, f, rx, ry, rz);
            if (GRAD) {
                // uexcl enters uelec with a minus sign (src/frclong.c:1061, :1142)
                double ug, fg;
                ewald_excl64(rsq, qb, R.C, &ug, &fg);
                G.gq[ii] -= ug;
                ewald_excl64(rsq, qa, R.C, &ug, &fg);
                G.gq[jj] -= ug;
            }
        }
    }
    rx = mic(rx, R.L[0]);
    ry = mic(ry, R.L[1]);
    rz = mic(rz, R.L[2]);
    rsq = rx * rx + ry * ry + rz * rz;
    if (rsq < R.rc2_el) {
        if (ct == MSPMS_ELEC_EWALD) {
            ewald_real64(rsq, qq, R.C, R.kappa, &u, &f);
        } else {
            wolf_real64(rsq, qq, R.C, R.kappa, R.rc_el, R.wvc1, R.wvc2, R.wfc1, R.wfc2, &u, &f);
            acc.vir += f * rsq;
        }
        acc.ureal += u;
        add_pair(fx, fy, fz, ii, jj, f, rx, ry, rz);
        if (GRAD) {
            double ug, fg;
            for (int side = 0; side < 2; side++) {
                const double qo = side ? qa : qb;
                if (ct == MSPMS_ELEC_EWALD)
                    ewald_real64(rsq, qo, R.C, R.kappa, &ug, &fg);
                else
                    wolf_real64(rsq, qo, R.C, R.kappa, R.rc_el, R.wvc1, R.wvc2, R.wfc1, R.wfc2,
                                &ug, &fg);
                G.gq[side ? jj : ii] += ug;
            }
        }
    }
}

__device__ inline bool is_long(const RepDev& R, double rx, double ry, double rz) {
    return fabs(rx) > R.L[0] * 0.5 || fabs(ry) > R.L[1] * 0.5 || fabs(rz) > R.L[2] * 0.5;
}

// an atom's share of the slot-ordered pair forces
__device__ inline void fold_slot(const SlotForces& S, int i, double& fx, double& fy,
                                 double& fz) {
    if (!S.f) return;
    const int g = S.gslot[i];
    if (g < 0) return;
    fx += S.f[g];
    fy += S.f[S.nslot + g];
    fz += S.f[2 * S.nslot + g];
}

// closing half kick of atom i (k_respa_final's arithmetic)
__device__ inline void kick_atom(const Kick& K, int i, const double* fx, const double* fy,
                                 const double* fz) {
#pragma clang fp contract(off)
    const double r_mass = K.rmass[i];  // 1.0 / aw, computed once
    const double deltby2 = K.rep_delt ? K.rep_delt[K.mols[K.atom_mol[i]].rep] / 2.0 : K.deltby2;
    const double a = K.vx[i] + (deltby2 * fx[i] * r_mass);
    const double b = K.vy[i] + (deltby2 * fy[i] * r_mass);
    const double c = K.vz[i] + (deltby2 * fz[i] * r_mass);
    K.vx[i] = a;
    K.vy[i] = b;
    K.vz[i] = c;
    K.kin[i] = (a * a + b * b + c * c) / r_mass;
}

template <bool GRAD>
__global__ void __launch_bounds__(256) k_intra(int nmol, const MolDev* __restrict__ mols,
                                               const RepDev* __restrict__ reps,
                                               const SiteDev* __restrict__ sites,
                                               const SpeciesDev* __restrict__ species,
                                               const TermDev* __restrict__ terms,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z, double* fx,
                                               double* fy, double* fz, double* epart,
                                               GradOut G, int natom_total,
                                               const double* __restrict__ frec, Kick K,
                                               int zero_first) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const MolDev M = mols[m];
    const RepDev& R = reps[M.rep];
    IntraAcc acc = {0.0, 0.0, 0.0, 0.0};
    bool same = true;  // isSameMole carried into the 1-4 section
    const int base = M.atom0;
    if (zero_first)  // this thread is the only writer of its molecule's atoms
        for (int k = 0; k < M.natom; k++) {
            fx[base + k] = 0.0;
            fy[base + k] = 0.0;
            fz[base + k] = 0.0;
        }
    if (frec) {
        // the reciprocal-space forces, computed into their own (3, N) buffer
        // beside the pair kernel, join the long forces here
        const size_t n = (size_t)natom_total;
        for (int k = 0; k < M.natom; k++) {
            fx[base + k] += frec[base + k];
            fy[base + k] += frec[n + base + k];
            fz[base + k] += frec[2 * n + base + k];
        }
    }
    for (int t = M.term0; t < M.term0 + M.nintra; t++) {
        const TermDev T = terms[t];
        const int mm = T.idx[0], nn = T.idx[1];
        const SiteDev a = sites[M.site0 + mm], b = sites[M.site0 + nn];
        const int ii = base + mm, jj = base + nn;
        double rx, ry, rz;
        {
#pragma clang fp contract(off)
            rx = x[ii] - x[jj];
            ry = y[ii] - y[jj];
            rz = z[ii] - z[jj];
        }
        const double qq = a.charge * b.charge;
        const bool lng = is_long(R, rx, ry, rz);
        if (T.kind != TERM_PAIR14) {
            // 1-2 / 1-3 ending pair (src/frclong.c:254-409, :438-593)
            same = !lng;
            if (lng) {
                if (a.ghost == MSPMS_GHOST_FULL || b.ghost == MSPMS_GHOST_FULL) continue;
                double mx = mic(rx, R.L[0]), my = mic(ry, R.L[1]), mz = mic(rz, R.L[2]);
                if (a.ghost != MSPMS_GHOST_LJ && b.ghost != MSPMS_GHOST_LJ)
                    intra_lj<GRAD>(R, a, b, ii, jj, mx, my, mz, 1.0, acc, fx, fy, fz, G);
                // restored through int variables (src/frclong.c:223): truncation
                rx = (double)(int)rx;
                ry = (double)(int)ry;
                rz = (double)(int)rz;
            }
            intra_elec<GRAD>(R, qq, a.charge, b.charge, ii, jj, rx, ry, rz, true, acc, fx, fy, fz,
                             G);
        } else {
            // 1-4 ending pair (src/frclong.c:603-787); `same` is whatever the
            // bond / angle sections left (:622, :632 compare, not assign)
            if (lng && (a.ghost == MSPMS_GHOST_FULL || b.ghost == MSPMS_GHOST_FULL)) continue;
            if (!same) {
                double mx = mic(rx, R.L[0]), my = mic(ry, R.L[1]), mz = mic(rz, R.L[2]);
                if (a.ghost != MSPMS_GHOST_LJ && b.ghost != MSPMS_GHOST_LJ)
                    intra_lj<GRAD>(R, a, b, ii, jj, mx, my, mz, 1.0, acc, fx, fy, fz, G);
                rx = (double)(int)rx;
                ry = (double)(int)ry;
                rz = (double)(int)rz;
            } else {
                intra_lj<GRAD>(R, a, b, ii, jj, rx, ry, rz, R.f0, acc, fx, fy, fz, G);
            }
            intra_elec<GRAD>(R, qq, a.charge, b.charge, ii, jj, rx, ry, rz, false, acc, fx, fy,
                             fz, G);
        }
    }
    double* ep = epart + (size_t)m * 4;
    ep[0] = acc.uvdw;
    ep[1] = acc.ureal;
    ep[2] = acc.uexcl;
    ep[3] = acc.vir;
    // this thread owns the molecule's atoms: their long forces are final
    if (K.vx)
        for (int k = 0; k < M.natom; k++)
            kick_atom(K, base + k, fx, fy, fz);
}

__global__ void __launch_bounds__(256) k_bonded(int nmol, const MolDev* __restrict__ mols,
                                                const SpeciesDev* __restrict__ species,
                                                const TermDev* __restrict__ terms,
                                                const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ z, double* fx,
                                                double* fy, double* fz, double* epart,
                                                Status* st) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const MolDev M = mols[m];
    const SpeciesDev S = species[M.species];
    const int base = M.atom0;
    for (int a = 0; a < M.natom; a++) {
        fx[base + a] = 0.0;
        fy[base + a] = 0.0;
        fz[base + a] = 0.0;
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0}, vir = 0.0;
    for (int t = S.bterm0; t < S.bterm0 + S.nbterm; t++) {
        const TermDev T = terms[t];
        double u = 0.0, v = 0.0;
        int sing = 0;
        int cls = bonded_term(T, base, x, y, z, fx, fy, fz, &u, &v, &sing);
        if (sing) atomicOr(&st->flags, ST_SINGULAR);
        e[cls] += u;
        vir += v;
    }
    double* ep = epart + (size_t)m * 5;
    ep[0] = e[0];
    ep[1] = e[1];
    ep[2] = e[2];
    ep[3] = e[3];
    ep[4] = vir;
}

// ---- deterministic block reductions ------------------------------------
// NV sums over the workgroup (at most 16 waves): a butterfly in each wave,
// then a 16-lane butterfly over the wave partials in wave 0.  The order is
// fixed, so the result is deterministic; the sums end in lanes 0..15 of wave
// 0 (read them in thread 0).  v is overwritten.  sh: NV * (number of waves).

template <typename T, int NV>
__device__ void block_sum_t(T (&v)[NV], T* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    butterfly<T, NV>(v, 32);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k * nw + wid] = v[k];
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = lane < nw ? sh[k * nw + lane] : T(0);
        butterfly<T, NV>(v, 8);
    }
    __syncthreads();
}

template <int NV>
__device__ void block_sum(double (&v)[NV], double* sh) {
    block_sum_t<double, NV>(v, sh);
}

template <int NV>
__device__ void block_sum_u64(unsigned long long (&v)[NV], unsigned long long* sh) {
    block_sum_t<unsigned long long, NV>(v, sh);
}

// The terms of one replica from its reduced partials d (ND slots, see
// k_finalize) and counts q; uvac: the vacuum energies, or null.
template <bool STEP>
__device__ inline void form_terms(const RepDev& R, int r, const double* d,
                                  const unsigned long long* q, const double* uvac,
                                  mspms_long_terms& t, mspms_short_terms& u, double& ukin2) {
#pragma clang fp contract(off)
    constexpr int ND = STEP ? 17 : 11;
    // full lists visit every inter-molecular pair twice
    t.uvdw = 0.5 * d[0] + d[4];
    t.ushift = (0.5 * d[1]) * R.shift4;   // src/frclong.c:985
    t.ureal = 0.5 * d[2] + d[5];
    t.uexcl = d[6];
    double vir = 0.5 * d[3] + d[7];
    t.ufourier = 0.0;
    t.uself = 0.0;
    t.uvacuum = 0.0;
    t.uelec = 0.0;
    if (R.charge_type == MSPMS_ELEC_EWALD) {
        t.ufourier = d[8] * (R.vfactor * R.C);                        // :1048
        t.uself = R.self_sum_q2 * (R.C * sqrt(R.kappa * R.kappa / MSPMS_PI));  // :1058
        t.uelec = t.ureal + t.ufourier - t.uself - t.uexcl;           // :1061
        if (uvac && R.ewald_bc == MSPMS_EWALD_BC_VACUUM) {
            t.uvacuum = uvac[r];
            t.uelec += t.uvacuum;                                     // :1125
        }
        vir += t.uelec;                                               // :1128
    } else if (R.charge_type == MSPMS_ELEC_WOLF) {
        t.uself = R.self_sum_q2 * (R.C * R.wvc3);                     // :1140
        t.uelec = t.ureal - t.uself - t.uexcl;                        // :1142
    } else if (R.charge_type == MSPMS_ELEC_SIMPLE_COULOMB) {
        // every molecule pair twice; the Coulomb virial is its energy
        // (src/elecfrc.c:134-135)
        t.uelec = 0.5 * d[10];
        vir += t.uelec;
    }
    t.usflj = d[9];
    t.uotherff = 0.0;
    t.uinter = t.uvdw + t.uelec + t.usflj + t.uotherff;               // :1223
    t.virial_inter = vir;
    t.n_lj_pairs = (int64_t)q[0];
    t.n_elec_pairs = (int64_t)q[1];
    t.lj_pair_hash = q[2];
    t.elec_pair_hash = q[3];
    u = mspms_short_terms{};
    ukin2 = 0.0;
    if (STEP) {
        u.ubond = d[11];
        u.uangle = d[12];
        u.udih = d[13];
        u.uimp = d[14];
        u.uintra = u.ubond + u.uangle + u.udih + u.uimp;  // src/frcshort.c:87
        u.virial_intra = d[15];
        u.usflj = 0.0;  // a step reports the wall in its long terms
        ukin2 = d[ND - 1] * 0.5;
    }
}

// Every partial of frclong into its terms, nb workgroups per replica (one
// partial of each kind per thread where the replica fits): each reduces a
// strided share of the replica's partials, every load of a pass issued
// before the first add, and stores its sums; the replica's last workgroup
// to finish (counter, reset by that workgroup) adds the nb sums in block
// order, so the result does not depend on arrival order.  STEP (one RESPA outer step) also
// reduces the bonded partials of the inner steps into the frcshort terms and
// the per-atom 2 x kinetic energies into ukin, in the same pass.  Slots of
// d[]: 0-3 pair (uvdw, shift, ureal, virial), 4-7 intramolecular (uvdw,
// ureal, uexcl, virial), 8 reciprocal, 9 solid-fluid, 10 simple Coulomb,
// 11-15 bonded (bond, angle, dihedral, improper, virial), 16 kinetic.
template <bool STEP>
__global__ void __launch_bounds__(FIN_T) k_finalize(
    const RepDev* __restrict__ reps, const int* __restrict__ col_cl0,
    const double* __restrict__ ep_pair, const unsigned long long* __restrict__ ip_pair,
    int have_pair, const double* __restrict__ ep_intra, int have_intra,
    const double* __restrict__ uk, int have_recip, const double* __restrict__ ep_sf, int have_sf,
    const double* __restrict__ uvac, const double* __restrict__ ep_coul,
    mspms_long_terms* __restrict__ out, int* lstate, const double* __restrict__ ep_short,
    mspms_short_terms* __restrict__ sout, const double* kin,  // written by the pass below
    double* __restrict__ ukin, FinBufs fb, int nb, FinAtoms A, const int* __restrict__ atom_site,
    const SiteDev* __restrict__ sites) {
#pragma clang fp contract(off)
    constexpr int ND = STEP ? 17 : 11;
    static_assert(ND <= FIN_ND, "finalize partial slots");
    __shared__ double rows[ND * FIN_T];
    __shared__ unsigned long long rowsu[4 * FIN_T];
    __shared__ double bsum[ND];
    __shared__ unsigned long long bsumu[4];
    __shared__ double tot[ND];
    __shared__ unsigned long long totu[4];
    __shared__ int last;
    if (A.fx) {
        // the step's closing pass over this replica's atoms, strided as the
        // kinetic partials below, so each thread reads back its own kicks
        const RepDev& Ra = reps[blockIdx.x / nb];
        const size_t n = (size_t)A.natom;
        for (int i = (blockIdx.x % nb) * FIN_T + threadIdx.x; i < Ra.natom; i += nb * FIN_T) {
            const int a = Ra.atom0 + i;
            double ax = A.fx[a], ay = A.fy[a], az = A.fz[a];
            fold_slot(A.S, a, ax, ay, az);
            A.fx[a] = ax + A.frec[a];
            A.fy[a] = ay + A.frec[n + a];
            A.fz[a] = az + A.frec[2 * n + a];
            if (A.K.vx) kick_atom(A.K, a, A.fx, A.fy, A.fz);
        }
    }
    if (fb.on && !*fb.on) {
        // the cluster kernels of this call have read the rebuild request
        if (blockIdx.x == 0 && threadIdx.x == 0 && lstate) lstate[LS_NEED] = 0;
        return;
    }
    const int r = blockIdx.x / nb, b = blockIdx.x % nb;
    const RepDev& R = reps[r];
    const int t0 = b * FIN_T + threadIdx.x, nt = nb * FIN_T;
    double d[ND];
#pragma unroll
    for (int k = 0; k < ND; k++) d[k] = 0.0;
    unsigned long long q[4] = {0, 0, 0, 0};
    const int c0 = have_pair ? col_cl0[R.col0] : 0;
    const int ncl = have_pair ? col_cl0[R.col0 + R.ncol] - c0 : 0;
    const int nmi = have_intra ? R.nmol : 0, nkr = have_recip ? R.nk : 0;
    const int nsf = have_sf ? R.natom : 0;
    const int nco = ep_coul && R.charge_type == MSPMS_ELEC_SIMPLE_COULOMB ? R.nmol : 0;
    const int nsh = STEP ? R.nmol : 0, nki = STEP && kin ? R.natom : 0;
    const int nmax = max(max(max(ncl, nmi), max(nkr, nsf)), max(nco, max(nsh, nki)));
    const double4 z4 = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int i = t0; i < nmax; i += nt) {
        // the loads of every kind first (one memory round trip per pass)
        const double4 e = i < ncl ? reinterpret_cast<const double4*>(ep_pair)[c0 + i] : z4;
        const ulonglong4 n = i < ncl ? reinterpret_cast<const ulonglong4*>(ip_pair)[c0 + i]
                                     : make_ulonglong4(0, 0, 0, 0);
        const double4 w = i < nmi ? reinterpret_cast<const double4*>(ep_intra)[R.mol0 + i] : z4;
        const double vk = i < nkr ? uk[R.k0 + i] : 0.0;
        const double vs = i < nsf ? ep_sf[R.atom0 + i] : 0.0;
        const double vc = i < nco ? ep_coul[R.mol0 + i] : 0.0;
        double vb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (i < nsh)
            for (int k = 0; k < 5; k++) vb[k] = ep_short[(size_t)(R.mol0 + i) * 5 + k];
        const double vt = i < nki ? kin[R.atom0 + i] : 0.0;
        d[0] += e.x;
        d[1] += e.y;
        d[2] += e.z;
        d[3] += e.w;
        q[0] += n.x;
        q[1] += n.y;
        q[2] += n.z;
        q[3] += n.w;
        d[4] += w.x;
        d[5] += w.y;
        d[6] += w.z;
        d[7] += w.w;
        d[8] += vk;
        d[9] += vs;
        d[10] += vc;
        if (STEP) {
            for (int k = 0; k < 5; k++) d[11 + k] += vb[k];
            d[ND - 1] += vt;
        }
    }
    block_sum_rows<double, ND, FIN_T>(d, rows, bsum);
    block_sum_rows<unsigned long long, 4, FIN_T>(q, rowsu, bsumu);
    if (STEP && fb.hist && fb.hist->on) {
        // a run recording every step's terms: the workgroup sums go to this
        // step's slot of the history of sums; k_hist_terms forms the terms
        // after the run (no fence or arrival count per step)
        if (threadIdx.x == 0) {
            const HistDev* h = fb.hist;
            const int t = *h->tick - 1;
            if (t >= 0 && t < h->cap) {
                double* P = h->part + (((size_t)t * h->nrep + r) * nb + b) * FIN_ND;
                unsigned long long* Q = h->partu + (((size_t)t * h->nrep + r) * nb + b) * 4;
                for (int k = 0; k < ND; k++) P[k] = bsum[k];
                for (int k = 0; k < 4; k++) Q[k] = bsumu[k];
                if (b == 0 && h->uvac && uvac) h->uvac[(size_t)t * h->nrep + r] = uvac[r];
            }
            // the cluster kernels of this call have read the rebuild request
            if (blockIdx.x == 0 && lstate) lstate[LS_NEED] = 0;
        }
        return;
    }
    if (threadIdx.x == 0) {
        double* P = fb.part + ((size_t)r * nb + b) * FIN_ND;
        unsigned long long* Q = fb.partu + ((size_t)r * nb + b) * 4;
        for (int k = 0; k < ND; k++) P[k] = bsum[k];
        for (int k = 0; k < 4; k++) Q[k] = bsumu[k];
        __threadfence();
        last = atomicAdd(&fb.count[r], 1) == nb - 1;
    }
    __syncthreads();
    if (!last) return;
    // the replica's last workgroup: the nb sums, read past the
    // (non-coherent) vector L1; wave w takes values w, w + 4, ...: lane bb
    // reads workgroup bb's sum (nb <= 64) and the wave adds them in a fixed
    // order, all loads of the workgroup in flight at once
    __threadfence();
    static_assert(FIN_BLOCKS <= 64, "one lane per workgroup sum");
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int k = w; k < ND + 4; k += FIN_T / 64) {
            if (k < ND) {
                const double v =
                    lane < nb ? __hip_atomic_load(fb.part + ((size_t)r * nb + lane) * FIN_ND + k,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : 0.0;
                const double s = wave_sum64(v);
                if (lane == 0) tot[k] = s;
            } else {
                const unsigned long long v =
                    lane < nb
                        ? __hip_atomic_load(fb.partu + ((size_t)r * nb + lane) * 4 + (k - ND),
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : 0ull;
                const unsigned long long s = wave_sum64(v);
                if (lane == 0) totu[k - ND] = s;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    fb.count[r] = 0;  // ready for the next call
    // the cluster kernels of this call have read the rebuild request
    if (lstate && r == 0) lstate[LS_NEED] = 0;
#pragma unroll
    for (int k = 0; k < ND; k++) d[k] = tot[k];
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = totu[k];

    mspms_long_terms t;
    mspms_short_terms u;
    double uk2;
    form_terms<STEP>(R, r, d, q, uvac, t, u, uk2);
    out[r] = t;
    if (STEP) {
        sout[r] = u;
        if (ukin) ukin[r] = uk2;
    }
}

// The terms of every step of a run that recorded histories (after it):
// one wave per (step, replica) adds the step's nb workgroup sums as the
// step's last workgroup would have (lane bb takes workgroup bb, wave sum),
// forms the terms and writes the histories; the last step's terms also go
// to the run's outputs.
__global__ void __launch_bounds__(64) k_hist_terms(const RepDev* __restrict__ reps, HistDev h,
                                                   int nb, int nsteps,
                                                   mspms_long_terms* __restrict__ lterms,
                                                   mspms_short_terms* __restrict__ sterms,
                                                   double* __restrict__ ukin) {
    constexpr int ND = 17;
    const int t = blockIdx.x / h.nrep, r = blockIdx.x % h.nrep, lane = threadIdx.x;
    double d[ND];
    unsigned long long q[4];
    const double* P = h.part + ((size_t)t * h.nrep + r) * nb * FIN_ND;
    const unsigned long long* Q = h.partu + ((size_t)t * h.nrep + r) * nb * 4;
#pragma unroll
    for (int k = 0; k < ND; k++) d[k] = wave_sum64(lane < nb ? P[lane * FIN_ND + k] : 0.0);
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = wave_sum64(lane < nb ? Q[lane * 4 + k] : 0ull);
    if (lane != 0) return;
    mspms_long_terms lt;
    mspms_short_terms st;
    double uk;
    // the step's vacuum energy is in slot r of its row of h.uvac
    form_terms<true>(reps[r], r, d, q, h.uvac ? h.uvac + (size_t)t * h.nrep : nullptr, lt, st,
                     uk);
    const size_t slot = (size_t)t * h.nrep + r;
    if (h.lh) h.lh[slot] = lt;
    if (h.sh) h.sh[slot] = st;
    if (h.kh) h.kh[slot] = uk;
    if (t == nsteps - 1) {
        lterms[r] = lt;
        sterms[r] = st;
        if (ukin) ukin[r] = uk;
    }
}

// the history descriptor of a device run, written in stream order
__global__ void k_set_hist(HistDev* d, HistDev v, int* hstep, int nrep) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d = v;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrep; r += gridDim.x * blockDim.x)
        hstep[r] = 0;
}

__global__ void __launch_bounds__(1024) k_finalize_short(const RepDev* __restrict__ reps,
                                                        const double* __restrict__ ep,
                                                        const double* __restrict__ ep_sf,
                                                        mspms_short_terms* __restrict__ out) {
    __shared__ double sh[6 * 16];
    const RepDev& R = reps[blockIdx.x];
    double p[6] = {0, 0, 0, 0, 0, 0};
    for (int m = R.mol0 + threadIdx.x; m < R.mol0 + R.nmol; m += blockDim.x)
        for (int k = 0; k < 5; k++) p[k] += ep[(size_t)m * 5 + k];
    if (ep_sf)
        for (int a = R.atom0 + threadIdx.x; a < R.atom0 + R.natom; a += blockDim.x) p[5] += ep_sf[a];
    block_sum<6>(p, sh);
    if (threadIdx.x != 0) return;
    mspms_short_terms t;
    t.ubond = p[0];
    t.uangle = p[1];
    t.udih = p[2];
    t.uimp = p[3];
    t.uintra = t.ubond + t.uangle + t.udih + t.uimp;  // src/frcshort.c:87
    t.virial_intra = p[4];
    t.usflj = p[5];
    out[blockIdx.x] = t;
}

// Parameter adjoint: one workgroup per (replica, site) work item
// (rep, global site, first atom of that site, molecule count).  Sums the
// per-atom derivatives over the replica's atoms of the site and adds the
// self-energy term: uelec carries -uself = -C coef sum q^2
// (src/frclong.c:1058, :1140), so d/dq_i = -2 C coef q_i.
__global__ void __launch_bounds__(256) k_grad_sites(
    const int4* __restrict__ work, const RepDev* __restrict__ reps,
    const SiteDev* __restrict__ sites, const SpeciesDev* __restrict__ species,
    const double* __restrict__ gq, const double* __restrict__ ge, const double* __restrict__ gs,
    int nsite, double* __restrict__ oq, double* __restrict__ oe, double* __restrict__ os) {
    __shared__ double sh[3 * 8];
    const int4 w = work[blockIdx.x];
    const RepDev& R = reps[w.x];
    const SiteDev S = sites[w.y];
    const int stride = species[S.species].natom;
    double p[3] = {0.0, 0.0, 0.0};
    for (int m = threadIdx.x; m < w.w; m += blockDim.x) {
        const int a = w.z + m * stride;
        p[0] += gq[a];
        p[1] += ge[a];
        p[2] += gs[a];
    }
    block_sum<3>(p, sh);
    if (threadIdx.x != 0) return;
    double coef = 0.0;
    if (R.charge_type == MSPMS_ELEC_EWALD)
        coef = R.C * sqrt(R.kappa * R.kappa / MSPMS_PI);
    else if (R.charge_type == MSPMS_ELEC_WOLF)
        coef = R.C * R.wvc3;
    const size_t o = (size_t)w.x * nsite + w.y;
    oq[o] = p[0] - 2.0 * coef * S.charge * (double)w.w;
    oe[o] = p[1];
    os[o] = p[2];
}

// Site k of molecule M rebuilt around the molecule's centre of mass wrapped
// into the primal box: cal_com_and_efg_one + reconstruct_from_com_one
// (src/misc.c:71-122) with the wrap of src/frclong.c:1084-1087.
__device__ inline void com_wrapped(const RepDev& R, const MolDev& M, const SpeciesDev& S,
                                   const SiteDev* sites, const double* x, const double* y,
                                   const double* z, double com[3], double comw[3]) {
#pragma clang fp contract(off)
    com[0] = com[1] = com[2] = 0.0;
    for (int k = 0; k < M.natom; k++) {
        const double aw = sites[S.site0 + k].aw;
        com[0] += x[M.atom0 + k] * aw;
        com[1] += y[M.atom0 + k] * aw;
        com[2] += z[M.atom0 + k] * aw;
    }
    for (int d = 0; d < 3; d++) {
        com[d] /= S.mw;
        comw[d] = com[d] - R.L[d] * rint(com[d] / R.L[d]);
    }
}

// Molecular centres of mass (unwrapped), cal_com_and_efg_one order.
__global__ void __launch_bounds__(256) k_mol_com(int nmol, const MolDev* __restrict__ mols,
                                                 const SpeciesDev* __restrict__ species,
                                                 const SiteDev* __restrict__ sites,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 double4* __restrict__ com) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const MolDev M = mols[m];
    const SpeciesDev S = species[M.species];
    double c[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < M.natom; k++) {
        const double aw = sites[S.site0 + k].aw;
        c[0] += x[M.atom0 + k] * aw;
        c[1] += y[M.atom0 + k] * aw;
        c[2] += z[M.atom0 + k] * aw;
    }
    com[m] = make_double4(c[0] / S.mw, c[1] / S.mw, c[2] / S.mw, 0.0);
}

// Simple Coulomb between molecules (xmole_coulomb_frc, src/elecfrc.c:67-157):
// molecule pairs whose minimum-image COM distance is inside rcutoffelec
// interact through all their non-GHOST_FULL site pairs, the site separation
// built on the image of the COM separation.  One wave per i-molecule, the
// lanes over the replica's other molecules (both orders of every pair are
// visited: forces on i only, energy halved in the finalize).  FP64.
template <int NA>
__global__ void __launch_bounds__(64) k_coulomb_mol(
    const RepDev* __restrict__ reps, const MolDev* __restrict__ mols,
    const SpeciesDev* __restrict__ species, const SiteDev* __restrict__ sites,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double4* __restrict__ com, double* fx, double* fy, double* fz,
    double* __restrict__ ep) {
#pragma clang fp contract(off)
    const int mi = blockIdx.x, lane = threadIdx.x;
    const MolDev Mi = mols[mi];
    const RepDev& R = reps[Mi.rep];
    const SpeciesDev Si = species[Mi.species];
    const double4 ci = com[mi];
    double ax[NA], ay[NA], az[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) ax[k] = ay[k] = az[k] = 0.0;
    double eu = 0.0;
    for (int mj = R.mol0 + lane; mj < R.mol0 + R.nmol; mj += 64) {
        if (mj == mi) continue;
        const MolDev Mj = mols[mj];
        const SpeciesDev Sj = species[Mj.species];
        const double4 cj = com[mj];
        double rxcm = ci.x - cj.x, rycm = ci.y - cj.y, rzcm = ci.z - cj.z;
        rxcm = rxcm - R.L[0] * rint(rxcm / R.L[0]);
        rycm = rycm - R.L[1] * rint(rycm / R.L[1]);
        rzcm = rzcm - R.L[2] * rint(rzcm / R.L[2]);
        const double rcmsq = rxcm * rxcm + rycm * rycm + rzcm * rzcm;
        if (!(rcmsq < R.rc2_el)) continue;
#pragma unroll
        for (int mm = 0; mm < NA; mm++) {
            if (mm >= Mi.natom) break;
            const SiteDev& A = sites[Si.site0 + mm];
            if (A.ghost == MSPMS_GHOST_FULL) continue;
            const int ii = Mi.atom0 + mm;
            const double exi = x[ii] - ci.x, eyi = y[ii] - ci.y, ezi = z[ii] - ci.z;
            for (int nn = 0; nn < Mj.natom; nn++) {
                const SiteDev& B = sites[Sj.site0 + nn];
                if (B.ghost == MSPMS_GHOST_FULL) continue;
                const int jj = Mj.atom0 + nn;
                const double rxij = exi + rxcm - (x[jj] - cj.x);
                const double ryij = eyi + rycm - (y[jj] - cj.y);
                const double rzij = ezi + rzcm - (z[jj] - cj.z);
                const double rijsq = rxij * rxij + ryij * ryij + rzij * rzij;
                const double rij = sqrt(rijsq);
                const double uij = R.C * A.charge * B.charge / rij;
                eu += uij;
                const double fij = uij / rijsq;
                ax[mm] += fij * rxij;
                ay[mm] += fij * ryij;
                az[mm] += fij * rzij;
            }
        }
    }
#pragma unroll
    for (int mm = 0; mm < NA; mm++) {
        if (mm >= Mi.natom) break;
        double gx = ax[mm], gy = ay[mm], gz = az[mm];
        for (int o = 32; o >= 1; o >>= 1) {
            gx += __shfl_xor(gx, o, 64);
            gy += __shfl_xor(gy, o, 64);
            gz += __shfl_xor(gz, o, 64);
        }
        if (lane == 0) {
            const int ii = Mi.atom0 + mm;
            fx[ii] += gx;
            fy[ii] += gy;
            fz[ii] += gz;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) eu += __shfl_xor(eu, o, 64);
    if (lane == 0) ep[mi] = eu;
}

// Ewald vacuum boundary term (src/frclong.c:1065-1126), one workgroup per
// replica: M = sum_i q_i r_i over the rebuilt sites, uvacuum = |M|^2 *
// (2 pi / 3V) C, F_i += -2 (2 pi / 3V) q_i C M.  The adjoint adds
// d uvacuum / d q_i = 2 (2 pi / 3V) C (M . r_i) to gq.
__global__ void __launch_bounds__(1024) k_vacuum(
    const RepDev* __restrict__ reps, const MolDev* __restrict__ mols,
    const SpeciesDev* __restrict__ species, const SiteDev* __restrict__ sites,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* fx, double* fy, double* fz, double* __restrict__ uvac, double* gq) {
#pragma clang fp contract(off)
    __shared__ double sh[3 * 16];
    __shared__ double dip[3];
    const RepDev& R = reps[blockIdx.x];
    if (R.ewald_bc != MSPMS_EWALD_BC_VACUUM) return;  // whole workgroup
    double p[3] = {0.0, 0.0, 0.0};
    for (int m = R.mol0 + threadIdx.x; m < R.mol0 + R.nmol; m += blockDim.x) {
        const MolDev M = mols[m];
        const SpeciesDev S = species[M.species];
        double com[3], comw[3];
        com_wrapped(R, M, S, sites, x, y, z, com, comw);
        for (int k = 0; k < M.natom; k++) {
            const int i = M.atom0 + k;
            const double q = sites[S.site0 + k].charge;
            p[0] = p[0] + q * ((x[i] - com[0]) + comw[0]);
            p[1] = p[1] + q * ((y[i] - com[1]) + comw[1]);
            p[2] = p[2] + q * ((z[i] - com[2]) + comw[2]);
        }
    }
    block_sum<3>(p, sh);
    if (threadIdx.x == 0) {
        dip[0] = p[0];
        dip[1] = p[1];
        dip[2] = p[2];
        double u = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        uvac[blockIdx.x] = u * (R.twopi_over_3v * R.C);
    }
    __syncthreads();
    const double qrx = dip[0], qry = dip[1], qrz = dip[2];
    for (int m = R.mol0 + threadIdx.x; m < R.mol0 + R.nmol; m += blockDim.x) {
        const MolDev M = mols[m];
        const SpeciesDev S = species[M.species];
        double com[3], comw[3];
        if (gq) com_wrapped(R, M, S, sites, x, y, z, com, comw);
        for (int k = 0; k < M.natom; k++) {
            const int i = M.atom0 + k;
            const double q = sites[S.site0 + k].charge;
            const double fij = -2.0 * R.twopi_over_3v * q * R.C;
            fx[i] += fij * qrx;
            fy[i] += fij * qry;
            fz[i] += fij * qrz;
            if (gq) {
                const double rx = (x[i] - com[0]) + comw[0], ry = (y[i] - com[1]) + comw[1],
                             rz = (z[i] - com[2]) + comw[2];
                gq[i] += 2.0 * R.twopi_over_3v * R.C * (qrx * rx + qry * ry + qrz * rz);
            }
        }
    }
}

}  // namespace

void launch_grad_sites(const LaunchCtx& c, const SpeciesDev* species, const int4* work,
                       int nwork, const double* gq, const double* ge, const double* gs,
                       double* out_q, double* out_e, double* out_s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(k_grad_sites, dim3(nwork), dim3(256), 0, c.stream, work, c.reps, c.sites,
                       species, gq, ge, gs, c.nsite, out_q, out_e, out_s);
}

void launch_intra(const LaunchCtx& c, const SpeciesDev* species, const TermDev* terms,
                  const double* x, const double* y, const double* z, double* fx, double* fy,
                  double* fz, double* epart, double* gq, double* ge, double* gs,
                  const double* frec, const Kick* kick, bool zero_first) {
    int g = (c.nmol + 255) / 256;
    const Kick K = kick ? *kick : Kick{};
    if (gq)
        hipLaunchKernelGGL(k_intra<true>, dim3(g), dim3(256), 0, c.stream, c.nmol, c.mols, c.reps,
                           c.sites, species, terms, x, y, z, fx, fy, fz, epart,
                           GradOut{gq, ge, gs}, c.natom, frec, K, zero_first ? 1 : 0);
    else
        hipLaunchKernelGGL(k_intra<false>, dim3(g), dim3(256), 0, c.stream, c.nmol, c.mols,
                           c.reps, c.sites, species, terms, x, y, z, fx, fy, fz, epart,
                           GradOut{nullptr, nullptr, nullptr}, c.natom, frec, K,
                           zero_first ? 1 : 0);
}

// fx += frec (3, N): the reciprocal forces where no k_intra runs
__global__ void __launch_bounds__(256) k_add_frec(int n, double* __restrict__ fx,
                                                  double* __restrict__ fy,
                                                  double* __restrict__ fz,
                                                  const double* __restrict__ frec,
                                                  const int* __restrict__ atom_site,
                                                  const SiteDev* __restrict__ sites, Kick K,
                                                  SlotForces S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = fx[i], ay = fy[i], az = fz[i];
    fold_slot(S, i, ax, ay, az);
    fx[i] = ax + frec[i];
    fy[i] = ay + frec[(size_t)n + i];
    fz[i] = az + frec[2 * (size_t)n + i];
    if (K.vx) kick_atom(K, i, fx, fy, fz);
}

__global__ void __launch_bounds__(256) k_fold_slots(int n, double* __restrict__ fx,
                                                    double* __restrict__ fy,
                                                    double* __restrict__ fz, SlotForces S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = fx[i], ay = fy[i], az = fz[i];
    fold_slot(S, i, ax, ay, az);
    fx[i] = ax;
    fy[i] = ay;
    fz[i] = az;
}

__global__ void __launch_bounds__(256) k_zero(size_t n, double* __restrict__ p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

void launch_add_frec(const LaunchCtx& c, double* fx, double* fy, double* fz,
                     const double* frec, const Kick* kick, const SlotForces* slots) {
    const Kick K = kick ? *kick : Kick{};
    const SlotForces S = slots ? *slots : SlotForces{};
    hipLaunchKernelGGL(k_add_frec, dim3((c.natom + 255) / 256), dim3(256), 0, c.stream, c.natom,
                       fx, fy, fz, frec, c.atom_site, c.sites, K, S);
}

void launch_fold_slots(const LaunchCtx& c, const SlotForces& S, double* fx, double* fy,
                       double* fz) {
    hipLaunchKernelGGL(k_fold_slots, dim3((c.natom + 255) / 256), dim3(256), 0, c.stream,
                       c.natom, fx, fy, fz, S);
}

// zero fill by a kernel (capturable; no memset nodes in the step graph)
void launch_zero(const LaunchCtx& c, double* p, size_t n) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, n, p);
}

void launch_bonded(const LaunchCtx& c, const SpeciesDev* species, const TermDev* terms,
                   const double* x, const double* y, const double* z, double* fx, double* fy,
                   double* fz, double* epart) {
    int g = (c.nmol + 255) / 256;
    hipLaunchKernelGGL(k_bonded, dim3(g), dim3(256), 0, c.stream, c.nmol, c.mols, species, terms,
                       x, y, z, fx, fy, fz, epart, c.status);
}

void launch_finalize_long(const LaunchCtx& c, const int* col_cl0, const double* ep_pair,
                          const unsigned long long* ip_pair, int have_pair,
                          const double* ep_intra, int have_intra, const double* uk,
                          int have_recip, const double* ep_sf, int have_sf,
                          const double* uvac, const double* ep_coul, void* terms,
                          int* lstate, const double* ep_short, void* sterms,
                          const double* kin, double* ukin, const FinBufs& fb,
                          const FinAtoms* atoms) {
    const dim3 grid(c.nrep * fb.nb);
    FinAtoms A{};
    if (atoms) A = *atoms;
    if (ep_short)
        hipLaunchKernelGGL(k_finalize<true>, grid, dim3(FIN_T), 0, c.stream, c.reps, col_cl0,
                           ep_pair, ip_pair, have_pair, ep_intra, have_intra, uk, have_recip,
                           ep_sf, have_sf, uvac, ep_coul, (mspms_long_terms*)terms, lstate,
                           ep_short, (mspms_short_terms*)sterms, kin, ukin, fb, fb.nb, A,
                           c.atom_site, c.sites);
    else
        hipLaunchKernelGGL(k_finalize<false>, grid, dim3(FIN_T), 0, c.stream, c.reps, col_cl0,
                           ep_pair, ip_pair, have_pair, ep_intra, have_intra, uk, have_recip,
                           ep_sf, have_sf, uvac, ep_coul, (mspms_long_terms*)terms, lstate,
                           nullptr, nullptr, nullptr, nullptr, fb, fb.nb, A, c.atom_site,
                           c.sites);
}

int coulomb_mol_supported(int max_mol_atoms) { return max_mol_atoms <= 16; }

void launch_coulomb_mol(const LaunchCtx& c, const SpeciesDev* species, const double* x,
                        const double* y, const double* z, double4* com, double* fx, double* fy,
                        double* fz, double* ep) {
    hipLaunchKernelGGL(k_mol_com, dim3((c.nmol + 255) / 256), dim3(256), 0, c.stream, c.nmol,
                       c.mols, species, c.sites, x, y, z, com);
    if (c.max_mol_atoms <= 4)
        hipLaunchKernelGGL(k_coulomb_mol<4>, dim3(c.nmol), dim3(64), 0, c.stream, c.reps, c.mols,
                           species, c.sites, x, y, z, com, fx, fy, fz, ep);
    else
        hipLaunchKernelGGL(k_coulomb_mol<16>, dim3(c.nmol), dim3(64), 0, c.stream, c.reps,
                           c.mols, species, c.sites, x, y, z, com, fx, fy, fz, ep);
}

void launch_vacuum(const LaunchCtx& c, const SpeciesDev* species, const double* x,
                   const double* y, const double* z, double* fx, double* fy, double* fz,
                   double* uvac, double* gq) {
    hipLaunchKernelGGL(k_vacuum, dim3(c.nrep), dim3(1024), 0, c.stream, c.reps, c.mols, species,
                       c.sites, x, y, z, fx, fy, fz, uvac, gq);
}

void launch_hist_terms(const LaunchCtx& c, const HistDev& h, int nsteps, void* lterms,
                       void* sterms, double* ukin) {
    hipLaunchKernelGGL(k_hist_terms, dim3(nsteps * h.nrep), dim3(64), 0, c.stream, c.reps, h, h.nb,
                       nsteps, (mspms_long_terms*)lterms, (mspms_short_terms*)sterms, ukin);
}

void launch_set_hist(const LaunchCtx& c, HistDev* d, const HistDev& v, int* hstep) {
    hipLaunchKernelGGL(k_set_hist, dim3((c.nrep + 255) / 256), dim3(256), 0, c.stream, d, v,
                       hstep, c.nrep);
}

void launch_finalize_short(const LaunchCtx& c, const double* epart, void* terms,
                           const double* ep_sf) {
    hipLaunchKernelGGL(k_finalize_short, dim3(c.nrep), dim3(1024), 0, c.stream, c.reps, epart,
                       ep_sf, (mspms_short_terms*)terms);
}
