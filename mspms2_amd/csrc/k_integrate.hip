// k_integrate.hip -- device-resident RESPA velocity Verlet (NVE) on gfx950.
// Reference: vver, src/md.c:22-83 (called from src/mspms2.c:99-132 and
// md_move, src/dispmd.c:42-46).
//
// One outer step:
//   k_respa_inner   per molecule (one thread): half kick with the long
//                   forces, then nstep_inner times {half kick + drift with
//                   the short forces, bonded forces of this molecule
//                   (frcshort's terms), half kick}.  The inner loop touches
//                   only the molecule's own atoms, so all nstep_inner
//                   substeps run in one launch with no grid-wide sync.
//   frclong         (the engine's long-range kernels)
//   k_respa_final   per atom: half kick with the new long forces and the
//                   kinetic energy partial (src/md.c:67-80)
// Per-atom arithmetic follows md.c expression for expression.
#include "engine.h"
#include "sf_eval.h"
#include "terms.h"
#include "wave.h"

namespace {

// The solid-fluid wall in the inner loop (the RESPA split): each inner step
// adds the wall's force on a GHOST_NONE site to its short-range force; the
// last inner step's energies go to usf[atom] (0 for other sites).
struct SfInner {
    SFDev sf;
    const RepDev* reps;
    double* usf;
    int on;
};

__device__ inline double sf_inner_site(const SfInner& W, const SiteDev& S, int rep, double x,
                                       double y, double z, double& fx, double& fy, double& fz,
                                       Status* st) {
    if (S.ghost != MSPMS_GHOST_NONE) return 0.0;
    int edom = 0;
    const double u = sfeval::sf_site(W.sf, W.reps[rep], S, x, y, z, fx, fy, fz, &edom);
    if (edom) atomicOr(&st->flags, ST_GRID_EDOM);
    return u;
}

// The start-of-step snapshot: while no pair-list overflow has been
// flagged, each step saves its starting positions and velocities, so that
// after an overflow the state at the end of the last good step can be put
// back.
struct SnapOut {
    double* p;   // (6, natom) or null
    int* count;
    int natom;
};

__device__ inline bool snap_wanted(const SnapOut& SN, const Status* st) {
    return SN.p && !(st->flags & ST_OVERFLOW);
}

__device__ inline void snap_atom(const SnapOut& SN, int i, double x, double y, double z,
                                 double vx, double vy, double vz) {
    const size_t n = (size_t)SN.natom;
    SN.p[i] = x;
    SN.p[n + i] = y;
    SN.p[2 * n + i] = z;
    SN.p[3 * n + i] = vx;
    SN.p[4 * n + i] = vy;
    SN.p[5 * n + i] = vz;
}

__device__ inline void snap_count(const SnapOut& SN) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(SN.count, 1);
}

// a replica's own outer step (HMC with delt adaptation): the host's
// derivation, src/initvars.c:605-607
__device__ inline void rep_steps(const double* rep_delt, int rep, int nstep_inner,
                                 double& deltby2, double& delts, double& deltsby2) {
#pragma clang fp contract(off)
    if (!rep_delt) return;
    const double d = rep_delt[rep];
    deltby2 = d / 2.0;
    delts = d / nstep_inner;
    deltsby2 = delts / 2.0;
}

__global__ void __launch_bounds__(256) k_respa_inner(
    int nmol, const MolDev* __restrict__ mols, const SpeciesDev* __restrict__ species,
    const TermDev* __restrict__ terms, const SiteDev* __restrict__ sites, double* x, double* y,
    double* z, double* vx, double* vy, double* vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl, double* fxs, double* fys,
    double* fzs, double deltby2, double delts, double deltsby2, int nstep_inner, double* epart,
    Status* st, const double* __restrict__ rep_delt, SfInner W, SnapOut SN) {
#pragma clang fp contract(off)
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const MolDev M = mols[m];
    const SpeciesDev S = species[M.species];
    rep_steps(rep_delt, M.rep, nstep_inner, deltby2, delts, deltsby2);
    const int a0 = M.atom0, a1 = M.atom0 + M.natom;
    if (snap_wanted(SN, st)) {
        for (int i = a0; i < a1; i++) snap_atom(SN, i, x[i], y[i], z[i], vx[i], vy[i], vz[i]);
        if (m == 0) snap_count(SN);
    }
    for (int i = a0; i < a1; i++) {
        double r_mass = 1.0 / sites[S.site0 + (i - a0)].aw;
        vx[i] += (deltby2 * fxl[i] * r_mass);
        vy[i] += (deltby2 * fyl[i] * r_mass);
        vz[i] += (deltby2 * fzl[i] * r_mass);
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0}, vir = 0.0;
    for (int ll = 0; ll < nstep_inner; ll++) {
        for (int i = a0; i < a1; i++) {
            double r_mass = 1.0 / sites[S.site0 + (i - a0)].aw;
            vx[i] += (deltsby2 * fxs[i] * r_mass);
            vy[i] += (deltsby2 * fys[i] * r_mass);
            vz[i] += (deltsby2 * fzs[i] * r_mass);
            x[i] = x[i] + delts * vx[i];
            y[i] = y[i] + delts * vy[i];
            z[i] = z[i] + delts * vz[i];
        }
        // frcshort of this molecule (src/frcshort.c:33-87)
        for (int i = a0; i < a1; i++) fxs[i] = fys[i] = fzs[i] = 0.0;
        e[0] = e[1] = e[2] = e[3] = 0.0;
        vir = 0.0;
        for (int t = S.bterm0; t < S.bterm0 + S.nbterm; t++) {
            double u = 0.0, v = 0.0;
            int sing = 0;
            int cls = bonded_term(terms[t], a0, x, y, z, fxs, fys, fzs, &u, &v, &sing);
            if (sing) atomicOr(&st->flags, ST_SINGULAR);
            e[cls] += u;
            vir += v;
        }
        if (W.on)
            for (int i = a0; i < a1; i++) {
                double gx = fxs[i], gy = fys[i], gz = fzs[i];
                const double u = sf_inner_site(W, sites[S.site0 + (i - a0)], M.rep, x[i], y[i],
                                               z[i], gx, gy, gz, st);
                fxs[i] = gx;
                fys[i] = gy;
                fzs[i] = gz;
                if (ll == nstep_inner - 1) W.usf[i] = u;
            }
        for (int i = a0; i < a1; i++) {
            double r_mass = 1.0 / sites[S.site0 + (i - a0)].aw;
            vx[i] += (deltsby2 * fxs[i] * r_mass);
            vy[i] += (deltsby2 * fys[i] * r_mass);
            vz[i] += (deltsby2 * fzs[i] * r_mass);
        }
    }
    double* ep = epart + (size_t)m * 5;
    ep[0] = e[0];
    ep[1] = e[1];
    ep[2] = e[2];
    ep[3] = e[3];
    ep[4] = vir;
}

// The same step with the molecule's atoms held in LDS for all substeps:
// one wave per workgroup, lane l owns molecule blockIdx.x * 64 + l and the
// slots [atom k][lane l] (stride 64, conflict-free across the wave).  No
// barriers: every lane touches only its own slots.  NA bounds the atoms per
// molecule.
template <int NA>
__global__ void __launch_bounds__(64) k_respa_inner_lds(
    int nmol, const MolDev* __restrict__ mols, const SpeciesDev* __restrict__ species,
    const TermDev* __restrict__ terms, const SiteDev* __restrict__ sites, double* x, double* y,
    double* z, double* vx, double* vy, double* vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl, double* fxs, double* fys,
    double* fzs, double deltby2, double delts, double deltsby2, int nstep_inner, double* epart,
    Status* st, const double* __restrict__ rep_delt, SfInner W, SnapOut SN) {
#pragma clang fp contract(off)
    __shared__ double lx[NA * 64], ly[NA * 64], lz[NA * 64], lvx[NA * 64], lvy[NA * 64],
        lvz[NA * 64], lfx[NA * 64], lfy[NA * 64], lfz[NA * 64], lrm[NA * 64];
    const int l = threadIdx.x;
    const int m = blockIdx.x * 64 + l;
    if (m >= nmol) return;
    const MolDev M = mols[m];
    const SpeciesDev S = species[M.species];
    rep_steps(rep_delt, M.rep, nstep_inner, deltby2, delts, deltsby2);
    const int na = M.natom, a0 = M.atom0;
    if (snap_wanted(SN, st)) {
        for (int k = 0; k < na; k++) {
            const int i = a0 + k;
            snap_atom(SN, i, x[i], y[i], z[i], vx[i], vy[i], vz[i]);
        }
        if (m == 0) snap_count(SN);
    }
    for (int k = 0; k < na; k++) {
        const int i = a0 + k, s = k * 64 + l;
        const double r_mass = 1.0 / sites[S.site0 + k].aw;
        lrm[s] = r_mass;
        lx[s] = x[i];
        ly[s] = y[i];
        lz[s] = z[i];
        lvx[s] = vx[i] + (deltby2 * fxl[i] * r_mass);
        lvy[s] = vy[i] + (deltby2 * fyl[i] * r_mass);
        lvz[s] = vz[i] + (deltby2 * fzl[i] * r_mass);
        lfx[s] = fxs[i];
        lfy[s] = fys[i];
        lfz[s] = fzs[i];
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0}, vir = 0.0;
    for (int ll = 0; ll < nstep_inner; ll++) {
        for (int k = 0; k < na; k++) {
            const int s = k * 64 + l;
            const double r_mass = lrm[s];
            lvx[s] += (deltsby2 * lfx[s] * r_mass);
            lvy[s] += (deltsby2 * lfy[s] * r_mass);
            lvz[s] += (deltsby2 * lfz[s] * r_mass);
            lx[s] = lx[s] + delts * lvx[s];
            ly[s] = ly[s] + delts * lvy[s];
            lz[s] = lz[s] + delts * lvz[s];
            lfx[s] = lfy[s] = lfz[s] = 0.0;
        }
        e[0] = e[1] = e[2] = e[3] = 0.0;
        vir = 0.0;
        for (int t = S.bterm0; t < S.bterm0 + S.nbterm; t++) {
            double u = 0.0, v = 0.0;
            int sing = 0;
            int cls = bonded_term(terms[t], l, lx, ly, lz, lfx, lfy, lfz, &u, &v, &sing, 64);
            if (sing) atomicOr(&st->flags, ST_SINGULAR);
            e[cls] += u;
            vir += v;
        }
        if (W.on)
            for (int k = 0; k < na; k++) {
                const int s = k * 64 + l;
                double gx = lfx[s], gy = lfy[s], gz = lfz[s];
                const double u = sf_inner_site(W, sites[S.site0 + k], M.rep, lx[s], ly[s], lz[s],
                                               gx, gy, gz, st);
                lfx[s] = gx;
                lfy[s] = gy;
                lfz[s] = gz;
                if (ll == nstep_inner - 1) W.usf[a0 + k] = u;
            }
        for (int k = 0; k < na; k++) {
            const int s = k * 64 + l;
            const double r_mass = lrm[s];
            lvx[s] += (deltsby2 * lfx[s] * r_mass);
            lvy[s] += (deltsby2 * lfy[s] * r_mass);
            lvz[s] += (deltsby2 * lfz[s] * r_mass);
        }
    }
    for (int k = 0; k < na; k++) {
        const int i = a0 + k, s = k * 64 + l;
        x[i] = lx[s];
        y[i] = ly[s];
        z[i] = lz[s];
        vx[i] = lvx[s];
        vy[i] = lvy[s];
        vz[i] = lvz[s];
        fxs[i] = lfx[s];
        fys[i] = lfy[s];
        fzs[i] = lfz[s];
    }
    double* ep = epart + (size_t)m * 5;
    ep[0] = e[0];
    ep[1] = e[1];
    ep[2] = e[2];
    ep[3] = e[3];
    ep[4] = vir;
}

// Molecules of up to NA <= TPM sites, TPM lanes per molecule (64 / TPM
// molecules per single-wave workgroup).  Lane j owns site j (kick, drift)
// and computes bonded terms j, j + TPM, ... into its own LDS copy of the
// molecule's forces; a site's force is then the sum of the TPM copies in
// lane order, so the chain of one inner step is ceil(terms / TPM) terms
// long instead of the molecule's whole term list.  The sums associate the
// terms' contributions differently from frcshort's running sum in term
// order (src/frcshort.c:47-79), a difference of rounding only.
template <int NA, int TPM>
__global__ void __launch_bounds__(64) k_respa_inner_tpm(
    int nmol, const MolDev* __restrict__ mols, const SpeciesDev* __restrict__ species,
    const TermDev* __restrict__ terms, const SiteDev* __restrict__ sites, double* x, double* y,
    double* z, double* vx, double* vy, double* vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl, double* fxs, double* fys,
    double* fzs, double deltby2, double delts, double deltsby2, int nstep_inner, double* epart,
    Status* st, const double* __restrict__ rep_delt, SfInner W, SnapOut SN) {
#pragma clang fp contract(off)
    static_assert(NA <= TPM && 64 % TPM == 0, "one site per lane");
    constexpr int MPW = 64 / TPM;
    __shared__ double lx[NA * MPW], ly[NA * MPW], lz[NA * MPW];
    __shared__ double cfx[TPM][NA * MPW], cfy[TPM][NA * MPW], cfz[TPM][NA * MPW];
    const int l = threadIdx.x, mi = l / TPM, j = l % TPM;
    const int m = blockIdx.x * MPW + mi;
    const bool live = m < nmol;
    int na = 0, a0 = 0, site0 = 0, bt0 = 0, nbt = 0, rep = 0;
    if (live) {
        const MolDev M = mols[m];
        const SpeciesDev S = species[M.species];
        na = M.natom;
        a0 = M.atom0;
        site0 = S.site0;
        bt0 = S.bterm0;
        nbt = S.nbterm;
        rep = M.rep;
        rep_steps(rep_delt, M.rep, nstep_inner, deltby2, delts, deltsby2);
    }
    double usf_last = 0.0;  // the wall's energy of this lane's site (last step)
    const bool own = live && j < na;  // this lane owns site j of its molecule
    const int sl = j * MPW + mi;      // its LDS slot
    double r_mass = 0.0, vxo = 0.0, vyo = 0.0, vzo = 0.0, fxo = 0.0, fyo = 0.0, fzo = 0.0;
    if (own && snap_wanted(SN, st)) {
        const int i = a0 + j;
        snap_atom(SN, i, x[i], y[i], z[i], vx[i], vy[i], vz[i]);
    }
    if (SN.p && m == 0 && j == 0 && snap_wanted(SN, st)) snap_count(SN);
    if (own) {
        const int i = a0 + j;
        r_mass = 1.0 / sites[site0 + j].aw;
        lx[sl] = x[i];
        ly[sl] = y[i];
        lz[sl] = z[i];
        vxo = vx[i] + (deltby2 * fxl[i] * r_mass);
        vyo = vy[i] + (deltby2 * fyl[i] * r_mass);
        vzo = vz[i] + (deltby2 * fzl[i] * r_mass);
        fxo = fxs[i];
        fyo = fys[i];
        fzo = fzs[i];
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0}, vir = 0.0;
    for (int ll = 0; ll < nstep_inner; ll++) {
        __syncthreads();  // the previous step's copies have been read
        if (own) {
            vxo += (deltsby2 * fxo * r_mass);
            vyo += (deltsby2 * fyo * r_mass);
            vzo += (deltsby2 * fzo * r_mass);
            lx[sl] = lx[sl] + delts * vxo;
            ly[sl] = ly[sl] + delts * vyo;
            lz[sl] = lz[sl] + delts * vzo;
        }
#pragma unroll
        for (int k = 0; k < NA; k++) {
            cfx[j][k * MPW + mi] = 0.0;
            cfy[j][k * MPW + mi] = 0.0;
            cfz[j][k * MPW + mi] = 0.0;
        }
        __syncthreads();
        e[0] = e[1] = e[2] = e[3] = 0.0;
        vir = 0.0;
        for (int t = bt0 + j; t < bt0 + nbt; t += TPM) {
            double u = 0.0, v = 0.0;
            int sing = 0;
            const int cls =
                bonded_term(terms[t], mi, lx, ly, lz, cfx[j], cfy[j], cfz[j], &u, &v, &sing, MPW);
            if (sing) atomicOr(&st->flags, ST_SINGULAR);
            e[cls] += u;
            vir += v;
        }
        __syncthreads();
        if (own) {
            double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int c = 0; c < TPM; c++) {
                gx += cfx[c][sl];
                gy += cfy[c][sl];
                gz += cfz[c][sl];
            }
            fxo = gx;
            fyo = gy;
            fzo = gz;
            if (W.on)
                usf_last = sf_inner_site(W, sites[site0 + j], rep, lx[sl], ly[sl], lz[sl], fxo, fyo,
                                         fzo, st);
            vxo += (deltsby2 * fxo * r_mass);
            vyo += (deltsby2 * fyo * r_mass);
            vzo += (deltsby2 * fzo * r_mass);
        }
    }
    if (own) {
        const int i = a0 + j;
        x[i] = lx[sl];
        y[i] = ly[sl];
        z[i] = lz[sl];
        vx[i] = vxo;
        vy[i] = vyo;
        vz[i] = vzo;
        fxs[i] = fxo;
        fys[i] = fyo;
        fzs[i] = fzo;
        if (W.on) W.usf[i] = usf_last;
    }
    // the last inner step's energies, summed over the molecule's lanes
    double ev[5] = {e[0], e[1], e[2], e[3], vir};
#pragma unroll
    for (int o = 1; o < TPM; o <<= 1)
#pragma unroll
        for (int k = 0; k < 5; k++) ev[k] += __shfl_xor(ev[k], o, 64);
    if (live && j == 0) {
        double* ep = epart + (size_t)m * 5;
        for (int k = 0; k < 5; k++) ep[k] = ev[k];
    }
}

// 1 / sqrt(x) within an ulp or two: the hardware estimate (v_rsq_f64) and
// two Newton steps, a shorter dependency chain than sqrt and a division.
__device__ inline double rsqrt_nr(double x) {
#pragma clang fp contract(off)
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * (1.5 - h * y * y);
    y = y * (1.5 - h * y * y);
    return y;
}

// bonded_term for the inner loop's latency chain: the harmonic bond and the
// harmonic angle with reciprocal square roots in place of sqrt and division
// (the same expressions otherwise; results differ from bonded_term's in the
// last bits); every other term is bonded_term's.  The caller's force copy
// holds this one term only and starts from zero, so the two fast terms
// store their forces (0 + f is f) instead of reading the copy back.
__device__ inline int bonded_term_inner(const TermDev& T, int base, const double* xx,
                                        const double* yy, const double* zz, double* fxs,
                                        double* fys, double* fzs, double* u, double* vir,
                                        int* singular, int stride) {
#pragma clang fp contract(off)
    const double* pr = T.para;
    if (T.kind == TERM_BOND && T.type == MSPMS_BOND_HARMONIC) {
        const int i1 = base + T.idx[0] * stride, i2 = base + T.idx[1] * stride;
        const double rx = xx[i1] - xx[i2], ry = yy[i1] - yy[i2], rz = zz[i1] - zz[i2];
        const double rsq = rx * rx + ry * ry + rz * rz, ri = rsqrt_nr(rsq), r = rsq * ri;
        const double dr = r - pr[1];
        *u = 0.5 * pr[0] * dr * dr;
        const double fij = -pr[0] * dr * ri;
        *vir = fij * rsq;
        fxs[i1] = fij * rx;
        fys[i1] = fij * ry;
        fzs[i1] = fij * rz;
        fxs[i2] = -(fij * rx);
        fys[i2] = -(fij * ry);
        fzs[i2] = -(fij * rz);
        return 0;
    }
    if (T.kind == TERM_ANGLE && T.type == MSPMS_ANGLE_HARMONIC) {
        const int ia = base + T.idx[0] * stride, ib = base + T.idx[1] * stride,
                  ic = base + T.idx[2] * stride;
        double xab = xx[ia] - xx[ib], yab = yy[ia] - yy[ib], zab = zz[ia] - zz[ib];
        double xbc = xx[ic] - xx[ib], ybc = yy[ic] - yy[ib], zbc = zz[ic] - zz[ib];
        const double rrab = rsqrt_nr(xab * xab + yab * yab + zab * zab);
        const double rrbc = rsqrt_nr(xbc * xbc + ybc * ybc + zbc * zbc);
        xab = xab * rrab;
        yab = yab * rrab;
        zab = zab * rrab;
        xbc = xbc * rrbc;
        ybc = ybc * rrbc;
        zbc = zbc * rrbc;
        const double cost = xab * xbc + yab * ybc + zab * zbc;
        const double theta = acos(cost);
        // 1 / max(1e-8, sin theta)
        const double s2 = 1.0 - cost * cost;
        const double rsint = s2 > 1.0e-16 ? rsqrt_nr(s2) : 1.0e8;
        const double dth = theta - pr[1];
        *u = 0.5 * pr[0] * dth * dth;
        const double gam = pr[0] * dth * rsint;
        const double fxa = gam * (xbc - xab * cost) * rrab, fya = gam * (ybc - yab * cost) * rrab,
                     fza = gam * (zbc - zab * cost) * rrab;
        const double fxc = gam * (xab - xbc * cost) * rrbc, fyc = gam * (yab - ybc * cost) * rrbc,
                     fzc = gam * (zab - zbc * cost) * rrbc;
        *vir = 0.0;
        fxs[ia] = fxa;
        fys[ia] = fya;
        fzs[ia] = fza;
        fxs[ib] = -fxa - fxc;
        fys[ib] = -fya - fyc;
        fzs[ib] = -fza - fzc;
        fxs[ic] = fxc;
        fys[ic] = fyc;
        fzs[ic] = fzc;
        return 1;
    }
    return bonded_term(T, base, xx, yy, zz, fxs, fys, fzs, u, vir, singular, stride);
}

// The inner loop with one bonded term per wave: a workgroup of NW waves
// takes 64 molecules (lane = molecule); wave w owns site w of each molecule
// (kicks, drift) and computes its term w into its own LDS copy of the
// molecule's forces, so that a wave runs one kind of term and no wave
// steps through the others' code (with one lane per site and term in a
// wave, bonds and angles alternate within the wave and each waits for the
// other).  A site's force is the sum of the NW copies in wave (= term)
// order, as k_respa_inner_tpm sums its lane copies.  Molecules of up to NW
// sites and NW terms.
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_respa_inner_tw(
    int nmol, const MolDev* __restrict__ mols, const SpeciesDev* __restrict__ species,
    const TermDev* __restrict__ terms, const SiteDev* __restrict__ sites, double* x, double* y,
    double* z, double* vx, double* vy, double* vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl, double* fxs, double* fys,
    double* fzs, double deltby2, double delts, double deltsby2, int nstep_inner, double* epart,
    Status* st, const double* __restrict__ rep_delt, SfInner W, SnapOut SN) {
#pragma clang fp contract(off)
    __shared__ double lx[NW * 64], ly[NW * 64], lz[NW * 64];
    __shared__ double cfx[NW][NW * 64], cfy[NW][NW * 64], cfz[NW][NW * 64];
    __shared__ double le[NW][5][64];
    const int w = threadIdx.x >> 6, mi = threadIdx.x & 63;
    const int m = blockIdx.x * 64 + mi;
    const bool live = m < nmol;
    int na = 0, a0 = 0, site0 = 0, nbt = 0, rep = 0;
    TermDev T{};
    if (live) {
        const MolDev M = mols[m];
        na = M.natom;
        a0 = M.atom0;
        site0 = M.site0;
        nbt = M.nbterm;
        rep = M.rep;
        if (w < nbt) T = terms[M.bterm0 + w];  // this wave's term, for every substep
        rep_steps(rep_delt, M.rep, nstep_inner, deltby2, delts, deltsby2);
    }
    const bool own = live && w < na;      // site w of molecule m
    const bool term = live && w < nbt;    // term w of molecule m
    const int sl = w * 64 + mi;
    double usf_last = 0.0;
    double r_mass = 0.0, vxo = 0.0, vyo = 0.0, vzo = 0.0, fxo = 0.0, fyo = 0.0, fzo = 0.0;
    double pxo = 0.0, pyo = 0.0, pzo = 0.0;  // the own site's position (LDS holds a copy)
    if (own && snap_wanted(SN, st)) {
        const int i = a0 + w;
        snap_atom(SN, i, x[i], y[i], z[i], vx[i], vy[i], vz[i]);
    }
    if (SN.p && m == 0 && w == 0 && snap_wanted(SN, st)) snap_count(SN);
    if (own) {
        const int i = a0 + w;
        r_mass = 1.0 / sites[site0 + w].aw;
        pxo = x[i];
        pyo = y[i];
        pzo = z[i];
        lx[sl] = pxo;
        ly[sl] = pyo;
        lz[sl] = pzo;
        vxo = vx[i] + (deltby2 * fxl[i] * r_mass);
        vyo = vy[i] + (deltby2 * fyl[i] * r_mass);
        vzo = vz[i] + (deltby2 * fzl[i] * r_mass);
        fxo = fxs[i];
        fyo = fys[i];
        fzo = fzs[i];
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0}, vir = 0.0;
    // a lane's entries of its wave's copy are written by that lane only.
    // The fast terms (harmonic bond and angle) store the same entries every
    // step, so the copy is cleared once; other terms accumulate into it and
    // clear it every step (in program order, no barrier needed)
    const bool stores = term && ((T.kind == TERM_BOND && T.type == MSPMS_BOND_HARMONIC) ||
                                 (T.kind == TERM_ANGLE && T.type == MSPMS_ANGLE_HARMONIC));
#pragma unroll
    for (int k = 0; k < NW; k++) {
        cfx[w][k * 64 + mi] = 0.0;
        cfy[w][k * 64 + mi] = 0.0;
        cfz[w][k * 64 + mi] = 0.0;
    }
    for (int ll = 0; ll < nstep_inner; ll++) {
        // the drift writes only this lane's site, which the previous step's
        // terms read before the barrier that ended them
        if (own) {
            vxo += (deltsby2 * fxo * r_mass);
            vyo += (deltsby2 * fyo * r_mass);
            vzo += (deltsby2 * fzo * r_mass);
            pxo = pxo + delts * vxo;
            pyo = pyo + delts * vyo;
            pzo = pzo + delts * vzo;
            lx[sl] = pxo;
            ly[sl] = pyo;
            lz[sl] = pzo;
        }
        __syncthreads();  // the positions are in; every copy has been read
        if (!stores && ll > 0) {
#pragma unroll
            for (int k = 0; k < NW; k++) {
                cfx[w][k * 64 + mi] = 0.0;
                cfy[w][k * 64 + mi] = 0.0;
                cfz[w][k * 64 + mi] = 0.0;
            }
        }
        e[0] = e[1] = e[2] = e[3] = 0.0;
        vir = 0.0;
        if (term) {
            double u = 0.0, v = 0.0;
            int sing = 0;
            const int cls =
                bonded_term_inner(T, mi, lx, ly, lz, cfx[w], cfy[w], cfz[w], &u, &v, &sing, 64);
            if (sing) atomicOr(&st->flags, ST_SINGULAR);
            e[cls] += u;
            vir += v;
        }
        __syncthreads();
        if (own) {
            double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int c = 0; c < NW; c++) {
                gx += cfx[c][sl];
                gy += cfy[c][sl];
                gz += cfz[c][sl];
            }
            fxo = gx;
            fyo = gy;
            fzo = gz;
            if (W.on)
                usf_last = sf_inner_site(W, sites[site0 + w], rep, pxo, pyo, pzo, fxo, fyo, fzo,
                                         st);
            vxo += (deltsby2 * fxo * r_mass);
            vyo += (deltsby2 * fyo * r_mass);
            vzo += (deltsby2 * fzo * r_mass);
        }
    }
    if (own) {
        const int i = a0 + w;
        x[i] = pxo;
        y[i] = pyo;
        z[i] = pzo;
        vx[i] = vxo;
        vy[i] = vyo;
        vz[i] = vzo;
        fxs[i] = fxo;
        fys[i] = fyo;
        fzs[i] = fzo;
        if (W.on) W.usf[i] = usf_last;
    }
    // the last inner step's energies, summed over the molecule's terms in
    // term order
    le[w][0][mi] = e[0];
    le[w][1][mi] = e[1];
    le[w][2][mi] = e[2];
    le[w][3][mi] = e[3];
    le[w][4][mi] = vir;
    __syncthreads();
    if (live && w == 0) {
        double* ep = epart + (size_t)m * 5;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < NW; c++) s += le[c][k][mi];
            ep[k] = s;
        }
    }
}

// The tricubic Hermite wall of one site by four lanes (q = lane & 3 of a
// quad): a lookup reads 4 contiguous 64-B pieces (the x-neighbours' 32-B
// records of the four (y, z) corner pairs); lane q loads the q-th 16 B of
// each, so one load instruction of the wave fetches 16 whole pieces where
// one lane per site fetches 64 scattered quarters.  Lane q's values are
// those of corner x offset q >> 1 and derivative bits 4 (q & 1) + k; its
// partial sums are added over the quad (DPP, every lane gets them).  The
// arithmetic is the lookup of sf_site (mylibtasos/forces.f90:6-173,
// cubicinterpolation.f90:9-116) with its terms summed in another order.
// All four lanes must be active.
__device__ inline double sf_tasos_quad(const SFDev& sf, const SiteDev& S, double xi, double yi,
                                       double zi, double& fxi, double& fyi, double& fzi,
                                       int* edom) {
    const int q = threadIdx.x & 3;
    const int ty = S.interp - 1;
    double xp = xi * sf.sigma_base - sf.x0, yp = yi * sf.sigma_base - sf.y0,
           zp = zi * sf.sigma_base - sf.z0;
    const double pz = sf.zperiod;
    zp = zp - floor(zp / pz) * pz;  // mylibtasos/pbc.f90:82-86
    const int ix = (int)floor(xp / sf.hx), iy = (int)floor(yp / sf.hy), iz = (int)floor(zp / sf.hz);
    const bool ok = ty >= 0 && ty < sf.ntype && ix >= 0 && iy >= 0 && iz >= 0 &&
                    ix <= sf.nx - 2 && iy <= sf.ny - 2 && iz <= sf.nz - 2;
    double fval = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (ok) {
        double bx[4], by[4], bz[4], dbx[4], dby[4], dbz[4];
        sfeval::hermite_1d((xp - ix * sf.hx) / sf.hx, sf.hx, bx, dbx);
        sfeval::hermite_1d((yp - iy * sf.hy) / sf.hy, sf.hy, by, dby);
        sfeval::hermite_1d((zp - iz * sf.hz) / sf.hz, sf.hz, bz, dbz);
        const size_t nn = (size_t)sf.nx * sf.ny * sf.nz;
        const float4* base = reinterpret_cast<const float4*>(sf.nodes) + (size_t)ty * 2 * nn;
        const int cx = q >> 1, half = q & 1;
        float4 rec[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {  // (cy, cz) = (c & 1, c >> 1)
            const size_t node =
                ((size_t)(iz + (c >> 1)) * sf.ny + (iy + (c & 1))) * sf.nx + (ix + cx);
            rec[c] = base[2 * node + half];
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int cy = c & 1, cz = c >> 1;
            const float vals[4] = {rec[c].x, rec[c].y, rec[c].z, rec[c].w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int ox = k & 1, oy = k >> 1, oz = half;
                const double v = (double)vals[k];
                const double wx = bx[cx + 2 * ox], wy = by[cy + 2 * oy], wz = bz[cz + 2 * oz];
                fval += v * wx * wy * wz;
                gx += v * dbx[cx + 2 * ox] * wy * wz;
                gy += v * wx * dby[cy + 2 * oy] * wz;
                gz += v * wx * wy * dbz[cz + 2 * oz];
            }
        }
    }
    // the quad's sums (lanes q ^ 1, then q ^ 2)
    fval += dpp64<0xB1>(fval);
    gx += dpp64<0xB1>(gx);
    gy += dpp64<0xB1>(gy);
    gz += dpp64<0xB1>(gz);
    fval += dpp64<0x4E>(fval);
    gx += dpp64<0x4E>(gx);
    gy += dpp64<0x4E>(gy);
    gz += dpp64<0x4E>(gz);
    if (!ok) {
        *edom = 1;
        return 0.0;
    }
    const double pe = (fval < 1.5 * sf.min_pe[ty]) ? -fval : fval;  // forces.f90:166-171
    const double fconv = sf.sigma_base / sf.epsilon_base;
    fxi += -gx * fconv;
    fyi += -gy * fconv;
    fzi += -gz * fconv;
    return pe / sf.epsilon_base;
}

// The inner loop of single-site molecules with the tricubic Hermite wall in
// it (the RESPA split): four lanes per site, 16 sites per wave.  All four
// lanes carry the site's kicks and drifts (the same operations on the same
// values, so the same results) and share the wall's lookup; lane 0 writes.
// Single-site molecules have no bonded terms.  Lanes past the last molecule
// work on the last one and write nothing, so every quad stays whole.
__global__ void __launch_bounds__(64) k_respa_inner_grid4(
    int nmol, const MolDev* __restrict__ mols, const SpeciesDev* __restrict__ species,
    const SiteDev* __restrict__ sites, double* x, double* y, double* z, double* vx, double* vy,
    double* vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl, double* fxs, double* fys,
    double* fzs, double deltby2, double delts, double deltsby2, int nstep_inner, double* epart,
    Status* st, const double* __restrict__ rep_delt, SfInner W, SnapOut SN) {
#pragma clang fp contract(off)
    const int q = threadIdx.x & 3;
    const int m0 = blockIdx.x * 16 + (threadIdx.x >> 2);
    const bool live = m0 < nmol;
    const int m = live ? m0 : nmol - 1;
    const MolDev M = mols[m];
    const int i = M.atom0;
    const SiteDev S = sites[species[M.species].site0];
    rep_steps(rep_delt, M.rep, nstep_inner, deltby2, delts, deltsby2);
    const bool writer = live && q == 0;
    if (writer && snap_wanted(SN, st)) snap_atom(SN, i, x[i], y[i], z[i], vx[i], vy[i], vz[i]);
    if (SN.p && m0 == 0 && q == 0 && snap_wanted(SN, st)) snap_count(SN);
    const double r_mass = 1.0 / S.aw;
    double xi = x[i], yi = y[i], zi = z[i];
    double vxo = vx[i] + (deltby2 * fxl[i] * r_mass);
    double vyo = vy[i] + (deltby2 * fyl[i] * r_mass);
    double vzo = vz[i] + (deltby2 * fzl[i] * r_mass);
    double fxo = fxs[i], fyo = fys[i], fzo = fzs[i];
    const bool wall = S.ghost == MSPMS_GHOST_NONE;
    double usf_last = 0.0;
    int edom = 0;
    for (int ll = 0; ll < nstep_inner; ll++) {
        vxo += (deltsby2 * fxo * r_mass);
        vyo += (deltsby2 * fyo * r_mass);
        vzo += (deltsby2 * fzo * r_mass);
        xi = xi + delts * vxo;
        yi = yi + delts * vyo;
        zi = zi + delts * vzo;
        // frcshort of a single-site molecule: no bonded term, the wall only
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const double u = sf_tasos_quad(W.sf, S, xi, yi, zi, gx, gy, gz, &edom);
        if (wall) {
            fxo = gx;
            fyo = gy;
            fzo = gz;
            usf_last = u;
        } else {
            fxo = fyo = fzo = 0.0;
            usf_last = 0.0;
        }
        vxo += (deltsby2 * fxo * r_mass);
        vyo += (deltsby2 * fyo * r_mass);
        vzo += (deltsby2 * fzo * r_mass);
    }
    if (edom && wall && writer) atomicOr(&st->flags, ST_GRID_EDOM);
    if (!writer) return;
    x[i] = xi;
    y[i] = yi;
    z[i] = zi;
    vx[i] = vxo;
    vy[i] = vyo;
    vz[i] = vzo;
    fxs[i] = fxo;
    fys[i] = fyo;
    fzs[i] = fzo;
    W.usf[i] = usf_last;
    double* ep = epart + (size_t)m * 5;
    for (int k = 0; k < 5; k++) ep[k] = 0.0;
}

__global__ void __launch_bounds__(256) k_respa_final(int natom, const int* __restrict__ atom_site,
                                                     const SiteDev* __restrict__ sites, double* vx,
                                                     double* vy, double* vz,
                                                     const double* __restrict__ fxl,
                                                     const double* __restrict__ fyl,
                                                     const double* __restrict__ fzl,
                                                     double deltby2, double* kin,
                                                     const double* __restrict__ rep_delt,
                                                     const int* __restrict__ atom_mol,
                                                     const MolDev* __restrict__ mols) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natom) return;
    if (rep_delt) deltby2 = rep_delt[mols[atom_mol[i]].rep] / 2.0;
    double r_mass = 1.0 / sites[atom_site[i]].aw;
    double a = vx[i] + (deltby2 * fxl[i] * r_mass);
    double b = vy[i] + (deltby2 * fyl[i] * r_mass);
    double c = vz[i] + (deltby2 * fzl[i] * r_mass);
    vx[i] = a;
    vy[i] = b;
    vz[i] = c;
    kin[i] = (a * a + b * b + c * c) / r_mass;
}

}  // namespace

void launch_respa_inner(const LaunchCtx& c, const SpeciesDev* species, const TermDev* terms,
                        double* x, double* y, double* z, double* vx, double* vy, double* vz,
                        const double* fxl, const double* fyl, const double* fzl, double* fxs,
                        double* fys, double* fzs, double deltby2, double delts,
                        double deltsby2, int nstep_inner, double* epart, const SFDev* sf,
                        double* usf) {
    SnapOut SN{c.snap, c.snap_count, c.natom};
    SfInner W{};
    if (sf) {
        W.sf = *sf;
        W.reps = c.reps;
        W.usf = usf;
        W.on = 1;
    }
#define RESPA_LDS(NA)                                                                        \
    hipLaunchKernelGGL(k_respa_inner_lds<NA>, dim3((c.nmol + 63) / 64), dim3(64), 0, c.stream,   \
                       c.nmol, c.mols, species, terms, c.sites, x, y, z, vx, vy, vz, fxl, fyl,   \
                       fzl, fxs, fys, fzs, deltby2, delts, deltsby2, nstep_inner, epart,        \
                       c.status, c.rep_delt, W, SN)
#define RESPA_TW(NW)                                                                         \
    hipLaunchKernelGGL(k_respa_inner_tw<NW>, dim3((c.nmol + 63) / 64), dim3(64 * NW), 0,          \
                       c.stream, c.nmol, c.mols, species, terms, c.sites, x, y, z, vx, vy, vz,    \
                       fxl, fyl, fzl, fxs, fys, fzs, deltby2, delts, deltsby2, nstep_inner,       \
                       epart, c.status, c.rep_delt, W, SN)
#define RESPA_TPM(NA, TPM)                                                                   \
    hipLaunchKernelGGL((k_respa_inner_tpm<NA, TPM>), dim3((c.nmol + 64 / TPM - 1) / (64 / TPM)), \
                       dim3(64), 0, c.stream, c.nmol, c.mols, species, terms, c.sites, x, y, z,  \
                       vx, vy, vz, fxl, fyl, fzl, fxs, fys, fzs, deltby2, delts, deltsby2,       \
                       nstep_inner, epart, c.status, c.rep_delt, W, SN)
    if (c.max_mol_atoms <= 1 && sf && sf->type == MSPMS_SF_TASOS) {
        // single-site molecules with the Hermite wall: four lanes per site
        hipLaunchKernelGGL(k_respa_inner_grid4, dim3((c.nmol + 15) / 16), dim3(64), 0, c.stream,
                           c.nmol, c.mols, species, c.sites, x, y, z, vx, vy, vz, fxl, fyl, fzl,
                           fxs, fys, fzs, deltby2, delts, deltsby2, nstep_inner, epart, c.status,
                           c.rep_delt, W, SN);
    } else if (c.max_mol_atoms <= 1) {
        RESPA_TPM(1, 1);  // one lane per single-site molecule
    } else if (c.max_mol_atoms <= 3 && c.max_mol_terms <= 3) {
        RESPA_TW(3);  // e.g. flexible three-site water: one term per wave
    } else if (c.max_mol_atoms <= 4 && c.max_mol_terms <= 4) {
        RESPA_TW(4);
    } else if (c.max_mol_atoms <= 4) {
        RESPA_TPM(4, 4);
    } else if (c.max_mol_atoms <= 8) {
        RESPA_LDS(8);
    } else if (c.max_mol_atoms <= 16) {
        RESPA_LDS(16);
    } else {
        int g = (c.nmol + 255) / 256;
        hipLaunchKernelGGL(k_respa_inner, dim3(g), dim3(256), 0, c.stream, c.nmol, c.mols,
                           species, terms, c.sites, x, y, z, vx, vy, vz, fxl, fyl, fzl, fxs, fys,
                           fzs, deltby2, delts, deltsby2, nstep_inner, epart, c.status,
                           c.rep_delt, W, SN);
    }
#undef RESPA_LDS
#undef RESPA_TW
#undef RESPA_TPM
}

void launch_respa_final(const LaunchCtx& c, double* vx, double* vy, double* vz,
                        const double* fxl, const double* fyl, const double* fzl,
                        double deltby2, double* kin) {
    int g = (c.natom + 255) / 256;
    hipLaunchKernelGGL(k_respa_final, dim3(g), dim3(256), 0, c.stream, c.natom, c.atom_site,
                       c.sites, vx, vy, vz, fxl, fyl, fzl, deltby2, kin, c.rep_delt, c.atom_mol,
                       c.mols);
}
