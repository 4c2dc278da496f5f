// k_npt.hip -- NPT RESPA step on gfx950: npt_respa with the Nose-Hoover
// thermostat / barostat operator npt_nh_operator (src/nptnh.c:10-234).
//
// Every phase that couples the atoms of a replica (sums of m v^2, scaling of
// all positions, the box) runs as one workgroup per replica, with the
// fixed-order workgroup sums of the NVT kernels; the box changes after every
// inner step, so this kernel also rewrites the replica's box-dependent
// tables that the force kernels read (RepDev: L, 1/L, cell sizes, and for
// Ewald 2 pi / L, 2 pi / V, 2 pi / 3V) as the reference rewrites its
// globals (src/nptnh.c:180-199), and k_kvec refreshes the Ewald weights
// exp(-B k^2) / k^2 of the new box before frclong.
//   k_npt_op<START>   operator (src/nptnh.c:10-130); outer half kick
//   k_npt_inner<FIRST> [previous inner step's closing half kick]; half kick,
//                     x := AA (AA x + delts v), box *= exp(delts vbs)
//   k_bonded          frcshort
//   ... nstep_inner times, then k_kvec, frclong and the finalize pass, then
//   k_npt_op<END>     last inner closing kick, outer half kick, operator
// FP64 in the reference's expression order; only the sums over atoms are
// taken in a tree order.
#include "engine.h"
#include "wave.h"
#include "../../include/mspms_gpu.h"

namespace {

constexpr int NPT_T = 1024;
constexpr double NPT_PI = 3.141592653589793;  // src/defs.h:98

__device__ double wg_sum_npt(double v, double* sh) {
    double a[1] = {v};
    butterfly<double, 1>(a, 32);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = a[0];
    __syncthreads();
    double t[1] = {sh[lane & 15]};
    butterfly<double, 1>(t, 8);
    __syncthreads();
    return t[0];
}

__device__ inline double rmass(const SiteDev* sites, const int* atom_site, int i) {
    return 1.0 / sites[atom_site[i]].aw;
}

// npt_nh_operator on the sum m v^2 of the current velocities; returns the
// velocity scale AA (every thread computes the same values; thread 0
// stores the state)
__device__ double npt_operator(NptDev& S, double mvsq, double vir_inter, double vir_intra,
                               int natom, double deltby2, double dt_outer2, double dt_outer4,
                               double dt_outer8, bool store, double* ukin_out) {
#pragma clang fp contract(off)
    const int nfree = S.nfree;
    const double treq = S.treq, preq = S.preq, Qts = S.Qts, Qbs = S.Qbs, boxv = S.boxv;
    const double N_plus1RT = (nfree + 1) * treq;
    const double one_3N = 1.0 + 3.0 / nfree;
    double vts = S.vts, vbs = S.vbs, rts = S.rts;
    const double pdiff = vir_inter + vir_intra + S.pljlrc * boxv * 3.0 - preq * boxv * 3.0;
    double Gts = (mvsq + Qbs * vbs * vbs - N_plus1RT) / Qts;
    vts = vts + dt_outer4 * Gts;
    const double BB = exp(-vts * dt_outer8);
    vbs = vbs * BB;
    double Gbs = (one_3N * mvsq + pdiff) / Qbs;
    vbs = vbs + dt_outer4 * Gbs;
    vbs = vbs * BB;
    rts = rts + deltby2 * vts;
    const double AA = exp(-dt_outer2 * (vts + one_3N * vbs));
    vbs = vbs * BB;
    mvsq = mvsq * AA * AA;
    Gbs = (one_3N * mvsq + pdiff) / Qbs;
    vbs = vbs + dt_outer4 * Gbs;
    vbs = vbs * BB;
    Gts = (mvsq + Qbs * vbs * vbs - N_plus1RT) / Qts;
    vts = vts + dt_outer4 * Gts;
    if (store) {
        S.vts = vts;
        S.vbs = vbs;
        S.rts = rts;
        S.utsbs = 0.5 * Qbs * vbs * vbs + 0.5 * Qts * vts * vts + (nfree + 1) * treq * rts +
                  preq * boxv;
        const double uk = mvsq / 2.0;
        S.ukin = uk;
        S.tinst = 2.0 * uk / nfree;
        const double pideal = natom / boxv * S.tinst;
        S.pinst = pideal + (vir_inter + vir_intra) / 3.0 / boxv + S.pljlrc;
        if (ukin_out) *ukin_out = uk;
    }
    return AA;
}

template <bool END>
__global__ void __launch_bounds__(NPT_T) k_npt_op(
    const RepDev* __restrict__ reps, const SiteDev* __restrict__ sites,
    const int* __restrict__ atom_site, NptDev* __restrict__ st, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl,
    const double* __restrict__ fxs, const double* __restrict__ fys,
    const double* __restrict__ fzs, const mspms_long_terms* __restrict__ lterms,
    const mspms_short_terms* __restrict__ sterms, double deltby2, double deltsby2,
    double dt_outer4, double dt_outer8, double* __restrict__ ukin) {
#pragma clang fp contract(off)
    __shared__ double sh[16];
    const int r = blockIdx.x;
    const RepDev& R = reps[r];
    NptDev S = st[r];
    const int a0 = R.atom0, a1 = R.atom0 + R.natom;
    double m2 = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += NPT_T) {
        const double r_mass = rmass(sites, atom_site, i);
        double a = vx[i], b = vy[i], c = vz[i];
        if (END) {
            // closing half kick of the last inner step (src/nptnh.c:201-210)
            // and the outer one with the new long forces (:214-223)
            a += (deltsby2 * fxs[i] * r_mass);
            b += (deltsby2 * fys[i] * r_mass);
            c += (deltsby2 * fzs[i] * r_mass);
            a += (deltby2 * fxl[i] * r_mass);
            b += (deltby2 * fyl[i] * r_mass);
            c += (deltby2 * fzl[i] * r_mass);
            vx[i] = a;
            vy[i] = b;
            vz[i] = c;
        }
        m2 += (a * a + b * b + c * c) / r_mass;
    }
    const double mvsq = wg_sum_npt(m2, sh);
    const double AA = npt_operator(S, mvsq, lterms[r].virial_inter, sterms[r].virial_intra,
                                   R.natom, deltby2, deltby2, dt_outer4, dt_outer8,
                                   threadIdx.x == 0, END && ukin ? ukin + r : nullptr);
    for (int i = a0 + threadIdx.x; i < a1; i += NPT_T) {
        double a = vx[i] * AA, b = vy[i] * AA, c = vz[i] * AA;
        if (!END) {
            // outer half kick (src/nptnh.c:141-150)
            const double r_mass = rmass(sites, atom_site, i);
            a += (deltby2 * fxl[i] * r_mass);
            b += (deltby2 * fyl[i] * r_mass);
            c += (deltby2 * fzl[i] * r_mass);
        }
        vx[i] = a;
        vy[i] = b;
        vz[i] = c;
    }
    if (threadIdx.x == 0) st[r] = S;
}

template <bool FIRST>
__global__ void __launch_bounds__(NPT_T) k_npt_inner(
    RepDev* __restrict__ reps, const SiteDev* __restrict__ sites,
    const int* __restrict__ atom_site, NptDev* __restrict__ st, double* __restrict__ x,
    double* __restrict__ y, double* __restrict__ z, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ fxs,
    const double* __restrict__ fys, const double* __restrict__ fzs, double delts,
    double deltsby2) {
#pragma clang fp contract(off)
    const int r = blockIdx.x;
    RepDev& R = reps[r];
    NptDev* S = st + r;
    const double vbs = S->vbs;
    const double AA = exp(deltsby2 * vbs);  // src/nptnh.c:157
    const int a0 = R.atom0, a1 = R.atom0 + R.natom;
    for (int i = a0 + threadIdx.x; i < a1; i += NPT_T) {
        const double r_mass = rmass(sites, atom_site, i);
        double a = vx[i], b = vy[i], c = vz[i];
        if (!FIRST) {
            // the previous inner step's closing half kick (src/nptnh.c:201-210)
            a += (deltsby2 * fxs[i] * r_mass);
            b += (deltsby2 * fys[i] * r_mass);
            c += (deltsby2 * fzs[i] * r_mass);
        }
        a += (deltsby2 * fxs[i] * r_mass);
        b += (deltsby2 * fys[i] * r_mass);
        c += (deltsby2 * fzs[i] * r_mass);
        vx[i] = a;
        vy[i] = b;
        vz[i] = c;
        double px = x[i] * AA, py = y[i] * AA, pz = z[i] * AA;
        px = px + delts * a;
        py = py + delts * b;
        pz = pz + delts * c;
        x[i] = px * AA;
        y[i] = py * AA;
        z[i] = pz * AA;
    }
    // every atom of the replica has read the old box before it changes
    __syncthreads();
    if (threadIdx.x != 0) return;
    // src/nptnh.c:179-199
    const double expfactor = exp(delts * vbs);
    const int nc[3] = {R.ncx, R.ncy, R.ncz};
    for (int d = 0; d < 3; d++) {
        R.L[d] = R.L[d] * expfactor;
        R.invL[d] = 1.0 / R.L[d];
        R.h[d] = R.L[d] / nc[d];
        R.inv_h[d] = 1.0 / R.h[d];
    }
    const double boxv = R.L[0] * R.L[1] * R.L[2];
    S->boxv = boxv;
    double p = 0.0;
    for (int k = 0; k < S->nlrc; k++) p += S->lrc[k] / boxv / boxv;
    S->pljlrc = p;
    if (R.charge_type == MSPMS_ELEC_EWALD) {
        R.vfactor = 2.0 * NPT_PI / boxv;
        for (int d = 0; d < 3; d++) R.twopi_l[d] = 2.0 * NPT_PI / R.L[d];
        R.twopi_over_3v = 2.0 * NPT_PI / 3.0 / boxv;
    }
}

// Ewald weights exp(-B k^2) / k^2 of the current box, one thread per k row
// (as the reference computes them in frclong, src/frclong.c:1003-1008)
__global__ void __launch_bounds__(256) k_kvec(int nrow, const KRow* __restrict__ rows,
                                              const RepDev* __restrict__ reps,
                                              double* __restrict__ kvec) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow) return;
    const KRow w = rows[t];
    const RepDev& R = reps[w.rep];
    const double rkx = R.twopi_l[0] * w.kx, rky = R.twopi_l[1] * w.ky;
    for (int kz = w.kz_lo; kz <= w.kz_hi; kz++) {
        const double rkz = R.twopi_l[2] * kz;
        const double rksq = rkx * rkx + rky * rky + rkz * rkz;
        // twice the weight: each k stands for k and -k (the half-space
        // enumeration of the reciprocal tables)
        kvec[w.k0 + (kz - w.kz_lo)] = 2.0 * (exp(-R.bfactor * rksq) / rksq);
    }
}

}  // namespace

void launch_npt_op(const LaunchCtx& c, NptDev* st, bool end, double* vel, const double* fl,
                   const double* fs, const void* lterms, const void* sterms, double deltby2,
                   double deltsby2, double dt_outer4, double dt_outer8, double* ukin) {
    const size_t n = (size_t)c.natom;
    auto* lt = (const mspms_long_terms*)lterms;
    auto* sm = (const mspms_short_terms*)sterms;
    if (end)
        hipLaunchKernelGGL(k_npt_op<true>, dim3(c.nrep), dim3(NPT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, vel, vel + n, vel + 2 * n, fl, fl + n,
                           fl + 2 * n, fs, fs + n, fs + 2 * n, lt, sm, deltby2, deltsby2,
                           dt_outer4, dt_outer8, ukin);
    else
        hipLaunchKernelGGL(k_npt_op<false>, dim3(c.nrep), dim3(NPT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, vel, vel + n, vel + 2 * n, fl, fl + n,
                           fl + 2 * n, fs, fs + n, fs + 2 * n, lt, sm, deltby2, deltsby2,
                           dt_outer4, dt_outer8, ukin);
}

void launch_npt_inner(const LaunchCtx& c, RepDev* reps, NptDev* st, bool first, double* pos,
                      double* vel, const double* fs, double delts, double deltsby2) {
    const size_t n = (size_t)c.natom;
    if (first)
        hipLaunchKernelGGL(k_npt_inner<true>, dim3(c.nrep), dim3(NPT_T), 0, c.stream, reps,
                           c.sites, c.atom_site, st, pos, pos + n, pos + 2 * n, vel, vel + n,
                           vel + 2 * n, fs, fs + n, fs + 2 * n, delts, deltsby2);
    else
        hipLaunchKernelGGL(k_npt_inner<false>, dim3(c.nrep), dim3(NPT_T), 0, c.stream, reps,
                           c.sites, c.atom_site, st, pos, pos + n, pos + 2 * n, vel, vel + n,
                           vel + 2 * n, fs, fs + n, fs + 2 * n, delts, deltsby2);
}

void launch_kvec(const LaunchCtx& c, const RecipBufs& rb, int nrow) {
    if (nrow <= 0) return;
    hipLaunchKernelGGL(k_kvec, dim3((nrow + 255) / 256), dim3(256), 0, c.stream, nrow, rb.rows,
                       c.reps, rb.kvec);
}
