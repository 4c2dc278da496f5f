// k_nvt.hip -- Nose-Hoover NVT RESPA step on gfx950: vver_nh_3
// (src/nvtnh.c:11-221), two thermostats (outer: ss, ps, qq; inner: sss,
// pss, qqs) as in Frenkel and Smit.
//
// The thermostats couple all atoms of a replica through sums of m v^2, and
// the implicit end-of-(sub)step velocity update is a Newton iteration whose
// every pass needs two sums over the replica's atoms.  One workgroup per
// replica runs each such phase; its sums are fixed-order block reductions
// whose result every thread holds bit-identically, so the iteration's
// control flow is uniform and the run deterministic:
//   k_nvt_drift<FIRST>   [outer half kick with ps (first substep)], sum of
//                        m v^2, inner half kick with pss, drift; then the
//                        sss, pss update (src/nvtnh.c:27-56)
//   k_bonded             frcshort (src/nvtnh.c:59)
//   k_nvt_solve<false>   implicit inner half kick: Newton on (v, pss)
//                        (src/nvtnh.c:61-157)
//   ... nstep_inner times, then frclong, then
//   k_nvt_solve<true>    ss, ps update (:161-162); implicit outer half kick:
//                        Newton on (v, ps) (:167-252); ukin, tinst and the
//                        thermostat energies (:254-263)
// Arithmetic is FP64 in the reference's expression order (no contraction);
// only the sums over atoms are taken in a different (tree) order.  The
// Newton loop ends as the reference's does: after 100 passes or when
// |(p_new - p_old) / p_new| <= 1e-10 -- the reference's per-atom test reads
// the unused element natom (loop index left at natom, zero), so 0/0 is
// never greater than the tolerance and only the thermostat test decides.
#include "engine.h"
#include "wave.h"
#include "../../include/mspms_gpu.h"

namespace {

constexpr int NVT_T = 1024;           // threads per replica workgroup
constexpr double NVT_ERR = 1.0e-10;   // src/nvtnh.c:19

// sum of v over the workgroup, in a fixed order, returned to every thread
// (each wave adds the 16 wave partials with the same tree)
__device__ double wg_sum(double v, double* sh) {
    double a[1] = {v};
    butterfly<double, 1>(a, 32);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = a[0];
    __syncthreads();
    double t[1] = {sh[lane & 15]};
    butterfly<double, 1>(t, 8);
    __syncthreads();  // sh free for the next sum
    return t[0];
}

__device__ inline double rmass(const SiteDev* sites, const int* atom_site, int i) {
    return 1.0 / sites[atom_site[i]].aw;
}

template <bool FIRST>
__global__ void __launch_bounds__(NVT_T) k_nvt_drift(
    const RepDev* __restrict__ reps, const SiteDev* __restrict__ sites,
    const int* __restrict__ atom_site, mspms_nvt_state* __restrict__ st, double* __restrict__ x,
    double* __restrict__ y, double* __restrict__ z, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ fxl,
    const double* __restrict__ fyl, const double* __restrict__ fzl,
    const double* __restrict__ fxs, const double* __restrict__ fys,
    const double* __restrict__ fzs, double deltby2, double delts, double deltsby2,
    double delts_sqby2, const int* __restrict__ active) {
#pragma clang fp contract(off)
    __shared__ double sh[16];
    if (active && !active[blockIdx.x]) return;  // a frozen replica (whole workgroup)
    const RepDev& R = reps[blockIdx.x];
    mspms_nvt_state* S = st + blockIdx.x;
    const double ps = S->ps, pss = S->pss;
    double m2 = 0.0;
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += NVT_T) {
        const double r_mass = rmass(sites, atom_site, i);
        double a = vx[i], b = vy[i], c = vz[i];
        if (FIRST) {
            a = a + deltby2 * (fxl[i] * r_mass - ps * a);
            b = b + deltby2 * (fyl[i] * r_mass - ps * b);
            c = c + deltby2 * (fzl[i] * r_mass - ps * c);
        }
        m2 = m2 + (a * a + b * b + c * c) / r_mass;
        a = a + deltsby2 * (fxs[i] * r_mass - pss * a);
        b = b + deltsby2 * (fys[i] * r_mass - pss * b);
        c = c + deltsby2 * (fzs[i] * r_mass - pss * c);
        vx[i] = a;
        vy[i] = b;
        vz[i] = c;
        x[i] = x[i] + delts * a;
        y[i] = y[i] + delts * b;
        z[i] = z[i] + delts * c;
    }
    const double sumv2 = wg_sum(m2, sh);
    if (threadIdx.x == 0) {
        S->sss = S->sss + pss * delts + (sumv2 - S->ggs * S->treq) * delts_sqby2 / S->qqs;
        S->pss = pss + (sumv2 - S->ggs * S->treq) * deltsby2 / S->qqs;
    }
}

// the implicit half kick v_new = v + dtby2 (f/m - p_new v_new) with the
// thermostat equation for p_new, solved by the reference's Newton
// iteration; vn: (3, N) scratch for the iterate
template <bool OUTER>
__global__ void __launch_bounds__(NVT_T) k_nvt_solve(
    const RepDev* __restrict__ reps, const SiteDev* __restrict__ sites,
    const int* __restrict__ atom_site, mspms_nvt_state* __restrict__ st,
    double* __restrict__ carry, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ fx, const double* __restrict__ fy,
    const double* __restrict__ fz, double* __restrict__ vn, size_t n, double dt, double dtby2,
    double dt_sqby2, double* __restrict__ ukin, const int* __restrict__ active) {
#pragma clang fp contract(off)
    __shared__ double sh[16];
    if (active && !active[blockIdx.x]) return;  // a frozen replica (whole workgroup)
    const int r = blockIdx.x;
    const RepDev& R = reps[r];
    mspms_nvt_state* S = st + r;
    const double treq = S->treq;
    const double q = OUTER ? S->qq : S->qqs, g = OUTER ? S->gg : S->ggs;
    double p;
    double ss = S->ss;
    if (OUTER) {
        // src/nvtnh.c:161-162, with the sum m v^2 the last inner solve left
        const double s2 = carry[r];
        ss = ss + S->ps * dt + (s2 - S->gg * treq) * dt_sqby2 / S->qq;
        p = S->ps + (s2 - S->gg * treq) * dtby2 / S->qq;
    } else {
        p = S->pss;
    }
    double* vnx = vn;
    double* vny = vn + n;
    double* vnz = vn + 2 * n;
    const int a0 = R.atom0, a1 = R.atom0 + R.natom;
    double m2 = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += NVT_T) {
        const double r_mass = rmass(sites, atom_site, i);
        const double a = vx[i], b = vy[i], c = vz[i];
        vnx[i] = a;
        vny[i] = b;
        vnz[i] = c;
        m2 = m2 + (a * a + b * b + c * c) / r_mass;
    }
    double sumv2 = wg_sum(m2, sh);
    double pn = p;
    bool ready = false;
    for (int iter = 0; !ready && iter < 100; iter++) {
        const double po = pn;
        double d = 0.0;
        for (int i = a0 + threadIdx.x; i < a1; i += NVT_T) {
            const double r_mass = rmass(sites, atom_site, i);
            const double ox = vnx[i], oy = vny[i], oz = vnz[i];
            const double bx = -dtby2 * (fx[i] * r_mass - po * ox) - (vx[i] - ox);
            const double by = -dtby2 * (fy[i] * r_mass - po * oy) - (vy[i] - oy);
            const double bz = -dtby2 * (fz[i] * r_mass - po * oz) - (vz[i] - oz);
            d = d + (ox * dt / q / r_mass) * bx;
            d = d + (oy * dt / q / r_mass) * by;
            d = d + (oz * dt / q / r_mass) * bz;
        }
        double delp = wg_sum(d, sh);
        const double di = -(po * dtby2 + 1.0);
        delp = delp - di * ((-sumv2 + g * treq) * dtby2 / q - (p - po));
        delp = delp / (-dt * dtby2 * sumv2 / q + di);
        m2 = 0.0;
        for (int i = a0 + threadIdx.x; i < a1; i += NVT_T) {
            const double r_mass = rmass(sites, atom_site, i);
            const double ox = vnx[i], oy = vny[i], oz = vnz[i];
            const double bx = -dtby2 * (fx[i] * r_mass - po * ox) - (vx[i] - ox);
            const double by = -dtby2 * (fy[i] * r_mass - po * oy) - (vy[i] - oy);
            const double bz = -dtby2 * (fz[i] * r_mass - po * oz) - (vz[i] - oz);
            const double nx = ox + (bx + dtby2 * ox * delp) / di;
            const double ny = oy + (by + dtby2 * oy * delp) / di;
            const double nz = oz + (bz + dtby2 * oz * delp) / di;
            vnx[i] = nx;
            vny[i] = ny;
            vnz[i] = nz;
            m2 = m2 + (nx * nx + ny * ny + nz * nz) / r_mass;
        }
        sumv2 = wg_sum(m2, sh);
        pn = po + delp;
        ready = !(fabs((pn - po) / pn) > NVT_ERR);
    }
    for (int i = a0 + threadIdx.x; i < a1; i += NVT_T) {
        vx[i] = vnx[i];
        vy[i] = vny[i];
        vz[i] = vnz[i];
    }
    if (threadIdx.x != 0) return;
    carry[r] = sumv2;
    if (OUTER) {
        S->ss = ss;
        S->ps = pn;
        const double uk = sumv2 / 2.0;
        S->ukin = uk;
        S->unhts = (pn * pn * S->qq) / 2.0 + S->gg * treq * ss;
        S->tinst = 2.0 * uk / S->nfree;
        if (ukin) ukin[r] = uk;
    } else {
        S->pss = pn;
        S->unhtss = (pn * pn * S->qqs) / 2.0 + S->ggs * treq * S->sss;
    }
}

}  // namespace

void launch_nvt_drift(const LaunchCtx& c, mspms_nvt_state* st, bool first, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz, const double* fl,
                      const double* fs, double deltby2, double delts, double deltsby2,
                      double delts_sqby2, const int* active) {
    const size_t n = (size_t)c.natom;
    if (first)
        hipLaunchKernelGGL(k_nvt_drift<true>, dim3(c.nrep), dim3(NVT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, x, y, z, vx, vy, vz, fl, fl + n, fl + 2 * n,
                           fs, fs + n, fs + 2 * n, deltby2, delts, deltsby2, delts_sqby2, active);
    else
        hipLaunchKernelGGL(k_nvt_drift<false>, dim3(c.nrep), dim3(NVT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, x, y, z, vx, vy, vz, fl, fl + n, fl + 2 * n,
                           fs, fs + n, fs + 2 * n, deltby2, delts, deltsby2, delts_sqby2, active);
}

void launch_nvt_solve(const LaunchCtx& c, mspms_nvt_state* st, bool outer, double* carry,
                      double* vx, double* vy, double* vz, const double* f, double* vn,
                      double dt, double dtby2, double dt_sqby2, double* ukin,
                      const int* active) {
    const size_t n = (size_t)c.natom;
    if (outer)
        hipLaunchKernelGGL(k_nvt_solve<true>, dim3(c.nrep), dim3(NVT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, carry, vx, vy, vz, f, f + n, f + 2 * n, vn, n,
                           dt, dtby2, dt_sqby2, ukin, active);
    else
        hipLaunchKernelGGL(k_nvt_solve<false>, dim3(c.nrep), dim3(NVT_T), 0, c.stream, c.reps,
                           c.sites, c.atom_site, st, carry, vx, vy, vz, f, f + n, f + 2 * n, vn, n,
                           dt, dtby2, dt_sqby2, ukin, active);
}
