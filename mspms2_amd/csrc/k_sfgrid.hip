// k_sfgrid.hip -- solid-fluid grid generation on gfx950: the reference's
// offline generator tools/sfgridgen.c:71-352 (energy, gradient and force
// Jacobian -- minus the Hessian -- of the LJ potential of a periodic solid
// at every node of a grid over one unit cell).
//
// One thread per node; the solid atoms sit in LDS.  Every node sums its
// (z image, atom) pairs in the generator's loop order in FP64 with the
// generator's expressions (no contraction), so the tables are the
// generator's bit for bit.  With the LJ switch on, a pair beyond rcuton
// adds the Jacobian terms left by the node's previous pair, as the
// generator's empty else branch does (tools/sfgridgen.c:254-256); they start
// at zero here (the generator reads uninitialised locals if the first pair of
// a node lies in the switching shell).
#include <hip/hip_runtime.h>
#include "sfgrid.h"

namespace {

constexpr int SFG_T = 256;
constexpr double SFG_RGAS = 8.314472;  // tools/sfgridgen.c:27

__global__ void __launch_bounds__(SFG_T) k_sfgridgen(SfgParams P, const double* __restrict__ sx,
                                                     const double* __restrict__ sy,
                                                     const double* __restrict__ sz,
                                                     double* __restrict__ out, size_t ntot) {
#pragma clang fp contract(off)
    __shared__ double lx[SFG_MAX_SOLID], ly[SFG_MAX_SOLID], lz[SFG_MAX_SOLID];
    for (int k = threadIdx.x; k < P.nsolid; k += SFG_T) {
        lx[k] = sx[k];
        ly[k] = sy[k];
        lz[k] = sz[k];
    }
    __syncthreads();
    const size_t index = (size_t)blockIdx.x * SFG_T + threadIdx.x;
    if (index >= ntot) return;
    // index = ii * ny * nz + jj * nz + kk (tools/sfgridgen.c:285)
    const int kk = (int)(index % P.nz);
    const int jj = (int)((index / P.nz) % P.ny);
    const int ii = (int)(index / ((size_t)P.nz * P.ny));
    const double fxx = P.x0 + P.itvl * ii;
    const double fyy = P.y0 + P.itvl * jj;
    const double fzz = P.z0 + P.itvl * kk;
    const double epsilonsf = P.epsilonsf, sig2 = P.sigmasf * P.sigmasf;
    double usf_vdw = 0.0, Dusfx = 0.0, Dusfy = 0.0, Dusfz = 0.0;
    double Dfxx = 0.0, Dfxy = 0.0, Dfxz = 0.0, Dfyx = 0.0, Dfyy = 0.0, Dfyz = 0.0;
    double Dfzx = 0.0, Dfzy = 0.0, Dfzz = 0.0;
    double dfxx = 0.0, dfxy = 0.0, dfxz = 0.0, dfyx = 0.0, dfyy = 0.0, dfyz = 0.0;
    double dfzx = 0.0, dfzy = 0.0, dfzz = 0.0;
    double factorLJswitch = 1.0;
    for (int pp = -P.nuc_z; pp <= P.nuc_z; pp++) {
        for (int nn = 0; nn < P.nsolid; nn++) {
            const double rxsf = fxx - lx[nn];
            const double rysf = fyy - ly[nn];
            double rzsf = fzz - lz[nn];
            rzsf += pp * P.uclz;
            const double rsfsq = rxsf * rxsf + rysf * rysf + rzsf * rzsf;
            if (!(rsfsq < P.rcutoffsq)) continue;
            if (P.lj_switch) {
                if (rsfsq < P.rcutonsq)
                    factorLJswitch = 1.0;
                else
                    factorLJswitch = (P.rcutoffsq - rsfsq) * (P.rcutoffsq - rsfsq) *
                                     (P.rcutoffsq + 2.0 * rsfsq - 3.0 * P.rcutonsq) /
                                     P.roff2_minus_ron2_cube;
            }
            const double r_rsfsq = sig2 / rsfsq;
            const double r_r6 = r_rsfsq * r_rsfsq * r_rsfsq;
            const double r_r12 = r_r6 * r_r6;
            const double r_r12_minus_r_r6 = r_r12 - r_r6;
            const double usf_vdw_temp = epsilonsf * r_r12_minus_r_r6;
            if (P.lj_switch)
                usf_vdw += usf_vdw_temp * factorLJswitch;
            else
                usf_vdw += usf_vdw_temp;
            double fsf;
            if (P.lj_switch && !(rsfsq < P.rcutonsq))
                fsf = 24.0 * epsilonsf * (r_r12_minus_r_r6 + r_r12) / rsfsq * factorLJswitch -
                      4.0 * usf_vdw_temp * 12.0 * (P.rcutoffsq - rsfsq) * (P.rcutonsq - rsfsq) /
                          P.roff2_minus_ron2_cube;
            else
                fsf = 24.0 * epsilonsf * (r_r12_minus_r_r6 + r_r12) / rsfsq;
            Dusfx -= fsf * rxsf;
            Dusfy -= fsf * rysf;
            Dusfz -= fsf * rzsf;
            if (!P.lj_switch || rsfsq < P.rcutonsq) {
                const double ddf =
                    96.0 * epsilonsf * (r_r12_minus_r_r6 + 6.0 * r_r12 - r_r6) / rsfsq / rsfsq;
                dfxx = -rxsf * ddf * rxsf + fsf;
                dfxy = -rxsf * ddf * rysf;
                dfxz = -rxsf * ddf * rzsf;
                dfyx = -rysf * ddf * rxsf;
                dfyy = -rysf * ddf * rysf + fsf;
                dfyz = -rysf * ddf * rzsf;
                dfzx = -rzsf * ddf * rxsf;
                dfzy = -rzsf * ddf * rysf;
                dfzz = -rzsf * ddf * rzsf + fsf;
            }
            Dfxx += dfxx;
            Dfxy += dfxy;
            Dfxz += dfxz;
            Dfyx += dfyx;
            Dfyy += dfyy;
            Dfyz += dfyz;
            Dfzx += dfzx;
            Dfzy += dfzy;
            Dfzz += dfzz;
        }
    }
    usf_vdw *= 4.0;
    const double v[13] = {usf_vdw, Dusfx, Dusfy, Dusfz, Dfxx, Dfxy, Dfxz,
                          Dfyx,    Dfyy,  Dfyz,  Dfzx,  Dfzy, Dfzz};
    for (int f = 0; f < 13; f++) out[(size_t)f * ntot + index] = v[f];
}

}  // namespace

extern "C" int mspms_sfgridgen_dims(const mspms_sfgridgen_spec* s, int* nx, int* ny, int* nz) {
    if (!s || !nx || !ny || !nz || !(s->grid_itvl > 0.0)) return MSPMS_EINVAL;
    // tools/sfgridgen.c:129-131
    *nx = (int)floor(s->uclx / s->grid_itvl) + 2;
    *ny = (int)floor(s->ucly / s->grid_itvl) + 2;
    *nz = (int)floor(s->uclz / s->grid_itvl) + 2;
    return 0;
}


// The generator constants of tools/sfgridgen.c:121-155 from a spec; 0 or an
// MSPMS_* error (recorded).
int sfgridgen_setup(const mspms_sfgridgen_spec* s, SfgParams* P, size_t* ntot) {
    if (!s || !s->sx || !s->sy || !s->sz) return mspms_fail(MSPMS_EINVAL, "null argument");
    if (s->s_natom < 1 || s->s_natom > SFG_MAX_SOLID)
        return mspms_fail(MSPMS_EINVAL, "1..1000 solid atoms (solid_atom_max)");
    if (!(s->rcutoff > 0.0) || !(s->uclz > 0.0))
        return mspms_fail(MSPMS_EINVAL, "cutoff and unit cell must be positive");
    int nx, ny, nz;
    if (mspms_sfgridgen_dims(s, &nx, &ny, &nz))
        return mspms_fail(MSPMS_EINVAL, "grid interval must be positive");
    P->rcutonsq = s->rcuton * s->rcuton;
    P->rcutoffsq = s->rcutoff * s->rcutoff;
    P->roff2_minus_ron2_cube = (P->rcutoffsq - P->rcutonsq) * (P->rcutoffsq - P->rcutonsq) *
                               (P->rcutoffsq - P->rcutonsq);
    P->sigmasf = (s->f_sigma + s->s_sigma) / 2.0;
    P->epsilonsf = sqrt(s->f_epsilon * s->s_epsilon) * SFG_RGAS;
    int nuc_z = 1;
    while (nuc_z * s->uclz < s->rcutoff) nuc_z++;
    const double xcenter = s->uclx / 2.0, ycenter = s->ucly / 2.0;
    P->x0 = xcenter - s->uclx / 2.0;
    P->y0 = ycenter - s->ucly / 2.0;
    P->z0 = 0.0;
    P->itvl = s->grid_itvl;
    P->uclz = s->uclz;
    P->nx = nx;
    P->ny = ny;
    P->nz = nz;
    P->nuc_z = nuc_z;
    P->nsolid = s->s_natom;
    P->lj_switch = s->isLJswitch ? 1 : 0;
    *ntot = (size_t)nx * ny * nz;
    return 0;
}

hipError_t sfgridgen_launch(const SfgParams& P, const double* sx, const double* sy,
                            const double* sz, double* out, size_t ntot, hipStream_t st) {
    hipLaunchKernelGGL(k_sfgridgen, dim3((unsigned)((ntot + SFG_T - 1) / SFG_T)), dim3(SFG_T), 0,
                       st, P, sx, sy, sz, out, ntot);
    return hipGetLastError();
}

extern "C" int mspms_sfgridgen(int device, const mspms_sfgridgen_spec* s, double* out) {
    SfgParams P;
    size_t ntot = 0;
    if (!out) return mspms_fail(MSPMS_EINVAL, "null argument");
    if (int e = sfgridgen_setup(s, &P, &ntot)) return e;
    if (hipSetDevice(device) != hipSuccess) return mspms_fail(MSPMS_EIO, "hipSetDevice");
    double *d_s = nullptr, *d_out = nullptr;
    const size_t ns = (size_t)s->s_natom;
    int rc = 0;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return mspms_fail(MSPMS_EIO, "stream");
    if (hipMalloc(&d_s, 3 * ns * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_out, 13 * ntot * sizeof(double)) != hipSuccess) {
        rc = mspms_fail(MSPMS_ENOMEM, "grid tables");
    } else if (hipMemcpyAsync(d_s, s->sx, ns * sizeof(double), hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(d_s + ns, s->sy, ns * sizeof(double), hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(d_s + 2 * ns, s->sz, ns * sizeof(double), hipMemcpyHostToDevice,
                              st)) {
        rc = mspms_fail(MSPMS_EIO, "upload");
    } else {
        if (sfgridgen_launch(P, d_s, d_s + ns, d_s + 2 * ns, d_out, ntot, st) != hipSuccess ||
            hipMemcpyAsync(out, d_out, 13 * ntot * sizeof(double), hipMemcpyDeviceToHost, st) ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = mspms_fail(MSPMS_EIO, "grid kernel");
    }
    if (d_s) (void)hipFree(d_s);
    if (d_out) (void)hipFree(d_out);
    (void)hipStreamDestroy(st);
    return rc;
}

