// k_sf.hip -- solid-fluid terms on gfx950 (src/frclong.c:1145-1209): one
// thread per site; only GHOST_NONE sites interact with the solid.  The
// per-site models (hypergeometric tube, explicit wall atoms, tricubic
// Hermite grid, myinterp cube polynomials) are sfeval::sf_site (sf_eval.h).
// Forces are added to the caller's force arrays (the site's own thread is
// the only writer); the energy goes to a per-site partial for the finalize.
// k_grid_node_major turns the Hermite grid of the C ABI into the device's
// node-major records.
#include "engine.h"
#include "sf_eval.h"

namespace {

__global__ void __launch_bounds__(256) k_sf(int natom, const int* __restrict__ atom_site,
                                            const int* __restrict__ atom_mol,
                                            const MolDev* __restrict__ mols,
                                            const SiteDev* __restrict__ sites,
                                            const RepDev* __restrict__ reps, SFDev sf,
                                            const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ z, double* fx,
                                            double* fy, double* fz, double* __restrict__ usf,
                                            Status* st) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natom) return;
    const SiteDev S = sites[atom_site[i]];
    if (S.ghost != MSPMS_GHOST_NONE) {
        usf[i] = 0.0;
        return;
    }
    const RepDev& R = reps[mols[atom_mol[i]].rep];
    double fxi = fx[i], fyi = fy[i], fzi = fz[i];
    int edom = 0;
    usf[i] = sfeval::sf_site(sf, R, S, x[i], y[i], z[i], fxi, fyi, fzi, &edom);
    if (edom) atomicOr(&st->flags, ST_GRID_EDOM);
    fx[i] = fxi;
    fy[i] = fyi;
    fz[i] = fzi;
}

// The Hermite grid from the C-ABI layout [type][field][node] (fields F, Fx,
// Fy, Fz, Fxy, Fxz, Fyz, Fxyz) to the device layout [type][node][o], o the
// derivative bits (x, y, z): F, Fx, Fy, Fxy, Fz, Fxz, Fyz, Fxyz, one 32-B
// record per node, so that a lookup reads 8 records instead of 64 scattered
// floats.
__global__ void __launch_bounds__(256) k_grid_node_major(const float* __restrict__ src,
                                                         float* __restrict__ dst, size_t nn,
                                                         int ntype) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn * (size_t)ntype) return;
    const size_t t = k / nn, node = k % nn;
    const int field_of[8] = {0, 1, 2, 4, 3, 5, 6, 7};
    float v[8];
#pragma unroll
    for (int o = 0; o < 8; o++) v[o] = src[(t * 8 + field_of[o]) * nn + node];
    float4* d = reinterpret_cast<float4*>(dst) + 2 * k;
    d[0] = make_float4(v[0], v[1], v[2], v[3]);
    d[1] = make_float4(v[4], v[5], v[6], v[7]);
}

}  // namespace

void launch_sf(const LaunchCtx& c, const SFDev& sf, const double* x, const double* y,
               const double* z, double* fx, double* fy, double* fz, double* usf) {
    int g = (c.natom + 255) / 256;
    hipLaunchKernelGGL(k_sf, dim3(g), dim3(256), 0, c.stream, c.natom, c.atom_site, c.atom_mol,
                       c.mols, c.sites, c.reps, sf, x, y, z, fx, fy, fz, usf, c.status);
}

void launch_grid_node_major(hipStream_t s, const float* src, float* dst, size_t nn, int ntype) {
    const size_t n = nn * (size_t)ntype;
    hipLaunchKernelGGL(k_grid_node_major, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src,
                       dst, nn, ntype);
}
