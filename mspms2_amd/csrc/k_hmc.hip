// k_hmc.hip -- batched hybrid Monte Carlo moves on gfx950, one independent
// chain per replica of the table.  Reference: hmc() (src/hmc.c:23-70),
// md_move (src/dispmd.c:12-103), velinit (src/initvars.c:151-232), the
// RANMAR generator and gaussran (src/random.c:14-138).
//
// Each replica owns a RANMAR stream (the state of Marsaglia and Zaman's
// lagged-Fibonacci generator combined with an arithmetic sequence, in
// James' formulation: 97 lags u[1..97], c/cd/cm, two cursors) and the
// cached second deviate of the polar Gaussian method.  The values are
// multiples of 2^-24, exact in FP64, so the stream is the reference's bit
// for bit.  One HMC move of every replica:
//   k_hmc_begin   draw the move-type number (hmc() draws one per move); on
//                 every move but the first, fresh Maxwell velocities drawn
//                 atom by atom, momentum removed, rescaled to treq exactly
//                 (velinit); save the positions
//   frclong + frcshort of the current positions (all replicas)
//   k_hmc_mark    H_old = uinter, uintra, ukin, ushift
//   nmd RESPA steps (vver)
//   k_hmc_end     dH = [(uinter - old) + (uintra - old) + (ukin - old) +
//                 (ushift - old)] / treq; accept if dH <= 0 or a uniform
//                 draw < exp(-dH); otherwise restore the positions; with
//                 per-chain steps, the equilibration-phase delt adjustment
// The random draws are sequential per replica (one lane); momentum,
// kinetic energy and the position save / restore use the whole workgroup.
#include "engine.h"
#include "../../include/mspms_gpu.h"

namespace {

__device__ inline double ranmar_next(RanmarDev& r) {
#pragma clang fp contract(off)
    double uni = r.u[r.i97] - r.u[r.j97];
    if (uni < 0.0) uni += 1.0;
    r.u[r.i97] = uni;
    if (--r.i97 == 0) r.i97 = 97;
    if (--r.j97 == 0) r.j97 = 97;
    r.c -= r.cd;
    if (r.c < 0.0) r.c += r.cm;
    uni -= r.c;
    if (uni < 0.0) uni += 1.0;
    return uni;
}

// polar method, the second deviate of each accepted pair kept for the next
// call (gaussran, src/random.c:113-138)
__device__ inline double gaussran(RanmarDev& r) {
#pragma clang fp contract(off)
    if (r.iset) {
        r.iset = 0;
        return r.gset;
    }
    double v1, v2, rsq;
    do {
        const double a = ranmar_next(r), b = ranmar_next(r);
        v1 = 2.0 * a - 1.0;
        v2 = 2.0 * b - 1.0;
        rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    const double fac = sqrt(-2.0 * log(rsq) / rsq);
    r.gset = v1 * fac;
    r.iset = 1;
    return v2 * fac;
}

template <int NV>
__device__ void block_sum_hmc(double (&v)[NV], double* sh, double* out) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = v[k];
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) sh[wid * NV + k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += sh[w * NV + k];
            out[k] = s;
        }
    __syncthreads();
}

// velinit (src/initvars.c:151-232) for one replica: lane 0 draws, the
// workgroup reduces and rescales
__device__ void velinit_rep(const RepDev& R, HmcRep& H, RanmarDev& rng, const double* aw_atom,
                            double* vx, double* vy, double* vz, double* sh, double* red) {
#pragma clang fp contract(off)
    if (threadIdx.x == 0) {
        const double stdvtmp = sqrt(H.treq);
        for (int i = R.atom0; i < R.atom0 + R.natom; i++) {
            const double stdv = stdvtmp / sqrt(aw_atom[i]);
            vx[i] = stdv * gaussran(rng);
            vy[i] = stdv * gaussran(rng);
            vz[i] = stdv * gaussran(rng);
        }
    }
    __syncthreads();
    double p[3] = {0.0, 0.0, 0.0};
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += blockDim.x) {
        p[0] += vx[i] * aw_atom[i];
        p[1] += vy[i] * aw_atom[i];
        p[2] += vz[i] * aw_atom[i];
    }
    block_sum_hmc<3>(p, sh, red);
    const double px = red[0] / H.system_mass, py = red[1] / H.system_mass,
                 pz = red[2] / H.system_mass;
    double k[1] = {0.0};
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += blockDim.x) {
        vx[i] -= px;
        vy[i] -= py;
        vz[i] -= pz;
        k[0] += aw_atom[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
    block_sum_hmc<1>(k, sh, red);
    const double tinst = 2.0 * (0.5 * red[0]) / H.nfree;
    const double scaling = sqrt(H.treq / tinst);
    k[0] = 0.0;
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += blockDim.x) {
        vx[i] *= scaling;
        vy[i] *= scaling;
        vz[i] *= scaling;
        k[0] += aw_atom[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
    block_sum_hmc<1>(k, sh, red);
    if (threadIdx.x == 0) H.ukin = 0.5 * red[0];
}

// start-up: velinit of every replica (the reference's init_vars draws the
// initial velocities from the freshly seeded stream)
__global__ void __launch_bounds__(256) k_hmc_init(const RepDev* __restrict__ reps,
                                                  HmcRep* __restrict__ hmc,
                                                  RanmarDev* __restrict__ rng,
                                                  const double* __restrict__ aw_atom,
                                                  double* vx, double* vy, double* vz,
                                                  const int* __restrict__ active) {
    __shared__ double sh[4 * 4];
    __shared__ double red[4];
    __shared__ RanmarDev r;
    const int rep = blockIdx.x;
    if (active && !active[rep]) return;  // whole workgroup
    if (threadIdx.x == 0) r = rng[rep];
    __syncthreads();
    velinit_rep(reps[rep], hmc[rep], r, aw_atom, vx, vy, vz, sh, red);
    if (threadIdx.x == 0) {
        rng[rep] = r;
        hmc[rep].first = 1;
    }
}

__global__ void __launch_bounds__(256) k_hmc_begin(const RepDev* __restrict__ reps,
                                                   HmcRep* __restrict__ hmc,
                                                   RanmarDev* __restrict__ rng,
                                                   const double* __restrict__ aw_atom,
                                                   const double* __restrict__ pos, double* pos_old,
                                                   double* vx, double* vy, double* vz,
                                                   int natom_total) {
    __shared__ double sh[4 * 4];
    __shared__ double red[4];
    __shared__ RanmarDev r;
    const int rep = blockIdx.x;
    const RepDev& R = reps[rep];
    HmcRep& H = hmc[rep];
    const bool first = H.first != 0;  // uniform
    if (threadIdx.x == 0) {
        r = rng[rep];
        (void)ranmar_next(r);  // the move-type draw of hmc() (src/hmc.c:25)
    }
    __syncthreads();
    if (!first) velinit_rep(R, H, r, aw_atom, vx, vy, vz, sh, red);
    const size_t n = (size_t)natom_total;
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += blockDim.x)
        for (int d = 0; d < 3; d++) pos_old[d * n + i] = pos[d * n + i];
    if (threadIdx.x == 0) {
        rng[rep] = r;
        H.first = 0;
    }
}

// H_old after the forces of the move's starting positions
__global__ void k_hmc_mark(int nrep, HmcRep* __restrict__ hmc,
                           const mspms_long_terms* __restrict__ lt,
                           const mspms_short_terms* __restrict__ st) {
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= nrep) return;
    HmcRep& H = hmc[rep];
    // with the wall split into frcshort, its energy belongs to uinter
    H.uinter_old = lt[rep].uinter + st[rep].usflj;
    H.uintra_old = st[rep].uintra;
    H.ushift_old = lt[rep].ushift;
    H.ukin_old = H.ukin;
}

__global__ void __launch_bounds__(256) k_hmc_end(const RepDev* __restrict__ reps,
                                                 HmcRep* __restrict__ hmc,
                                                 RanmarDev* __restrict__ rng,
                                                 const mspms_long_terms* __restrict__ lt,
                                                 const mspms_short_terms* __restrict__ st,
                                                 const double* __restrict__ ukin, double* pos,
                                                 const double* __restrict__ pos_old,
                                                 int natom_total, double* rep_delt) {
#pragma clang fp contract(off)
    __shared__ int accept;
    const int rep = blockIdx.x;
    const RepDev& R = reps[rep];
    HmcRep& H = hmc[rep];
    if (threadIdx.x == 0) {
        // src/dispmd.c:49-72
        const double du = (lt[rep].uinter - H.uinter_old) + (st[rep].uintra - H.uintra_old) +
                          (ukin[rep] - H.ukin_old) + (lt[rep].ushift - H.ushift_old);
        const double dH = du / H.treq;
        bool acc = dH <= 0.0;
        if (!acc) {
            RanmarDev r = rng[rep];
            acc = ranmar_next(r) < exp(-dH);
            rng[rep] = r;
        }
        H.moves++;
        if (acc) {
            H.accepted++;
            H.ukin = ukin[rep];
        } else {
            H.ukin = H.ukin_old;
        }
        H.last_dH = dH;
        accept = acc ? 1 : 0;
        // hmc()'s delt adjustment (src/hmc.c:47-58): md_move counted the
        // move (counts[20], counts[21], src/dispmd.c:55,74); once a cycle
        // is full, delt *= 1 - rreq_disp + accepted / cycle
        if (rep_delt && H.adj_cycle > 0) {
            H.adj_moves++;
            if (acc) H.adj_accepted++;
            if (H.adj_moves == H.adj_cycle) {
                const double rinst = H.adj_accepted * 1.0 / H.adj_cycle;
                rep_delt[rep] = rep_delt[rep] * (1.0 - H.rreq + rinst);
                H.adj_moves = 0;
                H.adj_accepted = 0;
            }
        }
    }
    __syncthreads();
    if (accept) return;
    const size_t n = (size_t)natom_total;
    for (int i = R.atom0 + threadIdx.x; i < R.atom0 + R.natom; i += blockDim.x)
        for (int d = 0; d < 3; d++) pos[d * n + i] = pos_old[d * n + i];
}

// n draws of one replica's stream (test access to the generator)
__global__ void k_hmc_draw(RanmarDev* rng, int rep, int n, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    RanmarDev r = rng[rep];
    for (int k = 0; k < n; k++) out[k] = ranmar_next(r);
    rng[rep] = r;
}

}  // namespace

void launch_hmc_init(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const double* aw_atom,
                     double* vel, const int* active) {
    const size_t n = (size_t)c.natom;
    hipLaunchKernelGGL(k_hmc_init, dim3(c.nrep), dim3(256), 0, c.stream, c.reps, hmc, rng,
                       aw_atom, vel, vel + n, vel + 2 * n, active);
}

void launch_hmc_begin(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const double* aw_atom,
                      const double* pos, double* pos_old, double* vel) {
    const size_t n = (size_t)c.natom;
    hipLaunchKernelGGL(k_hmc_begin, dim3(c.nrep), dim3(256), 0, c.stream, c.reps, hmc, rng,
                       aw_atom, pos, pos_old, vel, vel + n, vel + 2 * n, c.natom);
}

void launch_hmc_mark(const LaunchCtx& c, HmcRep* hmc, const void* lterms, const void* sterms) {
    hipLaunchKernelGGL(k_hmc_mark, dim3((c.nrep + 63) / 64), dim3(64), 0, c.stream, c.nrep, hmc,
                       (const mspms_long_terms*)lterms, (const mspms_short_terms*)sterms);
}

void launch_hmc_end(const LaunchCtx& c, HmcRep* hmc, RanmarDev* rng, const void* lterms,
                    const void* sterms, const double* ukin, double* pos,
                    const double* pos_old, double* rep_delt) {
    hipLaunchKernelGGL(k_hmc_end, dim3(c.nrep), dim3(256), 0, c.stream, c.reps, hmc, rng,
                       (const mspms_long_terms*)lterms, (const mspms_short_terms*)sterms, ukin,
                       pos, pos_old, c.natom, rep_delt);
}

void launch_hmc_draw(const LaunchCtx& c, RanmarDev* rng, int rep, int n, double* out) {
    hipLaunchKernelGGL(k_hmc_draw, dim3(1), dim3(64), 0, c.stream, rng, rep, n, out);
}
