// k_recip.hip -- Ewald reciprocal space (FP64) on gfx950.
// Reference: src/frclong.c:988-1048.  For every k = 2*pi*(kx/Lx, ky/Ly,
// kz/Lz) with 0 < kx^2+ky^2+kz^2 <= KSQMAX:
//   S(k) = sum_i q_i exp(i k.r_i),  U_k = kvec |S(k)|^2,
//   F_i += 2 q_i (S_r sin(k.r_i) - S_i cos(k.r_i)) Vf kvec C k.
// The reference evaluates cos/sin of k.r per (k, atom), twice.  Here
// exp(i k.r) = ex^kx ey^ky ez^kz with e = exp(i 2 pi x / L): three sincos
// per atom, then complex multiplies (|k| <= KMAX steps, error ~KMAX ulp).
//
//   k_sk     one workgroup per (replica, atom chunk of <= 64); a lane owns a
//            segment of SEG consecutive kz of one (kx, ky) row and
//            accumulates their S in registers over the chunk; per-atom
//            power tables live in LDS.  Partial S per chunk is written out
//            (no atomics).
//   k_sred   per k: deterministic sum over chunks, U_k, and the force
//            coefficients A_k, B_k = 2 Vf C kvec (S_r, S_i).
//   k_kforce one lane per atom over a range of rows; exp(i k.r) by
//            recurrence in registers; forces added with FP64 atomics
//            (one per atom per row range).
#include "engine.h"
#include "../../include/mspms_gpu.h"

namespace {

struct cplx {
    double re, im;
};
__device__ inline cplx cmul(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ inline cplx conjc(cplx a) { return {a.re, -a.im}; }
__device__ inline cplx cpow(cplx b, int n) {  // b^n, n >= 0
    cplx r = {1.0, 0.0};
    while (n > 0) {
        if (n & 1) r = cmul(r, b);
        b = cmul(b, b);
        n >>= 1;
    }
    return r;
}
__device__ inline cplx cpow_signed(cplx b, int n) {
    cplx r = cpow(b, n < 0 ? -n : n);
    return n < 0 ? conjc(r) : r;
}

// tab layout: [atom][n], n over x powers 0..kmx, y 0..kmy, z 0..kmz.
// A lane takes item (row r, segment g) with r fastest, so that a wave's
// lanes share a segment index over rows sorted by length; the segment
// starts at kz0 = kz_lo + g SEG, whose power comes from the z table.  A
// lane's serial chain is (SEG + 2 complex products) per atom, against
// (row length + 2) with one lane per row: there are only ~100 rows at
// kmax 8, so segments are what fills the waves.
template <int SEG>
__global__ void __launch_bounds__(1024) k_sk(const int4* __restrict__ work,
                                            const RepDev* __restrict__ reps,
                                            const KRow* __restrict__ rows_s,
                                            const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ z,
                                            const int* __restrict__ atom_site,
                                            const SiteDev* __restrict__ sites,
                                            double2* __restrict__ s_part, int nk_stride,
                                            double2* __restrict__ eatom, int natom) {
    extern __shared__ double2 tab[];
    __shared__ double qsh[64];
    const int4 w = work[blockIdx.x];
    const RepDev& R = reps[w.x];
    const int chunk = w.y, a0 = w.z, na = w.w;
    const int kmx = R.kmax[0], kmy = R.kmax[1], kmz = R.kmax[2];
    const int ntab = kmx + kmy + (kmz > 0 ? kmz : 1) + 3;  // z table keeps exp(i t)
    // tables: thread t -> (atom t&63, dim t>>6)
    for (int t = threadIdx.x; t < 3 * 64; t += blockDim.x) {
        const int a = t & 63, d = t >> 6;
        const int km = d == 0 ? kmx : (d == 1 ? kmy : (kmz > 0 ? kmz : 1));
        const int off = d == 0 ? 0 : (d == 1 ? kmx + 1 : kmx + kmy + 2);
        double2* row = tab + a * ntab + off;
        if (a < na) {
            const double* p = d == 0 ? x : (d == 1 ? y : z);
            double s, c;
            sincos(R.twopi_l[d] * p[a0 + a], &s, &c);
            eatom[(size_t)d * natom + a0 + a] = make_double2(c, s);  // for k_kforce
            cplx e = {c, s}, v = {1.0, 0.0};
            for (int n = 0; n <= km; n++) {
                row[n] = make_double2(v.re, v.im);
                v = cmul(v, e);
            }
            if (d == 0) qsh[a] = sites[atom_site[a0 + a]].charge;
        } else {
            for (int n = 0; n <= km; n++) row[n] = make_double2(0.0, 0.0);
            if (d == 0) qsh[a] = 0.0;
        }
    }
    __syncthreads();
    const int nseg = (2 * kmz + 1 + SEG - 1) / SEG;  // a row holds at most 2 kmz + 1 kz
    for (int it = threadIdx.x; it < R.nkrow * nseg; it += blockDim.x) {
        const int r = it % R.nkrow, g = it / R.nkrow;
        const KRow K = rows_s[R.krow0 + r];
        const int nk = min(SEG, K.nk - g * SEG);
        if (nk <= 0) continue;
        const int kz0 = K.kz_lo + g * SEG;
        double sr[SEG], si[SEG];
#pragma unroll
        for (int t = 0; t < SEG; t++) sr[t] = si[t] = 0.0;
        const int ax = K.kx < 0 ? -K.kx : K.kx, ay = K.ky < 0 ? -K.ky : K.ky;
        const int alo = kz0 < 0 ? -kz0 : kz0;
        int a = 0;
        // four atoms at a time: four independent kz recurrences in flight
        for (; a + 4 <= na; a += 4) {
            cplx c4[4], s4[4];
            double q4[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double2* ta = tab + (a + u) * ntab;
                const double2 tx = ta[ax], ty = ta[kmx + 1 + ay], tz = ta[kmx + kmy + 2 + alo],
                              t1 = ta[kmx + kmy + 2 + 1];
                const cplx cx = {tx.x, K.kx < 0 ? -tx.y : tx.y};
                const cplx cy = {ty.x, K.ky < 0 ? -ty.y : ty.y};
                const cplx cz = {tz.x, kz0 < 0 ? -tz.y : tz.y};
                s4[u] = {t1.x, t1.y};
                c4[u] = cmul(cmul(cx, cy), cz);
                q4[u] = qsh[a + u];
            }
#pragma unroll
            for (int t = 0; t < SEG; t++) {
                if (t < nk) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        sr[t] += q4[u] * c4[u].re;
                        si[t] += q4[u] * c4[u].im;
                        c4[u] = cmul(c4[u], s4[u]);
                    }
                }
            }
        }
        for (; a < na; a++) {
            const double2* ta = tab + a * ntab;
            double2 tx = ta[ax], ty = ta[kmx + 1 + ay], tz = ta[kmx + kmy + 2 + alo],
                    t1 = ta[kmx + kmy + 2 + 1];
            cplx cx = {tx.x, K.kx < 0 ? -tx.y : tx.y};
            cplx cy = {ty.x, K.ky < 0 ? -ty.y : ty.y};
            cplx cz = {tz.x, kz0 < 0 ? -tz.y : tz.y};
            const cplx step = {t1.x, t1.y};
            cplx c = cmul(cmul(cx, cy), cz);
            const double q = qsh[a];
#pragma unroll
            for (int t = 0; t < SEG; t++) {
                if (t < nk) {
                    sr[t] += q * c.re;
                    si[t] += q * c.im;
                    c = cmul(c, step);
                }
            }
        }
        double2* out = s_part + (size_t)chunk * nk_stride + (K.k0 - R.k0) + g * SEG;
#pragma unroll
        for (int t = 0; t < SEG; t++)
            if (t < nk) out[t] = make_double2(sr[t], si[t]);
    }
}

// 32 lanes per k: lane j sums chunks j, j+32, ... and a fixed shuffle tree
// combines the 32 partials (deterministic); 8 k per 256-thread workgroup.
__global__ void __launch_bounds__(256) k_sred(int nk_total, const int* __restrict__ k_rep,
                                              const RepDev* __restrict__ reps,
                                              const int* __restrict__ rep_chunk0,
                                              const double2* __restrict__ s_part, int nk_stride,
                                              const double* __restrict__ kvec,
                                              double2* __restrict__ ak, double* __restrict__ uk) {
    const int k = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int j = threadIdx.x & 31;
    const bool live = k < nk_total;
    const int kk = live ? k : nk_total - 1;
    const int r = k_rep[kk];
    const RepDev& R = reps[r];
    const int kl = kk - R.k0;
    double sr = 0.0, si = 0.0;
    if (live)
        for (int c = rep_chunk0[r] + j; c < rep_chunk0[r + 1]; c += 32) {
            double2 v = s_part[(size_t)c * nk_stride + kl];
            sr += v.x;
            si += v.y;
        }
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) {
        sr += __shfl_xor(sr, o, 64);
        si += __shfl_xor(si, o, 64);
    }
    if (!live || j != 0) return;
    const double kv = kvec[k];
    uk[k] = kv * (sr * sr + si * si);
    const double s = 2.0 * R.vfactor * R.C * kv;
    ak[k] = make_double2(s * sr, s * si);
}

// work: (rep, first atom, first row, end row); 64 atoms per wave.  Each
// lane keeps exp(i n 2pi z/L), n < NZ, in registers: the kz terms of a row
// are then independent products cxy * z^|kz| (conjugated for kz < 0)
// instead of a serial recurrence along kz.
// GRAD: also dU_fourier/dq_i = Vf C sum_k kvec 2 (S_r cos + S_i sin)
//     = sum_k (A_k cos + B_k sin), added into gq.
template <bool GRAD, int NZ>
__global__ void __launch_bounds__(64) k_kforce(const int4* __restrict__ work,
                                               const RepDev* __restrict__ reps,
                                               const KRow* __restrict__ rows,
                                               const double2* __restrict__ ak,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const int* __restrict__ atom_site,
                                               const SiteDev* __restrict__ sites, double* fx,
                                               double* fy, double* fz, double* gq,
                                               const double2* __restrict__ eatom, int natom) {
    const int4 w = work[blockIdx.x];
    const RepDev& R = reps[w.x];
    const int a = w.y + threadIdx.x;
    const bool live = a < R.atom0 + R.natom;
    const int ai = live ? a : R.atom0;
    // exp(i 2 pi x / L) per axis, computed once per atom by k_sk (the same
    // sincos), instead of once per (atom, row group)
    cplx e[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double2 v = eatom[(size_t)d * natom + ai];
        e[d] = {v.x, v.y};
    }
    cplx zp[NZ];
    zp[0] = {1.0, 0.0};
#pragma unroll
    for (int n = 1; n < NZ; n++) zp[n] = cmul(zp[n - 1], e[2]);
    double gx = 0.0, gy = 0.0, gz = 0.0, dq = 0.0;
    int pkx = 1 << 30, pky = 1 << 30;
    cplx cx = {1.0, 0.0}, cy = {1.0, 0.0};
    for (int r = w.z; r < w.w; r++) {
        const KRow K = rows[r];
        if (K.kx != pkx) {
            cx = cpow_signed(e[0], K.kx);
            pkx = K.kx;
            pky = 1 << 30;
        }
        if (K.ky == pky + 1)
            cy = cmul(cy, e[1]);
        else if (K.ky != pky)
            cy = cpow_signed(e[1], K.ky);
        pky = K.ky;
        const cplx cxy = cmul(cx, cy);
        double gsum = 0.0, gzr = 0.0;
        const double2* A = ak + K.k0 - K.kz_lo;  // A[kz] for kz in [lo, hi]
        if (K.kz_lo <= 0 && 0 <= K.kz_hi) {
            const double2 ab = A[0];
            const double g = ab.x * cxy.im - ab.y * cxy.re;
            gsum += g;
            if (GRAD) dq += ab.x * cxy.re + ab.y * cxy.im;
        }
#pragma unroll
        for (int n = 1; n < NZ; n++) {
            if (n >= K.kz_lo && n <= K.kz_hi) {  // kz = +n (uniform branch)
                const cplx c = cmul(cxy, zp[n]);
                const double2 ab = A[n];
                const double g = ab.x * c.im - ab.y * c.re;
                gsum += g;
                gzr += g * n;
                if (GRAD) dq += ab.x * c.re + ab.y * c.im;
            }
            if (-n >= K.kz_lo && -n <= K.kz_hi) {  // kz = -n
                const cplx c = cmul(cxy, conjc(zp[n]));
                const double2 ab = A[-n];
                const double g = ab.x * c.im - ab.y * c.re;
                gsum += g;
                gzr -= g * n;
                if (GRAD) dq += ab.x * c.re + ab.y * c.im;
            }
        }
        gx += gsum * K.kx;
        gy += gsum * K.ky;
        gz += gzr;
    }
    if (!live) return;
    const double q = sites[atom_site[a]].charge;
    atomicAdd(&fx[a], q * gx * R.twopi_l[0]);
    atomicAdd(&fy[a], q * gy * R.twopi_l[1]);
    atomicAdd(&fz[a], q * gz * R.twopi_l[2]);
    if (GRAD) atomicAdd(&gq[a], dq);
}

}  // namespace

int recip_nkz_supported(int kmaxz) { return kmaxz <= 16; }

void launch_recip_s(const LaunchCtx& c, const RecipBufs& rb, const double* x, const double* y,
                    const double* z) {
    // threads per chunk: one per (row, kz segment) item of the replica with
    // the most (99 rows x 4 segments of 5 at kmax 8 / KSQMAX 64), items
    // looped past 1024
    constexpr int SEG = 5;
    const int items = rb.nkrow_max * ((rb.nkz_max + SEG - 1) / SEG);
    const int sk_t = std::min(1024, std::max(64, (items + 63) / 64 * 64));
    size_t lds = (size_t)64 * rb.ntab_max * sizeof(double2);
    hipLaunchKernelGGL(k_sk<SEG>, dim3(rb.nwork_s), dim3(sk_t), lds, c.stream, rb.work_s, c.reps,
                       rb.rows_s, x, y, z, c.atom_site, c.sites, rb.s_part, rb.nk_stride,
                       rb.eatom, c.natom);
    hipLaunchKernelGGL(k_sred, dim3((rb.nk_total + 7) / 8), dim3(256), 0, c.stream,
                       rb.nk_total, rb.k_rep, c.reps, rb.rep_chunk0, rb.s_part, rb.nk_stride,
                       rb.kvec, rb.ak, rb.uk);
}

void launch_recip_f(const LaunchCtx& c, const RecipBufs& rb, const double* x, const double* y,
                    const double* z, double* fx, double* fy, double* fz, double* gq) {
    // z powers kept per lane: |kz| <= kmaxz, nkz_max = 2 kmaxz + 1
#define KFORCE(G, NZ)                                                                          \
    hipLaunchKernelGGL((k_kforce<G, NZ>), dim3(rb.nwork_f), dim3(64), 0, c.stream, rb.work_f,   \
                       c.reps, rb.rows, rb.ak, x, y, z, c.atom_site, c.sites, fx, fy, fz, gq,  \
                       rb.eatom, c.natom)
    if (gq) {
        if (rb.nkz_max <= 9)
            KFORCE(true, 5);
        else if (rb.nkz_max <= 17)
            KFORCE(true, 9);
        else
            KFORCE(true, 17);
    } else {
        if (rb.nkz_max <= 9)
            KFORCE(false, 5);
        else if (rb.nkz_max <= 17)
            KFORCE(false, 9);
        else
            KFORCE(false, 17);
    }
#undef KFORCE
}
