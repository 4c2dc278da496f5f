// k_myinterp.hip -- the "myinterp" solid-fluid model on gfx950: the cube
// polynomials that init_my_interp reads (src/sffrc.c:228-748) and
// get_values_from_grid evaluates (src/sffrc.c:951-1089).
//
// The reference ships no program that writes the 32-coefficient file; this
// file generates it from first principles:
//   k_mi_nodes  one thread per node: the fluid-solid LJ sum of
//               tools/sfgridgen.c (same nodes, images and cutoff) and its
//               derivatives up to third order, FP64;
//   k_mi_fit    one thread per cube and polynomial: the 32 coefficients from
//               the 8 corners' value, gradient and mixed second derivatives
//               through a fixed 32 x 56 map (cube-local coordinates), then
//               the exact change to the absolute Angstrom monomials the file
//               stores.
//   k_mi_cube_major  the engine's device copy of the tables, cube-major
//               ([type][cube][128], one contiguous KiB per cube) so that a
//               site's lookup reads eight full cache lines.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/mspms_gpu.h"
#include "engine.h"

namespace {

constexpr int MI_T = 256;
constexpr int MI_MAX_SOLID = 1000;     // solid_atom_max, tools/sfgridgen.c:26
constexpr double MI_RGAS = 8.314472;   // tools/sfgridgen.c:27
constexpr int MI_NQ = 7;               // value, 3 first, 3 mixed second derivatives
constexpr int MI_ND = 8 * MI_NQ;       // data per cube and polynomial

// exponents (x, y, z) of the 32 monomials in Amatrix_ele_assign_value order
// (src/sffrc.c:771-802), one hex digit each: every monomial of degree <= 3,
// then the degree-4 ones without a fourth power.  constexpr, so that the
// compiler folds them inside unrolled loops.
__host__ __device__ constexpr int mono_exp(int m, int ax) {
    constexpr int e[32] = {0x000, 0x100, 0x010, 0x001, 0x200, 0x110, 0x101, 0x020,
                           0x011, 0x002, 0x300, 0x210, 0x120, 0x201, 0x102, 0x030,
                           0x021, 0x012, 0x003, 0x111, 0x310, 0x220, 0x130, 0x301,
                           0x202, 0x103, 0x031, 0x022, 0x013, 0x211, 0x121, 0x112};
    return (e[m] >> (4 * (2 - ax))) & 0xf;
}
__host__ __device__ constexpr int mono_at(int a, int b, int c) {
    for (int k = 0; k < 32; k++)
        if (mono_exp(k, 0) == a && mono_exp(k, 1) == b && mono_exp(k, 2) == c) return k;
    return -1;
}

// ---------------------------------------------------------------------------
// per-node LJ sum and derivatives

struct MiNodeParams {
    double x0, y0, z0, itvl, sig2, eps4, uclz, rcutoffsq;
    int nx, ny, nz, nuc_z, nsolid;
};

__global__ void __launch_bounds__(MI_T) k_mi_nodes(MiNodeParams P, const double* __restrict__ sx,
                                                    const double* __restrict__ sy,
                                                    const double* __restrict__ sz,
                                                    double* __restrict__ out, size_t nnode) {
    __shared__ double lx[MI_MAX_SOLID], ly[MI_MAX_SOLID], lz[MI_MAX_SOLID];
    for (int k = threadIdx.x; k < P.nsolid; k += MI_T) {
        lx[k] = sx[k];
        ly[k] = sy[k];
        lz[k] = sz[k];
    }
    __syncthreads();
    const size_t node = (size_t)blockIdx.x * MI_T + threadIdx.x;
    if (node >= nnode) return;
    const int kk = (int)(node % P.nz);
    const int jj = (int)((node / P.nz) % P.ny);
    const int ii = (int)(node / ((size_t)P.nz * P.ny));
    const double px = P.x0 + P.itvl * ii, py = P.y0 + P.itvl * jj, pz = P.z0 + P.itvl * kk;
    // E, E_a, E_ab (xx yy zz xy xz yz), E_abc (xxy xxz xyy xyz xzz yyz yzz)
    double e = 0, ex = 0, ey = 0, ez = 0, exx = 0, eyy = 0, ezz = 0, exy = 0, exz = 0, eyz = 0;
    double exxy = 0, exxz = 0, exyy = 0, exyz = 0, exzz = 0, eyyz = 0, eyzz = 0;
    for (int pp = -P.nuc_z; pp <= P.nuc_z; pp++) {
        for (int nn = 0; nn < P.nsolid; nn++) {
            const double dx = px - lx[nn], dy = py - ly[nn], dz = pz - lz[nn] + pp * P.uclz;
            const double q = dx * dx + dy * dy + dz * dz;
            if (!(q < P.rcutoffsq)) continue;
            // U(q) = 4 eps (s^12 q^-6 - s^6 q^-3), q = r^2, and dU/dq .. d3U/dq3
            const double iq = 1.0 / q;
            const double s6 = P.sig2 * P.sig2 * P.sig2 * iq * iq * iq;
            const double s12 = s6 * s6;
            const double u0 = P.eps4 * (s12 - s6);
            const double u1 = P.eps4 * (-6.0 * s12 + 3.0 * s6) * iq;
            const double u2 = P.eps4 * (42.0 * s12 - 12.0 * s6) * iq * iq;
            const double u3 = P.eps4 * (-336.0 * s12 + 60.0 * s6) * iq * iq * iq;
            // d_a U = 2 u1 d_a; d_ab U = 4 u2 d_a d_b + 2 u1 delta_ab;
            // d_abc U = 8 u3 d_a d_b d_c + 4 u2 (delta_ab d_c + delta_ac d_b + delta_bc d_a)
            const double g1 = 2.0 * u1, g2 = 4.0 * u2, g3 = 8.0 * u3;
            e += u0;
            ex += g1 * dx;
            ey += g1 * dy;
            ez += g1 * dz;
            exx += g2 * dx * dx + g1;
            eyy += g2 * dy * dy + g1;
            ezz += g2 * dz * dz + g1;
            exy += g2 * dx * dy;
            exz += g2 * dx * dz;
            eyz += g2 * dy * dz;
            exxy += g3 * dx * dx * dy + g2 * dy;
            exxz += g3 * dx * dx * dz + g2 * dz;
            exyy += g3 * dx * dy * dy + g2 * dx;
            exyz += g3 * dx * dy * dz;
            exzz += g3 * dx * dz * dz + g2 * dx;
            eyyz += g3 * dy * dy * dz + g2 * dz;
            eyzz += g3 * dy * dz * dz + g2 * dy;
        }
    }
    // (E, Fx, Fy, Fz) x (value, d/dx, d/dy, d/dz, d2/dxdy, d2/dxdz, d2/dydz), F = -grad E
    const double v[28] = {e,    ex,    ey,    ez,    exy,   exz,   eyz,    // E
                          -ex,  -exx,  -exy,  -exz,  -exxy, -exxz, -exyz,  // Fx
                          -ey,  -exy,  -eyy,  -eyz,  -exyy, -exyz, -eyyz,  // Fy
                          -ez,  -exz,  -eyz,  -ezz,  -exyz, -exzz, -eyzz};  // Fz
#pragma unroll
    for (int q = 0; q < 28; q++) out[(size_t)q * nnode + node] = v[q];
}

// ---------------------------------------------------------------------------
// per-cube fit

struct MiFitParams {
    double x0, y0, z0, hx, hy, hz;
    int nx, ny, nz, ncx, ncy, ncz;
};

__global__ void __launch_bounds__(MI_T) k_mi_fit(MiFitParams P, const double* __restrict__ G,
                                                  const double* __restrict__ nd,
                                                  double* __restrict__ coef, int ncube) {
    const int t = blockIdx.x * MI_T + threadIdx.x;
    const int f = blockIdx.y;  // polynomial: E, Fx, Fy, Fz
    if (t >= ncube) return;
    const int iz = t % P.ncz, iy = (t / P.ncz) % P.ncy, ix = t / (P.ncz * P.ncy);
    const size_t nnode = (size_t)P.nx * P.ny * P.nz;
    const double* base = nd + (size_t)f * MI_NQ * nnode;
    // cube-local data: derivatives scaled to u = (x - xA)/hx etc.
    const double sc[MI_NQ] = {1.0, P.hx, P.hy, P.hz, P.hx * P.hy, P.hx * P.hz, P.hy * P.hz};
    double data[MI_ND];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const size_t node =
            ((size_t)(ix + (c & 1)) * P.ny + (iy + ((c >> 1) & 1))) * P.nz + (iz + ((c >> 2) & 1));
#pragma unroll
        for (int d = 0; d < MI_NQ; d++) data[c * MI_NQ + d] = base[(size_t)d * nnode + node] * sc[d];
    }
    double e[32];
#pragma unroll
    for (int m = 0; m < 32; m++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < MI_ND; j++) s += G[m * MI_ND + j] * data[j];
        e[m] = s;
    }
    // u^a v^b w^c = ((x - xA)/hx)^a ...: scale, then Taylor-shift each axis;
    // the monomial set is closed under both (every lower monomial is in it)
    const double ih[3] = {1.0 / P.hx, 1.0 / P.hy, 1.0 / P.hz};
#pragma unroll
    for (int m = 0; m < 32; m++) {
#pragma unroll
        for (int ax = 0; ax < 3; ax++)
#pragma unroll
            for (int a = 0; a < mono_exp(m, ax); a++) e[m] *= ih[ax];
    }
    const double shift[3] = {-(P.x0 + P.hx * ix), -(P.y0 + P.hy * iy), -(P.z0 + P.hz * iz)};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        double o[32];
#pragma unroll
        for (int m = 0; m < 32; m++) o[m] = 0.0;
        // (X + s)^p = sum_i C(p,i) X^i s^(p-i), X = the absolute coordinate
#pragma unroll
        for (int m = 0; m < 32; m++) {
            const int p = mono_exp(m, ax);
            double sp = 1.0;  // s^(p-i)
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                if (i > p) continue;
                const double binom = (p == 3 && (i == 1 || i == 2)) ? 3.0
                                     : (p == 2 && i == 1)           ? 2.0
                                                                    : 1.0;
                const int tgt = mono_at(ax == 0 ? i : mono_exp(m, 0), ax == 1 ? i : mono_exp(m, 1),
                                        ax == 2 ? i : mono_exp(m, 2));
                o[tgt] += e[m] * binom * sp;
                sp *= shift[ax];
            }
        }
#pragma unroll
        for (int m = 0; m < 32; m++) e[m] = o[m];
    }
#pragma unroll
    for (int m = 0; m < 32; m++) coef[((size_t)f * 32 + m) * ncube + t] = e[m];
}

// the 32 x 56 fit map in cube-local coordinates u, v, w in [0, 1]; data
// index c*7 + d, corner c = cx + 2 cy + 4 cz, d as the node data
std::vector<double> fit_matrix() {
    auto mono = [](int k, const double p[3], const int der[3]) {
        // d^der of monomial k at p
        double r = 1.0;
        for (int a = 0; a < 3; a++) {
            int e = mono_exp(k, a);
            for (int d = 0; d < der[a]; d++) {
                if (e == 0) return 0.0;
                r *= e;
                e--;
            }
            for (int j = 0; j < e; j++) r *= p[a];
        }
        return r;
    };
    const int gder[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int mder[3][3] = {{1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    // value / gradient conditions M (32 x 32)
    std::vector<double> M(32 * 32);
    for (int c = 0; c < 8; c++) {
        const double p[3] = {(double)(c & 1), (double)((c >> 1) & 1), (double)((c >> 2) & 1)};
        for (int d = 0; d < 4; d++)
            for (int k = 0; k < 32; k++) M[(c * 4 + d) * 32 + k] = mono(k, p, gder[d]);
    }
    // null space of M: u v (1-u)(1-v), u w (1-u)(1-w), v w (1-v)(1-w)
    double N[3][32] = {};
    const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int n = 0; n < 3; n++) {
        const int a = pr[n][0], b = pr[n][1];
        auto put = [&](int pa, int pb, double w) {
            int ex[3] = {0, 0, 0};
            ex[a] = pa;
            ex[b] = pb;
            N[n][mono_at(ex[0], ex[1], ex[2])] += w;
        };
        put(1, 1, 1.0);
        put(2, 1, -1.0);
        put(1, 2, -1.0);
        put(2, 2, 1.0);
    }
    // A = (M^T M + N N^T)^-1 M^T: the minimum-norm least-squares solution
    std::vector<double> S(32 * 64, 0.0);  // [S | M^T] reduced to [I | A]
    for (int i = 0; i < 32; i++) {
        for (int j = 0; j < 32; j++) {
            double s = 0.0;
            for (int r = 0; r < 32; r++) s += M[r * 32 + i] * M[r * 32 + j];
            for (int n = 0; n < 3; n++) s += N[n][i] * N[n][j];
            S[i * 64 + j] = s;
        }
        for (int r = 0; r < 32; r++) S[i * 64 + 32 + r] = M[r * 32 + i];
    }
    for (int col = 0; col < 32; col++) {
        int piv = col;
        for (int r = col + 1; r < 32; r++)
            if (std::fabs(S[r * 64 + col]) > std::fabs(S[piv * 64 + col])) piv = r;
        if (piv != col)
            for (int j = 0; j < 64; j++) std::swap(S[col * 64 + j], S[piv * 64 + j]);
        const double d = S[col * 64 + col];
        for (int j = 0; j < 64; j++) S[col * 64 + j] /= d;
        for (int r = 0; r < 32; r++) {
            if (r == col) continue;
            const double fct = S[r * 64 + col];
            if (fct != 0.0)
                for (int j = 0; j < 64; j++) S[r * 64 + j] -= fct * S[col * 64 + j];
        }
    }
    // G: value/gradient data through A; then add alpha_n N_n with alpha_n the
    // mean over corners of s_c (mixed data - mixed derivative of the
    // least-squares polynomial), s_c = the sign of N_n's mixed derivative
    std::vector<double> G(32 * MI_ND, 0.0);
    for (int m = 0; m < 32; m++)
        for (int c = 0; c < 8; c++)
            for (int d = 0; d < 4; d++) G[m * MI_ND + c * MI_NQ + d] = S[m * 64 + 32 + c * 4 + d];
    std::vector<double> W(3 * MI_ND, 0.0);  // alpha_n as a row over the data
    for (int n = 0; n < 3; n++) {
        for (int c = 0; c < 8; c++) {
            const double p[3] = {(double)(c & 1), (double)((c >> 1) & 1),
                                 (double)((c >> 2) & 1)};
            const double sgn = (1.0 - 2.0 * p[pr[n][0]]) * (1.0 - 2.0 * p[pr[n][1]]);
            W[n * MI_ND + c * MI_NQ + 4 + n] += sgn / 8.0;
            for (int k = 0; k < 32; k++) {
                const double dk = mono(k, p, mder[n]);
                if (dk == 0.0) continue;
                for (int j = 0; j < MI_ND; j++) W[n * MI_ND + j] -= sgn / 8.0 * dk * G[k * MI_ND + j];
            }
        }
    }
    for (int m = 0; m < 32; m++)
        for (int n = 0; n < 3; n++)
            if (N[n][m] != 0.0)
                for (int j = 0; j < MI_ND; j++) G[m * MI_ND + j] += N[n][m] * W[n * MI_ND + j];
    return G;
}

// ---------------------------------------------------------------------------
// cube-major device copy of the engine's tables

__global__ void k_mi_cube_major(const double* __restrict__ src, double* __restrict__ dst,
                                int ntype, int ncube) {
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)ntype * ncube * 128;
    if (o >= total) return;
    const int m = (int)(o % 128);
    const size_t tc = o / 128;
    const int cube = (int)(tc % ncube), t = (int)(tc / ncube);
    dst[o] = src[((size_t)t * 128 + m) * ncube + cube];
}

struct DevBuf {
    double* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

void launch_cube_major(hipStream_t s, const double* src, double* dst, int ntype, int ncube) {
    const size_t total = (size_t)ntype * ncube * 128;
    hipLaunchKernelGGL(k_mi_cube_major, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       src, dst, ntype, ncube);
}

// the engine's error reporter (engine.cpp)
int mspms_fail(int code, const char* what);

extern "C" int mspms_myinterp_nodes(int device, const mspms_sfgridgen_spec* s, double* out) {
    if (!s || !out || !s->sx || !s->sy || !s->sz)
        return mspms_fail(MSPMS_EINVAL, "null argument");
    if (s->s_natom < 1 || s->s_natom > MI_MAX_SOLID)
        return mspms_fail(MSPMS_EINVAL, "1..1000 solid atoms (solid_atom_max)");
    if (!(s->rcutoff > 0.0) || !(s->uclz > 0.0))
        return mspms_fail(MSPMS_EINVAL, "cutoff and unit cell must be positive");
    if (s->isLJswitch)
        return mspms_fail(MSPMS_EINVAL, "the myinterp generator takes the unswitched potential");
    int nx, ny, nz;
    if (mspms_sfgridgen_dims(s, &nx, &ny, &nz))
        return mspms_fail(MSPMS_EINVAL, "grid interval must be positive");
    MiNodeParams P;
    const double sig = (s->f_sigma + s->s_sigma) / 2.0;
    P.sig2 = sig * sig;
    P.eps4 = 4.0 * sqrt(s->f_epsilon * s->s_epsilon) * MI_RGAS;
    P.rcutoffsq = s->rcutoff * s->rcutoff;
    int nuc_z = 1;
    while (nuc_z * s->uclz < s->rcutoff) nuc_z++;  // tools/sfgridgen.c:146-148
    P.nuc_z = nuc_z;
    P.x0 = 0.0;  // xcenter - uclx/2 with xcenter = uclx/2 (tools/sfgridgen.c:150-155)
    P.y0 = 0.0;
    P.z0 = 0.0;
    P.itvl = s->grid_itvl;
    P.uclz = s->uclz;
    P.nx = nx;
    P.ny = ny;
    P.nz = nz;
    P.nsolid = s->s_natom;
    const size_t nnode = (size_t)nx * ny * nz, ns = (size_t)s->s_natom;
    if (hipSetDevice(device) != hipSuccess) return mspms_fail(MSPMS_EIO, "hipSetDevice");
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return mspms_fail(MSPMS_EIO, "stream");
    int rc = 0;
    {
        DevBuf d_s, d_out;
        if (hipMalloc(&d_s.p, 3 * ns * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_out.p, 28 * nnode * sizeof(double)) != hipSuccess) {
            rc = mspms_fail(MSPMS_ENOMEM, "node tables");
        } else if (hipMemcpyAsync(d_s.p, s->sx, ns * sizeof(double), hipMemcpyHostToDevice, st) ||
                   hipMemcpyAsync(d_s.p + ns, s->sy, ns * sizeof(double), hipMemcpyHostToDevice,
                                  st) ||
                   hipMemcpyAsync(d_s.p + 2 * ns, s->sz, ns * sizeof(double),
                                  hipMemcpyHostToDevice, st)) {
            rc = mspms_fail(MSPMS_EIO, "upload");
        } else {
            hipLaunchKernelGGL(k_mi_nodes, dim3((unsigned)((nnode + MI_T - 1) / MI_T)), dim3(MI_T),
                               0, st, P, d_s.p, d_s.p + ns, d_s.p + 2 * ns, d_out.p, nnode);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(out, d_out.p, 28 * nnode * sizeof(double), hipMemcpyDeviceToHost,
                               st) ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = mspms_fail(MSPMS_EIO, "node kernel");
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int mspms_myinterp_fit(int device, int nx, int ny, int nz, double x0, double y0,
                                  double z0, double hx, double hy, double hz,
                                  const double* node_data, double* coef) {
    if (!node_data || !coef) return mspms_fail(MSPMS_EINVAL, "null argument");
    if (nx < 2 || ny < 2 || nz < 2 || !(hx > 0) || !(hy > 0) || !(hz > 0))
        return mspms_fail(MSPMS_EINVAL, "at least 2 nodes per axis and positive intervals");
    const size_t nnode = (size_t)nx * ny * nz;
    const size_t ncube_l = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (ncube_l > (size_t)0x7fffffff) return mspms_fail(MSPMS_EINVAL, "too many cubes");
    const int ncube = (int)ncube_l;
    MiFitParams P{x0, y0, z0, hx, hy, hz, nx, ny, nz, nx - 1, ny - 1, nz - 1};
    const std::vector<double> G = fit_matrix();
    if (hipSetDevice(device) != hipSuccess) return mspms_fail(MSPMS_EIO, "hipSetDevice");
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return mspms_fail(MSPMS_EIO, "stream");
    int rc = 0;
    {
        DevBuf d_g, d_in, d_out;
        if (hipMalloc(&d_g.p, G.size() * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_in.p, 28 * nnode * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_out.p, 128 * ncube_l * sizeof(double)) != hipSuccess) {
            rc = mspms_fail(MSPMS_ENOMEM, "fit tables");
        } else if (hipMemcpyAsync(d_g.p, G.data(), G.size() * sizeof(double),
                                  hipMemcpyHostToDevice, st) ||
                   hipMemcpyAsync(d_in.p, node_data, 28 * nnode * sizeof(double),
                                  hipMemcpyHostToDevice, st)) {
            rc = mspms_fail(MSPMS_EIO, "upload");
        } else {
            hipLaunchKernelGGL(k_mi_fit, dim3((unsigned)((ncube + MI_T - 1) / MI_T), 4),
                               dim3(MI_T), 0, st, P, d_g.p, d_in.p, d_out.p, ncube);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(coef, d_out.p, 128 * ncube_l * sizeof(double),
                               hipMemcpyDeviceToHost, st) ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = mspms_fail(MSPMS_EIO, "fit kernel");
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}
