// k_pair.hip -- inter-molecular nonbonded pairs on gfx950: LJ + Ewald real
// space or Wolf, per cluster pair.  Reference: the atom-pair body of
// src/frclong.c:105-196 (ljfrc, ewald_real_frc, wolf_real_frc).
//
// A wave64 walks (part of) an i-cluster's row of the cluster-pair list.
// Lane (i = lane>>3, j = lane&7) evaluates one site pair per j-cluster;
// each i's force is reduced over its 8 lanes with wave shuffles at the end,
// and the energy / virial partials over the whole wave.  For the exact
// kernel the lists are full (both i->j and j->i), so every atom's force is
// written once, without atomics, in a fixed order: the result is
// deterministic.  The mixed kernel takes each cluster pair once and adds
// the j forces with FP64 atomics (half the pair terms; forces repeat to the
// last bits).  Energies and virial reach the finalize doubled and are halved
// there; pair counts and hashes are taken on i<j only.
//   k_pair     exact mode: one wave per i-cluster
//   k_pair_mx  mixed mode: WPC waves per i-cluster sharing its row, the
//              partials combined through LDS in wave order
//
// Precision modes:
//   FP64   separation from the unwrapped FP64 positions with the reference's
//          minimum image (d - L*rint(d/L), true division, no contraction),
//          reference-order FP64 terms: the (i<j, r<rc) pair sets are those of
//          src/frclong.c:114-125 bit for bit.
//   MIXED  separation = (FP32 offset_i - offset_j) + (cluster-centre delta
//          and image shift, FP64 rounded once to FP32); FP32 terms with the
//          polynomial erfc; FP64 accumulation of forces, energies, virial.
#include "engine.h"
#include "physics.h"
#include "wave.h"
#include "../../include/mspms_gpu.h"

namespace {

__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// pair-set hash key of (i, j), replica-local atom ids (the numbering a
// replica has as a system of its own)
__device__ inline uint64_t pair_key(int i, int j) {
    return splitmix64(((uint64_t)(uint32_t)i << 32) | (uint32_t)j);
}

// a lane's float from another lane of its row by a DPP pattern (all lanes
// active; row and bank masks full)
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// XCD-aware block order: consecutive i-clusters (spatial neighbours that
// share j-clusters) land on the same XCD and its L2.
__device__ inline int xcd_block(int b, int nb) {
    const int NX = 8;
    int per = (nb + NX - 1) / NX;
    int xcd = b % NX, k = b / NX;
    int lb = xcd * per + k;
    return lb;
}

// Exact mode (FP64, the reference's operation order), WPC waves per
// i-cluster combined through LDS in wave order; GRAD adds the per-atom
// parameter derivatives.
template <int ELEC, bool FP64, bool GRAD, int WPC>
__global__ void __launch_bounds__(64 * WPC) k_pair(
    const int* __restrict__ ncl_total, int maxcl, const RepDev* __restrict__ reps,
    const SiteDev* __restrict__ sites, const double4* __restrict__ cl_center,
    const float4* __restrict__ cl_box, const float4* __restrict__ plist,
    const int* __restrict__ plist_n, int cap, const int* __restrict__ slot_orig,
    const int2* __restrict__ slot_info, const double4* __restrict__ slot_x64,
    const float4* __restrict__ slot_off,
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    double* __restrict__ epart, unsigned long long* __restrict__ ipart, int want_hash,
    double* __restrict__ gq, double* __restrict__ ge, double* __restrict__ gs) {
    static_assert(FP64, "the mixed-precision path is k_pair_mx");
    __shared__ double sh_f[WPC][MSPMS_CL][6];
    __shared__ double sh_e[WPC][4];
    __shared__ unsigned long long sh_n[WPC][4];
    const int ci = xcd_block(blockIdx.x, gridDim.x);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ncl = min(*ncl_total, maxcl);
    if (ci >= ncl) return;  // whole workgroup
    const int isub = lane >> 3, jsub = lane & 7;

    const double4 Ci = cl_center[ci];
    const RepDev& R = reps[(int)Ci.w];
    const int islot = ci * MSPMS_CL + isub;
    const int2 infoi = slot_info[islot];
    const bool vi = infoi.y >= 0;
    const int oi = slot_orig[islot];
    SiteDev si;
    if (vi) si = sites[infoi.y];
    const bool lji = vi && si.ghost != MSPMS_GHOST_LJ;

    double afx = 0.0, afy = 0.0, afz = 0.0;
    double g_q = 0.0, g_e = 0.0, g_s = 0.0;  // d(pair energy)/d(own q, eps, sigma)
    double e_lj = 0.0, e_sh = 0.0, e_re = 0.0, vir = 0.0;
    unsigned long long n_lj = 0, n_el = 0, h_lj = 0, h_el = 0;

    const int nrow = plist_n[ci];
    const float4* row = plist + (size_t)ci * cap;

    // row pruning as in k_pair_mx: an entry whose current cluster boxes are
    // farther apart than the cutoff (with a margin of 1e-3 of it above FP32
    // rounding) holds no pair in range; the test is the same for every lane
    // of the wave, and the pairs that interact are still decided per pair
    const float rprune = (float)fmax(R.rc_lj, R.rc_el) * 1.001f;
    const float rprune2 = rprune * rprune;
    const float4 bim = cl_box[2 * ci], bih = cl_box[2 * ci + 1];
    {
        const double4 pi4 = slot_x64[islot];
        for (int e = w; e < nrow; e += WPC) {
            const float4 E = row[e];
            const uint32_t ent = __float_as_uint(E.w);
            const int cj = (int)(ent >> 7);
            if (cj != ci && !(ent & 64u)) {
                const float4 bjm = cl_box[2 * cj], bjh = cl_box[2 * cj + 1];
                const float gx = fmaxf(fabsf(E.x + (bim.x - bjm.x)) - (bih.x + bjh.x), 0.f);
                const float gy = fmaxf(fabsf(E.y + (bim.y - bjm.y)) - (bih.y + bjh.y), 0.f);
                const float gz = fmaxf(fabsf(E.z + (bim.z - bjm.z)) - (bih.z + bjh.z), 0.f);
                if (!(gx * gx + gy * gy + gz * gz < rprune2)) continue;
            }
            const int jslot = cj * MSPMS_CL + jsub;
            const int2 infoj = slot_info[jslot];
            if (!(vi && infoj.y >= 0 && infoj.x != infoi.x)) continue;
            const double4 pj4 = slot_x64[jslot];
            double rx, ry, rz, rsq;
            {
#pragma clang fp contract(off)
                rx = pi4.x - pj4.x;
                ry = pi4.y - pj4.y;
                rz = pi4.z - pj4.z;
                rx = rx - R.L[0] * rint(rx / R.L[0]);
                ry = ry - R.L[1] * rint(ry / R.L[1]);
                rz = rz - R.L[2] * rint(rz / R.L[2]);
                rsq = rx * rx + ry * ry + rz * rz;
            }
            const SiteDev sj = sites[infoj.y];
            const int oj = slot_orig[jslot];
            const bool count = oi < oj;
            if (lji && sj.ghost != MSPMS_GHOST_LJ && rsq < R.rc2_lj) {
                double u, f, ush;
                double sij, eij;
                {
#pragma clang fp contract(off)
                    sij = 0.5 * (si.sigma + sj.sigma);
                    eij = sqrt(si.epsilon * sj.epsilon);
                }
                lj64(rsq, sij, eij, R.switch_on, R.rc2_lj, R.rcon2, R.roff2_ron2_cube, R.shift1,
                     &u, &f, &ush);
                e_lj += u;
                e_sh += ush;
                vir += f * rsq;
                if (GRAD) {
                    double dds, dde;
                    lj64_dparam(rsq, sij, eij, R.switch_on, R.rc2_lj, R.rcon2, R.roff2_ron2_cube,
                                &dds, &dde);
                    g_s += 0.5 * dds;  // d sigma_ij / d sigma_i = 1/2
                    g_e += dde * deps_mix(eij, sj.epsilon);
                }
                afx += f * rx;
                afy += f * ry;
                afz += f * rz;
                if (count) {
                    n_lj++;
                    if (want_hash) h_lj += pair_key(oi - R.atom0, oj - R.atom0);
                }
            }
            if (ELEC != 0 && rsq < R.rc2_el) {
                double u, f;
                double qq;
                {
#pragma clang fp contract(off)
                    qq = si.charge * sj.charge;
                }
                if (ELEC == MSPMS_ELEC_EWALD) {
                    ewald_real64(rsq, qq, R.C, R.kappa, &u, &f);
                } else {
                    wolf_real64(rsq, qq, R.C, R.kappa, R.rc_el, R.wvc1, R.wvc2, R.wfc1, R.wfc2, &u,
                                &f);
                    vir += f * rsq;
                }
                e_re += u;
                if (GRAD) {
                    // the pair energy is linear in q_i: du/dq_i = u(q_i = 1)
                    double ug, fg;
                    if (ELEC == MSPMS_ELEC_EWALD)
                        ewald_real64(rsq, sj.charge, R.C, R.kappa, &ug, &fg);
                    else
                        wolf_real64(rsq, sj.charge, R.C, R.kappa, R.rc_el, R.wvc1, R.wvc2, R.wfc1,
                                    R.wfc2, &ug, &fg);
                    g_q += ug;
                }
                afx += f * rx;
                afy += f * ry;
                afz += f * rz;
                if (count) {
                    n_el++;
                    if (want_hash) h_el += pair_key(oi - R.atom0, oj - R.atom0);
                }
            }
        }
    }

    // per-atom partials: 8 j-lanes of each i, then the WPC waves in order
    double fr[6] = {afx, afy, afz, g_q, g_e, g_s};
    if (GRAD) {
        butterfly<double, 6>(fr, 4);
    } else {
        double f3[3] = {afx, afy, afz};
        butterfly<double, 3>(f3, 4);
        for (int d = 0; d < 3; d++) fr[d] = f3[d];
    }
    if (jsub == 0)
        for (int d = 0; d < 6; d++) sh_f[w][isub][d] = fr[d];
    double er[4] = {e_lj, e_sh, e_re, vir};
    unsigned long long nr[4] = {n_lj, n_el, h_lj, h_el};
    butterfly<double, 4>(er, 32);
    butterfly<unsigned long long, 4>(nr, 32);
    if (lane == 0) {
        for (int k = 0; k < 4; k++) {
            sh_e[w][k] = er[k];
            sh_n[w][k] = nr[k];
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < MSPMS_CL) {
        const int sl = ci * MSPMS_CL + t;
        if (slot_info[sl].y >= 0) {
            double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int ww = 0; ww < WPC; ww++)
                for (int d = 0; d < (GRAD ? 6 : 3); d++) g[d] += sh_f[ww][t][d];
            const int o = slot_orig[sl];
            fx[o] = g[0];
            fy[o] = g[1];
            fz[o] = g[2];
            if (GRAD) {
                gq[o] = g[3];
                ge[o] = g[4];
                gs[o] = g[5];
            }
        }
    } else if (t >= MSPMS_CL && t < MSPMS_CL + 4) {
        const int k = t - MSPMS_CL;
        double sum = 0.0;
        unsigned long long cnt = 0;
        for (int ww = 0; ww < WPC; ww++) {
            sum += sh_e[ww][k];
            cnt += sh_n[ww][k];
        }
        epart[(size_t)ci * 4 + k] = sum;
        ipart[(size_t)ci * 4 + k] = cnt;
    }
}

// Mixed precision, WPC waves per i-cluster: wave w of the workgroup takes
// row entries w, w + WPC, ...; the WPC partial forces and energies of each
// i are combined through LDS in wave order.  Per entry a lane loads two
// float4 (offset + charge, sigma / sqrt(eps) / molecule / flags) and the
// wave one FP64 cluster centre.
// HALF: each cluster pair sits in one row only (the self pair in its own,
// where lanes with j > i take each site pair once).  A lane's pair force
// goes to i as before and, negated, to j: per entry the eight i-lanes of
// each j are summed with shuffles and added with an FP64 atomic to j's slot
// in the slot-ordered force planes (one 64-B line per j-cluster and
// component); i's total is added atomically too.  k_list_check zeroed the
// planes; the first per-atom pass after this kernel folds them into the
// atoms' forces.  Energies and virial are doubled, since the finalize
// halves them.
// HASH: the pair hashes are kept (a compile-time switch: their registers
// would otherwise cost the kernel a wave per SIMD)
template <int ELEC, int WPC, bool HALF, bool HASH>
__global__ void __launch_bounds__(64 * WPC) k_pair_mx(
    const int* __restrict__ ncl_total, int maxcl, const RepDev* __restrict__ reps,
    const double4* __restrict__ cl_center, const float4* __restrict__ cl_box,
    const float4* __restrict__ plist, const int* __restrict__ plist_n, int cap,
    const int* __restrict__ slot_orig, const float4* __restrict__ slot_off,
    const float4* __restrict__ slot_par, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ fslot, int nslot,
    double* __restrict__ epart, unsigned long long* __restrict__ ipart) {
    constexpr bool want_hash = HASH;
    __shared__ double sh_f[WPC][MSPMS_CL][3];
    __shared__ double sh_e[WPC][4];
    __shared__ unsigned long long sh_n[WPC][4];
    // HALF: the j forces of a chunk's entries, added to the slot-ordered
    // forces after the chunk (one value per entry, j atom and component)
    __shared__ float sh_jf[HALF ? WPC : 1][64][3][MSPMS_CL];
    __shared__ int sh_jc[HALF ? WPC : 1][64];
    const int ci = xcd_block(blockIdx.x, gridDim.x);
    const int ncl = min(*ncl_total, maxcl);
    if (ci >= ncl) return;  // whole workgroup
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // HALF: i = lane & 7, so that the eight i-lanes of a j are adjacent and
    // their per-entry sum takes three DPP steps; otherwise i = lane >> 3
    const int isub = HALF ? (lane & 7) : (lane >> 3), jsub = HALF ? (lane >> 3) : (lane & 7);

    const double4 Ci = cl_center[ci];
    const RepDev& R = reps[(int)Ci.w];
    const int islot = ci * MSPMS_CL + isub;
    const float4 pi4 = slot_off[islot];
    const float4 qi4 = slot_par[islot];
    const int moli = __float_as_int(qi4.z), fli = __float_as_int(qi4.w);
    const bool vi = (fli & SLOT_VALID) != 0;
    const int oi = want_hash && vi ? slot_orig[islot] - R.atom0 : 0;
    const float q_i = pi4.w;
    const float rc2lj = (float)R.rc2_lj, rc2el = (float)R.rc2_el;
    const float Lx = (float)R.L[0], Ly = (float)R.L[1], Lz = (float)R.L[2];
    const float iLx = (float)R.invL[0], iLy = (float)R.invL[1], iLz = (float)R.invL[2];
    const float kappa = (float)R.kappa, Cf = (float)R.C, shift1 = (float)R.shift1;
    const float rcel = (float)R.rc_el, v1 = (float)R.wvc1, v2 = (float)R.wvc2,
                f1 = (float)R.wfc1, f2 = (float)R.wfc2;
    const float rcon2 = (float)R.rcon2, inv_den = (float)(1.0 / R.roff2_ron2_cube);
    // replica constants in registers: a reload inside the entry loop would
    // wait for the prefetched loads
    const bool sw_on = R.switch_on != 0;
    const int atom0 = R.atom0;
    // pruning of the row: an entry whose current cluster boxes are farther
    // apart than the cutoff (a margin of 1e-3 of it above FP32 rounding)
    // holds no pair in range and is skipped; the boxes are the clusters'
    // current extents about their build-time centres
    const float rprune = (float)fmax(R.rc_lj, R.rc_el) * 1.001f;
    const float rprune2 = rprune * rprune;
    const float4 bim = cl_box[2 * ci], bih = cl_box[2 * ci + 1];

    // FP64 accumulators, fed from FP32 partials that cover at most FLUSH row
    // entries of this wave (FLUSH pair terms per lane)
    double afx = 0.0, afy = 0.0, afz = 0.0;
    double e_lj = 0.0, e_sh = 0.0, e_re = 0.0, vir = 0.0;
    float pfx = 0.f, pfy = 0.f, pfz = 0.f, pe_lj = 0.f, pe_sh = 0.f, pe_re = 0.f, pvir = 0.f;
    unsigned int c_lj = 0, c_el = 0;
    unsigned long long h_lj = 0, h_el = 0;
    constexpr int FLUSH = 8;
    auto flush = [&]() {
        afx += (double)pfx;
        afy += (double)pfy;
        afz += (double)pfz;
        e_lj += (double)pe_lj;
        e_sh += (double)pe_sh;
        e_re += (double)pe_re;
        vir += (double)pvir;
        pfx = pfy = pfz = pe_lj = pe_sh = pe_re = pvir = 0.f;
    };

    // This wave's row entries (w, w + WPC, ...) are loaded 64 at a time, one
    // per lane, and read back with readlane.  The loop is unrolled by two
    // with two register sets: each set's j data are loaded (unconditionally,
    // the index clamped to the chunk) one entry before they are used, so at
    // every use at least the other set's three loads are in flight and the
    // wait does not cover them or the j-force atomics.
    const int nrow = plist_n[ci];
    const float4* row = plist + (size_t)ci * cap;
    const int nmine = nrow > w ? (nrow - w + WPC - 1) / WPC : 0;
    int it = 0;
    for (int base = 0; base < nmine; base += 64) {
        const float4 my_ent = base + lane < nmine ? row[w + (base + lane) * WPC]
                                                  : make_float4(0.f, 0.f, 0.f, 0.f);
        bool keep = base + lane < nmine;
        {
            const uint32_t ent = __float_as_uint(my_ent.w);
            const int cj = (int)(ent >> 7);
            if (keep && cj != ci && !(ent & 64u)) {
                const float4 bjm = cl_box[2 * cj], bjh = cl_box[2 * cj + 1];
                const float gx = fmaxf(fabsf(my_ent.x + (bim.x - bjm.x)) - (bih.x + bjh.x), 0.f);
                const float gy = fmaxf(fabsf(my_ent.y + (bim.y - bjm.y)) - (bih.y + bjh.y), 0.f);
                const float gz = fmaxf(fabsf(my_ent.z + (bim.z - bjm.z)) - (bih.z + bjh.z), 0.f);
                keep = gx * gx + gy * gy + gz * gz < rprune2;
            }
        }
        // the kept entries of this chunk, visited in row order
        uint64_t kept = __ballot(keep);
        const int cnt = __popcll(kept);
        if (cnt == 0) continue;
        int last = 0;
        auto next_lane = [&]() {
            if (kept) {
                last = __builtin_ctzll(kept);
                kept &= kept - 1;
            }
            return last;
        };
        auto entry_at = [&](int k) {
            return make_float4(
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_ent.x), k)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_ent.y), k)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_ent.z), k)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_ent.w), k)));
        };
        struct JSet {
            float4 sep, pj, qj;
            int oj;
        };
        auto fetch = [&](JSet& J) {
            J.sep = entry_at(next_lane());
            const int cn = (int)(__float_as_uint(J.sep.w) >> 7);
            J.pj = slot_off[cn * MSPMS_CL + jsub];
            J.qj = slot_par[cn * MSPMS_CL + jsub];
            J.oj = HALF && HASH ? slot_orig[cn * MSPMS_CL + jsub] : 0;
        };
        int kk = 0;  // entries of this chunk processed so far
        auto process = [&](const JSet& J) {
            if (it > 0 && it % FLUSH == 0) flush();  // uniform over the wave
            ++it;
            const float4 sep = J.sep;
            const uint32_t ent = __float_as_uint(sep.w);
            const float4 pj4 = J.pj, qj4 = J.qj;
            const int ojs = J.oj;  // j's atom (HALF and HASH)
            const int cj = (int)(ent >> 7);
            const int jslot = cj * MSPMS_CL + jsub;
            const int flj = __float_as_int(qj4.w);
            bool act = vi && (flj & SLOT_VALID) && __float_as_int(qj4.z) != moli;
            if (HALF && cj == ci) act = act && jsub > isub;
            float gx = 0.f, gy = 0.f, gz = 0.f;  // this lane's pair force on i
            bool hit = false;
            if (act) {
                float rx = sep.x + (pi4.x - pj4.x);
                float ry = sep.y + (pi4.y - pj4.y);
                float rz = sep.z + (pi4.z - pj4.z);
                if (ent & 64u) {
                    rx -= Lx * rintf(rx * iLx);
                    ry -= Ly * rintf(ry * iLy);
                    rz -= Lz * rintf(rz * iLz);
                }
                const float rsq = rx * rx + ry * ry + rz * rz;
                const bool in_lj = rsq < rc2lj && (fli & flj & SLOT_LJ);
                const bool in_el = ELEC != 0 && rsq < rc2el;
                if (in_lj || in_el) {
                    hit = true;
                    const bool count = HALF || islot < jslot;
                    int oj = 0, ka = oi, kb = 0;
                    if (want_hash) {
                        oj = (HALF ? ojs : slot_orig[jslot]) - atom0;
                        kb = oj;
                        if (HALF && oj < oi) {
                            ka = oj;
                            kb = oi;
                        }
                    }
                    const float rinv2 = __builtin_amdgcn_rcpf(rsq);  // v_rcp_f32 (~1 ulp)
                    float fsum = 0.f;
                    if (in_lj) {
                        const float sij = 0.5f * (qi4.x + qj4.x);
                        const float eij = qi4.y * qj4.y;
                        const float s2 = sij * sij;
                        float s6 = s2 * rinv2;
                        s6 = s6 * s6 * s6;
                        const float s12 = s6 * s6;
                        float u = 4.0f * eij * (s12 - s6);
                        float f = 24.0f * eij * (2.0f * s12 - s6) * rinv2;
                        float ush = 0.f;
                        if (sw_on) {
                            if (rsq >= rcon2) {
                                const float a = rc2lj - rsq;
                                const float sw = a * a * (rc2lj + 2.0f * rsq - 3.0f * rcon2) * inv_den;
                                f = f * sw - u * 12.0f * a * (rcon2 - rsq) * inv_den;
                                u *= sw;
                            }
                        } else {
                            const float sg6 = s2 * s2 * s2;
                            ush = eij * sg6 * (sg6 * shift1 - 1.0f);
                        }
                        pe_lj += u;
                        pe_sh += ush;
                        pvir += f * rsq;
                        fsum += f;
                        if (count) {
                            c_lj++;
                            if (want_hash) h_lj += pair_key(ka, kb);
                        }
                    }
                    if (in_el) {
                        const float rinv = rsqrtf(rsq);
                        const float r = rsq * rinv;
                        const float x = kappa * r;
                        const float ex = __expf(-x * x);
                        const float ec = ex * erfc_scaled_poly(x);
                        const float cq = Cf * q_i * pj4.w;
                        float u, f;
                        if (ELEC == MSPMS_ELEC_EWALD) {
                            u = cq * rinv * ec;
                            f = (u + cq * rinv * MSPMS_TWO_OVER_SQRTPI_F * x * ex) * rinv2;
                        } else {
                            u = cq * (ec * rinv + v1 + v2 * (r - rcel));
                            f = cq * rinv * (ec * rinv2 + f1 * ex * rinv + f2);
                            pvir += f * rsq;
                        }
                        pe_re += u;
                        fsum += f;
                        if (count) {
                            c_el++;
                            if (want_hash) h_el += pair_key(ka, kb);
                        }
                    }
                    gx = fsum * rx;
                    gy = fsum * ry;
                    gz = fsum * rz;
                    pfx += gx;
                    pfy += gy;
                    pfz += gz;
                }
            }
            if (HALF) {
                // j's share of this entry: the sum over its eight adjacent i-lanes
                // (lane ^ 1, lane ^ 2, then the mirror within the 8-lane half row)
                float v[3] = {gx, gy, gz};
    #pragma unroll
                for (int k = 0; k < 3; k++) {
                    v[k] += dpp_f32<0xB1>(v[k]);   // quad_perm [1,0,3,2]
                    v[k] += dpp_f32<0x4E>(v[k]);   // quad_perm [2,3,0,1]
                    v[k] += dpp_f32<0x141>(v[k]);  // row_half_mirror
                }
                // staged in LDS: the eight lanes of a j hold the same sum.
                // Atomics issued here, between one entry's j loads and the
                // next's, would sit under the wait for those loads.
                // (the force on i; j's is its negative, taken at the flush)
                if (isub == 0) {
                    sh_jf[w][kk][0][jsub] = v[0];
                    sh_jf[w][kk][1][jsub] = v[1];
                    sh_jf[w][kk][2][jsub] = v[2];
                }
                if (lane == 0) sh_jc[w][kk] = cj;
            }
            ++kk;
        };
        JSet A, B;
        fetch(A);
        for (int tt = 0; tt < cnt; tt += 2) {
            fetch(B);
            process(A);
            if (tt + 1 < cnt) {
                fetch(A);
                process(B);
            }
        }
        if (HALF) {
            // the chunk's j forces into the slot-ordered forces: eight
            // entries per instruction, one 64-B line each (LDS written and
            // read by this wave only, in order)
            __builtin_amdgcn_wave_barrier();
            for (int e0 = 0; e0 < cnt; e0 += 8) {
                const int e = e0 + (lane >> 3), a = lane & 7;
                const float f0 = e < cnt ? sh_jf[w][e][0][a] : 0.f;
                const float f1 = e < cnt ? sh_jf[w][e][1][a] : 0.f;
                const float f2 = e < cnt ? sh_jf[w][e][2][a] : 0.f;
                // j atoms without a pair in range (and padding slots) add nothing
                if (f0 != 0.f || f1 != 0.f || f2 != 0.f) {
                    const int sl = sh_jc[w][e] * MSPMS_CL + a;
                    atomicAdd(&fslot[sl], -(double)f0);
                    atomicAdd(&fslot[(size_t)nslot + sl], -(double)f1);
                    atomicAdd(&fslot[2 * (size_t)nslot + sl], -(double)f2);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    flush();
    double fr[3] = {afx, afy, afz};
    if (HALF) {
        // i's lanes are lane & 7 + 8 j: sum over lane bits 3..5
#pragma unroll
        for (int o = 32; o >= 8; o >>= 1) {
            double t[3];
#pragma unroll
            for (int k = 0; k < 3; k++) t[k] = __shfl_xor(fr[k], o, 64);
#pragma unroll
            for (int k = 0; k < 3; k++) fr[k] += t[k];
        }
    } else {
        butterfly<double, 3>(fr, 4);
    }
    if (jsub == 0) {
        sh_f[w][isub][0] = fr[0];
        sh_f[w][isub][1] = fr[1];
        sh_f[w][isub][2] = fr[2];
    }
    double er[4] = {e_lj, e_sh, e_re, vir};
    butterfly<double, 4>(er, 32);
    // counts fit 32 bits per wave; the 64-bit hashes only when asked for
    unsigned int cr[2] = {c_lj, c_el};
    butterfly<unsigned int, 2>(cr, 32);
    unsigned long long hr[2] = {h_lj, h_el};
    if (want_hash) butterfly<unsigned long long, 2>(hr, 32);
    if (lane == 0) {
        for (int k = 0; k < 4; k++) sh_e[w][k] = HALF ? 2.0 * er[k] : er[k];
        sh_n[w][0] = cr[0];
        sh_n[w][1] = cr[1];
        sh_n[w][2] = want_hash ? hr[0] : 0ull;
        sh_n[w][3] = want_hash ? hr[1] : 0ull;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < MSPMS_CL) {
        const int sl = ci * MSPMS_CL + t;
        if (__float_as_int(slot_par[sl].w) & SLOT_VALID) {
            double g[3] = {0.0, 0.0, 0.0};
            for (int ww = 0; ww < WPC; ww++)
                for (int d = 0; d < 3; d++) g[d] += sh_f[ww][t][d];
            if (HALF) {
                atomicAdd(&fslot[sl], g[0]);
                atomicAdd(&fslot[(size_t)nslot + sl], g[1]);
                atomicAdd(&fslot[2 * (size_t)nslot + sl], g[2]);
            } else {
                const int o = slot_orig[sl];
                fx[o] = g[0];
                fy[o] = g[1];
                fz[o] = g[2];
            }
        }
    } else if (t >= MSPMS_CL && t < MSPMS_CL + 4) {
        const int k = t - MSPMS_CL;
        double sum = 0.0;
        unsigned long long cnt = 0;
        for (int ww = 0; ww < WPC; ww++) {
            sum += sh_e[ww][k];
            cnt += sh_n[ww][k];
        }
        epart[(size_t)ci * 4 + k] = sum;
        ipart[(size_t)ci * 4 + k] = cnt;
    }
}

// one workgroup per i-cluster; the grid is padded to a multiple of 8 so the
// XCD remap stays a bijection
template <int ELEC, bool GRAD, int WPC>
void launch_64(const LaunchCtx& c, const ClusterBufs& b, double* fx, double* fy, double* fz,
               double* epart, unsigned long long* ipart, int want_hash, double* gq, double* ge,
               double* gs) {
    hipLaunchKernelGGL((k_pair<ELEC, true, GRAD, WPC>), dim3((b.maxcl + 7) / 8 * 8),
                       dim3(64 * WPC), 0, c.stream, b.ncl_total, b.maxcl, c.reps, c.sites,
                       b.cl_center, b.cl_box, b.plist, b.plist_n, b.cap, b.slot_orig, b.slot_info,
                       b.slot_x64, b.slot_off, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
}

template <int ELEC, int WPC, bool HALF, bool HASH>
void launch_mx_k(const LaunchCtx& c, const ClusterBufs& b, double* fx, double* fy, double* fz,
                 double* epart, unsigned long long* ipart) {
    hipLaunchKernelGGL((k_pair_mx<ELEC, WPC, HALF, HASH>), dim3((b.maxcl + 7) / 8 * 8),
                       dim3(64 * WPC), 0, c.stream, b.ncl_total, b.maxcl, c.reps, b.cl_center,
                       b.cl_box, b.plist, b.plist_n, b.cap, b.slot_orig, b.slot_off, b.slot_par,
                       fx, fy, fz, b.fslot, b.maxcl * MSPMS_CL, epart, ipart);
}

template <int ELEC, int WPC>
void launch_mx(const LaunchCtx& c, const ClusterBufs& b, double* fx, double* fy, double* fz,
               double* epart, unsigned long long* ipart, int want_hash) {
    if (b.half && want_hash)
        launch_mx_k<ELEC, WPC, true, true>(c, b, fx, fy, fz, epart, ipart);
    else if (b.half)
        launch_mx_k<ELEC, WPC, true, false>(c, b, fx, fy, fz, epart, ipart);
    else if (want_hash)
        launch_mx_k<ELEC, WPC, false, true>(c, b, fx, fy, fz, epart, ipart);
    else
        launch_mx_k<ELEC, WPC, false, false>(c, b, fx, fy, fz, epart, ipart);
}

template <int ELEC>
void launch_t(const LaunchCtx& c, const ClusterBufs& b, int precision, double* fx, double* fy,
              double* fz, double* epart, unsigned long long* ipart, int want_hash, double* gq,
              double* ge, double* gs) {
    if (gq)
        launch_64<ELEC, true, 4>(c, b, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
    else if (precision == MSPMS_PREC_FP64 && c.pair_wpc <= 1)
        launch_64<ELEC, false, 1>(c, b, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
    else if (precision == MSPMS_PREC_FP64 && c.pair_wpc == 2)
        launch_64<ELEC, false, 2>(c, b, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
    else if (precision == MSPMS_PREC_FP64 && c.pair_wpc <= 4)
        launch_64<ELEC, false, 4>(c, b, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
    else if (precision == MSPMS_PREC_FP64)
        launch_64<ELEC, false, 8>(c, b, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
    else if (c.pair_wpc <= 1)
        launch_mx<ELEC, 1>(c, b, fx, fy, fz, epart, ipart, want_hash);
    else if (c.pair_wpc == 2)
        launch_mx<ELEC, 2>(c, b, fx, fy, fz, epart, ipart, want_hash);
    else if (c.pair_wpc <= 4)
        launch_mx<ELEC, 4>(c, b, fx, fy, fz, epart, ipart, want_hash);
    else
        launch_mx<ELEC, 8>(c, b, fx, fy, fz, epart, ipart, want_hash);
}

}  // namespace

int pair_partial_count(const ClusterBufs& b) { return b.maxcl; }

// gq/ge/gs non-null: FP64 evaluation that also writes each atom's
// derivative of the pair energy with respect to its own charge, epsilon and
// sigma (the parameter adjoint)
void launch_pair(const LaunchCtx& c, const ClusterBufs& b, int precision, int charge_type,
                 double* fx, double* fy, double* fz, double* epart, unsigned long long* ipart,
                 int want_hash, double* gq, double* ge, double* gs) {
    if (charge_type == MSPMS_ELEC_EWALD)
        launch_t<MSPMS_ELEC_EWALD>(c, b, precision, fx, fy, fz, epart, ipart, want_hash, gq, ge,
                                   gs);
    else if (charge_type == MSPMS_ELEC_WOLF)
        launch_t<MSPMS_ELEC_WOLF>(c, b, precision, fx, fy, fz, epart, ipart, want_hash, gq, ge,
                                  gs);
    else
        launch_t<0>(c, b, precision, fx, fy, fz, epart, ipart, want_hash, gq, ge, gs);
}
