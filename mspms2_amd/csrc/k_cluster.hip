// k_cluster.hip -- spatial sort of atoms into 8-atom clusters and the
// cluster-pair list (gfx950).
//
// Replaces the O(N^2) molecule-pair loop of src/frclong.c:52-220 as the way
// candidate pairs are found; which pairs interact is still decided per pair
// with the reference's cutoff test (k_pair.hip).
//
// Pipeline (one launch each, all replicas at once):
//   k_list_check  per atom: displacement since the last build against half
//              the skin -> rebuild request; per slot: refresh of the slot
//              positions for a reused list (overwritten when it rebuilds)
// and, only when a rebuild is requested (the others return at once):
//   k_bin_scan per atom: wrap into the box, integer cell (cx,cy,cz), count;
//              the last workgroup: exclusive scans over cells and columns
//              (cells of a column are contiguous, ordered by cz); grids of
//              more than 16,384 cells: k_scan, one workgroup of 1024
//              threads, does the scans
//   k_scatter  per atom: cell member lists
//   k_place_bbox  one workgroup per column: exact rank of each atom inside
//              its cell by (wrapped z, index) -> slot in its column (a column
//              is cut into clusters of 8); then per cluster: bounding box,
//              FP32 offsets from the centre
//   k_plist    one wave per i-cluster: j-clusters whose boxes come within
//              rc_list, compacted with a wave ballot (deterministic order)
// The ranks make the final order independent of atomic arrival order.
//
// Reuse: a list built with rc_list = cutoff + skin holds every pair within
// the cutoff while no atom has moved more than skin/2 since the build (the
// pair's separation at the build was then below rc_list, so the boxes of the
// build met within rc_list, and the image recorded for the cluster pair is
// the pair's minimum image: the column window is narrower than the box, and
// the z image is marked ambiguous when Ei.z + Ej.z + rc_list >= Lz/2).  A
// reused slot keeps its build-time wrapped position and cluster centre; its
// offset becomes (wrapped - centre) + (x - x_build), continuous when the
// atom crosses a box face.
#include "engine.h"
#include "wave.h"
#include "../../include/mspms_gpu.h"

namespace {

__global__ void __launch_bounds__(64) k_list_check(
    int natom, int nslot, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ atom_site,
    const SiteDev* __restrict__ sites, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, const double4* __restrict__ atom_x0,
    const int* __restrict__ ncl_total, int maxcl, const int* __restrict__ slot_orig,
    const double4* __restrict__ slot_w64, const double4* __restrict__ cl_center,
    double4* __restrict__ slot_x64, float4* __restrict__ slot_off, float4* __restrict__ cl_box,
    double* __restrict__ fslot, int* lstate, double half_skin2, int force, int half,
    const HistDev* hist) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // a device run recording histories: this call begins the next step
    if (t == 0 && hist && hist->on) *hist->tick += 1;
    // The kernel is a short chain of dependent loads on a few waves per CU:
    // every load that depends on no other (list state, this thread's atom
    // and slot) is issued before any is used, and only the slot's atom
    // position waits for the slot's atom index.  Slot data read here before
    // the first build are never used (the list is not valid then).
    const bool in_atom = t < natom, in_slot = t < nslot;
    const int ls_valid = lstate[LS_VALID], ls_layout = lstate[LS_LAYOUT];
    const int ncl = min(*ncl_total, maxcl);
    const int site = in_atom ? atom_site[t] : 0;
    const double4 p0t = in_atom ? atom_x0[t] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double xt = in_atom ? x[t] : 0.0, yt = in_atom ? y[t] : 0.0, zt = in_atom ? z[t] : 0.0;
    const int o = in_slot ? slot_orig[t] : -1;
    const double4 w = in_slot ? slot_w64[t] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double4 C = in_slot ? cl_center[t / MSPMS_CL] : make_double4(0.0, 0.0, 0.0, 0.0);
    const bool valid = ls_valid != 0 && !force && ls_layout == half;
    const bool live_cl = valid && in_slot && t < MSPMS_CL * ncl && o >= 0;
    double4 p0o = make_double4(0.0, 0.0, 0.0, 0.0);
    double xo = 0.0, yo = 0.0, zo = 0.0;
    if (live_cl) {
        p0o = atom_x0[o];
        xo = x[o];
        yo = y[o];
        zo = z[o];
    }
    bool need = !valid;
    if (half && in_slot) {
        // the half-list kernel's slot-ordered forces start from zero
        fslot[t] = 0.0;
        fslot[(size_t)nslot + t] = 0.0;
        fslot[2 * (size_t)nslot + t] = 0.0;
    }
    if (in_atom) {
        const bool full_ghost = sites[site].ghost == MSPMS_GHOST_FULL;
        // a full ghost is not binned: its force starts from zero
        // (src/frclong.c:94,107); with one row per cluster pair every pair
        // force arrives by atomic addition, onto zero
        if (full_ghost || half) {
            fx[t] = 0.0;
            fy[t] = 0.0;
            fz[t] = 0.0;
        }
        if (!full_ghost && valid) {
            const double dx = xt - p0t.x, dy = yt - p0t.y, dz = zt - p0t.z;
            need = !(dx * dx + dy * dy + dz * dz <= half_skin2);  // NaN: rebuild
        }
    }
    // the cluster's current box over its valid slots (8 adjacent lanes)
    const float BIG = 3.0e38f;
    float lo[3] = {BIG, BIG, BIG}, hi[3] = {-BIG, -BIG, -BIG};
    if (live_cl) {
        slot_x64[t] = make_double4(xo, yo, zo, 0.0);
        const float4 off = make_float4((float)((w.x - C.x) + (xo - p0o.x)),
                                       (float)((w.y - C.y) + (yo - p0o.y)),
                                       (float)((w.z - C.z) + (zo - p0o.z)), (float)w.w);
        slot_off[t] = off;
        lo[0] = hi[0] = off.x;
        lo[1] = hi[1] = off.y;
        lo[2] = hi[2] = off.z;
    }
    // over the cluster's 8 adjacent lanes: lane ^ 1, lane ^ 2, then the
    // mirror within the 8-lane half row (DPP moves, no LDS crossbar)
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = fminf(lo[d], dpp32<0xB1>(lo[d]));
        hi[d] = fmaxf(hi[d], dpp32<0xB1>(hi[d]));
        lo[d] = fminf(lo[d], dpp32<0x4E>(lo[d]));
        hi[d] = fmaxf(hi[d], dpp32<0x4E>(hi[d]));
        lo[d] = fminf(lo[d], dpp32<0x141>(lo[d]));
        hi[d] = fmaxf(hi[d], dpp32<0x141>(hi[d]));
    }
    // the cluster's first lane writes its box
    if (valid && in_slot && t < MSPMS_CL * ncl && (t & (MSPMS_CL - 1)) == 0) {
        const int c = t / MSPMS_CL;
        if (lo[0] <= hi[0]) {
            cl_box[2 * c] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]),
                                        0.5f * (lo[2] + hi[2]), 0.f);
            cl_box[2 * c + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]),
                                            0.5f * (hi[2] - lo[2]), 0.f);
        } else {  // no valid slot: a point box
            cl_box[2 * c] = make_float4(0.f, 0.f, 0.f, 0.f);
            cl_box[2 * c + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    // one atomic per wave, and none once the request is seen set: atomics
    // on one address from every wave of a large system serialize
    if (__ballot(need) && (threadIdx.x & 63) == 0 &&
        !__hip_atomic_load(&lstate[LS_NEED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicOr(&lstate[LS_NEED], 1);
}

__device__ inline double wrap(double x, double L, double invL) {
    double w = x - L * floor(x * invL);
    if (w >= L) w -= L;
    if (w < 0.0) w += L;
    return w;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline void bin_atom(int i, const double* __restrict__ x, const double* __restrict__ y,
                      const double* __restrict__ z, const int* __restrict__ atom_site,
                      const int* __restrict__ atom_mol, const SiteDev* __restrict__ sites,
                      const MolDev* __restrict__ mols, const RepDev* __restrict__ reps,
                      int* __restrict__ atom_cell, int* __restrict__ atom_slot,
                      double* __restrict__ atom_wz, int* __restrict__ cell_count,
                      double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
                      int* __restrict__ atom_gslot, Status* st) {
    int s = atom_site[i];
    if (sites[s].ghost == MSPMS_GHOST_FULL) {
        // skipped by the pair loops (src/frclong.c:94,107); its force is
        // built from the reciprocal and other per-atom terms only
        atom_cell[i] = -1;
        atom_gslot[i] = -1;
        fx[i] = 0.0;
        fy[i] = 0.0;
        fz[i] = 0.0;
        return;
    }
    const RepDev& R = reps[mols[atom_mol[i]].rep];
    double xi = x[i], yi = y[i], zi = z[i];
    if (!(isfinite(xi) && isfinite(yi) && isfinite(zi))) {
        atomicOr(&st->flags, ST_NONFINITE);
        xi = yi = zi = 0.0;
    }
    double wx = wrap(xi, R.L[0], R.invL[0]);
    double wy = wrap(yi, R.L[1], R.invL[1]);
    double wz = wrap(zi, R.L[2], R.invL[2]);
    int cx = clampi((int)(wx * R.inv_h[0]), 0, R.ncx - 1);
    int cy = clampi((int)(wy * R.inv_h[1]), 0, R.ncy - 1);
    int cz = clampi((int)(wz * R.inv_h[2]), 0, R.ncz - 1);
    int cell = R.cell0 + ((cy * R.ncx + cx) * R.ncz + cz);
    atom_cell[i] = cell;
    atom_wz[i] = wz;
    atom_slot[i] = atomicAdd(&cell_count[cell], 1);
}

// Block-wide exclusive scan helper over 1024 threads.
template <int NT>
__device__ int block_exscan(int v, int* sh, int& total) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) sh[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int w = (lane < NT / 64) ? sh[lane] : 0;
        int wi = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(wi, o, 64);
            if (lane >= o) wi += t;
        }
        if (lane < NT / 64) sh[lane] = wi - w;
        if (lane == NT / 64 - 1) sh[NT / 64] = wi;
    }
    __syncthreads();
    int ex = incl - v + sh[wid];
    total = sh[NT / 64];
    __syncthreads();
    return ex;
}

// Cells per pass of scan_cells: NT * SCAN_PER<NT> counts staged in LDS.
template <int NT>
constexpr int SCAN_PER = NT >= 1024 ? 8 : 16;
// LDS index of staged cell q: one pad word per 32, so that the PER
// consecutive words a thread reads fall in different banks across the wave
__device__ inline int scan_pad(int q) { return q + (q >> 5); }
template <int NT>
constexpr int SCAN_LDS = NT * SCAN_PER<NT> + NT * SCAN_PER<NT> / 32;

// By one workgroup of NT threads: cell_start = exscan(cell_count),
// cell_count := 0; per column the atom count, cluster count, cluster start
// and cl_col.  stage: SCAN_LDS<NT> ints of LDS.
template <int NT>
__device__ void scan_cells(int ncell, int ncol, int* __restrict__ cell_count,
                           int* __restrict__ cell_start, const int* __restrict__ col_cell0,
                           int* __restrict__ col_cl0, int* __restrict__ cl_col,
                           int* __restrict__ ncl_total, int maxcl, Status* st, int* sh,
                           int* stage) {
    int carry = 0;
    // per pass NT * PER cells: read into LDS with coalesced loads, scanned
    // PER consecutive cells per thread, written back with coalesced stores
    // (a tube of 16 long boxes has ~200,000 cells)
    constexpr int PER = SCAN_PER<NT>;
    const int t = threadIdx.x;
    for (int base = 0; base < ncell; base += NT * PER) {
        const int n = min(NT * PER, ncell - base);
        for (int q = t; q < NT * PER; q += NT) stage[scan_pad(q)] = q < n ? cell_count[base + q] : 0;
        __syncthreads();
        int v[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            v[k] = stage[scan_pad(t * PER + k)];
            sum += v[k];
        }
        int tot;
        int run = carry + block_exscan<NT>(sum, sh, tot);  // its barriers order the reads
#pragma unroll
        for (int k = 0; k < PER; k++) {
            stage[scan_pad(t * PER + k)] = run;
            run += v[k];
        }
        __syncthreads();
        for (int q = t; q < n; q += NT) {
            cell_start[base + q] = stage[scan_pad(q)];
            cell_count[base + q] = 0;
        }
        __syncthreads();  // the next pass refills the stage
        carry += tot;
    }
    if (threadIdx.x == 0) cell_start[ncell] = carry;
    __syncthreads();
    carry = 0;
    for (int base = 0; base < ncol; base += NT) {
        int g = base + threadIdx.x;
        int n = 0;
        if (g < ncol) {
            int cnt = cell_start[col_cell0[g + 1]] - cell_start[col_cell0[g]];
            n = (cnt + MSPMS_CL - 1) / MSPMS_CL;
        }
        int tot;
        int ex = block_exscan<NT>(n, sh, tot);
        if (g < ncol) col_cl0[g] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        col_cl0[ncol] = carry;
        *ncl_total = carry;
        if (carry > maxcl) atomicOr(&st->flags, ST_OVERFLOW);
    }
    __syncthreads();
    // cluster -> column map
    for (int g = threadIdx.x; g < ncol; g += NT) {
        int a = col_cl0[g], b = col_cl0[g + 1];
        for (int c = a; c < b && c < maxcl; c++) cl_col[c] = g;
    }
}

// Per atom: wrap into the box, integer cell, count (bin_atom).  The last
// workgroup to finish (arrival count) scans the cells and columns
// (scan_cells): one launch instead of two.
__global__ void __launch_bounds__(256) k_bin_scan(
    int natom, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ atom_site,
    const int* __restrict__ atom_mol, const SiteDev* __restrict__ sites,
    const MolDev* __restrict__ mols, const RepDev* __restrict__ reps, int* __restrict__ atom_cell,
    int* __restrict__ atom_slot, double* __restrict__ atom_wz, int* cell_count,
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    int* __restrict__ atom_gslot, Status* st, const int* lstate, int ncell, int ncol,
    int* __restrict__ cell_start, const int* __restrict__ col_cell0, int* __restrict__ col_cl0,
    int* __restrict__ cl_col, int* __restrict__ ncl_total, int maxcl, int* ctr) {
    if (!lstate[LS_NEED]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < natom)
        bin_atom(i, x, y, z, atom_site, atom_mol, sites, mols, reps, atom_cell, atom_slot, atom_wz,
                 cell_count, fx, fy, fz, atom_gslot, st);
    if (!ctr) return;  // the cells are scanned by k_scan
    // arrival: this workgroup's counts are visible device-wide before its
    // count
    __shared__ int last;
    __shared__ int sh[32];
    __shared__ int stage[SCAN_LDS<256>];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ctr, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) *ctr = 0;  // for the next build
    scan_cells<256>(ncell, ncol, cell_count, cell_start, col_cell0, col_cl0, cl_col, ncl_total,
                    maxcl, st, sh, stage);
}

// The cell and column scans by one workgroup of 1024 threads, for grids of
// many cells (a workgroup of 256 at the end of k_bin_scan would take long).
__global__ void __launch_bounds__(1024) k_scan(int ncell, int ncol, int* cell_count,
                                               int* __restrict__ cell_start,
                                               const int* __restrict__ col_cell0,
                                               int* __restrict__ col_cl0, int* __restrict__ cl_col,
                                               int* __restrict__ ncl_total, int maxcl, Status* st,
                                               const int* lstate) {
    if (!lstate[LS_NEED]) return;
    __shared__ int sh[32];
    __shared__ int stage[SCAN_LDS<1024>];
    scan_cells<1024>(ncell, ncol, cell_count, cell_start, col_cell0, col_cl0, cl_col, ncl_total,
                     maxcl, st, sh, stage);
}

__global__ void k_scatter(int natom, const int* __restrict__ atom_cell,
                          const int* __restrict__ atom_slot, const int* __restrict__ cell_start,
                          int* __restrict__ cell_members, const int* lstate) {
    if (!lstate[LS_NEED]) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natom) return;
    int c = atom_cell[i];
    if (c < 0) return;
    cell_members[cell_start[c] + atom_slot[i]] = i;
}

// Rank inside the cell by (wrapped z, atom index); place into the column's
// slot range, which is 8 * (first cluster of the column) onwards.
__device__ inline void place_atom(int i, int col, int colstart, const double* __restrict__ x,
                                  const double* __restrict__ y, const double* __restrict__ z,
                                  const int* __restrict__ atom_cell,
                                  const double* __restrict__ atom_wz,
                                  const int* __restrict__ cell_start,
                                  const int* __restrict__ cell_members,
                                  const int* __restrict__ atom_mol,
                                  const int* __restrict__ atom_site,
                                  const MolDev* __restrict__ mols, const RepDev* __restrict__ reps,
                                  const int* __restrict__ col_cl0, int* __restrict__ slot_orig,
                                  int2* __restrict__ slot_info, double4* __restrict__ slot_x64,
                                  double4* slot_w64, double4* __restrict__ atom_x0,
                                  int* __restrict__ atom_gslot, const SiteDev* __restrict__ sites,
                                  int maxcl) {
    const int c = atom_cell[i];
    const RepDev& R = reps[mols[atom_mol[i]].rep];
    const double wz = atom_wz[i];
    const int a = cell_start[c], b = cell_start[c + 1];
    int rank = 0;
    for (int m = a; m < b; m++) {
        const int j = cell_members[m];
        const double wj = atom_wz[j];
        rank += (wj < wz) || (wj == wz && j < i);
    }
    const int slot = MSPMS_CL * col_cl0[col] + (a - colstart) + rank;
    if (slot >= MSPMS_CL * maxcl) {  // overflow flagged by the scan
        atom_gslot[i] = -1;
        return;
    }
    atom_gslot[i] = slot;
    double xi = x[i], yi = y[i], zi = z[i];
    if (!(isfinite(xi) && isfinite(yi) && isfinite(zi))) xi = yi = zi = 0.0;
    slot_orig[slot] = i;
    slot_info[slot] = make_int2(atom_mol[i], atom_site[i]);
    slot_x64[slot] = make_double4(x[i], y[i], z[i], 0.0);
    slot_w64[slot] = make_double4(wrap(xi, R.L[0], R.invL[0]), wrap(yi, R.L[1], R.invL[1]), wz,
                                  (double)sites[atom_site[i]].charge_f);
    atom_x0[i] = make_double4(x[i], y[i], z[i], 0.0);
}

// 8 lanes per cluster (slot t = 8 c + k; every lane of the wave takes part
// in the shuffles; c < c_end live): bounding box, centre, FP32 offsets;
// pads marked.
__device__ inline void bbox_slot(int t, int c_end, int col, int cnt, int cl0,
                                 const int* __restrict__ col_rep,
                                 const SiteDev* __restrict__ sites, int* __restrict__ slot_orig,
                                 int2* __restrict__ slot_info, const double4* slot_w64,
                                 float4* __restrict__ slot_off, float4* __restrict__ slot_par,
                                 double4* __restrict__ cl_center, float4* __restrict__ cl_ext,
                                 float4* __restrict__ cl_box) {
    const int c = t >> 3, k = t & 7;
    const bool live = c < c_end;
    const int in_cl = live ? min(MSPMS_CL, cnt - MSPMS_CL * (c - cl0)) : 0;
    const bool valid = live && k < in_cl;
    const int slot = MSPMS_CL * c + k;
    const double4 w = valid ? slot_w64[slot] : make_double4(0, 0, 0, 0);
    double lo[3], hi[3];
    const double BIG = 1e300;
    lo[0] = valid ? w.x : BIG;
    lo[1] = valid ? w.y : BIG;
    lo[2] = valid ? w.z : BIG;
    hi[0] = valid ? w.x : -BIG;
    hi[1] = valid ? w.y : -BIG;
    hi[2] = valid ? w.z : -BIG;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    }
    // the half extents, rounded up (every lane of the cluster holds them)
    const float4 ext = make_float4((float)(0.5 * (hi[0] - lo[0])) * 1.000001f + 1e-6f,
                                   (float)(0.5 * (hi[1] - lo[1])) * 1.000001f + 1e-6f,
                                   (float)(0.5 * (hi[2] - lo[2])) * 1.000001f + 1e-6f, 0.f);
    if (!live) return;
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]),
                 cz = 0.5 * (lo[2] + hi[2]);
    if (k == 0) {
        cl_center[c] = make_double4(cx, cy, cz, (double)col_rep[col]);
        cl_ext[c] = ext;
        cl_box[2 * c] = make_float4(0.f, 0.f, 0.f, 0.f);
        cl_box[2 * c + 1] = ext;
    }
    if (valid) {
        const int s = slot_info[slot].y;
        const int mol = slot_info[slot].x;
        const SiteDev& S = sites[s];
        const bool lj = S.ghost != MSPMS_GHOST_LJ;
        slot_off[slot] = make_float4((float)(w.x - cx), (float)(w.y - cy), (float)(w.z - cz),
                                     S.charge_f);
        slot_par[slot] = make_float4(S.sigma_f, lj ? S.sqrt_eps_f : 0.f, __int_as_float(mol),
                                     __int_as_float(SLOT_VALID | (lj ? SLOT_LJ : 0)));
    } else {
        slot_orig[slot] = -1;
        slot_info[slot] = make_int2(-1, -1);
        slot_off[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        slot_par[slot] = make_float4(0.f, 0.f, __int_as_float(-1), __int_as_float(0));
    }
}

// One workgroup per column: its atoms (the member lists of its cells, which
// are contiguous) take their slots (place_atom), then, after a workgroup
// barrier, its clusters take their boxes (bbox_slot).  A column's slots and
// clusters are its own, so no other workgroup's writes are read.
__global__ void __launch_bounds__(1024) k_place_bbox(
    int ncol, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ atom_cell,
    const double* __restrict__ atom_wz, const int* __restrict__ cell_start,
    const int* __restrict__ cell_members, const int* __restrict__ atom_mol,
    const int* __restrict__ atom_site, const MolDev* __restrict__ mols,
    const RepDev* __restrict__ reps, const int* __restrict__ col_cell0,
    const int* __restrict__ col_cl0, const int* __restrict__ col_rep,
    const int* __restrict__ ncl_total, int* __restrict__ slot_orig, int2* __restrict__ slot_info,
    double4* __restrict__ slot_x64, double4* slot_w64, double4* __restrict__ atom_x0,
    int* __restrict__ atom_gslot, const SiteDev* __restrict__ sites, float4* __restrict__ slot_off,
    float4* __restrict__ slot_par, double4* __restrict__ cl_center, float4* __restrict__ cl_ext,
    float4* __restrict__ cl_box, int maxcl, int* lstate) {
    if (!lstate[LS_NEED]) return;
    const int col = blockIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // usable for reuse unless the clusters overflowed (k_plist clears it
        // again on a row overflow)
        lstate[LS_VALID] = *ncl_total <= maxcl;
        lstate[LS_BUILDS] += 1;
    }
    const int m0 = cell_start[col_cell0[col]], m1 = cell_start[col_cell0[col + 1]];
    for (int m = m0 + threadIdx.x; m < m1; m += blockDim.x)
        place_atom(cell_members[m], col, m0, x, y, z, atom_cell, atom_wz, cell_start,
                   cell_members, atom_mol, atom_site, mols, reps, col_cl0, slot_orig, slot_info,
                   slot_x64, slot_w64, atom_x0, atom_gslot, sites, maxcl);
    __syncthreads();
    const int cl0 = col_cl0[col], c_end = min(col_cl0[col + 1], maxcl);
    // whole waves per pass, so the shuffles see all 64 lanes
    for (int t0 = MSPMS_CL * cl0 + (threadIdx.x & ~63); t0 < MSPMS_CL * c_end; t0 += blockDim.x)
        bbox_slot(t0 + (threadIdx.x & 63), c_end, col, m1 - m0, cl0, col_rep, sites, slot_orig,
                  slot_info, slot_w64, slot_off, slot_par, cl_center, cl_ext, cl_box);
}

// One wave per i-cluster.  Candidate columns cover the x/y range of the
// i-box widened by rc_list (periodic, each wrapped column at most once).
// A column's clusters are in wrapped-z order (cells by cz, atoms in a cell
// by wrapped z), so the j-boxes that can come within rc_list of the i-box
// lie in the cells that meet [Ci.z - Ei.z - rl, Ci.z + Ei.z + rl] (modulo
// Lz, with a small margin): their clusters, plus the one that straddles the
// window if it holds no atom, are the candidates (one or two
// index ranges per column, visited in ascending order).  A cluster whose z
// image is ambiguous (Ei.z + Ej.z + rl >= Lz/2) spans at least Lz minus the
// window, so its atoms meet the window too.  A window as long as the box
// (Ei.z + rl >= Lz/2) takes whole columns.  Each candidate is tested
// box-to-box; the window only drops clusters that fail that test, so the
// row is the same as with whole columns.  Entry layout:
//   bits 0-1 x image + 1, 2-3 y image + 1, 4-5 z image + 1,
//   bit 6 "ambiguous image: apply the minimum image per pair",
//   bits 7-31 j-cluster index,
// stored in .w of a float4 whose x, y, z hold the FP32 centre separation.
__global__ void __launch_bounds__(256) k_plist(const int* __restrict__ ncl_total, int maxcl,
                                               const int* __restrict__ cl_col,
                                               const int* __restrict__ col_cl0,
                                               const int* __restrict__ col_rep,
                                               const int* __restrict__ col_cell0,
                                               const int* __restrict__ cell_start,
                                               const RepDev* __restrict__ reps,
                                               const double4* __restrict__ cl_center,
                                               const float4* __restrict__ cl_ext,
                                               float4* __restrict__ plist,
                                               int* __restrict__ plist_n, int cap, Status* st,
                                               int* lstate, int half) {
    if (!lstate[LS_NEED]) return;
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    int ncl = min(*ncl_total, maxcl);
    if (wave >= ncl) return;
    int ci = wave;
    double4 Ci = cl_center[ci];
    float4 Ei = cl_ext[ci];
    const RepDev& R = reps[(int)Ci.w];
    const double rl = R.rc_list;
    int colx_lo = (int)floor((Ci.x - Ei.x - rl) * R.inv_h[0]);
    int colx_hi = (int)floor((Ci.x + Ei.x + rl) * R.inv_h[0]);
    int coly_lo = (int)floor((Ci.y - Ei.y - rl) * R.inv_h[1]);
    int coly_hi = (int)floor((Ci.y + Ei.y + rl) * R.inv_h[1]);
    // a range as wide as the box visits each wrapped column once; the image
    // is then not fixed per column and the pair kernel takes it per pair
    bool ambx = false, amby = false;
    if (colx_hi - colx_lo + 1 >= R.ncx) {
        colx_lo = 0;
        colx_hi = R.ncx - 1;
        ambx = true;
    }
    if (coly_hi - coly_lo + 1 >= R.ncy) {
        coly_lo = 0;
        coly_hi = R.ncy - 1;
        amby = true;
    }
    const double rl2 = rl * rl;
    // the z window in cells: [p1, q1], and [p2, ncz - 1] when it wraps.  An
    // atom's cell is floor(wz / h) (bin_atom), monotone in wz, so the cells of
    // the window's ends bound the cells of the atoms inside it; the margin mz
    // covers the FP32 widening of the cluster extents and the rounding of
    // the window's ends.
    const int ncz = R.ncz;
    const double Lz = R.L[2], ihz = R.inv_h[2], mz = 1e-5 * (Lz + rl + 1.0);
    const double zlo = Ci.z - (double)Ei.z - rl - mz, zhi = Ci.z + (double)Ei.z + rl + mz;
    bool zwin = (double)Ei.z + rl + mz < 0.5 * Lz;
    int p1 = 0, q1 = ncz - 1, p2 = ncz;
    if (zwin) {
        if (zlo < 0.0) {  // [0, zhi] and [zlo + Lz, Lz)
            q1 = (int)floor(zhi * ihz);
            p2 = (int)floor((zlo + Lz) * ihz);
        } else if (zhi >= Lz) {  // [0, zhi - Lz] and [zlo, Lz)
            q1 = (int)floor((zhi - Lz) * ihz);
            p2 = (int)floor(zlo * ihz);
        } else {
            p1 = clampi((int)floor(zlo * ihz), 0, ncz - 1);
            q1 = (int)floor(zhi * ihz);
        }
        q1 = clampi(q1, p1, ncz - 1);
        p2 = clampi(p2, 0, ncz);
        if (p2 <= q1 + 1) zwin = false;  // the two ranges meet: whole column
    }
    if (!zwin) {
        p1 = 0;
        q1 = ncz - 1;
        p2 = ncz;
    }
    int n = 0;
    float4* row = plist + (size_t)ci * cap;
    // candidate columns in (y, x) order, 64 per batch; their clusters are
    // flattened so that each pass of the wave tests 64 clusters whatever the
    // column heights (a column holds ~ncz clusters, far fewer than 64)
    const int ncx_c = colx_hi - colx_lo + 1, ncand = ncx_c * (coly_hi - coly_lo + 1);
    for (int cb = 0; cb < ncand; cb += 64) {
        const int nb = min(64, ncand - cb);
        // this lane's column: clusters [a, a + n1) and [a2, a2 + cnt - n1)
        int a = 0, a2 = 0, n1 = 0, cnt = 0, sx = 0, sy = 0;
        if (lane < nb) {
            const int q = cb + lane;
            const int cxu = colx_lo + q % ncx_c, cyu = coly_lo + q / ncx_c;
            const int cxw = ((cxu % R.ncx) + R.ncx) % R.ncx;
            const int cyw = ((cyu % R.ncy) + R.ncy) % R.ncy;
            sx = (cxu - cxw) / R.ncx;
            sy = (cyu - cyw) / R.ncy;
            const int col = R.col0 + cyw * R.ncx + cxw;
            // clusters past maxcl (an overflowed build) do not exist
            const int k0 = min(col_cl0[col], ncl), k1 = min(col_cl0[col + 1], ncl);
            if (zwin) {
                // cluster of the atom at column position m: k0 + m / 8
                const int c0 = col_cell0[col], m0 = cell_start[c0];
                a = min(col_cl0[col] + ((cell_start[c0 + p1] - m0) >> 3), k1);
                const int e1 = min(col_cl0[col] + ((cell_start[c0 + q1 + 1] - m0 + 7) >> 3), k1);
                n1 = max(e1 - a, 0);
                int e2 = e1;
                a2 = e1;
                if (p2 < ncz) {
                    a2 = max(min(col_cl0[col] + ((cell_start[c0 + p2] - m0) >> 3), k1), e1);
                    e2 = max(k1, a2);
                }
                cnt = n1 + (e2 - a2);
            } else {
                a = k0;
                n1 = max(k1 - k0, 0);
                a2 = k0 + n1;
                cnt = n1;
            }
        }
        // inclusive prefix of the column sizes over the lanes
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        for (int base = 0; base < total; base += 64) {
            const int f = base + lane;
            // column of flattened index f: number of columns with incl <= f
            int qc = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const int probe = qc + step - 1;
                const int v = __shfl(incl, probe & 63, 64);
                if (probe < nb && v <= f) qc += step;
            }
            const int qa = __shfl(a, qc & 63, 64), qcnt = __shfl(cnt, qc & 63, 64);
            const int qa2 = __shfl(a2, qc & 63, 64), qn1 = __shfl(n1, qc & 63, 64);
            const int qx = __shfl(sx, qc & 63, 64), qy = __shfl(sy, qc & 63, 64);
            const int qincl = __shfl(incl, qc & 63, 64);
            bool hit = false;
            uint32_t entry = 0;
            float4 sep = make_float4(0.f, 0.f, 0.f, 0.f);
            if (f < total) {
                const int o = f - (qincl - qcnt);
                const int cj = o < qn1 ? qa + o : qa2 + (o - qn1);
                double4 Cj = cl_center[cj];
                float4 Ej = cl_ext[cj];
                double dxc = Ci.x - (Cj.x + qx * R.L[0]);
                double dyc = Ci.y - (Cj.y + qy * R.L[1]);
                double dzc = Ci.z - Cj.z;
                double nz = rint(dzc * R.invL[2]);
                dzc -= nz * R.L[2];
                bool ambz = (double)Ei.z + (double)Ej.z + rl >= 0.5 * R.L[2];
                double gx = ambx ? 0.0 : fmax(0.0, fabs(dxc) - ((double)Ei.x + (double)Ej.x));
                double gy = amby ? 0.0 : fmax(0.0, fabs(dyc) - ((double)Ei.y + (double)Ej.y));
                double gz = ambz ? 0.0 : fmax(0.0, fabs(dzc) - ((double)Ei.z + (double)Ej.z));
                hit = gx * gx + gy * gy + gz * gz < rl2;
                // one row per cluster pair: same index parity in the lower
                // row, different parity in the higher (rows stay balanced)
                if (half && cj != ci && ((((ci ^ cj) & 1) == 0) == (ci > cj))) hit = false;
                int szi = (int)nz;  // j image: Cj + (sx, sy, sz) * L
                bool a2 = ambx || amby || ambz || qx < -1 || qx > 1 || qy < -1 || qy > 1 ||
                          szi < -1 || szi > 1;
                if (a2) szi = 0;
                entry = ((uint32_t)cj << 7) | ((a2 ? 1u : 0u) << 6) |
                        ((uint32_t)(szi + 1) << 4) |
                        ((uint32_t)((a2 ? 0 : qy) + 1) << 2) | (uint32_t)((a2 ? 0 : qx) + 1);
                // the centre separation the mixed kernel uses, rounded once
                // to FP32; the centres stay fixed while the list is reused
                if (a2) {
                    sep = make_float4((float)(Ci.x - Cj.x), (float)(Ci.y - Cj.y),
                                      (float)(Ci.z - Cj.z), 0.f);
                } else {
                    sep = make_float4((float)(Ci.x - (Cj.x + qx * R.L[0])),
                                      (float)(Ci.y - (Cj.y + qy * R.L[1])),
                                      (float)(Ci.z - (Cj.z + szi * R.L[2])), 0.f);
                }
                sep.w = __uint_as_float(entry);
            }
            uint64_t mask = __ballot(hit);
            int pre = __popcll(mask & ((1ull << lane) - 1ull));
            if (hit && n + pre < cap) row[n + pre] = sep;
            n += __popcll(mask);
        }
    }
    if (lane == 0) {
        if (ci == 0) lstate[LS_LAYOUT] = half;
        plist_n[ci] = min(n, cap);
        if (n > cap) {
            atomicOr(&st->flags, ST_OVERFLOW);
            lstate[LS_VALID] = 0;  // truncated row: never reuse
        }
        // the largest row: an atomic only when this row beats the largest
        // seen so far (every row's atomic on one address serialized the
        // build of a 120,000-row table)
        if (n > __hip_atomic_load(&st->max_list, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&st->max_list, n);
    }
}

}  // namespace

void launch_cluster_build(const LaunchCtx& c, ClusterBufs& b, const double* x, const double* y,
                          const double* z, double* fx, double* fy, double* fz, double half_skin2,
                          int force) {
    const int T = 256;
    int gA = (c.natom + T - 1) / T;
    const int nslot = b.maxcl * MSPMS_CL;
    // one-wave workgroups: the check is a short chain of dependent loads,
    // which more, smaller workgroups spread over more CUs (0.9 us per step)
    hipLaunchKernelGGL(k_list_check, dim3((std::max(c.natom, nslot) + 63) / 64), dim3(64), 0,
                       c.stream, c.natom, nslot, x, y, z, c.atom_site, c.sites, fx, fy, fz,
                       b.atom_x0, b.ncl_total, b.maxcl, b.slot_orig, b.slot_w64, b.cl_center,
                       b.slot_x64, b.slot_off, b.cl_box, b.fslot, b.lstate, half_skin2, force,
                       b.half, b.hist);
    // up to 16,384 cells the last binning workgroup scans them (one launch
    // less per step); larger grids take a workgroup of 1024 threads
    const bool scan_here = b.ncell <= 16384;
    hipLaunchKernelGGL(k_bin_scan, dim3(gA), dim3(T), 0, c.stream, c.natom, x, y, z, c.atom_site,
                       c.atom_mol, c.sites, c.mols, c.reps, b.atom_cell, b.atom_slot, b.atom_wz,
                       b.cell_count, fx, fy, fz, b.atom_gslot, c.status, b.lstate, b.ncell,
                       b.ncol, b.cell_start, b.col_cell0, b.col_cl0, b.cl_col, b.ncl_total,
                       b.maxcl, scan_here ? b.bin_ctr : nullptr);
    if (!scan_here)
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c.stream, b.ncell, b.ncol, b.cell_count,
                           b.cell_start, b.col_cell0, b.col_cl0, b.cl_col, b.ncl_total, b.maxcl,
                           c.status, b.lstate);
    hipLaunchKernelGGL(k_scatter, dim3(gA), dim3(T), 0, c.stream, c.natom, b.atom_cell,
                       b.atom_slot, b.cell_start, b.cell_members, b.lstate);
    // 1024 threads: a column of a long box holds thousands of atoms
    hipLaunchKernelGGL(k_place_bbox, dim3(b.ncol), dim3(1024), 0, c.stream, b.ncol, x, y, z,
                       b.atom_cell, b.atom_wz, b.cell_start, b.cell_members, c.atom_mol,
                       c.atom_site, c.mols, c.reps, b.col_cell0, b.col_cl0, b.col_rep,
                       b.ncl_total, b.slot_orig, b.slot_info, b.slot_x64, b.slot_w64, b.atom_x0,
                       b.atom_gslot, c.sites, b.slot_off, b.slot_par, b.cl_center, b.cl_ext,
                       b.cl_box, b.maxcl, b.lstate);
    int gP = (b.maxcl * 64 + T - 1) / T;
    hipLaunchKernelGGL(k_plist, dim3(gP), dim3(T), 0, c.stream, b.ncl_total, b.maxcl, b.cl_col,
                       b.col_cl0, b.col_rep, b.col_cell0, b.cell_start, c.reps, b.cl_center,
                       b.cl_ext, b.plist, b.plist_n, b.cap, c.status, b.lstate, b.half);
}
