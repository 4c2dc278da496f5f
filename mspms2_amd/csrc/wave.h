// wave.h -- wave64 butterfly reductions shared by the kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

// v[k] := sum of v[k] over the lanes of each aligned group of 2*from lanes
// (from = 32: the whole wave), every lane getting its group's sum.  Each
// step issues the shuffles of all NV values before the first add, so that
// their LDS-path latencies overlap; the order is fixed, so the sums are
// deterministic.
template <typename T, int NV>
__device__ inline void butterfly(T (&v)[NV], int from) {
#pragma unroll
    for (int o = from; o >= 1; o >>= 1) {
        T t[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) t[k] = __shfl_xor(v[k], o, 64);
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += t[k];
    }
}

// A 64-bit value moved between lanes by a DPP pattern (all lanes active;
// row and bank masks full), as two 32-bit moves.
template <int CTRL, typename T>
__device__ inline T dpp64(T v) {
    static_assert(sizeof(T) == 8, "64-bit values");
    union {
        T v;
        int h[2];
    } a, b;
    a.v = v;
    b.h[0] = __builtin_amdgcn_mov_dpp(a.h[0], CTRL, 0xF, 0xF, true);
    b.h[1] = __builtin_amdgcn_mov_dpp(a.h[1], CTRL, 0xF, 0xF, true);
    return b.v;
}

template <typename T>
__device__ inline T readlane64(T v, int lane) {
    union {
        T v;
        int h[2];
    } a, b;
    a.v = v;
    b.h[0] = __builtin_amdgcn_readlane(a.h[0], lane);
    b.h[1] = __builtin_amdgcn_readlane(a.h[1], lane);
    return b.v;
}

// Sum of v over the 64 lanes of a wave, in a fixed order: DPP steps within
// each row of 16 (quad_perm [1,0,3,2], quad_perm [2,3,0,1], half-row mirror,
// row mirror: every lane then holds its row's sum), then the four row sums
// as (r0 + r1) + (r2 + r3).  Short-latency VALU moves, where a shuffle
// butterfly goes through the LDS crossbar at each of its six steps.
template <typename T>
__device__ inline T wave_sum64(T v) {
    v += dpp64<0xB1>(v);
    v += dpp64<0x4E>(v);
    v += dpp64<0x141>(v);
    v += dpp64<0x140>(v);
    return (readlane64(v, 0) + readlane64(v, 16)) + (readlane64(v, 32) + readlane64(v, 48));
}

// Sums of NV values over the NT threads of a workgroup (NT a multiple of 64),
// in a fixed order: each thread's values go to LDS as rows of NT (rows:
// NV * NT); wave w sums rows w, w + NT/64, ..., a lane adding its NT/64
// entries of the row in order before the wave sum; the totals land in
// tot[NV].  Every wave reduces whole rows, so the work is spread over the
// waves instead of every wave reducing every value.  Ends with a barrier.
template <typename T, int NV, int NT>
__device__ void block_sum_rows(const T (&v)[NV], T* rows, T* tot) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) rows[k * NT + t] = v[k];
    __syncthreads();
    for (int k = w; k < NV; k += NT / 64) {
        T s = rows[k * NT + lane];
#pragma unroll
        for (int j = 64; j < NT; j += 64) s += rows[k * NT + j + lane];
        s = wave_sum64(s);
        if (lane == 0) tot[k] = s;
    }
    __syncthreads();
}

// A 32-bit float moved between lanes by a DPP pattern (all lanes active;
// row and bank masks full).
template <int CTRL>
__device__ inline float dpp32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
