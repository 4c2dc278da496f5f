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
