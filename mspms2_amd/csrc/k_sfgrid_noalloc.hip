// k_sfgrid_noalloc.hip -- the solid-fluid grid generator without any
// allocation: mspms_sfgridgen_noalloc stages the solid atoms in a fixed
// device array and writes the tables into the caller's device buffer, for
// callers that cannot allocate device memory.
#include <hip/hip_runtime.h>
#include <mutex>
#include "sfgrid.h"

// The solid atoms' staging buffer (3 x 1000 FP64, 24,000 bytes): one copy
// per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address comes from hipGetSymbolAddress,
// looked up once per device.
static __device__ __align__(256) double scratch[3 * SFG_MAX_SOLID];

namespace {

constexpr int SFG_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SFG_MAX_DEVICES] = {};

// scratch's address on `device`, which must be current
int scratch_on(int device, double** out) {
    if (device < 0 || device >= SFG_MAX_DEVICES)
        return mspms_fail(MSPMS_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        double* p = nullptr;
        if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return mspms_fail(MSPMS_EIO, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

}  // namespace

extern "C" int mspms_sfgridgen_noalloc(int device, const mspms_sfgridgen_spec* s, double* d_out,
                                       void* stream) {
    SfgParams P;
    size_t ntot = 0;
    if (!d_out) return mspms_fail(MSPMS_EINVAL, "null argument");
    // rejects more than 1000 solid atoms (the staging buffer's size) first
    if (int e = sfgridgen_setup(s, &P, &ntot)) return e;
    if (hipSetDevice(device) != hipSuccess) return mspms_fail(MSPMS_EIO, "hipSetDevice");
    double* d_s = nullptr;
    if (int e = scratch_on(device, &d_s)) return e;
    const size_t ns = (size_t)s->s_natom;
    const hipStream_t st = (hipStream_t)stream;
    // every call overwrites the atoms it uses: nothing has to start at zero
    if (hipMemcpyAsync(d_s, s->sx, ns * sizeof(double), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_s + ns, s->sy, ns * sizeof(double), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_s + 2 * ns, s->sz, ns * sizeof(double), hipMemcpyHostToDevice, st))
        return mspms_fail(MSPMS_EIO, "solid atoms upload");
    if (sfgridgen_launch(P, d_s, d_s + ns, d_s + 2 * ns, d_out, ntot, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return mspms_fail(MSPMS_EIO, "grid kernel");
    return 0;
}
