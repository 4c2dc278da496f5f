// sfgrid.h -- the solid-fluid grid generator's parameters and launcher
// (k_sfgrid.hip), shared by its entry points.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "../../include/mspms_gpu.h"

constexpr int SFG_MAX_SOLID = 1000;  // solid_atom_max, tools/sfgridgen.c:26

// tools/sfgridgen.c:121-155 in the generator's units (Angstrom, J/mol)
struct SfgParams {
    double x0, y0, z0, itvl;
    double sigmasf, epsilonsf, uclz;
    double rcutonsq, rcutoffsq, roff2_minus_ron2_cube;
    int nx, ny, nz, nuc_z, nsolid, lj_switch;
};

// the engine's error reporter (engine.cpp): records the message, returns code
int mspms_fail(int code, const char* what);
// validates a spec and derives its parameters and node count; 0 or an
// MSPMS_* code (recorded)
int sfgridgen_setup(const mspms_sfgridgen_spec* s, SfgParams* P, size_t* ntot);
// one thread per node on `st`: the 13 tables of every node into out
// (13 x ntot doubles, table-major) from device solid atoms sx, sy, sz
hipError_t sfgridgen_launch(const SfgParams& P, const double* sx, const double* sy,
                            const double* sz, double* out, size_t ntot, hipStream_t st);
