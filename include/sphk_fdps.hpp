// sphk_fdps.hpp -- the FDPS multiwalk functor API of the reference, served by
// libsphk.so.
//
// These are exactly the declarations of sph_yamauchi/run/use_gpu.hpp:16-64:
// FDPS's calcForceAllMultiWalk / calcForceAllAndWriteBackMultiWalk /
// calcForceMakingTreeMultiWalk call the Dispatch/Retrieve pairs below with
// tag_max = 1 and n_walk_limit = 200 (hdr_density.hpp:387-394,
// main.cpp:337-344, hdr_gravity.hpp:390-397).  A program that built against
// the reference's CUDA objects links libsphk.so instead; nothing on the
// caller's side changes.
//
// The caller owns the class definitions (hdr_density.hpp, hdr_hydro.hpp,
// hdr_gravity.hpp); libsphk.so reads them through layout-identical
// definitions (include/sphk_fdps_types.hpp), checked with static_asserts:
//   DensityEPI 80 B, DensityEPJ 64 B, Density 56 B, HydroEPI 144 B,
//   HydroEPJ 152 B, Hydro 72 B, GravityEPI 32 B, GravityEPJ 40 B,
//   PS::GravitySPJ 40 B, Gravity 40 B.
//
// Behaviour differences from the reference's CUDA build (documented, not
// hidden): precision is chosen at run time by the environment variable
// SPHK_PRECISION = "mixed" (default) | "fp64" instead of the USE_FLOAT /
// USE_DOUBLE build switch (Makefile:80-85); the "fp64" path is FP64 in every
// operation (the reference's double build still used rsqrtf/powf); errors
// print a message and abort, as the reference's assert/checkCudaErrors do.
// RP::KernelSupportRadiusMaximum is read from the caller's program when it
// defines RunParameter::KernelSupportRadiusMaximum (a weak reference), as
// hdr_densityGPU.cu:492-494,581 does; otherwise the value given to
// sphk_set_params on the adapter's context is used.
#pragma once

#include "sphk.h"

class DensityEPI;
class DensityEPJ;
class Density;
class HydroEPI;
class HydroEPJ;
class Hydro;
class GravityEPI;
class GravityEPJ;
namespace ParticleSimulator {
class GravitySPJ;
}
class Gravity;

void setup_gpu(int my_proc);
void reset_gpu(int my_proc);

int DispatchKernel_Dens(const int tag, const int n_walk, const DensityEPI **epi, const int *nip,
                        const DensityEPJ **epj, const int *njp);
int RetrieveKernel_Dens(const int tag, const int n_walk, const int *ni, Density **density);

int DispatchKernel_HydroForce(const int tag, const int n_walk, const HydroEPI **epi, const int *nip,
                              const HydroEPJ **epj, const int *njp);
int RetrieveKernel_HydroForce(const int tag, const int n_walk, const int *ni, Hydro **hydro);

int DispatchKernel_Gravity(const int tag, const int n_walk, const GravityEPI **epi, const int *nip,
                           const GravityEPJ **epj, const int *njp,
                           const ParticleSimulator::GravitySPJ **spj, const int *nsp);
int RetrieveKernel_Gravity(const int tag, const int n_walk, const int *ni, Gravity **gravity);

// The adapter's library context (created by setup_gpu), for callers that
// want to set parameters or read kernel timings through include/sphk.h.
extern "C" sphk_ctx *sphk_fdps_context(void);
