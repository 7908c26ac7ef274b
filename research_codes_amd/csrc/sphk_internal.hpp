// sphk_internal.hpp -- declarations shared between the kernel and runtime
// translation units of libsphk.so (not part of the public C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "sphk.h"

namespace sphk {

// Walk queue of a persistent launch: kXcd claim counters and one exit
// counter, each kQueueStride ints (128 bytes) apart; zero before the launch,
// zero again after it (the kernel's last workgroup resets it).
constexpr int kQueueStride = 32;
constexpr int kQueueInts = 9 * kQueueStride;

// queue == nullptr: one workgroup per walk; else a persistent launch that
// takes its walks from that queue.
hipError_t launch_density(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_dens_src &pi, const sphk_dens_src &pj,
                          const sphk_dens_out &out, double eta_ksrh, double hmax, bool mixed, int list_cap,
                          int *queue);

hipError_t launch_hydro(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                        const int32_t *j_idx, const sphk_hydro_src &pi, const sphk_hydro_src &pj,
                        const sphk_hydro_out &out, bool mixed, int *queue);

hipError_t launch_gravity(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_grav_src &pj, const int32_t *s_off,
                          const int32_t *s_idx, const sphk_grav_src &ps, const sphk_grav_src &pi,
                          const sphk_grav_out &out, bool mixed, bool uniform, int *queue);

// Records the error text returned by sphk_last_error() and passes the code through.
int set_error(int code, const char *fmt, ...);

}  // namespace sphk
